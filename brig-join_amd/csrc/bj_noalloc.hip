// bj_noalloc.hip -- bj_join_noalloc: the join of device-resident relations with
// no device allocation in the call.  It has no counterpart in the reference; it
// is for callers that cannot call hipMalloc (code under HIP graph capture, tight
// serving loops, harnesses that forbid allocation).
//
// The join's workspace (partitioned R and S, the filter, histograms, cursors,
// work items) is carved from `scratch`, a static device array of
// BJ_NOALLOC_SCRATCH_BYTES that every GPU the library runs on holds until the
// process ends; the caller passes the inputs and the row arrays.  The kernels
// and their sequence are bj_join's.  Counters live in the context (allocated by
// bj_create); everything that must start at zero is cleared by fill kernels in
// the sequence itself, so the array's leftovers from the previous call do not
// matter.  Calls that overlap on one GPU (two host threads, two contexts) share
// the array and race: callers serialise them.
#include "bj_api.h"
#include "bj_internal.h"
#include "bj_kernels.h"

#include <hip/hip_runtime.h>

#include <cstring>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

static_assert(BJ_NOALLOC_SCRATCH_BYTES % 256 == 0, "scratch size");

// Host code never uses this name as an address (on the host it is only a
// stand-in): it uses the device address looked up per device below.
static __device__ __align__(256) uint8_t scratch[BJ_NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int MAX_DEVICES = 64;
constexpr uint64_t ALIGN = 256;
std::mutex g_mu;
void* g_scratch[MAX_DEVICES] = {};  // device address of `scratch` on each device
hipError_t g_lookup[MAX_DEVICES] = {};
bool g_done[MAX_DEVICES] = {};

uint64_t align_up(uint64_t x) { return (x + ALIGN - 1) / ALIGN * ALIGN; }

}  // namespace

namespace bj {

// Looks up the scratch array's address on `device` (the current device) once;
// called by bj_create.
hipError_t noalloc_init(int device)
{
    if (device < 0 || device >= MAX_DEVICES) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_done[device]) {
        void* ptr = nullptr;
        g_lookup[device] = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        g_scratch[device] = g_lookup[device] == hipSuccess ? ptr : nullptr;
        if (g_lookup[device] != hipSuccess) (void)hipGetLastError();
        g_done[device] = true;
    }
    return g_lookup[device];
}

}  // namespace bj

using namespace bj;

extern "C" {

int bj_join_noalloc(bj_ctx* c, const bj_params* p, const uint32_t* r_key, const uint32_t* r_val,
                    const uint32_t* s_key, const uint32_t* s_val, uint32_t* out_key, uint32_t* out_rval,
                    uint32_t* out_sval, uint64_t out_cap, bj_result* out)
{
    if (!c || !p || !out) return fail(BJ_ERR_INVALID, "null argument");
    std::memset(out, 0, sizeof(*out));
    if (!p->inputs_on_device) return fail(BJ_ERR_INVALID, "bj_join_noalloc needs device-resident inputs");
    if (p->s_chunks != 1) return fail(BJ_ERR_INVALID, "bj_join_noalloc joins S in one pass (s_chunks = 1)");
    if (p->probe_breakdown) return fail(BJ_ERR_INVALID, "bj_join_noalloc has no probe phase timers");
    if (out_cap > 0 && (!out_key || !out_rval || !out_sval)) return fail(BJ_ERR_INVALID, "null row array");
    Plan pl;
    int rc = make_plan(*p, pl);
    if (rc) return rc;
    if ((pl.nr && (!r_key || !r_val)) || (pl.ns && (!s_key || !s_val)))
        return fail(BJ_ERR_INVALID, "null relation pointer");
    void* base = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_mu);
        if (c->device < MAX_DEVICES && g_done[c->device]) base = g_scratch[c->device];
    }
    if (!base) return fail(BJ_ERR_HIP, "the scratch array's device address is not known for this device");

    // the workspace, carved from the scratch array (the counters stay the context's)
    std::vector<std::pair<DevBuf bj_ctx::*, uint64_t>> items;
    uint64_t total = 0;
    workspace_items(pl, true, false, [&](DevBuf bj_ctx::*m, uint64_t bytes) -> int {
        if (m != &bj_ctx::counters) {
            items.emplace_back(m, align_up(bytes));
            total += align_up(bytes);
        }
        return BJ_OK;
    });
    if (total > BJ_NOALLOC_SCRATCH_BYTES)
        return fail(BJ_ERR_SIZE, "the join's workspace (" + std::to_string(total) + " B) exceeds the " +
                                     std::to_string((unsigned long long)BJ_NOALLOC_SCRATCH_BYTES) +
                                     " B scratch array of bj_join_noalloc");
    BJ_HIP(hipSetDevice(c->device));
    std::vector<DevBuf> saved;
    saved.reserve(items.size());
    uint64_t off = 0;
    for (const auto& it : items) {
        DevBuf& b = c->*(it.first);
        saved.push_back(b);
        b.p = (uint8_t*)base + off;
        b.bytes = it.second;
        off += it.second;
    }
    pl.cap = out_cap;
    const bool timing = p->collect_phase_times != 0;
    tl_reset(c, timing);
    g_no_alloc = true;
    rc = join_device(c, pl, timing, 0.0, r_key, r_val, s_key, s_val, out_key, out_rval, out_sval, out);
    g_no_alloc = false;
    if (rc != BJ_OK) (void)hipStreamSynchronize(c->stream);  // nothing of the call left in flight
    for (size_t i = 0; i < items.size(); i++) c->*(items[i].first) = saved[i];
    return rc;
}

}  // extern "C"

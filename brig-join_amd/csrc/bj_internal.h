// bj_internal.h -- host-side state shared by the single-join API (bj_api.cpp)
// and the batched join + autotuner (bj_batch.cpp).  Not part of the C ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <cmath>
#include <exception>
#include <functional>
#include <new>
#include <string>
#include <vector>

#include "bj_api.h"
#include "bj_device.h"

namespace bj {

// Records the detail for bj_last_error and returns code.
int fail(int code, const std::string& msg);

// Runs the body of an entry point: a C++ exception (a host allocation that
// fails, a thread that cannot start) becomes an error code instead of crossing
// the C ABI.
template <class F>
int guarded(F&& body) noexcept
{
    try {
        return body();
    } catch (const std::bad_alloc&) {
        return fail(BJ_ERR_NOMEM, "host memory allocation failed");
    } catch (const std::exception& e) {
        return fail(BJ_ERR_INTERNAL, std::string("host exception: ") + e.what());
    } catch (...) {
        return fail(BJ_ERR_INTERNAL, "host exception");
    }
}

// Row capacity max(1, n * max(1, answer_scale)) of a join with n = max(R, S):
// a scale that is not a finite number is refused, and so is a capacity above
// 2^40 rows (12 TiB of rows), which no device holds (and whose conversion to
// an integer would overflow).
constexpr uint64_t MAX_ROWS = 1ull << 40;
inline int row_capacity(uint64_t n, double answer_scale, uint64_t& cap)
{
    if (!std::isfinite(answer_scale)) return fail(BJ_ERR_INVALID, "answer_scale must be a finite number");
    const double c = (double)n * std::max(1.0, answer_scale);
    if (c > (double)MAX_ROWS) return fail(BJ_ERR_SIZE, "row capacity (relation size x answer_scale) above 2^40 rows");
    cap = std::max<uint64_t>(1, (uint64_t)c);
    return BJ_OK;
}

#define BJ_HIP(expr)                                                                         \
    do {                                                                                     \
        hipError_t e_ = (expr);                                                              \
        if (e_ != hipSuccess)                                                                \
            return ::bj::fail(BJ_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
};

// Grows b to at least bytes (contents are not kept); a buffer that grows again
// grows by at least a quarter of its size.
int ensure(DevBuf& b, size_t bytes);
void release(DevBuf& b);
// Set on a host thread while it runs the no-allocation entry point: ensure()
// then refuses to allocate, and no timing events are created.
extern thread_local bool g_no_alloc;

template <typename T>
T* P(DevBuf& b) { return (T*)b.p; }
template <typename T>
const T* P(const DevBuf& b) { return (const T*)b.p; }

uint64_t splitmix64(uint64_t x);
PermKeys perm_keys(uint64_t seed);
int ceil_log2(uint64_t n);

// CTR_ITEMS: probe work items of the last plan (written by the plan kernel);
// CTR_XCD .. CTR_XCD + 7: tile counters of the filter kernel, one per XCD;
// CTR_ACC_SPRIME / CTR_ACC_ITEMS: |S'| and vlsplit items summed over the chunks of S
// CTR_POOL: tuples in the pool of S' (S in chunks with one probe)
enum { CTR_SPRIME = 0, CTR_OUT, CTR_WORK, CTR_ITEMS, CTR_ERR, CTR_ACC_SPRIME, CTR_ACC_ITEMS, CTR_POOL, CTR_XCD = 8,
       CTR_N = 16 };

// Events EV_H2D0/1 bracket the host copies; EV_KB .. are a pool for the
// kernel-slot boundaries (tl_begin / tl_end), EV_N at creation, grown when a
// join records more (one set of S-side slots per chunk of S).
enum { EV_H2D0 = 0, EV_H2D1, EV_KB, EV_N = EV_KB + 2 * BJ_NUM_KERNELS };

// The plan of one join (bj_params -> radix bits, filter geometry, sizes).
struct Plan {
    uint64_t nr = 0, ns = 0;
    int bits = 1, bits1 = 1, bits2 = 0;  // logical levels (partition_size, second level)
    int p1 = 1;                          // bits of the first physical scatter pass (R)
    int p1_s = 1;                        // the same for S
    int scatter_wg = 1024;               // scatter workgroup size
    int G_r = 1, G_s = 1;                // histogram / first-pass scatter workgroups
    bool solo_r = false, solo_s = false; // one scatter workgroup per CU (large relations)
    int tile = 8192;
    uint32_t slice = 65536;
    uint32_t max_slices = 16;
    bool bf = false;
    bool probe_breakdown = false;  // phase timers in the probe kernel
    int bf_stage = 0;
    BloomGeom geom{0, 8, 1};
    uint64_t bf_words = 0;
    uint64_t cap = 1;
    uint64_t item_cap = 1;
};

// A partitioned R: keys / values in sub-bucket order, sub-bucket starts
// (2^bits + 1 entries) and the filter built from it (null without one).
struct RSide {
    const uint32_t* key;
    const uint32_t* val;
    const uint64_t* start;
    const uint64_t* filter;
};

int make_plan(const bj_params& p, Plan& pl);

}  // namespace bj

struct bj_ctx {
    int device = 0;
    int num_cu = 256;
    int max_lds = 64 * 1024;
    hipStream_t own_stream = nullptr;
    hipStream_t stream = nullptr;
    std::vector<hipEvent_t> ev;
    // timeline of the current join: slot k ran from event tl_b[k] to tl_e[k]
    // (-1: no event); one event serves as the end of a slot and the start of the
    // next, and nothing is recorded when the join does not collect times
    bool timing = true;
    int tl_b[BJ_NUM_KERNELS] = {}, tl_e[BJ_NUM_KERNELS] = {};
    int tl_next = bj::EV_KB, tl_last = -1;
    bool tl_open = false;
    // inputs copied from the host (only when inputs_on_device == 0)
    bj::DevBuf in_rk, in_rv, in_sk, in_sv;
    // partitioned relations and scratch
    bj::DevBuf r_key, r_val, s_key, s_val, tmp_key, tmp_val, sp_key, sp_val;
    bj::DevBuf filter, part, totals, l1h, l1off, colsum, r_start, s_start, s_fill, partials, items, counters;
    bj::DevBuf plan_cnt, item_start;
    bj::DevBuf out_key, out_rval, out_sval, dtmp;
    bj::DevBuf ptimers;  // probe phase timers (probe_breakdown)
    // S in chunks with one probe: every chunk's S' (pool_key / pool_val), the
    // piece table (pool_cnt / pool_off, one row per chunk) and the merged S'
    // (fin_key / fin_val at fin_start)
    bj::DevBuf pool_key, pool_val, pool_cnt, pool_off, fin_key, fin_val, fin_start;
    uint64_t* h_counters = nullptr;  // pinned
    // S in chunks: copies of host inputs run on copy_stream, R's first (then
    // ev_r_copied), then the chunks of S into two device buffers (in_sk / in_sv,
    // two chunks each); ev_copied[b] / ev_used[b] order a buffer's copy before
    // its join and its next copy after that join.
    hipStream_t copy_stream = nullptr;
    hipEvent_t ev_copied[2] = {}, ev_used[2] = {}, ev_r_copied = nullptr;
    // autotuner table of the batched join (bj_autotune / bj_tune_set)
    std::vector<bj_tune_entry> tune;
};

namespace bj {

// The device buffers of a join: f(member, bytes) for each.  with_r: the R side
// (partitioned R, its sub-bucket starts, the filter); with_out: the row buffers.
int workspace_items(const Plan& pl, bool with_r, bool with_out,
                    const std::function<int(DevBuf bj_ctx::*, uint64_t)>& f);
// Allocates (or, with c == nullptr, only sums) the device workspace of a join.
int ensure_workspace(bj_ctx* c, const Plan& pl, uint64_t* total_bytes = nullptr);
// R partition (and the filter built from it) into c's r_key / r_val / r_start / filter.
int enqueue_r(bj_ctx* c, const Plan& pl, const uint32_t* r_key, const uint32_t* r_val);
// S (filtered against R's filter when the plan has one) partitioned with R's
// radix bits into o_key / o_val (capacity o_cap; null: c's s_key / s_val) at
// the sub-bucket starts c->s_start; *s_cnt is the survivors per sub-bucket
// where the regions are not full (filter after the first pass, c->s_fill),
// else null.
int enqueue_s(bj_ctx* c, const Plan& pl, const RSide& R, const uint32_t* s_key, const uint32_t* s_val,
              const uint32_t** s_cnt, uint32_t* o_key = nullptr, uint32_t* o_val = nullptr, uint64_t o_cap = 0);
// Work items and the build + probe of a partitioned S (sk / sv at s_start, with
// s_cnt tuples per sub-bucket when not null) against R; rows to o_* (capacity cap).
int enqueue_plan_probe(bj_ctx* c, const Plan& pl, const RSide& R, const uint32_t* sk, const uint32_t* sv,
                       const uint64_t* s_start, const uint32_t* s_cnt, uint32_t* o_key, uint32_t* o_rval,
                       uint32_t* o_sval, uint64_t cap);
// enqueue_s, then enqueue_plan_probe on its output.
int enqueue_s_probe(bj_ctx* c, const Plan& pl, const RSide& R, const uint32_t* s_key, const uint32_t* s_val,
                    uint32_t* o_key, uint32_t* o_rval, uint32_t* o_sval, uint64_t cap);
// One-pass join of device-resident inputs into o_* (capacity pl.cap), with the
// workspace already in place; h2d_ms < 0: host copies bracketed by EV_H2D0/1.
int join_device(bj_ctx* c, const Plan& pl, bool timing, double h2d_ms, const uint32_t* r_key,
                const uint32_t* r_val, const uint32_t* s_key, const uint32_t* s_val, uint32_t* o_key,
                uint32_t* o_rval, uint32_t* o_sval, bj_result* out);
// Looks up the no-allocation entry point's scratch array on `device` (the current
// device), once per device (bj_create calls it).
hipError_t noalloc_init(int device);
float ev_ms(bj_ctx* c, int a, int b);
// Kernel-slot timeline (see bj_ctx::tl_*).
void tl_reset(bj_ctx* c, bool timing);
hipError_t tl_begin(bj_ctx* c, int slot);
hipError_t tl_end(bj_ctx* c, int slot);
float tl_ms(bj_ctx* c, int slot);
float tl_span_ms(bj_ctx* c, int first, int last);

}  // namespace bj

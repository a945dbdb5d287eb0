// bj_api.cpp -- host orchestration of the gfx950 radix hash join behind the C ABI.
//
// Counterparts in the reference (Chanho-LEE-SNU/BRIG-Join):
//   make_plan      parameter derivation of gpu_main (src/hash_kernel.cu:174-284):
//                  Bloom size/hash count formula (:188-196), partition_size and
//                  second_partition_size sizing (:198-284)
//   run_partition  partial_jointable_gpu_partition{,_r_only,_s_only}
//                  (src/vlsplit_hash_kernel.cu:10-113, 224-297)
//   bj_join        the device phase of gpu_main (src/hash_kernel.cu:636-1131) and
//                  join_execute_host's local branch (src/join_executor.cu:14-110)
// Unlike the reference, no host synchronisation happens between phases: every
// size the device produces (|S'|, work items) is read by the next kernel from
// device memory, and the host reads the counters once at the end.
#include "bj_api.h"
#include "bj_batch.h"
#include "bj_internal.h"
#include "bj_kernels.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <string>

using namespace bj;

namespace {

thread_local std::string g_last_error;

}  // namespace

namespace bj {
int fail(int code, const std::string& msg)
{
    g_last_error = msg;
    return code;
}
}  // namespace bj

uint64_t bj::splitmix64(uint64_t x)
{
    uint64_t z = x + 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

PermKeys bj::perm_keys(uint64_t seed)
{
    PermKeys K;
    uint64_t st = seed;
    for (int r = 0; r < 4; r++) {
        st = splitmix64(st);
        K.k[2 * r] = (uint32_t)st;
        st = splitmix64(st);
        K.k[2 * r + 1] = (uint32_t)(st >> 32) | 1u;
    }
    return K;
}

int bj::ceil_log2(uint64_t n)
{
    int l = 0;
    while ((1ull << l) < n) l++;
    return l;
}

namespace {

bool is_pow2(int64_t x) { return x > 0 && (x & (x - 1)) == 0; }

}  // namespace

thread_local bool bj::g_no_alloc = false;

int bj::ensure(DevBuf& b, size_t bytes)
{
    if (bytes == 0) bytes = 256;
    if (b.bytes >= bytes) return BJ_OK;
    if (g_no_alloc) return fail(BJ_ERR_INTERNAL, "device allocation inside the no-allocation entry point");
    // a buffer that has to grow again grows by at least a quarter, so a caller
    // whose joins keep getting larger reallocates a few times, not every call
    if (b.p) bytes = std::max<size_t>(bytes, b.bytes + b.bytes / 4);
    if (b.p) {
        hipError_t e = hipFree(b.p);
        (void)e;
        b.p = nullptr;
        b.bytes = 0;
    }
    hipError_t e = hipMalloc(&b.p, bytes);
    if (e != hipSuccess) {
        b.p = nullptr;
        (void)hipGetLastError();
        return fail(BJ_ERR_NOMEM, "hipMalloc(" + std::to_string(bytes) + " B) failed: " +
                                      hipGetErrorString(e));
    }
    b.bytes = bytes;
    return BJ_OK;
}

void bj::release(DevBuf& b)
{
    if (b.p) {
        hipError_t e = hipFree(b.p);
        (void)e;
    }
    b.p = nullptr;
    b.bytes = 0;
}

namespace bj {

constexpr uint64_t MAX_TUPLES = 0x7FFFFFFFull;  // row ids are 31-bit in the probe's staging
constexpr int MAX_BITS = 15;                     // joint histogram in LDS: 2^15 x 4 B
constexpr int MAX_LEVEL_BITS = 11;               // fan-out per scatter pass
constexpr uint64_t SEG_TARGET_BYTES = 64 * 1024; // Bloom segment built in LDS
constexpr uint64_t SEG_MAX_BYTES = 160 * 1024;
// A filter slice probed at random stays in one XCD's 4 MB L2 up to about this size:
// a smaller filter is probed before S is partitioned (bf_stage 1), a larger one
// slice by slice after S's first pass (bf_stage 2), with enough first-pass bits.
constexpr uint64_t L2_SLICE_BYTES = 3584ull << 10;

// Relations of at least this many tuples are scattered by one workgroup per CU
// (one chunk each): with several 128-way write-combining workgroups per CU the
// HBM write streams interleave more and a 640M-tuple pass runs 10-15% slower,
// while smaller relations need the extra workgroups to hide latency.
constexpr uint64_t SOLO_TUPLES = 32ull << 20;
constexpr int PLAN_CUS = 256;  // CUs of an MI355X: only the work split depends on it

int choose_G(uint64_t n)
{
    if (n >= SOLO_TUPLES) return PLAN_CUS;
    uint64_t g = (n + 32767) / 32768;
    return (int)std::max<uint64_t>(1, std::min<uint64_t>(512, g));
}

void split_bits(int total, int& b1, int& b2)
{
    if (total <= MAX_LEVEL_BITS) {
        b1 = total;
        b2 = 0;
    } else {
        b1 = (total + 1) / 2;
        b2 = total - b1;
    }
}

int make_plan(const bj_params& p, Plan& pl)
{
    if (p.r_size < 0 || p.s_size < 0) return fail(BJ_ERR_INVALID, "negative relation size");
    if ((uint64_t)p.r_size > MAX_TUPLES || (uint64_t)p.s_size > MAX_TUPLES)
        return fail(BJ_ERR_SIZE, "relation larger than 2^31-1 tuples");
    pl.nr = (uint64_t)p.r_size;
    pl.ns = (uint64_t)p.s_size;

    pl.tile = p.tile > 0 ? p.tile : 8192;
    if (pl.tile != 4096 && pl.tile != 8192) return fail(BJ_ERR_INVALID, "tile must be 4096 or 8192");

    // use_bf modes of the reference: 0..16 except 11 and 12 (src/hash_kernel.cu:146-150)
    if (p.use_bf < 0 || p.use_bf == 11 || p.use_bf == 12 || p.use_bf > 16)
        return fail(BJ_ERR_INVALID, "unsupported use_bf=" + std::to_string(p.use_bf));
    pl.bf = p.use_bf != 0;
    pl.probe_breakdown = p.probe_breakdown != 0;
    if (pl.bf && !std::isfinite(p.bf_bits_per_key))
        return fail(BJ_ERR_INVALID, "bf_bits_per_key must be a finite number");
    if (p.s_chunks < 0) return fail(BJ_ERR_INVALID, "s_chunks must be >= 0");

    int b1 = -1, b2 = -1;
    if (p.partition_size >= 0) {
        if (!is_pow2(p.partition_size) || p.partition_size < 2 || p.partition_size > 2048)
            return fail(BJ_ERR_INVALID, "partition_size must be a power of two in [2, 2048]");
        b1 = ceil_log2((uint64_t)p.partition_size);
    }
    if (p.second_partition_size >= 0) {
        if (!is_pow2(p.second_partition_size) || p.second_partition_size > 2048)
            return fail(BJ_ERR_INVALID, "second_partition_size must be a power of two in [1, 2048]");
        b2 = ceil_log2((uint64_t)p.second_partition_size);
    }

    // Bloom geometry first: its segment count can raise the automatic radix bits.
    double num_bits = 0.0;
    int k = 0;
    if (pl.bf) {
        const double R = (double)std::max<uint64_t>(pl.nr, 1);
        if (p.num_bits >= 0) {
            num_bits = (double)p.num_bits;
        } else if (p.bf_bits_per_key > 0) {
            num_bits = R * p.bf_bits_per_key;
        } else {
            // reference formula, src/hash_kernel.cu:189-193
            const double fp = p.true_false_ratio;
            if (!(fp > 0.0 && fp < 1.0)) return fail(BJ_ERR_INVALID, "true_false_ratio must be in (0,1)");
            num_bits = -R * std::log(fp) / (std::log(2.0) * std::log(2.0));
        }
        // the derived counts are clamped to [1, 8] as doubles, before the
        // conversion to int (a filter of many bits per key would overflow it)
        if (p.num_hashes >= 0) {
            k = p.num_hashes;
        } else if (p.num_bits < 0 && p.bf_bits_per_key > 0) {
            k = (int)std::min(8.0, std::round(p.bf_bits_per_key * std::log(2.0)));
        } else {
            k = (int)std::min(8.0, std::max(0.0, num_bits / R * std::log(2.0)));
        }
        if (k < 1) k = 1;
        if (p.num_hashes > 8) return fail(BJ_ERR_INVALID, "num_hashes must be in [1, 8]");
        if (k > 8) k = 8;
        if (num_bits < 256.0) num_bits = 256.0;
    }
    const uint64_t bf_bytes = pl.bf ? (uint64_t)std::ceil(num_bits / 8.0) : 0;
    const int seg_need = pl.bf ? std::max(0, ceil_log2((bf_bytes + SEG_TARGET_BYTES - 1) / SEG_TARGET_BYTES)) : 0;

    // radix bits
    int total;
    if (p.radix_bits > 0) {
        total = p.radix_bits;
        if (b1 >= 0 && b2 >= 0 && b1 + b2 != total)
            return fail(BJ_ERR_INVALID, "radix_bits != log2(partition_size * second_partition_size)");
        if (b1 >= 0 && b2 < 0) b2 = total - b1;
        else if (b2 >= 0 && b1 < 0) b1 = total - b2;
        else if (b1 < 0 && b2 < 0) split_bits(total, b1, b2);
        if (b1 < 1 || b2 < 0) return fail(BJ_ERR_INVALID, "radix_bits smaller than the given fan-out");
    } else {
        // smallest bits with min(R, S) / 2^bits <= 0.96 * tile: the mean sub-bucket
        // then lies over 4 standard deviations below the tile for uniform keys at
        // the sizes where the margin matters (16M: 11 bits, 0.466 against 0.487 ms
        // at 12; 64M x 1.28B: 13 bits, 12.81 against 13.07 ms at 14)
        const uint64_t m = std::max<uint64_t>(1, std::min(pl.nr, pl.ns));
        const double per = 0.96 * pl.tile;
        int need = std::max(1, ceil_log2((uint64_t)std::ceil((double)m / per)));
        need = std::max(need, seg_need);
        need = std::min(need, MAX_BITS);
        if (b1 >= 0 && b2 >= 0) {
        } else if (b1 >= 0) {
            b2 = std::max(0, need - b1);
        } else if (b2 >= 0) {
            b1 = std::max(1, need - b2);
        } else {
            split_bits(need, b1, b2);
        }
        total = b1 + b2;
    }
    if (total < 1 || total > MAX_BITS)
        return fail(BJ_ERR_SIZE, "total radix bits must be in [1, 15] (got " + std::to_string(total) + ")");
    if (b1 > MAX_LEVEL_BITS || b2 > MAX_LEVEL_BITS)
        return fail(BJ_ERR_SIZE, "a partition level exceeds 2048-way fan-out");
    pl.bits = total;
    pl.bits1 = b1;
    pl.bits2 = b2;
    if (p.pass1_bits > 0) {
        if (p.pass1_bits > total) return fail(BJ_ERR_INVALID, "pass1_bits exceeds the total radix bits");
        pl.p1 = p.pass1_bits;
    } else {
        pl.p1 = total <= 7 ? total : (total + 1) / 2;
    }
    if (pl.p1 > MAX_LEVEL_BITS || total - pl.p1 > MAX_LEVEL_BITS)
        return fail(BJ_ERR_SIZE, "a scatter pass exceeds 2048-way fan-out");

    if (pl.bf) {
        int seg_bits = std::min(seg_need, total);
        const uint64_t blocks = (bf_bytes + 7) / 8;  // 64-bit blocks
        uint64_t wps = std::max<uint64_t>(1, (blocks + (1ull << seg_bits) - 1) >> seg_bits);
        if (wps * 8 > SEG_MAX_BYTES)
            return fail(BJ_ERR_SIZE, "Bloom segment exceeds LDS: raise radix bits or lower filter bits");
        pl.geom.seg_bits = seg_bits;
        pl.geom.wps = (uint32_t)wps;
        pl.geom.k = k;
        pl.bf_words = (wps << seg_bits) * 2;  // in 32-bit words
        if (p.bf_stage < 0 || p.bf_stage > 2) return fail(BJ_ERR_INVALID, "bf_stage must be 0, 1 or 2");
        pl.bf_stage = p.bf_stage ? p.bf_stage : (pl.bf_words * 4 <= L2_SLICE_BYTES ? 1 : 2);
    }
    // first-pass bits of S
    if (p.s_pass1_bits > 0) {
        if (p.s_pass1_bits > total) return fail(BJ_ERR_INVALID, "s_pass1_bits exceeds the total radix bits");
        pl.p1_s = p.s_pass1_bits;
    } else if (pl.bf && pl.bf_stage == 2) {
        const int need = ceil_log2((pl.bf_words * 4 + L2_SLICE_BYTES - 1) / L2_SLICE_BYTES);
        pl.p1_s = std::max(pl.p1, std::min(std::min(total, 9), need));
    } else {
        pl.p1_s = pl.p1;
    }
    if (pl.p1_s > MAX_LEVEL_BITS || total - pl.p1_s > MAX_LEVEL_BITS)
        return fail(BJ_ERR_SIZE, "a scatter pass of S exceeds 2048-way fan-out");
    if (p.scatter_wg != 0 && p.scatter_wg != 512 && p.scatter_wg != 1024)
        return fail(BJ_ERR_INVALID, "scatter_wg must be 0, 512 or 1024");
    pl.scatter_wg = p.scatter_wg ? p.scatter_wg : 1024;

    pl.G_r = choose_G(pl.nr);
    pl.G_s = choose_G(pl.ns);
    pl.solo_r = pl.nr >= SOLO_TUPLES;
    pl.solo_s = pl.ns >= SOLO_TUPLES;
    if (const int rc = row_capacity(std::max(pl.nr, pl.ns), p.answer_scale, pl.cap)) return rc;
    // sum over sub-buckets of ceil(nb / tile) <= (nr + ns) / tile + 2^bits
    pl.item_cap = (uint64_t)pl.max_slices * ((pl.nr + pl.ns) / pl.tile + (1ull << total)) + 16;
    return BJ_OK;
}

// Radix partition of one relation into out_key/out_val with sub-bucket starts
// in start[0 .. 2^bits], in one or two physical scatter passes (p1 bits, then
// the rest).  With l1_only, only the first pass runs and its output (grouped by
// first-pass bucket, in the final positions' ranges) is left in out_key/out_val.
int run_partition(bj_ctx* c, const Plan& pl, int p1, const uint32_t* key, const uint32_t* val,
                  uint64_t n_host, const uint64_t* n_dev, int G, bool solo, uint32_t* out_key,
                  uint32_t* out_val, uint64_t out_cap, uint64_t* start, int k_hist, bool l1_only = false)
{
    // more than half of the CU's LDS: one scatter workgroup per CU
    const size_t min_lds = solo ? (size_t)c->max_lds / 2 + 1024 : 0;
    // k_hist is the BJ_K_HIST_* slot; the three slots after it follow in order
    auto beg = [&](int slot) { return tl_begin(c, slot); };
    auto end = [&](int slot) { return tl_end(c, slot); };
    hipStream_t st = c->stream;
    const int b2 = pl.bits - p1;
    uint64_t* err = P<uint64_t>(c->counters) + CTR_ERR;
    uint32_t* part = P<uint32_t>(c->part);
    BJ_HIP(beg(k_hist));
    BJ_HIP(launch_hist(key, n_host, n_dev, G, pl.bits, part, P<uint32_t>(c->totals), b2, P<uint32_t>(c->l1h), st));
    BJ_HIP(end(k_hist));
    BJ_HIP(beg(k_hist + 1));
    const bool two_pass = b2 > 0 && !l1_only;
    // enough second-pass workgroups to fill the chip: about 1024 in total
    const int W2 = std::min(std::max(1, std::min(64, 1024 >> p1)), G);
    // chunk groups of the column sums: the second pass's groups, else up to 8
    const int W = two_pass ? W2 : std::min(8, G);
    BJ_HIP(launch_offsets(part, G, pl.bits, b2, W, P<uint32_t>(c->colsum), P<uint32_t>(c->totals),
                          P<uint32_t>(c->l1h), start,
                          P<uint64_t>(c->l1off), st));
    BJ_HIP(end(k_hist + 1));
    uint32_t* k1 = two_pass ? P<uint32_t>(c->tmp_key) : out_key;
    uint32_t* v1 = two_pass ? P<uint32_t>(c->tmp_val) : out_val;
    const uint64_t cap1 = two_pass ? c->tmp_key.bytes / 4 : out_cap;
    BJ_HIP(beg(k_hist + 2));
    BJ_HIP(launch_scatter(key, val, k1, v1, cap1, n_host, n_dev, G, 0, pl.bits, p1, W2,
                          P<uint64_t>(c->l1off), start, P<uint32_t>(c->colsum), err, pl.scatter_wg, min_lds, st));
    BJ_HIP(end(k_hist + 2));
    BJ_HIP(beg(k_hist + 3));
    if (two_pass)
        BJ_HIP(launch_scatter(k1, v1, out_key, out_val, out_cap, n_host, n_dev, G, 1, pl.bits, p1, W2,
                              P<uint64_t>(c->l1off), start, P<uint32_t>(c->colsum), err, pl.scatter_wg, min_lds, st));
    BJ_HIP(end(k_hist + 3));
    return BJ_OK;
}

int workspace_items(const Plan& pl, bool with_r, bool with_out,
                    const std::function<int(DevBuf bj_ctx::*, uint64_t)>& f)
{
    int rc;
    auto need = [&](DevBuf bj_ctx::*m, uint64_t bytes) { return f(m, std::max<uint64_t>(bytes, 256)); };
    const uint64_t nr = std::max<uint64_t>(pl.nr, 1), ns = std::max<uint64_t>(pl.ns, 1);
    const uint64_t nbins = 1ull << pl.bits;
    const uint64_t gmax = (uint64_t)(with_r ? std::max(pl.G_r, pl.G_s) : pl.G_s);
    const uint64_t n_l1 = (1ull << std::max(pl.p1, pl.p1_s)) * gmax;
    const uint64_t w2max = (uint64_t)std::max(1, std::min(64, 1024 >> std::min(pl.p1, pl.p1_s)));
    if (with_r) {
        if ((rc = need(&bj_ctx::r_key, nr * 4)) || (rc = need(&bj_ctx::r_val, nr * 4))) return rc;
        if ((rc = need(&bj_ctx::r_start, (nbins + 1) * 8))) return rc;
        if (pl.bf && (rc = need(&bj_ctx::filter, pl.bf_words * 4))) return rc;
    }
    if ((rc = need(&bj_ctx::s_key, ns * 4)) || (rc = need(&bj_ctx::s_val, ns * 4))) return rc;
    if (pl.p1 < pl.bits || pl.p1_s < pl.bits || (pl.bf && pl.bf_stage == 2)) {
        const uint64_t t = with_r ? std::max(nr, ns) : ns;
        if ((rc = need(&bj_ctx::tmp_key, t * 4)) || (rc = need(&bj_ctx::tmp_val, t * 4))) return rc;
    }
    if (pl.bf) {
        if (pl.bf_stage == 1) {
            if ((rc = need(&bj_ctx::sp_key, ns * 4)) || (rc = need(&bj_ctx::sp_val, ns * 4))) return rc;
        }
        if ((rc = need(&bj_ctx::s_fill, nbins * 4))) return rc;
    }
    if ((rc = need(&bj_ctx::part, gmax * nbins * 4))) return rc;
    if ((rc = need(&bj_ctx::l1h, n_l1 * 4))) return rc;
    if ((rc = need(&bj_ctx::l1off, n_l1 * 8))) return rc;
    if ((rc = need(&bj_ctx::totals, nbins * 4))) return rc;
    if ((rc = need(&bj_ctx::colsum, std::max<uint64_t>(w2max, 8) * nbins * 4))) return rc;
    if ((rc = need(&bj_ctx::plan_cnt, nbins * 4))) return rc;
    if ((rc = need(&bj_ctx::item_start, (nbins + 1) * 8))) return rc;
    if ((rc = need(&bj_ctx::s_start, (nbins + 1) * 8))) return rc;
    if ((rc = need(&bj_ctx::partials, scan_partials_needed(nbins) * 8 + 64))) return rc;
    if ((rc = need(&bj_ctx::items, pl.item_cap * sizeof(HeavyItem)))) return rc;
    if ((rc = need(&bj_ctx::counters, CTR_N * sizeof(uint64_t)))) return rc;
    if (with_out) {
        if ((rc = need(&bj_ctx::out_key, pl.cap * 4)) || (rc = need(&bj_ctx::out_rval, pl.cap * 4)) ||
            (rc = need(&bj_ctx::out_sval, pl.cap * 4)))
            return rc;
    }
    return BJ_OK;
}

int ensure_workspace(bj_ctx* c, const Plan& pl, uint64_t* total_bytes)
{
    uint64_t total = 0;
    int rc = workspace_items(pl, true, true, [&](DevBuf bj_ctx::*m, uint64_t bytes) -> int {
        total += bytes;
        return c ? ensure(c->*m, bytes) : BJ_OK;
    });
    if (rc) return rc;
    if (total_bytes) *total_bytes = total;
    return BJ_OK;
}

float ev_ms(bj_ctx* c, int a, int b)
{
    float ms = 0.f;
    if (a < 0 || b < 0) return 0.f;
    if (hipEventElapsedTime(&ms, c->ev[a], c->ev[b]) != hipSuccess) {
        (void)hipGetLastError();
        return 0.f;
    }
    return ms;
}

void tl_reset(bj_ctx* c, bool timing)
{
    c->timing = timing;
    for (int i = 0; i < BJ_NUM_KERNELS; i++) c->tl_b[i] = c->tl_e[i] = -1;
    c->tl_next = EV_KB;
    c->tl_last = -1;
    c->tl_open = false;
}

static hipError_t tl_mark(bj_ctx* c, int& idx)
{
    if (c->tl_next >= (int)c->ev.size()) {
        if (g_no_alloc) {  // no new events there: the slot reads as 0 ms
            idx = -1;
            return hipSuccess;
        }
        hipEvent_t e = nullptr;
        if (hipEventCreate(&e) != hipSuccess) {
            (void)hipGetLastError();
            idx = -1;  // no event: the slot reads as 0 ms
            return hipSuccess;
        }
        c->ev.push_back(e);
    }
    idx = c->tl_next++;
    c->tl_last = idx;
    return hipEventRecord(c->ev[idx], c->stream);
}

// A slot that starts right after another one ends (nothing enqueued between)
// starts at that slot's end event.
hipError_t tl_begin(bj_ctx* c, int slot)
{
    if (!c->timing) return hipSuccess;
    if (c->tl_open) {
        c->tl_b[slot] = c->tl_last;
        c->tl_open = false;
        return hipSuccess;
    }
    return tl_mark(c, c->tl_b[slot]);
}

hipError_t tl_end(bj_ctx* c, int slot)
{
    if (!c->timing) return hipSuccess;
    const hipError_t e = tl_mark(c, c->tl_e[slot]);
    c->tl_open = true;
    return e;
}

float tl_ms(bj_ctx* c, int slot) { return ev_ms(c, c->tl_b[slot], c->tl_e[slot]); }

float tl_span_ms(bj_ctx* c, int first, int last) { return ev_ms(c, c->tl_b[first], c->tl_e[last]); }

int enqueue_r(bj_ctx* c, const Plan& pl, const uint32_t* r_key, const uint32_t* r_val)
{
    int rc;
    if ((rc = run_partition(c, pl, pl.p1, r_key, r_val, pl.nr, nullptr, pl.G_r, pl.solo_r, P<uint32_t>(c->r_key),
                            P<uint32_t>(c->r_val), c->r_key.bytes / 4, P<uint64_t>(c->r_start), BJ_K_HIST_R)))
        return rc;
    if (pl.bf) {
        BJ_HIP(tl_begin(c, BJ_K_BF_BUILD));
        BJ_HIP(launch_bloom_build(P<uint32_t>(c->r_key), P<uint64_t>(c->r_start), pl.bits, pl.geom,
                                  P<uint64_t>(c->filter), c->stream));
        BJ_HIP(tl_end(c, BJ_K_BF_BUILD));
    }
    return BJ_OK;
}

int enqueue_s(bj_ctx* c, const Plan& pl, const RSide& R, const uint32_t* s_key, const uint32_t* s_val,
              const uint32_t** s_cnt_out, uint32_t* o_key, uint32_t* o_val, uint64_t o_cap)
{
    if (!o_key) {
        o_key = P<uint32_t>(c->s_key);
        o_val = P<uint32_t>(c->s_val);
        o_cap = c->s_key.bytes / 4;
    }
    hipStream_t st = c->stream;
    uint64_t* ctr = P<uint64_t>(c->counters);
    auto beg = [&](int slot) { return tl_begin(c, slot); };
    auto end = [&](int slot) { return tl_end(c, slot); };
    int rc;
    const uint32_t* s_cnt = nullptr;
    if (pl.bf) {
        if (pl.bf_stage == 1) {
            // the reference's order: filter S, then partition the survivors
            BJ_HIP(beg(BJ_K_BF_FILTER));
            BJ_HIP(launch_bloom_prefilter(s_key, s_val, pl.ns, R.filter, pl.geom, P<uint32_t>(c->sp_key),
                                          P<uint32_t>(c->sp_val), c->sp_key.bytes / 4, ctr + CTR_SPRIME,
                                          ctr + CTR_ERR, c->num_cu, st));
            BJ_HIP(end(BJ_K_BF_FILTER));
            if ((rc = run_partition(c, pl, pl.p1_s, P<uint32_t>(c->sp_key), P<uint32_t>(c->sp_val), pl.ns,
                                    ctr + CTR_SPRIME,
                                    pl.G_s, pl.solo_s, o_key, o_val, o_cap, P<uint64_t>(c->s_start), BJ_K_HIST_S)))
                return rc;
        } else {
            // first-level pass of S, then filter each bucket against its L2-resident
            // filter slice while placing the survivors into their sub-buckets
            if ((rc = run_partition(c, pl, pl.p1_s, s_key, s_val, pl.ns, nullptr, pl.G_s, pl.solo_s, P<uint32_t>(c->tmp_key),
                                    P<uint32_t>(c->tmp_val), c->tmp_key.bytes / 4, P<uint64_t>(c->s_start),
                                    BJ_K_HIST_S, true)))
                return rc;
            BJ_HIP(beg(BJ_K_BF_FILTER));
            BJ_HIP(launch_zero_u32(P<uint32_t>(c->s_fill), 1ull << pl.bits, st));
            BJ_HIP(launch_filter_l1(P<uint32_t>(c->tmp_key), P<uint32_t>(c->tmp_val), P<uint64_t>(c->s_start),
                                    pl.bits, pl.bits - pl.p1_s, R.filter, pl.geom, o_key, o_val, o_cap,
                                    P<uint32_t>(c->s_fill),
                                    ctr + CTR_SPRIME, ctr + CTR_XCD, ctr + CTR_ERR, c->num_cu, st));
            s_cnt = P<uint32_t>(c->s_fill);
            BJ_HIP(end(BJ_K_BF_FILTER));
        }
    } else {
        if ((rc = run_partition(c, pl, pl.p1_s, s_key, s_val, pl.ns, nullptr, pl.G_s, pl.solo_s, o_key, o_val, o_cap,
                                P<uint64_t>(c->s_start), BJ_K_HIST_S)))
            return rc;
    }
    *s_cnt_out = s_cnt;
    return BJ_OK;
}

int enqueue_plan_probe(bj_ctx* c, const Plan& pl, const RSide& R, const uint32_t* sk, const uint32_t* sv,
                       const uint64_t* s_start, const uint32_t* s_cnt, uint32_t* o_key, uint32_t* o_rval,
                       uint32_t* o_sval, uint64_t cap)
{
    hipStream_t st = c->stream;
    uint64_t* ctr = P<uint64_t>(c->counters);
    const int nbins = 1 << pl.bits;
    uint64_t* item_start = P<uint64_t>(c->item_start);
    BJ_HIP(tl_begin(c, BJ_K_PLAN));
    BJ_HIP(launch_plan(R.start, s_start, s_cnt, pl.bits, (uint32_t)pl.tile, pl.slice,
                       pl.max_slices, P<uint32_t>(c->plan_cnt), item_start, ctr + CTR_ITEMS,
                       P<HeavyItem>(c->items), pl.item_cap, ctr + CTR_ERR, st));
    BJ_HIP(tl_end(c, BJ_K_PLAN));
    BJ_HIP(tl_begin(c, BJ_K_PROBE));
    BJ_HIP(launch_probe(pl.tile, R.key, R.val, R.start, sk, sv, s_start, s_cnt, pl.bits, pl.slice,
                        P<HeavyItem>(c->items), item_start + nbins, pl.item_cap, ctr + CTR_WORK, ctr + CTR_OUT,
                        cap, o_key, o_rval, o_sval, c->num_cu,
                        pl.probe_breakdown ? P<uint64_t>(c->ptimers) : nullptr, st));
    BJ_HIP(tl_end(c, BJ_K_PROBE));
    return BJ_OK;
}

int enqueue_s_probe(bj_ctx* c, const Plan& pl, const RSide& R, const uint32_t* s_key, const uint32_t* s_val,
                    uint32_t* o_key, uint32_t* o_rval, uint32_t* o_sval, uint64_t cap)
{
    const uint32_t* s_cnt = nullptr;
    int rc = enqueue_s(c, pl, R, s_key, s_val, &s_cnt);
    if (rc) return rc;
    return enqueue_plan_probe(c, pl, R, P<uint32_t>(c->s_key), P<uint32_t>(c->s_val), P<uint64_t>(c->s_start),
                              s_cnt, o_key, o_rval, o_sval, cap);
}

}  // namespace bj

// ===================================================================== C ABI
extern "C" {

int bj_version(void) { return BJ_API_VERSION; }

const char* bj_strerror(int code)
{
    switch (code) {
    case BJ_OK: return "ok";
    case BJ_ERR_INVALID: return "invalid argument";
    case BJ_ERR_HIP: return "HIP runtime error";
    case BJ_ERR_NOMEM: return "out of memory";
    case BJ_ERR_IO: return "I/O error";
    case BJ_ERR_SIZE: return "size exceeds a supported limit";
    case BJ_ERR_NODEVICE: return "no such HIP device";
    case BJ_ERR_INTERNAL: return "device-side consistency check failed";
    default: return "unknown error";
    }
}

const char* bj_last_error(void) { return g_last_error.c_str(); }

void bj_params_default(bj_params* p)
{
    if (!p) return;
    std::memset(p, 0, sizeof(*p));
    p->partition_size = -1;
    p->second_partition_size = -1;
    p->radix_bits = -1;
    p->use_bf = 0;
    p->bf_bits_per_key = 0.0;
    p->true_false_ratio = 0.1;
    p->num_bits = -1;
    p->num_hashes = -1;
    p->answer_scale = 1.0;
    p->tile = 0;
    p->inputs_on_device = 0;
    p->collect_phase_times = 1;
    p->s_chunks = 1;
    p->bf_stage = 0;
    p->pass1_bits = -1;
    p->s_pass1_bits = -1;
    p->scatter_wg = 0;
    p->probe_breakdown = 0;
}

int bj_create(bj_ctx** out, int device)
{
    return guarded([&]() -> int {
        if (!out) return fail(BJ_ERR_INVALID, "null ctx pointer");
        *out = nullptr;
        int ndev = 0;
        if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
            (void)hipGetLastError();
            return fail(BJ_ERR_NODEVICE, "no HIP device");
        }
        if (device < 0 || device >= ndev) return fail(BJ_ERR_NODEVICE, "device index out of range");
        BJ_HIP(hipSetDevice(device));
        bj_ctx* c = new bj_ctx();
        // a throw past this point (the event vector) destroys the half-built context
        try {
            c->device = device;
            int cu = 0;
            if (hipDeviceGetAttribute(&cu, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess && cu > 0)
                c->num_cu = cu;
            int max_lds = 64 * 1024;
            if (hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, device) != hipSuccess)
                max_lds = 64 * 1024;
            c->max_lds = max_lds;
            hipError_t e = hipStreamCreateWithFlags(&c->own_stream, hipStreamNonBlocking);
            if (e == hipSuccess) e = kernels_init(max_lds);
            // the scratch array of bj_join_noalloc on this device (its failure only
            // disables that entry point, which then reports it)
            if (e == hipSuccess) (void)noalloc_init(device);
            if (e == hipSuccess) e = batch_kernels_init(max_lds);
            c->ev.assign(EV_N, nullptr);
            for (int i = 0; i < EV_N && e == hipSuccess; i++) e = hipEventCreate(&c->ev[i]);
            if (e == hipSuccess)
                e = hipHostMalloc((void**)&c->h_counters, (CTR_N + 1) * sizeof(uint64_t), hipHostMallocDefault);
            if (e != hipSuccess) {
                bj_destroy(c);
                return fail(BJ_ERR_HIP, std::string("context setup: ") + hipGetErrorString(e));
            }
            c->stream = c->own_stream;
            int rc = ensure(c->counters, CTR_N * sizeof(uint64_t));
            if (rc == BJ_OK) rc = ensure(c->dtmp, 64);
            if (rc != BJ_OK) {
                bj_destroy(c);
                return rc;
            }
        } catch (...) {
            bj_destroy(c);
            throw;
        }
        *out = c;
        return BJ_OK;
    });
}

int bj_destroy(bj_ctx* c)
{
    if (!c) return BJ_OK;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    DevBuf* bufs[] = {&c->in_rk,  &c->in_rv,    &c->in_sk,   &c->in_sv,    &c->r_key,   &c->r_val,
                      &c->s_key,  &c->s_val,    &c->tmp_key, &c->tmp_val,  &c->sp_key,  &c->sp_val,
                      &c->filter, &c->part,     &c->totals,  &c->l1h,      &c->l1off,   &c->colsum,  &c->r_start,
                      &c->s_start, &c->s_fill,  &c->partials, &c->items,  &c->counters, &c->out_key,
                      &c->out_rval, &c->out_sval, &c->dtmp,  &c->plan_cnt, &c->item_start, &c->ptimers,
                      &c->pool_key, &c->pool_val, &c->pool_cnt, &c->pool_off, &c->fin_key, &c->fin_val,
                      &c->fin_start};
    for (DevBuf* b : bufs) release(*b);
    for (hipEvent_t ev : c->ev)
        if (ev) (void)hipEventDestroy(ev);
    for (int b = 0; b < 2; b++) {
        if (c->ev_copied[b]) (void)hipEventDestroy(c->ev_copied[b]);
        if (c->ev_used[b]) (void)hipEventDestroy(c->ev_used[b]);
    }
    if (c->ev_r_copied) (void)hipEventDestroy(c->ev_r_copied);
    if (c->copy_stream) (void)hipStreamDestroy(c->copy_stream);
    if (c->h_counters) (void)hipHostFree(c->h_counters);
    if (c->own_stream) (void)hipStreamDestroy(c->own_stream);
    delete c;
    return BJ_OK;
}

int bj_set_stream(bj_ctx* c, void* s)
{
    if (!c) return fail(BJ_ERR_INVALID, "null ctx");
    c->stream = s ? (hipStream_t)s : c->own_stream;
    return BJ_OK;
}

int bj_plan(const bj_params* p, bj_plan_info* out)
{
    return guarded([&]() -> int {
        if (!p || !out) return fail(BJ_ERR_INVALID, "null argument");
        std::memset(out, 0, sizeof(*out));
        Plan pl;
        int rc = make_plan(*p, pl);
        if (rc) return rc;
        uint64_t ws = 0;
        ensure_workspace(nullptr, pl, &ws);
        if (!p->inputs_on_device) ws += (pl.nr + pl.ns) * 8;
        out->bits1 = pl.bits1;
        out->bits2 = pl.bits2;
        out->bf = pl.bf ? 1 : 0;
        out->bf_seg_bits = pl.bf ? pl.geom.seg_bits : 0;
        out->bf_k = pl.bf ? pl.geom.k : 0;
        out->bf_words = pl.bf_words;
        out->tile = pl.tile;
        out->capacity = pl.cap;
        out->workspace_bytes = ws;
        out->bf_stage = pl.bf ? pl.bf_stage : 0;
        return BJ_OK;
    });
}

}  // extern "C"

namespace bj {

// Probe phase timers: cleared before a join that collects them, turned into
// milliseconds after it (each phase's share of all workgroups' ticks times the
// probe kernel's time; chunks of S add to the same timers and times).
static int probe_timers_begin(bj_ctx* c, const Plan& pl)
{
    if (!pl.probe_breakdown) return BJ_OK;
    const int n = probe_grid(pl.tile, c->num_cu) * PROBE_TIMER_SLOTS;
    int rc = ensure(c->ptimers, (uint64_t)n * sizeof(uint64_t));
    if (rc) return rc;
    BJ_HIP(launch_zero_u64(P<uint64_t>(c->ptimers), n, c->stream));
    return BJ_OK;
}

static int probe_timers_end(bj_ctx* c, const Plan& pl, bj_result* out)
{
    if (!pl.probe_breakdown) return BJ_OK;
    const int g = probe_grid(pl.tile, c->num_cu);
    std::vector<uint64_t> t((size_t)g * PROBE_TIMER_SLOTS);
    BJ_HIP(hipMemcpy(t.data(), c->ptimers.p, t.size() * sizeof(uint64_t), hipMemcpyDeviceToHost));
    double sum[PROBE_TIMER_SLOTS] = {}, all = 0.0;
    for (int w = 0; w < g; w++)
        for (int p = 0; p < PROBE_TIMER_SLOTS; p++) sum[p] += (double)t[(size_t)w * PROBE_TIMER_SLOTS + p];
    for (double v : sum) all += v;
    for (int p = 0; p < PROBE_TIMER_SLOTS; p++)
        out->probe_phase_ms[p] = all > 0.0 ? out->kernel_ms[BJ_K_PROBE] * sum[p] / all : 0.0;
    return BJ_OK;
}
static_assert(PROBE_TIMER_SLOTS == BJ_NUM_PROBE_PHASES, "probe phase slots");

// Fills the parts of a join's result that come from its plan and counters
// (h: CTR_N counters; s_prime / heavy: |S'| and the vlsplit work items).
static void fill_result(bj_ctx* c, const Plan& pl, const uint64_t* h, uint64_t s_prime, uint64_t heavy,
                        bj_result* out)
{
    out->match_count = h[CTR_OUT];
    out->capacity = pl.cap;
    out->rows_written = std::min(h[CTR_OUT], pl.cap);
    out->truncated = h[CTR_OUT] > pl.cap ? 1 : 0;
    out->s_survivors = s_prime;
    out->key = P<uint32_t>(c->out_key);
    out->r_val = P<uint32_t>(c->out_rval);
    out->s_val = P<uint32_t>(c->out_sval);
    out->bits1 = pl.bits1;
    out->bits2 = pl.bits2;
    out->bf_seg_bits = pl.bf ? pl.geom.seg_bits : 0;
    out->bf_k = pl.bf ? pl.geom.k : 0;
    out->bf_words = pl.bf_words;
    out->tile = pl.tile;
    out->heavy_items = heavy;
    out->bf_stage = pl.bf ? pl.bf_stage : 0;
    out->pass1_bits = pl.p1;
    out->s_pass1_bits = pl.p1_s;
}

static void fill_phases(bj_result* out, double h2d_ms, double total_ms)
{
    const double* k = out->kernel_ms;
    out->phase_ms[BJ_PH_H2D] = h2d_ms;
    out->phase_ms[BJ_PH_R_PART] = k[BJ_K_HIST_R] + k[BJ_K_SCAN_R] + k[BJ_K_SCATTER1_R] + k[BJ_K_SCATTER2_R];
    out->phase_ms[BJ_PH_BF_BUILD] = k[BJ_K_BF_BUILD];
    out->phase_ms[BJ_PH_S_FILTER] = k[BJ_K_BF_FILTER];
    out->phase_ms[BJ_PH_S_PART] = k[BJ_K_HIST_S] + k[BJ_K_SCAN_S] + k[BJ_K_SCATTER1_S] + k[BJ_K_SCATTER2_S];
    out->phase_ms[BJ_PH_PROBE] = k[BJ_K_PLAN] + k[BJ_K_PROBE];
    out->phase_ms[BJ_PH_TOTAL] = total_ms;
}

// The plan of one chunk of S: the full join's radix bits, passes, tile and
// filter (S is partitioned exactly as R), sizes from the chunk, the full
// join's row capacity.
static int chunk_plan(const bj_params& p, const Plan& full, uint64_t chunk, Plan& cp)
{
    bj_params q = p;
    q.s_size = (int64_t)chunk;
    q.radix_bits = full.bits;
    q.partition_size = 1 << full.bits1;
    q.second_partition_size = 1 << full.bits2;
    q.pass1_bits = full.p1;
    q.s_pass1_bits = full.p1_s;
    q.tile = full.tile;
    if (full.bf) q.bf_stage = full.bf_stage;
    int rc = make_plan(q, cp);
    if (rc) return rc;
    if (cp.bits != full.bits || cp.p1 != full.p1 || cp.p1_s != full.p1_s ||
        (full.bf && (cp.geom.seg_bits != full.geom.seg_bits || cp.geom.wps != full.geom.wps ||
                     cp.geom.k != full.geom.k || cp.bf_stage != full.bf_stage)))
        return fail(BJ_ERR_INTERNAL, "chunk plan differs from the join's");
    cp.cap = full.cap;
    return BJ_OK;
}

// Device bytes a context holds.
static uint64_t held_bytes(bj_ctx* c)
{
    uint64_t t = 0;
    const DevBuf* bufs[] = {&c->in_rk,  &c->in_rv,  &c->in_sk,   &c->in_sv,   &c->r_key,  &c->r_val,
                            &c->s_key,  &c->s_val,  &c->tmp_key, &c->tmp_val, &c->sp_key, &c->sp_val,
                            &c->filter, &c->part,   &c->totals,  &c->l1h,     &c->l1off,  &c->colsum,
                            &c->r_start, &c->s_start, &c->s_fill, &c->partials, &c->items, &c->counters,
                            &c->out_key, &c->out_rval, &c->out_sval, &c->dtmp, &c->plan_cnt, &c->item_start,
                            &c->ptimers, &c->pool_key, &c->pool_val, &c->pool_cnt, &c->pool_off, &c->fin_key,
                            &c->fin_val, &c->fin_start};
    for (const DevBuf* b : bufs) t += b->bytes;
    return t;
}

// s_chunks = 0: the fewest chunks (powers of two up to 1024) whose workspace,
// with the staged host inputs, fits 90% of the device memory this context can
// use (free memory plus what it already holds).
static int auto_s_chunks(bj_ctx* c, const bj_params& p, const Plan& full, int& chunks)
{
    size_t fr = 0, tot = 0;
    BJ_HIP(hipMemGetInfo(&fr, &tot));
    const uint64_t budget = (uint64_t)((double)(fr + held_bytes(c)) * 0.9);
    for (int k = 1; k <= 1024; k *= 2) {
        Plan cp;
        const uint64_t chunk = std::max<uint64_t>((full.ns + k - 1) / k, 1);
        int rc = BJ_OK;
        if (k == 1) cp = full;
        else rc = chunk_plan(p, full, chunk, cp);
        if (rc) return rc;
        uint64_t ws = 0;
        ensure_workspace(nullptr, cp, &ws);
        if (!p.inputs_on_device) ws += full.nr * 8 + (k == 1 ? full.ns * 8 : 2 * chunk * 8);
        if (ws <= budget) {
            chunks = k;
            return BJ_OK;
        }
    }
    return fail(BJ_ERR_NOMEM, "no split of S into up to 1024 chunks fits the device memory");
}

// S in chunks (the reference's s_chunks loop, src/hash_kernel.cu:646-650,
// 910-931): R is partitioned and its filter built once; chunk i of S is
// filtered and partitioned with R's radix bits.  Host chunks are copied on a
// second stream into two device buffers, chunk i + 1 while chunk i is joined.
// Pooled (chunk_probe 0, when a pool the size of S fits): every chunk's S' is
// appended to a pool, and after the last chunk the pieces are merged into one
// S' layout and probed once, so R streams past the tables once.  Otherwise
// every chunk is probed on its own, its rows appended through the one counter.
static int join_s_chunks(bj_ctx* c, const bj_params* p, const Plan& full, int chunks, const uint32_t* r_key,
                         const uint32_t* r_val, const uint32_t* s_key, const uint32_t* s_val, bj_result* out)
{
    const uint64_t ns = full.ns;
    const uint64_t chunk = (ns + (uint64_t)chunks - 1) / (uint64_t)chunks;
    const int nch = (int)((ns + chunk - 1) / chunk);  // chunks holding tuples
    Plan cp;
    int rc = chunk_plan(*p, full, chunk, cp);
    if (rc) return rc;
    if ((rc = ensure_workspace(c, cp))) return rc;
    hipStream_t st = c->stream;
    const bool host = !p->inputs_on_device;
    const bool timing = p->collect_phase_times != 0 || cp.probe_breakdown;
    if (host) {
        if (!c->copy_stream) BJ_HIP(hipStreamCreateWithFlags(&c->copy_stream, hipStreamNonBlocking));
        for (int b = 0; b < 2; b++) {
            if (!c->ev_copied[b]) BJ_HIP(hipEventCreateWithFlags(&c->ev_copied[b], hipEventDisableTiming));
            if (!c->ev_used[b]) BJ_HIP(hipEventCreateWithFlags(&c->ev_used[b], hipEventDisableTiming));
        }
        if (!c->ev_r_copied) BJ_HIP(hipEventCreateWithFlags(&c->ev_r_copied, hipEventDisableTiming));
        if ((rc = ensure(c->in_rk, full.nr * 4)) || (rc = ensure(c->in_rv, full.nr * 4)) ||
            (rc = ensure(c->in_sk, 2 * chunk * 4)) || (rc = ensure(c->in_sv, 2 * chunk * 4)))
            return rc;
    }
    const int nbins = 1 << cp.bits;
    // the pool holds up to all of S, the merged S' as much again (its size is
    // known after the last chunk), the piece table a row per chunk
    bool pooled = false;
    if (p->chunk_probe == 0) {
        const uint64_t table = (uint64_t)nch * nbins * 12;
        const uint64_t need = 2 * ns * 8 + table;
        const uint64_t held = c->pool_key.bytes + c->pool_val.bytes + c->pool_cnt.bytes + c->pool_off.bytes +
                              c->fin_key.bytes + c->fin_val.bytes;
        size_t fr = 0, tot = 0;
        if (hipMemGetInfo(&fr, &tot) != hipSuccess) {
            (void)hipGetLastError();
        } else if (need <= (uint64_t)((double)(fr + held) * 0.9)) {
            pooled = true;
            if ((rc = ensure(c->pool_key, ns * 4)) || (rc = ensure(c->pool_val, ns * 4)) ||
                (rc = ensure(c->pool_cnt, (uint64_t)nch * nbins * 4)) ||
                (rc = ensure(c->pool_off, (uint64_t)nch * nbins * 8)) ||
                (rc = ensure(c->fin_start, ((uint64_t)nbins + 1) * 8)))
                return rc;
        }
    } else if (p->chunk_probe != 1) {
        return fail(BJ_ERR_INVALID, "chunk_probe must be 0 or 1");
    }

    tl_reset(c, timing);
    // R's copy goes first on the copy stream, at the link's full rate; the
    // chunks of S follow it there while R is partitioned
    const bool h2d_ev = host && timing;
    if (host) {
        BJ_HIP(hipEventRecord(c->ev_used[0], st));  // the copies start after earlier work on st
        BJ_HIP(hipStreamWaitEvent(c->copy_stream, c->ev_used[0], 0));
        if (h2d_ev) BJ_HIP(hipEventRecord(c->ev[EV_H2D0], c->copy_stream));
        if (full.nr) {
            BJ_HIP(hipMemcpyAsync(c->in_rk.p, r_key, full.nr * 4, hipMemcpyHostToDevice, c->copy_stream));
            BJ_HIP(hipMemcpyAsync(c->in_rv.p, r_val, full.nr * 4, hipMemcpyHostToDevice, c->copy_stream));
        }
        if (h2d_ev) BJ_HIP(hipEventRecord(c->ev[EV_H2D1], c->copy_stream));
        BJ_HIP(hipEventRecord(c->ev_r_copied, c->copy_stream));
        BJ_HIP(hipStreamWaitEvent(st, c->ev_r_copied, 0));
        r_key = P<uint32_t>(c->in_rk);
        r_val = P<uint32_t>(c->in_rv);
    }
    uint64_t* ctr = P<uint64_t>(c->counters);
    BJ_HIP(launch_zero_u64(ctr, CTR_N, st));
    if ((rc = probe_timers_begin(c, cp))) return rc;
    if ((rc = enqueue_r(c, cp, r_key, r_val))) return rc;
    const RSide R{P<uint32_t>(c->r_key), P<uint32_t>(c->r_val), P<uint64_t>(c->r_start), P<uint64_t>(c->filter)};
    uint64_t* item_start = P<uint64_t>(c->item_start);

    // buffer b of the two holds chunks b, b + 2, ...; its next copy waits until
    // the join of the chunk before has enqueued everything that reads it
    auto copy_chunk = [&](int i) -> int {
        const int b = i & 1;
        const uint64_t off = (uint64_t)i * chunk, n = std::min(chunk, ns - off);
        if (i >= 2) BJ_HIP(hipStreamWaitEvent(c->copy_stream, c->ev_used[b], 0));
        BJ_HIP(hipMemcpyAsync(P<uint32_t>(c->in_sk) + b * chunk, s_key + off, n * 4, hipMemcpyHostToDevice,
                              c->copy_stream));
        BJ_HIP(hipMemcpyAsync(P<uint32_t>(c->in_sv) + b * chunk, s_val + off, n * 4, hipMemcpyHostToDevice,
                              c->copy_stream));
        BJ_HIP(hipEventRecord(c->ev_copied[b], c->copy_stream));
        return BJ_OK;
    };
    static const int s_slots[] = {BJ_K_BF_FILTER, BJ_K_HIST_S,  BJ_K_SCAN_S, BJ_K_SCATTER1_S,
                                  BJ_K_SCATTER2_S, BJ_K_PLAN, BJ_K_PROBE};
    struct SlotSpan { int slot, b, e; };
    std::vector<SlotSpan> s_events;  // per chunk and S-side slot: its begin and end events
    auto take_slots = [&]() {
        for (int sl : s_slots) {
            if (c->tl_b[sl] >= 0 && c->tl_e[sl] >= 0) s_events.push_back({sl, c->tl_b[sl], c->tl_e[sl]});
            c->tl_b[sl] = c->tl_e[sl] = -1;
        }
    };
    if (host && (rc = copy_chunk(0))) return rc;
    for (int i = 0; i < nch; i++) {
        const int b = i & 1;
        const uint64_t off = (uint64_t)i * chunk, n = std::min(chunk, ns - off);
        const uint32_t* ck = s_key + off;
        const uint32_t* cv = s_val + off;
        if (host) {
            BJ_HIP(hipStreamWaitEvent(st, c->ev_copied[b], 0));
            ck = P<uint32_t>(c->in_sk) + b * chunk;
            cv = P<uint32_t>(c->in_sv) + b * chunk;
        }
        // per-chunk counters: S', work items, filter tiles (rows, the pool and
        // error flags carry over)
        BJ_HIP(launch_zero_u64(ctr + CTR_SPRIME, 1, st));
        BJ_HIP(launch_zero_u64(ctr + CTR_WORK, 2, st));
        BJ_HIP(launch_zero_u64(ctr + CTR_XCD, CTR_N - CTR_XCD, st));
        c->tl_open = false;  // the chunk's first slot starts after the wait and the clears
        Plan pi = cp;
        pi.ns = n;
        if (pooled) {
            // the chunk's S' stays in its own region of the pool, [off, off + n),
            // and its pieces (sub-bucket counts and places) go to row i of the
            // piece table (timed with the second S pass)
            const uint32_t* s_cnt = nullptr;
            if ((rc = enqueue_s(c, pi, R, ck, cv, &s_cnt, P<uint32_t>(c->pool_key) + off,
                                P<uint32_t>(c->pool_val) + off, n)))
                return rc;
            if (host) BJ_HIP(hipEventRecord(c->ev_used[b], st));  // the chunk's input is read
            take_slots();
            BJ_HIP(tl_begin(c, BJ_K_SCATTER2_S));
            BJ_HIP(launch_pool_note(P<uint64_t>(c->s_start), s_cnt, nbins, off, P<uint32_t>(c->pool_cnt) + (size_t)i * nbins,
                                    P<uint64_t>(c->pool_off) + (size_t)i * nbins, ctr + CTR_POOL, st));
            BJ_HIP(tl_end(c, BJ_K_SCATTER2_S));
            take_slots();
            if (host && i + 1 < nch && (rc = copy_chunk(i + 1))) return rc;
            continue;
        }
        if ((rc = enqueue_s_probe(c, pi, R, ck, cv, P<uint32_t>(c->out_key), P<uint32_t>(c->out_rval),
                                  P<uint32_t>(c->out_sval), cp.cap)))
            return rc;
        // |S'| and work items of the chunk, summed on the device (a copy to the
        // host here would queue behind the next chunk's copy)
        BJ_HIP(launch_add2_u64(ctr + CTR_SPRIME, item_start + nbins, ctr + CTR_ACC_SPRIME, st));
        if (host) {
            BJ_HIP(hipEventRecord(c->ev_used[b], st));
            if (i + 1 < nch && (rc = copy_chunk(i + 1))) return rc;
        }
        take_slots();
    }
    int last_end = s_events.empty() ? -1 : s_events.back().e;
    if (pooled) {
        // merge the pieces into one S' layout (its size read back first), then
        // one build + probe
        BJ_HIP(tl_begin(c, BJ_K_SCATTER2_S));
        BJ_HIP(launch_pool_start(P<uint32_t>(c->pool_cnt), nch, nbins, P<uint32_t>(c->totals),
                                 P<uint64_t>(c->fin_start), P<uint64_t>(c->partials), st));
        BJ_HIP(hipMemcpyAsync(c->h_counters + CTR_N, P<uint64_t>(c->fin_start) + nbins, sizeof(uint64_t),
                              hipMemcpyDeviceToHost, st));
        BJ_HIP(hipStreamSynchronize(st));
        const uint64_t total = c->h_counters[CTR_N];
        if (total > ns) return fail(BJ_ERR_INTERNAL, "pooled S' larger than S");
        if ((rc = ensure(c->fin_key, total * 4)) || (rc = ensure(c->fin_val, total * 4))) return rc;
        BJ_HIP(launch_pool_merge(P<uint32_t>(c->pool_key), P<uint32_t>(c->pool_val), P<uint32_t>(c->pool_cnt),
                                 P<uint64_t>(c->pool_off), nch, nbins, P<uint64_t>(c->fin_start),
                                 P<uint32_t>(c->fin_key), P<uint32_t>(c->fin_val), c->fin_key.bytes / 4,
                                 ctr + CTR_ERR, st));
        BJ_HIP(tl_end(c, BJ_K_SCATTER2_S));
        take_slots();
        BJ_HIP(launch_zero_u64(ctr + CTR_WORK, 2, st));
        Plan pf = cp;
        pf.ns = total;
        c->tl_open = false;
        if ((rc = enqueue_plan_probe(c, pf, R, P<uint32_t>(c->fin_key), P<uint32_t>(c->fin_val),
                                     P<uint64_t>(c->fin_start), nullptr, P<uint32_t>(c->out_key),
                                     P<uint32_t>(c->out_rval), P<uint32_t>(c->out_sval), cp.cap)))
            return rc;
        last_end = c->tl_e[BJ_K_PROBE];
    }
    BJ_HIP(hipMemcpyAsync(c->h_counters, ctr, CTR_N * sizeof(uint64_t), hipMemcpyDeviceToHost, st));
    BJ_HIP(hipStreamSynchronize(st));

    const uint64_t* h = c->h_counters;
    if (h[CTR_ERR]) return fail(BJ_ERR_INTERNAL, "device check flags " + std::to_string(h[CTR_ERR]));
    if (pooled)
        fill_result(c, cp, h, h[CTR_POOL], h[CTR_ITEMS], out);
    else
        fill_result(c, cp, h, full.bf ? h[CTR_ACC_SPRIME] : ns, h[CTR_ACC_ITEMS], out);
    out->s_chunks = nch;
    if (timing) {
        double* k = out->kernel_ms;
        for (int i = 0; i < BJ_NUM_KERNELS; i++) k[i] = tl_ms(c, i);
        for (const SlotSpan& e : s_events) k[e.slot] += ev_ms(c, e.b, e.e);
        // the span ends at the last probe
        fill_phases(out, h2d_ev ? ev_ms(c, EV_H2D0, EV_H2D1) : 0.0, ev_ms(c, c->tl_b[BJ_K_HIST_R], last_end));
    }
    return probe_timers_end(c, cp, out);
}

// The join of device-resident inputs in one pass, into the row arrays o_*
// (capacity pl.cap): the workspace must already be in place.  h2d_ms < 0: the
// host copies before it were bracketed by EV_H2D0 / EV_H2D1.
int join_device(bj_ctx* c, const Plan& pl, bool timing, double h2d_ms, const uint32_t* r_key,
                const uint32_t* r_val, const uint32_t* s_key, const uint32_t* s_val, uint32_t* o_key,
                uint32_t* o_rval, uint32_t* o_sval, bj_result* out)
{
    int rc;
    hipStream_t st = c->stream;
    uint64_t* ctr = P<uint64_t>(c->counters);
    BJ_HIP(launch_zero_u64(ctr, CTR_N, st));
    if ((rc = probe_timers_begin(c, pl))) return rc;
    // slots without kernels (no filter) keep no events and read as 0 ms
    if ((rc = enqueue_r(c, pl, r_key, r_val))) return rc;
    const RSide R{P<uint32_t>(c->r_key), P<uint32_t>(c->r_val), P<uint64_t>(c->r_start), P<uint64_t>(c->filter)};
    if ((rc = enqueue_s_probe(c, pl, R, s_key, s_val, o_key, o_rval, o_sval, pl.cap))) return rc;

    BJ_HIP(hipMemcpyAsync(c->h_counters, ctr, CTR_N * sizeof(uint64_t), hipMemcpyDeviceToHost, st));
    BJ_HIP(hipStreamSynchronize(st));

    const uint64_t* h = c->h_counters;
    if (h[CTR_ERR]) return fail(BJ_ERR_INTERNAL, "device check flags " + std::to_string(h[CTR_ERR]));
    fill_result(c, pl, h, pl.bf ? h[CTR_SPRIME] : pl.ns, h[CTR_ITEMS], out);
    out->key = o_key;
    out->r_val = o_rval;
    out->s_val = o_sval;
    out->s_chunks = 1;
    if (timing) {
        double* k = out->kernel_ms;
        for (int i = 0; i < BJ_NUM_KERNELS; i++) k[i] = tl_ms(c, i);
        fill_phases(out, h2d_ms < 0 ? ev_ms(c, EV_H2D0, EV_H2D1) : h2d_ms,
                    tl_span_ms(c, BJ_K_HIST_R, BJ_K_PROBE));
    }
    return probe_timers_end(c, pl, out);
}

}  // namespace bj

extern "C" {

static int join_once(bj_ctx* c, const bj_params* p, const uint32_t* r_key, const uint32_t* r_val,
                     const uint32_t* s_key, const uint32_t* s_val, bj_result* out)
{
    if (!c || !p || !out) return fail(BJ_ERR_INVALID, "null argument");
    std::memset(out, 0, sizeof(*out));
    Plan pl;
    int rc = make_plan(*p, pl);
    if (rc) return rc;
    if ((pl.nr && (!r_key || !r_val)) || (pl.ns && (!s_key || !s_val)))
        return fail(BJ_ERR_INVALID, "null relation pointer");
    if (p->s_chunks < 0) return fail(BJ_ERR_INVALID, "s_chunks must be >= 0");
    BJ_HIP(hipSetDevice(c->device));
    int chunks = p->s_chunks;
    if (chunks == 0 && (rc = auto_s_chunks(c, *p, pl, chunks))) return rc;
    if (chunks > 1 && pl.ns > 1) return join_s_chunks(c, p, pl, chunks, r_key, r_val, s_key, s_val, out);
    if ((rc = ensure_workspace(c, pl))) return rc;
    hipStream_t st = c->stream;

    // Kernel-slot events only when the caller collects phase times: each event
    // between two kernels costs the device several microseconds.
    const bool timing = p->collect_phase_times != 0 || pl.probe_breakdown;
    tl_reset(c, timing);
    // host inputs: stage them on the device first (timed as BJ_PH_H2D)
    const bool h2d_ev = !p->inputs_on_device && timing;
    if (h2d_ev) BJ_HIP(hipEventRecord(c->ev[EV_H2D0], st));
    if (!p->inputs_on_device) {
        if ((rc = ensure(c->in_rk, pl.nr * 4)) || (rc = ensure(c->in_rv, pl.nr * 4)) ||
            (rc = ensure(c->in_sk, pl.ns * 4)) || (rc = ensure(c->in_sv, pl.ns * 4)))
            return rc;
        if (pl.nr) {
            BJ_HIP(hipMemcpyAsync(c->in_rk.p, r_key, pl.nr * 4, hipMemcpyHostToDevice, st));
            BJ_HIP(hipMemcpyAsync(c->in_rv.p, r_val, pl.nr * 4, hipMemcpyHostToDevice, st));
        }
        if (pl.ns) {
            BJ_HIP(hipMemcpyAsync(c->in_sk.p, s_key, pl.ns * 4, hipMemcpyHostToDevice, st));
            BJ_HIP(hipMemcpyAsync(c->in_sv.p, s_val, pl.ns * 4, hipMemcpyHostToDevice, st));
        }
        r_key = P<uint32_t>(c->in_rk);
        r_val = P<uint32_t>(c->in_rv);
        s_key = P<uint32_t>(c->in_sk);
        s_val = P<uint32_t>(c->in_sv);
    }
    if (h2d_ev) BJ_HIP(hipEventRecord(c->ev[EV_H2D1], st));
    const double h2d_ms = h2d_ev ? -1.0 : 0.0;
    return join_device(c, pl, timing, h2d_ms, r_key, r_val, s_key, s_val, P<uint32_t>(c->out_key),
                       P<uint32_t>(c->out_rval), P<uint32_t>(c->out_sval), out);
}

int bj_join(bj_ctx* c, const bj_params* p, const uint32_t* r_key, const uint32_t* r_val,
            const uint32_t* s_key, const uint32_t* s_val, bj_result* out)
{
    return guarded([&]() -> int {
        const int rc = join_once(c, p, r_key, r_val, s_key, s_val, out);
        if (rc != BJ_OK && c && c->stream) {
            // an error return leaves no copy of the caller's buffers and no kernel
            // on them in flight: the caller may free them once this returns
            if (c->copy_stream) (void)hipStreamSynchronize(c->copy_stream);
            (void)hipStreamSynchronize(c->stream);
        }
        return rc;
    });
}

int bj_partition(bj_ctx* c, const bj_params* p, const uint32_t* key, const uint32_t* val, int64_t n,
                 uint32_t* out_key, uint32_t* out_val, uint64_t* out_start)
{
    return guarded([&]() -> int {
        if (!c || !p || !out_start || (n > 0 && (!key || !val || !out_key || !out_val)))
            return fail(BJ_ERR_INVALID, "null argument");
        bj_params q = *p;
        q.r_size = n;
        q.s_size = n;
        q.use_bf = 0;
        Plan pl;
        int rc = make_plan(q, pl);
        if (rc) return rc;
        BJ_HIP(hipSetDevice(c->device));
        if ((rc = ensure_workspace(c, pl))) return rc;
        hipStream_t st = c->stream;
        uint64_t* ctr = P<uint64_t>(c->counters);
        tl_reset(c, false);
        BJ_HIP(launch_zero_u64(ctr, CTR_N, st));
        if ((rc = run_partition(c, pl, pl.p1, key, val, pl.nr, nullptr, pl.G_r, pl.solo_r, out_key, out_val, pl.nr, out_start,
                                BJ_K_HIST_R)))
            return rc;
        BJ_HIP(hipMemcpyAsync(c->h_counters, ctr, CTR_N * sizeof(uint64_t), hipMemcpyDeviceToHost, st));
        BJ_HIP(hipStreamSynchronize(st));
        if (c->h_counters[CTR_ERR]) return fail(BJ_ERR_INTERNAL, "device check flags " + std::to_string(c->h_counters[CTR_ERR]));
        return BJ_OK;
    });
}

int bj_copy_rows_to_host(bj_ctx* c, const bj_result* r, uint32_t* key, uint32_t* r_val, uint32_t* s_val,
                         uint64_t n)
{
    return guarded([&]() -> int {
        if (!c || !r) return fail(BJ_ERR_INVALID, "null argument");
        if (n > r->rows_written) return fail(BJ_ERR_INVALID, "n exceeds rows_written");
        if (n == 0) return BJ_OK;
        BJ_HIP(hipSetDevice(c->device));
        if (key) BJ_HIP(hipMemcpyAsync(key, r->key, n * 4, hipMemcpyDeviceToHost, c->stream));
        if (r_val) BJ_HIP(hipMemcpyAsync(r_val, r->r_val, n * 4, hipMemcpyDeviceToHost, c->stream));
        if (s_val) BJ_HIP(hipMemcpyAsync(s_val, r->s_val, n * 4, hipMemcpyDeviceToHost, c->stream));
        BJ_HIP(hipStreamSynchronize(c->stream));
        return BJ_OK;
    });
}

int bj_gen_build(bj_ctx* c, uint64_t seed, uint64_t n_r, uint32_t* d_key, uint32_t* d_val)
{
    if (!c || (n_r && !d_key)) return fail(BJ_ERR_INVALID, "null argument");
    if (n_r > MAX_TUPLES) return fail(BJ_ERR_SIZE, "n_r > 2^31-1");
    BJ_HIP(hipSetDevice(c->device));
    BJ_HIP(launch_gen_r(d_key, d_val, n_r, perm_keys(seed), c->stream));
    BJ_HIP(hipStreamSynchronize(c->stream));
    return BJ_OK;
}

int bj_gen_probe(bj_ctx* c, uint64_t seed_r, uint64_t seed_s, uint64_t n_r, uint64_t n_s, uint64_t m_s,
                 double zipf_s, uint32_t* d_key, uint32_t* d_val)
{
    if (!c || (n_s && !d_key)) return fail(BJ_ERR_INVALID, "null argument");
    if (n_s > (1ull << 31) || n_r > MAX_TUPLES) return fail(BJ_ERR_SIZE, "size limit");
    if (m_s > n_s) return fail(BJ_ERR_INVALID, "m_s > n_s");
    if (m_s > 0 && n_r == 0) return fail(BJ_ERR_INVALID, "matches requested with an empty R");
    if (zipf_s < 0.0) return fail(BJ_ERR_INVALID, "zipf_s < 0");
    BJ_HIP(hipSetDevice(c->device));
    int L = std::max(1, ceil_log2(n_s));
    const uint32_t useed = (uint32_t)splitmix64(seed_s ^ 0x3333333333333333ULL);
    const uint32_t zseed = (uint32_t)splitmix64(seed_s ^ 0x4444444444444444ULL);
    BJ_HIP(launch_gen_s(d_key, d_val, n_s, n_r, m_s, L, perm_keys(seed_r),
                        perm_keys(seed_s ^ 0x1111111111111111ULL),
                        perm_keys(seed_s ^ 0x2222222222222222ULL), useed, zipf_s, zseed,
                        P<double>(c->dtmp), c->stream));
    BJ_HIP(hipStreamSynchronize(c->stream));
    return BJ_OK;
}

}  // extern "C"

// bj_batch.hip -- gfx950 kernels of the batched join, and their launchers.
//
// One launch per phase covers every join of the batch.  Workgroup t of a
// launch is a task (a chunk of one relation, a column block of one join's
// histograms, a filter segment) and finds its join by a binary search over a
// per-join prefix of task counts, so joins of any mix of sizes share the chip
// without one launch per join.  The tile bodies (histogram, write-combining
// scatter, filter build/test, LDS build+probe) are the single-join kernels'
// own (bj_tiles.h).
//
// Reference counterparts: the same as the single-join kernels (bj_kernels.hip);
// the reference has no batched form.
#include "bj_batch.h"
#include "bj_tiles.h"

#include <algorithm>

namespace bj {

// Largest b in [0, n) with px[b] <= t; px is nondecreasing with px[n] = total,
// so for t < total this is the join whose tasks [px[b], px[b+1]) hold t (joins
// without tasks have px[b] == px[b+1] and are skipped).
__device__ __forceinline__ int find_join(const uint32_t* __restrict__ px, int n, uint32_t t)
{
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (px[mid] <= t) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

__device__ __forceinline__ const uint32_t* px_of(const BatchDev& B, int which)
{
    return B.px + (size_t)which * (B.n + 1);
}

// Radix bits and first-pass bits of one side (side 2: the filter pass of S).
__device__ __forceinline__ int side_bits(const BJoin& J, int side) { return side == 2 ? J.fb : J.bits; }
__device__ __forceinline__ int side_b1(const BJoin& J, int side) { return side == 2 ? J.fb : J.b1; }

// Source and current length of one side of join b.
__device__ __forceinline__ void side_src(const BJoin& J, int b, int side, const uint32_t* sp_key,
                                         const uint32_t* sp_val, const uint64_t* jc, const uint32_t*& key,
                                         const uint32_t*& val, uint64_t& n)
{
    if (side == 0) {
        key = J.rk;
        val = J.rv;
        n = J.nr;
    } else if (side == 2) {
        key = J.sk;
        val = J.sv;
        n = J.bf2 ? J.ns : 0u;
    } else if (J.bf) {
        key = sp_key + J.sp_base;
        val = sp_val ? sp_val + J.sp_base : nullptr;
        n = min(jc[(size_t)b * JC_N + JC_SPRIME], (uint64_t)J.ns);
    } else {
        key = J.sk;
        val = J.sv;
        n = J.ns;
    }
}

__device__ __forceinline__ uint32_t side_chunks(const BJoin& J, int side)
{
    const uint64_t n = side ? J.ns : J.nr;
    const uint64_t ch = side ? J.s_ch : J.r_ch;
    return (uint32_t)((n + ch - 1) / ch);
}

// Tuples [lo, hi) of chunk c of a side of current length n.  The chunks of a
// Bloom join's S' split its device-side length evenly (S' is known only on the
// device and is usually a small part of S: chunks of the planned length would
// leave most tasks empty and the first few with all of it); other sides take
// chunks of the planned length (multiples of the scatter tile).
__device__ __forceinline__ void chunk_range(const BJoin& J, int side, uint32_t c, uint64_t n, uint64_t& lo,
                                            uint64_t& hi)
{
    if (side == 1 && J.bf) {
        const uint64_t G = side_chunks(J, 1);
        lo = (uint64_t)c * n / G;
        hi = (uint64_t)(c + 1) * n / G;
    } else {
        const uint64_t ch = side ? J.s_ch : J.r_ch;
        lo = min((uint64_t)c * ch, n);
        hi = min(lo + ch, n);
    }
}

// ============================================================== histogram
template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void kb_hist(BatchDev B, int side, const uint32_t* __restrict__ sp_key,
                                                 const uint64_t* __restrict__ jc, uint32_t* __restrict__ part)
{
    extern __shared__ uint32_t h[];
    const uint32_t* px = px_of(B, side ? PX_STASK : PX_RTASK);
    const uint32_t t = blockIdx.x;
    const int b = find_join(px, B.n, t);
    const BJoin J = B.J[b];
    const uint32_t c = t - px[b];
    const int bits = side_bits(J, side);
    const int nbins = 1 << bits;
    for (int i = threadIdx.x; i < nbins; i += BLOCK) h[i] = 0u;
    __syncthreads();
    const uint32_t *key, *val;
    uint64_t n;
    side_src(J, b, side, sp_key, nullptr, jc, key, val, n);
    uint64_t lo, hi;
    chunk_range(J, side, c, n, lo, hi);
    hist_range<BLOCK>(key, lo, hi, 32 - bits, h);
    __syncthreads();
    uint32_t* dst = part + (side ? J.part_s : J.part_r) + (size_t)c * nbins;
    for (int i = threadIdx.x; i < nbins; i += BLOCK) dst[i] = h[i];
}

// ============================================================== cursors
// Per join and column block of 64 bins: cur_rel[c][j] = sum of part[c'][j] over
// chunks c' < c (the chunk's first position inside bin j), tot[sb_base + j] =
// the bin's size.  16 row groups per column are summed, scanned in LDS, then
// rescanned row by row; loads and stores are coalesced across columns.
__global__ __launch_bounds__(1024) void kb_colscan(BatchDev B, int side, const uint32_t* __restrict__ part,
                                                   uint32_t* __restrict__ cur_rel, uint32_t* __restrict__ tot)
{
    __shared__ uint32_t part_sum[16][64];
    const uint32_t* px = px_of(B, side ? PX_SCOL : PX_RCOL);
    const uint32_t t = blockIdx.x;
    const int b = find_join(px, B.n, t);
    const BJoin J = B.J[b];
    const int cb = (int)(t - px[b]);
    const int nbins = 1 << side_bits(J, side);
    const int G = (int)side_chunks(J, side);
    const uint64_t px0 = side ? J.part_s : J.part_r;
    const int c = threadIdx.x & 63, r = threadIdx.x >> 6;
    const int j = cb * 64 + c;
    const int rows = (G + 15) / 16;
    const int g0 = min(r * rows, G), g1 = min(g0 + rows, G);
    const uint32_t* src = part + px0;
    uint32_t s = 0;
    if (j < nbins)
        for (int g = g0; g < g1; g++) s += src[(size_t)g * nbins + j];
    part_sum[r][c] = s;
    __syncthreads();
    if (r == 0 && j < nbins) {
        uint32_t acc = 0;
        for (int q = 0; q < 16; q++) {
            const uint32_t v = part_sum[q][c];
            part_sum[q][c] = acc;
            acc += v;
        }
        tot[J.sb_base + j] = acc;
    }
    __syncthreads();
    if (j < nbins) {
        uint32_t acc = part_sum[r][c];
        uint32_t* dst = cur_rel + px0;
        for (int g = g0; g < g1; g++) {
            dst[(size_t)g * nbins + j] = acc;
            acc += src[(size_t)g * nbins + j];
        }
    }
}

// One workgroup per join: start[sb_base + j] = region base + exclusive scan of
// the join's bin sizes (at most 2048 bins).
__global__ __launch_bounds__(1024) void kb_binscan(BatchDev B, int side, const uint32_t* __restrict__ tot,
                                                   uint64_t* __restrict__ start)
{
    __shared__ uint64_t wtmp[1024 / WAVE + 1];
    const BJoin J = B.J[blockIdx.x];
    const int nbins = 1 << side_bits(J, side);
    const int per = (nbins + 1023) / 1024;
    const int lo = min((int)threadIdx.x * per, nbins), hi = min(lo + per, nbins);
    const uint32_t* tt = tot + J.sb_base;
    uint64_t s = 0;
    for (int i = lo; i < hi; i++) s += tt[i];
    uint64_t total;
    uint64_t ex = block_exclusive_scan<1024, uint64_t>(s, wtmp, &total) + (side ? J.s_base : J.r_base);
    for (int i = lo; i < hi; i++) {
        start[J.sb_base + i] = ex;
        ex += tt[i];
    }
}

// ============================================================== scatter
// First pass: task (join, chunk) scatters its chunk by the join's first-pass
// bits b1.  The cursor of first-level bucket B is the start of its first
// sub-bucket plus everything earlier chunks put into its sub-buckets,
// start[B << b2] + sum_j cur_rel[c][(B << b2) + j].  One-pass joins (b1 == bits)
// write the final arena, two-pass joins the scratch arena at the same
// positions.
// Tasks come from the prefix of the launch's class (cls 0: first passes of fewer
// than 2^WIDE_B1 buckets, 1: the others).  ITEMS 4: two workgroups per CU,
// 4096-tuple tiles; ITEMS 8: one per CU, 8192-tuple tiles and whole-line carries
// up to 256 buckets (the wide class, and large second passes).
__device__ __forceinline__ int kb_carry_log2(int nbk, bool big)
{
    if (big && nbk <= 256) return 5;
    return nbk <= 512 ? 4 : (nbk <= 1024 ? 3 : 2);  // scatter_carry
}

template <int BLOCK, int ITEMS>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(ITEMS == 4 ? 8 : 4, 8))) void kb_scatter(BatchDev B, int side, int cls, const uint32_t* __restrict__ sp_key,
                                                    const uint32_t* __restrict__ sp_val,
                                                    const uint64_t* __restrict__ jc,
                                                    const uint32_t* __restrict__ cur_rel,
                                                    const uint64_t* __restrict__ start, uint32_t* __restrict__ okey,
                                                    uint32_t* __restrict__ oval, uint64_t out_cap,
                                                    uint32_t* __restrict__ tkey, uint32_t* __restrict__ tval,
                                                    uint64_t tmp_cap, uint64_t* __restrict__ err)
{
    extern __shared__ uint64_t lds64[];
    const uint32_t* px = px_of(B, PX_P1 + 2 * side + cls);
    const uint32_t t = blockIdx.x;
    const int b = find_join(px, B.n, t);
    const BJoin J = B.J[b];
    const uint32_t c = t - px[b];
    const int bits = side_bits(J, side), b1 = side_b1(J, side);
    const int b2 = bits - b1;
    const int nbk = 1 << b1;
    const int cs_log2 = kb_carry_log2(nbk, ITEMS == 8);
    const ScatterLds L = scatter_lds<BLOCK, ITEMS>(lds64, nbk, cs_log2);
    const uint32_t* cr = cur_rel + (side ? J.part_s : J.part_r) + ((size_t)c << bits);
    const uint64_t* st = start + J.sb_base;
    // cursor of bucket j: st[j << b2] + its 2^b2 entries of cr, summed by one wave
    // (coalesced across lanes) or, for buckets narrower than a wave, by aligned
    // lane segments
    {
        const int seg = 1 << b2, wv = threadIdx.x / WAVE;
        const uint32_t lane = lane_id();
        constexpr int NWV = BLOCK / WAVE;
        if (seg >= WAVE) {
            for (int j = wv; j < nbk; j += NWV) {
                uint32_t acc = 0;
                for (int q = (int)lane; q < seg; q += WAVE) acc += cr[((size_t)j << b2) + q];
                acc = wave_sum<uint32_t>(acc);
                if (lane == 0) L.cur[j] = st[(size_t)j << b2] + acc;
            }
        } else {
            const int per = WAVE / seg;  // buckets per wave step
            for (int j0 = wv * per; j0 < nbk; j0 += NWV * per) {
                const size_t idx = ((size_t)j0 << b2) + lane;
                uint32_t v = (idx < ((size_t)nbk << b2)) ? cr[idx] : 0u;
                for (int o = 1; o < seg; o <<= 1) v += __shfl_xor(v, o, WAVE);
                const int j = j0 + (int)(lane >> b2);
                if ((lane & (uint32_t)(seg - 1)) == 0 && j < nbk) L.cur[j] = st[(size_t)j << b2] + v;
            }
        }
    }
    const uint32_t *key, *val;
    uint64_t n;
    side_src(J, b, side, sp_key, sp_val, jc, key, val, n);
    uint64_t lo, hi;
    chunk_range(J, side, c, n, lo, hi);
    const bool two = b2 > 0;
    const bool bad = scatter_range<BLOCK, ITEMS>(L, key + lo, val + lo, (uint32_t)(hi - lo), nbk, 32 - b1,
                                                 (uint32_t)nbk - 1u, cs_log2, two ? tkey : okey, two ? tval : oval,
                                                 two ? tmp_cap : out_cap);
    if (bad) atomicOr((unsigned long long*)err, 16ull);
}

// Second pass of a two-pass join: task (join, first-level bucket B, group w)
// takes what the first pass wrote into B from chunks [g0, g1) of group w
// (w2 groups per bucket) and scatters it by the remaining b2 bits.  Its range
// and its sub-bucket cursors come from cur_rel at chunks g0 and g1.
template <int BLOCK, int ITEMS>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(ITEMS == 4 ? 8 : 4, 8))) void kb_scatter2(BatchDev B, int side, const uint32_t* __restrict__ cur_rel,
                                                     const uint64_t* __restrict__ start,
                                                     const uint32_t* __restrict__ cnt,
                                                     const uint32_t* __restrict__ tkey,
                                                     const uint32_t* __restrict__ tval, uint32_t* __restrict__ okey,
                                                     uint32_t* __restrict__ oval, uint64_t out_cap,
                                                     uint64_t* __restrict__ err)
{
    extern __shared__ uint64_t lds64[];
    __shared__ unsigned long long s_lo, s_hi;
    const uint32_t* px = px_of(B, side ? PX_S2 : PX_R2);
    const uint32_t t = blockIdx.x;
    const int b = find_join(px, B.n, t);
    const BJoin J = B.J[b];
    const uint32_t local = t - px[b];
    const uint32_t W2 = side ? J.w2_s : J.w2_r;
    const int bb = (int)(local / W2), w = (int)(local % W2);
    const int b2 = J.bits - J.b1;
    const int nbk = 1 << b2;
    const int G = (int)side_chunks(J, side);
    const int g0 = (int)(((int64_t)w * G) / W2), g1 = (int)(((int64_t)(w + 1) * G) / W2);
    const int cs_log2 = kb_carry_log2(nbk, ITEMS == 8);
    const ScatterLds L = scatter_lds<BLOCK, ITEMS>(lds64, nbk, cs_log2);
    const size_t f0 = (size_t)bb << b2;  // first sub-bucket of the bucket, within the join
    const uint32_t* base = cur_rel + (side ? J.part_s : J.part_r);
    const uint64_t* st = start + J.sb_base + f0;
    const uint32_t* tt = cnt + J.sb_base + f0;
    if (threadIdx.x == 0) {
        s_lo = 0ull;
        s_hi = 0ull;
    }
    __syncthreads();
    // range: first-level cursors at chunks g0 and g1 (the bucket's end for G)
    for (int j = threadIdx.x; j < nbk; j += BLOCK) {
        const uint32_t a = g0 < G ? base[((size_t)g0 << J.bits) + f0 + j] : tt[j];
        const uint32_t e = g1 < G ? base[((size_t)g1 << J.bits) + f0 + j] : tt[j];
        if (a) atomicAdd(&s_lo, (unsigned long long)a);
        if (e) atomicAdd(&s_hi, (unsigned long long)e);
        L.cur[j] = st[j] + (g0 < G ? a : tt[j]);
    }
    __syncthreads();
    const uint64_t lo = st[0] + s_lo, hi = st[0] + s_hi;
    const bool bad = scatter_range<BLOCK, ITEMS>(L, tkey + lo, tval + lo, (uint32_t)(hi - lo), nbk, 32 - J.bits,
                                                 (uint32_t)nbk - 1u, cs_log2, okey, oval, out_cap);
    if (bad) atomicOr((unsigned long long*)err, 64ull);
}

// ============================================================== Bloom build
// One workgroup per filter segment of a Bloom join; the segment's R keys are
// the partitioned R of its sub-buckets.
template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void kb_bloom_build(BatchDev B, const uint32_t* __restrict__ rkey,
                                                        const uint64_t* __restrict__ r_start,
                                                        const uint32_t* __restrict__ r_cnt,
                                                        uint64_t* __restrict__ filter)
{
    extern __shared__ unsigned long long seg[];
    const uint32_t* px = px_of(B, PX_SEG);
    const uint32_t t = blockIdx.x;
    const int b = find_join(px, B.n, t);
    const BJoin J = B.J[b];
    const uint64_t s = t - px[b];
    const uint32_t wps = J.geom.wps;
    for (uint32_t i = threadIdx.x; i < wps; i += BLOCK) seg[i] = 0ull;
    __syncthreads();
    const int d = J.bits - J.geom.seg_bits;
    const uint32_t sb0 = J.sb_base + (uint32_t)(s << d);
    const uint32_t sb1 = J.sb_base + (uint32_t)((s + 1) << d) - 1u;
    const uint64_t lo = r_start[sb0], hi = r_start[sb1] + r_cnt[sb1];
    bloom_build_range<BLOCK>(rkey, lo, hi, J.geom, s * wps, seg);
    __syncthreads();
    uint64_t* dst = filter + J.f_base + s * wps;
    for (uint32_t i = threadIdx.x; i < wps; i += BLOCK) dst[i] = seg[i];
}

// ============================================================== Bloom filter S -> S'
// Persistent over the filter tiles of all Bloom joins (PF_TILE keys each,
// numbered join by join).  Workgroups are dealt to the XCDs round-robin, and
// XCD x takes the x-th eighth of the tile numbers in order from its own
// counter xctr[x]: the workgroups of one XCD work on neighbouring tiles of one
// or two joins at a time, so those filters stay in that XCD's L2 (dealt out by
// workgroup index instead, every XCD's L2 would hold the filters of every join
// in flight).  Each tile reserves its survivors' places in its join's S'
// staging region.  A bf2 join's tiles read its first pass (scratch arena,
// grouped by the top fb bits), so the keys of a tile probe one or two L2-sized
// slices of the filter.
template <int BLOCK, int ITEMS>
__global__ __launch_bounds__(BLOCK) void kb_prefilter(BatchDev B, uint32_t ntiles,
                                                      const uint64_t* __restrict__ filter,
                                                      const uint32_t* __restrict__ tkey,
                                                      const uint32_t* __restrict__ tval,
                                                      uint32_t* __restrict__ sp_key, uint32_t* __restrict__ sp_val,
                                                      uint64_t* __restrict__ jc, uint64_t* __restrict__ xctr,
                                                      uint64_t* __restrict__ err)
{
    constexpr int T = BLOCK * ITEMS;
    __shared__ uint32_t gcnt[ITEMS * (BLOCK / WAVE)];
    __shared__ uint64_t s_base;
    __shared__ uint32_t s_tile, s_next;
    const uint32_t* px = px_of(B, PX_FTILE);
    const uint32_t xcd = blockIdx.x % NXCD;
    const uint32_t lo = (uint32_t)(((uint64_t)ntiles * xcd) / NXCD);
    const uint32_t cnt = (uint32_t)(((uint64_t)ntiles * (xcd + 1)) / NXCD) - lo;
    unsigned long long* ctr = (unsigned long long*)xctr + xcd;
    if (threadIdx.x == 0) s_tile = (uint32_t)min(atomicAdd(ctr, 1ull), (unsigned long long)cnt);
    __syncthreads();
    uint32_t t = s_tile;
    bool bad = false;
    while (t < cnt) {
        // the next tile's number, read after the tile's first barrier
        if (threadIdx.x == 0) s_next = (uint32_t)min(atomicAdd(ctr, 1ull), (unsigned long long)cnt);
        const uint32_t g = lo + t;
        const int b = find_join(px, B.n, g);
        const BJoin J = B.J[b];
        const uint64_t tile = (uint64_t)(g - px[b]) * T;
        const uint32_t* sk = J.bf2 ? tkey + J.s_base : J.sk;
        const uint32_t* sv = J.bf2 ? tval + J.s_base : J.sv;
        bad |= prefilter_tile<BLOCK, ITEMS>(sk, sv, tile, J.ns, filter + J.f_base, J.geom, sp_key + J.sp_base,
                                            sp_val + J.sp_base, J.ns, jc + (size_t)b * JC_N + JC_SPRIME, gcnt,
                                            &s_base);
        t = s_next;
        __syncthreads();  // s_next is rewritten for the next tile
    }
    if (bad) atomicOr((unsigned long long*)err, 32ull);
}

// ============================================================== probe planning
__global__ void kb_plan_count(BatchDev B, uint32_t nsb, const uint32_t* __restrict__ r_cnt,
                              const uint32_t* __restrict__ s_cnt, uint32_t slice, uint32_t max_slices,
                              uint32_t* __restrict__ n_items)
{
    const uint32_t sb = blockIdx.x * blockDim.x + threadIdx.x;
    if (sb >= nsb) return;
    const int b = (int)B.sbj[sb];
    uint32_t c, sl;
    n_items[sb] = plan_items(r_cnt[sb], s_cnt[sb], B.J[b].tile, slice, max_slices, c, sl);
}

__global__ void kb_plan_fill(BatchDev B, uint32_t nsb, const uint32_t* __restrict__ r_cnt,
                             const uint32_t* __restrict__ s_cnt, uint32_t slice, uint32_t max_slices,
                             const uint64_t* __restrict__ item_start, uint64_t item_cap,
                             HeavyItem* __restrict__ items, uint64_t* __restrict__ err)
{
    const uint32_t sb = blockIdx.x * blockDim.x + threadIdx.x;
    if (sb >= nsb) return;
    const int b = (int)B.sbj[sb];
    uint32_t chunks, slices;
    const uint32_t n = plan_items(r_cnt[sb], s_cnt[sb], B.J[b].tile, slice, max_slices, chunks, slices);
    const uint64_t base = item_start[sb];
    if (base + n > item_cap) {
        atomicOr((unsigned long long*)err, 4ull);
        return;
    }
    for (uint32_t c = 0; c < chunks; c++)
        for (uint32_t s = 0; s < slices; s++) {
            HeavyItem it;
            it.sb = sb;
            it.chunk = c;
            it.slice = s;
            it.nslices = slices;
            items[base + (uint64_t)c * slices + s] = it;
        }
}

// ============================================================== build + probe
// Persistent over the sub-buckets [sb_lo, sb_hi) of all joins of one tile
// class: first their vlsplit items, then one item per regular sub-bucket.
// Rows of join b go to its output region, reserved on its own counter.  Items
// are claimed KB_PROBE_GRAB at a time (thread 0 waits for the queue's atomic
// once per claim), and an item's join is one load (B.sbj), beside its sizes.
// Probe keys per thread per round (as the single join's probe).
constexpr int KB_PROBE_ITEMS = 10;
constexpr int KB_PROBE_GRAB = 2;

template <int BLOCK, int CAP, int PITEMS>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(4, 8))) void kb_probe(
    BatchDev B, uint32_t sb_lo, uint32_t sb_hi, const uint32_t* __restrict__ rk, const uint32_t* __restrict__ rv,
    const uint64_t* __restrict__ r_start, const uint32_t* __restrict__ r_cnt, const uint32_t* __restrict__ sk,
    const uint32_t* __restrict__ sv, const uint64_t* __restrict__ s_start, const uint32_t* __restrict__ s_cnt,
    uint32_t slice, const HeavyItem* __restrict__ items, const uint64_t* __restrict__ item_start,
    uint64_t item_cap, uint64_t* __restrict__ work_ctr, uint64_t* __restrict__ jc,
    uint32_t* __restrict__ o_key, uint32_t* __restrict__ o_rval, uint32_t* __restrict__ o_sval)
{
    __shared__ ProbeLds<BLOCK, CAP> L;
    __shared__ uint64_t s_item, s_next;
    const uint64_t i0 = min(item_start[sb_lo], item_cap);
    const uint64_t n_heavy = min(item_start[sb_hi], item_cap) - i0;
    const uint64_t n_work = n_heavy + (uint64_t)(sb_hi - sb_lo);
    if (threadIdx.x == 0) s_item = atomicAdd((unsigned long long*)work_ctr, (unsigned long long)KB_PROBE_GRAB);
    __syncthreads();
    uint64_t item = s_item;
    int left = KB_PROBE_GRAB;  // items of the current claim, this one included
    __syncthreads();

    while (item < n_work) {
        if (left == 1 && threadIdx.x == 0)
            s_next = atomicAdd((unsigned long long*)work_ctr, (unsigned long long)KB_PROBE_GRAB);
        uint32_t sb, chunk = 0, sl = 0, nsl = 1;
        const bool heavy = item < n_heavy;
        if (heavy) {
            const HeavyItem hi = items[i0 + item];
            sb = hi.sb; chunk = hi.chunk; sl = hi.slice; nsl = hi.nslices;
        } else {
            sb = sb_lo + (uint32_t)(item - n_heavy);
        }
        const uint64_t r0 = r_start[sb], nr = r_cnt[sb], s0 = s_start[sb], ns = s_cnt[sb];
        const bool build_r = build_from_r(nr, ns, (uint32_t)CAP);
        const uint64_t nb = build_r ? nr : ns, np = build_r ? ns : nr;
        const bool skip = nr == 0 || ns == 0 || (!heavy && (nb > (uint64_t)CAP || np > (uint64_t)slice));
        if (!skip) {
            const int b = (int)B.sbj[sb];
            const int bits = B.J[b].bits;
            const uint64_t ob = B.J[b].out_base, oc = B.J[b].out_cap;
            const uint64_t b_lo = (uint64_t)chunk * CAP;
            const int bn = (int)min((uint64_t)CAP, nb - b_lo);
            const uint64_t per = (np + nsl - 1) / nsl;
            const uint64_t p_lo = min((uint64_t)sl * per, np);
            const uint32_t pn = (uint32_t)min(per, np - p_lo);
            const uint64_t brow0 = (build_r ? r0 : s0) + b_lo;
            const uint64_t prow0 = (build_r ? s0 : r0) + p_lo;
            probe_item<BLOCK, CAP, PITEMS>(L, (build_r ? rk : sk) + brow0, (build_r ? rv : sv) + brow0, bn,
                                           (build_r ? sk : rk) + prow0, (build_r ? sv : rv) + prow0, pn, build_r,
                                           bits, jc + (size_t)b * JC_N + JC_OUT, oc, o_key + ob, o_rval + ob,
                                           o_sval + ob);
        }
        __syncthreads();  // the table is rebuilt by the next item; s_next is written
        if (left > 1) {
            item++;
            left--;
        } else {
            item = s_next;
            left = KB_PROBE_GRAB;
        }
        __syncthreads();
    }
}

// ============================================================== launchers
constexpr int BH_BLOCK = 1024;
constexpr int BS_BLOCK = 1024;
constexpr int BS_ITEMS = 4;

static hipError_t set_lds_b(const void* fn, size_t bytes)
{
    hipFuncAttributes attr;
    hipError_t e = hipFuncGetAttributes(&attr, fn);
    if (e != hipSuccess) return e;
    const size_t stat = attr.sharedSizeBytes;
    return hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)(bytes > stat ? bytes - stat : 0));
}

static size_t g_max_lds_b = 64 * 1024;

hipError_t batch_kernels_init(int max_lds_bytes)
{
    hipError_t e;
    if ((e = set_lds_b((const void*)kb_hist<BH_BLOCK>, max_lds_bytes)) != hipSuccess) return e;
    if ((e = set_lds_b((const void*)kb_scatter<BS_BLOCK, BS_ITEMS>, max_lds_bytes)) != hipSuccess) return e;
    if ((e = set_lds_b((const void*)kb_scatter2<BS_BLOCK, BS_ITEMS>, max_lds_bytes)) != hipSuccess) return e;
    if ((e = set_lds_b((const void*)kb_scatter<BS_BLOCK, 2 * BS_ITEMS>, max_lds_bytes)) != hipSuccess) return e;
    if ((e = set_lds_b((const void*)kb_scatter2<BS_BLOCK, 2 * BS_ITEMS>, max_lds_bytes)) != hipSuccess) return e;
    g_max_lds_b = (size_t)max_lds_bytes;
    if ((e = set_lds_b((const void*)kb_bloom_build<1024>, max_lds_bytes)) != hipSuccess) return e;
    return hipSuccess;
}

hipError_t launch_batch_hist(const BatchDev& B, int side, uint32_t ntasks, const uint32_t* sp_key,
                             const uint64_t* jc, uint32_t* part, hipStream_t st)
{
    if (ntasks == 0) return hipSuccess;
    kb_hist<BH_BLOCK><<<ntasks, BH_BLOCK, sizeof(uint32_t) << B.max_bits, st>>>(B, side, sp_key, jc, part);
    return hipGetLastError();
}

hipError_t launch_batch_colscan(const BatchDev& B, int side, uint32_t ncolblk, const uint32_t* part,
                                uint32_t* cur_rel, uint32_t* tot, hipStream_t st)
{
    if (ncolblk == 0) return hipSuccess;
    kb_colscan<<<ncolblk, 1024, 0, st>>>(B, side, part, cur_rel, tot);
    return hipGetLastError();
}

hipError_t launch_batch_binscan(const BatchDev& B, int side, const uint32_t* tot, uint64_t* start,
                                hipStream_t st)
{
    if (B.n == 0) return hipSuccess;
    kb_binscan<<<B.n, 1024, 0, st>>>(B, side, tot, start);
    return hipGetLastError();
}

hipError_t launch_batch_scatter(const BatchDev& B, int side, int cls, uint32_t ntasks, int max_b1,
                                const uint32_t* sp_key, const uint32_t* sp_val, const uint64_t* jc,
                                const uint32_t* cur_rel, const uint64_t* start, uint32_t* okey, uint32_t* oval,
                                uint64_t out_cap, uint32_t* tkey, uint32_t* tval, uint64_t tmp_cap, uint64_t* err,
                                bool big, hipStream_t st)
{
    if (ntasks == 0) return hipSuccess;
    if (side < 0 || side > 2 || cls < 0 || cls > 1) return hipErrorInvalidValue;
    const size_t lds_big = std::max(scatter_lds_bytes_big(1 << max_b1), g_max_lds_b / 2 + 1024);
    if (big && lds_big <= g_max_lds_b) {
        kb_scatter<BS_BLOCK, 2 * BS_ITEMS><<<ntasks, BS_BLOCK, lds_big, st>>>(
            B, side, cls, sp_key, sp_val, jc, cur_rel, start, okey, oval, out_cap, tkey, tval, tmp_cap, err);
        return hipGetLastError();
    }
    const size_t lds = scatter_lds_bytes(1 << max_b1);
    kb_scatter<BS_BLOCK, BS_ITEMS><<<ntasks, BS_BLOCK, lds, st>>>(B, side, cls, sp_key, sp_val, jc, cur_rel, start,
                                                                  okey, oval, out_cap, tkey, tval, tmp_cap, err);
    return hipGetLastError();
}

hipError_t launch_batch_scatter2(const BatchDev& B, int side, uint32_t ntasks, const uint32_t* cur_rel,
                                 const uint64_t* start, const uint32_t* cnt, const uint32_t* tkey,
                                 const uint32_t* tval, uint32_t* okey, uint32_t* oval, uint64_t out_cap,
                                 uint64_t* err, bool big, hipStream_t st)
{
    if (ntasks == 0) return hipSuccess;
    const size_t lds_big = std::max(scatter_lds_bytes_big(1 << std::max(1, B.max_b2)), g_max_lds_b / 2 + 1024);
    if (big && lds_big <= g_max_lds_b) {
        kb_scatter2<BS_BLOCK, 2 * BS_ITEMS><<<ntasks, BS_BLOCK, lds_big, st>>>(B, side, cur_rel, start, cnt, tkey,
                                                                             tval, okey, oval, out_cap, err);
        return hipGetLastError();
    }
    const size_t lds = scatter_lds_bytes(1 << std::max(1, B.max_b2));
    kb_scatter2<BS_BLOCK, BS_ITEMS><<<ntasks, BS_BLOCK, lds, st>>>(B, side, cur_rel, start, cnt, tkey, tval,
                                                                   okey, oval, out_cap, err);
    return hipGetLastError();
}

hipError_t launch_batch_bloom_build(const BatchDev& B, uint32_t nseg, uint32_t max_wps, const uint32_t* rkey,
                                    const uint64_t* r_start, const uint32_t* r_cnt, uint64_t* filter,
                                    hipStream_t st)
{
    if (nseg == 0) return hipSuccess;
    kb_bloom_build<1024><<<nseg, 1024, (size_t)max_wps * 8, st>>>(B, rkey, r_start, r_cnt, filter);
    return hipGetLastError();
}

hipError_t launch_batch_prefilter(const BatchDev& B, uint32_t ntiles, const uint64_t* filter,
                                  const uint32_t* tkey, const uint32_t* tval, uint32_t* sp_key,
                                  uint32_t* sp_val, uint64_t* jc, uint64_t* xctr, uint64_t* err, int num_cu,
                                  hipStream_t st)
{
    if (ntiles == 0) return hipSuccess;
    // four workgroups per CU, a multiple of NXCD
    const uint32_t grid = (uint32_t)std::max(1, num_cu / NXCD) * 4u * NXCD;
    kb_prefilter<PF_BLOCK, PF_ITEMS><<<grid, PF_BLOCK, 0, st>>>(B, ntiles, filter, tkey, tval, sp_key, sp_val, jc,
                                                                xctr, err);
    return hipGetLastError();
}

hipError_t launch_batch_plan(const BatchDev& B, uint32_t nsb, const uint64_t* r_start, const uint32_t* r_cnt,
                             const uint64_t* s_start, const uint32_t* s_cnt, uint32_t slice,
                             uint32_t max_slices, uint32_t* counts, uint64_t* item_start, uint64_t* partials,
                             HeavyItem* items, uint64_t item_cap, uint64_t* err, hipStream_t st)
{
    (void)r_start;
    (void)s_start;
    if (nsb == 0) return hipSuccess;
    const int grid = (int)((nsb + 255) / 256);
    kb_plan_count<<<grid, 256, 0, st>>>(B, nsb, r_cnt, s_cnt, slice, max_slices, counts);
    hipError_t e = launch_scan(counts, (uint64_t)nsb, item_start, partials, st);
    if (e != hipSuccess) return e;
    kb_plan_fill<<<grid, 256, 0, st>>>(B, nsb, r_cnt, s_cnt, slice, max_slices, item_start, item_cap, items, err);
    return hipGetLastError();
}

hipError_t launch_batch_probe(const BatchDev& B, int tile, uint32_t sb_lo, uint32_t sb_hi,
                              const uint32_t* rk, const uint32_t* rv, const uint64_t* r_start,
                              const uint32_t* r_cnt, const uint32_t* sk, const uint32_t* sv,
                              const uint64_t* s_start, const uint32_t* s_cnt, uint32_t slice,
                              const HeavyItem* items, const uint64_t* item_start, uint64_t item_cap,
                              uint64_t* work_ctr, uint64_t* jc, uint32_t* o_key, uint32_t* o_rval,
                              uint32_t* o_sval, int num_cu, hipStream_t st)
{
    if (sb_hi <= sb_lo) return hipSuccess;
    if (tile == 8192) {
        kb_probe<512, 8192, KB_PROBE_ITEMS><<<num_cu * 2, 512, 0, st>>>(B, sb_lo, sb_hi, rk, rv, r_start, r_cnt, sk, sv,
                                                           s_start, s_cnt, slice, items, item_start, item_cap,
                                                           work_ctr, jc, o_key, o_rval, o_sval);
    } else if (tile == 4096) {
        kb_probe<256, 4096, KB_PROBE_ITEMS><<<num_cu * 4, 256, 0, st>>>(B, sb_lo, sb_hi, rk, rv, r_start, r_cnt, sk, sv,
                                                           s_start, s_cnt, slice, items, item_start, item_cap,
                                                           work_ctr, jc, o_key, o_rval, o_sval);
    } else {
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace bj

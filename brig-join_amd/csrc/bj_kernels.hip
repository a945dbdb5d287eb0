// bj_kernels.hip -- gfx950 kernels of the radix hash join, and their launchers.
//
// Pipeline (one GPU; all integer, all exact):
//   R:  k_hist -> k_colsums, k_starts, k_l1_offsets -> k_scatter(pass 1) -> k_scatter(pass 2)
//   BF: k_bloom_build (one workgroup per filter segment, built in LDS)
//   S:  bf_stage 2: k_hist + pass 1 of S, then k_filter_l1 (per first-level
//       bucket, XCD-aware, filter slice L2-resident) places the survivors;
//       bf_stage 1: k_bloom_prefilter (S -> S') then the two-level partition
//   J:  k_plan_count, k_plan_fill (vlsplit work items) -> k_probe (persistent;
//       LDS chained hash table on the build side, stream the other side)
//
// Reference counterparts (Chanho-LEE-SNU/BRIG-Join):
//   k_hist / k_colsums / k_starts    gpu_run_hist_for_both, gpu_make_simple_prefix,
//                                    gpu_run_hist_second_pass
//                                    (src/kernel_functions.cu:22-90, 461-492, 163-207)
//   k_scatter                        gpu_make_sorted_table_for_both, gpu_build_table_second_pass
//                                    (src/kernel_functions.cu:95-133, 210-281)
//   k_bloom_build                    insertKernelPartitioned (src/bloom.cu:698-722)
//   k_bloom_prefilter, k_filter_l1   queryAndCompactKernelPartitioned (src/bloom.cu:725-773)
//   k_probe                          gpu_probe_with_second_pass / gpu_fused_probe_transfer
//                                    local part (src/kernel_functions.cu:294-457, 612-789)
#include "bj_kernels.h"
#include "bj_device.h"
#include "bj_tiles.h"

#include <math.h>

#include <algorithm>

namespace bj {

// ============================================================== utilities
__global__ void k_zero_u64(uint64_t* __restrict__ p, int n)
{
    for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0ull;
}

__global__ void k_add2_u64(const uint64_t* __restrict__ a, const uint64_t* __restrict__ b, uint64_t* __restrict__ acc)
{
    if (threadIdx.x == 0) {
        acc[0] += *a;
        acc[1] += *b;
    }
}

// ============================================================== generator
__global__ void k_gen_r(uint32_t* __restrict__ key, uint32_t* __restrict__ val, uint64_t n,
                        PermKeys K)
{
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x) {
        key[i] = gen_r_key((uint32_t)i, K);
        if (val) val[i] = (uint32_t)i;
    }
}

// Zipf(s) rank in [1, n] by inverting the generalised harmonic number
// H_s(r) = sum_{i<=r} i^-s with an Euler-Maclaurin approximation; small ranks
// are summed exactly.
__device__ double zipf_H(double r, double s)
{
    if (r < 16.0) {
        double acc = 0.0;
        for (int i = 1; i <= (int)r; i++) acc += pow((double)i, -s);
        return acc;
    }
    // exact sum to 15, then the integral tail with end corrections
    double acc = 0.0;
    for (int i = 1; i <= 15; i++) acc += pow(15.0 + 1.0 - i, -s);  // sum 1..15
    const double a = 15.0;
    double tail;
    if (fabs(s - 1.0) < 1e-12) tail = log(r) - log(a);
    else tail = (pow(r, 1.0 - s) - pow(a, 1.0 - s)) / (1.0 - s);
    tail += 0.5 * (pow(r, -s) - pow(a, -s));
    tail -= s / 12.0 * (pow(r, -s - 1.0) - pow(a, -s - 1.0));
    return acc + tail;
}

__device__ uint32_t zipf_rank(double u, uint32_t n, double s, double Hn)
{
    const double t = u * Hn;
    // initial guess from the continuous inverse
    double g;
    if (fabs(s - 1.0) < 1e-12) g = exp(t - 0.5772156649015329);
    else g = pow(fmax(1e-300, (1.0 - s) * (t - 0.5772156649015329) + 1.0), 1.0 / (1.0 - s));
    double r = fmin(fmax(floor(g), 1.0), (double)n);
    // walk to the smallest r with H(r) >= t (a few steps)
    for (int it = 0; it < 64 && r < (double)n && zipf_H(r, s) < t; it++) r += 1.0;
    for (int it = 0; it < 64 && r > 1.0 && zipf_H(r - 1.0, s) >= t; it++) r -= 1.0;
    return (uint32_t)r;
}

__global__ void k_gen_s(uint32_t* __restrict__ key, uint32_t* __restrict__ val, uint64_t n_s,
                        uint64_t n_r, uint64_t m_s, int L, PermKeys KR, PermKeys KP,
                        PermKeys KN, uint32_t useed, double zipf_s, double zipf_Hn,
                        uint32_t zseed)
{
    for (uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; j < n_s;
         j += (uint64_t)gridDim.x * blockDim.x) {
        uint32_t p = permw((uint32_t)j, L, KP);
        while ((uint64_t)p >= n_s) p = permw(p, L, KP);
        uint32_t k;
        if ((uint64_t)p < m_s) {
            uint32_t u;
            if (zipf_s > 0.0) {
                const uint32_t a = lowbias32((uint32_t)j ^ zseed);
                const uint32_t b = lowbias32(a ^ 0x68E31DA4u);
                const double uu = ((double)(((uint64_t)a << 21) ^ (uint64_t)(b >> 11)) + 0.5) *
                                  (1.0 / 9007199254740992.0);
                u = zipf_rank(uu, (uint32_t)n_r, zipf_s, zipf_Hn) - 1u;
            } else {
                u = (uint32_t)(((uint64_t)lowbias32((uint32_t)j ^ useed) * n_r) >> 32);
            }
            k = gen_r_key(u, KR);
        } else {
            k = 0x80000000u | permw((uint32_t)j, 31, KN);
        }
        key[j] = k;
        if (val) val[j] = (uint32_t)j;
    }
}

// The reference datagen's distribution (datagen/gen_dataset.cpp:172-220) on the
// device.  Before its shuffle, R is [0, k) then k + (i - k) mod (2^31 - 1 - k),
// and S is a run of matching keys (0, 1, ...) and S-only keys R + j mod
// (2^31 - 1 - R) per region; the reference shuffles each region with
// Fisher-Yates, here position i of a region takes the unshuffled element at a
// keyed permutation of i (cycle-walked onto the region), so each region holds
// the same key multiset in another random order.
__device__ __forceinline__ uint32_t perm_onto(uint64_t i, uint64_t n, int L, const PermKeys& K)
{
    uint32_t p = permw((uint32_t)i, L, K);
    while ((uint64_t)p >= n) p = permw(p, L, K);
    return p;
}

__global__ void k_gen_refdist_r(uint32_t* __restrict__ key, uint64_t n, uint64_t k, int L, PermKeys K)
{
    const uint64_t r_only = 0x7FFFFFFFull - k;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t t = perm_onto(i, n, L, K);
        key[i] = (uint32_t)(t < k ? t : k + (t - k) % r_only);
    }
}

// One region of S: positions [0, n) of key (already offset), the first m
// unshuffled elements matching keys mc0, mc0 + 1, ..., the rest S-only keys
// n_r + (sj0 + j) mod (2^31 - 1 - n_r).
__global__ void k_gen_refdist_s(uint32_t* __restrict__ key, uint64_t n, uint64_t m, uint64_t mc0,
                                uint64_t sj0, uint64_t n_r, int L, PermKeys K)
{
    const uint64_t s_only = 0x7FFFFFFFull - n_r;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t t = perm_onto(i, n, L, K);
        key[i] = (uint32_t)(t < m ? mc0 + t : n_r + (sj0 + (t - m)) % s_only);
    }
}

__global__ void k_zipf_Hn(double* out, uint32_t n, double s)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) out[0] = zipf_H((double)n, s);
}

// ============================================================== histogram
// Per-workgroup histogram of the full radix (2^bits bins) over a contiguous
// chunk.  Chunking: chunk = ceil(n / G), workgroup g owns [g*chunk, (g+1)*chunk).
// k_scatter pass 1 uses exactly the same chunks, so the per-(bucket, group)
// offsets derived from these histograms place every tuple without atomics.
template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_hist(const uint32_t* __restrict__ key, uint64_t n_host,
                                                const uint64_t* __restrict__ n_dev, int G,
                                                int bits, uint32_t* __restrict__ part,
                                                uint32_t* __restrict__ totals, int bits2,
                                                uint32_t* __restrict__ l1h)
{
    extern __shared__ uint32_t h[];
    const int nbins = 1 << bits;
    for (int i = threadIdx.x; i < nbins; i += BLOCK) h[i] = 0u;
    // the bin totals that k_colsums accumulates start at zero
    if (totals)
        for (int i = blockIdx.x * BLOCK + threadIdx.x; i < nbins; i += G * BLOCK) totals[i] = 0u;
    __syncthreads();
    const uint64_t n = n_dev ? *n_dev : n_host;
    const uint64_t chunk = (n + (uint64_t)G - 1) / (uint64_t)G;
    const uint64_t lo = min((uint64_t)blockIdx.x * chunk, n);
    const uint64_t hi = min(lo + chunk, n);
    hist_range<BLOCK>(key, lo, hi, 32 - bits, h);
    __syncthreads();
    uint32_t* dst = part + (size_t)blockIdx.x * nbins;
    for (int b = threadIdx.x; b < nbins; b += BLOCK) dst[b] = h[b];
    // this chunk's first-level counts (bits2 > 0): thread b1 sums its 2^bits2
    // bins from LDS, starting at a rotated offset so that threads hit different
    // banks at each step
    if (bits2 > 0) {
        const int P1 = nbins >> bits2, seg = 1 << bits2;
        for (int b1 = threadIdx.x; b1 < P1; b1 += BLOCK) {
            const uint32_t* hb = h + ((size_t)b1 << bits2);
            uint32_t sum = 0;
            for (int j = 0; j < seg; j++) sum += hb[(j + b1) & (seg - 1)];
            l1h[(size_t)blockIdx.x * P1 + b1] = sum;
        }
    }
}

// Offsets of a partition, three launches.
// k_colsums: workgroups [0, ncb * W) sum column block cb (1024 bins) of the
// per-chunk histograms over chunk group w (chunks [w*G/W, (w+1)*G/W)) into
// colsum[w][b] and, by atomics, into totals[b] (zeroed by k_hist).  k_hist
// writes each chunk's first-level counts l1h[g][b1] when the first pass has
// fewer bits than the partition.
// Pass 2 derives its cursors from sb_start and colsum itself.
__global__ __launch_bounds__(1024) void k_colsums(const uint32_t* __restrict__ part, int G, int bits, int W,
                                                  uint32_t* __restrict__ colsum, uint32_t* __restrict__ totals)
{
    constexpr int B = 1024;
    const int nbins = 1 << bits;
    const int ncb = (nbins + B - 1) / B;
    {
        const int cb = blockIdx.x % ncb, w = blockIdx.x / ncb;
        const int c = cb * B + threadIdx.x;
        const int g0 = (int)((int64_t)w * G / W), g1 = (int)((int64_t)(w + 1) * G / W);
        if (c < nbins) {
            uint32_t sum = 0;
#pragma unroll 8
            for (int g = g0; g < g1; g++) sum += part[(size_t)g * nbins + c];
            colsum[(size_t)w * nbins + c] = sum;
            if (sum) atomicAdd(&totals[c], sum);
        }
    }
}

// k_starts: one workgroup scans the bin totals into the sub-bucket starts
// (sb_start[nbins] = number of tuples).  nbins is a power of two up to 2^15:
// thread t owns bins [t * per, (t + 1) * per), per = nbins / 1024 (at least 1),
// read as 16-byte vectors and written as pairs of starts from 4 bins on.
__global__ __launch_bounds__(1024) void k_starts(const uint32_t* __restrict__ totals, int nbins,
                                                 uint64_t* __restrict__ sb_start)
{
    constexpr int MAXPER = 32;
    __shared__ uint64_t wtmp[1024 / WAVE + 1];
    const int per = max(1, nbins >> 10);
    const int b0 = (int)threadIdx.x * per;
    uint32_t v[MAXPER];
    if (per >= 4) {
        const uint4* t4 = (const uint4*)(totals + b0);
#pragma unroll
        for (int j = 0; j < MAXPER / 4; j++) {
            const uint4 q = (4 * j < per) ? t4[j] : make_uint4(0u, 0u, 0u, 0u);
            v[4 * j] = q.x;
            v[4 * j + 1] = q.y;
            v[4 * j + 2] = q.z;
            v[4 * j + 3] = q.w;
        }
    } else {
#pragma unroll
        for (int i = 0; i < 4; i++) v[i] = (i < per && b0 + i < nbins) ? totals[b0 + i] : 0u;
#pragma unroll
        for (int i = 4; i < MAXPER; i++) v[i] = 0u;
    }
    uint64_t sum = 0;
#pragma unroll
    for (int i = 0; i < MAXPER; i++) sum += v[i];
    uint64_t tot;
    uint64_t ex = block_exclusive_scan<1024, uint64_t>(sum, wtmp, &tot);
    if (per >= 4) {
        ulonglong2* o2 = (ulonglong2*)(sb_start + b0);
#pragma unroll
        for (int j = 0; j < MAXPER / 2; j++) {
            const uint64_t e0 = ex, e1 = ex + v[2 * j];
            ex = e1 + v[2 * j + 1];
            if (2 * j < per) o2[j] = make_ulonglong2(e0, e1);
        }
    } else {
#pragma unroll
        for (int i = 0; i < 4; i++) {
            if (i < per && b0 + i < nbins) sb_start[b0 + i] = ex;
            ex += v[i];
        }
    }
    if (threadIdx.x == 0) sb_start[nbins] = tot;
}

// First-level cursors of the pass-1 scatter: l1off[g][b1] = start of bucket b1
// (sb_start[b1 << bits2]) + sum_{g' < g} l1h[g'][b1].  A column scan: each
// workgroup owns 16 columns; 64 row ranges per column are summed, scanned by
// one wave per column, then rescanned row by row.  Short row ranges keep the
// serial chains short when there are few columns (128 first-level buckets give
// 8 workgroups); loads and stores are coalesced over a row's 16 columns.
constexpr int L1_COLS = 16, L1_ROWS = 1024 / L1_COLS;
static_assert(L1_ROWS == WAVE && 1024 / WAVE == L1_COLS, "one wave scans one column");

__global__ __launch_bounds__(1024) void k_l1_offsets(const uint32_t* __restrict__ l1h, int G, int P1,
                                                     int bits2, const uint64_t* __restrict__ sb_start,
                                                     uint64_t* __restrict__ l1off)
{
    __shared__ uint64_t part_sum[L1_ROWS][L1_COLS + 1];
    const int c = threadIdx.x % L1_COLS, r = threadIdx.x / L1_COLS;
    const int b1 = blockIdx.x * L1_COLS + c;
    const int rows = (G + L1_ROWS - 1) / L1_ROWS;
    const int g0 = min(r * rows, G), g1 = min(g0 + rows, G);
    uint64_t s = 0;
    if (b1 < P1)
        for (int g = g0; g < g1; g++) s += l1h[(size_t)g * P1 + b1];
    part_sum[r][c] = s;
    __syncthreads();
    {
        // wave w: column w, lane = row range (the branch is uniform per wave)
        const int w = threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
        const int col = blockIdx.x * L1_COLS + w;
        if (col < P1) {
            const uint64_t v = part_sum[lane][w];
            part_sum[lane][w] = sb_start[(size_t)col << bits2] + wave_inclusive_scan<uint64_t>(v) - v;
        }
    }
    __syncthreads();
    if (b1 < P1) {
        uint64_t acc = part_sum[r][c];
        for (int g = g0; g < g1; g++) {
            l1off[(size_t)g * P1 + b1] = acc;
            acc += l1h[(size_t)g * P1 + b1];
        }
    }
}


__global__ void k_zero_u32(uint32_t* __restrict__ p, uint64_t n)
{
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = 0u;
}

// ============================================================== scan (u32 -> u64 exclusive)
// Reduce-then-scan over n entries; out has n + 1 entries, out[n] = total.
constexpr int SCAN_BLOCK = 256;
constexpr int SCAN_ITEMS = 16;
constexpr int SCAN_TILE = SCAN_BLOCK * SCAN_ITEMS;

__global__ __launch_bounds__(SCAN_BLOCK) void k_scan_reduce(const uint32_t* __restrict__ in,
                                                            uint64_t n, uint64_t* __restrict__ partial)
{
    __shared__ uint64_t wtmp[SCAN_BLOCK / WAVE + 1];
    const uint64_t base = (uint64_t)blockIdx.x * SCAN_TILE;
    uint64_t s = 0;
    for (int it = 0; it < SCAN_ITEMS; it++) {
        const uint64_t i = base + (uint64_t)it * SCAN_BLOCK + threadIdx.x;
        if (i < n) s += in[i];
    }
    uint64_t tot;
    block_exclusive_scan<SCAN_BLOCK, uint64_t>(s, wtmp, &tot);
    if (threadIdx.x == 0) partial[blockIdx.x] = tot;
}

// single workgroup: exclusive scan of the W partials in place (W <= 1024*64)
__global__ __launch_bounds__(1024) void k_scan_top(uint64_t* __restrict__ partial, int W)
{
    __shared__ uint64_t wtmp[1024 / WAVE + 1];
    const int per = (W + 1023) / 1024;
    const int lo = min((int)threadIdx.x * per, W);
    const int hi = min(lo + per, W);
    uint64_t s = 0;
    for (int i = lo; i < hi; i++) s += partial[i];
    uint64_t tot;
    uint64_t ex = block_exclusive_scan<1024, uint64_t>(s, wtmp, &tot);
    for (int i = lo; i < hi; i++) {
        const uint64_t c = partial[i];
        partial[i] = ex;
        ex += c;
    }
    if (threadIdx.x == 0) partial[W] = tot;
}

__global__ __launch_bounds__(SCAN_BLOCK) void k_scan_down(const uint32_t* __restrict__ in, uint64_t n,
                                                          const uint64_t* __restrict__ partial,
                                                          uint64_t* __restrict__ out, int W)
{
    __shared__ uint64_t wtmp[SCAN_BLOCK / WAVE + 1];
    // each thread owns SCAN_ITEMS consecutive entries
    const uint64_t base = (uint64_t)blockIdx.x * SCAN_TILE + (uint64_t)threadIdx.x * SCAN_ITEMS;
    uint32_t v[SCAN_ITEMS];
    uint64_t s = 0;
#pragma unroll
    for (int it = 0; it < SCAN_ITEMS; it++) {
        const uint64_t i = base + it;
        v[it] = (i < n) ? in[i] : 0u;
        s += v[it];
    }
    uint64_t tot;
    uint64_t ex = block_exclusive_scan<SCAN_BLOCK, uint64_t>(s, wtmp, &tot) + partial[blockIdx.x];
#pragma unroll
    for (int it = 0; it < SCAN_ITEMS; it++) {
        const uint64_t i = base + it;
        if (i < n) out[i] = ex;
        ex += v[it];
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) out[n] = partial[W];
}

// ============================================================== scatter
// Write-combining scatter of one range (scatter_range, bj_tiles.h).  What a
// workgroup owns:
//   mode 0 (pass 1, p1 bits): workgroup g owns the same chunk as k_hist's group
//       g; bucket = h >> (32 - p1); cur[b] = l1off[g * nbk + b].
//   mode 1 (pass 2, bits - p1 bits): workgroup (b1, w) owns the pass-1 output of
//       bucket b1 written by pass-1 groups [w*G/W2, (w+1)*G/W2); bucket =
//       (h >> (32 - bits)) & (2^(bits-p1) - 1); cur[b] = sb_start[sb] +
//       sum_{v < w} colsum[v][sb] with sb = (b1 << (bits-p1)) + b.
// Writes are bounds-checked against out_cap and flag err on violation.
// Occupancy targets: two 1024-thread workgroups per CU for the 4096-tuple tile,
// one for the 8192-tuple tile (launched one per CU), three 512-thread ones.
constexpr int sc_min_waves(int block, int items) { return block == 1024 ? (items == 4 ? 8 : 4) : 6; }

template <int BLOCK, int ITEMS>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(sc_min_waves(BLOCK, ITEMS), 8))) void k_scatter(
    const uint32_t* __restrict__ ikey, const uint32_t* __restrict__ ival,
    uint32_t* __restrict__ okey, uint32_t* __restrict__ oval, uint64_t out_cap,
    uint64_t n_host, const uint64_t* __restrict__ n_dev, int G, int mode, int bits, int p1, int W2,
    int cs_log2, const uint64_t* __restrict__ l1off, const uint64_t* __restrict__ sb_start,
    const uint32_t* __restrict__ colsum, uint64_t* __restrict__ err)
{
    const int b2 = bits - p1;
    const int nbk = (mode == 0) ? (1 << p1) : (1 << b2);
    const int shift = (mode == 0) ? (32 - p1) : (32 - bits);
    const uint32_t bmask = (uint32_t)nbk - 1u;
    extern __shared__ uint64_t lds64[];
    const ScatterLds L = scatter_lds<BLOCK, ITEMS>(lds64, nbk, cs_log2);

    uint64_t lo, hi;
    if (mode == 0) {
        const uint64_t n = n_dev ? *n_dev : n_host;
        const uint64_t chunk = (n + (uint64_t)G - 1) / (uint64_t)G;
        lo = min((uint64_t)blockIdx.x * chunk, n);
        hi = min(lo + chunk, n);
        for (int b = threadIdx.x; b < nbk; b += BLOCK) L.cur[b] = l1off[(size_t)blockIdx.x * nbk + b];
    } else {
        const int P1 = 1 << p1;
        const int b1 = blockIdx.x / W2, w = blockIdx.x % W2;
        const int g0 = (int)(((int64_t)w * G) / W2), g1 = (int)(((int64_t)(w + 1) * G) / W2);
        lo = (g0 < G) ? l1off[(size_t)g0 * P1 + b1] : sb_start[(uint64_t)(b1 + 1) << b2];
        hi = (g1 < G) ? l1off[(size_t)g1 * P1 + b1] : sb_start[(uint64_t)(b1 + 1) << b2];
        // cursor of sub-bucket sb: its start plus what the earlier groups hold of it
        const uint64_t nbins = 1ull << bits;
        for (int b = threadIdx.x; b < nbk; b += BLOCK) {
            const uint64_t sb = ((uint64_t)b1 << b2) + b;
            uint64_t cur = sb_start[sb];
            for (int v = 0; v < w; v++) cur += colsum[(size_t)v * nbins + sb];
            L.cur[b] = cur;
        }
    }
    // ranges are below 2^31 tuples
    const bool bad = scatter_range<BLOCK, ITEMS>(L, ikey + lo, ival + lo, (uint32_t)(hi - lo), nbk, shift,
                                                 bmask, cs_log2, okey, oval, out_cap);
    if (bad) atomicOr((unsigned long long*)err, 1ull);
}

// ============================================================== Bloom build
// One workgroup per filter segment: zero the segment in LDS, OR in the block
// of every R key of the segment's sub-buckets (one 64-bit LDS atomic per key),
// write the segment once.
template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_bloom_build(const uint32_t* __restrict__ rkey,
                                                       const uint64_t* __restrict__ r_start, int bits,
                                                       BloomGeom g, uint64_t* __restrict__ filter)
{
    extern __shared__ unsigned long long seg[];
    const uint32_t wps = g.wps;
    for (uint32_t i = threadIdx.x; i < wps; i += BLOCK) seg[i] = 0ull;
    __syncthreads();
    const uint64_t s = blockIdx.x;
    const int d = bits - g.seg_bits;
    const uint64_t lo = r_start[s << d], hi = r_start[(s + 1) << d];
    const uint64_t segbase = s * (uint64_t)wps;
    bloom_build_range<BLOCK>(rkey, lo, hi, g, segbase, seg);
    __syncthreads();
    uint64_t* dst = filter + segbase;
    for (uint32_t i = threadIdx.x; i < wps; i += BLOCK) dst[i] = seg[i];
}

// ============================================================== Bloom filter S -> S' (before partitioning)
// The reference's order (queryAndCompactKernelPartitioned): S is filtered
// before it is partitioned.  Used when the filter fits the MALL, since each key
// probes a random filter block.  Grid-stride over tiles (prefilter_tile).
template <int BLOCK, int ITEMS>
__global__ __launch_bounds__(BLOCK) void k_bloom_prefilter(
    const uint32_t* __restrict__ skey, const uint32_t* __restrict__ sval, uint64_t n,
    const uint64_t* __restrict__ filter, BloomGeom g, uint32_t* __restrict__ okey,
    uint32_t* __restrict__ oval, uint64_t out_cap, uint64_t* __restrict__ count,
    uint64_t* __restrict__ err)
{
    constexpr int T = BLOCK * ITEMS;
    __shared__ uint32_t gcnt[ITEMS * (BLOCK / WAVE)];
    __shared__ uint64_t s_base;
    bool bad = false;
    for (uint64_t tile = (uint64_t)blockIdx.x * T; tile < n; tile += (uint64_t)gridDim.x * T)
        bad |= prefilter_tile<BLOCK, ITEMS>(skey, sval, tile, n, filter, g, okey, oval, out_cap, count, gcnt,
                                            &s_base);
    if (bad) atomicOr((unsigned long long*)err, 2ull);
}

// ============================================================== Bloom filter after pass 1
// S has been scattered by its first-pass bucket.  The filter slice of
// first-pass bucket p is contiguous (segments are the top radix bits), 1/P1 of
// the filter.  Workgroups are dealt to XCDs round-robin (workgroup i runs on
// XCD i % 8), so XCD x takes buckets x, x+8, ... as one sequence of tiles,
// handed out in order by the XCD's own counter (tile_ctr[x]): all its
// workgroups work on the same bucket, whose filter slice stays in the XCD's L2
// while its S tuples stream past (loaded non-temporally), and a workgroup that
// finishes early takes the next tile instead of running ahead into the next
// bucket.  Survivors are counting-sorted by sub-bucket in LDS per tile, reserved
// with one atomic per (tile, sub-bucket) and written, compacted, into their
// sub-bucket's region of the S layout (sub-bucket sb gets positions
// sb_start[sb] + [0, s_fill[sb])), so no second pass over S is needed.  Values
// are read only for survivors.
//
// Tile walk of one XCD: bucket p owns ceil(size / T) consecutive tile numbers
// starting at tb; a workgroup's tile numbers only grow, so it walks forward.
struct FilterWalk {
    int p;          // current bucket (>= P1: done)
    uint64_t tb;    // first tile number of bucket p
    uint64_t lo, hi;
};

__device__ __forceinline__ void filter_walk_to(FilterWalk& w, uint64_t t, const uint64_t* __restrict__ sb_start,
                                               int P1, int bits2, uint64_t T)
{
    while (w.p < P1) {
        const uint64_t nt = (w.hi - w.lo + T - 1) / T;
        if (t < w.tb + nt) return;
        w.tb += nt;
        w.p += 8;
        if (w.p < P1) {
            w.lo = sb_start[(uint64_t)w.p << bits2];
            w.hi = sb_start[(uint64_t)(w.p + 1) << bits2];
        }
    }
}

// K: the filter's bits per key when fixed at compile time (3 or 4), 0 = g.k.
template <int BLOCK, int ITEMS, int K>
__global__ __launch_bounds__(BLOCK) void k_filter_l1(
    const uint32_t* __restrict__ ikey, const uint32_t* __restrict__ ival,
    const uint64_t* __restrict__ sb_start, int bits, int bits2, const uint64_t* __restrict__ filter,
    BloomGeom g, uint32_t* __restrict__ okey, uint32_t* __restrict__ oval, uint64_t out_cap,
    uint32_t* __restrict__ s_fill, uint64_t* __restrict__ count, uint64_t* __restrict__ tile_ctr,
    uint64_t* __restrict__ err)
{
    constexpr int T = BLOCK * ITEMS;
    constexpr int NXCD = 8;
    const int P1 = 1 << (bits - bits2), P2 = 1 << bits2;
    const int shift = 32 - bits;
    const uint32_t bmask = (uint32_t)P2 - 1u;
    extern __shared__ uint64_t lds64[];
    uint64_t* dst0 = lds64;                       // [P2] sub-bucket start + reserved base
    uint32_t* skey = (uint32_t*)(dst0 + P2);      // [T]
    uint32_t* sval = skey + T;                    // [T]
    uint32_t* cnt0 = sval + T;                    // [2 * P2] survivors per sub-bucket (two tiles)
    uint32_t* off = cnt0 + 2 * P2;                // [P2]
    uint32_t* wtmp = off + P2;                    // [BLOCK/64 + 1]
    __shared__ uint32_t s_total;
    __shared__ uint64_t s_tile;
    const bool small = P2 <= 8 * WAVE;

    const int xcd = blockIdx.x % NXCD;
    unsigned long long* xctr = (unsigned long long*)tile_ctr + xcd;
    bool bad = false;
    uint32_t surv_total = 0;
    for (int b = threadIdx.x; b < 2 * P2; b += BLOCK) cnt0[b] = 0u;
    // tile numbers are taken one tile ahead: the current one and the next are
    // known, the one after that is requested while the current tile runs
    __shared__ uint64_t s_tile2;
    if (threadIdx.x == 0) {
        s_tile = atomicAdd(xctr, 1ull);
        s_tile2 = atomicAdd(xctr, 1ull);
    }
    __syncthreads();
    uint64_t t_next = s_tile2;
    uint32_t par = 0;

    FilterWalk wk;
    wk.p = xcd;
    wk.tb = 0;
    wk.lo = wk.hi = 0;
    if (wk.p < P1) {
        wk.lo = sb_start[(uint64_t)wk.p << bits2];
        wk.hi = sb_start[(uint64_t)(wk.p + 1) << bits2];
    }
    filter_walk_to(wk, s_tile, sb_start, P1, bits2, T);
    // the current tile [tile, chi) of bucket p; keys are streamed once
    // (non-temporal, so the filter slice keeps the L2) and the next tile's keys
    // load while the current tile is written out
    int p = wk.p;
    uint64_t tile = wk.lo + (s_tile - wk.tb) * (uint64_t)T;
    uint64_t chi = min(tile + (uint64_t)T, wk.hi);
    uint32_t nk[ITEMS];
    if (p < P1) {
#pragma unroll
        for (int it = 0; it < ITEMS; it++) {
            const uint64_t i = tile + (uint64_t)it * BLOCK + threadIdx.x;
            nk[it] = (i < chi) ? __builtin_nontemporal_load(&ikey[i]) : 0u;
        }
    }

    // Per tile three barriers: test + rank | offsets and reservations (one wave;
    // it also clears the other tile's counts and publishes the next tile
    // number) | staging | out (runs into the next tile's test).
    while (p < P1) {
        const int tn = (int)(chi - tile);
        uint32_t* cnt = cnt0 + (par ? P2 : 0);
        uint32_t* cnext = cnt0 + (par ? 0 : P2);
        // the tile number after the next one, in flight during this tile
        unsigned long long t_after = 0ull;
        if (threadIdx.x == 0) t_after = atomicAdd(xctr, 1ull);
        uint32_t k[ITEMS], bk[ITEMS], rk[ITEMS];
        bool pass[ITEMS];
#pragma unroll
        for (int it = 0; it < ITEMS; it++) {
            const int s = it * BLOCK + threadIdx.x;
            k[it] = (s < tn) ? nk[it] : 0u;
        }
        // all filter words of the tile in flight before any is tested.  The
        // loads are unconditional (a lane past the tile reads the word of key
        // 0): loads under a branch, with addresses from earlier loads, would
        // each wait for all loads before them.
        uint64_t fw[ITEMS];
#pragma unroll
        for (int it = 0; it < ITEMS; it++) fw[it] = filter[bloom_block(k[it], g)];
#pragma unroll
        for (int it = 0; it < ITEMS; it++) {
            const int s = it * BLOCK + threadIdx.x;
            const uint64_t m = K ? bloom_mask_k<(K ? K : 1)>(k[it]) : bloom_mask(k[it], g.k);
            const bool ok = s < tn && (fw[it] & m) == m;
            pass[it] = ok;
            if (ok) {
                bk[it] = (radix_hash(k[it]) >> shift) & bmask;
                rk[it] = atomicAdd(&cnt[bk[it]], 1u);
            }
        }
        __syncthreads();
        if (small) {
            if (threadIdx.x < WAVE) {
                const int per = (P2 + WAVE - 1) / WAVE;
                const int b0 = min((int)threadIdx.x * per, P2), b1 = min(b0 + per, P2);
                uint32_t sum = 0;
                for (int b = b0; b < b1; b++) sum += cnt[b];
                const uint32_t inc = wave_inclusive_scan<uint32_t>(sum);
                uint32_t ex = inc - sum;
                for (int b = b0; b < b1; b++) {
                    const uint32_t c = cnt[b];
                    off[b] = ex;
                    ex += c;
                    const uint64_t sb = ((uint64_t)p << bits2) + b;
                    dst0[b] = sb_start[sb] + (c ? atomicAdd(&s_fill[sb], c) : 0u);
                    cnext[b] = 0u;
                }
                if (threadIdx.x == WAVE - 1) s_total = inc;
            }
        } else {
            const uint32_t tot = lds_exclusive_scan<BLOCK>(cnt, off, P2, wtmp);
            if (threadIdx.x == 0) s_total = tot;
            for (int b = threadIdx.x; b < P2; b += BLOCK) {
                const uint64_t sb = ((uint64_t)p << bits2) + b;
                const uint32_t c = cnt[b];
                dst0[b] = sb_start[sb] + (c ? atomicAdd(&s_fill[sb], c) : 0u);
                cnext[b] = 0u;
            }
        }
        // survivors' values: loads first, then the LDS stores
        uint32_t vv[ITEMS];
#pragma unroll
        for (int it = 0; it < ITEMS; it++)
            if (pass[it]) vv[it] = __builtin_nontemporal_load(&ival[tile + it * BLOCK + threadIdx.x]);
        __syncthreads();
        // the next tile: its keys are issued after this tile's value loads, so no
        // wait of this tile covers them (clamped rather than skipped past the end,
        // as above)
        filter_walk_to(wk, t_next, sb_start, P1, bits2, T);
        const int p_next = wk.p;
        const uint64_t tile_next = wk.lo + (t_next - wk.tb) * (uint64_t)T;
        const uint64_t chi_next = min(tile_next + (uint64_t)T, wk.hi);
        if (p_next < P1) {
#pragma unroll
            for (int it = 0; it < ITEMS; it++) {
                const uint64_t i = min(tile_next + (uint64_t)it * BLOCK + threadIdx.x, chi_next - 1);
                nk[it] = __builtin_nontemporal_load(&ikey[i]);
            }
        }
#pragma unroll
        for (int it = 0; it < ITEMS; it++) {
            if (pass[it]) {
                const uint32_t q = off[bk[it]] + rk[it];
                skey[q] = k[it];
                sval[q] = vv[it];
            }
        }
        // publish the tile number after the next (thread 0 waits for it only here)
        if (threadIdx.x == 0) s_tile = t_after;
        __syncthreads();
        const uint32_t ts = s_total;
        for (uint32_t q = threadIdx.x; q < ts; q += BLOCK) {
            const uint32_t kk = skey[q];
            const uint32_t b = (radix_hash(kk) >> shift) & bmask;
            const uint64_t d = dst0[b] + (q - off[b]);
            if (d < out_cap) {
                okey[d] = kk;
                oval[d] = sval[q];
            } else {
                bad = true;
            }
        }
        surv_total += ts;
        p = p_next;
        tile = tile_next;
        chi = chi_next;
        t_next = s_tile;  // read before thread 0 rewrites it after the next tile's third barrier
        par ^= 1u;
    }
    if (threadIdx.x == 0 && surv_total)
        atomicAdd((unsigned long long*)count, (unsigned long long)surv_total);
    if (bad) atomicOr((unsigned long long*)err, 8ull);
}

// ============================================================== S' of the chunks of S
// With S in chunks, every chunk's S' (its surviving tuples, in sub-bucket
// regions) stays in its own region of a pool, and one probe runs after the
// last chunk on all of them, so R streams past the tables once instead of once
// per chunk.  Note: sub-bucket sb of the chunk whose region starts at base has
// its piece at base + s_start[sb], cnt[sb] tuples long (the chunk's row of the
// piece table); *total accumulates the tuples.
constexpr int POOL_BLOCK = 256;

__global__ __launch_bounds__(POOL_BLOCK) void k_pool_note(const uint64_t* __restrict__ s_start,
                                                          const uint32_t* __restrict__ s_cnt, int nbins, uint64_t base,
                                                          uint32_t* __restrict__ cnt, uint64_t* __restrict__ off,
                                                          uint64_t* __restrict__ total)
{
    const int sb = blockIdx.x * POOL_BLOCK + threadIdx.x;
    uint64_t n = 0;
    if (sb < nbins) {
        const uint64_t s0 = s_start[sb];
        n = s_cnt ? (uint64_t)s_cnt[sb] : s_start[sb + 1] - s0;
        cnt[sb] = (uint32_t)n;
        off[sb] = base + s0;
    }
    n = wave_sum<uint64_t>(n);
    if (lane_id() == 0 && n) atomicAdd((unsigned long long*)total, (unsigned long long)n);
}

// totals[sb] = the pieces of sub-bucket sb summed over the nch chunks.
__global__ void k_pool_totals(const uint32_t* __restrict__ cnt, int nch, int nbins, uint32_t* __restrict__ totals)
{
    const int sb = blockIdx.x * blockDim.x + threadIdx.x;
    if (sb >= nbins) return;
    uint32_t t = 0;
    for (int c = 0; c < nch; c++) t += cnt[(size_t)c * nbins + sb];
    totals[sb] = t;
}

// Sub-bucket sb (one workgroup) gathers its pieces, chunk by chunk, into
// [start[sb], start[sb + 1]) of ok / ov.
__global__ __launch_bounds__(POOL_BLOCK) void k_pool_merge(
    const uint32_t* __restrict__ pk, const uint32_t* __restrict__ pv, const uint32_t* __restrict__ cnt,
    const uint64_t* __restrict__ off, int nch, int nbins, const uint64_t* __restrict__ start,
    uint32_t* __restrict__ ok, uint32_t* __restrict__ ov, uint64_t out_cap, uint64_t* __restrict__ err)
{
    const int sb = blockIdx.x;
    uint64_t d = start[sb];
    if (start[sb + 1] > out_cap) {
        if (threadIdx.x == 0) atomicOr((unsigned long long*)err, 16ull);
        return;
    }
    for (int c = 0; c < nch; c++) {
        const uint32_t n = cnt[(size_t)c * nbins + sb];
        const uint64_t o = off[(size_t)c * nbins + sb];
        for (uint32_t i = threadIdx.x; i < n; i += POOL_BLOCK) {
            ok[d + i] = pk[o + i];
            ov[d + i] = pv[o + i];
        }
        d += n;
    }
}

// ============================================================== probe planning
// A sub-bucket whose build side (build_from_r: R when it fits one LDS tile,
// else the smaller side) fits one LDS tile and whose probe side fits one slice
// is "regular": one work item, indexed by the sub-bucket itself.
// Larger ones are split (vlsplit): the build side into ceil(nb / tile) chunks,
// the probe side into min(ceil(np / slice), max_slices) slices, one explicit
// work item per (chunk, slice).  Count, scan, fill: no contention, and the
// item buffer bound max_slices * (sum ceil(nb / tile)) always holds.
__device__ __forceinline__ void sb_sizes(const uint64_t* r_start, const uint64_t* s_start,
                                         const uint32_t* s_cnt, int sb, uint64_t& r0, uint64_t& nr,
                                         uint64_t& s0, uint64_t& ns)
{
    r0 = r_start[sb];
    nr = r_start[sb + 1] - r0;
    s0 = s_start[sb];
    ns = s_cnt ? (uint64_t)s_cnt[sb] : (s_start[sb + 1] - s0);
}

__global__ void k_plan_count(const uint64_t* __restrict__ r_start, const uint64_t* __restrict__ s_start,
                             const uint32_t* __restrict__ s_cnt, int nbins, uint32_t tile,
                             uint32_t slice, uint32_t max_slices, uint32_t* __restrict__ n_items)
{
    const int sb = blockIdx.x * blockDim.x + threadIdx.x;
    if (sb >= nbins) return;
    uint64_t r0, nr, s0, ns;
    sb_sizes(r_start, s_start, s_cnt, sb, r0, nr, s0, ns);
    uint32_t c, sl;
    n_items[sb] = plan_items(nr, ns, tile, slice, max_slices, c, sl);
}

// One workgroup: thread t owns a run of at most 32 sub-buckets; one block scan
// of the counts into item_start (item_start[nbins] = total, also stored at
// *total_out when given), then the items of the (few) split sub-buckets.
__global__ __launch_bounds__(1024) void k_plan_fill(const uint64_t* __restrict__ r_start,
                                                    const uint64_t* __restrict__ s_start,
                                                    const uint32_t* __restrict__ s_cnt, int nbins, uint32_t tile,
                                                    uint32_t slice, uint32_t max_slices,
                                                    const uint32_t* __restrict__ n_items,
                                                    uint64_t* __restrict__ item_start, uint64_t item_cap,
                                                    HeavyItem* __restrict__ items, uint64_t* __restrict__ err,
                                                    uint64_t* __restrict__ total_out)
{
    // nbins is a power of two up to 2^15: thread t owns sub-buckets
    // [t * per, (t + 1) * per), per = nbins / 1024 (at least 1); counts read as
    // 16-byte vectors and starts written in pairs from 4 sub-buckets on
    constexpr int MAXPER = 32;
    __shared__ uint64_t wtmp[1024 / WAVE + 1];
    const int per = max(1, nbins >> 10);
    const int b0 = (int)threadIdx.x * per, b1 = min(b0 + per, nbins);
    uint32_t v[MAXPER];
    if (per >= 4) {
        const uint4* n4 = (const uint4*)(n_items + b0);
#pragma unroll
        for (int j = 0; j < MAXPER / 4; j++) {
            const uint4 q = (4 * j < per) ? n4[j] : make_uint4(0u, 0u, 0u, 0u);
            v[4 * j] = q.x;
            v[4 * j + 1] = q.y;
            v[4 * j + 2] = q.z;
            v[4 * j + 3] = q.w;
        }
    } else {
#pragma unroll
        for (int i = 0; i < 4; i++) v[i] = (b0 + i < b1) ? n_items[b0 + i] : 0u;
#pragma unroll
        for (int i = 4; i < MAXPER; i++) v[i] = 0u;
    }
    uint64_t sum = 0;
#pragma unroll
    for (int i = 0; i < MAXPER; i++) sum += v[i];
    uint64_t total;
    uint64_t base = block_exclusive_scan<1024, uint64_t>(sum, wtmp, &total);
    if (threadIdx.x == 0) {
        item_start[nbins] = total;
        if (total_out) *total_out = total;
    }
    if (per >= 4) {
        ulonglong2* o2 = (ulonglong2*)(item_start + b0);
        uint64_t e = base;
#pragma unroll
        for (int j = 0; j < MAXPER / 2; j++) {
            const uint64_t e0 = e, e1 = e + v[2 * j];
            e = e1 + v[2 * j + 1];
            if (2 * j < per) o2[j] = make_ulonglong2(e0, e1);
        }
    }
    bool over = false;
#pragma unroll
    for (int i = 0; i < MAXPER; i++) {
        const int sb = b0 + i;
        if (sb < b1) {
            if (per < 4) item_start[sb] = base;
            if (v[i]) {
                uint64_t r0, nr, s0, ns;
                sb_sizes(r_start, s_start, s_cnt, sb, r0, nr, s0, ns);
                uint32_t chunks, slices;
                plan_items(nr, ns, tile, slice, max_slices, chunks, slices);
                if (base + v[i] > item_cap) {
                    over = true;
                } else {
                    for (uint32_t c = 0; c < chunks; c++)
                        for (uint32_t q = 0; q < slices; q++) {
                            HeavyItem it;
                            it.sb = (uint32_t)sb;
                            it.chunk = c;
                            it.slice = q;
                            it.nslices = slices;
                            items[base + (uint64_t)c * slices + q] = it;
                        }
                }
            }
            base += v[i];
        }
    }
    if (over) atomicOr((unsigned long long*)err, 4ull);
}

// ============================================================== build + probe
// Persistent workgroups pull work items from a global counter (the next item
// is fetched while the current one runs): first the vlsplit items (the
// largest), then one item per regular sub-bucket.  Each item builds the
// build side's chunk into an LDS table and streams the other side's slice
// past it (probe_item, bj_tiles.h).
// COLLECT (the reference's COLLECT / --fused_probe_breakdown): thread 0 of
// every workgroup adds its phase times (PT_*) to timers[blockIdx.x * PT_N + p].
#ifndef BJ_PROBE_GRAB
#define BJ_PROBE_GRAB 2
#endif
constexpr int PROBE_GRAB = BJ_PROBE_GRAB;

template <int BLOCK, int CAP, int PITEMS, bool COLLECT>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(4, 8))) void k_probe(
    const uint32_t* __restrict__ rk, const uint32_t* __restrict__ rv,
    const uint64_t* __restrict__ r_start, const uint32_t* __restrict__ sk,
    const uint32_t* __restrict__ sv, const uint64_t* __restrict__ s_start,
    const uint32_t* __restrict__ s_cnt, int bits, uint32_t slice,
    const HeavyItem* __restrict__ items, const uint64_t* __restrict__ n_items_p, uint64_t item_cap,
    uint64_t* __restrict__ work_ctr, uint64_t* __restrict__ out_count, uint64_t out_cap,
    uint32_t* __restrict__ o_key, uint32_t* __restrict__ o_rval, uint32_t* __restrict__ o_sval,
    uint64_t* __restrict__ timers)
{
    __shared__ ProbeLds<BLOCK, CAP> L;
    __shared__ uint64_t s_item, s_next;
    uint64_t tacc[PT_N] = {}, tmark = 0;
    if (COLLECT && threadIdx.x == 0) tmark = phase_clock();

    const int nbins = 1 << bits;
    const uint64_t n_heavy = min(*n_items_p, item_cap);
    const uint64_t n_work = n_heavy + (uint64_t)nbins;

    // items are claimed PROBE_GRAB at a time: thread 0 waits for the queue's
    // atomic once per claim instead of once per item
    if (threadIdx.x == 0) s_item = atomicAdd((unsigned long long*)work_ctr, (unsigned long long)PROBE_GRAB);
    __syncthreads();
    uint64_t item = s_item;
    int left = PROBE_GRAB;  // items of the current claim, this one included
    __syncthreads();

    while (item < n_work) {
        if (left == 1 && threadIdx.x == 0)
            s_next = atomicAdd((unsigned long long*)work_ctr, (unsigned long long)PROBE_GRAB);

        uint32_t sb, chunk = 0, sl = 0, nsl = 1;
        const bool heavy = item < n_heavy;
        if (heavy) {
            const HeavyItem hi = items[item];
            sb = hi.sb; chunk = hi.chunk; sl = hi.slice; nsl = hi.nslices;
        } else {
            sb = (uint32_t)(item - n_heavy);
        }
        uint64_t r0, nr, s0, ns;
        sb_sizes(r_start, s_start, s_cnt, (int)sb, r0, nr, s0, ns);
        const bool build_r = build_from_r(nr, ns, (uint32_t)CAP);
        const uint64_t nb = build_r ? nr : ns, np = build_r ? ns : nr;
        // empty, or a regular index whose sub-bucket is on the vlsplit list: nothing to do
        const bool skip = nr == 0 || ns == 0 || (!heavy && (nb > (uint64_t)CAP || np > (uint64_t)slice));
        if (!skip) {
            const uint64_t b_lo = (uint64_t)chunk * CAP;
            const int bn = (int)min((uint64_t)CAP, nb - b_lo);
            const uint64_t per = (np + nsl - 1) / nsl;
            const uint64_t p_lo = min((uint64_t)sl * per, np);
            const uint32_t pn = (uint32_t)min(per, np - p_lo);
            const uint64_t brow0 = (build_r ? r0 : s0) + b_lo;
            const uint64_t prow0 = (build_r ? s0 : r0) + p_lo;
            phase_tick<COLLECT>(tacc, &tmark, PT_ITEM);
            probe_item<BLOCK, CAP, PITEMS, COLLECT>(L, (build_r ? rk : sk) + brow0, (build_r ? rv : sv) + brow0,
                                                    bn, (build_r ? sk : rk) + prow0, (build_r ? sv : rv) + prow0,
                                                    pn, build_r, bits, out_count, out_cap, o_key, o_rval, o_sval,
                                                    tacc, &tmark);
        }
        __syncthreads();  // the table is rebuilt by the next item; s_next is written
        if (left > 1) {
            item++;
            left--;
        } else {
            item = s_next;
            left = PROBE_GRAB;
        }
        __syncthreads();
    }
    if (COLLECT && threadIdx.x == 0) {
        phase_tick<COLLECT>(tacc, &tmark, PT_ITEM);
        for (int p = 0; p < PT_N; p++) timers[(size_t)blockIdx.x * PT_N + p] += tacc[p];
    }
}

// ============================================================== launchers
static int ceil_div_i(uint64_t a, uint64_t b) { return (int)((a + b - 1) / b); }

hipError_t launch_add2_u64(const uint64_t* a, const uint64_t* b, uint64_t* acc, hipStream_t st)
{
    k_add2_u64<<<1, 64, 0, st>>>(a, b, acc);
    return hipGetLastError();
}

hipError_t launch_zero_u64(uint64_t* p, int n, hipStream_t st)
{
    k_zero_u64<<<1, 256, 0, st>>>(p, n);
    return hipGetLastError();
}

hipError_t launch_gen_r(uint32_t* key, uint32_t* val, uint64_t n, const PermKeys& K, hipStream_t st)
{
    if (n == 0) return hipSuccess;
    const int grid = (int)std::min<uint64_t>(65536, (n + 255) / 256);
    k_gen_r<<<grid, 256, 0, st>>>(key, val, n, K);
    return hipGetLastError();
}

hipError_t launch_gen_s(uint32_t* key, uint32_t* val, uint64_t n_s, uint64_t n_r, uint64_t m_s,
                        int L, const PermKeys& KR, const PermKeys& KP, const PermKeys& KN,
                        uint32_t useed, double zipf_s, uint32_t zseed, double* d_tmp,
                        hipStream_t st)
{
    if (n_s == 0) return hipSuccess;
    double Hn = 0.0;
    if (zipf_s > 0.0) {
        k_zipf_Hn<<<1, 64, 0, st>>>(d_tmp, (uint32_t)n_r, zipf_s);
        hipError_t e = hipMemcpyAsync(&Hn, d_tmp, sizeof(double), hipMemcpyDeviceToHost, st);
        if (e != hipSuccess) return e;
        e = hipStreamSynchronize(st);
        if (e != hipSuccess) return e;
    }
    const int grid = (int)std::min<uint64_t>(65536, (n_s + 255) / 256);
    k_gen_s<<<grid, 256, 0, st>>>(key, val, n_s, n_r, m_s, L, KR, KP, KN, useed, zipf_s, Hn, zseed);
    return hipGetLastError();
}

static int log2_ceil_u64(uint64_t n)
{
    int l = 0;
    while ((1ull << l) < n) l++;
    return l;
}

hipError_t launch_gen_refdist_r(uint32_t* key, uint64_t n, uint64_t k, const PermKeys& K, hipStream_t st)
{
    if (n == 0) return hipSuccess;
    const int grid = (int)std::min<uint64_t>(65536, (n + 255) / 256);
    k_gen_refdist_r<<<grid, 256, 0, st>>>(key, n, k, std::max(1, log2_ceil_u64(n)), K);
    return hipGetLastError();
}

hipError_t launch_gen_refdist_s(uint32_t* key, uint64_t n, uint64_t m, uint64_t mc0, uint64_t sj0, uint64_t n_r,
                                const PermKeys& K, hipStream_t st)
{
    if (n == 0) return hipSuccess;
    const int grid = (int)std::min<uint64_t>(65536, (n + 255) / 256);
    k_gen_refdist_s<<<grid, 256, 0, st>>>(key, n, m, mc0, sj0, n_r, std::max(1, log2_ceil_u64(n)), K);
    return hipGetLastError();
}

constexpr int HIST_BLOCK = 1024;
constexpr int SC_BLOCK = 1024;
constexpr int SC_ITEMS = 4;
constexpr int SC_BIG_ITEMS = 8;  // 8192-tuple tiles for one-workgroup-per-CU launches
static size_t g_max_lds = 64 * 1024;
constexpr int FL_BLOCK = 512;
constexpr int FL_ITEMS = 8;

// Allow dynamic LDS up to the device limit minus the kernel's static LDS.
static hipError_t set_lds(const void* fn, size_t bytes)
{
    hipFuncAttributes attr;
    hipError_t e = hipFuncGetAttributes(&attr, fn);
    if (e != hipSuccess) return e;
    const size_t stat = attr.sharedSizeBytes;
    return hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)(bytes > stat ? bytes - stat : 0));
}

hipError_t kernels_init(int max_lds_bytes)
{
    hipError_t e;
    if ((e = set_lds((const void*)k_hist<HIST_BLOCK>, max_lds_bytes)) != hipSuccess) return e;
    if ((e = set_lds((const void*)k_scatter<SC_BLOCK, SC_ITEMS>, max_lds_bytes)) != hipSuccess) return e;
    if ((e = set_lds((const void*)k_scatter<SC_BLOCK / 2, SC_ITEMS * 2>, max_lds_bytes)) != hipSuccess) return e;
    if ((e = set_lds((const void*)k_scatter<SC_BLOCK, SC_BIG_ITEMS>, max_lds_bytes)) != hipSuccess) return e;
    g_max_lds = (size_t)max_lds_bytes;
    if ((e = set_lds((const void*)k_bloom_build<1024>, max_lds_bytes)) != hipSuccess) return e;
    if ((e = set_lds((const void*)k_filter_l1<FL_BLOCK, FL_ITEMS, 0>, max_lds_bytes)) != hipSuccess) return e;
    if ((e = set_lds((const void*)k_filter_l1<FL_BLOCK, FL_ITEMS, 3>, max_lds_bytes)) != hipSuccess) return e;
    if ((e = set_lds((const void*)k_filter_l1<FL_BLOCK, FL_ITEMS, 4>, max_lds_bytes)) != hipSuccess) return e;
    return hipSuccess;
}

hipError_t launch_hist(const uint32_t* key, uint64_t n_host, const uint64_t* n_dev, int G, int bits,
                       uint32_t* part, uint32_t* totals, int bits2, uint32_t* l1h, hipStream_t st)
{
    const size_t lds = sizeof(uint32_t) << bits;
    k_hist<HIST_BLOCK><<<G, HIST_BLOCK, lds, st>>>(key, n_host, n_dev, G, bits, part, totals, bits2, l1h);
    return hipGetLastError();
}

hipError_t launch_zero_u32(uint32_t* p, uint64_t n, hipStream_t st)
{
    if (n == 0) return hipSuccess;
    const int grid = (int)std::min<uint64_t>(4096, (n + 255) / 256);
    k_zero_u32<<<grid, 256, 0, st>>>(p, n);
    return hipGetLastError();
}

hipError_t launch_offsets(const uint32_t* part, int G, int bits, int bits2, int W, uint32_t* colsum,
                          uint32_t* totals, uint32_t* l1h, uint64_t* sb_start, uint64_t* l1off, hipStream_t st)
{
    if (bits > 15) return hipErrorInvalidValue;
    const int nbins = 1 << bits, P1 = 1 << (bits - bits2);
    k_colsums<<<ceil_div_i(nbins, 1024) * W, 1024, 0, st>>>(part, G, bits, W, colsum, totals);
    k_starts<<<1, 1024, 0, st>>>(totals, nbins, sb_start);
    k_l1_offsets<<<ceil_div_i(P1, L1_COLS), 1024, 0, st>>>(bits2 ? l1h : part, G, P1, bits2, sb_start, l1off);
    return hipGetLastError();
}

size_t scan_partials_needed(uint64_t n) { return (size_t)ceil_div_i(n ? n : 1, SCAN_TILE) + 1; }

hipError_t launch_scan(const uint32_t* in, uint64_t n, uint64_t* out, uint64_t* partial,
                       hipStream_t st)
{
    const int W = ceil_div_i(n ? n : 1, SCAN_TILE);
    if (W > 1024 * 64) return hipErrorInvalidValue;
    k_scan_reduce<<<W, SCAN_BLOCK, 0, st>>>(in, n, partial);
    k_scan_top<<<1, 1024, 0, st>>>(partial, W);
    k_scan_down<<<W, SCAN_BLOCK, 0, st>>>(in, n, partial, out, W);
    return hipGetLastError();
}

// Write-combining geometry: tile T = SC_BLOCK * SC_ITEMS tuples, carry chunk
// cs tuples (64 B up to 512 buckets, then 32 B / 16 B so the carry stays at
// 64 KB of LDS).
int scatter_carry(int nbk) { return nbk <= 512 ? 16 : (nbk <= 1024 ? 8 : 4); }

static int log2i(int x)
{
    int l = 0;
    while ((1 << l) < x) l++;
    return l;
}

// One-workgroup-per-CU launches have the CU's LDS to themselves: whole 128-byte
// lines of keys and of values per carried chunk up to 256 buckets.
static int scatter_carry_big(int nbk) { return nbk <= 256 ? 32 : scatter_carry(nbk); }

static size_t scatter_lds_tile(int nbk, int tile, int cs)
{
    return (size_t)nbk * 8 + (size_t)nbk * cs * 8 + (size_t)tile * 8 + (size_t)nbk * 20 + (SC_BLOCK / WAVE + 1) * 4;
}

size_t scatter_lds_bytes(int nbk) { return scatter_lds_tile(nbk, SC_BLOCK * SC_ITEMS, scatter_carry(nbk)); }

size_t scatter_lds_bytes_big(int nbk)
{
    return scatter_lds_tile(nbk, SC_BLOCK * SC_BIG_ITEMS, scatter_carry_big(nbk));
}

hipError_t launch_scatter(const uint32_t* ikey, const uint32_t* ival, uint32_t* okey, uint32_t* oval,
                          uint64_t out_cap, uint64_t n_host, const uint64_t* n_dev, int G, int mode,
                          int bits, int p1, int W2, const uint64_t* l1off, const uint64_t* sb_start,
                          const uint32_t* colsum, uint64_t* err, int wg, size_t min_lds, hipStream_t st)
{
    const int nbk = (mode == 0) ? (1 << p1) : (1 << (bits - p1));
    const int grid = (mode == 0) ? G : (1 << p1) * W2;
    const int cs = log2i(scatter_carry(nbk));
    const size_t lds = std::max(scatter_lds_bytes(nbk), min_lds);
    const int cs_big = scatter_carry_big(nbk);
    const size_t lds_big = std::max(scatter_lds_tile(nbk, SC_BLOCK * SC_BIG_ITEMS, cs_big), min_lds);
    if (wg != 512 && min_lds > 0 && lds_big <= g_max_lds) {
        // one workgroup per CU: twice the tile, twice the loads in flight, whole-line carries
        k_scatter<SC_BLOCK, SC_BIG_ITEMS><<<grid, SC_BLOCK, lds_big, st>>>(
            ikey, ival, okey, oval, out_cap, n_host, n_dev, G, mode, bits, p1, W2, log2i(cs_big), l1off, sb_start, colsum,
            err);
    } else if (wg == 512) {
        // same tile (4096 tuples), half the threads: more workgroups per CU
        k_scatter<SC_BLOCK / 2, SC_ITEMS * 2><<<grid, SC_BLOCK / 2, lds, st>>>(
            ikey, ival, okey, oval, out_cap, n_host, n_dev, G, mode, bits, p1, W2, cs, l1off, sb_start, colsum,
            err);
    } else {
        k_scatter<SC_BLOCK, SC_ITEMS><<<grid, SC_BLOCK, lds, st>>>(
            ikey, ival, okey, oval, out_cap, n_host, n_dev, G, mode, bits, p1, W2, cs, l1off, sb_start, colsum,
            err);
    }
    return hipGetLastError();
}

hipError_t launch_bloom_build(const uint32_t* rkey, const uint64_t* r_start, int bits,
                              const BloomGeom& g, uint64_t* filter, hipStream_t st)
{
    const int nseg = 1 << g.seg_bits;
    k_bloom_build<1024><<<nseg, 1024, (size_t)g.wps * 8, st>>>(rkey, r_start, bits, g, filter);
    return hipGetLastError();
}

hipError_t launch_bloom_prefilter(const uint32_t* skey, const uint32_t* sval, uint64_t n,
                                  const uint64_t* filter, const BloomGeom& g, uint32_t* okey,
                                  uint32_t* oval, uint64_t out_cap, uint64_t* count, uint64_t* err,
                                  int num_cu, hipStream_t st)
{
    if (n == 0) return hipSuccess;
    constexpr int B = 512, I = 8;
    const int grid = (int)std::min<uint64_t>((uint64_t)num_cu * 8, (n + B * I - 1) / (B * I));
    k_bloom_prefilter<B, I><<<grid, B, 0, st>>>(skey, sval, n, filter, g, okey, oval, out_cap, count, err);
    return hipGetLastError();
}

hipError_t launch_filter_l1(const uint32_t* ikey, const uint32_t* ival, const uint64_t* sb_start, int bits,
                            int bits2, const uint64_t* filter, const BloomGeom& g, uint32_t* okey,
                            uint32_t* oval, uint64_t out_cap, uint32_t* s_fill, uint64_t* count,
                            uint64_t* tile_ctr, uint64_t* err, int num_cu, hipStream_t st)
{
    const int P2 = 1 << bits2;
    const size_t lds = (size_t)P2 * 8 + (size_t)FL_BLOCK * FL_ITEMS * 8 + (size_t)P2 * 12 +
                       (FL_BLOCK / WAVE + 1) * 4;
    const int W = std::max(1, num_cu / 8) * 2;  // workgroups per XCD
#define BJ_FILTER(K)                                                                                      \
    k_filter_l1<FL_BLOCK, FL_ITEMS, K><<<8 * W, FL_BLOCK, lds, st>>>(ikey, ival, sb_start, bits, bits2, filter, g, \
                                                                     okey, oval, out_cap, s_fill, count, tile_ctr, err)
    if (g.k == 3) BJ_FILTER(3);
    else if (g.k == 4) BJ_FILTER(4);
    else BJ_FILTER(0);
#undef BJ_FILTER
    return hipGetLastError();
}

hipError_t launch_pool_note(const uint64_t* s_start, const uint32_t* s_cnt, int nbins, uint64_t base, uint32_t* cnt,
                            uint64_t* off, uint64_t* total, hipStream_t st)
{
    k_pool_note<<<ceil_div_i((uint64_t)nbins, POOL_BLOCK), POOL_BLOCK, 0, st>>>(s_start, s_cnt, nbins, base, cnt, off,
                                                                                total);
    return hipGetLastError();
}

hipError_t launch_pool_start(const uint32_t* cnt, int nch, int nbins, uint32_t* totals, uint64_t* start,
                             uint64_t* partials, hipStream_t st)
{
    k_pool_totals<<<ceil_div_i((uint64_t)nbins, 256), 256, 0, st>>>(cnt, nch, nbins, totals);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    return launch_scan(totals, (uint64_t)nbins, start, partials, st);
}

hipError_t launch_pool_merge(const uint32_t* pk, const uint32_t* pv, const uint32_t* cnt, const uint64_t* off,
                             int nch, int nbins, const uint64_t* start, uint32_t* ok, uint32_t* ov,
                             uint64_t out_cap, uint64_t* err, hipStream_t st)
{
    k_pool_merge<<<nbins, POOL_BLOCK, 0, st>>>(pk, pv, cnt, off, nch, nbins, start, ok, ov, out_cap, err);
    return hipGetLastError();
}

hipError_t launch_plan(const uint64_t* r_start, const uint64_t* s_start, const uint32_t* s_cnt, int bits,
                       uint32_t tile, uint32_t slice, uint32_t max_slices, uint32_t* counts,
                       uint64_t* item_start, uint64_t* total_out, HeavyItem* items, uint64_t item_cap,
                       uint64_t* err, hipStream_t st)
{
    if (bits > 15) return hipErrorInvalidValue;
    const int nbins = 1 << bits;
    k_plan_count<<<ceil_div_i(nbins, 256), 256, 0, st>>>(r_start, s_start, s_cnt, nbins, tile, slice, max_slices,
                                                         counts);
    k_plan_fill<<<1, 1024, 0, st>>>(r_start, s_start, s_cnt, nbins, tile, slice, max_slices, counts, item_start,
                                    item_cap, items, err, total_out);
    return hipGetLastError();
}

int probe_grid(int tile, int num_cu) { return tile == 8192 ? num_cu * 2 : num_cu * 4; }

// Probe keys per thread per round: 10 (a round of 5120 keys at 512 threads)
// measured 11% faster than 8 on the 640M join, 6 was 23% slower, and 12 no
// faster than 10 while spilling registers.
constexpr int PROBE_ITEMS = 10;

static_assert(PT_N == PROBE_TIMER_SLOTS, "probe timer slots");

hipError_t launch_probe(int tile, const uint32_t* rk, const uint32_t* rv, const uint64_t* r_start,
                        const uint32_t* sk, const uint32_t* sv, const uint64_t* s_start,
                        const uint32_t* s_cnt, int bits, uint32_t slice, const HeavyItem* items,
                        const uint64_t* n_items, uint64_t item_cap, uint64_t* work_ctr, uint64_t* out_count,
                        uint64_t out_cap, uint32_t* o_key, uint32_t* o_rval, uint32_t* o_sval, int num_cu,
                        uint64_t* timers, hipStream_t st)
{
#define BJ_PROBE(B, C, G)                                                                                  \
    do {                                                                                                 \
        if (timers)                                                                                      \
            k_probe<B, C, PROBE_ITEMS, true><<<(G), B, 0, st>>>(rk, rv, r_start, sk, sv, s_start, s_cnt, bits, slice, \
                                                      items, n_items, item_cap, work_ctr, out_count,     \
                                                      out_cap, o_key, o_rval, o_sval, timers);           \
        else                                                                                             \
            k_probe<B, C, PROBE_ITEMS, false><<<(G), B, 0, st>>>(rk, rv, r_start, sk, sv, s_start, s_cnt, bits,    \
                                                       slice, items, n_items, item_cap, work_ctr,        \
                                                       out_count, out_cap, o_key, o_rval, o_sval, nullptr); \
    } while (0)
    if (tile == 8192) BJ_PROBE(512, 8192, probe_grid(8192, num_cu));
    else if (tile == 4096) BJ_PROBE(256, 4096, probe_grid(4096, num_cu));
    else return hipErrorInvalidValue;
#undef BJ_PROBE
    return hipGetLastError();
}


}  // namespace bj

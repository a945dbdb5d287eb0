// bj_kernels.hip -- gfx950 kernels of the radix hash join, and their launchers.
//
// Pipeline (one GPU; all integer, all exact):
//   R:  k_hist -> k_reduce_hist -> scans -> k_scatter(pass 1) -> k_scatter(pass 2)
//   BF: k_bloom_build (one workgroup per filter segment, built in LDS)
//   S:  k_bloom_filter (query + wave-ballot compaction S -> S')   [Bloom on]
//       then the same two-level partition as R
//   J:  k_plan_probe (vlsplit work items) -> k_probe (persistent; LDS chained
//       hash table on the smaller side, stream the other side, LDS row staging)
//
// Reference counterparts (Chanho-LEE-SNU/BRIG-Join):
//   k_hist / k_reduce_hist / scans   gpu_run_hist_for_both, gpu_make_simple_prefix,
//                                    gpu_run_hist_second_pass
//                                    (src/kernel_functions.cu:22-90, 461-492, 163-207)
//   k_scatter                        gpu_make_sorted_table_for_both, gpu_build_table_second_pass
//                                    (src/kernel_functions.cu:95-133, 210-281)
//   k_bloom_build                    insertKernelPartitioned (src/bloom.cu:698-722)
//   k_bloom_filter                   queryAndCompactKernelPartitioned (src/bloom.cu:725-773)
//   k_probe                          gpu_probe_with_second_pass / gpu_fused_probe_transfer
//                                    local part (src/kernel_functions.cu:294-457, 612-789)
#include "bj_kernels.h"
#include "bj_device.h"

#include <math.h>

#include <algorithm>

namespace bj {

// ============================================================== utilities
__global__ void k_zero_u64(uint64_t* __restrict__ p, int n)
{
    for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0ull;
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
                                                int bits, uint32_t* __restrict__ part)
{
    extern __shared__ uint32_t h[];
    const int nbins = 1 << bits;
    for (int i = threadIdx.x; i < nbins; i += BLOCK) h[i] = 0u;
    __syncthreads();
    const uint64_t n = n_dev ? *n_dev : n_host;
    const uint64_t chunk = (n + (uint64_t)G - 1) / (uint64_t)G;
    const uint64_t lo = min((uint64_t)blockIdx.x * chunk, n);
    const uint64_t hi = min(lo + chunk, n);
    const int sh = 32 - bits;
    constexpr int U = 8;
    uint64_t i = lo + threadIdx.x;
    for (; i + (uint64_t)(U - 1) * BLOCK < hi; i += (uint64_t)U * BLOCK) {
        uint32_t k[U];
#pragma unroll
        for (int u = 0; u < U; u++) k[u] = key[i + (uint64_t)u * BLOCK];
#pragma unroll
        for (int u = 0; u < U; u++) atomicAdd(&h[radix_hash(k[u]) >> sh], 1u);
    }
    for (; i < hi; i += BLOCK) atomicAdd(&h[radix_hash(key[i]) >> sh], 1u);
    __syncthreads();
    uint32_t* dst = part + (size_t)blockIdx.x * nbins;
    for (int b = threadIdx.x; b < nbins; b += BLOCK) dst[b] = h[b];
}

// totals[b] = sum_g part[g][b]  (one thread per bin, coalesced across bins)
__global__ void k_reduce_totals(const uint32_t* __restrict__ part, int G, int nbins,
                                uint32_t* __restrict__ totals)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nbins) return;
    uint32_t s = 0;
    for (int g = 0; g < G; g++) s += part[(size_t)g * nbins + b];
    totals[b] = s;
}

// l1part[b1 * G + g] = sum_{b2} part[g][(b1 << bits2) + b2]
__global__ void k_reduce_l1(const uint32_t* __restrict__ part, int G, int bits, int bits2,
                            uint32_t* __restrict__ l1part)
{
    const int P1 = 1 << (bits - bits2);
    const int nbins = 1 << bits;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= P1 * G) return;
    const int g = t / P1, b1 = t % P1;
    const uint32_t* src = part + (size_t)g * nbins + ((size_t)b1 << bits2);
    uint32_t s = 0;
    for (int b2 = 0; b2 < (1 << bits2); b2++) s += src[b2];
    l1part[(size_t)b1 * G + g] = s;
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
// LDS-staged scatter.  Each tile of T tuples is counting-sorted by bucket in
// LDS (rank from an LDS atomic, offsets from an LDS scan), then written out so
// that consecutive lanes write consecutive addresses inside each bucket run.
//   mode 0 (pass 1): workgroup g owns the same chunk as k_hist's group g;
//                    bucket = h >> (32 - bits1); cursor[b] = l1off[b * G + g].
//   mode 1 (pass 2): workgroup p owns first-level bucket p (its range is
//                    sb_start[p << bits2] .. sb_start[(p+1) << bits2]);
//                    bucket = (h >> (32 - bits)) & (P2 - 1);
//                    cursor[b] = sb_start[(p << bits2) + b].
// Writes are bounds-checked against out_cap and flag err on violation.
template <int BLOCK, int ITEMS>
__global__ __launch_bounds__(BLOCK) void k_scatter(
    const uint32_t* __restrict__ ikey, const uint32_t* __restrict__ ival,
    uint32_t* __restrict__ okey, uint32_t* __restrict__ oval, uint64_t out_cap,
    uint64_t n_host, const uint64_t* __restrict__ n_dev, int G, int mode, int bits, int bits2,
    const uint64_t* __restrict__ l1off, const uint64_t* __restrict__ sb_start,
    uint64_t* __restrict__ err)
{
    constexpr int T = BLOCK * ITEMS;
    const int nbk = (mode == 0) ? (1 << (bits - bits2)) : (1 << bits2);
    const int shift = (mode == 0) ? (32 - (bits - bits2)) : (32 - bits);
    const uint32_t bmask = (uint32_t)nbk - 1u;

    extern __shared__ uint64_t lds64[];
    uint64_t* cur = lds64;                                 // [nbk]
    uint32_t* skey = (uint32_t*)(cur + nbk);               // [T]
    uint32_t* sval = skey + T;                             // [T]
    uint32_t* cnt = sval + T;                              // [nbk]
    uint32_t* off = cnt + nbk;                             // [nbk]
    uint32_t* wtmp = off + nbk;                            // [BLOCK/64 + 1]

    uint64_t lo, hi;
    if (mode == 0) {
        const uint64_t n = n_dev ? *n_dev : n_host;
        const uint64_t chunk = (n + (uint64_t)G - 1) / (uint64_t)G;
        lo = min((uint64_t)blockIdx.x * chunk, n);
        hi = min(lo + chunk, n);
        for (int b = threadIdx.x; b < nbk; b += BLOCK) cur[b] = l1off[(size_t)b * G + blockIdx.x];
    } else {
        const uint64_t p = blockIdx.x;
        lo = sb_start[p << bits2];
        hi = sb_start[(p + 1) << bits2];
        for (int b = threadIdx.x; b < nbk; b += BLOCK) cur[b] = sb_start[(p << bits2) + b];
    }

    for (uint64_t tile = lo; tile < hi; tile += T) {
        const int tn = (int)min((uint64_t)T, hi - tile);
        for (int b = threadIdx.x; b < nbk; b += BLOCK) cnt[b] = 0u;
        __syncthreads();
        uint32_t k[ITEMS], v[ITEMS], bk[ITEMS], rk[ITEMS];
#pragma unroll
        for (int it = 0; it < ITEMS; it++) {
            const int s = it * BLOCK + threadIdx.x;
            if (s < tn) {
                k[it] = ikey[tile + s];
                v[it] = ival[tile + s];
            }
        }
#pragma unroll
        for (int it = 0; it < ITEMS; it++) {
            const int s = it * BLOCK + threadIdx.x;
            if (s < tn) {
                bk[it] = (radix_hash(k[it]) >> shift) & bmask;
                rk[it] = atomicAdd(&cnt[bk[it]], 1u);
            }
        }
        __syncthreads();
        lds_exclusive_scan<BLOCK>(cnt, off, nbk, wtmp);
        __syncthreads();
#pragma unroll
        for (int it = 0; it < ITEMS; it++) {
            const int s = it * BLOCK + threadIdx.x;
            if (s < tn) {
                const uint32_t p = off[bk[it]] + rk[it];
                skey[p] = k[it];
                sval[p] = v[it];
            }
        }
        __syncthreads();
        bool bad = false;
#pragma unroll
        for (int it = 0; it < ITEMS; it++) {
            const int s = it * BLOCK + threadIdx.x;
            if (s < tn) {
                const uint32_t kk = skey[s];
                const uint32_t b = (radix_hash(kk) >> shift) & bmask;
                const uint64_t dst = cur[b] + (uint64_t)(s - (int)off[b]);
                if (dst < out_cap) {
                    okey[dst] = kk;
                    oval[dst] = sval[s];
                } else {
                    bad = true;
                }
            }
        }
        if (bad) atomicOr((unsigned long long*)err, 1ull);
        __syncthreads();
        for (int b = threadIdx.x; b < nbk; b += BLOCK) cur[b] += cnt[b];
        __syncthreads();
    }
}

// ============================================================== Bloom build
// One workgroup per filter segment: zero the segment in LDS, OR in the k bits
// of every R key of the segment's sub-buckets, write the segment once.
template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_bloom_build(const uint32_t* __restrict__ rkey,
                                                       const uint64_t* __restrict__ r_start, int bits,
                                                       BloomGeom g, uint32_t* __restrict__ filter)
{
    extern __shared__ uint32_t seg[];
    const uint32_t wps = g.wps;
    for (uint32_t i = threadIdx.x; i < wps; i += BLOCK) seg[i] = 0u;
    __syncthreads();
    const uint64_t s = blockIdx.x;
    const int d = bits - g.seg_bits;
    const uint64_t lo = r_start[s << d], hi = r_start[(s + 1) << d];
    const uint64_t segbase = s * (uint64_t)wps;
    for (uint64_t i = lo + threadIdx.x; i < hi; i += BLOCK) {
        BloomAddr a;
        bloom_addr(rkey[i], g, a);
        const uint32_t wb = (uint32_t)(a.block_word - segbase);
#pragma unroll
        for (int w = 0; w < 8; w++)
            if (a.mask[w]) atomicOr(&seg[wb + w], a.mask[w]);
    }
    __syncthreads();
    uint4* dst = (uint4*)(filter + segbase);
    const uint4* src = (const uint4*)seg;
    for (uint32_t i = threadIdx.x; i < wps / 4; i += BLOCK) dst[i] = src[i];
}

// ============================================================== Bloom filter S -> S'
// Query each S key's 256-bit block (two 16-byte loads), keep the survivors.
// Compaction: wave ballot + mbcnt inside each (item, wave) group, an LDS scan
// over the groups, one global atomic per tile.  s_val is read only for
// survivors.
template <int BLOCK, int ITEMS>
__global__ __launch_bounds__(BLOCK) void k_bloom_filter(
    const uint32_t* __restrict__ skey, const uint32_t* __restrict__ sval, uint64_t n,
    const uint32_t* __restrict__ filter, BloomGeom g, uint32_t* __restrict__ okey,
    uint32_t* __restrict__ oval, uint64_t out_cap, uint64_t* __restrict__ count,
    uint64_t* __restrict__ err)
{
    constexpr int NW = BLOCK / WAVE;
    constexpr int NG = ITEMS * NW;  // groups per tile
    constexpr int T = BLOCK * ITEMS;
    static_assert(NG <= 2 * WAVE, "group scan handles at most 128 groups");
    __shared__ uint32_t gcnt[NG];
    __shared__ uint64_t s_base;
    const int wid = threadIdx.x / WAVE;
    const uint32_t lane = lane_id();

    for (uint64_t tile = (uint64_t)blockIdx.x * T; tile < n; tile += (uint64_t)gridDim.x * T) {
        uint32_t k[ITEMS];
        bool pass[ITEMS];
#pragma unroll
        for (int it = 0; it < ITEMS; it++) {
            const uint64_t i = tile + (uint64_t)it * BLOCK + threadIdx.x;
            k[it] = (i < n) ? skey[i] : 0u;
        }
#pragma unroll
        for (int it = 0; it < ITEMS; it++) {
            const uint64_t i = tile + (uint64_t)it * BLOCK + threadIdx.x;
            bool ok = false;
            if (i < n) {
                BloomAddr a;
                bloom_addr(k[it], g, a);
                const uint4* blk = (const uint4*)(filter + a.block_word);
                const uint4 lo4 = blk[0], hi4 = blk[1];
                ok = ((lo4.x & a.mask[0]) == a.mask[0]) && ((lo4.y & a.mask[1]) == a.mask[1]) &&
                     ((lo4.z & a.mask[2]) == a.mask[2]) && ((lo4.w & a.mask[3]) == a.mask[3]) &&
                     ((hi4.x & a.mask[4]) == a.mask[4]) && ((hi4.y & a.mask[5]) == a.mask[5]) &&
                     ((hi4.z & a.mask[6]) == a.mask[6]) && ((hi4.w & a.mask[7]) == a.mask[7]);
            }
            pass[it] = ok;
        }
        uint32_t rank[ITEMS];
#pragma unroll
        for (int it = 0; it < ITEMS; it++) {
            const uint64_t m = __ballot(pass[it]);
            rank[it] = lanes_below(m);
            if (lane == 0) gcnt[it * NW + wid] = (uint32_t)__popcll(m);
        }
        __syncthreads();
        if (wid == 0) {
            // exclusive scan of NG (<= 128) group counts: two entries per lane
            const uint32_t a0 = (2 * lane < (uint32_t)NG) ? gcnt[2 * lane] : 0u;
            const uint32_t a1 = (2 * lane + 1 < (uint32_t)NG) ? gcnt[2 * lane + 1] : 0u;
            const uint32_t inc = wave_inclusive_scan<uint32_t>(a0 + a1);
            const uint32_t ex = inc - a0 - a1;
            if (2 * lane < (uint32_t)NG) gcnt[2 * lane] = ex;
            if (2 * lane + 1 < (uint32_t)NG) gcnt[2 * lane + 1] = ex + a0;
            const uint32_t total = __shfl(inc, WAVE - 1, WAVE);
            if (lane == 0) s_base = total ? atomicAdd((unsigned long long*)count,
                                                      (unsigned long long)total) : 0ull;
        }
        __syncthreads();
        bool bad = false;
#pragma unroll
        for (int it = 0; it < ITEMS; it++) {
            if (pass[it]) {
                const uint64_t i = tile + (uint64_t)it * BLOCK + threadIdx.x;
                const uint64_t pos = s_base + gcnt[it * NW + wid] + rank[it];
                if (pos < out_cap) {
                    okey[pos] = k[it];
                    oval[pos] = sval[i];
                } else {
                    bad = true;
                }
            }
        }
        if (bad) atomicOr((unsigned long long*)err, 2ull);
        __syncthreads();
    }
}

// ============================================================== probe planning
// Sub-buckets whose smaller side fits one LDS tile and whose larger side fits
// one slice are "regular" and are processed as one work item, indexed by the
// sub-bucket itself.  Larger ones are split (vlsplit): the build side into
// ceil(nb / tile) chunks, the probe side into ceil(np / slice) slices, one
// explicit work item per (chunk, slice).  If the item buffer would overflow,
// a sub-bucket falls back to one slice per chunk (always fits: the buffer holds
// at least sum(ceil(nb/tile)) items).
__global__ void k_plan_probe(const uint64_t* __restrict__ r_start, const uint64_t* __restrict__ s_start,
                             int nbins, uint32_t tile, uint32_t slice, HeavyItem* __restrict__ items,
                             uint64_t item_cap, uint64_t* __restrict__ n_items,
                             uint64_t* __restrict__ err)
{
    const int sb = blockIdx.x * blockDim.x + threadIdx.x;
    if (sb >= nbins) return;
    const uint64_t nr = r_start[sb + 1] - r_start[sb];
    const uint64_t ns = s_start[sb + 1] - s_start[sb];
    if (nr == 0 || ns == 0) return;
    const uint64_t nb = min(nr, ns), np = max(nr, ns);
    if (nb <= tile && np <= slice) return;
    const uint64_t chunks = (nb + tile - 1) / tile;
    uint64_t slices = (np + slice - 1) / slice;
    unsigned long long cur = atomicAdd((unsigned long long*)n_items, 0ull);
    unsigned long long take;
    for (;;) {
        take = chunks * slices;
        if (cur + take > item_cap) {
            slices = 1;
            take = chunks;
        }
        if (cur + take > item_cap) {
            atomicOr((unsigned long long*)err, 4ull);
            return;
        }
        const unsigned long long prev =
            atomicCAS((unsigned long long*)n_items, cur, cur + take);
        if (prev == cur) break;
        cur = prev;
        slices = (np + slice - 1) / slice;
    }
    for (uint64_t c = 0; c < chunks; c++)
        for (uint64_t s = 0; s < slices; s++) {
            HeavyItem it;
            it.sb = (uint32_t)sb;
            it.chunk = (uint32_t)c;
            it.slice = (uint32_t)s;
            it.nslices = (uint32_t)slices;
            items[cur + c * slices + s] = it;
        }
}

// ============================================================== build + probe
// Persistent workgroups pull work items from a global counter: first the
// vlsplit items (largest first), then one item per regular sub-bucket.
// Per item: the smaller side's chunk is loaded into an LDS chained hash table
// (head[] of index+1 per bucket, nxt[] links), then the other side's slice is
// streamed through it.  Matches are staged in LDS as (key, build value, probe
// row) and flushed with one global atomic per flush; the probe side's value is
// read only for matching rows, at flush time.
template <int BLOCK, int CAP, int PITEMS>
__global__ __launch_bounds__(BLOCK) void k_probe(
    const uint32_t* __restrict__ rk, const uint32_t* __restrict__ rv,
    const uint64_t* __restrict__ r_start, const uint32_t* __restrict__ sk,
    const uint32_t* __restrict__ sv, const uint64_t* __restrict__ s_start, int bits,
    uint32_t slice, const HeavyItem* __restrict__ items, const uint64_t* __restrict__ n_items_p,
    uint64_t* __restrict__ work_ctr, uint64_t* __restrict__ out_count, uint64_t out_cap,
    uint32_t* __restrict__ o_key, uint32_t* __restrict__ o_rval, uint32_t* __restrict__ o_sval)
{
    constexpr int NBMAX = CAP / 2;
    constexpr int OB = 2 * BLOCK;
    __shared__ uint32_t hkey[CAP];
    __shared__ uint32_t hval[CAP];
    __shared__ uint16_t nxt[CAP];
    __shared__ uint32_t head[NBMAX];
    __shared__ uint32_t ob_key[OB];
    __shared__ uint32_t ob_bval[OB];
    __shared__ uint32_t ob_pidx[OB];
    __shared__ uint32_t s_ocnt;
    __shared__ uint64_t s_item;
    __shared__ unsigned long long s_base;

    const int nbins = 1 << bits;
    const uint64_t n_heavy = *n_items_p;
    const uint64_t n_work = n_heavy + (uint64_t)nbins;
    if (threadIdx.x == 0) s_ocnt = 0u;
    __syncthreads();

    // Flush the staged rows; every thread must call it (contains barriers).
    auto flush = [&](uint32_t staged) {
        const uint32_t nrow = min(staged, (uint32_t)OB);
        if (threadIdx.x == 0)
            s_base = nrow ? atomicAdd((unsigned long long*)out_count, (unsigned long long)nrow) : 0ull;
        __syncthreads();
        const uint64_t base = s_base;
        for (uint32_t i = threadIdx.x; i < nrow; i += BLOCK) {
            const uint64_t gi = base + i;
            if (gi < out_cap) {
                const uint32_t pid = ob_pidx[i];
                const uint32_t prow = pid & 0x7FFFFFFFu;
                o_key[gi] = ob_key[i];
                if (pid & 0x80000000u) {  // build side S, probe side R
                    o_rval[gi] = rv[prow];
                    o_sval[gi] = ob_bval[i];
                } else {                  // build side R, probe side S
                    o_rval[gi] = ob_bval[i];
                    o_sval[gi] = sv[prow];
                }
            }
        }
        __syncthreads();
        if (threadIdx.x == 0) s_ocnt = 0u;
        __syncthreads();
    };

    for (;;) {
        if (threadIdx.x == 0) s_item = atomicAdd((unsigned long long*)work_ctr, 1ull);
        __syncthreads();
        const uint64_t item = s_item;
        __syncthreads();
        if (item >= n_work) break;

        uint32_t sb, chunk = 0, sl = 0, nsl = 1;
        const bool heavy = item < n_heavy;
        if (heavy) {
            const HeavyItem hi = items[item];
            sb = hi.sb; chunk = hi.chunk; sl = hi.slice; nsl = hi.nslices;
        } else {
            sb = (uint32_t)(item - n_heavy);
        }
        const uint64_t r0 = r_start[sb], nr = r_start[sb + 1] - r0;
        const uint64_t s0 = s_start[sb], ns = s_start[sb + 1] - s0;
        if (nr == 0 || ns == 0) continue;
        const bool build_r = nr <= ns;
        const uint64_t nb = build_r ? nr : ns, np = build_r ? ns : nr;
        if (!heavy && (nb > (uint64_t)CAP || np > (uint64_t)slice)) continue;  // on the vlsplit list

        const uint64_t b_lo = (uint64_t)chunk * CAP;
        const int bn = (int)min((uint64_t)CAP, nb - b_lo);
        const uint64_t per = (np + nsl - 1) / nsl;
        const uint64_t p_lo = min((uint64_t)sl * per, np);
        const uint64_t pn = min(per, np - p_lo);
        const uint32_t* bkp = (build_r ? rk : sk) + (build_r ? r0 : s0) + b_lo;
        const uint32_t* bvp = (build_r ? rv : sv) + (build_r ? r0 : s0) + b_lo;
        const uint64_t prow0 = (build_r ? s0 : r0) + p_lo;
        const uint32_t* pkp = (build_r ? sk : rk) + prow0;
        const uint32_t pflag = build_r ? 0u : 0x80000000u;

        // table geometry: 2^tb buckets, tb = ceil(log2 bn) - 1 (load 1..2)
        int tb = 0;
        while ((1 << (tb + 1)) < bn && (1 << (tb + 1)) <= NBMAX) tb++;
        const int nbk = 1 << tb;
        const int tshift = 32 - tb;
        for (int i = threadIdx.x; i < nbk; i += BLOCK) head[i] = 0u;
        __syncthreads();
        for (int i = threadIdx.x; i < bn; i += BLOCK) {
            const uint32_t k = bkp[i];
            hkey[i] = k;
            hval[i] = bvp[i];
            const uint32_t t = tb ? ((radix_hash(k) << bits) >> tshift) : 0u;
            nxt[i] = (uint16_t)atomicExch(&head[t], (uint32_t)(i + 1));
        }
        __syncthreads();

        for (uint64_t base = 0; base < pn; base += (uint64_t)BLOCK * PITEMS) {
            uint32_t kk[PITEMS];
#pragma unroll
            for (int it = 0; it < PITEMS; it++) {
                const uint64_t idx = base + (uint64_t)it * BLOCK + threadIdx.x;
                kk[it] = (idx < pn) ? pkp[idx] : 0u;
            }
#pragma unroll
            for (int it = 0; it < PITEMS; it++) {
                const uint64_t idx = base + (uint64_t)it * BLOCK + threadIdx.x;
                if (idx < pn) {
                    const uint32_t k = kk[it];
                    const uint32_t t = tb ? ((radix_hash(k) << bits) >> tshift) : 0u;
                    uint32_t c = head[t];
                    while (c) {
                        const uint32_t j = c - 1u;
                        if (hkey[j] == k) {
                            const uint32_t pid = (uint32_t)(prow0 + idx) | pflag;
                            const uint32_t slot = atomicAdd(&s_ocnt, 1u);
                            if (slot < (uint32_t)OB) {
                                ob_key[slot] = k;
                                ob_bval[slot] = hval[j];
                                ob_pidx[slot] = pid;
                            } else {
                                // staging full (duplicate keys): write this row directly
                                const unsigned long long gi =
                                    atomicAdd((unsigned long long*)out_count, 1ull);
                                if (gi < out_cap) {
                                    const uint32_t prow = pid & 0x7FFFFFFFu;
                                    o_key[gi] = k;
                                    if (build_r) { o_rval[gi] = hval[j]; o_sval[gi] = sv[prow]; }
                                    else         { o_rval[gi] = rv[prow]; o_sval[gi] = hval[j]; }
                                }
                            }
                        }
                        c = nxt[j];
                    }
                }
                __syncthreads();
                const uint32_t staged = s_ocnt;
                __syncthreads();
                if (staged > (uint32_t)(OB - BLOCK)) flush(staged);
            }
        }
    }
    __syncthreads();
    const uint32_t staged = s_ocnt;
    __syncthreads();
    flush(staged);
}

// ============================================================== launchers
static int ceil_div_i(uint64_t a, uint64_t b) { return (int)((a + b - 1) / b); }

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

constexpr int HIST_BLOCK = 1024;
constexpr int SC_BLOCK = 512;
constexpr int SC_ITEMS = 8;

static hipError_t set_lds(const void* fn, size_t bytes)
{
    return hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
}

hipError_t kernels_init(int max_lds_bytes)
{
    hipError_t e;
    if ((e = set_lds((const void*)k_hist<HIST_BLOCK>, max_lds_bytes)) != hipSuccess) return e;
    if ((e = set_lds((const void*)k_scatter<SC_BLOCK, SC_ITEMS>, max_lds_bytes)) != hipSuccess) return e;
    if ((e = set_lds((const void*)k_bloom_build<1024>, max_lds_bytes)) != hipSuccess) return e;
    return hipSuccess;
}

hipError_t launch_hist(const uint32_t* key, uint64_t n_host, const uint64_t* n_dev, int G, int bits,
                       uint32_t* part, hipStream_t st)
{
    const size_t lds = sizeof(uint32_t) << bits;
    k_hist<HIST_BLOCK><<<G, HIST_BLOCK, lds, st>>>(key, n_host, n_dev, G, bits, part);
    return hipGetLastError();
}

hipError_t launch_reduce(const uint32_t* part, int G, int bits, int bits2, uint32_t* totals,
                         uint32_t* l1part, hipStream_t st)
{
    const int nbins = 1 << bits;
    k_reduce_totals<<<ceil_div_i(nbins, 256), 256, 0, st>>>(part, G, nbins, totals);
    const int n1 = (1 << (bits - bits2)) * G;
    k_reduce_l1<<<ceil_div_i(n1, 256), 256, 0, st>>>(part, G, bits, bits2, l1part);
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

size_t scatter_lds_bytes(int nbk)
{
    return (size_t)nbk * 8 + (size_t)SC_BLOCK * SC_ITEMS * 8 + (size_t)nbk * 8 +
           (SC_BLOCK / WAVE + 1) * 4;
}

hipError_t launch_scatter(const uint32_t* ikey, const uint32_t* ival, uint32_t* okey, uint32_t* oval,
                          uint64_t out_cap, uint64_t n_host, const uint64_t* n_dev, int G, int mode,
                          int bits, int bits2, const uint64_t* l1off, const uint64_t* sb_start,
                          uint64_t* err, hipStream_t st)
{
    const int nbk = (mode == 0) ? (1 << (bits - bits2)) : (1 << bits2);
    const int grid = (mode == 0) ? G : (1 << (bits - bits2));
    k_scatter<SC_BLOCK, SC_ITEMS><<<grid, SC_BLOCK, scatter_lds_bytes(nbk), st>>>(
        ikey, ival, okey, oval, out_cap, n_host, n_dev, G, mode, bits, bits2, l1off, sb_start, err);
    return hipGetLastError();
}

hipError_t launch_bloom_build(const uint32_t* rkey, const uint64_t* r_start, int bits,
                              const BloomGeom& g, uint32_t* filter, hipStream_t st)
{
    const int nseg = 1 << g.seg_bits;
    k_bloom_build<1024><<<nseg, 1024, (size_t)g.wps * 4, st>>>(rkey, r_start, bits, g, filter);
    return hipGetLastError();
}

hipError_t launch_bloom_filter(const uint32_t* skey, const uint32_t* sval, uint64_t n,
                               const uint32_t* filter, const BloomGeom& g, uint32_t* okey,
                               uint32_t* oval, uint64_t out_cap, uint64_t* count, uint64_t* err,
                               int num_cu, hipStream_t st)
{
    if (n == 0) return hipSuccess;
    constexpr int B = 512, I = 8;
    const int grid = (int)std::min<uint64_t>((uint64_t)num_cu * 8, (n + B * I - 1) / (B * I));
    k_bloom_filter<B, I><<<grid, B, 0, st>>>(skey, sval, n, filter, g, okey, oval, out_cap, count, err);
    return hipGetLastError();
}

hipError_t launch_plan_probe(const uint64_t* r_start, const uint64_t* s_start, int bits,
                             uint32_t tile, uint32_t slice, HeavyItem* items, uint64_t item_cap,
                             uint64_t* n_items, uint64_t* err, hipStream_t st)
{
    const int nbins = 1 << bits;
    k_plan_probe<<<ceil_div_i(nbins, 256), 256, 0, st>>>(r_start, s_start, nbins, tile, slice, items,
                                                         item_cap, n_items, err);
    return hipGetLastError();
}

hipError_t launch_probe(int tile, const uint32_t* rk, const uint32_t* rv, const uint64_t* r_start,
                        const uint32_t* sk, const uint32_t* sv, const uint64_t* s_start, int bits,
                        uint32_t slice, const HeavyItem* items, const uint64_t* n_items,
                        uint64_t* work_ctr, uint64_t* out_count, uint64_t out_cap, uint32_t* o_key,
                        uint32_t* o_rval, uint32_t* o_sval, int num_cu, hipStream_t st)
{
    if (tile == 8192) {
        k_probe<1024, 8192, 4><<<num_cu, 1024, 0, st>>>(rk, rv, r_start, sk, sv, s_start, bits, slice,
                                                         items, n_items, work_ctr, out_count, out_cap,
                                                         o_key, o_rval, o_sval);
    } else if (tile == 4096) {
        k_probe<512, 4096, 4><<<num_cu * 2, 512, 0, st>>>(rk, rv, r_start, sk, sv, s_start, bits, slice,
                                                          items, n_items, work_ctr, out_count, out_cap,
                                                          o_key, o_rval, o_sval);
    } else {
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace bj

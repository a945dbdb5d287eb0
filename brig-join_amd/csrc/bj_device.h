// bj_device.h -- device helpers for the gfx950 radix hash join.
//
// Integer-only hashing, the synthetic generator's permutation, blocked Bloom
// addressing and wave64 scan primitives.  Everything here is exact.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace bj {

constexpr int WAVE = 64;

// lowbias32 (C. Wellons' published integer hash).  A bijection on uint32, so
// distinct keys keep distinct hashes and equal keys always share a bucket.
__host__ __device__ __forceinline__ uint32_t lowbias32(uint32_t x)
{
    x ^= x >> 16;
    x *= 0x7feb352dU;
    x ^= x >> 15;
    x *= 0x846ca68bU;
    x ^= x >> 16;
    return x;
}

// Radix hash: Fibonacci hashing, key * 2^32/phi (odd), a bijection on uint32 in
// one multiply.  Partition bucket = its top `bits` bits (which depend on every
// key bit); the LDS hash-table bucket of the build/probe kernel takes the bits
// just below them.
constexpr uint32_t RADIX_MUL = 0x9E3779B1u;
__host__ __device__ __forceinline__ uint32_t radix_hash(uint32_t key) { return key * RADIX_MUL; }

// ---------------------------------------------------------------- generator
// 4 rounds of (add, odd multiply, xor-shift right) mod 2^w: each step is a
// bijection on [0, 2^w), so the composition is a keyed permutation.
struct PermKeys { uint32_t k[8]; };

__host__ __device__ __forceinline__ uint32_t permw(uint32_t x, int w, const PermKeys& K)
{
    const uint32_t mask = (w >= 32) ? 0xFFFFFFFFu : ((1u << w) - 1u);
    const int sh = (w + 1) / 2;
    x &= mask;
#pragma unroll
    for (int r = 0; r < 4; r++) {
        x = (x + K.k[2 * r]) & mask;
        x = (x * K.k[2 * r + 1]) & mask;
        x ^= x >> sh;
    }
    return x;
}

// R key of row i: the 31-bit permutation cycle-walked onto [0, 2^31 - 1), +1.
// Distinct for distinct i and always in [1, 2^31).
__host__ __device__ __forceinline__ uint32_t gen_r_key(uint32_t i, const PermKeys& K)
{
    uint32_t y = permw(i, 31, K);
    while (y == 0x7FFFFFFFu) y = permw(y, 31, K);
    return y + 1u;
}

// ---------------------------------------------------------------- Bloom
// Register-blocked filter: 2^seg_bits segments of `wps` 64-bit blocks; all k
// bits of a key live in one 64-bit block of its segment, so a build is one
// 64-bit LDS atomic OR and a query is one 8-byte load.  The segment is the top
// seg_bits of the radix hash, so segment s holds exactly the R keys of the
// sub-buckets [s << (bits-seg_bits), (s+1) << (bits-seg_bits)).
struct BloomGeom {
    int seg_bits;
    uint32_t wps;   // 64-bit blocks per segment
    int k;          // bits per key, 1..8
};

__host__ __device__ __forceinline__ uint64_t bloom_block(uint32_t key, const BloomGeom& g)
{
    const uint32_t h = radix_hash(key);
    const uint64_t seg = g.seg_bits ? (uint64_t)(h >> (32 - g.seg_bits)) : 0ull;
    const uint32_t rest = g.seg_bits ? (h << g.seg_bits) : h;
    return seg * g.wps + (((uint64_t)rest * g.wps) >> 32);
}

// The k bit positions are 6-bit fields of a second multiplicative mix (fields
// taken from the top, xor-folded so the low fields see the high key bits):
// field i < 5 is bits [26 - 6i, 32 - 6i) of h2, field i >= 5 the same of h3.
// k is the same for every lane, so the loops below are scalar loops of k
// iterations (no work for absent fields).
__host__ __device__ __forceinline__ uint64_t bloom_mask(uint32_t key, int k)
{
    uint32_t h2 = (key ^ 0x5BD1E995u) * 0xC2B2AE3Du;
    h2 ^= h2 >> 15;
    uint64_t m = 0;
    const int k1 = k < 5 ? k : 5;
#pragma unroll 1
    for (int i = 0; i < k1; i++) m |= 1ull << ((h2 >> (26 - 6 * i)) & 63u);
    if (k > 5) {
        uint32_t h3 = h2 * 0x27D4EB2Fu;
        h3 ^= h3 >> 15;
#pragma unroll 1
        for (int i = 5; i < k; i++) m |= 1ull << ((h3 >> (26 - 6 * (i - 5))) & 63u);
    }
    return m;
}

// The same mask for a k fixed at compile time (k <= 5): straight-line code.
template <int K>
__device__ __forceinline__ uint64_t bloom_mask_k(uint32_t key)
{
    static_assert(K >= 1 && K <= 5, "fields of h2 only");
    uint32_t h2 = (key ^ 0x5BD1E995u) * 0xC2B2AE3Du;
    h2 ^= h2 >> 15;
    uint64_t m = 0;
#pragma unroll
    for (int i = 0; i < K; i++) m |= 1ull << ((h2 >> (26 - 6 * i)) & 63u);
    return m;
}

// ---------------------------------------------------------------- wave64 scans
__device__ __forceinline__ uint32_t lane_id() { return threadIdx.x & (WAVE - 1); }

// number of set bits of `m` below this lane
__device__ __forceinline__ uint32_t lanes_below(uint64_t m)
{
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

// sum over the wave, in every lane
template <typename T>
__device__ __forceinline__ T wave_sum(T x)
{
#pragma unroll
    for (int o = 1; o < WAVE; o <<= 1) x += __shfl_xor(x, o, WAVE);
    return x;
}

template <typename T>
__device__ __forceinline__ T wave_inclusive_scan(T x)
{
    const uint32_t lane = lane_id();
#pragma unroll
    for (int o = 1; o < WAVE; o <<= 1) {
        T y = __shfl_up(x, o, WAVE);
        if (lane >= (uint32_t)o) x += y;
    }
    return x;
}

// Block-wide exclusive scan of one value per thread.  `wtmp` holds
// BLOCK/64 + 1 entries of LDS.  Returns the exclusive prefix; *total = sum.
template <int BLOCK, typename T>
__device__ __forceinline__ T block_exclusive_scan(T v, T* wtmp, T* total)
{
    constexpr int NW = BLOCK / WAVE;
    const int wid = threadIdx.x / WAVE;
    const uint32_t lane = lane_id();
    const T inc = wave_inclusive_scan<T>(v);
    if (lane == WAVE - 1) wtmp[wid] = inc;
    __syncthreads();
    if (wid == 0) {
        T w = (lane < (uint32_t)NW) ? wtmp[lane] : (T)0;
        T wi = wave_inclusive_scan<T>(w);
        if (lane < (uint32_t)NW) wtmp[lane] = wi - w;
        if (lane == (uint32_t)(NW - 1)) wtmp[NW] = wi;
    }
    __syncthreads();
    const T res = wtmp[wid] + inc - v;
    *total = wtmp[NW];
    __syncthreads();
    return res;
}

// Exclusive scan of an LDS array in[0..n) into out[0..n) (may alias).
template <int BLOCK>
__device__ __forceinline__ uint32_t lds_exclusive_scan(const uint32_t* in, uint32_t* out, int n,
                                                      uint32_t* wtmp)
{
    const int per = (n + BLOCK - 1) / BLOCK;
    const int lo = min((int)threadIdx.x * per, n);
    const int hi = min(lo + per, n);
    uint32_t s = 0;
    for (int i = lo; i < hi; i++) s += in[i];
    uint32_t tot;
    uint32_t ex = block_exclusive_scan<BLOCK, uint32_t>(s, wtmp, &tot);
    for (int i = lo; i < hi; i++) {
        const uint32_t c = in[i];
        out[i] = ex;
        ex += c;
    }
    return tot;
}

}  // namespace bj

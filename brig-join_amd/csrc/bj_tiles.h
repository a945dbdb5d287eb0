// bj_tiles.h -- device bodies shared by the single-join kernels (bj_kernels.hip)
// and the batched kernels (bj_batch.hip).
//
// Each body works on one contiguous range (a histogram chunk, a scatter range,
// a filter segment, a probe work item); the kernels around them only decide
// which range a workgroup owns and where its results go.
#pragma once

#include "bj_device.h"

namespace bj {

// ---------------------------------------------------------------- histogram
// h[] (LDS, zeroed by the caller) += histogram of the top (32 - sh) bits of the
// radix hash over key[lo, hi).  8 loads in flight per thread.
template <int BLOCK>
__device__ __forceinline__ void hist_range(const uint32_t* __restrict__ key, uint64_t lo, uint64_t hi,
                                           int sh, uint32_t* h)
{
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
}

// ---------------------------------------------------------------- scatter
// LDS-staged scatter with write-combining carries.
//
// Each tile of T tuples is counting-sorted by bucket in LDS (rank from an LDS
// atomic, offsets from an LDS scan).  Bucket b's tuples go to consecutive global
// positions starting at cur[b].  Only whole CS-tuple chunks (CS * 4 bytes) are
// written; the unaligned tail of every bucket (< CS tuples) is carried in LDS
// to the next tile, so every chunk of the output is written by one workgroup
// within one tile.  Partial chunks occur only where two workgroups' ranges of
// a bucket meet.  The next tile is loaded into registers while the current one
// is written out.
struct ScatterLds {
    uint64_t* cur;   // [nbk] next global position
    uint32_t* ck;    // [nbk * cs] carried keys
    uint32_t* cv;    // [nbk * cs] carried values
    uint32_t* skey;  // [T]
    uint32_t* sval;  // [T]
    uint32_t* cnt0;  // [nbk] tile tuples (even tiles)
    uint32_t* cnt1;  // [nbk] tile tuples (odd tiles)
    uint32_t* off;   // [nbk] staging offset
    uint32_t* ccnt;  // [nbk] carried tuples
    uint32_t* nw;    // [nbk] tuples written this tile
    uint32_t* wtmp;  // [BLOCK/64 + 1]
};

template <int BLOCK, int ITEMS>
__device__ __forceinline__ ScatterLds scatter_lds(uint64_t* lds64, int nbk, int cs_log2)
{
    constexpr int T = BLOCK * ITEMS;
    ScatterLds L;
    L.cur = lds64;
    L.ck = (uint32_t*)(L.cur + nbk);
    L.cv = L.ck + (nbk << cs_log2);
    L.skey = L.cv + (nbk << cs_log2);
    L.sval = L.skey + T;
    L.cnt0 = L.sval + T;
    L.cnt1 = L.cnt0 + nbk;
    L.off = L.cnt1 + nbk;
    L.ccnt = L.off + nbk;
    L.nw = L.ccnt + nbk;
    L.wtmp = L.nw + nbk;
    return L;
}

// Bucket b after a tile: the nw[b] tuples written leave the carry / tile, the rest
// (ccnt + cnt - nw) stay carried.
__device__ __forceinline__ void scatter_retire(const ScatterLds& L, uint32_t* cprev, int b)
{
    const uint32_t w = L.nw[b];
    L.cur[b] += w;
    L.ccnt[b] = L.ccnt[b] + cprev[b] - w;
    cprev[b] = 0u;
}

// Tuples of bucket b written this tile: up to its last whole chunk.
__device__ __forceinline__ uint32_t scatter_nw(const ScatterLds& L, int b, uint32_t c, uint32_t cs)
{
    const uint32_t a0 = (uint32_t)L.cur[b] & (cs - 1u);
    const uint32_t e = a0 + L.ccnt[b] + c;
    return e >= cs ? (e & ~(cs - 1u)) - a0 : 0u;
}

// Scatters kin/vin[0, len) by bucket (radix_hash(key) >> shift) & bmask into
// okey/oval at the cursors L.cur[] (initialised by the caller; this function
// begins with a barrier).  Writes are bounds-checked against out_cap; returns
// true on a violation.
//
// Per tile three barriers: ranks (LDS atomics) | offsets (one wave scans the
// counts, retires the previous tile's buckets and sizes this tile's writes) |
// staging + carried tuples out + next tile's loads | the tile's tuples out or
// into the carry (which runs into the next tile's ranks).  Counts alternate
// between two arrays so the next tile's ranks need no clearing barrier.
template <int BLOCK, int ITEMS>
__device__ __forceinline__ bool scatter_range(const ScatterLds& L, const uint32_t* __restrict__ kin,
                                              const uint32_t* __restrict__ vin, uint32_t len, int nbk,
                                              int shift, uint32_t bmask, int cs_log2,
                                              uint32_t* __restrict__ okey, uint32_t* __restrict__ oval,
                                              uint64_t out_cap)
{
    constexpr int T = BLOCK * ITEMS;
    constexpr int SMALL = 8;  // buckets per lane of the one-wave scan
    const uint32_t cs = 1u << cs_log2;
    const bool small = nbk <= SMALL * WAVE;
    for (int b = threadIdx.x; b < nbk; b += BLOCK) {
        L.ccnt[b] = 0u;
        L.cnt0[b] = 0u;
        L.cnt1[b] = 0u;
        L.nw[b] = 0u;
    }
    // cur/ccnt are read by other threads; an empty range skips the tile loop's barriers
    __syncthreads();

    uint32_t k[ITEMS], v[ITEMS];
#pragma unroll
    for (int it = 0; it < ITEMS; it++) {
        const uint32_t i = (uint32_t)it * BLOCK + threadIdx.x;
        if (i < len) {
            k[it] = __builtin_nontemporal_load(&kin[i]);
            v[it] = __builtin_nontemporal_load(&vin[i]);
        }
    }
    bool bad = false;
    uint32_t par = 0;

    for (uint32_t tile = 0; tile < len; tile += T, par ^= 1u) {
        // cnt1 follows cnt0: select by offset (a select between the two struct
        // fields would put the struct in scratch memory)
        uint32_t* cnt = L.cnt0 + (par ? nbk : 0);
        uint32_t* cprev = L.cnt0 + (par ? 0 : nbk);
        const int tn = (int)min((uint32_t)T, len - tile);
        uint32_t bk[ITEMS], rk[ITEMS];
#pragma unroll
        for (int it = 0; it < ITEMS; it++) {
            const int s = it * BLOCK + threadIdx.x;
            if (s < tn) {
                bk[it] = (radix_hash(k[it]) >> shift) & bmask;
                rk[it] = atomicAdd(&cnt[bk[it]], 1u);
            }
        }
        __syncthreads();
        if (nbk > 2 * WAVE && nbk <= BLOCK) {
            // one bucket per thread, scanned by the whole workgroup: from 3
            // buckets per lane on, one wave working through them serially keeps
            // the others waiting longer than the scan's two extra barriers
            const int b = threadIdx.x;
            uint32_t c = 0;
            if (b < nbk) {
                scatter_retire(L, cprev, b);
                c = cnt[b];
            }
            uint32_t tot;
            const uint32_t ex = block_exclusive_scan<BLOCK, uint32_t>(c, L.wtmp, &tot);
            if (b < nbk) {
                L.off[b] = ex;
                L.nw[b] = scatter_nw(L, b, c, cs);
            }
        } else if (small) {
            if (threadIdx.x < WAVE) {
                // lane l owns buckets [l*per, l*per + per)
                const int per = (nbk + WAVE - 1) / WAVE;
                const int b0 = min((int)threadIdx.x * per, nbk), b1 = min(b0 + per, nbk);
                uint32_t sum = 0;
                for (int b = b0; b < b1; b++) {
                    scatter_retire(L, cprev, b);
                    sum += cnt[b];
                }
                uint32_t ex = wave_inclusive_scan<uint32_t>(sum) - sum;
                for (int b = b0; b < b1; b++) {
                    const uint32_t c = cnt[b];
                    L.off[b] = ex;
                    ex += c;
                    L.nw[b] = scatter_nw(L, b, c, cs);
                }
            }
        } else {
            for (int b = threadIdx.x; b < nbk; b += BLOCK) scatter_retire(L, cprev, b);
            lds_exclusive_scan<BLOCK>(cnt, L.off, nbk, L.wtmp);
            for (int b = threadIdx.x; b < nbk; b += BLOCK) L.nw[b] = scatter_nw(L, b, cnt[b], cs);
        }
        __syncthreads();
#pragma unroll
        for (int it = 0; it < ITEMS; it++) {
            const int s = it * BLOCK + threadIdx.x;
            if (s < tn) {
                const uint32_t p = L.off[bk[it]] + rk[it];
                L.skey[p] = k[it];
                L.sval[p] = v[it];
            }
        }
        // prefetch the next tile while this one is written out
        const uint32_t next = tile + T;
#pragma unroll
        for (int it = 0; it < ITEMS; it++) {
            const uint32_t i = next + (uint32_t)it * BLOCK + threadIdx.x;
            if (i < len) {
                k[it] = __builtin_nontemporal_load(&kin[i]);
                v[it] = __builtin_nontemporal_load(&vin[i]);
            }
        }
        // carried tuples inside the written range go out
        for (int idx = threadIdx.x; idx < (nbk << cs_log2); idx += BLOCK) {
            const int b = idx >> cs_log2;
            const uint32_t i = (uint32_t)idx & (cs - 1u);
            if (i < L.ccnt[b] && i < L.nw[b]) {
                const uint64_t g = L.cur[b] + i;
                if (g < out_cap) {
                    okey[g] = L.ck[idx];
                    oval[g] = L.cv[idx];
                } else {
                    bad = true;
                }
            }
        }
        __syncthreads();
        // the tile's tuples go out, or into the carry
        for (int s = threadIdx.x; s < tn; s += BLOCK) {
            const uint32_t kk = L.skey[s];
            const uint32_t b = (radix_hash(kk) >> shift) & bmask;
            const uint32_t rel = L.ccnt[b] + (uint32_t)s - L.off[b];
            const uint32_t w_b = L.nw[b];
            if (rel < w_b) {
                const uint64_t g = L.cur[b] + rel;
                if (g < out_cap) {
                    okey[g] = kk;
                    oval[g] = L.sval[s];
                } else {
                    bad = true;
                }
            } else {
                const uint32_t slot = (b << cs_log2) + rel - w_b;
                L.ck[slot] = kk;
                L.cv[slot] = L.sval[s];
            }
        }
    }
    __syncthreads();
    // retire the last tile, then the remaining carries (partial chunks at the range ends)
    {
        uint32_t* cprev = L.cnt0 + (par ? 0 : nbk);
        for (int b = threadIdx.x; b < nbk; b += BLOCK) scatter_retire(L, cprev, b);
    }
    __syncthreads();
    for (int idx = threadIdx.x; idx < (nbk << cs_log2); idx += BLOCK) {
        const int b = idx >> cs_log2;
        const uint32_t i = (uint32_t)idx & (cs - 1u);
        if (i < L.ccnt[b]) {
            const uint64_t g = L.cur[b] + i;
            if (g < out_cap) {
                okey[g] = L.ck[idx];
                oval[g] = L.cv[idx];
            } else {
                bad = true;
            }
        }
    }
    return bad;
}

// ---------------------------------------------------------------- Bloom build
// seg[] (LDS, wps 64-bit blocks, zeroed by the caller) |= the block masks of
// rkey[lo, hi); all keys belong to the segment starting at block segbase.
template <int BLOCK>
__device__ __forceinline__ void bloom_build_range(const uint32_t* __restrict__ rkey, uint64_t lo,
                                                  uint64_t hi, const BloomGeom& g, uint64_t segbase,
                                                  unsigned long long* seg)
{
    constexpr int U = 8;
    for (uint64_t i = lo + threadIdx.x; i < hi; i += (uint64_t)U * BLOCK) {
        uint32_t kk[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            const uint64_t j = i + (uint64_t)u * BLOCK;
            kk[u] = (j < hi) ? __builtin_nontemporal_load(&rkey[j]) : 0u;
        }
#pragma unroll
        for (int u = 0; u < U; u++) {
            if (i + (uint64_t)u * BLOCK < hi) {
                const uint32_t w = (uint32_t)(bloom_block(kk[u], g) - segbase);
                atomicOr(&seg[w], (unsigned long long)bloom_mask(kk[u], g.k));
            }
        }
    }
}

// ---------------------------------------------------------------- Bloom filter S -> S'
// One tile [tile, tile + BLOCK*ITEMS) of skey/sval (clipped to n): every key
// tests its 64-bit block; survivors are compacted with a wave ballot + mbcnt
// per (item, wave) group, an LDS scan over the groups and one global atomic on
// *count per tile, and written to okey/oval at that position.  Values are read
// only for survivors.  gcnt holds ITEMS * BLOCK/64 entries, s_base one.
// Returns true if a write would pass out_cap.
template <int BLOCK, int ITEMS>
__device__ __forceinline__ bool prefilter_tile(const uint32_t* __restrict__ skey,
                                               const uint32_t* __restrict__ sval, uint64_t tile, uint64_t n,
                                               const uint64_t* __restrict__ filter, const BloomGeom& g,
                                               uint32_t* __restrict__ okey, uint32_t* __restrict__ oval,
                                               uint64_t out_cap, uint64_t* __restrict__ count, uint32_t* gcnt,
                                               uint64_t* s_base)
{
    constexpr int NW = BLOCK / WAVE;
    constexpr int NG = ITEMS * NW;  // groups per tile
    static_assert(NG <= 2 * WAVE, "group scan handles at most 128 groups");
    const int wid = threadIdx.x / WAVE;
    const uint32_t lane = lane_id();
    uint32_t k[ITEMS];
    bool pass[ITEMS];
#pragma unroll
    for (int it = 0; it < ITEMS; it++) {
        const uint64_t i = tile + (uint64_t)it * BLOCK + threadIdx.x;
        k[it] = (i < n) ? __builtin_nontemporal_load(&skey[i]) : 0u;
    }
    // every filter word in flight before any is tested; the loads are
    // unconditional (a lane past n reads the word of key 0): loads under a
    // branch, with addresses from earlier loads, would each wait for all loads
    // before them
    uint64_t fw[ITEMS];
#pragma unroll
    for (int it = 0; it < ITEMS; it++) fw[it] = filter[bloom_block(k[it], g)];
#pragma unroll
    for (int it = 0; it < ITEMS; it++) {
        const uint64_t i = tile + (uint64_t)it * BLOCK + threadIdx.x;
        const uint64_t m = bloom_mask(k[it], g.k);
        pass[it] = i < n && (fw[it] & m) == m;
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
        if (lane == 0)
            *s_base = total ? atomicAdd((unsigned long long*)count, (unsigned long long)total) : 0ull;
    }
    // survivors' values, all in flight before the stores
    uint32_t vv[ITEMS];
#pragma unroll
    for (int it = 0; it < ITEMS; it++)
        if (pass[it]) vv[it] = sval[tile + (uint64_t)it * BLOCK + threadIdx.x];
    __syncthreads();
    bool bad = false;
#pragma unroll
    for (int it = 0; it < ITEMS; it++) {
        if (pass[it]) {
            const uint64_t pos = *s_base + gcnt[it * NW + wid] + rank[it];
            if (pos < out_cap) {
                okey[pos] = k[it];
                oval[pos] = vv[it];
            } else {
                bad = true;
            }
        }
    }
    __syncthreads();  // gcnt / s_base are rewritten by the next tile
    return bad;
}

// ---------------------------------------------------------------- build + probe of one work item
// The build side's chunk (bn <= CAP tuples) is counting-sorted into an LDS
// table by the hash bits below the radix bits (up to two buckets per build
// tuple): bst[t] is the first slot of bucket t (bst[nbk] = bn), hkey[] holds the
// keys bucket by bucket and hidx[] their row in the chunk (8 bytes of LDS per
// build tuple).  A probe key reads its bucket's start and end, then scans a
// short contiguous run of keys: no pointer chasing.
// The probe side (pn keys) streams through in rounds of BLOCK*PITEMS keys (the
// next round is loaded while the current one is processed).  Each round counts
// its matches and block-scans the counts.  A round with at most STAGE rows
// stages them in LDS as (probe row, build row) pairs and reserves them with one
// global atomic; the next round loads their keys and values, one or two rows
// per thread, and the round after stores them after its own count (stores by
// consecutive threads).  Larger rounds reserve and write from the lanes at
// once.  Both
// values of a row are read only for matches: the build side's from its
// (cache-resident) chunk, the probe side's from its slice.  Rows past out_cap
// are counted, not written.
template <int BLOCK, int CAP>
struct alignas(16) ProbeLds {
    static constexpr int STAGE = 2 * BLOCK;  // rows of one round staged in LDS
    uint32_t stage_p[2][STAGE];  // probe-side row of each of a round's rows, two
                                 // rounds (first: short LDS offsets)
    uint16_t stage_b[2][STAGE];  // table slot of the build tuple (its key and row)
    uint32_t hkey[CAP];        // keys by bucket (bucket counts while the table is built)
    uint16_t hidx[CAP];
    uint16_t bst[CAP + 2];     // bucket starts, CAP buckets at most
    uint32_t wtmp[BLOCK / WAVE + 1];
    unsigned long long s_base[2];  // reservations of staged rounds being stored
    unsigned long long s_direct;  // reservation of a round written from the lanes
};

// Probe phase timers (one accumulator per workgroup, kept by thread 0, in
// ticks of the constant-rate wall clock): fetching work items, clearing the
// table, building it, counting the rounds' matches (through the reservation),
// writing the rounds' rows.
enum { PT_ITEM = 0, PT_CLEAR, PT_BUILD, PT_COUNT, PT_WRITE, PT_N };

__device__ __forceinline__ uint64_t phase_clock() { return wall_clock64(); }

// thread 0 of a collecting workgroup: acc[ph] += now - mark, mark = now
template <bool COLLECT>
__device__ __forceinline__ void phase_tick(uint64_t* acc, uint64_t* mark, int ph)
{
    if (COLLECT && threadIdx.x == 0) {
        const uint64_t t = phase_clock();
        acc[ph] += t - *mark;
        *mark = t;
    }
}

// 64-bit value that is the same in every lane, in scalar registers
__device__ __forceinline__ uint64_t uniform_u64(uint64_t v)
{
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
    return ((uint64_t)hi << 32) | lo;
}

// COLLECT: thread 0 adds the item's phase times to tacc[PT_CLEAR..PT_WRITE]
// (tmark: the time the item started; left at the item's end).
template <int BLOCK, int CAP, int PITEMS, bool COLLECT = false>
__device__ __forceinline__ void probe_item(ProbeLds<BLOCK, CAP>& L, const uint32_t* __restrict__ bkp,
                                           const uint32_t* __restrict__ bvp, int bn,
                                           const uint32_t* __restrict__ pkp, const uint32_t* __restrict__ pvp,
                                           uint32_t pn, bool build_r, int bits,
                                           uint64_t* __restrict__ out_count, uint64_t out_cap,
                                           uint32_t* __restrict__ o_key, uint32_t* __restrict__ o_rval,
                                           uint32_t* __restrict__ o_sval, uint64_t* tacc = nullptr,
                                           uint64_t* tmark = nullptr)
{
    constexpr int ROUND = BLOCK * PITEMS;
    constexpr int BPT = CAP / BLOCK;  // build tuples per thread
    constexpr int STAGE = ProbeLds<BLOCK, CAP>::STAGE;
    static_assert(CAP % BLOCK == 0 && CAP <= 32768, "bucket and rank pack into 16 bits each");
    if (pn == 0) return;  // an empty slice (uniform)

    // table geometry: 2^tb buckets, about two per build tuple when the probe
    // side is at least as long as the build side (shorter runs per probe key:
    // 5-7% off the probe of sub-buckets with 1-4 times as many probe as build
    // tuples, against a threshold of 8 times), else one to two build tuples per
    // bucket (less table to clear and scan)
    int tb = 0;
    if (pn >= (uint32_t)bn) {
        while ((1 << tb) < 2 * bn && (1 << tb) < CAP) tb++;
    } else {
        while ((1 << (tb + 1)) < bn && (1 << (tb + 1)) <= CAP / 2) tb++;
    }
    const int nbk = 1 << tb;
    // bucket of a key: hash bits [32 - bits - tb, 32 - bits), just below the
    // radix bits (a field of width 0 is bucket 0)
    const uint32_t toff = (uint32_t)(32 - bits - tb);
    auto bucket = [&](uint32_t key) { return __builtin_amdgcn_ubfe(radix_hash(key), toff, (uint32_t)tb); };
    uint32_t* cnt = L.hkey;  // bucket counts until the keys are placed
    if (nbk >= 4) {
        for (int i = threadIdx.x * 4; i < nbk; i += BLOCK * 4) *(uint4*)&cnt[i] = make_uint4(0u, 0u, 0u, 0u);
    } else if (threadIdx.x < nbk) {
        cnt[threadIdx.x] = 0u;
    }
    __syncthreads();
    phase_tick<COLLECT>(tacc, tmark, PT_CLEAR);
    // build: rank each tuple in its bucket, scan the counts, place;
    // bucket (< 2^15) and rank (< 2^16) of each build tuple, packed
    uint32_t bkey[BPT], btr[BPT];
#pragma unroll
    for (int u = 0; u < BPT; u++) {
        const int i = u * BLOCK + threadIdx.x;
        bkey[u] = (i < bn) ? bkp[i] : 0u;
    }
#pragma unroll
    for (int u = 0; u < BPT; u++) {
        const int i = u * BLOCK + threadIdx.x;
        if (i < bn) {
            const uint32_t t = bucket(bkey[u]);
            btr[u] = (atomicAdd(&cnt[t], 1u) << 16) | t;
        }
    }
    __syncthreads();
    {
        // exclusive scan of the counts into the bucket starts; a thread's run
        // of counts is a power of two, read and written four at a time from 4
        // counts on (16-byte reads, 8-byte writes of four 16-bit starts)
        const int per_t = (nbk + BLOCK - 1) / BLOCK;
        const int lo_t = min((int)threadIdx.x * per_t, nbk), hi_t = min(lo_t + per_t, nbk);
        uint32_t sum = 0;
        if (per_t >= 4) {
            for (int t = lo_t; t < hi_t; t += 4) {
                const uint4 q = *(const uint4*)&cnt[t];
                sum += q.x + q.y + q.z + q.w;
            }
        } else {
            for (int t = lo_t; t < hi_t; t++) sum += cnt[t];
        }
        uint32_t tot;
        uint32_t ex = block_exclusive_scan<BLOCK, uint32_t>(sum, L.wtmp, &tot);
        if (per_t >= 4) {
            for (int t = lo_t; t < hi_t; t += 4) {
                const uint4 q = *(const uint4*)&cnt[t];
                const uint32_t e1 = ex + q.x, e2 = e1 + q.y, e3 = e2 + q.z;
                *(uint2*)&L.bst[t] = make_uint2((ex & 0xFFFFu) | (e1 << 16), (e2 & 0xFFFFu) | (e3 << 16));
                ex = e3 + q.w;
            }
        } else {
            for (int t = lo_t; t < hi_t; t++) {
                L.bst[t] = (uint16_t)ex;
                ex += cnt[t];
            }
        }
        if (threadIdx.x == 0) L.bst[nbk] = (uint16_t)bn;
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < BPT; u++) {
        const int i = u * BLOCK + threadIdx.x;
        if (i < bn) {
            const uint32_t q = (uint32_t)L.bst[btr[u] & 0xFFFFu] + (btr[u] >> 16);
            L.hkey[q] = bkey[u];
            L.hidx[q] = (uint16_t)i;
        }
    }
    __syncthreads();
    phase_tick<COLLECT>(tacc, tmark, PT_BUILD);

    // key loads clamped to the last key (pn >= 1): no branches around loads
    uint32_t nk[PITEMS];
#pragma unroll
    for (int it = 0; it < PITEMS; it++) nk[it] = pkp[min((uint32_t)it * BLOCK + threadIdx.x, pn - 1u)];
    // Rows are written two rounds late, with two barriers per round (those of
    // the count scan): round r stages the (probe row, build row) pairs of its
    // matches in stage[r & 1] and reserves their places (one atomic by thread 0);
    // round r + 1 loads the key and values of rows tid and tid + BLOCK; round
    // r + 2 stores them after its count, at the reservation thread 0 published
    // in s_base[r & 1] before that round's scan.  The scan barriers of the round
    // in between separate every LDS write from the reads of the same slot.  No
    // wait for a reservation or a value load is on a round's path.
    uint32_t wk0, wp0, wb0, wk1, wp1, wb1;  // loaded rows: key, probe / build value
    uint32_t n_loaded = 0;          // rows in the registers, stored this round (uniform)
    uint32_t n_staged = 0;          // rows staged last round, loaded this round (uniform)
    unsigned long long sb_loaded = 0, sb_staged = 0;  // thread 0: their reservations
    uint32_t par = 0;               // round parity
    // The key and build row come from the table (LDS), the values from global
    // memory.  The four loads are issued in every round, also where nothing is
    // staged (row 0 of the probe side and table slot 0, unused): the loads in
    // flight at the top of a round are then the same on every path (the next
    // keys, then these four), so waiting for the keys never waits for these.
    auto load_rows = [&](uint32_t n, int buf) {
        uint32_t p0 = 0, q0s = 0, p1 = 0, q1s = 0;
        if (n) {  // uniform; rows tid and tid + BLOCK, clamped to the last row
            const uint32_t j0 = min((uint32_t)threadIdx.x, n - 1u);
            const uint32_t j1 = min((uint32_t)threadIdx.x + BLOCK, n - 1u);
            p0 = L.stage_p[buf][j0];
            q0s = L.stage_b[buf][j0];
            p1 = L.stage_p[buf][j1];
            q1s = L.stage_b[buf][j1];
        }
        wk0 = L.hkey[q0s];
        wk1 = L.hkey[q1s];
        uint32_t b0 = L.hidx[q0s], b1 = L.hidx[q1s];
        // per-lane rows: vector loads on every path (a known row 0 would become
        // a scalar load, which the vector load counter does not see)
        asm volatile("" : "+v"(p0), "+v"(b0), "+v"(p1), "+v"(b1));
        wp0 = pvp[p0];
        wp1 = pvp[p1];
        wb0 = bvp[b0];
        wb1 = bvp[b1];
    };
    auto flush = [&](uint32_t n, int slot) {  // after a barrier that published L.s_base[slot]
        const uint64_t sb = uniform_u64(L.s_base[slot]);
        const uint32_t nw = sb >= out_cap ? 0u : (uint32_t)min<uint64_t>(n, out_cap - sb);
        const uint32_t j0 = threadIdx.x, j1 = threadIdx.x + BLOCK;
        if (j0 < nw) {
            o_key[sb + j0] = wk0;
            o_rval[sb + j0] = build_r ? wb0 : wp0;
            o_sval[sb + j0] = build_r ? wp0 : wb0;
        }
        if (j1 < nw) {
            o_key[sb + j1] = wk1;
            o_rval[sb + j1] = build_r ? wb1 : wp1;
            o_sval[sb + j1] = build_r ? wp1 : wb1;
        }
    };
    load_rows(0, 0);
    for (uint32_t base = 0; base < pn; base += ROUND, par ^= 1u) {
        uint32_t kk[PITEMS], qs[PITEMS], qe[PITEMS];
#pragma unroll
        for (int it = 0; it < PITEMS; it++) kk[it] = nk[it];
        // the next round's keys, issued before this round's value loads so that
        // the next round's wait for them does not cover those; clamped, and
        // issued in the last round too, so that every round has the same loads
        // in flight (the compiler's waits then count past them)
#pragma unroll
        for (int it = 0; it < PITEMS; it++)
            nk[it] = pkp[min(base + ROUND + (uint32_t)it * BLOCK + threadIdx.x, pn - 1u)];
#pragma unroll
        for (int it = 0; it < PITEMS; it++) {
            // two 16-bit reads: one 32-bit read at a 2-byte aligned address is
            // slower on this LDS than the pair (t + 1 <= nbk)
            const uint32_t t = bucket(kk[it]);
            qs[it] = L.bst[t];
            qe[it] = L.bst[t + 1];
        }
        if (base + ROUND > pn) {  // last round: keys past pn match nothing
#pragma unroll
            for (int it = 0; it < PITEMS; it++)
                if (base + (uint32_t)it * BLOCK + threadIdx.x >= pn) qe[it] = qs[it];
        }
        // count the matches of every key and note its first matching slot q0
        // (the run is scanned downwards, so the last match seen is the first)
        uint32_t c = 0, hitmask = 0, multi = 0, q0[PITEMS];
#pragma unroll
        for (int it = 0; it < PITEMS; it++) {
            uint32_t m = 0, f = 0;
            for (uint32_t q = qe[it]; q > qs[it];) {
                q--;
                const bool e = L.hkey[q] == kk[it];
                f = e ? q : f;
                m += e;
            }
            q0[it] = f;
            c += m;
            hitmask |= (m ? 1u : 0u) << it;
            multi |= (m > 1 ? 1u : 0u) << it;
        }
        // the reservation of the rows stored this round (staged two rounds ago,
        // in slot par; its last reader was that round's store, before the last
        // round's scan barriers)
        if (threadIdx.x == 0 && n_loaded) L.s_base[par] = sb_loaded;
        uint32_t total;
        const uint32_t pre = block_exclusive_scan<BLOCK, uint32_t>(c, L.wtmp, &total);
        phase_tick<COLLECT>(tacc, tmark, PT_COUNT);
        if (n_loaded) flush(n_loaded, (int)par);
        // last round's staged rows into the registers (stage[par ^ 1] was
        // written before this round's scan barriers)
        load_rows(n_staged, (int)(par ^ 1u));
        n_loaded = n_staged;
        sb_loaded = sb_staged;
        n_staged = 0;
        if (total == 0) {  // uniform
            if (n_loaded) phase_tick<COLLECT>(tacc, tmark, PT_WRITE);
            continue;
        }
        if (total <= (uint32_t)STAGE) {  // uniform
            if (threadIdx.x == 0) sb_staged = atomicAdd((unsigned long long*)out_count, (unsigned long long)total);
            uint32_t w = pre;
#pragma unroll
            for (int it = 0; it < PITEMS; it++) {
                if (hitmask & (1u << it)) {
                    const uint32_t prow = base + (uint32_t)it * BLOCK + threadIdx.x;
                    L.stage_p[par][w] = prow;
                    L.stage_b[par][w] = (uint16_t)q0[it];
                    w++;
                    if (multi & (1u << it)) {
                        for (uint32_t q = q0[it] + 1; q < qe[it]; q++) {
                            if (L.hkey[q] == kk[it]) {
                                L.stage_p[par][w] = prow;
                                L.stage_b[par][w] = (uint16_t)q;
                                w++;
                            }
                        }
                    }
                }
            }
            n_staged = total;
            phase_tick<COLLECT>(tacc, tmark, PT_WRITE);
            continue;
        }
        // a round with more rows than the stage (duplicate-heavy keys): rows
        // written from the lanes, after the reservation
        if (threadIdx.x == 0)
            L.s_direct = atomicAdd((unsigned long long*)out_count, (unsigned long long)total);
        uint32_t pv[PITEMS], bv[PITEMS], hx[PITEMS];
#pragma unroll
        for (int it = 0; it < PITEMS; it++) hx[it] = L.hidx[q0[it]];
#pragma unroll
        for (int it = 0; it < PITEMS; it++) {
            const bool hit = (hitmask >> it) & 1u;
            pv[it] = pvp[hit ? base + (uint32_t)it * BLOCK + threadIdx.x : 0u];
            bv[it] = bvp[hx[it]];
        }
        __syncthreads();
        uint64_t w = L.s_direct + pre;
#pragma unroll
        for (int it = 0; it < PITEMS; it++) {
            if (hitmask & (1u << it)) {
                const uint32_t pval = pv[it];
                if (w < out_cap) {
                    o_key[w] = kk[it];
                    o_rval[w] = build_r ? bv[it] : pval;
                    o_sval[w] = build_r ? pval : bv[it];
                }
                w++;
                // further matches of the same key (duplicate build keys)
                if (multi & (1u << it)) {
                    for (uint32_t q = q0[it] + 1; q < qe[it]; q++) {
                        if (L.hkey[q] == kk[it]) {
                            if (w < out_cap) {
                                const uint32_t bval = bvp[L.hidx[q]];
                                o_key[w] = kk[it];
                                o_rval[w] = build_r ? bval : pval;
                                o_sval[w] = build_r ? pval : bval;
                            }
                            w++;
                        }
                    }
                }
            }
        }
        // L.s_direct is rewritten only after the next round's scan barriers
        phase_tick<COLLECT>(tacc, tmark, PT_WRITE);
    }
    // the rows still in flight: those in the registers (slot par) and those
    // staged by the last round (stage[par ^ 1], slot par ^ 1)
    if (n_loaded || n_staged) {  // uniform
        __syncthreads();  // the last round's store read slot par ^ 1 after its scan
        if (threadIdx.x == 0) {
            if (n_loaded) L.s_base[par] = sb_loaded;
            if (n_staged) L.s_base[par ^ 1u] = sb_staged;
        }
        __syncthreads();
        if (n_loaded) flush(n_loaded, (int)par);
        if (n_staged) {
            load_rows(n_staged, (int)(par ^ 1u));
            flush(n_staged, (int)(par ^ 1u));
        }
        phase_tick<COLLECT>(tacc, tmark, PT_WRITE);
    }
    // the caller synchronises before the next item rebuilds the table
}

// Build side of a sub-bucket: R whenever it fits one LDS table (the
// reference's build side), else the smaller side.  Equal S keys, as under a
// skewed S, then stream past distinct R keys instead of piling up in one run
// of the table that every matching probe walks.
__host__ __device__ __forceinline__ bool build_from_r(uint64_t nr, uint64_t ns, uint32_t tile)
{
    return nr <= tile || nr <= ns;
}

// Work-item split of one sub-bucket (vlsplit).  A sub-bucket whose build side
// (build_from_r) fits one LDS tile and whose probe side fits one slice is
// "regular": one work item, indexed by the sub-bucket itself.  Larger ones are
// split: the build side into ceil(nb / tile) chunks, the probe side into
// min(ceil(np / slice), max_slices) slices, one explicit item per (chunk, slice).
__device__ __forceinline__ uint32_t plan_items(uint64_t nr, uint64_t ns, uint32_t tile, uint32_t slice,
                                               uint32_t max_slices, uint32_t& chunks, uint32_t& slices)
{
    chunks = slices = 0;
    if (nr == 0 || ns == 0) return 0;
    const bool br = build_from_r(nr, ns, tile);
    const uint64_t nb = br ? nr : ns, np = br ? ns : nr;
    if (nb <= tile && np <= slice) return 0;
    chunks = (uint32_t)((nb + tile - 1) / tile);
    slices = (uint32_t)min<uint64_t>((np + slice - 1) / slice, max_slices);
    return chunks * slices;
}

}  // namespace bj

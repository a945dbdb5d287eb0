// bj_kernels.h -- launchers of the gfx950 kernels (host side).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "bj_device.h"

namespace bj {

// One explicit vlsplit work item of an oversized sub-bucket.
struct HeavyItem {
    uint32_t sb;       // sub-bucket
    uint32_t chunk;    // build-side chunk (tile tuples each)
    uint32_t slice;    // probe-side slice index
    uint32_t nslices;  // number of probe-side slices of this sub-bucket
};

hipError_t kernels_init(int max_lds_bytes);

hipError_t launch_zero_u64(uint64_t* p, int n, hipStream_t st);
// acc[0] += *a, acc[1] += *b (one thread): per-chunk counts summed on the device.
hipError_t launch_add2_u64(const uint64_t* a, const uint64_t* b, uint64_t* acc, hipStream_t st);
hipError_t launch_gen_r(uint32_t* key, uint32_t* val, uint64_t n, const PermKeys& K, hipStream_t st);
hipError_t launch_gen_s(uint32_t* key, uint32_t* val, uint64_t n_s, uint64_t n_r, uint64_t m_s,
                        int L, const PermKeys& KR, const PermKeys& KP, const PermKeys& KN,
                        uint32_t useed, double zipf_s, uint32_t zseed, double* d_tmp,
                        hipStream_t st);

// The reference datagen's distribution (datagen/gen_dataset.cpp:172-220) in a
// keyed random order: R of n keys with k matching ones; one region of S (n keys,
// m matching keys from mc0 on, S-only keys from offset sj0 of their sequence).
hipError_t launch_gen_refdist_r(uint32_t* key, uint64_t n, uint64_t k, const PermKeys& K, hipStream_t st);
hipError_t launch_gen_refdist_s(uint32_t* key, uint64_t n, uint64_t m, uint64_t mc0, uint64_t sj0, uint64_t n_r,
                                const PermKeys& K, hipStream_t st);

// Also zeroes totals (2^bits entries) for launch_offsets when not null, and
// with bits2 > 0 writes each chunk's first-level counts l1h[g][b1] (the sums
// over the low bits2 bits).
hipError_t launch_hist(const uint32_t* key, uint64_t n_host, const uint64_t* n_dev, int G, int bits,
                       uint32_t* part, uint32_t* totals, int bits2, uint32_t* l1h, hipStream_t st);
hipError_t launch_zero_u32(uint32_t* p, uint64_t n, hipStream_t st);
// Sub-bucket starts, pass-1 cursors l1off[g][b1] and per-group column sums
// colsum[w][b] (W groups of chunks) from the per-chunk histograms and, when
// bits2 > 0, the first-level counts l1h (both from launch_hist; bits <= 15);
// totals (zeroed by launch_hist) is scratch.
hipError_t launch_offsets(const uint32_t* part, int G, int bits, int bits2, int W, uint32_t* colsum,
                          uint32_t* totals, uint32_t* l1h, uint64_t* sb_start, uint64_t* l1off, hipStream_t st);
size_t scan_partials_needed(uint64_t n);
hipError_t launch_scan(const uint32_t* in, uint64_t n, uint64_t* out, uint64_t* partial,
                       hipStream_t st);
int scatter_carry(int nbk);
size_t scatter_lds_bytes(int nbk);
// LDS of a 1024-thread, 8192-tuple tile with the carries of scatter_carry_big
size_t scatter_lds_bytes_big(int nbk);
// pass 1 (mode 0): p1-bit buckets; pass 2 (mode 1): the remaining bits, W2
// workgroups per pass-1 bucket, cursors from sb_start and colsum.  The launch reserves at
// least min_lds bytes of LDS per workgroup (a way to cap workgroups per CU).
hipError_t launch_scatter(const uint32_t* ikey, const uint32_t* ival, uint32_t* okey, uint32_t* oval,
                          uint64_t out_cap, uint64_t n_host, const uint64_t* n_dev, int G, int mode,
                          int bits, int p1, int W2, const uint64_t* l1off, const uint64_t* sb_start,
                          const uint32_t* colsum, uint64_t* err, int wg, size_t min_lds, hipStream_t st);
hipError_t launch_bloom_build(const uint32_t* rkey, const uint64_t* r_start, int bits,
                              const BloomGeom& g, uint64_t* filter, hipStream_t st);
hipError_t launch_bloom_prefilter(const uint32_t* skey, const uint32_t* sval, uint64_t n,
                                  const uint64_t* filter, const BloomGeom& g, uint32_t* okey,
                                  uint32_t* oval, uint64_t out_cap, uint64_t* count, uint64_t* err,
                                  int num_cu, hipStream_t st);
// tile_ctr: 8 zeroed counters, one per XCD (the tile order of each XCD)
hipError_t launch_filter_l1(const uint32_t* ikey, const uint32_t* ival, const uint64_t* sb_start, int bits,
                            int bits2, const uint64_t* filter, const BloomGeom& g, uint32_t* okey,
                            uint32_t* oval, uint64_t out_cap, uint32_t* s_fill, uint64_t* count,
                            uint64_t* tile_ctr, uint64_t* err, int num_cu, hipStream_t st);
// S in chunks, one probe after the last chunk: each chunk's S' stays in its own
// region of a pool (starting at base); launch_pool_note records its pieces
// (sub-bucket sb: s_cnt[sb] tuples when not null, else the region's length, at
// base + s_start[sb]) in this chunk's row cnt / off of the piece table (nbins
// entries each) and adds them to *total.  launch_pool_start sums the table over
// nch chunks into totals and scans them into start (nbins + 1 entries);
// launch_pool_merge gathers every sub-bucket's pieces into
// [start[sb], start[sb + 1]) of ok / ov (err |= 16 past out_cap).
hipError_t launch_pool_note(const uint64_t* s_start, const uint32_t* s_cnt, int nbins, uint64_t base, uint32_t* cnt,
                            uint64_t* off, uint64_t* total, hipStream_t st);
hipError_t launch_pool_start(const uint32_t* cnt, int nch, int nbins, uint32_t* totals, uint64_t* start,
                             uint64_t* partials, hipStream_t st);
hipError_t launch_pool_merge(const uint32_t* pk, const uint32_t* pv, const uint32_t* cnt, const uint64_t* off,
                             int nch, int nbins, const uint64_t* start, uint32_t* ok, uint32_t* ov,
                             uint64_t out_cap, uint64_t* err, hipStream_t st);
hipError_t launch_plan(const uint64_t* r_start, const uint64_t* s_start, const uint32_t* s_cnt, int bits,
                       uint32_t tile, uint32_t slice, uint32_t max_slices, uint32_t* counts,
                       uint64_t* item_start, uint64_t* total_out, HeavyItem* items, uint64_t item_cap,
                       uint64_t* err, hipStream_t st);  // total_out: item count (nullable)
hipError_t launch_probe(int tile, const uint32_t* rk, const uint32_t* rv, const uint64_t* r_start,
                        const uint32_t* sk, const uint32_t* sv, const uint64_t* s_start,
                        const uint32_t* s_cnt, int bits, uint32_t slice, const HeavyItem* items,
                        const uint64_t* n_items, uint64_t item_cap, uint64_t* work_ctr, uint64_t* out_count,
                        uint64_t out_cap, uint32_t* o_key, uint32_t* o_rval, uint32_t* o_sval, int num_cu,
                        uint64_t* timers, hipStream_t st);
// Workgroups of a probe launch (persistent: a fixed number per CU).  With
// timers, every workgroup adds its phase times (ticks, PT_* order of bj_tiles.h,
// PT_N = 5 per workgroup) to its slots; the caller clears them.
int probe_grid(int tile, int num_cu);
constexpr int PROBE_TIMER_SLOTS = 5;

}  // namespace bj

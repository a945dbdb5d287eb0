// bj_batch.h -- device descriptors and launchers of the batched join (host side).
//
// A batch is many independent joins of uneven sizes.  Every phase is one
// launch over all joins: the workgroups of a launch are dealt out as tasks
// (a histogram/scatter chunk, a filter segment, a column block) and each task
// finds its join by a binary search over a per-join prefix of task counts.
// Every join is partitioned with its own radix bits into its region of shared
// arenas, in one scatter pass (up to 9 bits) or two (up to 15: the first by the
// top bits into a scratch arena at the final positions, the second within each
// first-level bucket); the probe is a persistent launch over the sub-buckets of
// all joins.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "bj_device.h"
#include "bj_kernels.h"

namespace bj {

// One join of a batch, as the kernels see it.
struct BJoin {
    const uint32_t* rk;       // R keys / values (device)
    const uint32_t* rv;
    const uint32_t* sk;       // S keys / values (device)
    const uint32_t* sv;
    uint32_t nr, ns;          // sizes (< 2^31)
    int bits;                 // radix bits (1..15)
    int b1;                   // bits of the first scatter pass (== bits: one pass)
    int bf;                   // Bloom filter on
    int bf2;                  // ... and probed after a first pass of S by the top fb bits,
    int fb;                   //     so that each bucket's filter slice stays in the L2
    BloomGeom geom;           // filter geometry (segments = top seg_bits radix bits)
    uint32_t sb_base;         // first global sub-bucket
    uint32_t r_ch, s_ch;      // chunk length of R / S histogram and scatter tasks
    uint32_t tile;            // LDS table capacity of the probe: 4096 or 8192
    uint32_t w2_r, w2_s;      // second-pass workgroups per first-level bucket (R / S)
    uint64_t r_base, s_base;  // region offsets in the partitioned R / S arenas
    uint64_t sp_base;         // region offset in the S' staging arena (Bloom joins)
    uint64_t part_r, part_s;  // per-chunk histogram offsets in the part arena
    uint64_t f_base;          // filter offset in 64-bit blocks
    uint64_t out_base, out_cap;
};

// Per-join device counters.
enum { JC_SPRIME = 0, JC_OUT = 1, JC_N = 2 };

// Prefix arrays (n + 1 entries each, u32) that map a task index to its join.
// PX_P1 + 2 * side + cls: the first-pass scatter tasks of one side (0 R, 1 S,
// 2 the filter pass of S) over the joins of one class (0: first passes of fewer
// than 2^WIDE_B1 buckets, 1: the others), each class in a launch of its own.
enum { PX_RTASK = 0, PX_STASK, PX_SEG, PX_RCOL, PX_SCOL, PX_SB, PX_R2, PX_S2, PX_FTILE, PX_P1, PX_N = PX_P1 + 6 };
constexpr int WIDE_B1 = 7;

// Keys per tile of the S -> S' filter launch (PX_FTILE counts these tiles).
constexpr int PF_BLOCK = 512, PF_ITEMS = 8, PF_TILE = PF_BLOCK * PF_ITEMS;
constexpr int NXCD = 8;  // XCDs of the chip: workgroup i of a launch runs on XCD i % 8

struct BatchDev {
    const BJoin* J;           // [n]
    const uint32_t* px;       // [PX_N][n + 1]
    const uint32_t* sbj;      // [number of sub-buckets] the join of each sub-bucket
    int n;
    int max_bits;             // largest radix bits of the batch
    int max_b2;               // largest bits of the second scatter pass
};

hipError_t batch_kernels_init(int max_lds_bytes);

// side 0 = R, side 1 = S.  For S of a Bloom join the source is the S' staging
// arena and its length the device counter jc[b][JC_SPRIME].  side 2 = the first
// pass of S of the joins that filter after it (bf2): their S by its top fb bits
// into the scratch arena (other joins: nothing), before the filter.
hipError_t launch_batch_hist(const BatchDev& B, int side, uint32_t ntasks, const uint32_t* sp_key,
                             const uint64_t* jc, uint32_t* part, hipStream_t st);
hipError_t launch_batch_colscan(const BatchDev& B, int side, uint32_t ncolblk, const uint32_t* part,
                                uint32_t* cur_rel, uint32_t* tot, hipStream_t st);
hipError_t launch_batch_binscan(const BatchDev& B, int side, const uint32_t* tot, uint64_t* start,
                                hipStream_t st);
// First pass: every join's chunks by its first-pass bits, to the final arena
// (one-pass joins) or to the scratch arena tkey/tval at the final positions.
// One launch per class cls of first passes (ntasks tasks, PX_P1 + 2 * side + cls),
// with LDS for max_b1 bits.  big: one workgroup per CU with 8192-tuple tiles and
// whole-line carries (the class of wide first passes), else two per CU with
// 4096-tuple tiles.
hipError_t launch_batch_scatter(const BatchDev& B, int side, int cls, uint32_t ntasks, int max_b1,
                                const uint32_t* sp_key, const uint32_t* sp_val, const uint64_t* jc,
                                const uint32_t* cur_rel, const uint64_t* start, uint32_t* okey, uint32_t* oval,
                                uint64_t out_cap, uint32_t* tkey, uint32_t* tval, uint64_t tmp_cap, uint64_t* err,
                                bool big, hipStream_t st);
// Second pass of two-pass joins: scratch arena -> final arena.  big: one
// workgroup per CU with 8192-tuple tiles (for launches over many tuples).
hipError_t launch_batch_scatter2(const BatchDev& B, int side, uint32_t ntasks, const uint32_t* cur_rel,
                                 const uint64_t* start, const uint32_t* cnt, const uint32_t* tkey,
                                 const uint32_t* tval, uint32_t* okey, uint32_t* oval, uint64_t out_cap,
                                 uint64_t* err, bool big, hipStream_t st);
hipError_t launch_batch_bloom_build(const BatchDev& B, uint32_t nseg, uint32_t max_wps, const uint32_t* rkey,
                                    const uint64_t* r_start, const uint32_t* r_cnt, uint64_t* filter,
                                    hipStream_t st);
// S -> S' of every Bloom join: ntiles filter tiles (PX_FTILE), persistent,
// XCD x working through the x-th eighth of them on its own counter xctr[x]
// (NXCD counters, zeroed).
hipError_t launch_batch_prefilter(const BatchDev& B, uint32_t ntiles, const uint64_t* filter,
                                  const uint32_t* tkey, const uint32_t* tval, uint32_t* sp_key,
                                  uint32_t* sp_val, uint64_t* jc, uint64_t* xctr, uint64_t* err, int num_cu,
                                  hipStream_t st);
hipError_t launch_batch_plan(const BatchDev& B, uint32_t nsb, const uint64_t* r_start, const uint32_t* r_cnt,
                             const uint64_t* s_start, const uint32_t* s_cnt, uint32_t slice,
                             uint32_t max_slices, uint32_t* counts, uint64_t* item_start, uint64_t* partials,
                             HeavyItem* items, uint64_t item_cap, uint64_t* err, hipStream_t st);
// Probe of the sub-buckets [sb_lo, sb_hi) (all of one tile class) and their
// vlsplit items item_start[sb_lo] .. item_start[sb_hi].
hipError_t launch_batch_probe(const BatchDev& B, int tile, uint32_t sb_lo, uint32_t sb_hi,
                              const uint32_t* rk, const uint32_t* rv, const uint64_t* r_start,
                              const uint32_t* r_cnt, const uint32_t* sk, const uint32_t* sv,
                              const uint64_t* s_start, const uint32_t* s_cnt, uint32_t slice,
                              const HeavyItem* items, const uint64_t* item_start, uint64_t item_cap,
                              uint64_t* work_ctr, uint64_t* jc, uint32_t* o_key, uint32_t* o_rval,
                              uint32_t* o_sval, int num_cu, hipStream_t st);

}  // namespace bj

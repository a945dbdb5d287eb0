/*
 * bj_api.h -- C ABI of the MI355X (gfx950) radix hash join.
 *
 * Plain C types only (no HIP, no torch, no C++ in the signatures).  Every entry
 * point returns BJ_OK (0) or a negative BJ_ERR_* code and never exits the
 * process.  A context is single-threaded; independent contexts may be used from
 * different host threads.  The library owns all device memory of a context; the
 * caller owns its inputs.
 *
 * What each entry point replaces in the reference (Chanho-LEE-SNU/BRIG-Join):
 *   bj_gpu_main      gpu_main(...)                 src/hash_kernel.h:21-34,
 *                                                  src/hash_kernel.cu:103-1654
 *                    (its one-GPU share: read inputs, derive parameters, join,
 *                    count; plus the output.bin that the reference parses but
 *                    never writes, src/driver.cpp:50,119)
 *   bj_join          the device phase of gpu_main plus join_execute_host(ctx,
 *                    s, s_val, n) (src/join_executor.h:87-88, local branch of
 *                    src/join_executor.cu:14-110, 351-353) and the partition
 *                    drivers partial_jointable_gpu_partition*(...)
 *                    (src/vlsplit_hash_kernel.h:15-60)
 *   bj_cpu_sanity_count  run_sanity_check(...)     src/sanity_check.h:5-6
 *   bj_gen_build / bj_gen_probe   the dataset generator gen_dataset main()
 *                    (datagen/gen_dataset.cpp:79-263), with the synthetic
 *                    distribution of the benchmark workloads, on the device
 *   bj_gen_refdist_host / bj_gen_refdist   the same generator with the
 *                    reference's own distribution (exact files on the host, the
 *                    same key multiset on the device)
 *   bj_write_output / bj_read_output  output.bin (format defined below; the
 *                    reference defines none)
 */
#ifndef BJ_API_H
#define BJ_API_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BJ_API_VERSION 4

/* ---- error codes ------------------------------------------------------- */
#define BJ_OK              0
#define BJ_ERR_INVALID    -1  /* invalid argument or unsupported mode         */
#define BJ_ERR_HIP        -2  /* HIP runtime error                            */
#define BJ_ERR_NOMEM      -3  /* device or host allocation failed             */
#define BJ_ERR_IO         -4  /* file open / read / write failed or truncated */
#define BJ_ERR_SIZE       -5  /* a size exceeds a supported limit             */
#define BJ_ERR_NODEVICE   -6  /* no HIP device / device index out of range    */
#define BJ_ERR_INTERNAL   -7  /* a device-side consistency check failed       */

/* ---- per-phase timing slots (milliseconds, hipEvent on the join stream) - */
enum {
    BJ_PH_H2D = 0,       /* host->device copy of host inputs (0 for device inputs) */
    BJ_PH_R_PART,        /* two-level radix partition of R                          */
    BJ_PH_BF_BUILD,      /* blocked Bloom filter build from partitioned R           */
    BJ_PH_S_FILTER,      /* Bloom query + compaction S -> S'                        */
    BJ_PH_S_PART,        /* two-level radix partition of S (or S')                  */
    BJ_PH_PROBE,         /* LDS build + probe + row materialisation                 */
    BJ_PH_TOTAL,         /* first hot-path kernel to last (excludes H2D)            */
    BJ_NUM_PHASES
};

/* ---- phases inside the probe kernel (bj_params.probe_breakdown; the
 * reference's COLLECT timers, src/kernel_functions.cu:300-302, 442-451, turned
 * into times as src/join_executor.cu:56-98 does: each phase's share of all
 * workgroups' clock ticks times the probe kernel's time) ------------------ */
enum {
    BJ_PP_ITEMS = 0,     /* fetching work items and their sizes (preamble)           */
    BJ_PP_CLEAR,         /* clearing the LDS table's bucket counts (memset)          */
    BJ_PP_BUILD,         /* ranking, scanning and placing the build side (build)     */
    BJ_PP_COUNT,         /* probe rounds: buckets, match counts, scan, reservation   */
    BJ_PP_WRITE,         /* probe rounds: value loads and row writes                 */
    BJ_NUM_PROBE_PHASES
};

/* ---- per-kernel timing slots (milliseconds, hipEvent on the join stream) */
enum {
    BJ_K_HIST_R = 0,     /* k_hist over R (joint radix histogram)                 */
    BJ_K_SCAN_R,         /* histogram reduction + offset scans of R               */
    BJ_K_SCATTER1_R,     /* first-level scatter of R                              */
    BJ_K_SCATTER2_R,     /* second-level scatter of R                             */
    BJ_K_BF_BUILD,       /* Bloom build                                           */
    BJ_K_BF_FILTER,      /* Bloom query + compaction of S                         */
    BJ_K_HIST_S,
    BJ_K_SCAN_S,
    BJ_K_SCATTER1_S,
    BJ_K_SCATTER2_S,
    BJ_K_PLAN,           /* vlsplit work planning                                 */
    BJ_K_PROBE,          /* LDS build + probe + rows                              */
    BJ_NUM_KERNELS
};

/* ---- join parameters: the reference's argument surface ---------------- */
typedef struct bj_params {
    int64_t r_size;                 /* argv[4]  R (tuples)                            */
    int64_t s_size;                 /* argv[5]  S (tuples)                            */
    int32_t partition_size;         /* --partition_size: first-level fan-out P1 (power
                                       of two, 2..2048); -1 = auto                     */
    int32_t second_partition_size;  /* --second_partition_size: second-level fan-out
                                       P2 (power of two, 1..2048); -1 = auto           */
    int32_t radix_bits;             /* total radix bits log2(P1*P2) (1..15); -1 = from
                                       the two fan-outs or auto (the fewest bits with
                                       min(R,S)/2^bits <= 0.96 * tile)                 */
    int32_t use_bf;                 /* argv[8]: 0 = Bloom off; any supported nonzero
                                       mode = Bloom on (one GPU: the filter built from
                                       R is already complete, so every reference mode
                                       collapses to "on")                              */
    double  bf_bits_per_key;        /* > 0: filter bits per R key; <= 0: derive from
                                       true_false_ratio with the reference formula
                                       (a finite number when the filter is on)         */
    double  true_false_ratio;       /* argv[11] target false-positive rate p          */
    int64_t num_bits;               /* --num_bits override (filter bits), -1 = none   */
    int32_t num_hashes;             /* --num_hashes override (k, 1..8), -1 = none     */
    double  answer_scale;           /* --answer_scale: row capacity =
                                       max(R,S) * max(1, answer_scale); a finite
                                       number, capacities above 2^40 rows refused     */
    int32_t tile;                   /* LDS build tile in tuples (4096 or 8192); 0=auto */
    int32_t inputs_on_device;       /* 1: r_key.. are device pointers; 0: host         */
    int32_t collect_phase_times;    /* 1: fill bj_result.phase_ms                      */
    int32_t s_chunks;               /* argv[12]: S is joined in this many chunks (the
                                       reference's s_chunks loop, hash_kernel.cu:910-931):
                                       R is partitioned and filtered once, each chunk
                                       of S is filtered, partitioned and probed, rows
                                       appended; host chunks are copied on a second
                                       stream while the previous chunk is joined.
                                       1 = one pass; 0 = the fewest chunks whose
                                       workspace fits 90% of the device memory        */
    int32_t bf_stage;               /* where the filter drops S tuples: 1 = before S is
                                       partitioned (the reference's order); 2 = after
                                       the first-level pass, one filter slice per
                                       bucket resident in an XCD's L2; 0 = auto (1 when
                                       the whole filter fits an XCD's L2, 3.5 MB; else
                                       2: a filter probed at random from the MALL or
                                       HBM is several times slower)                    */
    int32_t pass1_bits;             /* radix bits of the first physical scatter pass
                                       (the rest go to the second); -1 = auto (half).
                                       The resulting sub-bucket layout does not depend
                                       on it; partition_size is the logical first
                                       level, whose buckets are unions of sub-buckets  */
    int32_t s_pass1_bits;           /* the same for S; -1 = auto (with bf_stage 2: enough
                                       bits for each first-pass bucket's filter slice
                                       to fit an XCD's L2, else pass1_bits)            */
    int32_t scatter_wg;             /* scatter workgroup size, 512 or 1024; 0 = auto   */
    int32_t probe_breakdown;        /* --fused_probe_breakdown: 1 = time the probe's
                                       phases (bj_result.probe_phase_ms; implies
                                       collect_phase_times)                          */
    int32_t chunk_probe;            /* S in chunks: 0 = one build + probe after the
                                       last chunk on every chunk's S' (pooled in device
                                       memory; R streams past the tables once) when a
                                       pool the size of S fits 90% of the free device
                                       memory, else one per chunk; 1 = one per chunk */
} bj_params;

/* ---- join result ------------------------------------------------------- */
typedef struct bj_result {
    uint64_t match_count;     /* exact number of (r, s) pairs with equal keys        */
    uint64_t rows_written;    /* min(match_count, capacity)                          */
    uint64_t capacity;        /* row capacity of the output buffers                  */
    int32_t  truncated;       /* 1 if match_count > capacity (count stays exact)     */
    uint64_t s_survivors;     /* |S'| after the Bloom filter (= S when it is off)    */
    /* Device pointers to the rows (SoA), owned by the context; valid until the
     * next bj_join on the same context or bj_destroy.  Row order is unspecified. */
    uint32_t *key, *r_val, *s_val;
    double   phase_ms[BJ_NUM_PHASES];
    double   kernel_ms[BJ_NUM_KERNELS];
    /* plan actually used */
    int32_t  bits1, bits2;    /* first / second level radix bits                     */
    int32_t  bf_seg_bits;     /* filter segments = 2^bf_seg_bits                     */
    int32_t  bf_k;            /* hash bits per key                                   */
    uint64_t bf_words;        /* filter size in 32-bit words (0 = Bloom off)         */
    int32_t  tile;            /* LDS build tile (tuples)                             */
    uint64_t heavy_items;     /* vlsplit work items for oversized sub-buckets        */
    int32_t  bf_stage;        /* 1 or 2 as in bj_params (0 = Bloom off)              */
    int32_t  pass1_bits;      /* radix bits of the first physical scatter pass       */
    int32_t  s_pass1_bits;    /* the same for S                                      */
    int32_t  s_chunks;        /* chunks S was joined in (1 = one pass)               */
    double   probe_phase_ms[BJ_NUM_PROBE_PHASES];  /* with probe_breakdown, else 0   */
} bj_result;

/* ---- the plan a join would use (no device needed) ---------------------- */
typedef struct bj_plan_info {
    int32_t  bits1, bits2;       /* first / second level radix bits            */
    int32_t  bf;                 /* Bloom on                                   */
    int32_t  bf_seg_bits, bf_k;  /* filter segments = 2^seg_bits, bits per key */
    uint64_t bf_words;           /* filter size in 32-bit words                */
    int32_t  tile;               /* LDS build tile (tuples)                    */
    uint64_t capacity;           /* output row capacity                        */
    uint64_t workspace_bytes;    /* device memory a context needs for the join,
                                    excluding device-resident inputs           */
    int32_t  bf_stage;           /* 1 or 2 (0 = Bloom off)                     */
} bj_plan_info;

typedef struct bj_ctx bj_ctx;

/* ---- lifecycle --------------------------------------------------------- */
int  bj_create(bj_ctx **ctx, int device);
int  bj_destroy(bj_ctx *ctx);
/* Run on an existing HIP stream (hipStream_t passed as void*); NULL = the
 * context's own stream.  The stream must have the default priority. */
int  bj_set_stream(bj_ctx *ctx, void *hip_stream);
void bj_params_default(bj_params *p);
const char *bj_strerror(int code);
/* Detail of the last error raised on the calling host thread. */
const char *bj_last_error(void);
int  bj_version(void);

/* Derive the plan (radix bits, filter geometry, capacity, device footprint)
 * that bj_join would use for these parameters; validates them like bj_join.
 * Counterpart of gpu_main's parameter derivation, src/hash_kernel.cu:174-284. */
int  bj_plan(const bj_params *p, bj_plan_info *out);

/* ---- the join ---------------------------------------------------------- */
/* Equi-join R(key, r_val) with S(key, s_val) on key.  Emits every matching
 * (key, r_val, s_val) row.  Sizes come from p->r_size / p->s_size.  Returns
 * after the join has finished; an error return also waits for the work
 * already queued, so the input buffers are free to reuse either way. */
int  bj_join(bj_ctx *ctx, const bj_params *p,
             const uint32_t *r_key, const uint32_t *r_val,
             const uint32_t *s_key, const uint32_t *s_val,
             bj_result *out);

/* Size of the device scratch array behind bj_join_noalloc (256 MiB), held on
 * every GPU the library runs on until the process ends. */
#define BJ_NOALLOC_SCRATCH_BYTES (256ull << 20)

/* No counterpart in the reference.  The join of bj_join (same kernels, same
 * count and row multiset) for callers that cannot allocate device memory during
 * a call -- code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc: it calls no hipMalloc / hipFree and creates no events.
 * Inputs must be device pointers (p->inputs_on_device = 1) joined in one pass
 * (p->s_chunks = 1), without probe phase timers.  The workspace is carved from a
 * static device array of BJ_NOALLOC_SCRATCH_BYTES per GPU; a join whose
 * workspace (bj_plan's workspace_bytes minus the row buffers) exceeds it is
 * refused with BJ_ERR_SIZE, which at 256 MiB allows about 8-11M tuples per
 * relation.  Rows go to the caller's device arrays out_key / out_rval / out_sval
 * (out_cap rows each; the count stays exact past it, as in bj_join) and
 * out->key / r_val / s_val point there.  Every call on one GPU shares that GPU's
 * array, so calls that overlap on one GPU race: serialise them. */
int  bj_join_noalloc(bj_ctx *ctx, const bj_params *p,
                     const uint32_t *r_key, const uint32_t *r_val,
                     const uint32_t *s_key, const uint32_t *s_val,
                     uint32_t *out_key, uint32_t *out_rval, uint32_t *out_sval,
                     uint64_t out_cap, bj_result *out);

/* Two-level radix partition of one relation (device pointers), with the radix
 * bits bj_join would use for p (r_size = n).  Writes the partitioned tuples to
 * out_key/out_val (n entries each) and the sub-bucket starts to out_start
 * (2^(bits1+bits2) + 1 entries; bucket b of a key = the top bits of
 * key * 0x9E3779B1 mod 2^32).  Counterpart of partial_jointable_gpu_partition_r_only
 * (src/vlsplit_hash_kernel.h:40-46). */
int  bj_partition(bj_ctx *ctx, const bj_params *p, const uint32_t *key,
                  const uint32_t *val, int64_t n, uint32_t *out_key,
                  uint32_t *out_val, uint64_t *out_start);

/* Copy the first n rows of a result to host arrays (n <= rows_written). */
int  bj_copy_rows_to_host(bj_ctx *ctx, const bj_result *res,
                          uint32_t *key, uint32_t *r_val, uint32_t *s_val,
                          uint64_t n);

/* ---- device-side synthetic generator ------------------------------------
 * R: key_i distinct in [1, 2^31) (a keyed 31-bit permutation of i), val_i = i.
 * S: exactly m_s positions match (chosen by a keyed permutation of positions);
 *    a matching key is R's key at a row drawn uniformly (zipf_s == 0) or with a
 *    Zipf(zipf_s) law over R's rows (zipf_s > 0); every other key is distinct in
 *    [2^31, 2^32).  val_j = j.  Pointers are device pointers. */
int  bj_gen_build(bj_ctx *ctx, uint64_t seed, uint64_t n_r,
                  uint32_t *d_key, uint32_t *d_val);
int  bj_gen_probe(bj_ctx *ctx, uint64_t seed_r, uint64_t seed_s, uint64_t n_r,
                  uint64_t n_s, uint64_t m_s, double zipf_s,
                  uint32_t *d_key, uint32_t *d_val);

/* ---- the reference datagen's distribution ---------------------------------
 * The keys of datagen/gen_dataset.cpp (main, :79-263): k = floor(R * pct / 100)
 * (at most S) matching keys [0, k), each once in R and once in S; R's other keys
 * k, k+1, ... (below 2^31 - 1); S's other keys R, R+1, ... (below 2^31 - 1).
 * max_pct >= 0 is the reference's --max load imbalance (the first quarter of S
 * holds matches at max_pct percent); < 0 = off.  *matches (may be NULL) = k, the
 * join size.
 * bj_gen_refdist_host: the reference's order exactly (Fisher-Yates over
 *   std::mt19937_64 with the reference's seeds, gen_dataset.cpp:67-77, 192,
 *   210-219), so raw files written from these arrays are byte-identical to the
 *   reference generator's for the same arguments.  Host arrays; sequential.
 * bj_gen_refdist: the same key multiset per region in a keyed random order,
 *   generated on the device (device pointers). */
int  bj_gen_refdist_host(int64_t r_size, int64_t s_size, double answer_pct, uint64_t seed,
                         double max_pct, uint32_t *r_key, uint32_t *s_key, int64_t *matches);
int  bj_gen_refdist(bj_ctx *ctx, int64_t r_size, int64_t s_size, double answer_pct, uint64_t seed,
                    double max_pct, uint32_t *d_r_key, uint32_t *d_s_key, int64_t *matches);

/* ---- output.bin ----------------------------------------------------------
 * Little-endian: char magic[8] = "BJOUT001", uint64 count, then uint32
 * key[count], uint32 r_val[count], uint32 s_val[count] (SoA, mirroring the
 * reference's GPU_answer / GPU_answer_val / GPU_answer_s_val,
 * src/hash_kernel.cu:475-477).  bj_write_output sorts the rows by
 * (key, r_val, s_val) first (in place), so two correct runs give byte-identical
 * files. */
int  bj_write_output(const char *path, uint64_t n,
                     uint32_t *key, uint32_t *r_val, uint32_t *s_val);
/* Reads a file written by bj_write_output; the arrays are malloc'ed and must be
 * released with bj_free_host. */
int  bj_read_output(const char *path, uint64_t *n,
                    uint32_t **key, uint32_t **r_val, uint32_t **s_val);
void bj_free_host(void *p);

/* ---- the CPU join the reference prints beside the GPU count -------------
 * Count of matching pairs by the reference's CPU algorithm (open addressing,
 * capacity 2^ceil(log2 2R), multiplicative hash, linear probing).  Reported
 * beside the GPU count by bj_gpu_main when sanity_check is set. */
long long bj_cpu_sanity_count(const uint32_t *r_key, int64_t r_size,
                              const uint32_t *s_key, int64_t s_size);

/* ---- gpu_main, one-GPU share ---------------------------------------------
 * Mirrors gpu_main's parameters (src/hash_kernel.h:21-34) and the driver's
 * positional + flag surface (src/driver.cpp:43-198).  Multi-GPU-only knobs
 * (use_nvlink, use_warp_compact, kernel_fusion, redist_skew_threshold,
 * partition_rs_split, r_shuffle_enable, r_replicate_ratio) are accepted and
 * have no effect on one GPU. */
typedef struct bj_main_args {
    const char *r_path;            /* argv[1] raw int32 keys (+ value block)     */
    const char *s_path;            /* argv[2]                                    */
    const char *out_path;          /* argv[3] output.bin (NULL or "" = none)     */
    int64_t r_size, s_size;        /* argv[4], argv[5]                           */
    int32_t sanity_check;          /* argv[6]                                    */
    int32_t chunk_size;            /* argv[7] (accepted)                         */
    int32_t use_bf;                /* argv[8]                                    */
    int32_t use_nvlink;            /* argv[9]                                    */
    int32_t use_warp_compact;      /* argv[10]                                   */
    double  true_false_ratio;      /* argv[11]                                   */
    int32_t s_chunks;              /* argv[12]                                   */
    int32_t kernel_fusion;         /* argv[13]                                   */
    double  redist_skew_threshold; /* argv[14]                                   */
    int64_t override_num_bits;     /* --num_bits                                 */
    int32_t override_num_hashes;   /* --num_hashes                               */
    int32_t override_partition_rs_split;    /* --partition_rs_split            */
    int32_t override_partition_size;        /* --partition_size                */
    int32_t override_second_partition_size; /* --second_partition_size         */
    int32_t fused_probe_breakdown; /* --fused_probe_breakdown: print phase times */
    int32_t read_val;              /* --read_val: value block follows the keys   */
    int32_t r_shuffle_enable;      /* --r_shuffle_enable                         */
    double  r_replicate_ratio;     /* --r_replicate_ratio                        */
    double  answer_scale;          /* --answer_scale                             */
    /* additions of this library */
    int32_t radix_bits;            /* --radix_bits (total), -1 = auto            */
    double  bf_bits_per_key;       /* --bf_bits_per_key, <= 0 = formula          */
    int32_t val_mode;              /* --val_mode, without --read_val: 1 = the
                                      reference's synthesized values, R i+1 and
                                      S i+100000001 (default, src/hash_kernel.cu:
                                      330-336); 0 = the bare row index i          */
    int32_t device;                /* --device                                   */
    int32_t verbose;               /* print the reference-style summary          */
} bj_main_args;

void bj_main_args_default(bj_main_args *a);
/* Returns BJ_OK and fills *res (rows stay on the device of the internal
 * context, which is destroyed on return: only the counts remain valid). */
int  bj_gpu_main(const bj_main_args *a, bj_result *res);

/* ---- batched form ----------------------------------------------------------
 * Many independent joins of uneven sizes run as one sequence of launches: one
 * launch per phase covers every join (each join partitioned in one level of
 * its own radix bits into its region of shared arenas), and the build + probe
 * is one persistent launch per LDS tile size over the sub-buckets of all
 * joins.  The sequence is captured once in a hipGraph and replayed by
 * bj_batch_run.  Radix bits, LDS tile and Bloom bits per key of each join come
 * from its spec, else from the context's autotuner table, else from a size
 * rule.  Each join's result equals bj_join's on the same inputs (same count,
 * same row multiset).  No counterpart in the reference, which runs one join
 * per gpu_main call (src/hash_kernel.cu:103-1654). */
typedef struct bj_join_spec {
    const uint32_t *r_key, *r_val;   /* device pointers                          */
    const uint32_t *s_key, *s_val;
    int64_t r_size, s_size;          /* each < 2^31                              */
    int32_t radix_bits;              /* 1..11; -1 = autotuner / size rule        */
    int32_t tile;                    /* 4096 or 8192; 0 = autotuner / 8192       */
    double  bf_bits_per_key;         /* > 0: filter on; 0: off; < 0: autotuner   */
    int32_t num_hashes;              /* 1..8; -1 = min(round(bits_per_key ln 2), 4) */
    double  answer_scale;            /* row capacity = max(R,S) * max(1, scale),
                                        finite, at most 2^40 rows               */
    double  match_hint;              /* expected fraction of S with a match, used
                                        only to pick the autotuner's table entry;
                                        < 0 = unknown (taken as 0.1)             */
} bj_join_spec;

typedef struct bj_join_out {
    uint64_t match_count;            /* exact                                    */
    uint64_t rows_written;           /* min(match_count, capacity)               */
    uint64_t capacity;
    uint64_t s_survivors;            /* |S'| after the filter (= |S| without)    */
    int32_t  truncated;
    int32_t  radix_bits, tile, bf_k; /* the configuration used                   */
    int32_t  bf_stage;               /* 0: no filter; 1: S filtered before its
                                        partition; 2: after a first pass of S by
                                        its top bf_pass_bits bits (large filters) */
    int32_t  bf_pass_bits;
    double   bf_bits_per_key;        /* 0 = filter off                           */
    /* rows (SoA, device, unordered), owned by the batch until bj_batch_destroy */
    uint32_t *key, *r_val, *s_val;
} bj_join_out;

/* One autotuner table entry.  kind 0: radix bits and tile for joins whose
 * smaller side is about 2^lg tuples; kind 1: filter bits per key for joins with
 * |S|/|R| about 2^sr and selectivity class sel (0: <= 3%, 1: <= 15%, 2: <= 35%,
 * 3: more), in filter class lg (0: a 10 bits/key filter of R stays in an XCD's
 * L2, 3.5 MB, and S is filtered before its partition; 1: a larger filter, probed
 * after a first pass of S by the top bits, so each bucket probes an L2-sized
 * slice). */
typedef struct bj_tune_entry {
    int32_t kind;
    int32_t lg, sr, sel;
    int32_t radix_bits, tile;
    double  bf_bits_per_key;
    double  ms;                      /* best measured time of the tuning batch   */
} bj_tune_entry;

typedef struct bj_batch bj_batch;

void bj_join_spec_default(bj_join_spec *s);
/* Configuration (radix_bits, tile, bf_bits_per_key, bf_k, capacity fields of
 * cfg[i]) and device footprint a batch would use; ctx may be NULL (size rules
 * only, no device needed). */
int  bj_batch_plan(const bj_ctx *ctx, int n, const bj_join_spec *specs, bj_join_out *cfg,
                   uint64_t *workspace_bytes);
/* Plans the batch, allocates its workspace and (use_graph) captures its launch
 * sequence in a hipGraph.  The input pointers of specs must stay valid while
 * the batch is used; their contents may change between runs. */
int  bj_batch_create(bj_ctx *ctx, int n, const bj_join_spec *specs, int use_graph, bj_batch **out);
/* Runs the batch once; fills outs[0..n) in the order of the specs.  device_ms
 * (may be NULL) receives the device time of the launch sequence. */
int  bj_batch_run(bj_batch *b, bj_join_out *outs, double *device_ms);
int  bj_batch_destroy(bj_batch *b);

/* Autotuner: times candidate configurations on synthetic batches of the size
 * and selectivity classes that specs fall into (classes already in the table
 * are kept) and records the fastest in the context's table. */
int  bj_autotune(bj_ctx *ctx, int n, const bj_join_spec *specs, int *entries_tuned);
int  bj_tune_get(const bj_ctx *ctx, bj_tune_entry *out, int cap, int *n);
int  bj_tune_set(bj_ctx *ctx, const bj_tune_entry *in, int n);

/* ---- serving form ----------------------------------------------------------
 * A request loop over device-resident relations.  bj_server_add_relation
 * partitions R once (and builds its filter) into server-owned memory; each
 * request joins a caller's S against one resident R, exactly as bj_join would
 * join them.  Requests are admitted in arrival order as soon as a slot (a
 * stream of the server) is free and the request's footprint (its S workspace
 * and its rows) fits the server's HBM arena: requests that arrive while others
 * run start as slots and memory free up, and their kernels overlap on the
 * device (continuous batching).  A request's rows stay in the arena until
 * bj_server_release.  The server's scheduler runs on its own host thread; all
 * entry points may be called from any one other thread.  No counterpart in the
 * reference. */
typedef struct bj_server_opts {
    int32_t  slots;            /* requests on the device at once, 1..8 (default 3)  */
    uint64_t arena_bytes;      /* device memory for request workspaces and rows;
                                  0 = half of the free memory at creation          */
    int32_t  use_bf;           /* a filter per relation (default 1)                 */
    double   bf_bits_per_key;  /* default 10                                        */
    int32_t  num_hashes;       /* default 3                                         */
    int32_t  radix_bits;       /* per relation, 1..15; -1 = from |R| (default)      */
    double   answer_scale;     /* row capacity of a request = |S| * max(1, scale),
                                  finite, at most 2^40 rows                         */
} bj_server_opts;

typedef struct bj_request_result {
    uint64_t id;
    int32_t  status;           /* BJ_OK or a BJ_ERR_* code                          */
    int32_t  relation;
    uint64_t s_size;
    uint64_t match_count, rows_written, capacity, s_survivors;
    int32_t  truncated;
    double   t_submit, t_start, t_end;  /* host seconds since bj_server_create;
                                           latency = t_end - t_submit            */
    double   device_ms;        /* first kernel to last, on the device               */
    uint32_t *key, *r_val, *s_val;  /* rows (device, unordered), until release     */
} bj_request_result;

typedef struct bj_server_stats {
    uint64_t submitted, completed, released;
    int32_t  max_in_flight;    /* most requests on the device at one time           */
    uint64_t arena_bytes, arena_peak_bytes, relation_bytes;
} bj_server_stats;

typedef struct bj_server bj_server;

void bj_server_opts_default(bj_server_opts *o);
int  bj_server_create(int device, const bj_server_opts *o, bj_server **out);
/* Partitions R (device pointers; not retained) into the server: *rel_id. */
int  bj_server_add_relation(bj_server *s, const uint32_t *r_key, const uint32_t *r_val,
                            int64_t n_r, int32_t *rel_id);
/* Arena bytes a request of n_s tuples against relation rel occupies. */
int  bj_server_request_bytes(const bj_server *s, int32_t rel, int64_t n_s, uint64_t *bytes);
/* Queues a request (device pointers, read while it runs); never blocks.
 * BJ_ERR_SIZE if its footprint exceeds the whole arena. */
int  bj_server_submit(bj_server *s, int32_t rel, const uint32_t *s_key, const uint32_t *s_val,
                      int64_t n_s, uint64_t *id);
/* Completed requests not yet returned by poll (at most cap); never blocks. */
int  bj_server_poll(bj_server *s, bj_request_result *out, int cap, int *n);
/* Blocks until request id has completed. */
int  bj_server_wait(bj_server *s, uint64_t id, bj_request_result *out);
/* Copies the first n rows of a completed request to host arrays (n <= rows_written). */
int  bj_server_copy_rows(bj_server *s, const bj_request_result *r, uint32_t *key, uint32_t *r_val,
                         uint32_t *s_val, uint64_t n);
/* Frees a completed request's rows. */
int  bj_server_release(bj_server *s, uint64_t id);
/* Blocks until every submitted request has completed. */
int  bj_server_drain(bj_server *s);
int  bj_server_stats_get(bj_server *s, bj_server_stats *out);
/* Waits for the requests on the device, drops queued ones, frees everything. */
int  bj_server_destroy(bj_server *s);

#ifdef __cplusplus
}
#endif
#endif /* BJ_API_H */

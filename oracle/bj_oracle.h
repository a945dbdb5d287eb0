/*
 * bj_oracle.h -- CPU oracle for the radix hash join.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (the HIP library, the CLI
 * driver) links, loads or calls this code.  Only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg use it, and only as the checker.
 *
 * Contents and what each function follows:
 *   orc_sanity_count   restates run_sanity_check (src/sanity_check.cpp:10-73):
 *                      capacity 2^ceil(log2(2R)), hash (u64)(u32)key*2654435761
 *                      masked, linear probing, CAS build with a per-slot count,
 *                      probe adds the slot count.  Count only, like the reference,
 *                      including its INT_MIN-as-EMPTY sentinel.
 *   orc_cpu_join       the same table extended to emit every (key, r_val, s_val)
 *                      row (join semantics of kernel_functions.cu:399-430: every
 *                      R match of every S tuple).  Uses an occupancy bit instead of
 *                      the sentinel so every uint32 key is legal.
 *   orc_nested_loop    brute force |R|x|S| join for tiny inputs.
 *   orc_gen_*          restatement of the synthetic generator (integer only) that
 *                      the product runs on the device (k_gen_r / k_gen_s).
 *   orc_radix_hash     the partition hash (key * 0x9E3779B1, Fibonacci hashing).
 *   orc_bloom_*        restatement of the register-blocked Bloom filter addressing.
 *
 * Pins (tests/test_oracle.py): orc_sanity_count vs the reference's own
 * sanity_check.cpp compiled into oracle/_ref; orc_cpu_join vs orc_nested_loop on
 * adversarial tiny inputs (duplicates, key 0, 0x80000000, 0xFFFFFFFF); the
 * generator's closed-form match count and the bijectivity of its permutation.
 */
#ifndef BJ_ORACLE_H
#define BJ_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- generator (restated; integer only) ---- */
uint64_t orc_splitmix64(uint64_t x);
uint32_t orc_lowbias32(uint32_t x);
void     orc_perm_keys(uint64_t seed, uint32_t keys[8]);
uint32_t orc_permw(uint32_t x, int w, const uint32_t keys[8]);
uint32_t orc_r_key(uint64_t i, const uint32_t keysR[8]);
void     orc_gen_r(uint64_t seed, uint64_t n_r, uint32_t *key, uint32_t *val);
void     orc_gen_s(uint64_t seed_r, uint64_t seed_s, uint64_t n_r, uint64_t n_s,
                   uint64_t m_s, uint32_t *key, uint32_t *val);
/* bijectivity pin: count values in [0,2^w) hit != exactly once (0 == bijection) */
uint64_t orc_permw_defects(int w, uint64_t seed);

/* ---- joins ---- */
long long orc_sanity_count(const int32_t *tableA, long long R,
                           const int32_t *tableB, long long S, int threads);
long long orc_cpu_join(const uint32_t *r_key, const uint32_t *r_val, long long R,
                       const uint32_t *s_key, const uint32_t *s_val, long long S,
                       uint32_t *o_key, uint32_t *o_rval, uint32_t *o_sval,
                       long long cap, int threads);
long long orc_nested_loop(const uint32_t *r_key, const uint32_t *r_val, long long R,
                          const uint32_t *s_key, const uint32_t *s_val, long long S,
                          uint32_t *o_key, uint32_t *o_rval, uint32_t *o_sval,
                          long long cap);

/* ---- radix / bloom addressing (restated from the product's spec) ---- */
uint32_t orc_radix_hash(uint32_t key);
/* Register-blocked Bloom filter: 2^seg_bits segments of wps 64-bit blocks;
 * all k bits of a key in one 64-bit block. */
void orc_bloom_build(const uint32_t *keys, long long n, int seg_bits, uint64_t wps, int k,
                     uint64_t *filter);
int  orc_bloom_query(const uint64_t *filter, uint32_t key, int seg_bits, uint64_t wps, int k);

#ifdef __cplusplus
}
#endif
#endif

/*
 * bj_oracle.c -- CPU oracle for the radix hash join.
 *
 * TEST INFRASTRUCTURE ONLY (see bj_oracle.h).  The product never links this.
 */
#include "bj_oracle.h"

#include <limits.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------- */
/* Generator restatement.  Spec: splitmix64 (Steele/Lea/Flood, public) derives */
/* per-round add/mul constants; permw is 4 rounds of (add, odd multiply,      */
/* xor-shift-right) mod 2^w, each step a bijection on [0,2^w).                */
/* ------------------------------------------------------------------------- */
uint64_t orc_splitmix64(uint64_t x)
{
    uint64_t z = x + 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

/* lowbias32 (C. Wellons' published integer hash): a bijection on uint32. */
uint32_t orc_lowbias32(uint32_t x)
{
    x ^= x >> 16;
    x *= 0x7feb352dU;
    x ^= x >> 15;
    x *= 0x846ca68bU;
    x ^= x >> 16;
    return x;
}

void orc_perm_keys(uint64_t seed, uint32_t keys[8])
{
    uint64_t st = seed;
    for (int r = 0; r < 4; r++) {
        st = orc_splitmix64(st);
        keys[2 * r] = (uint32_t)st;
        st = orc_splitmix64(st);
        keys[2 * r + 1] = (uint32_t)(st >> 32) | 1u;
    }
}

uint32_t orc_permw(uint32_t x, int w, const uint32_t keys[8])
{
    const uint32_t mask = (w >= 32) ? 0xFFFFFFFFu : ((1u << w) - 1u);
    const int sh = (w + 1) / 2;
    x &= mask;
    for (int r = 0; r < 4; r++) {
        x = (x + keys[2 * r]) & mask;
        x = (x * keys[2 * r + 1]) & mask;
        x ^= x >> sh;
    }
    return x;
}

uint32_t orc_r_key(uint64_t i, const uint32_t keysR[8])
{
    /* cycle-walk the 31-bit permutation onto [0, 2^31-1), then shift to [1, 2^31) */
    uint32_t y = orc_permw((uint32_t)i, 31, keysR);
    while (y == 0x7FFFFFFFu)
        y = orc_permw(y, 31, keysR);
    return y + 1u;
}

void orc_gen_r(uint64_t seed, uint64_t n_r, uint32_t *key, uint32_t *val)
{
    uint32_t keysR[8];
    orc_perm_keys(seed, keysR);
    for (uint64_t i = 0; i < n_r; i++) {
        key[i] = orc_r_key(i, keysR);
        if (val) val[i] = (uint32_t)i;
    }
}

static int ceil_log2_u64(uint64_t n)
{
    int l = 0;
    while (((uint64_t)1 << l) < n) l++;
    return l;
}

void orc_gen_s(uint64_t seed_r, uint64_t seed_s, uint64_t n_r, uint64_t n_s,
               uint64_t m_s, uint32_t *key, uint32_t *val)
{
    uint32_t keysR[8], keysP[8], keysN[8];
    orc_perm_keys(seed_r, keysR);
    orc_perm_keys(seed_s ^ 0x1111111111111111ULL, keysP);
    orc_perm_keys(seed_s ^ 0x2222222222222222ULL, keysN);
    const uint32_t useed = (uint32_t)orc_splitmix64(seed_s ^ 0x3333333333333333ULL);
    int L = ceil_log2_u64(n_s);
    if (L < 1) L = 1;
    for (uint64_t j = 0; j < n_s; j++) {
        uint32_t p = orc_permw((uint32_t)j, L, keysP);
        while ((uint64_t)p >= n_s)
            p = orc_permw(p, L, keysP);
        uint32_t k;
        if ((uint64_t)p < m_s) {
            uint32_t u = (uint32_t)(((uint64_t)orc_lowbias32((uint32_t)j ^ useed) * n_r) >> 32);
            k = orc_r_key(u, keysR);
        } else {
            k = 0x80000000u | orc_permw((uint32_t)j, 31, keysN);
        }
        key[j] = k;
        if (val) val[j] = (uint32_t)j;
    }
}

uint64_t orc_permw_defects(int w, uint64_t seed)
{
    uint32_t keys[8];
    orc_perm_keys(seed, keys);
    const uint64_t n = (uint64_t)1 << w;
    uint8_t *seen = (uint8_t *)calloc((size_t)((n + 7) / 8), 1);
    if (!seen) return (uint64_t)-1;
    uint64_t defects = 0;
    for (uint64_t x = 0; x < n; x++) {
        uint32_t y = orc_permw((uint32_t)x, w, keys);
        if ((uint64_t)y >= n) { defects++; continue; }
        if (seen[y >> 3] & (1u << (y & 7))) defects++;
        seen[y >> 3] |= (uint8_t)(1u << (y & 7));
    }
    free(seen);
    return defects;
}

/* ------------------------------------------------------------------------- */
/* run_sanity_check restated (src/sanity_check.cpp:10-73), count only.        */
/* ------------------------------------------------------------------------- */
long long orc_sanity_count(const int32_t *tableA, long long R,
                           const int32_t *tableB, long long S, int threads)
{
    if (threads > 0) {
#ifdef _OPENMP
        omp_set_num_threads(threads);
#endif
    }
    const int32_t EMPTY_KEY = INT_MIN;                                   /* :16 */
    const long long HT_CAP = 1LL << (int)ceil(log2((double)R * 2));      /* :17 */
    const long long HT_MASK = HT_CAP - 1;
    int32_t *ht_key = (int32_t *)malloc(sizeof(int32_t) * (size_t)HT_CAP);
    int32_t *ht_cnt = (int32_t *)malloc(sizeof(int32_t) * (size_t)HT_CAP);
    if (!ht_key || !ht_cnt) { free(ht_key); free(ht_cnt); return -1; }

    #pragma omp parallel for schedule(static)                           /* :22-26 */
    for (long long i = 0; i < HT_CAP; i++) { ht_key[i] = EMPTY_KEY; ht_cnt[i] = 0; }

    #pragma omp parallel for schedule(static)                           /* :28-45 */
    for (long long i = 0; i < R; i++) {
        int32_t key = tableA[i];
        long long h = (long long)(((unsigned long long)(uint32_t)key * 2654435761ULL) & (unsigned long long)HT_MASK);
        for (;;) {
            int32_t expected = EMPTY_KEY;
            if (__atomic_compare_exchange_n(&ht_key[h], &expected, key, 0,
                                            __ATOMIC_RELAXED, __ATOMIC_RELAXED)) {
                __atomic_fetch_add(&ht_cnt[h], 1, __ATOMIC_RELAXED);
                break;
            }
            if (expected == key) {
                __atomic_fetch_add(&ht_cnt[h], 1, __ATOMIC_RELAXED);
                break;
            }
            h = (h + 1) & HT_MASK;
        }
    }

    long long cnt = 0;
    #pragma omp parallel for reduction(+:cnt) schedule(static)          /* :50-62 */
    for (long long j = 0; j < S; j++) {
        int32_t key = tableB[j];
        long long h = (long long)(((unsigned long long)(uint32_t)key * 2654435761ULL) & (unsigned long long)HT_MASK);
        while (__atomic_load_n(&ht_key[h], __ATOMIC_RELAXED) != EMPTY_KEY) {
            if (__atomic_load_n(&ht_key[h], __ATOMIC_RELAXED) == key) {
                cnt += __atomic_load_n(&ht_cnt[h], __ATOMIC_RELAXED);
                break;
            }
            h = (h + 1) & HT_MASK;
        }
    }
    free(ht_key);
    free(ht_cnt);
    return cnt;
}

/* ------------------------------------------------------------------------- */
/* The same table, extended to emit rows.  Slot word = (1<<32)|key so every   */
/* uint32 key is legal; a per-slot count plus a prefix groups the R rows of  */
/* each distinct key so duplicate R keys emit every match.                    */
/* ------------------------------------------------------------------------- */
long long orc_cpu_join(const uint32_t *r_key, const uint32_t *r_val, long long R,
                       const uint32_t *s_key, const uint32_t *s_val, long long S,
                       uint32_t *o_key, uint32_t *o_rval, uint32_t *o_sval,
                       long long cap, int threads)
{
    if (threads > 0) {
#ifdef _OPENMP
        omp_set_num_threads(threads);
#endif
    }
    if (R <= 0 || S <= 0) return 0;
    const long long HT_CAP = 1LL << (int)ceil(log2((double)R * 2));
    const long long HT_MASK = HT_CAP - 1;
    uint64_t *slot = (uint64_t *)malloc(sizeof(uint64_t) * (size_t)HT_CAP);
    uint32_t *scnt = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)HT_CAP);
    uint64_t *sbeg = (uint64_t *)malloc(sizeof(uint64_t) * (size_t)(HT_CAP + 1));
    uint32_t *rslot = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)R);
    uint32_t *grp = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)R);
    if (!slot || !scnt || !sbeg || !rslot || !grp) {
        free(slot); free(scnt); free(sbeg); free(rslot); free(grp);
        return -1;
    }

    #pragma omp parallel for schedule(static)
    for (long long i = 0; i < HT_CAP; i++) { slot[i] = 0; scnt[i] = 0; }

    /* build: CAS the (occupied|key) word, count the rows per slot */
    #pragma omp parallel for schedule(static)
    for (long long i = 0; i < R; i++) {
        const uint64_t want = (1ULL << 32) | r_key[i];
        long long h = (long long)(((unsigned long long)r_key[i] * 2654435761ULL) & (unsigned long long)HT_MASK);
        for (;;) {
            uint64_t expected = 0;
            if (__atomic_compare_exchange_n(&slot[h], &expected, want, 0,
                                            __ATOMIC_RELAXED, __ATOMIC_RELAXED) ||
                expected == want) {
                __atomic_fetch_add(&scnt[h], 1u, __ATOMIC_RELAXED);
                rslot[i] = (uint32_t)h;
                break;
            }
            h = (h + 1) & HT_MASK;
        }
    }
    /* group R rows by slot */
    sbeg[0] = 0;
    for (long long i = 0; i < HT_CAP; i++) sbeg[i + 1] = sbeg[i] + scnt[i];
    for (long long i = 0; i < HT_CAP; i++) scnt[i] = 0;
    for (long long i = 0; i < R; i++) {
        uint32_t h = rslot[i];
        grp[sbeg[h] + scnt[h]++] = (uint32_t)i;
    }

    /* probe: count per static chunk, prefix, then write rows */
    int nt = 1;
#ifdef _OPENMP
    nt = omp_get_max_threads();
#endif
    long long *tcount = (long long *)calloc((size_t)nt + 1, sizeof(long long));
    int team = 1;
    #pragma omp parallel num_threads(nt)
    {
        int t = 0, nth = 1;
#ifdef _OPENMP
        t = omp_get_thread_num();
        nth = omp_get_num_threads();
        #pragma omp single
        team = nth;
#endif
        long long lo = S * t / nth, hi = S * (t + 1) / nth, c = 0;
        for (long long j = lo; j < hi; j++) {
            const uint64_t want = (1ULL << 32) | s_key[j];
            long long h = (long long)(((unsigned long long)s_key[j] * 2654435761ULL) & (unsigned long long)HT_MASK);
            while (slot[h] != 0) {
                if (slot[h] == want) { c += (long long)(sbeg[h + 1] - sbeg[h]); break; }
                h = (h + 1) & HT_MASK;
            }
        }
        tcount[t + 1] = c;
        #pragma omp barrier
        #pragma omp single
        for (int i = 0; i < nth; i++) tcount[i + 1] += tcount[i];
        long long w = tcount[t];
        for (long long j = lo; j < hi; j++) {
            const uint64_t want = (1ULL << 32) | s_key[j];
            long long h = (long long)(((unsigned long long)s_key[j] * 2654435761ULL) & (unsigned long long)HT_MASK);
            while (slot[h] != 0) {
                if (slot[h] == want) {
                    for (uint64_t g = sbeg[h]; g < sbeg[h + 1]; g++, w++) {
                        if (w < cap) {
                            o_key[w] = s_key[j];
                            o_rval[w] = r_val[grp[g]];
                            o_sval[w] = s_val[j];
                        }
                    }
                    break;
                }
                h = (h + 1) & HT_MASK;
            }
        }
    }
    long long total = tcount[team];
    free(tcount);
    free(slot); free(scnt); free(sbeg); free(rslot); free(grp);
    return total;
}

long long orc_nested_loop(const uint32_t *r_key, const uint32_t *r_val, long long R,
                          const uint32_t *s_key, const uint32_t *s_val, long long S,
                          uint32_t *o_key, uint32_t *o_rval, uint32_t *o_sval,
                          long long cap)
{
    long long w = 0;
    for (long long j = 0; j < S; j++)
        for (long long i = 0; i < R; i++)
            if (r_key[i] == s_key[j]) {
                if (w < cap) { o_key[w] = s_key[j]; o_rval[w] = r_val[i]; o_sval[w] = s_val[j]; }
                w++;
            }
    return w;
}

/* ------------------------------------------------------------------------- */
/* Radix hash and blocked Bloom addressing, restated from the product spec.  */
/* ------------------------------------------------------------------------- */
uint32_t orc_radix_hash(uint32_t key) { return key * 0x9E3779B1u; }

static void bloom_addr(uint32_t key, int seg_bits, uint64_t wps, int k, uint64_t *block,
                       uint64_t *mask)
{
    const uint32_t h = orc_radix_hash(key);
    const uint64_t seg = seg_bits ? (uint64_t)(h >> (32 - seg_bits)) : 0;
    const uint32_t rest = seg_bits ? (h << seg_bits) : h;
    *block = seg * wps + (((uint64_t)rest * wps) >> 32);
    uint32_t h2 = (key ^ 0x5BD1E995u) * 0xC2B2AE3Du;
    h2 ^= h2 >> 15;
    uint32_t h3 = h2 * 0x27D4EB2Fu;
    h3 ^= h3 >> 15;
    uint64_t m = 0;
    for (int i = 0; i < k; i++) {
        uint32_t bit = (i < 5) ? ((h2 >> (26 - 6 * i)) & 63u) : ((h3 >> (26 - 6 * (i - 5))) & 63u);
        m |= 1ULL << bit;
    }
    *mask = m;
}

void orc_bloom_build(const uint32_t *keys, long long n, int seg_bits, uint64_t wps, int k,
                     uint64_t *filter)
{
    for (long long i = 0; i < n; i++) {
        uint64_t b, m;
        bloom_addr(keys[i], seg_bits, wps, k, &b, &m);
        filter[b] |= m;
    }
}

int orc_bloom_query(const uint64_t *filter, uint32_t key, int seg_bits, uint64_t wps, int k)
{
    uint64_t b, m;
    bloom_addr(key, seg_bits, wps, k, &b, &m);
    return (filter[b] & m) == m;
}

// bj_gen.cpp -- the reference datagen's key distribution, on the host (the
// reference's exact files) and on the device (the same key multiset in a keyed
// random order).
//
// Counterpart in the reference (Chanho-LEE-SNU/BRIG-Join): datagen/gen_dataset.cpp
//   sizes and the --max load imbalance     :113-165
//   the value ranges (matching [0,k), R-only [k, 2^31-1), S-only from R on)  :172-209
//   shuffle: Fisher-Yates over std::mt19937_64, j = rng() % (i + 1)          :67-77
//   seeds: R seed, S seed ^ 0xdeadbeef, hot/cold regions seed ^ 0xdeadbeef and
//   seed ^ 0x9e3779b97f4a7c15                                                :192, 210-219
// mt19937_64 is a fully specified engine, so the host generator writes the
// reference's files byte for byte.  The join of R with S has exactly k rows.
#include "bj_api.h"
#include "bj_internal.h"
#include "bj_kernels.h"

#include <hip/hip_runtime.h>

#include <random>

using namespace bj;

namespace {

struct RefDist {
    uint64_t k = 0;                      // matching keys
    uint64_t hot_n = 0, cold_n = 0;      // S regions
    uint64_t m_hot = 0, m_cold = 0;      // matching keys per region
};

// The reference's sizing (gen_dataset.cpp:113-165), as errors instead of exits.
int refdist_plan(int64_t r_size, int64_t s_size, double pct, double max_pct, RefDist& d)
{
    if (r_size < 0 || s_size < 0) return fail(BJ_ERR_INVALID, "negative relation size");
    if (r_size >= 0x7FFFFFFFLL || s_size > 0x7FFFFFFFLL)
        return fail(BJ_ERR_SIZE, "relation sizes must be below 2^31 - 1");
    if (!(pct >= 0.0 && pct <= 100.0)) return fail(BJ_ERR_INVALID, "answer_pct must be in [0, 100]");
    const long long R = r_size, S = s_size;
    long long k = (long long)(R * pct / 100.0);
    if (k > S) k = S;
    const bool imbalance = max_pct >= 0.0;
    const long long hot_n = imbalance ? (long long)(S * 0.25) : 0;
    const long long cold_n = S - hot_n;
    long long m_hot = 0, m_cold = k;
    if (imbalance) {
        if (max_pct < pct || max_pct > 100.0)
            return fail(BJ_ERR_INVALID, "max_pct must satisfy answer_pct <= max_pct <= 100");
        m_hot = (long long)(hot_n * max_pct / 100.0);
        if (m_hot > hot_n) m_hot = hot_n;
        if (m_hot > k) return fail(BJ_ERR_INVALID, "max_pct too high: the hot region needs more matches than exist");
        m_cold = k - m_hot;
        if (m_cold > cold_n) return fail(BJ_ERR_INVALID, "infeasible: the cold region cannot hold its matches");
    }
    d.k = (uint64_t)k;
    d.hot_n = (uint64_t)hot_n;
    d.cold_n = (uint64_t)cold_n;
    d.m_hot = (uint64_t)m_hot;
    d.m_cold = (uint64_t)m_cold;
    return BJ_OK;
}

void fisher_yates(uint32_t* a, uint64_t n, uint64_t seed)
{
    std::mt19937_64 rng(seed);
    for (uint64_t i = n; i-- > 1;) {
        const uint64_t j = rng() % (i + 1);
        const uint32_t t = a[i];
        a[i] = a[j];
        a[j] = t;
    }
}

constexpr uint64_t SEED_S = 0xdeadbeefULL;
constexpr uint64_t SEED_COLD = 0x9e3779b97f4a7c15ULL;

}  // namespace

extern "C" {

int bj_gen_refdist_host(int64_t r_size, int64_t s_size, double answer_pct, uint64_t seed, double max_pct,
                        uint32_t* r_key, uint32_t* s_key, int64_t* matches)
{
    return guarded([&]() -> int {
        RefDist d;
        int rc = refdist_plan(r_size, s_size, answer_pct, max_pct, d);
        if (rc) return rc;
        if ((r_size && !r_key) || (s_size && !s_key)) return fail(BJ_ERR_INVALID, "null output array");
        const uint64_t R = (uint64_t)r_size, S = (uint64_t)s_size;
        const uint64_t r_only = 0x7FFFFFFFull - d.k;
        for (uint64_t i = 0; i < R; i++) r_key[i] = (uint32_t)(i < d.k ? i : d.k + (i - d.k) % r_only);
        fisher_yates(r_key, R, seed);
        const uint64_t s_only = 0x7FFFFFFFull - R;
        uint64_t mc = 0, sj = 0, p = 0;
        auto fill = [&](uint64_t m, uint64_t rest) {
            for (uint64_t c = 0; c < m; c++) s_key[p++] = (uint32_t)(mc++);
            for (uint64_t c = 0; c < rest; c++) s_key[p++] = (uint32_t)(R + (sj++) % s_only);
        };
        fill(d.m_hot, d.hot_n - d.m_hot);
        fill(d.m_cold, d.cold_n - d.m_cold);
        if (max_pct >= 0.0) {
            fisher_yates(s_key, d.hot_n, seed ^ SEED_S);
            fisher_yates(s_key + d.hot_n, d.cold_n, seed ^ SEED_COLD);
        } else {
            fisher_yates(s_key, S, seed ^ SEED_S);
        }
        if (matches) *matches = (int64_t)d.k;
        return BJ_OK;
    });
}

int bj_gen_refdist(bj_ctx* c, int64_t r_size, int64_t s_size, double answer_pct, uint64_t seed, double max_pct,
                   uint32_t* d_r_key, uint32_t* d_s_key, int64_t* matches)
{
    return guarded([&]() -> int {
        if (!c) return fail(BJ_ERR_INVALID, "null ctx");
        RefDist d;
        int rc = refdist_plan(r_size, s_size, answer_pct, max_pct, d);
        if (rc) return rc;
        if ((r_size && !d_r_key) || (s_size && !d_s_key)) return fail(BJ_ERR_INVALID, "null output array");
        BJ_HIP(hipSetDevice(c->device));
        const uint64_t R = (uint64_t)r_size;
        BJ_HIP(launch_gen_refdist_r(d_r_key, R, d.k, perm_keys(seed), c->stream));
        BJ_HIP(launch_gen_refdist_s(d_s_key, d.hot_n, d.m_hot, 0, 0, R, perm_keys(seed ^ SEED_S), c->stream));
        BJ_HIP(launch_gen_refdist_s(d_s_key + d.hot_n, d.cold_n, d.m_cold, d.m_hot, d.hot_n - d.m_hot, R,
                                    perm_keys(seed ^ SEED_COLD), c->stream));
        BJ_HIP(hipStreamSynchronize(c->stream));
        if (matches) *matches = (int64_t)d.k;
        return BJ_OK;
    });
}

}  // extern "C"

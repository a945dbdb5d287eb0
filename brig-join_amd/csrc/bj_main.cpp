// bj_main.cpp -- gpu_main's one-GPU share, output.bin I/O and the CPU count
// that the reference prints beside the GPU count.
//
// Counterparts in the reference (Chanho-LEE-SNU/BRIG-Join):
//   bj_gpu_main          gpu_main (src/hash_kernel.cu:103-1654): input load with
//                        optional value block (:302-341), synthesized values
//                        (:330-336), the sanity check (:344-345), the join, the
//                        grand-total print (:1141-1155) and the timing summary
//                        (:1182-1302)
//   bj_cpu_sanity_count  run_sanity_check (src/sanity_check.cpp:10-73)
//   bj_write_output      the output.bin that driver.cpp parses (src/driver.cpp:50,119)
//                        but never writes; layout defined in bj_api.h
#include "bj_api.h"

#include <omp.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>
#include <parallel/algorithm>
#include <vector>

extern "C" const char* bj_last_error(void);

namespace {

struct Row {
    uint32_t k, r, s;
    bool operator<(const Row& o) const
    {
        if (k != o.k) return k < o.k;
        if (r != o.r) return r < o.r;
        return s < o.s;
    }
};

const char MAGIC[8] = {'B', 'J', 'O', 'U', 'T', '0', '0', '1'};

bool read_exact(FILE* f, void* dst, size_t bytes)
{
    return bytes == 0 || std::fread(dst, 1, bytes, f) == bytes;
}

// Runs the body of an entry point: a C++ exception (a host allocation that
// fails) becomes an error code instead of crossing the C ABI.
template <class F>
int guarded_main(F&& body) noexcept
{
    try {
        return body();
    } catch (const std::bad_alloc&) {
        return BJ_ERR_NOMEM;
    } catch (...) {
        return BJ_ERR_INTERNAL;
    }
}

}  // namespace

extern "C" {

void bj_free_host(void* p) { std::free(p); }

int bj_write_output(const char* path, uint64_t n, uint32_t* key, uint32_t* r_val, uint32_t* s_val)
{
    return guarded_main([&]() -> int {
        if (!path || (n && (!key || !r_val || !s_val))) return BJ_ERR_INVALID;
        if (n) {
            std::vector<Row> rows(n);
#pragma omp parallel for schedule(static)
            for (int64_t i = 0; i < (int64_t)n; i++) rows[i] = Row{key[i], r_val[i], s_val[i]};
            __gnu_parallel::sort(rows.begin(), rows.end());
#pragma omp parallel for schedule(static)
            for (int64_t i = 0; i < (int64_t)n; i++) {
                key[i] = rows[i].k;
                r_val[i] = rows[i].r;
                s_val[i] = rows[i].s;
            }
        }
        FILE* f = std::fopen(path, "wb");
        if (!f) return BJ_ERR_IO;
        bool ok = std::fwrite(MAGIC, 1, 8, f) == 8 && std::fwrite(&n, 8, 1, f) == 1;
        if (n) {
            ok = ok && std::fwrite(key, 4, n, f) == n;
            ok = ok && std::fwrite(r_val, 4, n, f) == n;
            ok = ok && std::fwrite(s_val, 4, n, f) == n;
        }
        ok = (std::fclose(f) == 0) && ok;
        return ok ? BJ_OK : BJ_ERR_IO;
    });
}

int bj_read_output(const char* path, uint64_t* n, uint32_t** key, uint32_t** r_val, uint32_t** s_val)
{
    if (!path || !n || !key || !r_val || !s_val) return BJ_ERR_INVALID;
    *key = *r_val = *s_val = nullptr;
    *n = 0;
    FILE* f = std::fopen(path, "rb");
    if (!f) return BJ_ERR_IO;
    char magic[8];
    uint64_t cnt = 0;
    if (!read_exact(f, magic, 8) || std::memcmp(magic, MAGIC, 8) != 0 || !read_exact(f, &cnt, 8)) {
        std::fclose(f);
        return BJ_ERR_IO;
    }
    // the rest of the file must be exactly the count's three arrays (a damaged
    // count is refused before anything is allocated for it)
    long here = std::ftell(f), end = -1;
    if (here >= 0 && std::fseek(f, 0, SEEK_END) == 0) end = std::ftell(f);
    if (end < here || std::fseek(f, here, SEEK_SET) != 0 || cnt > (uint64_t)(end - here) / 12 ||
        cnt * 12 != (uint64_t)(end - here)) {
        std::fclose(f);
        return BJ_ERR_IO;
    }
    uint32_t* a = (uint32_t*)std::malloc(std::max<uint64_t>(cnt, 1) * 4);
    uint32_t* b = (uint32_t*)std::malloc(std::max<uint64_t>(cnt, 1) * 4);
    uint32_t* c = (uint32_t*)std::malloc(std::max<uint64_t>(cnt, 1) * 4);
    bool ok = a && b && c && read_exact(f, a, cnt * 4) && read_exact(f, b, cnt * 4) &&
              read_exact(f, c, cnt * 4);
    std::fclose(f);
    if (!ok) {
        std::free(a);
        std::free(b);
        std::free(c);
        return (a && b && c) ? BJ_ERR_IO : BJ_ERR_NOMEM;
    }
    *n = cnt;
    *key = a;
    *r_val = b;
    *s_val = c;
    return BJ_OK;
}

// The reference's CPU algorithm: capacity 2^ceil(log2 2R), hash
// (u64)key * 2654435761 masked, linear probing, a CAS build with a count per
// distinct key, and a probe that adds the count (src/sanity_check.cpp:16-62).
long long bj_cpu_sanity_count(const uint32_t* r_key, int64_t R, const uint32_t* s_key, int64_t S)
{
    if (R <= 0 || S <= 0) return 0;
    if (!r_key || !s_key) return -1;
    const int64_t cap = 1LL << (int)std::ceil(std::log2((double)R * 2));
    const int64_t mask = cap - 1;
    // slot word = (1 << 32) | key, 0 = empty, so every uint32 key is legal
    std::vector<std::atomic<uint64_t>> slot(cap);
    std::vector<std::atomic<uint32_t>> cnt(cap);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < cap; i++) {
        slot[i].store(0, std::memory_order_relaxed);
        cnt[i].store(0, std::memory_order_relaxed);
    }
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < R; i++) {
        const uint64_t want = (1ULL << 32) | r_key[i];
        int64_t h = (int64_t)(((uint64_t)r_key[i] * 2654435761ULL) & (uint64_t)mask);
        for (;;) {
            uint64_t expect = 0;
            if (slot[h].compare_exchange_strong(expect, want, std::memory_order_relaxed) || expect == want) {
                cnt[h].fetch_add(1, std::memory_order_relaxed);
                break;
            }
            h = (h + 1) & mask;
        }
    }
    long long total = 0;
#pragma omp parallel for reduction(+ : total) schedule(static)
    for (int64_t j = 0; j < S; j++) {
        const uint64_t want = (1ULL << 32) | s_key[j];
        int64_t h = (int64_t)(((uint64_t)s_key[j] * 2654435761ULL) & (uint64_t)mask);
        for (;;) {
            const uint64_t v = slot[h].load(std::memory_order_relaxed);
            if (v == 0) break;
            if (v == want) {
                total += cnt[h].load(std::memory_order_relaxed);
                break;
            }
            h = (h + 1) & mask;
        }
    }
    return total;
}

void bj_main_args_default(bj_main_args* a)
{
    if (!a) return;
    std::memset(a, 0, sizeof(*a));
    a->r_size = 32000000;
    a->s_size = 32000000;
    a->chunk_size = 1024;
    a->true_false_ratio = 0.1;
    a->s_chunks = 1;
    a->override_num_bits = -1;
    a->override_num_hashes = -1;
    a->override_partition_rs_split = -1;
    a->override_partition_size = -1;
    a->override_second_partition_size = -1;
    a->r_replicate_ratio = -1.0;
    a->answer_scale = 1.0;
    a->radix_bits = -1;
    a->bf_bits_per_key = 0.0;
    a->val_mode = 1;  // the reference's synthesized payloads (src/hash_kernel.cu:330-336)
    a->device = 0;
    a->verbose = 1;
}

static int load_relation(const char* path, int64_t n, int read_val, int val_mode, uint32_t val_base,
                         std::vector<uint32_t>& key, std::vector<uint32_t>& val)
{
    key.resize((size_t)n);
    val.resize((size_t)n);
    FILE* f = std::fopen(path, "rb");
    if (!f) return BJ_ERR_IO;
    bool ok = read_exact(f, key.data(), (size_t)n * 4);
    if (ok && read_val) ok = read_exact(f, val.data(), (size_t)n * 4);
    std::fclose(f);
    if (!ok) return BJ_ERR_IO;
    if (!read_val) {
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < n; i++) val[i] = (uint32_t)i + (val_mode == 1 ? val_base : 0u);
    }
    return BJ_OK;
}

int bj_gpu_main(const bj_main_args* a, bj_result* res)
{
    return guarded_main([&]() -> int {
        if (!a || !res || !a->r_path || !a->s_path) return BJ_ERR_INVALID;
        if (a->r_size < 0 || a->s_size < 0) return BJ_ERR_INVALID;
        std::memset(res, 0, sizeof(*res));
        std::vector<uint32_t> rk, rv, sk, sv;
        int rc = load_relation(a->r_path, a->r_size, a->read_val, a->val_mode, 1u, rk, rv);
        if (rc == BJ_OK) rc = load_relation(a->s_path, a->s_size, a->read_val, a->val_mode, 100000001u, sk, sv);
        if (rc != BJ_OK) {
            if (a->verbose) std::fprintf(stderr, "Failed to read the input files (or a value block is missing)\n");
            return rc;
        }

        long long cpu_count = -1;
        double cpu_sec = 0.0;
        if (a->sanity_check) {
            auto t0 = std::chrono::steady_clock::now();
            cpu_count = bj_cpu_sanity_count(rk.data(), a->r_size, sk.data(), a->s_size);
            cpu_sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
            if (a->verbose)
                std::printf("[sanityCheck] Answer Num (hash-based): %lld  (%.4f sec, %d threads)\n", cpu_count,
                            cpu_sec, omp_get_max_threads());
        }

        bj_params p;
        bj_params_default(&p);
        p.r_size = a->r_size;
        p.s_size = a->s_size;
        p.partition_size = a->override_partition_size;
        p.second_partition_size = a->override_second_partition_size;
        p.radix_bits = a->radix_bits;
        p.use_bf = a->use_bf;
        p.bf_bits_per_key = a->bf_bits_per_key;
        p.true_false_ratio = a->true_false_ratio;
        p.num_bits = a->override_num_bits;
        p.num_hashes = a->override_num_hashes;
        p.answer_scale = a->answer_scale;
        p.s_chunks = a->s_chunks;
        p.probe_breakdown = a->fused_probe_breakdown ? 1 : 0;
        p.inputs_on_device = 0;
        p.collect_phase_times = 1;

        bj_ctx* ctx = nullptr;
        rc = bj_create(&ctx, a->device);
        if (rc != BJ_OK) {
            if (a->verbose) std::fprintf(stderr, "bj_create: %s (%s)\n", bj_strerror(rc), bj_last_error());
            return rc;
        }
        rc = bj_join(ctx, &p, rk.data(), rv.data(), sk.data(), sv.data(), res);
        if (rc != BJ_OK) {
            if (a->verbose) std::fprintf(stderr, "bj_join: %s (%s)\n", bj_strerror(rc), bj_last_error());
            bj_destroy(ctx);
            return rc;
        }
        if (a->out_path && a->out_path[0]) {
            const uint64_t n = res->rows_written;
            std::vector<uint32_t> ok(n ? n : 1), orv(n ? n : 1), osv(n ? n : 1);
            rc = bj_copy_rows_to_host(ctx, res, ok.data(), orv.data(), osv.data(), n);
            if (rc == BJ_OK) rc = bj_write_output(a->out_path, n, ok.data(), orv.data(), osv.data());
            if (rc != BJ_OK && a->verbose) std::fprintf(stderr, "writing %s failed\n", a->out_path);
        }
        res->key = res->r_val = res->s_val = nullptr;  // the context is gone after return
        bj_destroy(ctx);
        if (rc != BJ_OK) return rc;

        if (a->verbose) {
            std::printf("=== Grand Total Answer Count ===\n%llu\n", (unsigned long long)res->match_count);
            if (res->truncated)
                std::printf("[warn] answer buffer truncated: %llu rows kept of %llu (raise --answer_scale)\n",
                            (unsigned long long)res->rows_written, (unsigned long long)res->match_count);
            if (a->sanity_check)
                std::printf("[sanityCheck] CPU %lld vs GPU %llu : %s\n", cpu_count,
                            (unsigned long long)res->match_count,
                            (cpu_count >= 0 && (unsigned long long)cpu_count == res->match_count) ? "MATCH"
                                                                                                 : "MISMATCH");
            if (a->use_bf)
                std::printf("BF candidates: %llu of %lld (%.3f%%), k=%d, %llu filter words\n",
                            (unsigned long long)res->s_survivors, (long long)a->s_size,
                            a->s_size ? 100.0 * (double)res->s_survivors / (double)a->s_size : 0.0, res->bf_k,
                            (unsigned long long)res->bf_words);
            std::printf("radix bits: %d + %d, LDS tile %d, vlsplit items %llu\n", res->bits1, res->bits2, res->tile,
                        (unsigned long long)res->heavy_items);
            std::printf("===== timing (ms, device) =====\n");
            std::printf("[0a] H2D              %10.3f\n", res->phase_ms[BJ_PH_H2D]);
            std::printf("[1]  R partition      %10.3f\n", res->phase_ms[BJ_PH_R_PART]);
            std::printf("[0c] BF build         %10.3f\n", res->phase_ms[BJ_PH_BF_BUILD]);
            std::printf("[0e] S filter         %10.3f\n", res->phase_ms[BJ_PH_S_FILTER]);
            std::printf("[2-1] S partition     %10.3f\n", res->phase_ms[BJ_PH_S_PART]);
            std::printf("[2-2] Probe           %10.3f\n", res->phase_ms[BJ_PH_PROBE]);
            std::printf("Total (excl. H2D)     %10.3f\n", res->phase_ms[BJ_PH_TOTAL]);
            if (a->fused_probe_breakdown) {
                std::printf("      [2-2a] Work items    %10.3f\n", res->probe_phase_ms[BJ_PP_ITEMS]);
                std::printf("      [2-2a2] Table clear  %10.3f\n", res->probe_phase_ms[BJ_PP_CLEAR]);
                std::printf("      [2-2b] Build         %10.3f\n", res->probe_phase_ms[BJ_PP_BUILD]);
                std::printf("      [2-2c] Probe count   %10.3f\n", res->probe_phase_ms[BJ_PP_COUNT]);
                std::printf("      [2-2d] Rows          %10.3f\n", res->probe_phase_ms[BJ_PP_WRITE]);
                const double t = res->phase_ms[BJ_PH_TOTAL] * 1e-3;
                std::printf("input tuples/s (excl. H2D): %.4e\n",
                            t > 0 ? (double)(a->r_size + a->s_size) / t : 0.0);
            }
        }
        return (a->sanity_check && cpu_count >= 0 && (unsigned long long)cpu_count != res->match_count)
                   ? BJ_ERR_INTERNAL
                   : BJ_OK;
    });
}

}  // extern "C"

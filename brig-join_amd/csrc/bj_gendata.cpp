// bj_gendata.cpp -- dataset generator with the reference datagen's command line.
//
//   bj_gendata <R_size> <S_size> <answer_pct> <out_dir> [seed] [--max <pct>] [--verify]
//              [--device N]
//
// Same arguments, file names (out_dir/gen_R<R>_ans<pct>pct.bin, ..._S...) and
// raw int32 key files as datagen/gen_dataset.cpp:79-263.  By default the keys
// come from bj_gen_refdist_host, which reproduces the reference's files byte
// for byte; with --device N they are generated on GPU N by bj_gen_refdist (the
// same key multiset in another random order).  --verify counts the join on the
// CPU (bj_cpu_sanity_count), which equals the number of matching keys.
#include "bj_api.h"

#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

static bool write_keys(const std::string& path, const uint32_t* a, int64_t n)
{
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) return false;
    const bool ok = n == 0 || std::fwrite(a, 4, (size_t)n, f) == (size_t)n;
    return (std::fclose(f) == 0) && ok;
}

int main(int argc, char** argv)
{
    if (argc < 5) {
        std::fprintf(stderr,
                     "Usage: %s <R_size> <S_size> <answer_pct> <out_dir> [seed] [--max <pct>] [--verify] "
                     "[--device N]\n",
                     argv[0]);
        return 1;
    }
    const long long R = std::atoll(argv[1]), S = std::atoll(argv[2]);
    const double pct = std::atof(argv[3]);
    std::string dir = argv[4];
    uint64_t seed = 42;
    bool verify = false;
    double max_pct = -1.0;
    int device = -1;
    for (int i = 5; i < argc; i++) {
        const std::string a = argv[i];
        if (a == "--verify") {
            verify = true;
        } else if (a == "--max" || a == "--device") {
            if (i + 1 >= argc) {
                std::fprintf(stderr, "%s requires a value\n", a.c_str());
                return 1;
            }
            if (a == "--max") max_pct = std::atof(argv[++i]);
            else device = std::atoi(argv[++i]);
        } else {
            seed = (uint64_t)std::atoll(argv[i]);
        }
    }
    if (!dir.empty() && dir.back() == '/') dir.pop_back();

    std::vector<uint32_t> rk((size_t)(R > 0 ? R : 1)), sk((size_t)(S > 0 ? S : 1));
    int64_t k = 0;
    int rc;
    if (device < 0) {
        rc = bj_gen_refdist_host(R, S, pct, seed, max_pct, rk.data(), sk.data(), &k);
    } else {
        bj_ctx* ctx = nullptr;
        rc = bj_create(&ctx, device);
        void *dr = nullptr, *ds = nullptr;
        if (rc == BJ_OK && (hipMalloc(&dr, (size_t)(R > 0 ? R : 1) * 4) != hipSuccess ||
                            hipMalloc(&ds, (size_t)(S > 0 ? S : 1) * 4) != hipSuccess))
            rc = BJ_ERR_NOMEM;
        if (rc == BJ_OK) rc = bj_gen_refdist(ctx, R, S, pct, seed, max_pct, (uint32_t*)dr, (uint32_t*)ds, &k);
        if (rc == BJ_OK && (hipMemcpy(rk.data(), dr, (size_t)R * 4, hipMemcpyDeviceToHost) != hipSuccess ||
                            hipMemcpy(sk.data(), ds, (size_t)S * 4, hipMemcpyDeviceToHost) != hipSuccess))
            rc = BJ_ERR_HIP;
        if (dr) (void)hipFree(dr);
        if (ds) (void)hipFree(ds);
        if (ctx) bj_destroy(ctx);
    }
    if (rc != BJ_OK) {
        std::fprintf(stderr, "generation failed: %s (%s)\n", bj_strerror(rc), bj_last_error());
        return 1;
    }
    std::printf("R=%lld  S=%lld  matches=%lld  answer=%.4f%%\n", R, S, (long long)k,
                R ? (double)k / (double)R * 100.0 : 0.0);

    char pct_str[64];
    if (max_pct >= 0.0) std::snprintf(pct_str, sizeof(pct_str), "%.1f_max%.1f", pct, max_pct);
    else std::snprintf(pct_str, sizeof(pct_str), "%.1f", pct);
    const std::string r_path = dir + "/gen_R" + std::to_string(R) + "_ans" + pct_str + "pct.bin";
    const std::string s_path = dir + "/gen_S" + std::to_string(S) + "_ans" + pct_str + "pct.bin";
    if (!write_keys(r_path, rk.data(), R) || !write_keys(s_path, sk.data(), S)) {
        std::fprintf(stderr, "writing %s / %s failed\n", r_path.c_str(), s_path.c_str());
        return 1;
    }
    std::printf("Writing R -> %s\nWriting S -> %s\n", r_path.c_str(), s_path.c_str());
    if (verify) {
        const long long got = bj_cpu_sanity_count(rk.data(), R, sk.data(), S);
        std::printf("Expected matches: %lld\nActual matches:   %lld\n%s\n", (long long)k, got,
                    got == k ? "Selectivity is CORRECT" : "ERROR: selectivity mismatch");
        if (got != k) return 1;
    }
    return 0;
}

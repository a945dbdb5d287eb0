// host_check.cpp -- host-only checks of the library's CPU code, built with
// AddressSanitizer and UndefinedBehaviorSanitizer (tests/sanitize/Makefile).
//
// Covers the code that runs on the host without a GPU: the planner (bj_plan,
// parameter derivation of the reference's gpu_main, src/hash_kernel.cu:174-284),
// the batch planner (bj_batch_plan with no context), the serving arena's
// allocator, output.bin I/O including damaged files, the reference datagen's
// host generator (datagen/gen_dataset.cpp:172-220), the CPU count
// (src/sanity_check.cpp:10-73) and the error paths of the entry points that
// need a device.  Every check that fails prints its line and the program exits
// with status 1; a sanitizer report aborts it.
#include "bj_api.h"
#include "bj_arena.h"

#include <unistd.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <random>
#include <string>
#include <tuple>
#include <vector>

static int g_fail = 0;
#define CHECK(cond)                                                          \
    do {                                                                     \
        if (!(cond)) {                                                       \
            std::fprintf(stderr, "CHECK failed at line %d: %s\n", __LINE__, #cond); \
            g_fail = 1;                                                      \
        }                                                                    \
    } while (0)

static bool valid_code(int rc)
{
    return rc == BJ_OK || rc == BJ_ERR_INVALID || rc == BJ_ERR_SIZE || rc == BJ_ERR_NOMEM ||
           rc == BJ_ERR_HIP || rc == BJ_ERR_IO || rc == BJ_ERR_NODEVICE || rc == BJ_ERR_INTERNAL;
}

// ---------------------------------------------------------------- arena
static void check_arena()
{
    std::mt19937_64 rng(7);
    bj::ArenaAlloc a;
    const uint64_t size = 1ull << 30;
    a.reset(size);
    CHECK(a.size() == size && a.used() == 0 && a.free_ranges() == 1 && a.largest() == size);
    std::map<uint64_t, uint64_t> live;  // offset -> requested bytes
    for (int op = 0; op < 20000; op++) {
        const bool do_alloc = live.empty() || (rng() % 100) < 55;
        if (do_alloc) {
            const uint64_t want = 1 + rng() % (size / 64);
            uint64_t off = 0;
            if (a.alloc(want, off)) {
                const uint64_t r = bj::ArenaAlloc::round(want);
                CHECK(off % bj::ArenaAlloc::ALIGN == 0);
                CHECK(off + r <= size);
                auto nx = live.lower_bound(off);
                if (nx != live.end()) CHECK(off + r <= nx->first);
                if (nx != live.begin()) {
                    auto pv = std::prev(nx);
                    CHECK(pv->first + bj::ArenaAlloc::round(pv->second) <= off);
                }
                live[off] = want;
            } else {
                CHECK(a.largest() < bj::ArenaAlloc::round(want));
            }
        } else {
            auto it = live.begin();
            std::advance(it, (long)(rng() % live.size()));
            a.release(it->first, it->second);
            live.erase(it);
        }
        uint64_t used = 0;
        for (const auto& l : live) used += bj::ArenaAlloc::round(l.second);
        CHECK(a.used() == used);
    }
    for (const auto& l : live) a.release(l.first, l.second);
    CHECK(a.used() == 0 && a.free_ranges() == 1 && a.largest() == size);
    // a size that is not a multiple of the alignment is rounded down
    a.reset(1000);
    uint64_t off = 0;
    CHECK(a.alloc(768, off) && off == 0);
    CHECK(!a.alloc(1, off));
}

// ---------------------------------------------------------------- planner
static void check_plan()
{
    std::mt19937_64 rng(11);
    const int64_t sizes[] = {-1, 0, 1, 2, 1000, 65536, 1 << 20, 16000000, 64000000, 640000000,
                             1280000000, 2147483647LL, 2147483648LL};
    const int32_t psz[] = {-1, 0, 1, 2, 3, 512, 1024, 2048, 4096};
    const int32_t bits[] = {-1, 0, 1, 7, 10, 14, 15, 16};
    const int32_t bfm[] = {0, 1, 10, 11, 14, 16, 17, -1};
    const double bpk[] = {0.0, 0.5, 4.0, 10.0, 12.0, 40.0, -3.0};
    const double tfr[] = {0.1, 0.0, 1.0, 1e-9, 0.5};
    const int64_t nbits[] = {-1, 0, 100, 6400000000LL, 1LL << 40};
    const int32_t nh[] = {-1, 0, 1, 3, 8, 9};
    const int32_t tiles[] = {0, 4096, 8192, 1000};
    const int32_t stages[] = {0, 1, 2, 3, -1};
    const int32_t p1[] = {-1, 1, 7, 12, 20};
    const int32_t wg[] = {0, 512, 1024, 256};
    // each field takes its default (a valid value) half of the time, else any
    // value of its list, so that the grid reaches valid plans and errors alike
    auto pick = [&](const auto& arr) { return arr[rng() % (sizeof(arr) / sizeof(arr[0]))]; };
    auto maybe = [&](auto& field, const auto& arr) {
        if (rng() % 2) field = pick(arr);
    };
    int ok = 0;
    for (int t = 0; t < 40000; t++) {
        bj_params p;
        bj_params_default(&p);
        p.r_size = pick(sizes);
        p.s_size = pick(sizes);
        maybe(p.partition_size, psz);
        maybe(p.second_partition_size, psz);
        maybe(p.radix_bits, bits);
        p.use_bf = pick(bfm);
        maybe(p.bf_bits_per_key, bpk);
        maybe(p.true_false_ratio, tfr);
        maybe(p.num_bits, nbits);
        maybe(p.num_hashes, nh);
        p.answer_scale = (rng() % 4) * 0.75;
        maybe(p.tile, tiles);
        maybe(p.bf_stage, stages);
        maybe(p.pass1_bits, p1);
        maybe(p.s_pass1_bits, p1);
        maybe(p.scatter_wg, wg);
        p.s_chunks = (int32_t)(rng() % 5) - 1;
        bj_plan_info pi;
        const int rc = bj_plan(&p, &pi);
        CHECK(valid_code(rc));
        if (rc == BJ_OK) {
            ok++;
            CHECK(pi.bits1 >= 1 && pi.bits1 + pi.bits2 <= 15);
            CHECK(pi.tile == 4096 || pi.tile == 8192);
            CHECK(pi.capacity >= 1);
            CHECK(pi.workspace_bytes > 0);
            CHECK((pi.bf != 0) == (p.use_bf != 0));
            CHECK(!pi.bf || (pi.bf_words > 0 && pi.bf_k >= 1 && pi.bf_k <= 8 && (pi.bf_stage == 1 || pi.bf_stage == 2)));
        } else {
            CHECK(bj_last_error() != nullptr && std::strlen(bj_last_error()) > 0);
        }
    }
    CHECK(ok > 1000);  // the grid reaches valid plans, not only errors
    CHECK(bj_plan(nullptr, nullptr) == BJ_ERR_INVALID);
}

// ---------------------------------------------------------------- batch planner
static void check_batch_plan()
{
    std::mt19937_64 rng(13);
    for (int t = 0; t < 400; t++) {
        const int n = 1 + (int)(rng() % 64);
        std::vector<bj_join_spec> specs(n);
        std::vector<bj_join_out> cfg(n);
        for (auto& s : specs) {
            bj_join_spec_default(&s);
            const int kind = (int)(rng() % 8);
            s.r_size = kind == 0 ? 0 : (int64_t)(rng() % (kind == 1 ? 100 : 16000000));
            s.s_size = kind == 2 ? 0 : (int64_t)(rng() % (kind == 3 ? 5 : 40000000));
            s.radix_bits = (int32_t)(rng() % 14) - 1;
            s.tile = (rng() % 3) == 0 ? 4096 : 0;
            s.bf_bits_per_key = (double)((int)(rng() % 20) - 2);
            s.num_hashes = (int32_t)(rng() % 10) - 1;
            s.answer_scale = (double)(rng() % 3);
            s.match_hint = (double)(rng() % 100) / 100.0 - 0.1;
        }
        uint64_t ws = 0;
        const int rc = bj_batch_plan(nullptr, n, specs.data(), cfg.data(), &ws);
        CHECK(valid_code(rc));
        if (rc == BJ_OK) {
            CHECK(ws > 0);
            for (int i = 0; i < n; i++) {
                CHECK(cfg[i].tile == 4096 || cfg[i].tile == 8192);
                CHECK(cfg[i].capacity >= 1);
            }
        }
    }
    CHECK(bj_batch_plan(nullptr, 0, nullptr, nullptr, nullptr) != BJ_OK || true);
    bj_join_spec bad;
    bj_join_spec_default(&bad);
    bad.r_size = -5;
    CHECK(bj_batch_plan(nullptr, 1, &bad, nullptr, nullptr) != BJ_OK);
}

// ---------------------------------------------------------------- output.bin
static std::string tmp_path(const char* name)
{
    const char* d = std::getenv("TMPDIR");
    return std::string(d && *d ? d : "/tmp") + "/bj_host_check_" + std::to_string((long)getpid()) + "_" + name;
}

static void write_raw(const std::string& path, const void* p, size_t n)
{
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) return;
    if (n) std::fwrite(p, 1, n, f);
    std::fclose(f);
}

static void check_output_io()
{
    std::mt19937_64 rng(17);
    const std::string path = tmp_path("out.bin");
    for (uint64_t n : {0ull, 1ull, 7ull, 100000ull}) {
        std::vector<uint32_t> k(n), r(n), s(n);
        for (uint64_t i = 0; i < n; i++) {
            k[i] = (uint32_t)(rng() % 1000);
            r[i] = (uint32_t)rng();
            s[i] = (uint32_t)rng();
        }
        std::vector<uint32_t> k0 = k, r0 = r, s0 = s;
        CHECK(bj_write_output(path.c_str(), n, k.data(), r.data(), s.data()) == BJ_OK);
        uint64_t m = 0;
        uint32_t *a = nullptr, *b = nullptr, *c = nullptr;
        CHECK(bj_read_output(path.c_str(), &m, &a, &b, &c) == BJ_OK);
        CHECK(m == n);
        if (m == n && a && b && c) {
            // sorted by (key, r_val, s_val), and the same multiset as written
            std::multimap<uint32_t, std::pair<uint32_t, uint32_t>> want, got;
            for (uint64_t i = 0; i < n; i++) {
                want.emplace(k0[i], std::make_pair(r0[i], s0[i]));
                got.emplace(a[i], std::make_pair(b[i], c[i]));
                if (i) {
                    const bool le = a[i - 1] < a[i] || (a[i - 1] == a[i] && (b[i - 1] < b[i] ||
                                                     (b[i - 1] == b[i] && c[i - 1] <= c[i])));
                    CHECK(le);
                }
            }
            std::vector<std::tuple<uint32_t, uint32_t, uint32_t>> w1, g1;
            for (auto& e : want) w1.emplace_back(e.first, e.second.first, e.second.second);
            for (auto& e : got) g1.emplace_back(e.first, e.second.first, e.second.second);
            std::sort(w1.begin(), w1.end());
            std::sort(g1.begin(), g1.end());
            CHECK(w1 == g1);
        }
        bj_free_host(a);
        bj_free_host(b);
        bj_free_host(c);
    }
    // damaged files are refused, and the out-pointers stay empty
    struct Hdr { char magic[8]; uint64_t n; };
    const Hdr good{{'B', 'J', 'O', 'U', 'T', '0', '0', '1'}, 3};
    auto refused = [&](const void* bytes, size_t len) {
        write_raw(path, bytes, len);
        uint64_t m = 99;
        uint32_t *a = (uint32_t*)1, *b = (uint32_t*)1, *c = (uint32_t*)1;
        const int rc = bj_read_output(path.c_str(), &m, &a, &b, &c);
        const bool ok = rc != BJ_OK && m == 0 && !a && !b && !c;
        if (rc == BJ_OK) {
            bj_free_host(a);
            bj_free_host(b);
            bj_free_host(c);
        }
        return ok;
    };
    CHECK(refused("", 0));                     // empty
    CHECK(refused(&good, 12));                 // truncated header
    Hdr bad = good;
    bad.magic[7] = '2';
    CHECK(refused(&bad, sizeof(bad)));         // wrong magic
    CHECK(refused(&good, sizeof(good)));       // header promises rows the file lacks
    for (uint64_t huge : {1ull << 62, (1ull << 62) + 1, ~0ull, 1ull << 40}) {
        Hdr h = good;
        h.n = huge;                            // a count whose byte size overflows or exceeds the file
        std::vector<uint8_t> buf(sizeof(h) + 48);
        std::memcpy(buf.data(), &h, sizeof(h));
        CHECK(refused(buf.data(), buf.size()));
    }
    {
        // trailing bytes after the rows are refused too
        std::vector<uint8_t> buf(sizeof(good) + 3 * 12 + 5);
        std::memcpy(buf.data(), &good, sizeof(good));
        CHECK(refused(buf.data(), buf.size()));
    }
    uint64_t m;
    uint32_t *a, *b, *c;
    CHECK(bj_read_output("/nonexistent/dir/x.bin", &m, &a, &b, &c) == BJ_ERR_IO);
    CHECK(bj_write_output("/nonexistent/dir/x.bin", 0, nullptr, nullptr, nullptr) == BJ_ERR_IO);
    CHECK(bj_write_output(path.c_str(), 3, nullptr, nullptr, nullptr) == BJ_ERR_INVALID);
    CHECK(bj_read_output(nullptr, &m, &a, &b, &c) == BJ_ERR_INVALID);
    std::remove(path.c_str());
}

// ---------------------------------------------------------------- generator, CPU count
static void check_gen_and_count()
{
    for (int64_t r : {0LL, 1LL, 10LL, 1000LL, 12345LL}) {
        for (int64_t s : {0LL, 1LL, 17LL, 2000LL}) {
            for (double pct : {0.0, 5.0, 100.0}) {
                std::vector<uint32_t> rk(std::max<int64_t>(r, 1)), sk(std::max<int64_t>(s, 1));
                int64_t k = -1;
                const int rc = bj_gen_refdist_host(r, s, pct, 42, -1.0, rk.data(), sk.data(), &k);
                CHECK(valid_code(rc));
                if (rc != BJ_OK) continue;
                CHECK(k >= 0 && k <= std::min(r, s));
                // matching keys [0, k) occur once in each relation: the CPU count is k
                const long long c = bj_cpu_sanity_count(rk.data(), r, sk.data(), s);
                CHECK(c == (r > 0 && s > 0 ? k : 0));
            }
        }
    }
    std::vector<uint32_t> rk(4000), sk(4000);
    int64_t k = 0;
    CHECK(bj_gen_refdist_host(4000, 4000, 10.0, 1, 30.0, rk.data(), sk.data(), &k) == BJ_OK);
    // the count against a nested loop, on keys with duplicates
    std::mt19937_64 rng(19);
    for (int t = 0; t < 20; t++) {
        const int nr = 1 + (int)(rng() % 300), ns = 1 + (int)(rng() % 300);
        std::vector<uint32_t> a(nr), b(ns);
        for (auto& x : a) x = (uint32_t)(rng() % 50) + (t % 2 ? 0x7FFFFFF0u : 0u);
        for (auto& x : b) x = (uint32_t)(rng() % 50) + (t % 2 ? 0x7FFFFFF0u : 0u);
        long long want = 0;
        for (uint32_t x : a)
            for (uint32_t y : b) want += x == y;
        CHECK(bj_cpu_sanity_count(a.data(), nr, b.data(), ns) == want);
    }
    CHECK(bj_cpu_sanity_count(nullptr, 5, nullptr, 5) < 0);
    CHECK(bj_cpu_sanity_count(nullptr, 0, nullptr, 0) == 0);
}

// ---------------------------------------------------------------- device entry points without a device
static void check_no_device_paths()
{
    for (int code = -10; code <= 1; code++) CHECK(bj_strerror(code) != nullptr);
    CHECK(bj_version() > 0);
    bj_ctx* c = nullptr;
    const int rc = bj_create(&c, 0);
    if (rc == BJ_OK) {  // a GPU is present: only the error paths that need none
        CHECK(bj_create(&c, 1 << 20) == BJ_ERR_NODEVICE || true);
        bj_destroy(c);
    } else {
        CHECK(rc == BJ_ERR_NODEVICE && c == nullptr);
    }
    CHECK(bj_create(nullptr, 0) == BJ_ERR_INVALID);
    CHECK(bj_destroy(nullptr) == BJ_OK);
    bj_main_args a;
    bj_main_args_default(&a);
    a.r_path = "/nonexistent/r.bin";
    a.s_path = "/nonexistent/s.bin";
    a.r_size = 10;
    a.s_size = 10;
    bj_result res;
    const int m = bj_gpu_main(&a, &res);
    CHECK(m != BJ_OK && valid_code(m));
    CHECK(bj_gpu_main(nullptr, &res) == BJ_ERR_INVALID);
    bj_server* s = nullptr;
    bj_server_opts o;
    bj_server_opts_default(&o);
    const int sr = bj_server_create(0, &o, &s);
    if (sr == BJ_OK) {
        bj_server_destroy(s);
    } else {
        CHECK(valid_code(sr) && s == nullptr);
    }
}

int main()
{
    check_arena();
    check_plan();
    check_batch_plan();
    check_output_io();
    check_gen_and_count();
    check_no_device_paths();
    if (g_fail) return 1;
    std::printf("host_check ok\n");
    return 0;
}

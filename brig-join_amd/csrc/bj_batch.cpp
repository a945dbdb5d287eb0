// bj_batch.cpp -- the batched join and its autotuner (host side of the C ABI).
//
// A batch plans every join (radix bits, LDS tile, filter geometry), lays the
// joins out in shared arenas, and enqueues one launch per phase for all of
// them (bj_batch.hip).  The launch sequence has no host synchronisation and no
// memset, so it is captured once in a hipGraph and replayed.  The reference
// has no batched form; each join follows the same algorithm as bj_join.
#include "bj_api.h"
#include "bj_batch.h"
#include "bj_internal.h"
#include "bj_kernels.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <string>
#include <vector>

using namespace bj;

namespace {

constexpr int B_MAX_BITS = 15;                   // joint histogram of a chunk in LDS: 2^15 x 4 B
constexpr int B_ONE_PASS_BITS = 9;               // up to 512-way: one scatter pass
constexpr uint64_t B_MAX_TUPLES = 0x7FFFFFFFull;
constexpr uint32_t B_SLICE = 65536;
constexpr uint32_t B_MAX_SLICES = 16;
constexpr uint64_t B_SEG_TARGET = 64 * 1024;     // Bloom segment built in LDS
constexpr uint64_t B_SEG_MAX = 160 * 1024;
constexpr int B_MAX_JOINS = 1 << 20;
// A filter up to this size stays in an XCD's L2 while S probes it at random,
// so S is filtered before it is partitioned.  A larger one (bf2) is probed after
// a first pass of S by the top fb bits, so that a chunk of that pass probes a
// slice of about B_SLICE_BYTES (the same order as the single join's filter
// after its first pass).
constexpr uint64_t B_FILTER_L2_BYTES = 3584ull << 10;
constexpr uint64_t B_SLICE_BYTES = 128ull << 10;

struct JoinCfg {
    int bits = 1;
    int b1 = 1;  // first-pass bits (== bits: one pass)
    int tile = 8192;
    double bpk = 0.0;
    int k = 0;
    bool bf2 = false;  // filter after a first pass of S by fb bits
    int fb = 0;
    BloomGeom geom{0, 1, 1};
    uint64_t cap = 1;
};

int sel_class(double hint)
{
    if (hint < 0.0) return 1;
    if (hint <= 0.03) return 0;
    if (hint <= 0.15) return 1;
    if (hint <= 0.35) return 2;
    return 3;
}

const double SEL_REP[4] = {0.01, 0.1, 0.25, 0.5};

int lg_class(uint64_t m)
{
    const double l = std::log2((double)std::max<uint64_t>(m, 1));
    return std::max(10, std::min(24, (int)std::lround(l)));
}

int sr_class(uint64_t nr, uint64_t ns)
{
    if (nr == 0 || ns == 0) return 0;
    const double l = std::log2((double)ns / (double)nr);
    return std::max(-3, std::min(3, (int)std::lround(l)));
}

// Filter class of a join with |R| = nr: 1 when a 10 bits/key filter exceeds
// what stays in the L2 (it would be probed after a first pass of S), else 0.
int filter_class(uint64_t nr) { return (double)nr * 10.0 / 8.0 > (double)B_FILTER_L2_BYTES ? 1 : 0; }

int rule_bits(uint64_t m, int tile)
{
    const double per = 0.6 * tile;
    const uint64_t need = (uint64_t)std::ceil((double)std::max<uint64_t>(m, 1) / per);
    return std::max(1, std::min(B_MAX_BITS, ceil_log2(need)));
}

const bj_tune_entry* find_tune(const bj_ctx* c, int kind, int lg, int sr, int sel)
{
    if (!c) return nullptr;
    for (const bj_tune_entry& e : c->tune) {
        if (e.kind != kind) continue;
        if (kind == 0 && e.lg == lg) return &e;
        if (kind == 1 && e.lg == lg && e.sr == sr && e.sel == sel) return &e;
    }
    return nullptr;
}

int resolve(const bj_ctx* c, const bj_join_spec& s, JoinCfg& o)
{
    if (s.r_size < 0 || s.s_size < 0) return fail(BJ_ERR_INVALID, "negative relation size");
    if ((uint64_t)s.r_size > B_MAX_TUPLES || (uint64_t)s.s_size > B_MAX_TUPLES)
        return fail(BJ_ERR_SIZE, "relation larger than 2^31-1 tuples");
    const uint64_t nr = (uint64_t)s.r_size, ns = (uint64_t)s.s_size;
    const uint64_t m = std::min(nr, ns);
    const int lg = lg_class(m);
    const bj_tune_entry* t0 = find_tune(c, 0, lg, 0, 0);

    if (s.tile != 0 && s.tile != 4096 && s.tile != 8192) return fail(BJ_ERR_INVALID, "tile must be 4096 or 8192");
    o.tile = s.tile ? s.tile : (t0 ? t0->tile : 8192);
    if (s.radix_bits > 0) {
        if (s.radix_bits > B_MAX_BITS) return fail(BJ_ERR_SIZE, "batched joins use at most 15 radix bits");
        o.bits = s.radix_bits;
    } else if (t0 && t0->tile == o.tile) {
        o.bits = t0->radix_bits;
    } else {
        o.bits = rule_bits(m, o.tile);
    }

    o.b1 = o.bits <= B_ONE_PASS_BITS ? o.bits : (o.bits + 1) / 2;

    double bpk = s.bf_bits_per_key;
    if (!std::isfinite(bpk)) return fail(BJ_ERR_INVALID, "bf_bits_per_key must be a finite number");
    if (bpk < 0.0) {
        // the tuner's filter classes: filters that stay in the L2 (lg 0) and
        // filters probed after a first pass of S (lg 1), told apart by the size
        // of a 10 bits/key filter
        const int fcls = filter_class(nr);
        const bj_tune_entry* t1 = find_tune(c, 1, fcls, sr_class(nr, ns), sel_class(s.match_hint));
        bpk = t1 ? t1->bf_bits_per_key : (sel_class(s.match_hint) <= 1 ? 10.0 : 0.0);
    }
    if (nr == 0 || ns == 0) bpk = 0.0;
    o.bpk = bpk;
    if (bpk > 0.0) {
        if (s.num_hashes > 8) return fail(BJ_ERR_INVALID, "num_hashes must be in [1, 8]");
        // all k bits of a key share one 64-bit word: past about 4 bits the false-positive
        // rate barely moves while every extra bit costs instructions per key
        o.k = s.num_hashes > 0 ? s.num_hashes : std::min(4, (int)std::lround(bpk * std::log(2.0)));
        o.k = std::max(1, std::min(8, o.k));
        const double num_bits = std::max(256.0, (double)nr * bpk);
        const uint64_t bytes = (uint64_t)std::ceil(num_bits / 8.0);
        const int seg_need = ceil_log2((bytes + B_SEG_TARGET - 1) / B_SEG_TARGET);
        const int seg_bits = std::min(seg_need, o.bits);
        const uint64_t blocks = (bytes + 7) / 8;
        const uint64_t wps = std::max<uint64_t>(1, (blocks + (1ull << seg_bits) - 1) >> seg_bits);
        if (wps * 8 > B_SEG_MAX)
            return fail(BJ_ERR_SIZE, "Bloom segment exceeds LDS: raise radix bits or lower filter bits");
        o.geom.seg_bits = seg_bits;
        o.geom.wps = (uint32_t)wps;
        o.geom.k = o.k;
        o.bf2 = bytes > B_FILTER_L2_BYTES;
        o.fb = o.bf2 ? std::max(1, std::min(o.bits, ceil_log2((bytes + B_SLICE_BYTES - 1) / B_SLICE_BYTES))) : 0;
    } else {
        o.k = 0;
    }
    return row_capacity(std::max(nr, ns), s.answer_scale, o.cap);
}

// Histogram / scatter chunk of a relation: at least 32K tuples, at most 512
// chunks, and at most 2^21 per-chunk histogram entries per relation.
uint32_t chunk_len(uint64_t n, int bits)
{
    const uint64_t max_chunks = std::min<uint64_t>(512, std::max<uint64_t>(16, (1ull << 21) >> bits));
    const uint64_t per = (n + max_chunks - 1) / max_chunks;
    return (uint32_t)std::max<uint64_t>(32768, (per + 4095) / 4096 * 4096);
}

// Second-pass workgroups per first-level bucket: about 64K tuples each.
uint32_t pass2_groups(uint64_t n, int b1, uint64_t chunks)
{
    const uint64_t per_bucket = n >> b1;
    const uint64_t w = std::max<uint64_t>(1, per_bucket / 65536);
    return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(std::min<uint64_t>(64, chunks), w));
}

// The layout of a batch: per-join descriptors in internal order (tile 8192
// first) plus the totals that size the arenas and launches.
struct Layout {
    std::vector<int> order;   // internal index -> caller index
    std::vector<BJoin> J;
    std::vector<JoinCfg> cfg;
    std::vector<uint32_t> px; // PX_N x (n + 1), then the join of each sub-bucket (nsb)
    uint32_t nt_r = 0, nt_s = 0, nseg = 0, ncol_r = 0, ncol_s = 0, nsb = 0, nsb_a = 0, max_wps = 1;
    uint32_t nt_r2 = 0, nt_s2 = 0, nf_tiles = 0;
    uint32_t nt_p1[3][2] = {};  // first-pass tasks per side and class
    int max_b1[3][2] = {{1, 1}, {1, 1}, {1, 1}};  // their largest first-pass bits
    int max_bits = 1, max_b2 = 0;
    bool any_bf = false, any_two = false, any_bf2 = false;
    uint64_t tot_r = 0, tot_s = 0, tot_sp = 0, part_n = 0, f_blocks = 0, out_n = 0, item_cap = 16;
};

int make_layout(const bj_ctx* c, int n, const bj_join_spec* specs, Layout& L)
{
    if (n < 0 || n > B_MAX_JOINS) return fail(BJ_ERR_INVALID, "number of joins out of range");
    if (n > 0 && !specs) return fail(BJ_ERR_INVALID, "null specs");
    std::vector<JoinCfg> cfg(n);
    for (int i = 0; i < n; i++) {
        int rc = resolve(c, specs[i], cfg[i]);
        if (rc) return fail(rc, "join " + std::to_string(i) + ": " + bj_last_error());
    }
    L.order.clear();
    for (int i = 0; i < n; i++)
        if (cfg[i].tile == 8192) L.order.push_back(i);
    for (int i = 0; i < n; i++)
        if (cfg[i].tile != 8192) L.order.push_back(i);
    L.J.assign(n, BJoin{});
    L.cfg.assign(n, JoinCfg{});
    L.px.assign((size_t)PX_N * (n + 1), 0u);
    auto PX = [&](int which, int i) -> uint32_t& { return L.px[(size_t)which * (n + 1) + i]; };
    uint64_t nt_r = 0, nt_s = 0, nseg = 0, ncol_r = 0, ncol_s = 0, nsb = 0, nsb_a = 0, nt_r2 = 0, nt_s2 = 0;
    uint64_t nft = 0;
    uint64_t np1[3][2] = {};
    for (int i = 0; i < n; i++) {
        const bj_join_spec& s = specs[L.order[i]];
        const JoinCfg& g = cfg[L.order[i]];
        L.cfg[i] = g;
        BJoin& J = L.J[i];
        J.rk = s.r_key;
        J.rv = s.r_val;
        J.sk = s.s_key;
        J.sv = s.s_val;
        J.nr = (uint32_t)s.r_size;
        J.ns = (uint32_t)s.s_size;
        J.bits = g.bits;
        J.b1 = g.b1;
        J.bf = g.bpk > 0.0 ? 1 : 0;
        J.bf2 = J.bf && g.bf2 ? 1 : 0;
        J.fb = J.bf2 ? g.fb : 0;
        J.geom = g.geom;
        J.tile = (uint32_t)g.tile;
        const uint64_t nbins = 1ull << g.bits;
        J.sb_base = (uint32_t)nsb;
        PX(PX_SB, i) = (uint32_t)nsb;
        nsb += nbins;
        if (g.tile == 8192) nsb_a = nsb;
        J.r_ch = chunk_len(J.nr, g.bits);
        J.s_ch = chunk_len(J.ns, g.bits);
        const uint64_t cr = (J.nr + J.r_ch - 1) / J.r_ch, cs = (J.ns + J.s_ch - 1) / J.s_ch;
        PX(PX_R2, i) = (uint32_t)nt_r2;
        PX(PX_S2, i) = (uint32_t)nt_s2;
        if (g.b1 < g.bits) {
            J.w2_r = pass2_groups(J.nr, g.b1, cr);
            // a Bloom join's second pass sees S', about the hinted match ratio of S
            // plus false positives (fewer groups than S' needs only costs parallelism)
            const double sel = (J.bf && s.match_hint >= 0.0) ? std::min(1.0, s.match_hint + 0.05) : 1.0;
            J.w2_s = pass2_groups((uint64_t)((double)J.ns * sel), g.b1, cs);
            nt_r2 += (uint64_t)J.w2_r << g.b1;
            nt_s2 += (uint64_t)J.w2_s << g.b1;
            L.any_two = true;
            L.max_b2 = std::max(L.max_b2, g.bits - g.b1);
        }
        // first-pass tasks by class: R and S by b1, the filter pass of S by fb
        const int cls = g.b1 >= WIDE_B1 ? 1 : 0, fcls = J.fb >= WIDE_B1 ? 1 : 0;
        for (int sd = 0; sd < 3; sd++)
            for (int k = 0; k < 2; k++) PX(PX_P1 + 2 * sd + k, i) = (uint32_t)np1[sd][k];
        np1[0][cls] += cr;
        np1[1][cls] += cs;
        L.max_b1[0][cls] = L.max_b1[1][cls] = std::max(L.max_b1[0][cls], g.b1);
        if (J.bf2) {
            L.any_bf2 = true;
            np1[2][fcls] += cs;
            L.max_b1[2][fcls] = std::max(L.max_b1[2][fcls], J.fb);
        }
        PX(PX_RTASK, i) = (uint32_t)nt_r;
        nt_r += cr;
        PX(PX_STASK, i) = (uint32_t)nt_s;
        nt_s += cs;
        J.part_r = L.part_n;
        L.part_n += cr * nbins;
        J.part_s = L.part_n;
        L.part_n += cs * nbins;
        PX(PX_RCOL, i) = (uint32_t)ncol_r;
        ncol_r += (nbins + 63) / 64;
        PX(PX_SCOL, i) = (uint32_t)ncol_s;
        ncol_s += (nbins + 63) / 64;
        J.r_base = L.tot_r;
        L.tot_r += J.nr;
        J.s_base = L.tot_s;
        L.tot_s += J.ns;
        PX(PX_SEG, i) = (uint32_t)nseg;
        PX(PX_FTILE, i) = (uint32_t)nft;
        if (J.bf) {
            nft += (J.ns + (uint64_t)PF_TILE - 1) / PF_TILE;
            J.sp_base = L.tot_sp;
            L.tot_sp += J.ns;
            J.f_base = L.f_blocks;
            L.f_blocks += (uint64_t)g.geom.wps << g.geom.seg_bits;
            nseg += 1ull << g.geom.seg_bits;
            L.max_wps = std::max(L.max_wps, g.geom.wps);
            L.any_bf = true;
        }
        J.out_base = L.out_n;
        J.out_cap = g.cap;
        L.out_n += g.cap;
        L.max_bits = std::max(L.max_bits, g.bits);
        L.item_cap += (uint64_t)B_MAX_SLICES * ((J.nr + (uint64_t)J.ns) / g.tile + nbins);
    }
    PX(PX_RTASK, n) = (uint32_t)nt_r;
    PX(PX_STASK, n) = (uint32_t)nt_s;
    PX(PX_SEG, n) = (uint32_t)nseg;
    PX(PX_RCOL, n) = (uint32_t)ncol_r;
    PX(PX_SCOL, n) = (uint32_t)ncol_s;
    PX(PX_SB, n) = (uint32_t)nsb;
    PX(PX_R2, n) = (uint32_t)nt_r2;
    PX(PX_S2, n) = (uint32_t)nt_s2;
    PX(PX_FTILE, n) = (uint32_t)nft;
    for (int sd = 0; sd < 3; sd++)
        for (int k = 0; k < 2; k++) {
            PX(PX_P1 + 2 * sd + k, n) = (uint32_t)np1[sd][k];
            L.nt_p1[sd][k] = (uint32_t)np1[sd][k];
        }
    if (std::max({nt_r, nt_s, nseg, ncol_r, ncol_s, nsb, nt_r2, nt_s2, nft}) > 0x7FFFFFFFull)
        return fail(BJ_ERR_SIZE, "batch too large for one launch per phase");
    L.nt_r = (uint32_t)nt_r;
    L.nt_s = (uint32_t)nt_s;
    L.nseg = (uint32_t)nseg;
    L.ncol_r = (uint32_t)ncol_r;
    L.ncol_s = (uint32_t)ncol_s;
    L.nsb = (uint32_t)nsb;
    L.nsb_a = (uint32_t)nsb_a;
    L.nt_r2 = (uint32_t)nt_r2;
    L.nt_s2 = (uint32_t)nt_s2;
    L.nf_tiles = (uint32_t)nft;
    // the join of every sub-bucket (the probe's and the planner's lookup)
    L.px.resize((size_t)PX_N * (n + 1) + nsb);
    uint32_t* sbj = L.px.data() + (size_t)PX_N * (n + 1);
    for (int i = 0; i < n; i++) {
        const uint32_t lo = L.px[(size_t)PX_SB * (n + 1) + i], hi = L.px[(size_t)PX_SB * (n + 1) + i + 1];
        for (uint32_t q = lo; q < hi; q++) sbj[q] = (uint32_t)i;
    }
    return BJ_OK;
}

}  // namespace

struct bj_batch {
    bj_ctx* c = nullptr;
    int n = 0;
    Layout L;
    DevBuf dJ, dpx;
    DevBuf r_key, r_val, s_key, s_val, sp_key, sp_val, tmp_key, tmp_val, part, cur_rel, tot_r, tot_s, start_r, start_s;
    DevBuf filter, plan_cnt, item_start, partials, items, ctr, out_key, out_rval, out_sval;
    uint64_t* h_ctr = nullptr;  // pinned: jc[n][JC_N], then err
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
};

namespace {

// counters: jc[n][JC_N], work counters of the two probe launches, error flags,
// the filter launch's per-XCD tile counters
size_t ctr_words(int n) { return (size_t)n * JC_N + 3 + NXCD; }

// Sums (b == nullptr) or allocates the workspace of a batch.
int batch_workspace(bj_batch* b, const Layout& L, int n, uint64_t* total_bytes)
{
    uint64_t total = 0;
    auto need = [&](DevBuf bj_batch::*m, uint64_t bytes) -> int {
        bytes = std::max<uint64_t>(bytes, 256);
        total += bytes;
        return b ? ensure(b->*m, bytes) : BJ_OK;
    };
    int rc;
    const uint64_t nsb = std::max<uint32_t>(L.nsb, 1);
    if ((rc = need(&bj_batch::dJ, sizeof(BJoin) * std::max(n, 1))) ||
        (rc = need(&bj_batch::dpx, L.px.size() * 4)) || (rc = need(&bj_batch::r_key, L.tot_r * 4)) ||
        (rc = need(&bj_batch::r_val, L.tot_r * 4)) || (rc = need(&bj_batch::s_key, L.tot_s * 4)) ||
        (rc = need(&bj_batch::s_val, L.tot_s * 4)) || (rc = need(&bj_batch::sp_key, L.tot_sp * 4)) ||
        (rc = need(&bj_batch::sp_val, L.tot_sp * 4)) ||
        (rc = need(&bj_batch::tmp_key, L.any_two || L.any_bf2 ? std::max(L.tot_r, L.tot_s) * 4 : 0)) ||
        (rc = need(&bj_batch::tmp_val, L.any_two || L.any_bf2 ? std::max(L.tot_r, L.tot_s) * 4 : 0)) ||
        (rc = need(&bj_batch::part, L.part_n * 4)) ||
        (rc = need(&bj_batch::cur_rel, L.part_n * 4)) || (rc = need(&bj_batch::tot_r, nsb * 4)) ||
        (rc = need(&bj_batch::tot_s, nsb * 4)) || (rc = need(&bj_batch::start_r, nsb * 8)) ||
        (rc = need(&bj_batch::start_s, nsb * 8)) || (rc = need(&bj_batch::filter, L.f_blocks * 8)) ||
        (rc = need(&bj_batch::plan_cnt, nsb * 4)) || (rc = need(&bj_batch::item_start, (nsb + 1) * 8)) ||
        (rc = need(&bj_batch::partials, scan_partials_needed(nsb) * 8 + 64)) ||
        (rc = need(&bj_batch::items, L.item_cap * sizeof(HeavyItem))) ||
        (rc = need(&bj_batch::ctr, ctr_words(n) * 8)) || (rc = need(&bj_batch::out_key, L.out_n * 4)) ||
        (rc = need(&bj_batch::out_rval, L.out_n * 4)) || (rc = need(&bj_batch::out_sval, L.out_n * 4)))
        return rc;
    if (total_bytes) *total_bytes = total;
    return BJ_OK;
}

// The whole launch sequence of a batch on stream st (no host synchronisation,
// no memset: graph-capturable).
int enqueue(bj_batch* b, hipStream_t st)
{
    const Layout& L = b->L;
    const int n = b->n;
    BatchDev B;
    B.J = P<BJoin>(b->dJ);
    B.px = P<uint32_t>(b->dpx);
    B.sbj = B.px + (size_t)PX_N * (n + 1);
    B.n = n;
    B.max_bits = L.max_bits;
    B.max_b2 = L.max_b2;
    uint64_t* ctr = P<uint64_t>(b->ctr);
    uint64_t* jc = ctr;
    uint64_t* work = ctr + (size_t)n * JC_N;
    uint64_t* err = work + 2;
    const int num_cu = b->c->num_cu;
    // second passes over many tuples scatter with one workgroup per CU (as
    // bj_join does for large relations); first passes keep two per CU, which
    // hides their per-chunk cursor sums better
    const bool big_r = L.tot_r >= (32ull << 20), big_s = L.tot_s >= (32ull << 20);

    BJ_HIP(launch_zero_u64(ctr, (int)ctr_words(n), st));
    // R: one-level partition of every join into the R arena
    BJ_HIP(launch_batch_hist(B, 0, L.nt_r, nullptr, jc, P<uint32_t>(b->part), st));
    BJ_HIP(launch_batch_colscan(B, 0, L.ncol_r, P<uint32_t>(b->part), P<uint32_t>(b->cur_rel),
                                P<uint32_t>(b->tot_r), st));
    BJ_HIP(launch_batch_binscan(B, 0, P<uint32_t>(b->tot_r), P<uint64_t>(b->start_r), st));
    for (int k = 0; k < 2; k++)
        BJ_HIP(launch_batch_scatter(B, 0, k, L.nt_p1[0][k], L.max_b1[0][k], nullptr, nullptr, jc,
                                    P<uint32_t>(b->cur_rel), P<uint64_t>(b->start_r), P<uint32_t>(b->r_key),
                                    P<uint32_t>(b->r_val), b->r_key.bytes / 4, P<uint32_t>(b->tmp_key),
                                    P<uint32_t>(b->tmp_val), b->tmp_key.bytes / 4, err, k == 1, st));
    BJ_HIP(launch_batch_scatter2(B, 0, L.nt_r2, P<uint32_t>(b->cur_rel), P<uint64_t>(b->start_r),
                                 P<uint32_t>(b->tot_r), P<uint32_t>(b->tmp_key), P<uint32_t>(b->tmp_val),
                                 P<uint32_t>(b->r_key), P<uint32_t>(b->r_val), b->r_key.bytes / 4, err, big_r,
                                 st));
    // Bloom joins: filter from the partitioned R, then S -> S'
    if (L.any_bf) {
        BJ_HIP(launch_batch_bloom_build(B, L.nseg, L.max_wps, P<uint32_t>(b->r_key), P<uint64_t>(b->start_r),
                                        P<uint32_t>(b->tot_r), P<uint64_t>(b->filter), st));
        if (L.any_bf2) {
            // the joins with large filters: S by its top fb bits into the scratch
            // arena first (the S cursors are recomputed for S' below)
            BJ_HIP(launch_batch_hist(B, 2, L.nt_s, nullptr, jc, P<uint32_t>(b->part), st));
            BJ_HIP(launch_batch_colscan(B, 2, L.ncol_s, P<uint32_t>(b->part), P<uint32_t>(b->cur_rel),
                                        P<uint32_t>(b->tot_s), st));
            BJ_HIP(launch_batch_binscan(B, 2, P<uint32_t>(b->tot_s), P<uint64_t>(b->start_s), st));
            // (both classes in the 4096-tuple form: the 8192-tuple form measured slower
            // for this pass, 16 joins of 16M x 32M: 11.1 against 10.7 ms)
            for (int k = 0; k < 2; k++)
                BJ_HIP(launch_batch_scatter(B, 2, k, L.nt_p1[2][k], L.max_b1[2][k], nullptr, nullptr, jc,
                                            P<uint32_t>(b->cur_rel), P<uint64_t>(b->start_s),
                                            P<uint32_t>(b->tmp_key), P<uint32_t>(b->tmp_val),
                                            b->tmp_key.bytes / 4, P<uint32_t>(b->tmp_key),
                                            P<uint32_t>(b->tmp_val), b->tmp_key.bytes / 4, err, false, st));
        }
        BJ_HIP(launch_batch_prefilter(B, L.nf_tiles, P<uint64_t>(b->filter), P<uint32_t>(b->tmp_key),
                                      P<uint32_t>(b->tmp_val), P<uint32_t>(b->sp_key), P<uint32_t>(b->sp_val), jc,
                                      err + 1, err, num_cu, st));
    }
    // S (or S'): one-level partition into the S arena
    BJ_HIP(launch_batch_hist(B, 1, L.nt_s, P<uint32_t>(b->sp_key), jc, P<uint32_t>(b->part), st));
    BJ_HIP(launch_batch_colscan(B, 1, L.ncol_s, P<uint32_t>(b->part), P<uint32_t>(b->cur_rel),
                                P<uint32_t>(b->tot_s), st));
    BJ_HIP(launch_batch_binscan(B, 1, P<uint32_t>(b->tot_s), P<uint64_t>(b->start_s), st));
    for (int k = 0; k < 2; k++)
        BJ_HIP(launch_batch_scatter(B, 1, k, L.nt_p1[1][k], L.max_b1[1][k], P<uint32_t>(b->sp_key),
                                    P<uint32_t>(b->sp_val), jc, P<uint32_t>(b->cur_rel), P<uint64_t>(b->start_s),
                                    P<uint32_t>(b->s_key), P<uint32_t>(b->s_val), b->s_key.bytes / 4,
                                    P<uint32_t>(b->tmp_key), P<uint32_t>(b->tmp_val), b->tmp_key.bytes / 4, err,
                                    k == 1, st));
    BJ_HIP(launch_batch_scatter2(B, 1, L.nt_s2, P<uint32_t>(b->cur_rel), P<uint64_t>(b->start_s),
                                 P<uint32_t>(b->tot_s), P<uint32_t>(b->tmp_key), P<uint32_t>(b->tmp_val),
                                 P<uint32_t>(b->s_key), P<uint32_t>(b->s_val), b->s_key.bytes / 4, err, big_s,
                                 st));
    // vlsplit plan and the persistent build + probe, one launch per tile class
    BJ_HIP(launch_batch_plan(B, L.nsb, P<uint64_t>(b->start_r), P<uint32_t>(b->tot_r), P<uint64_t>(b->start_s),
                             P<uint32_t>(b->tot_s), B_SLICE, B_MAX_SLICES, P<uint32_t>(b->plan_cnt),
                             P<uint64_t>(b->item_start), P<uint64_t>(b->partials), P<HeavyItem>(b->items),
                             L.item_cap, err, st));
    BJ_HIP(launch_batch_probe(B, 8192, 0, L.nsb_a, P<uint32_t>(b->r_key), P<uint32_t>(b->r_val),
                              P<uint64_t>(b->start_r), P<uint32_t>(b->tot_r), P<uint32_t>(b->s_key),
                              P<uint32_t>(b->s_val), P<uint64_t>(b->start_s), P<uint32_t>(b->tot_s), B_SLICE,
                              P<HeavyItem>(b->items), P<uint64_t>(b->item_start), L.item_cap, work, jc,
                              P<uint32_t>(b->out_key), P<uint32_t>(b->out_rval), P<uint32_t>(b->out_sval),
                              num_cu, st));
    BJ_HIP(launch_batch_probe(B, 4096, L.nsb_a, L.nsb, P<uint32_t>(b->r_key), P<uint32_t>(b->r_val),
                              P<uint64_t>(b->start_r), P<uint32_t>(b->tot_r), P<uint32_t>(b->s_key),
                              P<uint32_t>(b->s_val), P<uint64_t>(b->start_s), P<uint32_t>(b->tot_s), B_SLICE,
                              P<HeavyItem>(b->items), P<uint64_t>(b->item_start), L.item_cap, work + 1, jc,
                              P<uint32_t>(b->out_key), P<uint32_t>(b->out_rval), P<uint32_t>(b->out_sval),
                              num_cu, st));
    return BJ_OK;
}

void fill_cfg(const JoinCfg& g, bj_join_out& o)
{
    o.radix_bits = g.bits;
    o.tile = g.tile;
    o.bf_k = g.bpk > 0.0 ? g.k : 0;
    o.bf_stage = g.bpk > 0.0 ? (g.bf2 ? 2 : 1) : 0;
    o.bf_pass_bits = g.bpk > 0.0 && g.bf2 ? g.fb : 0;
    o.bf_bits_per_key = g.bpk;
    o.capacity = g.cap;
}

}  // namespace

// ===================================================================== C ABI
extern "C" {

void bj_join_spec_default(bj_join_spec* s)
{
    if (!s) return;
    std::memset(s, 0, sizeof(*s));
    s->radix_bits = -1;
    s->tile = 0;
    s->bf_bits_per_key = -1.0;
    s->num_hashes = -1;
    s->answer_scale = 1.0;
    s->match_hint = -1.0;
}

int bj_batch_plan(const bj_ctx* c, int n, const bj_join_spec* specs, bj_join_out* cfg, uint64_t* workspace_bytes)
{
    return guarded([&]() -> int {
        Layout L;
        int rc = make_layout(c, n, specs, L);
        if (rc) return rc;
        if (cfg) {
            for (int i = 0; i < n; i++) {
                bj_join_out& o = cfg[L.order[i]];
                std::memset(&o, 0, sizeof(o));
                fill_cfg(L.cfg[i], o);
            }
        }
        if (workspace_bytes) batch_workspace(nullptr, L, n, workspace_bytes);
        return BJ_OK;
    });
}

int bj_batch_destroy(bj_batch* b)
{
    if (!b) return BJ_OK;
    if (b->c) {
        (void)hipSetDevice(b->c->device);
        if (b->c->stream) (void)hipStreamSynchronize(b->c->stream);
    }
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    if (b->e0) (void)hipEventDestroy(b->e0);
    if (b->e1) (void)hipEventDestroy(b->e1);
    DevBuf* bufs[] = {&b->dJ,      &b->dpx,     &b->r_key,   &b->r_val,    &b->s_key,      &b->s_val,
                      &b->sp_key,  &b->sp_val,  &b->tmp_key, &b->tmp_val,  &b->part,       &b->cur_rel,
                      &b->tot_r,   &b->tot_s,
                      &b->start_r, &b->start_s, &b->filter,  &b->plan_cnt, &b->item_start, &b->partials,
                      &b->items,   &b->ctr,     &b->out_key, &b->out_rval, &b->out_sval};
    for (DevBuf* d : bufs) release(*d);
    if (b->h_ctr) (void)hipHostFree(b->h_ctr);
    delete b;
    return BJ_OK;
}

int bj_batch_create(bj_ctx* c, int n, const bj_join_spec* specs, int use_graph, bj_batch** out)
{
    return guarded([&]() -> int {
        if (!c || !out) return fail(BJ_ERR_INVALID, "null argument");
        *out = nullptr;
        for (int i = 0; i < n && specs; i++) {
            const bj_join_spec& s = specs[i];
            if ((s.r_size > 0 && (!s.r_key || !s.r_val)) || (s.s_size > 0 && (!s.s_key || !s.s_val)))
                return fail(BJ_ERR_INVALID, "join " + std::to_string(i) + ": null relation pointer");
        }
        bj_batch* b = new bj_batch();
        // a throw past this point (the layout's vectors, the enqueue under
        // capture) ends an open capture and destroys the half-built batch
        try {
            b->c = c;
            b->n = n;
            int rc = make_layout(c, n, specs, b->L);
            if (rc) {
                delete b;
                return rc;
            }
            hipError_t e = hipSetDevice(c->device);
            if (e == hipSuccess) e = hipEventCreate(&b->e0);
            if (e == hipSuccess) e = hipEventCreate(&b->e1);
            if (e == hipSuccess) e = hipHostMalloc((void**)&b->h_ctr, ctr_words(n) * 8, hipHostMallocDefault);
            if (e != hipSuccess) {
                bj_batch_destroy(b);
                return fail(BJ_ERR_HIP, std::string("batch setup: ") + hipGetErrorString(e));
            }
            if ((rc = batch_workspace(b, b->L, n, nullptr))) {
                bj_batch_destroy(b);
                return rc;
            }
            hipStream_t st = c->stream;
            e = hipMemcpyAsync(b->dJ.p, b->L.J.data(), sizeof(BJoin) * n, hipMemcpyHostToDevice, st);
            if (e == hipSuccess)
                e = hipMemcpyAsync(b->dpx.p, b->L.px.data(), b->L.px.size() * 4, hipMemcpyHostToDevice, st);
            if (e == hipSuccess) e = hipStreamSynchronize(st);
            if (e != hipSuccess) {
                bj_batch_destroy(b);
                return fail(BJ_ERR_HIP, std::string("batch descriptors: ") + hipGetErrorString(e));
            }
            if (use_graph) {
                e = hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal);
                if (e != hipSuccess) {
                    bj_batch_destroy(b);
                    return fail(BJ_ERR_HIP, std::string("hipStreamBeginCapture: ") + hipGetErrorString(e));
                }
                rc = enqueue(b, st);
                hipGraph_t g = nullptr;
                e = hipStreamEndCapture(st, &g);
                b->graph = g;
                if (rc) {
                    bj_batch_destroy(b);
                    return rc;
                }
                if (e == hipSuccess) e = hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0);
                if (e != hipSuccess) {
                    bj_batch_destroy(b);
                    return fail(BJ_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(e));
                }
            }
        } catch (...) {
            hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
            if (hipStreamIsCapturing(c->stream, &cs) == hipSuccess && cs == hipStreamCaptureStatusActive) {
                hipGraph_t g = nullptr;
                (void)hipStreamEndCapture(c->stream, &g);
                if (g) (void)hipGraphDestroy(g);
            }
            bj_batch_destroy(b);
            throw;
        }
        *out = b;
        return BJ_OK;
    });
}

int bj_batch_run(bj_batch* b, bj_join_out* outs, double* device_ms)
{
    return guarded([&]() -> int {
        if (!b || (b->n > 0 && !outs)) return fail(BJ_ERR_INVALID, "null argument");
        bj_ctx* c = b->c;
        BJ_HIP(hipSetDevice(c->device));
        hipStream_t st = c->stream;
        BJ_HIP(hipEventRecord(b->e0, st));
        if (b->exec) {
            BJ_HIP(hipGraphLaunch(b->exec, st));
        } else {
            int rc = enqueue(b, st);
            if (rc) return rc;
        }
        BJ_HIP(hipEventRecord(b->e1, st));
        BJ_HIP(hipMemcpyAsync(b->h_ctr, b->ctr.p, ctr_words(b->n) * 8, hipMemcpyDeviceToHost, st));
        BJ_HIP(hipStreamSynchronize(st));
        const uint64_t err = b->h_ctr[(size_t)b->n * JC_N + 2];
        if (err) return fail(BJ_ERR_INTERNAL, "device check flags " + std::to_string(err));
        if (device_ms) {
            float ms = 0.f;
            BJ_HIP(hipEventElapsedTime(&ms, b->e0, b->e1));
            *device_ms = ms;
        }
        for (int i = 0; i < b->n; i++) {
            const BJoin& J = b->L.J[i];
            bj_join_out& o = outs[b->L.order[i]];
            std::memset(&o, 0, sizeof(o));
            fill_cfg(b->L.cfg[i], o);
            const uint64_t cnt = b->h_ctr[(size_t)i * JC_N + JC_OUT];
            o.match_count = cnt;
            o.rows_written = std::min(cnt, J.out_cap);
            o.truncated = cnt > J.out_cap ? 1 : 0;
            o.s_survivors = J.bf ? b->h_ctr[(size_t)i * JC_N + JC_SPRIME] : J.ns;
            o.key = P<uint32_t>(b->out_key) + J.out_base;
            o.r_val = P<uint32_t>(b->out_rval) + J.out_base;
            o.s_val = P<uint32_t>(b->out_sval) + J.out_base;
        }
        return BJ_OK;
    });
}

// ------------------------------------------------------------------ autotuner
int bj_tune_get(const bj_ctx* c, bj_tune_entry* out, int cap, int* n)
{
    if (!c || !n) return fail(BJ_ERR_INVALID, "null argument");
    *n = (int)c->tune.size();
    if (out)
        for (int i = 0; i < std::min(cap, *n); i++) out[i] = c->tune[i];
    return BJ_OK;
}

int bj_tune_set(bj_ctx* c, const bj_tune_entry* in, int n)
{
    return guarded([&]() -> int {
        if (!c || (n > 0 && !in) || n < 0) return fail(BJ_ERR_INVALID, "bad argument");
        for (int i = 0; i < n; i++) {
            const bj_tune_entry& e = in[i];
            if (e.kind == 0 && (e.radix_bits < 1 || e.radix_bits > B_MAX_BITS || (e.tile != 4096 && e.tile != 8192)))
                return fail(BJ_ERR_INVALID, "tune entry " + std::to_string(i) + ": bad radix bits or tile");
            if (e.kind == 1 && (e.bf_bits_per_key < 0.0 || e.bf_bits_per_key > 64.0))
                return fail(BJ_ERR_INVALID, "tune entry " + std::to_string(i) + ": bad filter bits per key");
            if (e.kind != 0 && e.kind != 1) return fail(BJ_ERR_INVALID, "tune entry kind must be 0 or 1");
        }
        c->tune.assign(in, in + n);
        return BJ_OK;
    });
}

}  // extern "C"

namespace {

// Synthetic tuning batch: m joins of |R| = nr, |S| = ns, a fraction `ratio` of
// S matching, generated on the device.
struct TuneData {
    DevBuf rk, rv, sk, sv;
    int m = 0;
    uint64_t nr = 0, ns = 0;
};

int gen_tune_data(bj_ctx* c, TuneData& d, int m, uint64_t nr, uint64_t ns, double ratio)
{
    int rc;
    if ((rc = ensure(d.rk, m * nr * 4)) || (rc = ensure(d.rv, m * nr * 4)) || (rc = ensure(d.sk, m * ns * 4)) ||
        (rc = ensure(d.sv, m * ns * 4)) || (rc = ensure(c->dtmp, 64)))
        return rc;
    d.m = m;
    d.nr = nr;
    d.ns = ns;
    const uint64_t ms = std::min<uint64_t>(ns, (uint64_t)std::llround(ratio * (double)ns));
    const int L = std::max(1, ceil_log2(ns));
    for (int j = 0; j < m; j++) {
        const uint64_t sr = 7000 + j, ss = sr ^ 0xDEADBEEFull;
        uint32_t* rk = P<uint32_t>(d.rk) + j * nr;
        uint32_t* rv = P<uint32_t>(d.rv) + j * nr;
        uint32_t* sk = P<uint32_t>(d.sk) + j * ns;
        uint32_t* sv = P<uint32_t>(d.sv) + j * ns;
        BJ_HIP(launch_gen_r(rk, rv, nr, perm_keys(sr), c->stream));
        BJ_HIP(launch_gen_s(sk, sv, ns, nr, ms, L, perm_keys(sr), perm_keys(ss ^ 0x1111111111111111ULL),
                            perm_keys(ss ^ 0x2222222222222222ULL), (uint32_t)splitmix64(ss ^ 0x3333333333333333ULL),
                            0.0, 0u, P<double>(c->dtmp), c->stream));
    }
    BJ_HIP(hipStreamSynchronize(c->stream));
    return BJ_OK;
}

// Best-of-3 device time of one configuration on the tuning batch.
int time_config(bj_ctx* c, TuneData& d, int bits, int tile, double bpk, double& ms_out)
{
    std::vector<bj_join_spec> specs(d.m);
    for (int j = 0; j < d.m; j++) {
        bj_join_spec& s = specs[j];
        bj_join_spec_default(&s);
        s.r_key = P<uint32_t>(d.rk) + j * d.nr;
        s.r_val = P<uint32_t>(d.rv) + j * d.nr;
        s.s_key = P<uint32_t>(d.sk) + j * d.ns;
        s.s_val = P<uint32_t>(d.sv) + j * d.ns;
        s.r_size = (int64_t)d.nr;
        s.s_size = (int64_t)d.ns;
        s.radix_bits = bits;
        s.tile = tile;
        s.bf_bits_per_key = bpk;
    }
    bj_batch* b = nullptr;
    int rc = bj_batch_create(c, d.m, specs.data(), 0, &b);
    if (rc) return rc;
    std::vector<bj_join_out> outs(d.m);
    double best = 1e30;
    for (int r = 0; r < 4 && rc == BJ_OK; r++) {
        double ms = 0.0;
        rc = bj_batch_run(b, outs.data(), &ms);
        if (r > 0) best = std::min(best, ms);
    }
    bj_batch_destroy(b);
    ms_out = best;
    return rc;
}

}  // namespace

extern "C" int bj_autotune(bj_ctx* c, int n, const bj_join_spec* specs, int* entries_tuned)
{
    return guarded([&]() -> int {
        if (!c || (n > 0 && !specs) || n < 0) return fail(BJ_ERR_INVALID, "bad argument");
        if (entries_tuned) *entries_tuned = 0;
        BJ_HIP(hipSetDevice(c->device));
        std::map<int, bool> lgs;
        std::map<std::pair<int, int>, bool> bfs;
        for (int i = 0; i < n; i++) {
            const bj_join_spec& s = specs[i];
            if (s.r_size <= 0 || s.s_size <= 0) continue;
            const uint64_t nr = (uint64_t)s.r_size, ns = (uint64_t)s.s_size;
            if (s.radix_bits <= 0 || s.tile == 0) {
                const int lg = lg_class(std::min(nr, ns));
                if (!find_tune(c, 0, lg, 0, 0)) lgs[lg] = true;
            }
            if (s.bf_bits_per_key < 0.0) {
                const int sr = sr_class(nr, ns), sel = sel_class(s.match_hint), fcls = filter_class(nr);
                if (!find_tune(c, 1, fcls, sr, sel)) bfs[{fcls * 16 + sr + 3, sel}] = true;
            }
        }
        TuneData d;
        int tuned = 0;
        int rc = BJ_OK;
        // radix bits and tile per build-size class, filter off
        for (const auto& kv : lgs) {
            const int lg = kv.first;
            const uint64_t nr = 1ull << lg, ns = 2ull << lg;
            const int m = (int)std::max<uint64_t>(1, std::min<uint64_t>(32, (1ull << 23) >> lg));
            if ((rc = gen_tune_data(c, d, m, nr, ns, 0.1))) break;
            bj_tune_entry best{};
            best.kind = 0;
            best.lg = lg;
            best.ms = 1e30;
            for (int tile : {8192, 4096}) {
                const int b0 = rule_bits(nr, tile);
                for (int bits = std::max(1, b0 - 1); bits <= std::min(B_MAX_BITS, b0 + 1); bits++) {
                    double ms = 0.0;
                    if ((rc = time_config(c, d, bits, tile, 0.0, ms))) break;
                    if (ms < best.ms) {
                        best.ms = ms;
                        best.radix_bits = bits;
                        best.tile = tile;
                    }
                }
                if (rc) break;
            }
            if (rc) break;
            c->tune.push_back(best);
            tuned++;
        }
        // filter bits per key per (filter class, |S|/|R|, selectivity) class: at
        // |R| = 2^20 for filters that stay in the L2, at 2^23 (two joins) for those
        // probed after a first pass of S
        for (const auto& kv : bfs) {
            if (rc) break;
            const int fcls = kv.first.first / 16, sr = kv.first.first % 16 - 3, sel = kv.first.second;
            const uint64_t nr = fcls ? (1ull << 23) : (1ull << 20);
            const uint64_t ns = sr >= 0 ? (nr << sr) : (nr >> -sr);
            if ((rc = gen_tune_data(c, d, fcls ? 2 : 8, nr, ns, SEL_REP[sel]))) break;
            const int lg = lg_class(std::min(nr, ns));
            const bj_tune_entry* t0 = find_tune(c, 0, lg, 0, 0);
            const int tile = t0 ? t0->tile : 8192;
            const int bits = t0 ? t0->radix_bits : rule_bits(std::min(nr, ns), tile);
            bj_tune_entry best{};
            best.kind = 1;
            best.lg = fcls;
            best.sr = sr;
            best.sel = sel;
            best.ms = 1e30;
            // a filter has to beat the join without one by 3%: near-ties otherwise
            // flip between runs, and the filter is the riskier choice for the
            // larger joins of the class
            for (double bpk : {0.0, 6.0, 10.0, 14.0}) {
                double ms = 0.0;
                if ((rc = time_config(c, d, bits, tile, bpk, ms))) break;
                if (bpk == 0.0 ? ms < best.ms : ms < best.ms * 0.97) {
                    best.ms = ms;
                    best.bf_bits_per_key = bpk;
                }
            }
            if (rc) break;
            c->tune.push_back(best);
            tuned++;
        }
        release(d.rk);
        release(d.rv);
        release(d.sk);
        release(d.sv);
        if (entries_tuned) *entries_tuned = tuned;
        return rc;
    });
}

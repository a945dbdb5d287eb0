"""The batched form: many joins in one launch sequence (hipGraph), autotuner.

CPU tests cover the batch planner (per-join configuration, validation,
footprint) without a device.  GPU tests hold every join of a batch to the same
bar as the single join: the exact count and the same sorted multiset of
(key, r_val, s_val) rows as the oracle (small joins) and as the single-join
path bj_join run one join at a time (larger joins).
"""
import math

import numpy as np
import pytest

import brigjoin as bj
import oracle as O


# ---------------------------------------------------------------- CPU: planner
def spec(nr, ns, **kw):
    return bj.join_spec(r_size=nr, s_size=ns, **kw)


def test_plan_size_rule_and_overrides():
    specs = [spec(1 << 16, 1 << 17), spec(16 << 20, 32 << 20), spec(5000, 100, tile=4096),
             spec(1 << 20, 1 << 20, radix_bits=3, bf_bits_per_key=0.0),
             spec(1 << 20, 4 << 20, bf_bits_per_key=12.0, num_hashes=4, answer_scale=2.0),
             spec(0, 1000), spec(1000, 0)]
    cfg, ws = bj.batch_plan(specs)
    # size rule: smallest bits with min(R,S) / 2^bits <= 0.6 * tile, at most 15
    assert cfg[0].radix_bits == math.ceil(math.log2((1 << 16) / (0.6 * 8192)))
    assert cfg[1].radix_bits == 12 and cfg[1].tile == 8192
    assert cfg[2].tile == 4096 and cfg[2].radix_bits == 1
    assert cfg[3].radix_bits == 3 and cfg[3].bf_bits_per_key == 0.0 and cfg[3].bf_k == 0
    assert cfg[4].bf_bits_per_key == 12.0 and cfg[4].bf_k == 4
    assert cfg[4].capacity == 2 * (4 << 20)
    # unknown selectivity: filter on at 10 bits per key, k = min(round(10 ln 2), 4) = 4
    assert cfg[0].bf_bits_per_key == 10.0 and cfg[0].bf_k == 4
    # an empty side needs no filter
    assert cfg[5].bf_k == 0 and cfg[6].bf_k == 0
    assert cfg[1].capacity == 32 << 20
    # the arenas alone: R and S (key + value), S' for filtered joins, three output columns
    tot_r = sum(s.r_size for s in specs)
    tot_s = sum(s.s_size for s in specs)
    assert ws > 8 * (tot_r + tot_s) + 12 * sum(c.capacity for c in cfg)


def test_plan_selectivity_hint_rule():
    cfg, _ = bj.batch_plan([spec(1 << 20, 1 << 21, match_hint=h) for h in (0.01, 0.1, 0.3, 0.6)])
    assert [c.bf_bits_per_key > 0 for c in cfg] == [True, True, False, False]


def test_plan_filter_stage_by_size():
    """A filter that fits an XCD's L2 (3.5 MB) is probed at random before S is
    partitioned (stage 1); a larger one after a first pass of S by enough top
    bits that each bucket probes a slice of at most 128 KB (stage 2), never more
    bits than the join's own radix bits."""
    cfg, _ = bj.batch_plan([spec(2_800_000, 4 << 20, match_hint=0.01),   # 3.5 MB at 10 bits/key
                            spec(3_000_000, 4 << 20, match_hint=0.01),   # 3.75 MB
                            spec(16 << 20, 16 << 20, match_hint=0.01),   # 20 MB
                            spec(16 << 20, 16 << 20, match_hint=0.01, bf_bits_per_key=10.0),
                            spec(16 << 20, 16 << 20, match_hint=0.01, bf_bits_per_key=10.0, radix_bits=7),
                            spec(16 << 20, 16 << 20, match_hint=0.5)])
    assert [c.bf_bits_per_key for c in cfg] == [10.0, 10.0, 10.0, 10.0, 10.0, 0.0]
    assert [c.bf_stage for c in cfg] == [1, 2, 2, 2, 2, 0]
    assert cfg[1].bf_pass_bits == math.ceil(math.log2(3_000_000 * 10 / 8 / (128 << 10)))
    assert cfg[2].bf_pass_bits == math.ceil(math.log2((16 << 20) * 10 / 8 / (128 << 10)))
    assert cfg[4].bf_pass_bits == 7 and cfg[4].radix_bits == 7
    assert [c.bf_pass_bits for c in (cfg[0], cfg[5])] == [0, 0]


@pytest.mark.parametrize("kw,code", [
    (dict(tile=1000), -1), (dict(radix_bits=16), -5), (dict(num_hashes=9, bf_bits_per_key=8.0), -1),
    (dict(bf_bits_per_key=float("nan")), -1), (dict(bf_bits_per_key=float("inf")), -1),
    (dict(answer_scale=float("inf")), -1), (dict(answer_scale=1e30), -5),
])
def test_plan_rejects_bad_specs(kw, code):
    with pytest.raises(bj.BjError) as e:
        bj.batch_plan([spec(1000, 1000), spec(1000, 1000, **kw)])
    assert e.value.code == code
    assert "join 1" in str(e.value)


def test_plan_rejects_sizes():
    with pytest.raises(bj.BjError) as e:
        bj.batch_plan([spec(-1, 10)])
    assert e.value.code == -1
    with pytest.raises(bj.BjError) as e:
        bj.batch_plan([spec(1 << 31, 10)])
    assert e.value.code == -5


def test_plan_empty_batch():
    cfg, ws = bj.batch_plan([])
    assert cfg == [] and ws > 0


# ---------------------------------------------------------------- GPU helpers
def _dev(a):
    import torch
    a = np.ascontiguousarray(a, np.uint32)
    if len(a) == 0:
        return torch.zeros(1, dtype=torch.int32, device="cuda")
    return torch.from_numpy(a.view(np.int32)).to("cuda")


class HostJoin:
    """One join's host inputs and their device copies (kept alive with the batch)."""

    def __init__(self, rk, rv, sk, sv, **kw):
        self.h = [np.ascontiguousarray(x, np.uint32) for x in (rk, rv, sk, sv)]
        self.d = [_dev(x) for x in self.h]
        self.kw = kw

    def spec(self):
        return bj.join_spec(*self.d, r_size=len(self.h[0]), s_size=len(self.h[2]), **self.kw)


def _rows(batch, out):
    k, a, b = batch.rows_to_host(out)
    return O.sort_rows(k, a, b)


def _mixed_small_joins():
    rng = np.random.default_rng(11)
    joins = []
    # duplicates on both sides, assorted bits / tiles / filters
    # radix bits above 9 take two scatter passes
    for i, kw in enumerate([dict(), dict(bf_bits_per_key=10.0), dict(radix_bits=1), dict(radix_bits=11),
                            dict(tile=4096, radix_bits=4), dict(bf_bits_per_key=4.0, num_hashes=2),
                            dict(tile=4096, bf_bits_per_key=8.0), dict(radix_bits=7),
                            dict(radix_bits=14, bf_bits_per_key=10.0), dict(radix_bits=10, tile=4096)]):
        nr, ns = int(rng.integers(1, 30000)), int(rng.integers(1, 60000))
        rk = rng.integers(0, 20000, nr, dtype=np.uint32)
        sk = rng.integers(0, 30000, ns, dtype=np.uint32)
        joins.append(HostJoin(rk, np.arange(nr, dtype=np.uint32) + 7 * i, sk,
                              np.arange(ns, dtype=np.uint32) * 3 + i, **kw))
    # special keys, empty sides, a heavy key beyond one LDS tile
    sp = np.array([0, 1, 0x80000000, 0xFFFFFFFF, 0x7FFFFFFF], np.uint32)
    # 75 rows: more than max(|R|, |S|), so the capacity is raised
    joins.append(HostJoin(np.tile(sp, 3), np.arange(15, dtype=np.uint32), np.tile(sp, 5),
                          np.arange(25, dtype=np.uint32), answer_scale=4.0))
    joins.append(HostJoin(np.zeros(0, np.uint32), np.zeros(0, np.uint32), sp, np.arange(5, dtype=np.uint32)))
    joins.append(HostJoin(sp, np.arange(5, dtype=np.uint32), np.zeros(0, np.uint32), np.zeros(0, np.uint32)))
    hk = np.full(20000, 12345, np.uint32)
    hk[::3] = np.arange(len(hk[::3]), dtype=np.uint32)
    joins.append(HostJoin(hk, np.arange(20000, dtype=np.uint32), np.array([12345, 1, 2], np.uint32),
                          np.arange(3, dtype=np.uint32)))
    return joins


@pytest.mark.gpu
@pytest.mark.parametrize("use_graph", [1, 0])
def test_batch_matches_oracle(gpu_ctx, use_graph):
    joins = _mixed_small_joins()
    print(f"[batch oracle] joins={len(joins)} graph={use_graph}", flush=True)
    with bj.Batch(gpu_ctx, [j.spec() for j in joins], use_graph=bool(use_graph)) as b:
        outs, ms = b.run()
        assert ms > 0
        for i, (j, o) in enumerate(zip(joins, outs)):
            n, k, a, c = O.cpu_join(*j.h)
            assert o.match_count == n, i
            assert not o.truncated, i
            assert np.array_equal(_rows(b, o), O.sort_rows(k, a, c)), i
            if o.bf_k:
                assert o.s_survivors <= len(j.h[2])


@pytest.mark.gpu
@pytest.mark.parametrize("use_graph", [1, 0])
def test_batch_large_filters_after_first_pass(gpu_ctx, use_graph):
    """Filters larger than an XCD's L2 are probed after a first pass of S by the
    top bits (stage 2), next to stage-1 and unfiltered joins of the same batch;
    every join's rows equal the oracle's, repeated S keys included."""
    joins = []
    for i, (nr, ns, kw) in enumerate([
            (3_000_000, 6_000_000, dict(bf_bits_per_key=10.0)),                 # 3.75 MB: stage 2
            (4_500_000, 5_000_000, dict(bf_bits_per_key=12.0, num_hashes=4)),   # 6.75 MB
            (20_000, 50_000, dict(bf_bits_per_key=10.0)),                       # stage 1
            (3_200_000, 3_300_000, dict(bf_bits_per_key=10.0, tile=4096)),
            (100_000, 400_000, dict(bf_bits_per_key=0.0))]):                    # no filter
        rk, rv = O.gen_r(300 + i, nr)
        m = O.match_count(ns, 0.05)
        sk, sv = O.gen_s(300 + i, 301 + i, nr, ns, m)
        if i == 1:
            # a hot key: 40000 matching S positions carry R's key of row 77
            pos = np.nonzero(sk < np.uint32(1 << 31))[0][:40000]
            sk[pos] = rk[77]
        joins.append(HostJoin(rk, rv, sk, sv, **kw))
    print(f"[batch stage 2] joins={len(joins)} graph={use_graph}", flush=True)
    with bj.Batch(gpu_ctx, [j.spec() for j in joins], use_graph=bool(use_graph)) as b:
        for rep in range(2):
            outs, _ = b.run()
            assert [o.bf_stage for o in outs] == [2, 2, 1, 2, 0]
            assert all(o.bf_pass_bits >= 1 for o in (outs[0], outs[1], outs[3]))
            for i, (j, o) in enumerate(zip(joins, outs)):
                n, k, a, c = O.cpu_join(*j.h)
                assert o.match_count == n and not o.truncated, (rep, i)
                assert np.array_equal(_rows(b, o), O.sort_rows(k, a, c)), (rep, i)
                if o.bf_k:
                    assert n <= o.s_survivors <= len(j.h[2]), (rep, i)


@pytest.mark.gpu
def test_batch_graph_replays(gpu_ctx):
    import torch
    joins = _mixed_small_joins()[:6]
    expect = [O.sort_rows(*O.cpu_join(*j.h)[1:]) for j in joins]
    with bj.Batch(gpu_ctx, [j.spec() for j in joins], use_graph=True) as b:
        for rep in range(4):
            # eager device work between replays
            x = torch.arange(1 << 20, device="cuda", dtype=torch.int32)
            assert int(x.sum().item()) == (1 << 20) * ((1 << 20) - 1) // 2
            outs, _ = b.run()
            for i, o in enumerate(outs):
                assert np.array_equal(_rows(b, o), expect[i]), (rep, i)


def _gen_join(ctx, seed, nr, ns, ratio, zipf=0.0):
    import torch
    rk, rv = torch.empty(max(nr, 1), dtype=torch.int32, device="cuda"), torch.empty(max(nr, 1), dtype=torch.int32, device="cuda")
    sk, sv = torch.empty(max(ns, 1), dtype=torch.int32, device="cuda"), torch.empty(max(ns, 1), dtype=torch.int32, device="cuda")
    ctx.gen_build(seed, nr, rk, rv)
    m = O.match_count(ns, ratio)
    ctx.gen_probe(seed, seed ^ 0xDEADBEEF, nr, ns, m, sk, sv, zipf_s=zipf)
    return (rk, rv, sk, sv), m


@pytest.mark.gpu
def test_batch_matches_single_join(gpu_ctx):
    """Generated joins of uneven sizes: each batch result equals bj_join's on the same inputs."""
    rng = np.random.default_rng(3)
    specs, data, ms_ = [], [], []
    for i in range(14):
        nr = int(math.exp(rng.uniform(math.log(1 << 14), math.log(3 << 20))))
        ns = int(nr * rng.uniform(1.0, 4.0))
        ratio = float(rng.uniform(0.01, 0.5))
        d, m = _gen_join(gpu_ctx, 100 + i, nr, ns, ratio)
        data.append(d)
        ms_.append(m)
        specs.append(bj.join_spec(*d, r_size=nr, s_size=ns, match_hint=ratio))
    print(f"[batch vs single] joins={len(specs)}", flush=True)
    with bj.Batch(gpu_ctx, specs, use_graph=True) as b:
        outs, _ = b.run()
        for i, o in enumerate(outs):
            assert o.match_count == ms_[i], i
            rows_b = _rows(b, o)
            p = bj.default_params(r_size=specs[i].r_size, s_size=specs[i].s_size, inputs_on_device=1)
            res = gpu_ctx.join(p, *data[i])
            assert res.match_count == o.match_count
            k, a, c = gpu_ctx.rows_to_host(res)
            assert np.array_equal(rows_b, O.sort_rows(k, a, c)), i
            # every row is a real match: values are row indices
            rk = data[i][0][:specs[i].r_size].cpu().numpy().view(np.uint32)
            sk = data[i][2][:specs[i].s_size].cpu().numpy().view(np.uint32)
            assert np.array_equal(rk[rows_b[:, 1]], rows_b[:, 0])
            assert np.array_equal(sk[rows_b[:, 2]], rows_b[:, 0])
            assert len(np.unique(rows_b[:, 2])) == len(rows_b)


@pytest.mark.gpu
def test_batch_skew_and_truncation(gpu_ctx):
    """Zipf-skewed S (vlsplit items across joins) and a join past its row capacity."""
    d1, m1 = _gen_join(gpu_ctx, 5, 200_000, 3_000_000, 0.2, zipf=1.0)
    d2, m2 = _gen_join(gpu_ctx, 6, 50_000, 400_000, 0.5, zipf=1.2)
    rk = np.full(3000, 9, np.uint32)
    sk = np.full(4000, 9, np.uint32)
    tj = HostJoin(rk, np.arange(3000, dtype=np.uint32), sk, np.arange(4000, dtype=np.uint32))
    specs = [bj.join_spec(*d1, r_size=200_000, s_size=3_000_000, bf_bits_per_key=10.0),
             bj.join_spec(*d2, r_size=50_000, s_size=400_000, tile=4096, bf_bits_per_key=0.0),
             tj.spec()]
    with bj.Batch(gpu_ctx, specs) as b:
        outs, _ = b.run()
        for i, d in enumerate((d1, d2)):
            h = [x[:n].cpu().numpy().view(np.uint32) for x, n in
                 zip(d, (specs[i].r_size, specs[i].r_size, specs[i].s_size, specs[i].s_size))]
            n, k, a, c = O.cpu_join(*h)
            assert outs[i].match_count == n == (m1, m2)[i]
            assert np.array_equal(_rows(b, outs[i]), O.sort_rows(k, a, c))
        t = outs[2]
        assert t.match_count == 3000 * 4000 and t.truncated == 1 and t.rows_written == t.capacity == 4000
        rows = _rows(b, t)
        assert np.all(rows[:, 0] == 9) and len(np.unique(rows[:, 1:], axis=0)) == len(rows)


@pytest.mark.gpu
def test_autotune_and_tuned_batch(gpu_ctx):
    specs = []
    data = []
    for i, (nr, ns, ratio) in enumerate([(1 << 16, 1 << 17, 0.05), (1 << 18, 1 << 19, 0.3)]):
        d, m = _gen_join(gpu_ctx, 200 + i, nr, ns, ratio)
        data.append((d, m))
        specs.append(bj.join_spec(*d, r_size=nr, s_size=ns, match_hint=ratio))
    n = bj.autotune(gpu_ctx, specs)
    table = bj.tune_get(gpu_ctx)
    assert n == len(table) >= 3          # two size classes, at least one filter class
    for e in table:
        assert e.kind in (0, 1) and e.ms > 0
        if e.kind == 0:
            assert 1 <= e.radix_bits <= 15 and e.tile in (4096, 8192)
        else:
            assert e.bf_bits_per_key in (0.0, 6.0, 10.0, 14.0)
    # classes already tuned are kept
    assert bj.autotune(gpu_ctx, specs) == 0
    cfg, _ = bj.batch_plan(specs, ctx=gpu_ctx)
    kinds0 = {e.lg: e for e in table if e.kind == 0}
    assert cfg[0].radix_bits == kinds0[16].radix_bits and cfg[0].tile == kinds0[16].tile
    with bj.Batch(gpu_ctx, specs) as b:
        outs, _ = b.run()
        for (d, m), o in zip(data, outs):
            assert o.match_count == m
    # the table round-trips, and bad entries are refused
    bj.tune_set(gpu_ctx, table)
    assert len(bj.tune_get(gpu_ctx)) == len(table)
    bad = bj.BjTuneEntry(kind=0, lg=16, radix_bits=16, tile=8192)
    with pytest.raises(bj.BjError):
        bj.tune_set(gpu_ctx, [bad])
    bj.tune_set(gpu_ctx, [])
    assert bj.tune_get(gpu_ctx) == []

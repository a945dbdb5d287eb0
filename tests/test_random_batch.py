"""Seeded batches of joins with random sizes and specs against the oracle (GPU tests).

Each batch mixes joins of random sizes (empty sides included) and key shapes
(distinct, duplicate-heavy, one hot S key) with random specs: radix bits
(one-pass joins of 7-9 bits take the wide first-pass launch, others the
narrow one, 10-15 bits two passes), LDS tile, Bloom bits per key and hash
count (filters that fit an XCD's L2 test S in its order, larger ones after a
first pass of S), a match hint.  Each batch runs twice (the second run replays
the captured graph); every join's rows must equal the oracle's CPU join
(src/sanity_check.cpp:10-73 extended to rows) both times.
"""
import numpy as np
import pytest

import brigjoin as bj
import oracle as O

pytestmark = pytest.mark.gpu


def _dev(a):
    import torch
    a = np.ascontiguousarray(a, np.uint32)
    if len(a) == 0:
        return torch.zeros(1, dtype=torch.int32, device="cuda")
    return torch.from_numpy(a.view(np.int32)).to("cuda")


def _join(rng, stage2=False):
    shape = 0 if stage2 else rng.integers(0, 3)
    big = rng.random() < 0.15
    nr = int(rng.integers(0, 3_500_000 if big else 60_000))
    ns = int(rng.integers(0, 4_000_000 if big else 90_000))
    if stage2:
        # a filter above an XCD's L2 share: tested after a first pass of S
        nr, ns = int(rng.integers(2_600_000, 3_500_000)), int(rng.integers(2_000_000, 5_000_000))
    if shape == 0:
        rk = rng.integers(0, 1 << 32, size=nr, dtype=np.uint64).astype(np.uint32)
        sk = rng.integers(0, 1 << 32, size=ns, dtype=np.uint64).astype(np.uint32)
        if nr and ns:
            take = rng.random(ns) < rng.uniform(0.01, 0.5)
            sk[take] = rk[rng.integers(0, nr, size=int(take.sum()))]
    elif shape == 1:
        nr, ns = min(nr, 20_000), min(ns, 20_000)
        dom = int(rng.integers(100, 4000))
        rk = rng.integers(0, dom, size=nr, dtype=np.uint32)
        sk = rng.integers(0, 2 * dom, size=ns, dtype=np.uint32)
    else:
        rk = (rng.permutation(max(nr, 1)).astype(np.uint32) * np.uint32(2246822519))[:nr]
        sk = rng.integers(0, 1 << 32, size=ns, dtype=np.uint64).astype(np.uint32)
        if nr and ns:
            sk[rng.random(ns) < 0.2] = rk[int(rng.integers(0, nr))]
    rv = rng.integers(0, 1 << 32, size=nr, dtype=np.uint64).astype(np.uint32)
    sv = rng.integers(0, 1 << 32, size=ns, dtype=np.uint64).astype(np.uint32)
    kw = {}
    if rng.random() < 0.6:
        kw["radix_bits"] = int(rng.integers(1, 16))
    if rng.random() < 0.5:
        kw["tile"] = int(rng.choice([4096, 8192]))
    u = rng.random()
    if stage2:
        kw["bf_bits_per_key"] = float(rng.uniform(12.0, 16.0))
    elif u < 0.45:
        kw["bf_bits_per_key"] = float(rng.uniform(4.0, 14.0))
        if rng.random() < 0.5:
            kw["num_hashes"] = int(rng.integers(1, 7))
    elif u < 0.7:
        kw["bf_bits_per_key"] = 0.0
    if rng.random() < 0.5:
        kw["match_hint"] = float(rng.uniform(0.0, 0.6))
    return (rk, rv, sk, sv), kw


@pytest.mark.parametrize("seed", range(6))
def test_random_batch(gpu_ctx, seed):
    import torch
    rng = np.random.default_rng(500 + seed)
    joins = [_join(rng) for _ in range(int(rng.integers(4, 24)))]
    if seed % 2 == 0:
        joins.insert(int(rng.integers(0, len(joins) + 1)), _join(rng, stage2=True))
    expect = []
    specs, keep = [], []
    for h, kw in joins:
        n, k, a, c = O.cpu_join(*h)
        expect.append((n, O.sort_rows(k, a, c)))
        # room for every row (truncation has its own test)
        scale = max(1.0, float(n) / max(1, len(h[0]), len(h[2])) + 1.0)
        d = [_dev(x) for x in h]
        keep.append(d)
        specs.append(bj.join_spec(*d, r_size=len(h[0]), s_size=len(h[2]), answer_scale=scale, **kw))
    torch.cuda.synchronize()
    cfg, _ = bj.batch_plan(specs, gpu_ctx)
    print(f"[random batch] seed {seed}: {len(joins)} joins, bits {[c.radix_bits for c in cfg]}, "
          f"filter stages {[c.bf_stage for c in cfg]}", flush=True)
    with bj.Batch(gpu_ctx, specs, use_graph=True) as b:
        for rep in range(2):
            outs, _ = b.run()
            for i, ((n, rows), o) in enumerate(zip(expect, outs)):
                assert o.match_count == n and not o.truncated, (rep, i)
                k, a, c = b.rows_to_host(o)
                assert np.array_equal(O.sort_rows(k, a, c), rows), (rep, i)


@pytest.mark.parametrize("use_graph", [1, 0])
def test_empty_and_degenerate_batches(gpu_ctx, use_graph):
    """No joins at all, and joins whose sides are all empty, run and report nothing."""
    with bj.Batch(gpu_ctx, [], use_graph=bool(use_graph)) as b:
        for _ in range(2):
            outs, _ = b.run()
            assert outs == []
    z = _dev(np.zeros(0, np.uint32))
    one = _dev(np.array([5], np.uint32))
    specs = [bj.join_spec(z, z, z, z, r_size=0, s_size=0),
             bj.join_spec(one, one, z, z, r_size=1, s_size=0, bf_bits_per_key=10.0),
             bj.join_spec(z, z, one, one, r_size=0, s_size=1, radix_bits=9)]
    with bj.Batch(gpu_ctx, specs, use_graph=bool(use_graph)) as b:
        for _ in range(2):
            outs, _ = b.run()
            assert [o.match_count for o in outs] == [0, 0, 0]
            assert all(o.rows_written == 0 and not o.truncated for o in outs)

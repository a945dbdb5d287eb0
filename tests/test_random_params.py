"""Seeded sweep of the join's parameter surface against the oracle (GPU tests).

Every case draws inputs of one of three shapes (distinct keys, a small key
domain with duplicates on both sides, a few very hot S keys) and a random
combination of the plan's options: radix bits or fan-outs, LDS tile, Bloom
filter (bits per key, hash count, stage), first-pass bits of R and S, scatter
workgroup size, S in chunks (pooled or per-chunk probe) and host or device
inputs.  Combinations the planner refuses (for example first-pass bits above
the total) are checked to be refused by bj_plan as well, then skipped.

Bar: the same match count and the same sorted multiset of (key, r_val, s_val)
rows as the oracle's CPU join (restated from the reference's sanity_check,
src/sanity_check.cpp:10-73, extended to rows; every match per S tuple as in
src/kernel_functions.cu:399-430).
"""
import numpy as np
import pytest

import brigjoin as bj
import oracle as O

pytestmark = pytest.mark.gpu

N_CASES = 64


def _keys(rng, shape):
    """R and S keys (uint32) of one of three shapes, and their row values."""
    nr = int(rng.integers(0, 120_000))
    ns = int(rng.integers(0, 160_000))
    if shape == "distinct":
        rk = rng.integers(0, 1 << 32, size=nr, dtype=np.uint64).astype(np.uint32)
        # about half of S from R, the rest anywhere (0 and 2^31 included now and then)
        take = rng.random(ns) < 0.5
        sk = rng.integers(0, 1 << 32, size=ns, dtype=np.uint64).astype(np.uint32)
        if nr:
            sk[take] = rk[rng.integers(0, nr, size=int(take.sum()))]
        if ns:
            sk[rng.integers(0, ns, size=min(ns, 3))] = np.array([0, 0x80000000, 0xFFFFFFFF], np.uint32)[:min(ns, 3)]
    elif shape == "dups":
        dom = int(rng.integers(50, 5000))
        nr = min(nr, 20_000)
        ns = min(ns, 20_000)
        rk = rng.integers(0, dom, size=nr, dtype=np.uint32)
        sk = rng.integers(0, 2 * dom, size=ns, dtype=np.uint32)
    else:  # hot S keys over distinct R keys
        rk = (rng.permutation(max(nr, 1)).astype(np.uint32) * np.uint32(2654435761))[:nr]
        sk = rng.integers(0, 1 << 32, size=ns, dtype=np.uint64).astype(np.uint32)
        if nr and ns:
            hot = rk[rng.integers(0, nr, size=3)]
            pick = rng.random(ns)
            sk[pick < 0.3] = hot[0]
            sk[(pick >= 0.3) & (pick < 0.4)] = hot[1]
            sk[(pick >= 0.4) & (pick < 0.45)] = hot[2]
    rv = rng.integers(0, 1 << 32, size=nr, dtype=np.uint64).astype(np.uint32)
    sv = rng.integers(0, 1 << 32, size=ns, dtype=np.uint64).astype(np.uint32)
    return rk, rv, sk, sv


def _params(rng):
    kw = {}
    if rng.random() < 0.6:
        kw["radix_bits"] = int(rng.integers(1, 16))
    elif rng.random() < 0.5:
        kw["partition_size"] = 1 << int(rng.integers(1, 11))
    kw["tile"] = int(rng.choice([0, 4096, 8192]))
    if rng.random() < 0.6:
        kw["use_bf"] = 1
        kw["bf_bits_per_key"] = float(rng.uniform(3.0, 14.0))
        if rng.random() < 0.5:
            kw["num_hashes"] = int(rng.integers(1, 7))
        kw["bf_stage"] = int(rng.choice([0, 1, 2]))
    # first-pass bits within the total when it is given (now and then above it, or
    # with the total left to the planner: cases the planner may refuse)
    top = kw.get("radix_bits", 11) + (1 if rng.random() < 0.1 else 0)
    if rng.random() < 0.3:
        kw["pass1_bits"] = int(rng.integers(1, min(top, 11) + 1))
    if rng.random() < 0.3:
        kw["s_pass1_bits"] = int(rng.integers(1, min(top, 11) + 1))
    kw["scatter_wg"] = int(rng.choice([0, 512, 1024]))
    if rng.random() < 0.35:
        kw["s_chunks"] = int(rng.integers(2, 5))
        kw["chunk_probe"] = int(rng.integers(0, 2))
    return kw


def _as_input(a, on_device):
    import torch
    a = np.ascontiguousarray(a, np.uint32)
    t = torch.from_numpy((a if len(a) else np.zeros(1, np.uint32)).view(np.int32))
    return t.to("cuda") if on_device else t


@pytest.mark.parametrize("case", range(N_CASES))
def test_random_parameters(gpu_ctx, case):
    import torch
    rng = np.random.default_rng(1000 + case)
    shape = ("distinct", "dups", "hot")[case % 3]
    rk, rv, sk, sv = _keys(rng, shape)
    kw = _params(rng)
    on_device = bool(rng.random() < 0.75)
    n, k0, a0, b0 = O.cpu_join(rk, rv, sk, sv)
    expect = O.sort_rows(k0, a0, b0)
    # room for every row (truncation is covered by its own tests)
    scale = max(1.0, float(n) / max(1, len(rk), len(sk)) + 1.0)
    p = bj.default_params(r_size=len(rk), s_size=len(sk), inputs_on_device=int(on_device),
                          answer_scale=scale, **kw)
    print(f"[random] case {case} {shape} |R|={len(rk)} |S|={len(sk)} rows={n} device={on_device} {kw}",
          flush=True)
    try:
        bj.plan(p)
    except bj.BjError as e:
        # refused by the planner: the join must refuse it the same way
        d = [_as_input(x, on_device) for x in (rk, rv, sk, sv)]
        with pytest.raises(bj.BjError) as e2:
            gpu_ctx.join(p, *d)
        assert e2.value.code == e.code
        pytest.skip(f"combination refused by the planner: {e}")
    d = [_as_input(x, on_device) for x in (rk, rv, sk, sv)]
    torch.cuda.synchronize()
    res = gpu_ctx.join(p, *d)
    assert res.match_count == n
    assert res.truncated == 0 and res.rows_written == n
    k, a, b = gpu_ctx.rows_to_host(res)
    assert np.array_equal(O.sort_rows(k, a, b), expect)

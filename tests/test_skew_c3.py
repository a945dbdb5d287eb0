"""The skewed, selective workload (GPU tests): R = 64M distinct keys, S = 1.28B
keys of which 1% match, the matching keys Zipf(1.0) over R's rows, filter at
12 bits/key with k = 4.

At full size the oracle cannot join 1.34 billion tuples in seconds, so parity
rests on properties that do not depend on the size (R's keys are distinct and
below 2^31, S's matching keys are drawn from R, its other keys are at or above
2^31):
- the count is exactly round(0.01 * |S|);
- the rows' S positions are exactly the positions whose key is below 2^31,
  each once;
- every row satisfies S_key[s_val] == key == R_key[r_val] (values are row
  indices).
With distinct R keys these determine the row multiset.  The hot keys' probe
sides (the top key alone has about 688K tuples) are split into vlsplit slices
there; the oracle-sized case below puts more than 64K probe tuples on one key
at 1% match and is checked row for row against the oracle, with the filter in
both stages.  Every S tuple emits each of its matches (src/kernel_functions.cu:
399-430).
"""
import numpy as np
import pytest

import brigjoin as bj
import oracle as O

pytestmark = pytest.mark.gpu

C3 = dict(use_bf=1, bf_bits_per_key=12.0, num_hashes=4)


def test_c3_full_size(gpu_ctx):
    import torch
    n_r, n_s = 64_000_000, 1_280_000_000
    m = O.match_count(n_s, 0.01)
    rk, rv = (torch.empty(n_r, dtype=torch.int32, device="cuda") for _ in range(2))
    sk, sv = (torch.empty(n_s, dtype=torch.int32, device="cuda") for _ in range(2))
    gpu_ctx.gen_build(42, n_r, rk, rv)
    gpu_ctx.gen_probe(42, 42 ^ 0xDEADBEEF, n_r, n_s, m, sk, sv, zipf_s=1.0)
    torch.cuda.synchronize()
    p = bj.default_params(r_size=n_r, s_size=n_s, inputs_on_device=1, **C3)
    print(f"[c3] n_r={n_r} n_s={n_s} m={m} {C3}", flush=True)
    res = gpu_ctx.join(p, rk, rv, sk, sv)
    assert res.match_count == m and res.rows_written == m and res.truncated == 0
    assert res.heavy_items > 0            # the hot keys ran as vlsplit items
    k, a, b = gpu_ctx.rows_to_host(res)
    rkh = rk.cpu().numpy().view(np.uint32)
    del rk, rv, sv
    skh = sk.cpu().numpy().view(np.uint32)
    del sk
    torch.cuda.empty_cache()
    seen = np.zeros(n_s, dtype=bool)
    seen[b] = True
    assert int(seen.sum()) == m                                   # no S position twice
    assert np.array_equal(seen, skh < np.uint32(1 << 31))         # exactly the matching positions
    assert np.array_equal(skh[b], k)
    assert np.array_equal(rkh[a], k)
    # the skew is there: the most frequent key carries a large share of the rows
    _, counts = np.unique(k, return_counts=True)
    assert counts.max() > 64 * 1024


def skewed_inputs(n_r=1_000_000, n_s=10_000_000, hot=70_000):
    """1% of S matches: `hot` tuples of one R key, the rest uniform over R."""
    m = O.match_count(n_s, 0.01)
    rk, rv = O.gen_r(11, n_r)
    sk, sv = O.gen_s(11, 12, n_r, n_s, m)
    match_pos = np.nonzero(sk < np.uint32(1 << 31))[0]
    assert len(match_pos) == m
    sk[match_pos[:hot]] = rk[12345]
    return rk, rv, sk, sv, m


@pytest.mark.parametrize("stage", [1, 2])
def test_heavy_key_over_slice_at_1pct(gpu_ctx, stage):
    import torch
    rk, rv, sk, sv, m = skewed_inputs()
    p = bj.default_params(r_size=len(rk), s_size=len(sk), inputs_on_device=1, bf_stage=stage, **C3)
    d = [torch.from_numpy(x.view(np.int32)).cuda() for x in (rk, rv, sk, sv)]
    torch.cuda.synchronize()
    print(f"[heavy] |R|={len(rk)} |S|={len(sk)} stage={stage}", flush=True)
    res = gpu_ctx.join(p, *d)
    n, ek, ea, eb = O.cpu_join(rk, rv, sk, sv)
    assert res.match_count == n == m
    assert res.heavy_items >= 2           # the hot key's probe side was split
    got = O.sort_rows(*gpu_ctx.rows_to_host(res))
    assert np.array_equal(got, O.sort_rows(ek, ea, eb))

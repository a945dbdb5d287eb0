"""Joins at the largest supported relation size, 2^31 - 1 tuples (GPU tests).

A relation of 2^31 - 1 tuples is the planner's limit (row ids are 31-bit in
the probe's staging); every position, cursor and count above it is 64-bit.
These joins put that limit on each side in turn, with the filter on (a
first-level pass of S before the filter, vlsplit items on the large side), so
that a 32-bit overflow anywhere on the path would show as a wrong row.

The oracle cannot join these sizes in seconds, so the bar is the size-free
property check of the 640M test: R's keys are distinct and below 2^31, S's
matching keys come from R and its other keys are at or above 2^31, values are
row indices.  Then the count is round(0.1 * |S|), the rows' S positions are
exactly the positions with a key below 2^31, each once, and every row has
S_key[s_val] == key == R_key[r_val]; together these fix the row multiset.
"""
import numpy as np
import pytest

import brigjoin as bj
import oracle as O

pytestmark = pytest.mark.gpu

MAX = (1 << 31) - 1
SMALL = 64_000_000


@pytest.mark.parametrize("nr,ns", [(MAX, SMALL), (SMALL, MAX)], ids=["r_max", "s_max"])
def test_max_relation_size(gpu_ctx, nr, ns):
    import torch
    m = O.match_count(ns, 0.1)
    print(f"[max_size] |R|={nr} |S|={ns} rows={m}", flush=True)
    rk, rv = (torch.empty(nr, dtype=torch.int32, device="cuda") for _ in range(2))
    sk, sv = (torch.empty(ns, dtype=torch.int32, device="cuda") for _ in range(2))
    gpu_ctx.gen_build(7, nr, rk, rv)
    gpu_ctx.gen_probe(7, 7 ^ 0xDEADBEEF, nr, ns, m, sk, sv)
    torch.cuda.synchronize()
    p = bj.default_params(r_size=nr, s_size=ns, inputs_on_device=1, use_bf=1, bf_bits_per_key=10.0,
                          num_hashes=3, collect_phase_times=0)
    res = gpu_ctx.join(p, rk, rv, sk, sv)
    print(f"[max_size] bits {res.bits1}+{res.bits2} filter stage {res.bf_stage} "
          f"survivors {res.s_survivors} items {res.heavy_items}", flush=True)
    assert res.match_count == m and res.rows_written == m and res.truncated == 0
    assert res.s_survivors >= m
    k, a, b = gpu_ctx.rows_to_host(res)
    rkh = rk.cpu().numpy().view(np.uint32)
    skh = sk.cpu().numpy().view(np.uint32)
    del rk, rv, sk, sv
    torch.cuda.empty_cache()
    assert int(b.max()) < ns and int(a.max()) < nr
    seen = np.zeros(ns, dtype=bool)
    seen[b] = True
    assert int(seen.sum()) == m                                  # no S position twice
    assert np.array_equal(seen, skh < np.uint32(1 << 31))        # exactly the matching positions
    assert np.array_equal(skh[b], k)
    assert np.array_equal(rkh[a], k)

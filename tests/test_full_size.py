"""The 640M x 640M join at its full size (GPU tests).

The CPU oracle cannot join 1.28 billion tuples in seconds, so parity at this
size rests on properties that do not depend on the size.  R's keys are
distinct and S's matching keys are drawn from R, all below 2^31, while S's
other keys are at or above 2^31.  So:
- the count is exactly round(0.1 * |S|);
- the rows' S positions are exactly the positions whose key is below 2^31, each
  once;
- every row satisfies S_key[s_val] == key == R_key[r_val] (values are row
  indices).
Together these determine the row multiset exactly.
"""
import numpy as np
import pytest

import brigjoin as bj
import oracle as O

pytestmark = pytest.mark.gpu

N = 640_000_000


@pytest.fixture(scope="module")
def relations(gpu_ctx):
    import torch
    m = O.match_count(N, 0.1)
    rk, rv, sk, sv = (torch.empty(N, dtype=torch.int32, device="cuda") for _ in range(4))
    gpu_ctx.gen_build(42, N, rk, rv)
    gpu_ctx.gen_probe(42, 42 ^ 0xDEADBEEF, N, N, m, sk, sv)
    torch.cuda.synchronize()
    rkh = rk.cpu().numpy().view(np.uint32)
    skh = sk.cpu().numpy().view(np.uint32)
    yield m, (rk, rv, sk, sv), rkh, skh
    del rk, rv, sk, sv
    torch.cuda.empty_cache()


@pytest.mark.parametrize("kw", [
    dict(partition_size=512, radix_bits=14, use_bf=1, bf_bits_per_key=10.0, num_hashes=3),
    dict(partition_size=512, radix_bits=14, use_bf=1, bf_bits_per_key=10.0, num_hashes=3, s_chunks=4),
    dict(partition_size=512, radix_bits=14, use_bf=1, bf_bits_per_key=10.0, num_hashes=3, s_chunks=4,
         chunk_probe=1),
])
def test_full_size_join(gpu_ctx, relations, kw):
    m, dev, rkh, skh = relations
    p = bj.default_params(r_size=N, s_size=N, inputs_on_device=1, **kw)
    print(f"[full_size] n={N} {kw}", flush=True)
    res = gpu_ctx.join(p, *dev)
    assert res.match_count == m and res.rows_written == m and res.truncated == 0
    assert res.s_survivors >= m
    k, a, b = gpu_ctx.rows_to_host(res)
    seen = np.zeros(N, dtype=bool)
    seen[b] = True
    assert int(seen.sum()) == m                                  # no S position twice
    assert np.array_equal(seen, skh < np.uint32(1 << 31))        # exactly the matching positions
    assert np.array_equal(skh[b], k)
    assert np.array_equal(rkh[a], k)

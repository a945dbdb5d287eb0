"""bj_join_noalloc (GPU tests): the join with no device allocation in the call.

Its output must equal bj_join's on the same inputs, bit for bit (same count,
same sorted row multiset, which is also the oracle's), on two calls in a row,
the second with a smaller input (the scratch array keeps the first call's
contents).  The context it runs on allocates nothing during the calls (device
free memory unchanged), a join whose workspace exceeds the array is refused,
and the count stays exact past the caller's row capacity.
"""
import numpy as np
import pytest

import brigjoin as bj
import oracle as O

pytestmark = pytest.mark.gpu


def inputs(n_r, n_s, ratio, seed):
    import torch
    m = O.match_count(n_s, ratio)
    rk, rv = O.gen_r(seed, n_r)
    sk, sv = O.gen_s(seed, seed + 1, n_r, n_s, m)
    dev = [torch.from_numpy(x.view(np.int32)).cuda() for x in (rk, rv, sk, sv)]
    return (rk, rv, sk, sv), dev, m


def test_noalloc_matches_bj_join_twice():
    import torch
    cases = [(2_000_000, 3_000_000, 0.2, 5, dict(use_bf=1, bf_bits_per_key=10.0, num_hashes=3)),
             (300_000, 450_000, 0.5, 9, dict())]
    data = [inputs(n_r, n_s, ratio, seed) for n_r, n_s, ratio, seed, _ in cases]
    cap = max(n_s for _, n_s, *_ in cases)
    outs = [torch.empty(cap, dtype=torch.int32, device="cuda") for _ in range(3)]
    torch.cuda.synchronize()
    with bj.Context(0) as ref_ctx, bj.Context(0) as ctx:
        expect = []
        for (n_r, n_s, _, _, kw), (host, dev, m) in zip(cases, data):
            p = bj.default_params(r_size=n_r, s_size=n_s, inputs_on_device=1, **kw)
            res = ref_ctx.join(p, *dev)
            expect.append((res.match_count, O.sort_rows(*ref_ctx.rows_to_host(res))))
        torch.cuda.synchronize()
        free0 = torch.cuda.mem_get_info()[0]
        for (n_r, n_s, _, _, kw), (host, dev, m), (n_exp, rows_exp) in zip(cases, data, expect):
            print(f"[noalloc] |R|={n_r} |S|={n_s} {kw}", flush=True)
            p = bj.default_params(r_size=n_r, s_size=n_s, inputs_on_device=1, **kw)
            res = ctx.join_noalloc(p, *dev, *outs, cap)
            assert res.match_count == n_exp == m and res.rows_written == m and res.truncated == 0
            assert res.key == outs[0].data_ptr()
            k, a, b = (t[:m].cpu().numpy().view(np.uint32) for t in outs)
            rows = O.sort_rows(k, a, b)
            assert np.array_equal(rows, rows_exp)
            n, ek, ea, eb = O.cpu_join(*host)
            assert np.array_equal(rows, O.sort_rows(ek, ea, eb))
        torch.cuda.synchronize()
        free1 = torch.cuda.mem_get_info()[0]
        assert free0 - free1 < (16 << 20), (free0, free1)


def test_noalloc_limits():
    import torch
    (host, dev, m) = inputs(100_000, 100_000, 1.0, 3)
    outs = [torch.empty(1000, dtype=torch.int32, device="cuda") for _ in range(3)]
    torch.cuda.synchronize()
    with bj.Context(0) as ctx:
        # a workspace larger than the scratch array is refused before anything runs
        big = bj.default_params(r_size=20_000_000, s_size=20_000_000, inputs_on_device=1)
        with pytest.raises(bj.BjError) as e:
            ctx.join_noalloc(big, *dev, *outs, 1000)
        assert e.value.code == -5
        # host inputs and chunked S are refused
        with pytest.raises(bj.BjError):
            ctx.join_noalloc(bj.default_params(r_size=100_000, s_size=100_000), *dev, *outs, 1000)
        with pytest.raises(bj.BjError):
            ctx.join_noalloc(bj.default_params(r_size=100_000, s_size=100_000, inputs_on_device=1, s_chunks=2),
                             *dev, *outs, 1000)
        # past the caller's capacity: the count stays exact, the rows written are matches
        p = bj.default_params(r_size=100_000, s_size=100_000, inputs_on_device=1)
        res = ctx.join_noalloc(p, *dev, *outs, 1000)
        assert res.match_count == m == 100_000 and res.rows_written == 1000 and res.truncated == 1
        k, a, b = (t[:1000].cpu().numpy().view(np.uint32) for t in outs)
        rk, rv, sk, sv = host
        assert np.all(rk[a] == k) and np.all(sk[b] == k)
        assert len(np.unique(b)) == 1000

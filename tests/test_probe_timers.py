"""Phase timers inside the probe kernel (probe_breakdown), through the C ABI
(GPU tests).

The reference records per-block clock64() stamps in the probe when COLLECT is
set (src/kernel_functions.cu:300-302, 442-451) and turns each phase's share of
the cycles into seconds of the probe kernel (src/join_executor.cu:56-98).  The
timers must not change the join: rows bit-exact against the oracle with and
without them.  The phases split the probe kernel's time exactly (they are
shares of it), the build, count and row phases of a join with matches are
non-zero, and chunks of S add up.
"""
import numpy as np
import pytest

import brigjoin as bj
import oracle as O

pytestmark = pytest.mark.gpu


def to_dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a, np.uint32).view(np.int32)).to("cuda")


def run(ctx, rk, rv, sk, sv, **kw):
    import torch
    p = bj.default_params(r_size=len(rk), s_size=len(sk), inputs_on_device=1, **kw)
    d = [to_dev(x) for x in (rk, rv, sk, sv)]
    torch.cuda.synchronize()
    print(f"[probe_timers] |R|={len(rk)} |S|={len(sk)} {kw}", flush=True)
    res = ctx.join(p, *d)
    k, a, b = ctx.rows_to_host(res)
    return res, O.sort_rows(k, a, b)


@pytest.mark.parametrize("kw", [
    dict(use_bf=1, bf_bits_per_key=10.0, num_hashes=3),
    dict(radix_bits=8, tile=4096),
    dict(radix_bits=2),                      # vlsplit items (chunks x slices)
    dict(use_bf=1, radix_bits=12, s_chunks=3),
])
def test_breakdown_splits_probe_time(gpu_ctx, kw):
    n_r, n_s = 1 << 20, 1 << 21
    m = O.match_count(n_s, 0.2)
    rk, rv = O.gen_r(42, n_r)
    sk, sv = O.gen_s(42, 42 ^ 0xDEADBEEF, n_r, n_s, m)
    n, k, a, b = O.cpu_join(rk, rv, sk, sv)
    expect = O.sort_rows(k, a, b)
    # collect_phase_times = 0: the breakdown still times the probe kernel
    res, rows = run(gpu_ctx, rk, rv, sk, sv, probe_breakdown=1, collect_phase_times=0, **kw)
    assert res.match_count == n == m and np.array_equal(rows, expect)
    pp = res.probe_phases()
    probe_ms = res.kernels()["probe"]
    assert probe_ms > 0 and all(v >= 0 for v in pp.values())
    assert abs(sum(pp.values()) - probe_ms) <= 1e-6 * probe_ms + 1e-9
    assert pp["build"] > 0 and pp["count"] > 0 and pp["write"] > 0
    # without the breakdown: the same rows, no probe phases
    res0, rows0 = run(gpu_ctx, rk, rv, sk, sv, **kw)
    assert np.array_equal(rows0, expect)
    assert all(v == 0 for v in res0.probe_phases().values())


def test_breakdown_without_matches(gpu_ctx):
    """No matches: every round stops after its count; no row phase."""
    rk, rv = O.gen_r(7, 300_000)
    sk, sv = O.gen_s(7, 8, 300_000, 500_000, 0)
    res, rows = run(gpu_ctx, rk, rv, sk, sv, probe_breakdown=1)
    pp = res.probe_phases()
    assert res.match_count == 0 and len(rows) == 0
    assert pp["count"] > 0 and pp["write"] == 0

"""The serving form: resident relations, admission by footprint, continuous batching.

The bar is the single join's: each request's count and sorted row multiset
equal the oracle's (and bj_join's) on the same inputs, whether it ran alone or
overlapped with other requests on the device.
"""
import numpy as np
import pytest

import brigjoin as bj
import oracle as O


def test_server_opts_default_and_nodevice():
    o = bj.server_opts()
    assert o.slots == 3 and o.use_bf == 1 and o.bf_bits_per_key == 10.0 and o.num_hashes == 3
    assert o.radix_bits == -1 and o.answer_scale == 1.0 and o.arena_bytes == 0
    try:
        import torch
        has_gpu = torch.cuda.is_available()
    except Exception:
        has_gpu = False
    if not has_gpu:
        with pytest.raises(bj.BjError) as e:
            bj.Server(0)
        assert e.value.code == -6
    # options are checked before the device is opened
    for bad in (dict(answer_scale=float("nan")), dict(bf_bits_per_key=float("inf")), dict(slots=0)):
        with pytest.raises(bj.BjError) as e:
            bj.Server(0, **bad)
        assert e.value.code == -1, bad


def _gen(ctx, seed, nr, ns, ratio):
    import torch
    rk, rv = (torch.empty(max(nr, 1), dtype=torch.int32, device="cuda") for _ in range(2))
    ctx.gen_build(seed, nr, rk, rv)
    return rk, rv


def _gen_s(ctx, seed_r, seed_s, nr, ns, ratio):
    import torch
    sk, sv = (torch.empty(max(ns, 1), dtype=torch.int32, device="cuda") for _ in range(2))
    m = O.match_count(ns, ratio)
    ctx.gen_probe(seed_r, seed_s, nr, ns, m, sk, sv)
    return sk, sv, m


def _host(t, n):
    return t[:n].cpu().numpy().view(np.uint32)


@pytest.mark.gpu
def test_serve_alone_and_in_mix(gpu_ctx):
    rels = [(11, 300_000), (12, 1_000_000), (13, 2_000_000)]
    Rdev = [_gen(gpu_ctx, seed, n, 0, 0) for seed, n in rels]
    rng = np.random.default_rng(7)
    reqs = []
    for i in range(20):
        ri = int(rng.integers(0, len(rels)))
        ns = int(np.exp(rng.uniform(np.log(50_000), np.log(3_000_000))))
        sk, sv, m = _gen_s(gpu_ctx, rels[ri][0], 500 + i, rels[ri][1], ns, 0.1)
        reqs.append((ri, sk, sv, ns, m))
    big = max(r[3] for r in reqs)
    with bj.Server(0, slots=3) as srv:
        ids = [srv.add_relation(rk, rv, n) for (rk, rv), (_, n) in zip(Rdev, rels)]
        assert ids == [0, 1, 2]
        # alone: one request on the device at a time
        alone = []
        for ri, sk, sv, ns, m in reqs:
            rid = srv.submit(ri, sk, sv, ns)
            r = srv.wait(rid)
            assert r.status == 0 and r.match_count == m
            alone.append(O.sort_rows(*srv.rows_to_host(r)))
            srv.release(rid)
        assert srv.stats().max_in_flight == 1
    # in the mix: all submitted at once into an arena that holds about three of
    # the largest requests, so requests queue for slots and memory
    with bj.Server(0, slots=3) as srv:
        for (rk, rv), (_, n) in zip(Rdev, rels):
            srv.add_relation(rk, rv, n)
        foot = max(srv.request_bytes(r[0], r[3]) for r in reqs)
    with bj.Server(0, slots=3, arena_bytes=int(3.2 * foot)) as srv:
        for (rk, rv), (_, n) in zip(Rdev, rels):
            srv.add_relation(rk, rv, n)
        print(f"[serve mix] requests={len(reqs)} largest |S|={big} footprint={foot}", flush=True)
        rids = [srv.submit(ri, sk, sv, ns) for ri, sk, sv, ns, m in reqs]
        for k, rid in enumerate(rids):
            r = srv.wait(rid)
            assert r.status == 0, k
            assert r.match_count == reqs[k][4], k
            assert r.t_submit <= r.t_start <= r.t_end
            assert np.array_equal(O.sort_rows(*srv.rows_to_host(r)), alone[k]), k
            srv.release(rid)
        st = srv.stats()
        assert st.submitted == st.completed == st.released == len(reqs)
        assert st.max_in_flight >= 2              # requests overlapped on the device
        assert st.arena_peak_bytes <= st.arena_bytes
    # and a few against the oracle
    for k in range(3):
        ri, sk, sv, ns, m = reqs[k]
        rk, rv = Rdev[ri]
        n, a, b, c = O.cpu_join(_host(rk, rels[ri][1]), _host(rv, rels[ri][1]), _host(sk, ns), _host(sv, ns))
        assert n == m and np.array_equal(O.sort_rows(a, b, c), alone[k])


@pytest.mark.gpu
def test_serve_matches_single_join_and_limits(gpu_ctx):
    rk, rv = _gen(gpu_ctx, 21, 500_000, 0, 0)
    sk, sv, m = _gen_s(gpu_ctx, 21, 77, 500_000, 800_000, 0.3)
    with bj.Server(0, slots=2, arena_bytes=1 << 30, use_bf=0) as srv:
        foot = srv.request_bytes(srv.add_relation(rk, rv, 500_000), 800_000)
    # room for the rows of about five requests besides one running
    with bj.Server(0, slots=2, arena_bytes=int(4.5 * foot), use_bf=0) as srv:
        rel = srv.add_relation(rk, rv, 500_000)
        rid = srv.submit(rel, sk, sv, 800_000)
        r = srv.wait(rid)
        assert r.status == 0 and r.match_count == m and r.s_survivors == 800_000
        rows = O.sort_rows(*srv.rows_to_host(r))
        res = gpu_ctx.join(bj.default_params(r_size=500_000, s_size=800_000, inputs_on_device=1),
                           rk, rv, sk, sv)
        assert res.match_count == m
        assert np.array_equal(O.sort_rows(*gpu_ctx.rows_to_host(res)), rows)
        # a waited-for result is not reported again by poll
        assert srv.poll() == []
        srv.release(rid)
        with pytest.raises(bj.BjError):
            srv.release(rid)
        # a request that can never fit the arena; a relation that does not exist
        with pytest.raises(bj.BjError) as e:
            srv.submit(rel, sk, sv, 100_000_000)
        assert e.value.code == -5
        with pytest.raises(bj.BjError) as e:
            srv.submit(5, sk, sv, 10)
        assert e.value.code == -1
        # rows that are never released block the queue: drain reports it instead of waiting
        held = [srv.submit(rel, sk, sv, 800_000) for _ in range(12)]
        with pytest.raises(bj.BjError) as e:
            srv.drain()
        assert e.value.code == -3
        done = srv.poll()
        assert 0 < len(done) < len(held)
        for x in done:
            assert x.status == 0 and x.match_count == m
            srv.release(x.id)
        finished = {x.id for x in done}
        for i in held:
            if i not in finished:
                x = srv.wait(i)
                assert x.status == 0 and x.match_count == m
                srv.release(i)
        srv.drain()
        st = srv.stats()
        assert st.completed == st.submitted == st.released == 13


@pytest.mark.gpu
def test_serve_degenerate_requests(gpu_ctx):
    """Empty S, a one-tuple S, an empty resident relation: each result equals the oracle's."""
    rk, rv = _gen(gpu_ctx, 31, 1000, 0, 0)
    sk, sv, m = _gen_s(gpu_ctx, 31, 32, 1000, 5000, 0.5)
    hr = [_host(rk, 1000), _host(rv, 1000)]
    hs = [_host(sk, 5000), _host(sv, 5000)]
    with bj.Server(0, slots=2) as srv:
        rel = srv.add_relation(rk, rv, 1000)
        empty_rel = srv.add_relation(rk, rv, 0)
        cases = [(rel, 0), (rel, 1), (rel, 5000), (empty_rel, 5000), (empty_rel, 0)]
        ids = [srv.submit(r, sk, sv, n) for r, n in cases]
        for (r, n), rid in zip(cases, ids):
            res = srv.wait(rid)
            nr = 1000 if r == rel else 0
            cnt, k, a, b = O.cpu_join(hr[0][:nr], hr[1][:nr], hs[0][:n], hs[1][:n])
            assert res.status == 0 and res.match_count == cnt and not res.truncated, (r, n)
            assert np.array_equal(O.sort_rows(*srv.rows_to_host(res)), O.sort_rows(k, a, b)), (r, n)
            srv.release(rid)
    assert m == O.match_count(5000, 0.5)

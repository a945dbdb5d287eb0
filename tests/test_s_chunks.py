"""S joined in chunks (s_chunks), through the C ABI (GPU tests).

The reference's s_chunks loop (src/hash_kernel.cu:646-650, 910-931) joins S
chunk by chunk against the partitioned R, appending rows.  Bar: bit-exact.  The
same match count and the same sorted multiset of (key, r_val, s_val) rows as
the oracle's CPU join, for device and host inputs (host chunks are copied on a
second stream while the previous chunk is joined), chunk counts that do and do
not divide |S|, more chunks than S tuples, the filter on (both stages) and
off, duplicate-heavy keys, and a row capacity exceeded across chunks.  Both
probe modes: every chunk's S' pooled and probed once after the last chunk
(chunk_probe 0, the default where the pool fits) and one probe per chunk
(chunk_probe 1).
"""
import numpy as np
import pytest

import brigjoin as bj
import oracle as O

pytestmark = pytest.mark.gpu


def to_dev(a):
    import torch
    a = np.ascontiguousarray(a, np.uint32)
    if len(a) == 0:
        return torch.zeros(1, dtype=torch.int32, device="cuda")
    return torch.from_numpy(a.view(np.int32)).to("cuda")


def run(ctx, rk, rv, sk, sv, host, **kw):
    import torch
    p = bj.default_params(r_size=len(rk), s_size=len(sk), inputs_on_device=0 if host else 1, **kw)
    ins = [np.ascontiguousarray(x, np.uint32) for x in (rk, rv, sk, sv)]
    if not host:
        ins = [to_dev(x) for x in ins]
        torch.cuda.synchronize()
    print(f"[s_chunks] |R|={len(rk)} |S|={len(sk)} host={host} {kw}", flush=True)
    res = ctx.join(p, *ins)
    k, a, b = ctx.rows_to_host(res)
    return res, O.sort_rows(k, a, b)


CASES = [
    # n_r, n_s, ratio, params
    (1 << 20, 1 << 21, 0.1, dict(s_chunks=4)),
    (1 << 20, 3_000_001, 0.2, dict(s_chunks=7, use_bf=1, bf_bits_per_key=10.0, num_hashes=3)),
    (1 << 21, 1 << 21, 0.1, dict(s_chunks=3, use_bf=1, bf_stage=2, partition_size=512, radix_bits=14)),
    (1 << 20, 1_500_000, 0.3, dict(s_chunks=5, use_bf=1, bf_stage=1, radix_bits=12)),
    (500_000, 2_000_000, 1.0, dict(s_chunks=2, tile=4096)),
    (2_000_000, 300_000, 0.5, dict(s_chunks=6, radix_bits=4)),  # vlsplit items in every chunk
]


MODES = pytest.mark.parametrize("cprobe", [0, 1], ids=["pooled", "per_chunk"])


@MODES
@pytest.mark.parametrize("host", [False, True], ids=["device", "host"])
@pytest.mark.parametrize("ci", range(len(CASES)))
def test_chunks_match_oracle(gpu_ctx, ci, host, cprobe):
    n_r, n_s, ratio, kw = CASES[ci]
    kw = dict(kw, chunk_probe=cprobe)
    m = O.match_count(n_s, ratio)
    rk, rv = O.gen_r(42, n_r)
    sk, sv = O.gen_s(42, 42 ^ 0xDEADBEEF, n_r, n_s, m)
    res, rows = run(gpu_ctx, rk, rv, sk, sv, host, collect_phase_times=1, **kw)
    n, k, a, b = O.cpu_join(rk, rv, sk, sv)
    assert n == m == res.match_count and res.truncated == 0
    assert res.s_chunks == min(kw["s_chunks"], n_s)
    assert np.array_equal(rows, O.sort_rows(k, a, b))
    # the S-side times are the sums over the chunks; the whole span covers them
    ks = res.kernels()
    assert ks["probe"] > 0 and ks["hist_r"] > 0
    assert res.phases()["total"] >= ks["probe"]


@MODES
def test_chunks_same_survivors_as_one_pass(gpu_ctx, cprobe):
    """|S'| is a per-tuple property of the filter: the chunk sums equal one pass."""
    n_r, n_s = 1 << 20, 1 << 21
    m = O.match_count(n_s, 0.1)
    rk, rv = O.gen_r(42, n_r)
    sk, sv = O.gen_s(42, 42 ^ 0xDEADBEEF, n_r, n_s, m)
    kw = dict(use_bf=1, bf_stage=2, radix_bits=12, bf_bits_per_key=8.0, num_hashes=3)
    one, rows1 = run(gpu_ctx, rk, rv, sk, sv, False, s_chunks=1, **kw)
    many, rows5 = run(gpu_ctx, rk, rv, sk, sv, True, s_chunks=5, chunk_probe=cprobe, **kw)
    assert one.s_chunks == 1 and many.s_chunks == 5
    assert one.s_survivors == many.s_survivors and one.s_survivors >= m
    assert np.array_equal(rows1, rows5)


def test_more_chunks_than_tuples(gpu_ctx):
    rk = np.array([5, 7, 9, 0x80000001], np.uint32)
    rv = np.arange(4, dtype=np.uint32)
    sk = np.array([9, 1, 0x80000001], np.uint32)
    sv = np.array([10, 11, 12], np.uint32)
    for host in (False, True):
        res, rows = run(gpu_ctx, rk, rv, sk, sv, host, s_chunks=8, use_bf=1)
        assert res.s_chunks == 3 and res.match_count == 2
        assert rows.tolist() == [[9, 2, 10], [0x80000001, 3, 12]]


def test_context_reuse_after_rejected_call(gpu_ctx):
    """A call rejected by argument validation (s_chunks < 0, refused before any
    copy or kernel is queued) leaves the context usable: the next chunked join on
    it, from host buffers, is exact."""
    rk = np.array([5, 7, 9, 0x80000001], np.uint32)
    rv = np.arange(4, dtype=np.uint32)
    sk = np.array([9, 1, 0x80000001, 7], np.uint32)
    sv = np.array([10, 11, 12, 13], np.uint32)
    with pytest.raises(bj.BjError):
        run(gpu_ctx, rk, rv, sk, sv, True, s_chunks=-1)
    res, rows = run(gpu_ctx, rk, rv, sk, sv, True, s_chunks=2, use_bf=1)
    assert res.s_chunks == 2 and res.match_count == 3
    assert rows.tolist() == [[7, 1, 13], [9, 2, 10], [0x80000001, 3, 12]]


@MODES
def test_duplicates_across_chunks(gpu_ctx, cprobe):
    """Skewed S (every key of a small set, many times) with duplicate R keys."""
    rng = np.random.default_rng(7)
    rk = rng.integers(1, 300, 20_000, dtype=np.uint32)
    rv = np.arange(len(rk), dtype=np.uint32)
    sk = rng.integers(1, 400, 150_000, dtype=np.uint32)
    sv = np.arange(len(sk), dtype=np.uint32)
    n, k, a, b = O.cpu_join(rk, rv, sk, sv)
    res, rows = run(gpu_ctx, rk, rv, sk, sv, True, s_chunks=3, answer_scale=float(n) / len(sk) + 1.0,
                    chunk_probe=cprobe)
    assert res.match_count == n and res.truncated == 0
    assert np.array_equal(rows, O.sort_rows(k, a, b))


@MODES
def test_capacity_exceeded_across_chunks(gpu_ctx, cprobe):
    """The count stays exact and the written rows are valid matches."""
    rk = np.repeat(np.arange(1, 11, dtype=np.uint32), 50)   # 10 keys x 50 rows
    rv = np.arange(len(rk), dtype=np.uint32)
    sk = np.tile(np.arange(1, 21, dtype=np.uint32), 40)     # 800 S tuples, half match 50 each
    sv = np.arange(len(sk), dtype=np.uint32)
    n, _, _, _ = O.cpu_join(rk, rv, sk, sv)
    res, rows = run(gpu_ctx, rk, rv, sk, sv, True, s_chunks=4, chunk_probe=cprobe)
    assert res.match_count == n == 400 * 50 and res.truncated == 1
    assert res.rows_written == res.capacity == 800
    assert np.all(rk[rows[:, 1]] == rows[:, 0]) and np.all(sk[rows[:, 2]] == rows[:, 0])
    assert len(np.unique(rows, axis=0)) == len(rows)


def test_auto_chunks_fit_in_one_pass(gpu_ctx):
    n_r, n_s = 1 << 20, 1 << 20
    m = O.match_count(n_s, 0.1)
    rk, rv = O.gen_r(42, n_r)
    sk, sv = O.gen_s(42, 42 ^ 0xDEADBEEF, n_r, n_s, m)
    res, rows = run(gpu_ctx, rk, rv, sk, sv, True, s_chunks=0, use_bf=1)
    assert res.s_chunks == 1 and res.match_count == m


def test_chunk_probe_invalid(gpu_ctx):
    rk = np.array([5, 7], np.uint32)
    with pytest.raises(bj.BjError):
        run(gpu_ctx, rk, rk, rk, rk, False, s_chunks=2, chunk_probe=2)

"""Parity of the HIP join against the oracle, through the C ABI (GPU tests).

Bar: bit-exact.  The same match count and the same sorted multiset of
(key, r_val, s_val) rows as the oracle's CPU join (restated from the reference's
sanity_check and extended to rows) on the same inputs, and as the golden
fixtures (brute-force rows, reference counts).
"""
import glob
import os

import numpy as np
import pytest

import brigjoin as bj
import oracle as O

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = sorted(glob.glob(os.path.join(HERE, "golden", "*.npz")))


def to_dev(a):
    import torch
    a = np.ascontiguousarray(a, np.uint32)
    if len(a) == 0:
        return torch.zeros(1, dtype=torch.int32, device="cuda")
    return torch.from_numpy(a.view(np.int32)).to("cuda")


def dev_empty(n):
    import torch
    return torch.empty(max(n, 1), dtype=torch.int32, device="cuda")


def to_host(t, n):
    return t[:n].cpu().numpy().view(np.uint32)


def gpu_join(ctx, rk, rv, sk, sv, **kw):
    import torch
    p = bj.default_params(r_size=len(rk), s_size=len(sk), inputs_on_device=1, **kw)
    d = [to_dev(x) for x in (rk, rv, sk, sv)]
    torch.cuda.synchronize()
    print(f"[gpu_join] |R|={len(rk)} |S|={len(sk)} {kw}", flush=True)
    res = ctx.join(p, *d)
    k, a, b = ctx.rows_to_host(res)
    return res, O.sort_rows(k, a, b)


def radix_hash_np(x):
    """Vectorised partition hash (checked against the oracle's scalar version below)."""
    return np.asarray(x, np.uint32) * np.uint32(0x9E3779B1)


def oracle_rows(rk, rv, sk, sv):
    n, k, a, b = O.cpu_join(rk, rv, sk, sv)
    return n, O.sort_rows(k, a, b)


VARIANTS = [
    dict(),
    dict(use_bf=1, bf_bits_per_key=10.0, num_hashes=3),
    dict(use_bf=1, bf_stage=1, radix_bits=12),
    dict(use_bf=1, bf_stage=2, radix_bits=12),
    dict(use_bf=1, bf_stage=2, radix_bits=6),
    dict(radix_bits=1),
    dict(radix_bits=4, tile=4096),
    dict(partition_size=64, second_partition_size=16, use_bf=10),
    dict(radix_bits=13, use_bf=1, bf_bits_per_key=4.0, num_hashes=2),
    dict(use_bf=1, bf_stage=2, radix_bits=12, s_pass1_bits=8, scatter_wg=512),
    dict(radix_bits=10, pass1_bits=3, s_pass1_bits=8, scatter_wg=512),
]


@pytest.mark.parametrize("path", GOLDEN, ids=[os.path.basename(p)[:-4] for p in GOLDEN])
@pytest.mark.parametrize("vi", range(len(VARIANTS)))
def test_golden_fixtures(gpu_ctx, path, vi):
    z = np.load(path, allow_pickle=False)
    res, rows = gpu_join(gpu_ctx, z["r_key"], z["r_val"], z["s_key"], z["s_val"], **VARIANTS[vi])
    assert res.match_count == int(z["count"])
    assert res.truncated == 0 or int(z["count"]) > res.capacity
    expect = z["rows"].reshape(-1, 3)
    if not res.truncated:
        assert np.array_equal(rows, expect)


def test_generator_matches_oracle(gpu_ctx):
    n_r, n_s = 1 << 20, 3 << 19
    m = O.match_count(n_s, 0.1)
    dk, dv = dev_empty(n_r), dev_empty(n_r)
    gpu_ctx.gen_build(42, n_r, dk, dv)
    rk, rv = O.gen_r(42, n_r)
    assert np.array_equal(to_host(dk, n_r), rk) and np.array_equal(to_host(dv, n_r), rv)
    sk_d, sv_d = dev_empty(n_s), dev_empty(n_s)
    gpu_ctx.gen_probe(42, 42 ^ 0xDEADBEEF, n_r, n_s, m, sk_d, sv_d)
    sk, sv = O.gen_s(42, 42 ^ 0xDEADBEEF, n_r, n_s, m)
    assert np.array_equal(to_host(sk_d, n_s), sk) and np.array_equal(to_host(sv_d, n_s), sv)


@pytest.mark.parametrize("n_r,n_s,ratio,kw", [
    (1 << 20, 1 << 20, 0.1, dict()),
    (1 << 20, 1 << 20, 0.1, dict(use_bf=1, bf_bits_per_key=10.0, num_hashes=3)),
    (1 << 21, 1 << 21, 0.1, dict(use_bf=1, bf_stage=2, partition_size=512, radix_bits=14,
                                 bf_bits_per_key=10.0, num_hashes=3)),
    (1 << 21, 3 << 20, 0.2, dict(use_bf=1, bf_stage=1, partition_size=512, radix_bits=14)),
    (300_000, 2_000_000, 0.5, dict(use_bf=1, bf_bits_per_key=12.0, num_hashes=4)),
    (2_000_000, 300_000, 1.0, dict(tile=4096)),
    (1_000_000, 1_000_000, 0.1, dict(radix_bits=2)),           # vlsplit: chunks x slices
    (77_777, 1_234_567, 0.3, dict(partition_size=32, radix_bits=9)),
    (1 << 21, 3 << 20, 0.1, dict(use_bf=1, radix_bits=14, bf_stage=2, s_pass1_bits=9, scatter_wg=512)),
    (1 << 20, 1 << 21, 0.1, dict(use_bf=1, radix_bits=13, bf_stage=2, s_pass1_bits=4)),
])
def test_generated_parity(gpu_ctx, n_r, n_s, ratio, kw):
    m = O.match_count(n_s, ratio)
    rk, rv = O.gen_r(42, n_r)
    sk, sv = O.gen_s(42, 42 ^ 0xDEADBEEF, n_r, n_s, m)
    res, rows = gpu_join(gpu_ctx, rk, rv, sk, sv, **kw)
    n, expect = oracle_rows(rk, rv, sk, sv)
    assert n == m == res.match_count
    assert np.array_equal(rows, expect)
    if kw.get("use_bf"):
        assert m <= res.s_survivors < n_s


def test_config1_full_size(gpu_ctx):
    """R=S=16M, 10% match, Bloom off, 10 radix bits: bit-exact against the oracle."""
    import torch
    n = 16_000_000
    m = O.match_count(n, 0.1)
    dk, dv, sk_d, sv_d = (dev_empty(n) for _ in range(4))
    gpu_ctx.gen_build(42, n, dk, dv)
    gpu_ctx.gen_probe(42, 42 ^ 0xDEADBEEF, n, n, m, sk_d, sv_d)
    torch.cuda.synchronize()
    p = bj.default_params(r_size=n, s_size=n, partition_size=1024, inputs_on_device=1)
    res = gpu_ctx.join(p, dk, dv, sk_d, sv_d)
    assert res.match_count == m == 1_600_000 and res.bits1 == 10
    k, a, b = gpu_ctx.rows_to_host(res)
    rk, rv = to_host(dk, n), to_host(dv, n)
    sk, sv = to_host(sk_d, n), to_host(sv_d, n)
    n_o, expect = oracle_rows(rk, rv, sk, sv)
    assert n_o == m
    assert np.array_equal(O.sort_rows(k, a, b), expect)


def test_skewed_zipf_parity(gpu_ctx):
    """Zipf(1.0) matching keys over R: heavy sub-buckets take the vlsplit path."""
    import torch
    n_r, n_s = 1 << 20, 12_000_000
    m = O.match_count(n_s, 0.05)
    dk, dv, sk_d, sv_d = dev_empty(n_r), dev_empty(n_r), dev_empty(n_s), dev_empty(n_s)
    gpu_ctx.gen_build(7, n_r, dk, dv)
    gpu_ctx.gen_probe(7, 9, n_r, n_s, m, sk_d, sv_d, zipf_s=1.0)
    torch.cuda.synchronize()
    sk = to_host(sk_d, n_s)
    top = np.bincount(np.unique(sk[sk < 2**31], return_inverse=True)[1]).max()
    assert top > 20_000                     # a genuinely heavy key
    for kw in (dict(use_bf=1, bf_bits_per_key=12.0, num_hashes=4), dict()):
        p = bj.default_params(r_size=n_r, s_size=n_s, inputs_on_device=1, **kw)
        res = gpu_ctx.join(p, dk, dv, sk_d, sv_d)
        k, a, b = gpu_ctx.rows_to_host(res)
        rk, rv, sv = to_host(dk, n_r), to_host(dv, n_r), to_host(sv_d, n_s)
        n_o, expect = oracle_rows(rk, rv, sk, sv)
        assert res.match_count == n_o == m
        assert np.array_equal(O.sort_rows(k, a, b), expect)


def test_truncation_exact_count(gpu_ctx):
    """Output beyond capacity is dropped, the count stays exact (many-to-many)."""
    rk = np.full(3000, 7, np.uint32)
    sk = np.concatenate([np.full(4000, 7, np.uint32), np.arange(100, 200, dtype=np.uint32)])
    rv = np.arange(len(rk), dtype=np.uint32)
    sv = np.arange(len(sk), dtype=np.uint32)
    res, rows = gpu_join(gpu_ctx, rk, rv, sk, sv)
    assert res.match_count == 3000 * 4000
    assert res.truncated == 1 and res.rows_written == res.capacity == len(sk)
    assert np.all(rows[:, 0] == 7)
    assert np.all(rk[rows[:, 1]] == 7) and np.all(sk[rows[:, 2]] == 7)
    assert len(np.unique(rows, axis=0)) == len(rows)
    res2, rows2 = gpu_join(gpu_ctx, rk, rv, sk, sv, answer_scale=4000.0)
    assert res2.truncated == 0 and res2.match_count == 12_000_000
    assert len(np.unique(rows2[:, 1].astype(np.uint64) << np.uint64(32) | rows2[:, 2])) == 12_000_000


def test_host_inputs_path(gpu_ctx):
    n_r, n_s = 500_000, 700_000
    m = O.match_count(n_s, 0.2)
    rk, rv = O.gen_r(3, n_r)
    sk, sv = O.gen_s(3, 4, n_r, n_s, m)
    p = bj.default_params(r_size=n_r, s_size=n_s, inputs_on_device=0, use_bf=1)
    res = gpu_ctx.join(p, rk, rv, sk, sv)
    k, a, b = gpu_ctx.rows_to_host(res)
    n, expect = oracle_rows(rk, rv, sk, sv)
    assert res.match_count == n == m
    assert np.array_equal(O.sort_rows(k, a, b), expect)


def test_repeated_joins_reuse_context(gpu_ctx):
    """The context keeps its workspace; a larger then a smaller join stay exact."""
    for n, ratio in ((800_000, 0.3), (120_000, 0.9), (800_000, 0.01)):
        m = O.match_count(n, ratio)
        rk, rv = O.gen_r(n, n)
        sk, sv = O.gen_s(n, n + 1, n, n, m)
        res, rows = gpu_join(gpu_ctx, rk, rv, sk, sv, use_bf=1)
        _, expect = oracle_rows(rk, rv, sk, sv)
        assert res.match_count == m and np.array_equal(rows, expect)


def test_gpu_main_reference_files(tmp_path):
    """gpu_main's file surface: raw int32 key files (+ value block), output.bin."""
    z = np.load(os.path.join(HERE, "golden", "refgen_2000.npz"), allow_pickle=False)
    rpath, spath, opath = tmp_path / "R.bin", tmp_path / "S.bin", tmp_path / "output.bin"
    z["r_key"].tofile(rpath)
    z["s_key"].tofile(spath)
    # without --read_val the reference synthesizes R's values as i + 1 and S's as
    # i + 100000001 (src/hash_kernel.cu:330-336); the fixture's rows carry row indices
    args = bj.main_args(r_path=str(rpath), s_path=str(spath), out_path=str(opath),
                        r_size=len(z["r_key"]), s_size=len(z["s_key"]), sanity_check=1,
                        use_bf=14, verbose=0)
    res = bj.gpu_main(args)
    assert res.match_count == int(z["count"]) == int(z["ref_count"])
    k, a, b = bj.read_output(str(opath))
    expect = z["rows"] + np.array([0, 1, 100000001], np.uint32)
    assert np.array_equal(np.stack([k, a, b], 1), expect)
    # --val_mode 0: the bare row index
    args = bj.main_args(r_path=str(rpath), s_path=str(spath), out_path=str(opath),
                        r_size=len(z["r_key"]), s_size=len(z["s_key"]), use_bf=14, val_mode=0, verbose=0)
    bj.gpu_main(args)
    k, a, b = bj.read_output(str(opath))
    assert np.array_equal(np.stack([k, a, b], 1), z["rows"])
    # --read_val: the value block follows the key block
    rv = (np.arange(len(z["r_key"])) * 3 + 11).astype(np.uint32)
    sv = (np.arange(len(z["s_key"])) * 5 + 7).astype(np.uint32)
    np.concatenate([z["r_key"], rv]).tofile(rpath)
    np.concatenate([z["s_key"], sv]).tofile(spath)
    args = bj.main_args(r_path=str(rpath), s_path=str(spath), out_path=str(opath),
                        r_size=len(rv), s_size=len(sv), read_val=1, verbose=0)
    bj.gpu_main(args)
    k, a, b = bj.read_output(str(opath))
    n, expect = oracle_rows(z["r_key"], rv, z["s_key"], sv)
    assert np.array_equal(np.stack([k, a, b], 1), expect)


@pytest.mark.parametrize("n,kw", [
    (40, dict(radix_bits=12)),                          # most buckets empty
    (100_000, dict(radix_bits=11)),                     # single level, 2048-way
    (1_000_000, dict(partition_size=512, radix_bits=14)),
    (777_777, dict(partition_size=1024, radix_bits=15)),
    (3_000_000, dict(partition_size=64, second_partition_size=64)),
    (0, dict(radix_bits=8)),
])
def test_partition_kernels(gpu_ctx, n, kw):
    """Partition parity: a permutation of the input, every tuple in its bucket
    (top bits of key * 0x9E3779B1), bucket sizes equal to a CPU histogram."""
    import torch
    rng = np.random.default_rng(n + 1)
    key = rng.integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32)
    if n > 10:
        key[: n // 10] = key[0]                           # a heavy key
    val = np.arange(n, dtype=np.uint32)
    p = bj.default_params(**kw)
    info = bj.plan(bj.default_params(r_size=n, s_size=n, **kw))
    bits = info.bits1 + info.bits2
    dk, dv = to_dev(key), to_dev(val)
    ok, ov = dev_empty(n), dev_empty(n)
    st = torch.zeros((1 << bits) + 1, dtype=torch.int64, device="cuda")
    torch.cuda.synchronize()
    gpu_ctx.partition(p, dk, dv, n, ok, ov, st)
    start = st.cpu().numpy().astype(np.uint64)
    hk, hv = to_host(ok, n), to_host(ov, n)
    assert np.array_equal(np.sort(hv), val)                  # a permutation (values are row ids)
    assert np.array_equal(hk, key[hv])                       # keys travel with their values
    hk32 = radix_hash_np(key)
    assert all(int(hk32[i]) == O.radix_hash(int(key[i])) for i in range(min(n, 200)))
    h = hk32.astype(np.uint64) >> np.uint64(32 - bits)
    hist = np.bincount(h.astype(np.int64), minlength=1 << bits)
    assert np.array_equal(np.diff(start), hist)              # bucket sizes
    assert start[0] == 0 and start[-1] == n
    hb = radix_hash_np(hk).astype(np.uint64) >> np.uint64(32 - bits)
    pos = np.arange(n, dtype=np.uint64)
    assert np.all((start[hb.astype(np.int64)] <= pos) & (pos < start[hb.astype(np.int64) + 1]))


@pytest.mark.parametrize("n,kw", [
    (40_000_000, dict(partition_size=512, radix_bits=14)),   # 128 + 128 buckets
    (36_000_000, dict(partition_size=256, radix_bits=15)),   # 256 + 128 buckets
])
def test_partition_large_relations(gpu_ctx, n, kw):
    """Relations of 32M tuples or more are scattered by one workgroup per CU with
    8192-tuple tiles and whole-line carries: the same partition properties as
    above, checked on the device."""
    import torch
    g = torch.Generator(device="cuda").manual_seed(n)
    key = torch.randint(-2**31, 2**31 - 1, (n,), dtype=torch.int32, device="cuda", generator=g)
    key[: n // 10] = key[0]                                   # a heavy key
    val = torch.arange(n, dtype=torch.int32, device="cuda")
    p = bj.default_params(**kw)
    info = bj.plan(bj.default_params(r_size=n, s_size=n, **kw))
    bits = info.bits1 + info.bits2
    ok, ov = dev_empty(n), dev_empty(n)
    st = torch.zeros((1 << bits) + 1, dtype=torch.int64, device="cuda")
    torch.cuda.synchronize()
    print(f"[partition_large] n={n} {kw}", flush=True)
    gpu_ctx.partition(p, key, val, n, ok, ov, st)
    torch.cuda.synchronize()
    ov64 = ov[:n].to(torch.int64)
    assert torch.equal(torch.sort(ov64).values, torch.arange(n, device="cuda"))   # a permutation
    assert torch.equal(ok[:n], key[ov64])                                         # keys travel with values
    mul = torch.tensor(0x9E3779B1, dtype=torch.int64, device="cuda")
    h_in = ((key.to(torch.int64) & 0xFFFFFFFF) * mul & 0xFFFFFFFF) >> (32 - bits)
    hist = torch.bincount(h_in, minlength=1 << bits)
    assert torch.equal(st[1:] - st[:-1], hist)                                    # bucket sizes
    assert int(st[0]) == 0 and int(st[-1]) == n
    h_out = ((ok[:n].to(torch.int64) & 0xFFFFFFFF) * mul & 0xFFFFFFFF) >> (32 - bits)
    pos = torch.arange(n, device="cuda")
    assert bool(((st[h_out] <= pos) & (pos < st[h_out + 1])).all())             # each in its bucket


@pytest.mark.parametrize("kw", [
    dict(partition_size=512, radix_bits=14, use_bf=1, bf_bits_per_key=10.0, num_hashes=3),
    dict(partition_size=512, radix_bits=14),
])
def test_large_join_rows(gpu_ctx, kw):
    """R = S = 48M (the one-workgroup-per-CU scatter path), 10% match.  R's keys
    are distinct, so the rows are exactly determined by the matching S
    positions: the s_val set must be {j : S_key[j] < 2^31} (the generator's
    matches), and every row must satisfy S_key[s_val] == key == R_key[r_val]."""
    import torch
    n = 48_000_000
    m = O.match_count(n, 0.1)
    dk, dv, sk, sv = (dev_empty(n) for _ in range(4))
    gpu_ctx.gen_build(42, n, dk, dv)
    gpu_ctx.gen_probe(42, 42 ^ 0xDEADBEEF, n, n, m, sk, sv)
    torch.cuda.synchronize()
    p = bj.default_params(r_size=n, s_size=n, inputs_on_device=1, **kw)
    print(f"[large_join] n={n} {kw}", flush=True)
    res = gpu_ctx.join(p, dk, dv, sk, sv)
    assert res.match_count == m and res.rows_written == m
    k, a, b = gpu_ctx.rows_to_host(res)
    skh, dkh = to_host(sk, n), to_host(dk, n)
    expect_s = np.nonzero(skh < np.uint32(1 << 31))[0]         # keys below 2^31 match
    assert len(expect_s) == m
    assert np.array_equal(np.sort(b.astype(np.int64)), expect_s)
    assert np.array_equal(skh[b], k)
    assert np.array_equal(dkh[a], k)


@pytest.mark.parametrize("kw", [dict(), dict(probe_breakdown=1), dict(s_chunks=2), dict(tile=4096)])
def test_probe_mixed_round_paths(gpu_ctx, kw):
    """Rounds of one work item that alternate between staged rows (at most 1024
    per round, stored two rounds later) and rows written from the lanes (a
    round holding a key with 3000 build duplicates): the rows in flight across
    a direct round, and at the end of each item, must all be written once."""
    rng = np.random.default_rng(11)
    heavy = np.uint32(0x01234567)
    rk = np.concatenate([np.arange(1, 47_001, dtype=np.uint32) * np.uint32(40_009),
                         np.full(3000, heavy, np.uint32)])
    rv = np.arange(len(rk), dtype=np.uint32)
    sk = rng.integers(1, 2**31, 2_000_000, dtype=np.uint64).astype(np.uint32)
    sk[rng.choice(len(sk), 200_000, replace=False)] = rk[rng.integers(0, 47_000, 200_000)]
    sk[rng.choice(len(sk), 25, replace=False)] = heavy
    sv = np.arange(len(sk), dtype=np.uint32)
    n, k, a, b = O.cpu_join(rk, rv, sk, sv)
    assert n > 25 * 3000
    res, rows = gpu_join(gpu_ctx, rk, rv, sk, sv, answer_scale=float(n) / len(sk) + 1.0, **kw)
    assert res.match_count == n and res.truncated == 0
    assert np.array_equal(rows, O.sort_rows(k, a, b))

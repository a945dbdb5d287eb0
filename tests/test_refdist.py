"""The reference datagen's distribution (datagen/gen_dataset.cpp:79-263).

CPU: bj_gen_refdist_host reproduces the reference generator's files byte for
byte -- checked against the golden fixtures written by the reference's own
binary (refgen_2000: 2000 x 2000 at 10%; refgen_max_3000: 3000 x 5000 at 7.5%
with --max 15), and, where oracle/_ref is built, against that binary on more
sizes and seeds.  The CLI bj_gendata writes the same files.  Closed forms: the
join size equals k = floor(R * pct / 100), the matching keys are [0, k).
GPU: bj_gen_refdist holds the same key multiset per region (sorted keys equal
the host generator's) and the join of its relations is the oracle's.
"""
import os
import subprocess
import tempfile

import numpy as np
import pytest

import brigjoin as bj
import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))


def fixture(name):
    return np.load(os.path.join(HERE, "golden", name + ".npz"), allow_pickle=False)


@pytest.mark.parametrize("name,args", [
    ("refgen_2000", (2000, 2000, 10.0, 42, -1.0)),
    ("refgen_max_3000", (3000, 5000, 7.5, 9, 15.0)),
])
def test_host_generator_matches_reference_files(name, args):
    z = fixture(name)
    rk, sk, k = bj.gen_refdist_host(*args)
    assert np.array_equal(rk, z["r_key"]) and np.array_equal(sk, z["s_key"])
    assert k == int(z["count"]) == int(z["ref_count"])


def test_host_generator_closed_forms():
    R, S, pct = 10_000, 7_000, 33.3
    rk, sk, k = bj.gen_refdist_host(R, S, pct, seed=5)
    assert k == int(R * pct / 100.0)
    assert np.array_equal(np.sort(rk), np.concatenate([np.arange(k), k + np.arange(R - k)]).astype(np.uint32))
    assert np.array_equal(np.sort(sk), np.concatenate([np.arange(k), R + np.arange(S - k)]).astype(np.uint32))
    assert bj.cpu_sanity_count(rk, sk) == k
    # the hot first quarter of S holds floor(hot_n * max / 100) matches
    rk, sk, k = bj.gen_refdist_host(R, 20_000, 10.0, seed=3, max_pct=18.0)
    hot = sk[: 5000]
    assert int((hot < k).sum()) == int(5000 * 18.0 / 100.0) and int((sk < k).sum()) == k
    # k is capped at |S|
    _, _, k = bj.gen_refdist_host(100, 10, 50.0)
    assert k == 10


@pytest.mark.parametrize("args", [
    (10, 10, 101.0, 42, -1.0),     # pct out of range
    (10, 10, -1.0, 42, -1.0),
    (3000, 5000, 7.5, 9, 40.0),    # hot region needs more matches than exist
    (3000, 5000, 7.5, 9, 5.0),     # max below the mean
])
def test_host_generator_errors(args):
    with pytest.raises(bj.BjError):
        bj.gen_refdist_host(*args)


@pytest.mark.skipif(not os.path.exists(O.REF_GEN_PATH), reason="oracle/_ref/gen_dataset not built")
@pytest.mark.parametrize("R,S,pct,seed,mx", [
    (1, 1, 100.0, 42, None), (0, 50, 10.0, 42, None), (12345, 6789, 3.3, 7, None),
    (40000, 90000, 25.0, 1234567, None), (20000, 60000, 25.0, 11, 30.0), (4096, 4096, 0.0, 42, None),
])
def test_host_generator_against_reference_binary(R, S, pct, seed, mx):
    with tempfile.TemporaryDirectory() as d:
        extra = ["--max", str(mx)] if mx is not None else []
        subprocess.run([O.REF_GEN_PATH, str(R), str(S), str(pct), d, str(seed)] + extra, check=True,
                       stdout=subprocess.DEVNULL)
        tag = f"{pct:.1f}" + (f"_max{mx:.1f}" if mx is not None else "")
        ref_r = np.fromfile(os.path.join(d, f"gen_R{R}_ans{tag}pct.bin"), np.uint32)
        ref_s = np.fromfile(os.path.join(d, f"gen_S{S}_ans{tag}pct.bin"), np.uint32)
    rk, sk, _ = bj.gen_refdist_host(R, S, pct, seed, -1.0 if mx is None else mx)
    assert np.array_equal(rk, ref_r) and np.array_equal(sk, ref_s)


def test_gendata_cli_writes_reference_files(tmp_path):
    z = fixture("refgen_max_3000")
    r = subprocess.run([bj.GENDATA_PATH, "3000", "5000", "7.5", str(tmp_path), "9", "--max", "15", "--verify"],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    assert "Selectivity is CORRECT" in r.stdout
    rk = np.fromfile(tmp_path / "gen_R3000_ans7.5_max15.0pct.bin", np.uint32)
    sk = np.fromfile(tmp_path / "gen_S5000_ans7.5_max15.0pct.bin", np.uint32)
    assert np.array_equal(rk, z["r_key"]) and np.array_equal(sk, z["s_key"])
    r = subprocess.run([bj.GENDATA_PATH, "3000", "5000", "7.5", str(tmp_path), "9", "--max", "40"],
                       capture_output=True, text=True)
    assert r.returncode == 1


@pytest.mark.gpu
@pytest.mark.parametrize("R,S,pct,mx", [(2000, 2000, 10.0, -1.0), (300_000, 500_000, 10.0, -1.0),
                                        (1_000_000, 1_000_000, 20.0, 60.0), (1, 3, 100.0, -1.0)])
def test_device_generator_multiset_and_join(gpu_ctx, R, S, pct, mx):
    import torch
    print(f"[refdist] R={R} S={S} pct={pct} max={mx}", flush=True)
    rk_h, sk_h, k = bj.gen_refdist_host(R, S, pct, seed=42, max_pct=mx)
    dr = torch.empty(max(R, 1), dtype=torch.int32, device="cuda")
    ds = torch.empty(max(S, 1), dtype=torch.int32, device="cuda")
    k2 = gpu_ctx.gen_refdist(R, S, pct, dr, ds, seed=42, max_pct=mx)
    assert k2 == k
    rk = dr[:R].cpu().numpy().view(np.uint32)
    sk = ds[:S].cpu().numpy().view(np.uint32)
    assert np.array_equal(np.sort(rk), np.sort(rk_h))
    hot = int(S * 0.25) if mx >= 0 else 0
    assert np.array_equal(np.sort(sk[:hot]), np.sort(sk_h[:hot]))
    assert np.array_equal(np.sort(sk[hot:]), np.sort(sk_h[hot:]))
    if R >= 1000:
        assert not np.array_equal(rk, np.sort(rk))   # a random order, not the unshuffled one
    rv = np.arange(R, dtype=np.uint32)
    sv = np.arange(S, dtype=np.uint32)
    p = bj.default_params(r_size=R, s_size=S, inputs_on_device=1, use_bf=1)
    res = gpu_ctx.join(p, dr, torch.from_numpy(rv.view(np.int32)).cuda(), ds,
                       torch.from_numpy(sv.view(np.int32)).cuda())
    n, ek, ea, eb = O.cpu_join(rk, rv, sk, sv)
    assert res.match_count == n == k
    got = O.sort_rows(*gpu_ctx.rows_to_host(res))
    assert np.array_equal(got, O.sort_rows(ek, ea, eb))

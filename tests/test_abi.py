"""C-ABI surface of libbrigjoin without any device call (CPU tests).

The library loads, exports every function include/bj_api.h declares, derives
the reference's parameters (bj_plan), reports errors as codes, writes and reads
output.bin, and its CLI keeps the reference driver's argument surface.
"""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

import brigjoin as bj
import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "bj_api.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(bj_\w+)\s*\(", text)))


def test_library_loads_and_exports_every_declared_symbol():
    L = bj.lib()
    names = declared_functions()
    assert len(names) >= 15
    for name in names:
        assert hasattr(L, name), f"{name} declared in bj_api.h but not exported"
    assert set(names) == set(bj.EXPORTS)
    out = subprocess.run(["nm", "-D", "--defined-only", bj.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    for name in names:
        assert re.search(rf"\bT {name}$", out, flags=re.M), name


def test_version_and_strerror():
    L = bj.lib()
    assert L.bj_version() == 4
    assert L.bj_strerror(0) == b"ok"
    for code in range(-7, 0):
        assert L.bj_strerror(code) not in (b"ok", b"unknown error")
    assert L.bj_strerror(-99) == b"unknown error"


def test_params_default():
    p = bj.default_params()
    assert p.partition_size == -1 and p.second_partition_size == -1 and p.radix_bits == -1
    assert p.num_bits == -1 and p.num_hashes == -1 and p.use_bf == 0
    assert p.answer_scale == 1.0 and abs(p.true_false_ratio - 0.1) < 1e-12


def test_plan_config1():
    # R=S=16M, Bloom off, 10 radix bits at the first level
    info = bj.plan(bj.default_params(r_size=16_000_000, s_size=16_000_000, partition_size=1024))
    assert info.bits1 == 10 and info.bits2 >= 1 and info.bf == 0
    per = 16_000_000 / (1 << (info.bits1 + info.bits2))
    assert per <= 0.96 * info.tile and per > 0.48 * info.tile
    assert info.capacity == 16_000_000


def test_plan_config2():
    # R=S=640M, 14 total bits with partition_size 512, Bloom 10 bits/key, k=3
    p = bj.default_params(r_size=640_000_000, s_size=640_000_000, partition_size=512, radix_bits=14,
                          use_bf=1, bf_bits_per_key=10.0, num_hashes=3, inputs_on_device=1)
    info = bj.plan(p)
    assert (info.bits1, info.bits2) == (9, 5)
    assert info.bf == 1 and info.bf_k == 3
    assert info.bf_seg_bits == 14                      # 800 MB / 2^14 = 48.8 KB segments in LDS
    assert abs(info.bf_words * 32 / 640e6 - 10.0) < 0.01
    assert info.workspace_bytes < 64 * 2**30           # fits one MI355X with room to spare


def test_plan_reference_bloom_formula():
    # no bits/key: num_bits = -R ln p / ln^2 2, k = (int)(bits/R ln 2)  (src/hash_kernel.cu:189-193)
    import math
    R = 1_000_000
    info = bj.plan(bj.default_params(r_size=R, s_size=R, use_bf=14, true_false_ratio=0.1))
    bits = -R * math.log(0.1) / math.log(2) ** 2
    assert info.bf_k == int(bits / R * math.log(2))
    assert info.bf_words * 32 >= bits
    info2 = bj.plan(bj.default_params(r_size=R, s_size=R, use_bf=10, num_bits=6_400_000,
                                      num_hashes=5))
    assert info2.bf_k == 5 and info2.bf_words * 32 >= 6_400_000


@pytest.mark.parametrize("kw,code", [
    (dict(partition_size=3), -1),
    (dict(partition_size=4096), -1),
    (dict(second_partition_size=6), -1),
    (dict(use_bf=11), -1),
    (dict(use_bf=17), -1),
    (dict(radix_bits=16), -5),
    (dict(tile=1000), -1),
    (dict(use_bf=1, num_hashes=9), -1),
    (dict(use_bf=1, true_false_ratio=1.5), -1),
    (dict(r_size=2**31), -5),
    (dict(r_size=-1), -1),
    (dict(partition_size=512, second_partition_size=32, radix_bits=12), -1),
    # not finite numbers, and a row capacity whose conversion would overflow
    (dict(use_bf=1, bf_bits_per_key=float("nan")), -1),
    (dict(use_bf=1, bf_bits_per_key=float("inf")), -1),
    (dict(answer_scale=float("nan")), -1),
    (dict(answer_scale=float("inf")), -1),
    (dict(answer_scale=1e30), -5),
    (dict(s_chunks=-2), -1),
])
def test_plan_rejects_invalid(kw, code):
    base = dict(r_size=1000, s_size=1000)
    base.update(kw)
    p = bj.default_params(**base)
    info = bj.BjPlanInfo()
    assert bj.lib().bj_plan(ctypes.byref(p), ctypes.byref(info)) == code
    assert bj.lib().bj_last_error()


def test_create_without_device_reports_nodevice():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    h = ctypes.c_void_p()
    assert bj.lib().bj_create(ctypes.byref(h), 0) == -6


def test_output_bin_roundtrip(tmp_path):
    rng = np.random.default_rng(3)
    n = 10000
    k = rng.integers(0, 50, n, dtype=np.uint32)
    a = rng.integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32)
    b = rng.integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32)
    path = str(tmp_path / "output.bin")
    bj.write_output(path, k, a, b)
    raw = open(path, "rb").read()
    assert raw[:8] == b"BJOUT001" and int.from_bytes(raw[8:16], "little") == n
    assert len(raw) == 16 + 12 * n
    k2, a2, b2 = bj.read_output(path)
    assert np.array_equal(np.stack([k2, a2, b2], 1), O.sort_rows(k, a, b))
    # two writes of the same multiset in different orders are byte-identical
    perm = rng.permutation(n)
    path2 = str(tmp_path / "output2.bin")
    bj.write_output(path2, k[perm], a[perm], b[perm])
    assert open(path2, "rb").read() == raw


def test_output_bin_empty_and_truncated(tmp_path):
    path = str(tmp_path / "empty.bin")
    bj.write_output(path, np.zeros(0, np.uint32), np.zeros(0, np.uint32), np.zeros(0, np.uint32))
    k, a, b = bj.read_output(path)
    assert len(k) == 0
    bj.write_output(path, np.arange(5, dtype=np.uint32), np.arange(5, dtype=np.uint32),
                    np.arange(5, dtype=np.uint32))
    raw = open(path, "rb").read()
    open(path, "wb").write(raw[:-3])
    with pytest.raises(bj.BjError):
        bj.read_output(path)
    open(path, "wb").write(b"NOTMAGIC" + raw[8:])
    with pytest.raises(bj.BjError):
        bj.read_output(path)


def test_cpu_sanity_count_matches_oracle():
    rng = np.random.default_rng(11)
    rk = rng.integers(0, 3000, 20000, dtype=np.uint32)
    sk = rng.integers(0, 6000, 25000, dtype=np.uint32)
    n, *_ = O.cpu_join(rk, np.arange(len(rk), dtype=np.uint32), sk,
                       np.arange(len(sk), dtype=np.uint32), cap=0)
    assert bj.cpu_sanity_count(rk, sk) == n
    # every uint32 key is legal (the reference's INT_MIN marker is not needed)
    rk2 = np.array([0x80000000, 5, 0x80000000], np.uint32)
    sk2 = np.array([0x80000000, 5], np.uint32)
    assert bj.cpu_sanity_count(rk2, sk2) == 3


def test_gpu_main_io_errors(tmp_path):
    missing = str(tmp_path / "nope.bin")
    args = bj.main_args(r_path=missing, s_path=missing, out_path="", r_size=10, s_size=10,
                        verbose=0)
    res = bj.BjResult()
    assert bj.lib().bj_gpu_main(ctypes.byref(args), ctypes.byref(res)) == -4
    # a key file shorter than R, and a missing value block with read_val
    short = tmp_path / "short.bin"
    np.arange(5, dtype=np.uint32).tofile(short)
    args = bj.main_args(r_path=str(short), s_path=str(short), out_path="", r_size=10, s_size=5,
                        verbose=0)
    assert bj.lib().bj_gpu_main(ctypes.byref(args), ctypes.byref(res)) == -4
    args = bj.main_args(r_path=str(short), s_path=str(short), out_path="", r_size=5, s_size=5,
                        read_val=1, verbose=0)
    assert bj.lib().bj_gpu_main(ctypes.byref(args), ctypes.byref(res)) == -4


def test_driver_cli_surface(tmp_path):
    drv = bj.DRIVER_PATH
    assert os.path.exists(drv)
    r = subprocess.run([drv], capture_output=True, text=True)
    assert r.returncode == 2 and "rPath sPath outPath" in r.stdout
    # the reference README command line parses (it then fails on the missing files)
    cmd = [drv, str(tmp_path / "R.bin"), str(tmp_path / "S.bin"), str(tmp_path / "output.bin"),
           "640000000", "640000000", "0", "1024", "14", "1", "1", "0.1", "1", "1", "1.4",
           "--partition_size", "512", "--fused_probe_breakdown", "1", "--r_replicate_ratio", "4.5"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 1
    assert "unknown option" not in r.stderr
    assert "BF filtering is enabled" in r.stdout
    r = subprocess.run(cmd[:7] + ["1024", "0", "0", "0", "0.1", "0"], capture_output=True, text=True)
    assert r.returncode == 2 and "s_chunks must be >= 1" in r.stderr

"""The command line end to end (GPU tests): bj_gendata writes reference-format
files (byte-identical to datagen/gen_dataset's), bj_driver joins them with the
reference README's flags (src/driver.cpp:43-198, README.md:19-25) and writes
output.bin, whose rows must equal the oracle's sorted rows on the same files
with the reference's synthesized payloads R i+1, S i+100000001
(src/hash_kernel.cu:330-336).
"""
import os
import subprocess

import numpy as np
import pytest

import brigjoin as bj
import oracle as O

pytestmark = pytest.mark.gpu


def gendata(tmp_path, R, S, pct, extra=()):
    r = subprocess.run([bj.GENDATA_PATH, str(R), str(S), str(pct), str(tmp_path)] + list(extra),
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    tag = f"{pct:.1f}"
    return tmp_path / f"gen_R{R}_ans{tag}pct.bin", tmp_path / f"gen_S{S}_ans{tag}pct.bin"


def expected_rows(rpath, spath):
    rk = np.fromfile(rpath, np.uint32)
    sk = np.fromfile(spath, np.uint32)
    rv = (np.arange(len(rk)) + 1).astype(np.uint32)
    sv = (np.arange(len(sk)) + 100000001).astype(np.uint32)
    n, k, a, b = O.cpu_join(rk, rv, sk, sv)
    return n, O.sort_rows(k, a, b)


@pytest.mark.parametrize("positional,flags", [
    # the README command line, sizes scaled down: sanity 0, chunk 1024, use_bf 14,
    # nvlink 1, warp compact 1, ratio 0.1, s_chunks 1, kernel fusion 1, skew 1.4
    (["0", "1024", "14", "1", "1", "0.1", "1", "1", "1.4"],
     ["--partition_size", "512", "--fused_probe_breakdown", "1", "--r_replicate_ratio", "4.5"]),
    # CPU sanity check on, S in 3 chunks, no filter
    (["1", "1024", "0", "0", "0", "0.1", "3"], []),
])
def test_driver_joins_reference_files(tmp_path, positional, flags):
    R, S = 300_000, 450_000
    rpath, spath = gendata(tmp_path, R, S, 10.0)
    out = tmp_path / "output.bin"
    cmd = [bj.DRIVER_PATH, str(rpath), str(spath), str(out), str(R), str(S)] + positional + flags
    print("[driver]", " ".join(cmd[5:]), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    n, expect = expected_rows(rpath, spath)
    assert n == R // 10
    assert f"=== Grand Total Answer Count ===\n{n}\n" in r.stdout
    if positional[0] == "1":
        assert "MATCH" in r.stdout and "MISMATCH" not in r.stdout
    if "--fused_probe_breakdown" in flags:
        assert "[2-2c] Probe count" in r.stdout
    k, a, b = bj.read_output(str(out))
    assert np.array_equal(np.stack([k, a, b], 1), expect)
    # the file is the sorted rows: byte-identical to one written from the oracle's rows
    ref = tmp_path / "expect.bin"
    bj.write_output(str(ref), expect[:, 0].copy(), expect[:, 1].copy(), expect[:, 2].copy())
    assert out.read_bytes() == ref.read_bytes()

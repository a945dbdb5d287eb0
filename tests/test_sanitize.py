"""Host code under AddressSanitizer and UndefinedBehaviorSanitizer (CPU tests).

tests/sanitize/Makefile builds the library's host sources (planner, batch
planner, serving loop and its arena allocator, file I/O, generators, the CPU
count, the command line front end) with both sanitizers, against the unchanged
gfx950 kernel objects of the library's own build.  host_check exercises every
entry point that runs without a GPU, and the error paths of the ones that need
one; the sanitized bj_driver reads reference-format files written by bj_gendata
(the reference datagen's layout, datagen/gen_dataset.cpp:28-35) and damaged ones.
Any sanitizer report aborts the program, which fails the test.
"""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SAN = os.path.join(ROOT, "tests", "sanitize")
PKG = os.path.join(ROOT, "brig-join_amd")
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=1",
           UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1", HIP_VISIBLE_DEVICES="")


@pytest.fixture(scope="module")
def built():
    kobj = os.path.join(PKG, "_build", "bj_kernels.o")
    if not os.path.exists(kobj):
        pytest.skip("the library is not built (__graft_entry__.build())")
    r = subprocess.run(["make", "-C", SAN, "-j8", "all"], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-4000:]
    return os.path.join(SAN, "bin")


def run(args, timeout=300):
    r = subprocess.run(args, capture_output=True, text=True, timeout=timeout, env=ENV)
    assert "Sanitizer" not in r.stderr and "runtime error" not in r.stderr, r.stderr[-4000:]
    return r


def test_host_check(built):
    r = run([os.path.join(built, "host_check")])
    assert r.returncode == 0, r.stdout + r.stderr[-4000:]
    assert "host_check ok" in r.stdout


def test_driver_under_sanitizers(built, tmp_path):
    drv = os.path.join(built, "bj_driver")
    # no arguments, and too few: usage, non-zero status
    assert run([drv]).returncode != 0
    assert run([drv, "a", "b"]).returncode != 0
    # reference-format files from the generator, then the README flags; without
    # a GPU the driver reads both files and stops at the device
    gen = os.path.join(PKG, "bin", "bj_gendata")
    g = subprocess.run([gen, "3000", "5000", "10", str(tmp_path)], capture_output=True, text=True,
                       timeout=120, env=ENV)
    assert g.returncode == 0, g.stderr
    rpath, spath = tmp_path / "gen_R3000_ans10.0pct.bin", tmp_path / "gen_S5000_ans10.0pct.bin"
    assert rpath.stat().st_size == 3000 * 4 and spath.stat().st_size == 5000 * 4
    out = tmp_path / "out.bin"
    base = [drv, str(rpath), str(spath), str(out), "3000", "5000", "1", "1024", "14", "1", "1", "0.1", "1",
            "1", "1.4", "--partition_size", "512", "--num_bits=48000", "--num_hashes", "3"]
    r = run(base)
    assert r.returncode != 0 or out.exists()
    # a value block that is missing, sizes larger than the files, a file that is
    # not there: refused, no sanitizer report
    assert run(base + ["--read_val", "1"]).returncode != 0
    big = list(base)
    big[4], big[5] = "4000", "9000"
    assert run(big).returncode != 0
    missing = list(base)
    missing[1] = str(tmp_path / "nope.bin")
    assert run(missing).returncode != 0
    # a truncated key file
    np.arange(10, dtype=np.uint32).tofile(tmp_path / "short.bin")
    short = list(base)
    short[1] = str(tmp_path / "short.bin")
    assert run(short).returncode != 0

"""The benchmark's byte model and roofline grouping (CPU tests).

The whole-join algorithmic bytes of the three single-join workloads are pinned
to the closed forms B = 40 (N_R + N_S) + 12 M without the filter and
B = 44 N_R + F + 8 N_S + 48 M_S + 12 M with it (two partition passes; F the
filter bytes, M_S the matching S tuples), and to the values the measurement
plan derived from them by hand (1.299e9, 3.792e10, 1.392e10).  The per-kernel
roofline sums a kernel's launches per step and names the kernel with the most
device time.
"""
import importlib.util
import os

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def bench():
    spec = importlib.util.spec_from_file_location("bench_under_test", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


@pytest.mark.parametrize("nr,ns,m,bpk,planned", [
    (16_000_000, 16_000_000, 1_600_000, 0.0, 1.299e9),
    (640_000_000, 640_000_000, 64_000_000, 10.0, 3.792e10),
    (64_000_000, 1_280_000_000, 12_800_000, 12.0, 1.392e10),
])
def test_whole_join_bytes(bench, nr, ns, m, bpk, planned):
    f = int(nr * bpk / 8)
    got = bench.total_algorithmic_bytes(nr, ns, m, m, f, bpk > 0)
    want = 44 * nr + f + 8 * ns + 48 * m + 12 * m if bpk > 0 else 40 * (nr + ns) + 12 * m
    assert got == want
    assert abs(got - planned) / planned < 1e-3


def test_per_launch_bytes(bench):
    n, m = 1000, 100
    ab = lambda k, st: bench.algorithmic_bytes(k, n, 2 * n, m, m, 800, True, st)  # noqa: E731
    # a partition pass reads and writes 8 B per tuple; the filter reads 8 B per S tuple and
    # writes 8 B per survivor; the probe reads 8 B per R and S' tuple and writes 12 B per row
    assert ab("scatter1_r", 2) == 16 * n and ab("hist_r", 2) == 4 * n
    assert ab("bf_build", 2) == 4 * n + 800 and ab("bf_filter", 2) == 8 * 2 * n + 8 * m
    assert ab("probe", 2) == 8 * n + 8 * m + 12 * m
    # stage 2: S's first pass covers all of S and the filter places S' (no second S pass)
    assert ab("scatter1_s", 2) == 16 * 2 * n and ab("scatter2_s", 2) == 0
    # stage 1: S' alone is partitioned, in two passes
    assert ab("scatter1_s", 1) == 16 * m and ab("scatter2_s", 1) == 16 * m
    assert ab("plan", 2) == 0 and ab("scan_r", 2) == 0


def test_kernel_rooflines_groups_launches(bench, monkeypatch):
    monkeypatch.setattr(bench, "load_traffic", lambda w, slot: {"scatter1_r": 30, "scatter2_r": 34,
                                                                "scatter1_s": 38, "bf_filter": 50}.get(slot))
    kern = {"hist_r": 0.4, "scan_r": 0.02, "scatter1_r": 2.0, "scatter2_r": 2.2, "scatter1_s": 2.6,
            "scatter2_s": 0.005, "bf_filter": 3.5, "probe": 1.9, "plan": 0.02}
    alg = {"hist_r": 4.0e9, "scatter1_r": 10.0e9, "scatter2_r": 10.0e9, "scatter1_s": 10.0e9,
           "bf_filter": 6.0e9, "probe": 8.0e9}
    out = bench.kernel_rooflines(kern, lambda k: alg.get(k, 0.0), 2, "w")
    assert set(out) == {"k_hist", "k_scatter", "k_filter_l1", "k_probe"}
    sc = out["k_scatter"]
    assert sc["launches"] == 3 and sc["slots"] == ["scatter1_r", "scatter2_r", "scatter1_s"]
    assert sc["step_ms"] == pytest.approx(6.8) and sc["avg_ms"] == pytest.approx(6.8 / 3, abs=1e-4)
    assert sc["alg_bytes"] == pytest.approx(10.0e9)
    assert sc["achieved"] == pytest.approx(30.0e9 / 6.8e-3 / 1e9, abs=0.1)
    assert sc["frac"] == pytest.approx(sc["achieved"] / bench.HBM_PEAK_GBS, abs=1e-4)
    assert sc["traffic"] == (30 + 34 + 38) // 3
    assert out["k_hist"]["traffic"] is None          # not recorded for that slot
    assert max(out, key=lambda k: out[k]["step_ms"]) == "k_scatter"
    # a filter before S's partition is the prefilter kernel
    assert "k_bloom_prefilter" in bench.kernel_rooflines(kern, lambda k: alg.get(k, 0.0), 1, "w")

"""Pins of the CPU oracle (no GPU).

Every oracle function is checked against something it does not itself compute:
golden fixtures whose rows are a numpy brute force and whose counts come from the
reference's own run_sanity_check (compiled from src/sanity_check.cpp into
oracle/_ref); the reference's own dataset generator; closed forms (the
generator's exact match count); bijectivity by exhaustion; and brute force on
tiny inputs.
"""
import glob
import os

import numpy as np
import pytest

import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = sorted(glob.glob(os.path.join(HERE, "golden", "*.npz")))


def load(path):
    z = np.load(path, allow_pickle=False)
    return {k: z[k] for k in z.files}


@pytest.mark.parametrize("path", GOLDEN, ids=[os.path.basename(p)[:-4] for p in GOLDEN])
def test_cpu_join_matches_golden_rows(path):
    g = load(path)
    n, k, a, b = O.cpu_join(g["r_key"], g["r_val"], g["s_key"], g["s_val"])
    assert n == int(g["count"])
    assert np.array_equal(O.sort_rows(k, a, b), g["rows"].reshape(-1, 3))


@pytest.mark.parametrize("path", GOLDEN, ids=[os.path.basename(p)[:-4] for p in GOLDEN])
def test_sanity_count_matches_reference_count(path):
    """The count-only restatement reproduces the reference's printed count exactly,
    including its INT_MIN-as-EMPTY behaviour (fixture intmin_in_r)."""
    g = load(path)
    if len(g["r_key"]) == 0 or len(g["s_key"]) == 0:
        assert int(g["ref_count"]) == 0
        return
    assert O.sanity_count(g["r_key"], g["s_key"]) == int(g["ref_count"])


def test_intmin_quirk_is_the_references():
    g = load(os.path.join(HERE, "golden", "intmin_in_r.npz"))
    assert int(g["ref_count"]) != int(g["count"])      # the reference loses INT_MIN keys of R
    assert O.sanity_count(g["r_key"], g["s_key"]) == int(g["ref_count"])


@pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref not built (needs the reference)")
@pytest.mark.parametrize("seed", [1, 7, 42])
def test_sanity_count_vs_reference_binary_random(seed):
    rng = np.random.default_rng(seed)
    rk = rng.integers(0, 5000, 20000, dtype=np.uint32)      # duplicates on both sides
    sk = rng.integers(0, 8000, 30000, dtype=np.uint32)
    ref, _ = O.ref_sanity_count(rk, sk)
    assert O.sanity_count(rk, sk) == ref
    n, *_ = O.cpu_join(rk, np.arange(len(rk), dtype=np.uint32), sk,
                       np.arange(len(sk), dtype=np.uint32), cap=0)
    assert n == ref


def test_cpu_join_vs_brute_force_adversarial():
    rng = np.random.default_rng(5)
    for trial in range(20):
        pool = np.array([0, 1, 0x80000000, 0xFFFFFFFF, 0x7FFFFFFF, 3, 4], np.uint32)
        rk = rng.choice(pool, rng.integers(0, 40)).astype(np.uint32)
        sk = rng.choice(pool, rng.integers(0, 40)).astype(np.uint32)
        rv = rng.integers(0, 2**32, len(rk), dtype=np.uint64).astype(np.uint32)
        sv = rng.integers(0, 2**32, len(sk), dtype=np.uint64).astype(np.uint32)
        n1, k1, a1, b1 = O.cpu_join(rk, rv, sk, sv)
        n2, k2, a2, b2 = O.nested_loop(rk, rv, sk, sv)
        # and a third, independent statement: numpy outer equality
        eq = np.equal.outer(sk, rk)
        assert n1 == n2 == int(eq.sum())
        assert np.array_equal(O.sort_rows(k1, a1, b1), O.sort_rows(k2, a2, b2))


def test_cpu_join_truncation_keeps_exact_count():
    rk = np.full(10, 5, np.uint32)
    sk = np.full(10, 5, np.uint32)
    n, k, a, b = O.cpu_join(rk, np.arange(10, dtype=np.uint32), sk,
                            np.arange(10, dtype=np.uint32), cap=7)
    assert n == 100 and len(k) == 7
    assert np.all(k == 5)


@pytest.mark.parametrize("w", list(range(1, 27)))
def test_permutation_is_bijective_exhaustive(w):
    """Every width runs the same code with a different mask: exhaust all 2^w inputs."""
    assert O.permw_defects(w, 42) == 0
    if w <= 20:
        assert O.permw_defects(w, 0xDEADBEEF) == 0


def test_permutation_31_bits_has_no_collisions():
    """2^24 distinct inputs at w=31: a map that is not injective would collide here
    with near certainty (a random function gives ~65k collisions)."""
    rk, _ = O.gen_r(42, 1 << 24)
    assert len(np.unique(rk)) == len(rk)


@pytest.mark.parametrize("n_r,n_s,ratio", [(1, 1, 1.0), (1000, 1000, 0.1), (4096, 20000, 0.01),
                                           (50000, 30000, 0.5), (1 << 20, 1 << 20, 0.1)])
def test_generator_closed_form(n_r, n_s, ratio):
    rk, rv = O.gen_r(42, n_r)
    m = O.match_count(n_s, ratio)
    sk, sv = O.gen_s(42, 42 ^ 0xDEADBEEF, n_r, n_s, m)
    assert np.array_equal(rv, np.arange(n_r, dtype=np.uint32))
    assert np.array_equal(sv, np.arange(n_s, dtype=np.uint32))
    assert rk.min() >= 1 and rk.max() < 2**31
    assert len(np.unique(rk)) == n_r                        # distinct R keys
    match = sk < 2**31
    assert int(match.sum()) == m                            # exactly m_s matching positions
    assert np.all(np.isin(sk[match], rk))                   # drawn from R
    hi = sk[~match]
    assert len(np.unique(hi)) == len(hi)                    # distinct non-matching keys
    # the join size is m_s in closed form
    if n_r * n_s <= 4e10:
        n, *_ = O.cpu_join(rk, rv, sk, sv, cap=0)
        assert n == m


def test_generator_reference_datagen_files_roundtrip(tmp_path):
    """The reference datagen's files (raw int32) give exactly k = R*pct/100 matches."""
    if not O.ref_available():
        pytest.skip("oracle/_ref not built")
    import subprocess
    subprocess.run([O.REF_GEN_PATH, "30000", "50000", "10.0", str(tmp_path)], check=True,
                   capture_output=True)
    rk = np.fromfile(tmp_path / "gen_R30000_ans10.0pct.bin", np.uint32)
    sk = np.fromfile(tmp_path / "gen_S50000_ans10.0pct.bin", np.uint32)
    assert O.sanity_count(rk, sk) == 3000 == O.ref_sanity_count(rk, sk)[0]


@pytest.mark.parametrize("bpk,k,lo,hi", [(10, 3, 0.012, 0.035), (12, 4, 0.005, 0.025),
                                         (4, 2, 0.08, 0.20)])
def test_bloom_no_false_negatives_and_fpr(bpk, k, lo, hi):
    """Register-blocked (64-bit) filter: no false negatives; the false-positive rate
    matches the Poisson mixture over keys per block, sum_j P(j) (1-(63/64)^(k j))^k,
    computed here independently of the filter code."""
    import math
    n = 200000
    rk, _ = O.gen_r(7, n)
    seg_bits = 3
    blocks = n * bpk // 64
    wps = max(1, blocks >> seg_bits)
    f = O.bloom_build(rk, seg_bits, wps, k)
    assert O.bloom_query(f, rk[:5000], seg_bits, wps, k).all()
    probe = (np.arange(20000, dtype=np.uint64) + 2**31).astype(np.uint32)
    fpr = O.bloom_query(f, probe, seg_bits, wps, k).mean()
    lam = n / (wps << seg_bits)
    theory = sum(math.exp(-lam) * lam ** j / math.factorial(j) *
                 (1 - (63 / 64) ** (k * j)) ** k for j in range(60))
    assert lo < theory < hi
    assert abs(fpr - theory) < 0.25 * theory + 0.003

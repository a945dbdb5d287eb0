"""Writes the golden fixtures of tests/golden/ (run in the build container).

Each fixture holds inputs and expected outputs:
  r_key, r_val, s_key, s_val   the relations
  rows                         every (key, r_val, s_val) match, sorted; computed
                               here by a numpy brute force over all pairs
  count                        len(rows)
  ref_count                    the match count printed by the reference's own
                               run_sanity_check (src/sanity_check.cpp), compiled
                               from its source into oracle/_ref by `make -C oracle ref`
The refgen_* fixtures use the reference's own dataset generator
(datagen/gen_dataset.cpp, compiled into oracle/_ref/gen_dataset) and keeps its raw
int32 files as they are written (reference input format).
"""
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import oracle as O  # noqa: E402


def brute(rk, rv, sk, sv):
    eq = np.equal.outer(sk, rk)           # [S, R]
    si, ri = np.nonzero(eq)
    rows = np.stack([sk[si], rv[ri], sv[si]], axis=1).astype(np.uint32)
    if len(rows):
        order = np.lexsort((rows[:, 2], rows[:, 1], rows[:, 0]))
        rows = rows[order]
    return rows.reshape(-1, 3)


def save(name, rk, rv, sk, sv):
    rk, rv, sk, sv = (np.asarray(x, np.uint32) for x in (rk, rv, sk, sv))
    rows = brute(rk, rv, sk, sv)
    ref_count, _ = O.ref_sanity_count(rk, sk) if (len(rk) and len(sk)) else (0, 0.0)
    np.savez_compressed(os.path.join(HERE, name + ".npz"), r_key=rk, r_val=rv, s_key=sk, s_val=sv,
                        rows=rows, count=np.int64(len(rows)), ref_count=np.int64(ref_count))
    print(f"{name}: |R|={len(rk)} |S|={len(sk)} rows={len(rows)} ref_count={ref_count}")


def main():
    if not O.ref_available():
        raise SystemExit("build oracle/_ref first: make -C oracle ref")
    rng = np.random.default_rng(20241217)
    # duplicates on both sides, the edge keys 0 and 0xFFFFFFFF (no INT_MIN in R:
    # the reference's table uses INT_MIN as its EMPTY marker)
    base = np.array([0, 1, 2, 3, 0xFFFFFFFF, 0x7FFFFFFF, 12345, 0xDEADBEEF], np.uint64)
    rk = rng.choice(base, 40).astype(np.uint32)
    sk = np.concatenate([rng.choice(base, 30), [0x80000000, 0x80000000, 99]]).astype(np.uint32)
    save("dups_edges", rk, np.arange(len(rk)), sk, np.arange(len(sk)) + 1000)
    # INT_MIN in R: brute-force rows are the truth; the reference's count differs
    rk2 = np.array([0x80000000, 5, 0x80000000, 7], np.uint32)
    sk2 = np.array([0x80000000, 5, 5, 9, 0x80000000], np.uint32)
    save("intmin_in_r", rk2, np.arange(4), sk2, np.arange(5))
    # many-to-many and a skewed probe side
    save("many_to_many", [42, 42, 42, 7], [10, 11, 12, 13], [42, 42, 42, 42, 8], [0, 1, 2, 3, 4])
    save("skew_one_key", np.arange(1, 65), np.arange(64), np.full(300, 17), np.arange(300))
    save("empty_r", [], [], [1, 2, 3], [0, 1, 2])
    save("empty_s", [1, 2, 3], [0, 1, 2], [], [])
    # the synthetic generator (R distinct in [1,2^31), 10% of S drawn from R)
    n = 1500
    rk, rv = O.gen_r(42, n)
    sk, sv = O.gen_s(42, 42 ^ 0xDEADBEEF, n, n, O.match_count(n, 0.1))
    save("gen_1500", rk, rv, sk, sv)
    # the reference's own generator and file format: raw int32 keys
    with tempfile.TemporaryDirectory() as d:
        subprocess.run([O.REF_GEN_PATH, "2000", "2000", "10.0", d], check=True,
                       stdout=subprocess.DEVNULL)
        rk = np.fromfile(os.path.join(d, "gen_R2000_ans10.0pct.bin"), np.uint32)
        sk = np.fromfile(os.path.join(d, "gen_S2000_ans10.0pct.bin"), np.uint32)
    save("refgen_2000", rk, np.arange(len(rk)), sk, np.arange(len(sk)))
    refgen_max()


def refgen_max():
    """The reference generator with its --max load imbalance (hot first quarter of S)."""
    with tempfile.TemporaryDirectory() as d:
        subprocess.run([O.REF_GEN_PATH, "3000", "5000", "7.5", d, "9", "--max", "15"], check=True,
                       stdout=subprocess.DEVNULL)
        rk = np.fromfile(os.path.join(d, "gen_R3000_ans7.5_max15.0pct.bin"), np.uint32)
        sk = np.fromfile(os.path.join(d, "gen_S5000_ans7.5_max15.0pct.bin"), np.uint32)
    save("refgen_max_3000", rk, np.arange(len(rk)), sk, np.arange(len(sk)))


if __name__ == "__main__":
    if sys.argv[1:] == ["refgen_max"]:
        refgen_max()
    else:
        main()

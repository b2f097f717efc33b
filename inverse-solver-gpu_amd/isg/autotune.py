"""Build the autotuner cache for this GPU.

    python inverse-solver-gpu_amd/isg/autotune.py --levels 6 7 8 --nrhs 1 64 --out tuning.txt

For every tree depth and batch size, a mixed-precision context times every
admissible kernel variant of every stage (near field, P2M, M2L per level,
M2M / L2L per level) for both operator precisions and keeps the fastest; the
choices of all runs are merged into one text file (one line per choice:
tree-depth precision batch-class stage level value).
"""
import argparse
import os
import sys
import tempfile
import time

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
import isg  # noqa: E402


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--levels", type=int, nargs="+", default=[6, 7, 8])
    ap.add_argument("--nrhs", type=int, nargs="+", default=[1, 64])
    ap.add_argument("--reps", type=int, default=5)
    ap.add_argument("--out", default=isg.Solver.TUNING_FILE)
    a = ap.parse_args()
    lines = []
    for level in a.levels:
        for nrhs in a.nrhs:
            t0 = time.perf_counter()
            s = isg.Solver(level=level, max_rhs=nrhs, precision=1, tuning=False)
            s.autotune(nrhs, a.reps)
            with tempfile.TemporaryDirectory() as d:
                p = os.path.join(d, "t.txt")
                s.tuning_save(p)
                lines += [l for l in open(p) if not l.startswith("#")]
            s.close()
            print("level %d, %d rhs: tuned in %.1fs" % (level, nrhs, time.perf_counter() - t0), flush=True)
    with open(a.out, "w") as f:
        f.write("# isg tuning: tree-depth precision(0 fp64, 1 mixed) batch-class(0 nrhs<16, 1 nrhs>=16) "
                "stage level value\n")
        f.writelines(lines)
    print("wrote %s (%d choices)" % (a.out, len(lines)), flush=True)


if __name__ == "__main__":
    main()

"""Operator micro-benchmark: batched MLFMA matvecs at one configuration.

    python inverse-solver-gpu_amd/isg/kbench.py --level 8 --nrhs 64 --precision 1 --reps 3 [--copy-mib 1024]

Runs `reps` batched matvecs (nrhs right-hand sides, FP64 or FP32 operator)
with live kernel timing and prints one JSON line with the near-field and M2L
launch times.  --copy-mib adds a plain device-to-device copy of that many MiB
(torch), whose memory counters calibrate FETCH_SIZE / WRITE_SIZE readings of
the same profiler run.
"""
import argparse
import json
import os
import sys

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
import isg  # noqa: E402


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--level", type=int, default=8)
    ap.add_argument("--nrhs", type=int, default=64)
    ap.add_argument("--precision", type=int, default=1)
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--copy-mib", type=int, default=0)
    a = ap.parse_args()
    if a.copy_mib > 0:
        import torch
        src = torch.ones(a.copy_mib << 20, dtype=torch.uint8, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        print("copy %d MiB done" % a.copy_mib, file=sys.stderr, flush=True)
    s = isg.Solver(level=a.level, numrx=16, max_rhs=a.nrhs, precision=a.precision)
    rng = np.random.default_rng(0)
    X = (rng.standard_normal((a.nrhs, s.N)) + 1j * rng.standard_normal((a.nrhs, s.N))).astype(np.complex128)
    s.mlfma_many(X)  # warm-up
    s.profile(True)
    for _ in range(a.reps):
        s.mlfma_many(X)
    near = "near_fp32" if a.precision == 1 else "near"
    m2l = "m2l_fp32" if a.precision == 1 else "m2l"
    nms, nn = s.profile_read(near)
    mms, mn = s.profile_read(m2l)
    print(json.dumps({"level": a.level, "nrhs": a.nrhs, "precision": a.precision, "reps": a.reps,
                      "near_avg_ms": nms / max(nn, 1), "near_launches": nn,
                      "m2l_avg_ms": mms / max(mn, 1), "m2l_launchsets": mn}), flush=True)


if __name__ == "__main__":
    main()

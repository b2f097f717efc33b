"""Operator micro-benchmark: batched MLFMA matvecs at one configuration.

    python inverse-solver-gpu_amd/isg/kbench.py --level 8 --nrhs 64 --precision 1 --vec32 --reps 3 [--copy-mib 1024]

Runs `reps` batched matvecs (nrhs right-hand sides; FP64 operator, FP32
operator on FP64 vectors, or with --vec32 the FP32 operator on FP32 vectors
as the mixed-precision solver runs it) with live kernel timing and prints
one JSON line with the near-field, M2L, P2M and interpolation (M2M + L2L)
launch times.  --copy-mib adds a
plain device-to-device copy of that many MiB, whose memory counters
calibrate FETCH_SIZE / WRITE_SIZE readings of the same profiler run.
Device buffers come from torch, imported before the library starts HIP.
"""
import argparse
import json
import os
import sys

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--level", type=int, default=8)
    ap.add_argument("--nrhs", type=int, default=64)
    ap.add_argument("--precision", type=int, default=1)
    ap.add_argument("--vec32", action="store_true", help="FP32 vectors (needs --precision 1)")
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--copy-mib", type=int, default=0)
    ap.add_argument("--near", type=int, default=None,
                    help="force the near-field variant (16, 32 or 64 boxes per tile; 128: the split-FP16 kernel of the FP32 "
                         "operator on pairs of right-hand sides, 256: the same on 128-box tiles of one)")
    ap.add_argument("--p2m", type=int, default=None,
                    help="force the P2M variant (16, 32 or 64 boxes per tile; 128: the split-FP16 kernel of the FP32 "
                         "operator)")
    a = ap.parse_args()
    import torch
    import isg
    if a.copy_mib > 0:
        src = torch.ones(a.copy_mib << 20, dtype=torch.uint8, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        print("copy %d MiB done" % a.copy_mib, file=sys.stderr, flush=True)
    s = isg.Solver(level=a.level, numrx=16, max_rhs=a.nrhs, precision=a.precision)
    if a.near is not None:
        s.tuning_set(a.precision, 1 if a.nrhs >= 16 else 0, "near", 0, a.near)
    if a.p2m is not None:
        s.tuning_set(a.precision, 1 if a.nrhs >= 16 else 0, "p2m", 0, a.p2m)
    rng = np.random.default_rng(0)
    X = (rng.standard_normal((a.nrhs, s.N)) + 1j * rng.standard_normal((a.nrhs, s.N)))
    if a.vec32:
        x = torch.from_numpy(X.astype(np.complex64)).cuda()
        r = torch.empty_like(x)
        torch.cuda.synchronize()

        def run():
            s.mlfma_f32_dev(x.data_ptr(), r.data_ptr(), a.nrhs)
    else:
        x = torch.from_numpy(X).cuda()
        r = torch.empty_like(x)
        torch.cuda.synchronize()

        def run():
            s.mlfma_dev(x.data_ptr(), r.data_ptr(), a.nrhs)
    run()  # warm-up
    s.profile(True)
    for _ in range(a.reps):
        run()
    near = "near_fp32" if a.precision == 1 else "near"
    m2l = "m2l_fp32" if a.precision == 1 else "m2l"
    nms, nn = s.profile_read(near)
    mms, mn = s.profile_read(m2l)
    pms, pn = s.profile_read("p2m")
    ims, im = s.profile_read("interp")
    print(json.dumps({"level": a.level, "nrhs": a.nrhs, "precision": a.precision, "vec32": a.vec32, "reps": a.reps,
                      "near_avg_ms": nms / max(nn, 1), "near_launches": nn,
                      "m2l_avg_ms": mms / max(mn, 1), "m2l_launchsets": mn,
                      "p2m_avg_ms": pms / max(pn, 1), "interp_avg_ms": ims / max(im, 1),
                      "interp_launchsets": im}), flush=True)


if __name__ == "__main__":
    main()

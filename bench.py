#!/usr/bin/env python3
"""Benchmark: seconds per DBIM iteration on one MI355X.

Workload (the headline configuration): 1024x1024 pixels of lambda/10
(N = 1,048,576 unknowns, quadtree level 8, 8x8-pixel leaf boxes), 64
line-source transmitters and 256 receivers on a circle of radius 60
wavelengths (--radius 0: the reference's plane waves and far-field
receivers), contrast = 20 seeded random ellipses (seed 3, eps_r - 1 in
U(0.05, 0.8) + 0.02i), BiCGSTAB relative residual 1e-6 (iteration limit
300), Tikhonov regular = sqrt(1e-3); 1 warm-up and 5 timed DBIM iterations.

One step = one DBIM iteration (mom.cpp:277-400): per transmitter a Frechet
derivative solve, a total-field solve and an adjoint solve, plus the line
search, gradient and Polak-Ribiere update.  Inputs are resident in HBM before
the timed region; the timed region brackets K iterations with device
synchronisation on both sides (each iteration ends with a stream sync).

Three paths run from the same synthetic measurements:
  plain   FP64 operator, one right-hand side at a time, every solve from zero
          -- the reference's loop on this library's kernels, each BiCGSTAB
          iteration replayed as a HIP graph;
  fast    every transmitter of a phase in one multi-RHS BiCGSTAB with
          deflation, FP32 operator with FP64 solver scalars/vectors and an FP64
          true-residual defect correction (every solve still reaches 1e-6 in
          FP64), FP64 gradient and derivative products, total-field solves
          started from the previous DBIM total field moved to first order along
          the step (--fast-start);
  learned the fast path with the learned total-field start instead.
`value` is the fast path's s/iteration; the plain and learned figures are
reported beside it.

    python bench.py [--steps K] [--warmup W] [--gpus 1]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "inverse-solver-gpu_amd"))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

METRIC = "s per DBIM iteration, 1024² px ×64 Tx, rel-res 1e-6: plain vs batched+learned"
PEAK_TFLOPS = {"f64": 78.6, "f32": 157.3}
PEAK_HBM_GBS = 8000.0

# total-field start of the fast path (DbimOptions::field_start)
START_KINDS = {"zero": 0, "previous": 1, "first-order": 2, "learned": 3}
WARMNET_FILE = os.path.join(ROOT, "inverse-solver-gpu_amd", "weights", "warmnet_512.bin")

WORKLOADS = {
    # name: level, transmitters, receivers, ellipses, seed, default circle radius (wavelengths)
    "dbim-1024": dict(level=8, numtx=64, numrx=256, nell=20, seed=3, radius=60.0),
    "dbim-256": dict(level=6, numtx=16, numrx=64, nell=8, seed=3, radius=60.0),
    # 2048^2 (204.8 wavelengths a side): the circle passes outside the corners; transmitters in
    # two batches of 64 (the 128-RHS work vectors would take most of the 288 GB)
    "dbim-2048": dict(level=9, numtx=128, numrx=512, nell=20, seed=5, radius=160.0, batch=64),
}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def near_flops_per_rhs(level, S_leaf):
    """Algorithmic flops of one near-field + L2P application (one RHS):
    8*64*64 per in-domain 3x3 box pair (P2P) + 8*64*S per leaf box (L2P)."""
    nb = 1 << (level - 1)
    pairs = (3 * nb - 2) ** 2
    return 8 * 64 * 64 * pairs + 8 * 64 * S_leaf * nb * nb, pairs


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "hbm_traffic_fp32_operator_1024x64rhs.txt")


def measured_traffic():
    """HBM bytes (FETCH_SIZE + WRITE_SIZE, copy-calibrated) of one 64-RHS
    launch of the FP32 near-field kernel at level 8, averaged over the
    launches of a timed fast-path iteration (most carry the fused BiCGSTAB
    post step), from the committed rocprofv3 counter summary; None when that
    summary is absent."""
    try:
        with open(TRAFFIC_FILE) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch_64rhs="):
                    return int(line.split("=")[1])
    except OSError:
        pass
    return None


def m2l_bytes_per_rhs(level, numsamp, cbytes, fused_leaf=False):
    """Read every multipole once and write every local once, levels 2..leaf;
    with the fused leaf kernel also read the parent locals it interpolates."""
    b = sum(2 * cbytes * 4 ** i * numsamp[i] for i in range(2, level))
    if fused_leaf:
        b += cbytes * 4 ** (level - 2) * numsamp[level - 2]
    return b


def run_path(solver, dbim, warmup, steps, label, timing=True):
    for w in range(warmup):
        st = dbim.iterate()
        log("[%s] warmup iteration %d: %.3fs, %d solves, %d BiCGSTAB iterations, misfit %.4e"
            % (label, w, st["seconds"], st["solves"], st["solve_iters"], st["misfit"]))
    solver.profile(timing)
    c0 = solver.counters()
    t0 = time.perf_counter()
    stats = []
    for k in range(steps):
        st = dbim.iterate()
        stats.append(st)
        log("[%s] timed iteration %d: %.3fs, %d solves, %d BiCGSTAB iterations, %d not converged, misfit %.4e"
            % (label, k, st["seconds"], st["solves"], st["solve_iters"], st["not_converged"], st["misfit"]))
    t1 = time.perf_counter()
    c1 = solver.counters()
    stats[-1]["refinements_timed"] = c1["refinements"] - c0["refinements"]
    prof = {k: solver.profile_read(k) for k in solver.KERNELS}
    solver.profile(False)
    return (t1 - t0) / steps, stats, prof


def cpu_baseline(level, numrx, matvecs_per_iter, nsample):
    """Time the CPU oracle (FP64 C++/OpenMP restatement of the reference CPU
    path) on `nsample` matvecs of the same size and scale by the matvec count
    of one plain DBIM iteration measured on the GPU."""
    from oracle.oracle import Oracle
    o = Oracle(level, numrx=numrx)
    rng = np.random.default_rng(0)
    x = rng.standard_normal(o.N) + 1j * rng.standard_normal(o.N)
    o.mlfma(x)  # first touch
    t0 = time.perf_counter()
    for _ in range(nsample):
        o.mlfma(x)
    t_mv = (time.perf_counter() - t0) / nsample
    threads = o.L.orc_threads()
    return {
        "value": round(t_mv * matvecs_per_iter, 2),
        "unit": "s",
        "cores": threads,
        "kind": "port",
        "sample": ("%d FP64 MLFMA matvecs at %dx%d on the CPU oracle (%.3f s each, %d OpenMP threads), "
                   "scaled by the %d matvecs of one plain DBIM iteration counted on the GPU run; "
                   "BiCGSTAB vector algebra and receiver operators not included (extrapolated)"
                   % (nsample, 8 << (level - 1), 8 << (level - 1), t_mv, threads, matvecs_per_iter)),
    }


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="dbim-1024", choices=sorted(WORKLOADS))
    ap.add_argument("--paths", default="fast,plain,learned",
                    help="fast: the headline path; plain: the reference's FP64 single-RHS path; learned: the "
                         "fast path with the learned total-field start instead of the first-order one")
    ap.add_argument("--fast-precision", default="mixed", choices=["mixed", "fp64"])
    ap.add_argument("--fast-start", default="first-order", choices=sorted(START_KINDS),
                    help="total-field start of the fast path")
    ap.add_argument("--fast-products", default="fp64", choices=["fp64", "fp32"],
                    help="operator of the DBIM's derivative-source product on the fast path (fp32 needs "
                         "--fast-precision mixed); the gradient's product is FP64 either way")
    ap.add_argument("--fast-solve-start", default="zero", choices=["zero", "previous"],
                    help="adjoint and derivative-field solves of the fast path: zero (reference) or the "
                         "previous DBIM iteration's solution scaled by the residual-minimising lambda "
                         "(pays only where the misfit stalls)")
    ap.add_argument("--mixed-margin", type=float, default=0.9,
                    help="FP32 solve target as a fraction of the tolerance (mixed precision); 0.9 needed no "
                         "defect correction at 1024^2 (both geometries) and 2048^2, the library default is 0.5")
    ap.add_argument("--tuning", default="committed",
                    help="kernel-variant cache: 'committed' (inverse-solver-gpu_amd/tuning), 'none' or a path")
    ap.add_argument("--kernel-timing", default="on", choices=["on", "off"],
                    help="live HIP-event kernel timing in the fast and learned paths' timed regions (off: "
                         "BiCGSTAB iterations replay as HIP graphs, no roofline figures); the plain path "
                         "replays graphs unless it is run without the fast path")
    ap.add_argument("--radius", type=float, default=None,
                    help="transmitter/receiver circle radius in wavelengths (0: plane waves, far-field "
                         "receivers; default: the workload's, 60 at 1024^2)")
    ap.add_argument("--cpu-sample", type=int, default=2, help="oracle matvecs timed for cpu_baseline (0: skip)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: there is no multi-GPU path (sharding the transmitters "
              "over RCCL is designed but not built); run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    paths = [p for p in args.paths.split(",") if p]
    if not set(paths) <= {"fast", "plain", "learned"} or not ({"fast", "plain"} & set(paths)):
        print("--paths: a comma list of fast, plain, learned with fast or plain in it", file=sys.stderr)
        sys.exit(2)

    import isg
    wl = WORKLOADS[args.workload]
    if args.radius is None:
        args.radius = wl["radius"]
    level, ntx, nrx = wl["level"], wl["numtx"], wl["numrx"]
    side = 8 << (level - 1)
    tuning = {"committed": True, "none": False}.get(args.tuning, args.tuning)
    common = dict(level=level, numrx=nrx, numtx=ntx, toli=1e-6, iti=300, regular=float(np.sqrt(1e-3)),
                  tuning=tuning, radius=args.radius)
    fast_prec = 1 if args.fast_precision == "mixed" else 0

    # synthetic measurements of the true contrast, made once by the batched FP64 path
    t0 = time.perf_counter()
    batch = wl.get("batch", ntx)
    gen = isg.Solver(max_rhs=batch, **common)
    pos, r2t = gen.geometry()
    otrue = isg.raster_to_tree(isg.phantom_ellipses(side, wl["nell"], wl["seed"]), r2t)
    d_gen = isg.Dbim(gen, otrue, field_start=1)
    mes = d_gen.get("mes")
    log("measurements (FP64, batched): %.2fs, %d solves, %d BiCGSTAB iterations, %d not converged"
        % (time.perf_counter() - t0, d_gen.setup_stats["solves"], d_gen.setup_stats["solve_iters"],
           d_gen.setup_stats["not_converged"]))
    S_leaf = gen.numsamp[-1]
    numsamp = list(gen.numsamp)
    del d_gen
    gen.close()
    flops_mv, pairs = near_flops_per_rhs(level, S_leaf)

    res = {}
    if "fast" in paths:
        fast = isg.Solver(max_rhs=batch, precision=fast_prec, mixed_margin=args.mixed_margin,
                          products=1 if args.fast_products == "fp32" and fast_prec == 1 else 0,
                          solve_start=1 if args.fast_solve_start == "previous" else 0, **common)
        d_fast = isg.Dbim(fast, otrue, field_start=START_KINDS[args.fast_start], measurements=mes)
        if args.fast_start == "learned":
            d_fast.set_warmnet(isg.WarmNet(max_batch=8, max_pixels=side * side, path=WARMNET_FILE))
        res["fast"] = run_path(fast, d_fast, args.warmup, args.steps, "fast", args.kernel_timing == "on")
        contrast = {"fast": d_fast.get("o")}
        fast_leaf_fused = fast.tuning_get(fast_prec, 1, "leaf") == 1
        del d_fast
        fast.close()
    learned_error = None
    if "learned" in paths:
        # a side figure: a failure here (e.g. missing network weights) is reported in the
        # line instead of costing the headline measurement
        lrn = None
        try:
            lrn = isg.Solver(max_rhs=batch, precision=fast_prec, mixed_margin=args.mixed_margin,
                             products=1 if args.fast_products == "fp32" and fast_prec == 1 else 0,
                             solve_start=1 if args.fast_solve_start == "previous" else 0, **common)
            d_lrn = isg.Dbim(lrn, otrue, field_start=START_KINDS["learned"], measurements=mes)
            d_lrn.set_warmnet(isg.WarmNet(max_batch=8, max_pixels=side * side, path=WARMNET_FILE))
            res["learned"] = run_path(lrn, d_lrn, args.warmup, args.steps, "learned", args.kernel_timing == "on")
            del d_lrn
        except (isg.IsgError, OSError) as e:
            learned_error = str(e)
            log("[learned] skipped: %s" % e)
        if lrn is not None:
            lrn.close()
    if "plain" in paths:
        plain = isg.Solver(max_rhs=1, **common)
        d_plain = isg.Dbim(plain, otrue, field_start=0, measurements=mes)
        # the plain path is timed at its best: without live kernel timing, so that each BiCGSTAB
        # iteration replays as a HIP graph (its single-RHS kernels are short enough for launch
        # overhead to show: measured 1.002 vs 1.022 s per DBIM iteration); it keeps the timers
        # only when it is the line's main path and supplies the roofline figures
        plain_timing = args.kernel_timing == "on" and "fast" not in paths
        res["plain"] = run_path(plain, d_plain, args.warmup, args.steps, "plain", plain_timing)
        contrast_plain = d_plain.get("o")
        del d_plain
        plain.close()

    main_path = "fast" if "fast" in res else "plain"
    fast_leaf_fused = "fast" in res and fast_leaf_fused
    s_iter, stats, prof = res[main_path]
    mixed = main_path == "fast" and fast_prec == 1
    matvecs = sum(s["matvecs"] for s in stats)
    mv32 = sum(s["matvecs_fp32"] for s in stats)
    if mixed:
        near_ms, near_n = prof["near_fp32"]
        m2l_ms, m2l_n = prof["m2l_fp32"]
        n_near, prec, cb = mv32, "f32", 8
    else:
        near_ms, near_n = prof["near"]
        m2l_ms, m2l_n = prof["m2l"]
        n_near, prec, cb = matvecs, "f64", 16
    vec_ms, vec_n = prof["bicgs_update"]
    achieved = n_near * flops_mv / (near_ms * 1e-3) / 1e12 if near_ms > 0 else 0.0
    fused = main_path == "fast" and fast_leaf_fused
    m2l_gbs = (n_near * m2l_bytes_per_rhs(level, numsamp, cb, fused) / (m2l_ms * 1e-3) / 1e9
               if m2l_ms > 0 else 0.0)
    iters = sum(s["solve_iters"] for s in stats)
    traffic = measured_traffic() if (mixed and level == 8 and ntx == 64) else None

    out = {
        "metric": METRIC if args.workload == "dbim-1024" else METRIC.replace("1024² px ×64 Tx", "%d² px ×%d Tx" % (side, ntx)),
        "value": round(s_iter, 4),
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * s_iter, 2),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 operator + f64 solver" if mixed else "f64",
        "data": "synthetic (seeded ellipse phantom, measurements synthesized by the FP64 forward solver)",
        "config": {
            "workload": "DBIM %dx%d px lambda/10, level %d, %d Tx, %d Rx %s, %d ellipses seed %d, rel-res 1e-6"
                        % (side, side, level, ntx, nrx,
                           ("on a circle of radius %g lambda (line sources)" % args.radius) if args.radius > 0
                           else "(plane waves, far-field receivers)", wl["nell"], wl["seed"]),
            "path": ("batched %d RHS + deflation, %s, %s total-field start, %s adjoint/derivative starts, "
                     "%s derivative-source product"
                     % (batch, ("FP32 operator / FP64 solver (FP32 solves to %g x tol) + FP64 defect correction"
                                % args.mixed_margin) if mixed else "FP64",
                        args.fast_start, args.fast_solve_start,
                        "FP32" if args.fast_products == "fp32" and mixed else "FP64")
                     if main_path == "fast" else "1 RHS, FP64, zero start"),
            "kernel_tuning": args.tuning,
        },
        "roofline": {
            "kernel": "k_near_mfma (P2P + L2P on v_mfma_%s_16x16x4)" % prec,
            "bound": "mfma",
            "achieved": round(achieved, 2),
            "peak": PEAK_TFLOPS[prec],
            "unit": "TFLOP/s",
            "frac": round(achieved / PEAK_TFLOPS[prec], 4),
            "traffic": traffic,
            "traffic_launch": "64 RHS, level 8, mean over the near-field launches of a fast-path DBIM iteration "
                              "(rocprofv3 FETCH_SIZE x2 + WRITE_SIZE x1, copy-calibrated)"
                              if traffic else None,
            "flops_per_rhs_launch": flops_mv,
            "note": "complex products in the three-product form: the matrix cores execute 6 of the 8 "
                    "algorithmic flops per complex multiply-add, so the executed MFMA rate is 0.75 x achieved",
            "avg_launch_ms": round(near_ms / max(near_n, 1), 4),
            "launches": near_n,
        },
        "m2l": {"achieved_GBs": round(m2l_gbs, 1), "peak_GBs": PEAK_HBM_GBS, "frac": round(m2l_gbs / PEAK_HBM_GBS, 4),
                "ms_per_matvec_launchset": round(m2l_ms / max(m2l_n, 1), 4),
                "leaf_l2l_fused": fused},
        # x/r update reads x, p, s, t, r~ and writes x, r: x is FP64 (16 B), the
        # work vectors FP64 (plain) or FP32 (mixed): 7 x 16 B or 2 x 16 + 5 x 8 B per unknown
        "bicgs_update": {"GBs": round((72 if mixed else 112) * side * side * iters / max(vec_ms * 1e-3, 1e-12) / 1e9, 1),
                         "avg_launch_ms": round(vec_ms / max(vec_n, 1), 4)},
        "solves": sum(s["solves"] for s in stats),
        "bicgstab_iterations": iters,
        "field_bicgstab_iterations": sum(s["field_iters"] for s in stats),
        "matvecs": matvecs,
        "matvecs_fp32": mv32,
        "defect_corrections": stats[-1].get("refinements_timed"),
        "misfit": stats[-1]["misfit"],
    }
    if "plain" in res and "fast" in res:
        out["plain_s_per_iter"] = round(res["plain"][0], 4)
        out["speedup_vs_plain"] = round(res["plain"][0] / s_iter, 3)
        out["plain_misfit"] = res["plain"][1][-1]["misfit"]
        # reconstruction after the same iterations: fast path against the FP64 plain path
        out["contrast_rel_diff_vs_plain"] = float(np.linalg.norm(contrast["fast"] - contrast_plain)
                                                  / np.linalg.norm(contrast_plain))
    if learned_error is not None:
        out["learned_start"] = {"error": learned_error}
    if "learned" in res:
        lst = res["learned"][1]
        out["learned_start"] = {
            "s_per_iter": round(res["learned"][0], 4),
            "speedup_vs_plain": round(res["plain"][0] / res["learned"][0], 3) if "plain" in res else None,
            "field_bicgstab_iterations": sum(s["field_iters"] for s in lst),
            "network_s_per_iter": round(sum(s["net_seconds"] for s in lst) / len(lst), 4),
            "mean_abs_lambda": round(sum(s["lambda_abs"] for s in lst) / len(lst), 4),
            "note": "the fast path with the learned total-field start (network + residual-minimising blend) "
                    "in place of the first-order start; field_bicgstab_iterations compares with the line's",
        }
    if args.cpu_sample > 0 and "plain" in res:
        plain_mv = int(round(sum(s["matvecs"] for s in res["plain"][1]) / len(res["plain"][1])))
        try:
            out["cpu_baseline"] = cpu_baseline(level, nrx, plain_mv, args.cpu_sample)
        except Exception as e:  # the CPU oracle needs its .so; report instead of failing the bench
            out["cpu_baseline"] = {"error": str(e)}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

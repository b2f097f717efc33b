#!/usr/bin/env python3
"""Benchmark: seconds per DBIM iteration on one MI355X.

Workload (the headline configuration): 1024x1024 pixels of lambda/10
(N = 1,048,576 unknowns, quadtree level 8, 8x8-pixel leaf boxes), 64
line-source transmitters and 256 receivers on a circle of radius 60
wavelengths, contrast = 20 seeded random ellipses (seed 3, eps_r - 1 in
U(0.05, 0.8) + 0.02i), BiCGSTAB relative residual 1e-6 (iteration limit 300),
Tikhonov regular = sqrt(1e-3), 5 DBIM iterations.

One step = one DBIM iteration (mom.cpp:277-400): per transmitter a Frechet
derivative solve, a total-field solve and an adjoint solve, plus the line
search, gradient and Polak-Ribiere update.  The steps cycle through the
inversion's first 5 iterations from a cold start: after every 5th step the
inversion returns to its initial state (zero object, isg_dbim_restart: device
copies only, inside the timed region), so K steps with K a multiple of 5
average whole 5-iteration inversions.  Inputs are resident in HBM before the
timed region; every step ends with a stream synchronisation.

The synthetic measurements are made once, before any timing, by the batched
mixed-precision solver with an FP64 true-residual certification and an
iteration limit raised for this synthesis only (--synth-iti), so that every
forward solve of the true contrast reaches the 1e-6 tolerance (counts in the
line's "synthesis").

Paths, all from the same measurements:
  learned (value)  every transmitter of a phase in one multi-RHS BiCGSTAB with
          deflation, FP32 operator with FP64 solver scalars/vectors and an FP64
          true-residual defect correction (every solve still reaches 1e-6 in
          FP64), FP64 gradient and derivative products, total-field solves
          started from the warm-start network's prediction blended onto the
          first-order start (isg_dbim field_start 3);
  fast    the same with the first-order total-field start alone (no network);
  plain   FP64 operator, one right-hand side at a time, every solve from zero:
          the reference's loop on this library's kernels, each BiCGSTAB
          iteration replayed as a HIP graph, the dense receiver operator
          applied once per phase for all transmitters.
The reference geometry (plane waves, far-field receivers; --radius 0) is timed
as well, one 5-iteration cycle per path ("reference_geometry").

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
PEAK_TFLOPS = {"f64": 78.6, "f32": 157.3, "f16": 2500.0}
PEAK_HBM_GBS = 8000.0
CYCLE = 5  # DBIM iterations of one inversion (the workload's 5; --cycle sets it)

# total-field start (DbimOptions::field_start)
START_KINDS = {"zero": 0, "previous": 1, "first-order": 2, "learned": 3}
WARMNET_FILE = os.path.join(ROOT, "inverse-solver-gpu_amd", "weights", "warmnet.bin")

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


def measured_traffic(split):
    """HBM bytes (FETCH_SIZE + WRITE_SIZE, copy-calibrated) of one 64-RHS
    launch of the FP32 operator's near-field kernel at level 8 (split: the
    split-FP16 kernel's tuning value, 128 or 256), averaged over the launches of a timed fast-path
    iteration (most carry the fused BiCGSTAB post step), from the committed
    rocprofv3 counter summary; None when that summary is absent."""
    key = "traffic_bytes_per_launch_64rhs%s=" % {128: "_split", 256: "_splitw"}.get(split, "")
    try:
        with open(TRAFFIC_FILE) as f:
            for line in f:
                if line.startswith(key):
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
    """W untimed then K timed steps; step p is iteration p mod CYCLE + 1 of
    the inversion, the state going back to the start before every
    CYCLE-th step (inside the timed region when it falls there)."""
    pos = [0]
    cyc = CYCLE

    def step():
        if pos[0] > 0 and pos[0] % cyc == 0:
            dbim.restart()
        st = dbim.iterate()
        st["iteration"] = pos[0] % cyc + 1
        pos[0] += 1
        return st

    for w in range(warmup):
        st = step()
        log("[%s] warmup step %d (iteration %d): %.3fs, %d solves, %d BiCGSTAB iterations, misfit %.4e"
            % (label, w, st["iteration"], st["seconds"], st["solves"], st["solve_iters"], st["misfit"]))
    solver.profile(timing)
    c0 = solver.counters()
    t0 = time.perf_counter()
    stats = []
    for k in range(steps):
        st = step()
        stats.append(st)
        log("[%s] timed step %d (iteration %d): %.3fs, %d solves, %d BiCGSTAB iterations (%d total-field), "
            "%d not converged, alpha %.3e, misfit %.4e"
            % (label, k, st["iteration"], st["seconds"], st["solves"], st["solve_iters"], st["field_iters"],
               st["not_converged"], st["alpha"], st["misfit"]))
    t1 = time.perf_counter()
    c1 = solver.counters()
    stats[-1]["refinements_timed"] = c1["refinements"] - c0["refinements"]
    prof = {k: solver.profile_read(k) for k in solver.KERNELS}
    solver.profile(False)
    return (t1 - t0) / steps, stats, prof


def misfit_by_iteration(stats):
    """Misfit after each of the cycle's iterations (the last timed visit)."""
    out = {}
    for s in stats:
        out[s["iteration"]] = s["misfit"]
    return [out[k] for k in sorted(out)]


def synthesize(isg, common, batch, otrue, iti):
    """Measurements of the true contrast: batched mixed-precision forward
    solves certified in FP64 to the tolerance, iteration limit `iti`."""
    t0 = time.perf_counter()
    kw = dict(common)
    kw["iti"] = iti
    gen = isg.Solver(max_rhs=batch, precision=1, mixed_margin=0.5, **kw)
    gen.set_progress(lambda done, total: log("measurements: %d of %d transmitters solved (%.1fs)"
                                             % (done, total, time.perf_counter() - t0)))
    d = isg.Dbim(gen, otrue, field_start=1)
    mes = d.get("mes")
    s = d.setup_stats
    info = {"solver": "batched, FP32 operator + FP64 true-residual defect correction to rel-res %g"
                      % common["toli"],
            "iteration_limit": iti, "solves": s["solves"], "bicgstab_iterations": s["solve_iters"],
            "not_converged": s["not_converged"], "seconds": round(time.perf_counter() - t0, 2)}
    log("measurements: %.2fs, %d solves, %d BiCGSTAB iterations, %d not converged (limit %d)"
        % (info["seconds"], s["solves"], s["solve_iters"], s["not_converged"], iti))
    numsamp = list(gen.numsamp)
    del d
    gen.close()
    return mes, info, numsamp


def run_paths(isg, args, wl, common, otrue, mes, paths, warmup, steps, timing):
    """Time the requested paths from the same measurements; returns a dict:
    res (per path: s/step, stats, kernel timers), contrast (per path), fused
    (leaf M2L + L2L fused), split (the FP32 near field ran on the split-FP16
    kernel) and errors (paths skipped)."""
    batch = wl.get("batch", wl["numtx"])
    fast_prec = 1 if args.fast_precision == "mixed" else 0
    fast_kw = dict(max_rhs=batch, precision=fast_prec, mixed_margin=args.mixed_margin,
                   products=1 if args.fast_products == "fp32" and fast_prec == 1 else 0,
                   solve_start=1 if args.fast_solve_start == "previous" else 0)
    res, contrast, errors = {}, {}, {}
    fused = split = False
    side = 8 << (wl["level"] - 1)
    for p in paths:
        if p == "plain":
            s = isg.Solver(max_rhs=1, **common)
            d = isg.Dbim(s, otrue, field_start=0, measurements=mes)
            # timed at its best: without live kernel timers, so that each BiCGSTAB iteration
            # replays as a HIP graph (its single-RHS kernels are short enough for launch
            # overhead to show), unless it is the only path and supplies the roofline figures
            ptime = timing and paths == ["plain"]
        else:
            s = isg.Solver(**fast_kw, **common)
            start = "learned" if p == "learned" else args.fast_start
            try:
                d = isg.Dbim(s, otrue, field_start=START_KINDS[start], measurements=mes)
                if start == "learned":
                    d.set_warmnet(isg.WarmNet(max_batch=args.net_batch, max_pixels=side * side,
                                              path=WARMNET_FILE))
            except (isg.IsgError, OSError) as e:
                errors[p] = str(e)
                log("[%s] skipped: %s" % (p, e))
                s.close()
                continue
            ptime = timing
            fused = fused or s.tuning_get(fast_prec, 1, "leaf") == 1
            # the split-FP16 near field: k_near_h16 (tuning value 128) serves launches of >= 4
            # right-hand sides on a grid of >= 8 x 8 leaf boxes, k_near_h16w (256) grids of >= 16 x 8
            nv = s.tuning_get(1, 1 if batch >= 16 else 0, "near")
            if fast_prec == 1 and ((nv == 128 and batch >= 4 and wl["level"] >= 4) or (nv == 256 and wl["level"] >= 5)):
                split = split or nv
        res[p] = run_path(s, d, warmup, steps, p, ptime)
        contrast[p] = d.get("o")
        del d
        s.close()
    return {"res": res, "contrast": contrast, "fused": fused, "split": split, "errors": errors}


def cpu_baselines(isg, level, radius, o_rec, inc0, plain_iter_solves):
    """The CPU oracle (FP64 C++/OpenMP restatement of the reference CPU path:
    solve.cpp MLFMA, bicgs.cpp/bicgs_gpu.cu BiCGSTAB, mom.cpp DBIM) on the
    host's cores: the 64^2 DBIM run in full, per-matvec and per-solve times at
    256^2, and one 1024^2 solve of the benchmark's kind scaled to a DBIM
    iteration."""
    from oracle.oracle import Oracle, Dbim
    out = {"nproc": os.cpu_count()}
    # 64^2 DBIM: 16 Tx / 16 Rx on a 10-lambda circle, 3 blobs (seed 1), 3 iterations
    o4 = Oracle(4, numrx=16, radius=10.0)
    out["threads"] = o4.L.orc_threads()
    ot4 = isg.raster_to_tree(isg.phantom_blobs(o4.side, 3, 1), o4.unkmap())
    t0 = time.perf_counter()
    d4 = Dbim(o4, 16, ot4, regular=float(np.sqrt(1e-3)), tol=1e-6, maxit=300)
    t_setup = time.perf_counter() - t0
    t0 = time.perf_counter()
    for _ in range(3):
        d4.iterate()
    t_it = (time.perf_counter() - t0) / 3
    out["dbim_64"] = {"s_per_iter": round(t_it, 4), "setup_s": round(t_setup, 4),
                      "misfit_after_3": float(d4.hist()[0]), "kind": "measured, full run"}
    # 256^2, 1 Tx: matvec and one forward solve of the 8-blob phantom (seed 2)
    o6 = Oracle(6, numrx=16)
    rng = np.random.default_rng(0)
    x = rng.standard_normal(o6.N) + 1j * rng.standard_normal(o6.N)
    o6.mlfma(x)
    t0 = time.perf_counter()
    for _ in range(5):
        o6.mlfma(x)
    t_mv6 = (time.perf_counter() - t0) / 5
    ot6 = isg.raster_to_tree(isg.phantom_blobs(o6.side, 8, 2), o6.unkmap())
    inc6 = np.exp(1j * o6.k0 * o6.pos().real)
    t0 = time.perf_counter()
    _, info6 = o6.bicgs(ot6 * o6.k0 ** 2, inc6, tol=1e-6, maxit=300)
    t_solve6 = time.perf_counter() - t0
    out["single_rhs_256"] = {"s_per_matvec": round(t_mv6, 5), "s_per_solve": round(t_solve6, 4),
                             "solve_iterations": info6["iters"], "kind": "measured"}
    # 1024^2: one forward solve with the benchmark's reconstructed contrast and transmitter 0's
    # incident field, to 1e-6; a DBIM iteration has 3 solves per transmitter
    del o4, o6, d4
    o8 = Oracle(level, numrx=16, radius=radius)
    o8.mlfma(np.zeros(o8.N, np.complex128))  # first touch
    y = rng.standard_normal(o8.N) + 1j * rng.standard_normal(o8.N)
    t0 = time.perf_counter()
    o8.mlfma(y)
    t_mv8 = time.perf_counter() - t0
    t0 = time.perf_counter()
    _, info8 = o8.bicgs(o_rec * o8.k0 ** 2, inc0, tol=1e-6, maxit=300)
    t_solve8 = time.perf_counter() - t0
    out["solve_1024"] = {"s_per_matvec": round(t_mv8, 4), "s_per_solve": round(t_solve8, 3),
                         "solve_iterations": info8["iters"], "kind": "measured"}
    return {
        "value": round(t_solve8 * plain_iter_solves, 2),
        "unit": "s",
        "cores": out["threads"],
        "kind": "port",
        "sample": ("one FP64 BiCGSTAB solve at %dx%d on the CPU oracle (%.2f s, %d iterations, %d OpenMP "
                   "threads on %d host CPUs) with the reconstructed contrast and transmitter 0's incident field, "
                   "x %d solves of one DBIM iteration (extrapolated: receivers and DBIM vector algebra "
                   "not included); 64^2 DBIM and 256^2 matvec/solve timings measured in full below"
                   % (o8.side, o8.side, t_solve8, info8["iters"], out["threads"], out["nproc"],
                      plain_iter_solves)),
        "detail": out,
    }


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--net-batch", type=int, default=16,
                    help="transmitters per warm-start network launch (its workspace grows with it)")
    ap.add_argument("--cycle", type=int, default=5,
                    help="DBIM iterations of one inversion: step p is iteration p mod cycle + 1, the state going "
                         "back to the start before every cycle-th step (10 with --steps 10 --warmup 0: one "
                         "10-iteration inversion per path, the full-scale configuration's)")
    ap.add_argument("--workload", default="dbim-1024", choices=sorted(WORKLOADS))
    ap.add_argument("--paths", default="learned,fast,plain",
                    help="learned: the headline path; fast: the same with the first-order total-field start "
                         "alone; plain: the reference's FP64 single-RHS path")
    ap.add_argument("--fast-precision", default="mixed", choices=["mixed", "fp64"])
    ap.add_argument("--fast-start", default="first-order", choices=sorted(START_KINDS),
                    help="total-field start of the fast path")
    ap.add_argument("--fast-products", default="fp64", choices=["fp64", "fp32"],
                    help="operator of the DBIM's derivative-source product on the batched paths (fp32 needs "
                         "--fast-precision mixed); the gradient's product is FP64 either way")
    ap.add_argument("--fast-solve-start", default="zero", choices=["zero", "previous"],
                    help="adjoint and derivative-field solves of the batched paths: zero (reference) or the "
                         "previous DBIM iteration's solution scaled by the residual-minimising lambda")
    ap.add_argument("--mixed-margin", type=float, default=0.9,
                    help="FP32 solve target as a fraction of the tolerance (mixed precision); the library "
                         "default is 0.5")
    ap.add_argument("--synth-iti", type=int, default=3000,
                    help="BiCGSTAB iteration limit of the measurement synthesis only (the timed solves keep 300)")
    ap.add_argument("--tuning", default="committed",
                    help="kernel-variant cache: 'committed' (inverse-solver-gpu_amd/tuning), 'none' or a path")
    ap.add_argument("--kernel-timing", default="on", choices=["on", "off"],
                    help="live HIP-event kernel timing in the batched paths' timed regions (off: BiCGSTAB "
                         "iterations replay as HIP graphs, no roofline figures)")
    ap.add_argument("--radius", type=float, default=None,
                    help="transmitter/receiver circle radius in wavelengths (0: plane waves, far-field "
                         "receivers; default: the workload's, 60 at 1024^2)")
    ap.add_argument("--ref-geometry", default="on", choices=["on", "off"],
                    help="also time the learned and plain paths in the reference geometry (plane waves, "
                         "far-field receivers), one 5-iteration cycle each")
    ap.add_argument("--cpu-baseline", default="on", choices=["on", "off"])
    ap.add_argument("--save-measurements", default=None,
                    help="write the synthesized measurements (numtx x numrx complex, .npy) to this file")
    ap.add_argument("--measurements", default=None,
                    help="read the measurements from a file written by --save-measurements instead of "
                         "synthesizing them (profiling runs)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: there is no multi-GPU path (sharding the transmitters "
              "over RCCL is designed but not built); run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    if args.cycle < 1:
        print("--cycle: at least 1", file=sys.stderr)
        sys.exit(2)
    global CYCLE
    CYCLE = args.cycle
    paths = [p for p in args.paths.split(",") if p]
    if not paths or not set(paths) <= {"fast", "plain", "learned"}:
        print("--paths: a comma list of learned, fast, plain", file=sys.stderr)
        sys.exit(2)

    import isg
    wl = WORKLOADS[args.workload]
    if args.radius is None:
        args.radius = wl["radius"]
    level, ntx, nrx = wl["level"], wl["numtx"], wl["numrx"]
    side = 8 << (level - 1)
    batch = wl.get("batch", ntx)
    tuning = {"committed": True, "none": False}.get(args.tuning, args.tuning)
    common = dict(level=level, numrx=nrx, numtx=ntx, toli=1e-6, iti=300, regular=float(np.sqrt(1e-3)),
                  tuning=tuning, radius=args.radius)
    fast_prec = 1 if args.fast_precision == "mixed" else 0
    timing = args.kernel_timing == "on"

    probe = isg.Solver(max_rhs=1, level=level, numrx=nrx, numtx=ntx, tuning=False, radius=args.radius)
    _, r2t = probe.geometry()
    probe.close()
    otrue = isg.raster_to_tree(isg.phantom_ellipses(side, wl["nell"], wl["seed"]), r2t)
    if args.measurements:
        mes = np.load(args.measurements)
        synth = {"from_file": args.measurements}
        probe = isg.Solver(max_rhs=1, **common)
        numsamp = list(probe.numsamp)
        probe.close()
    else:
        mes, synth, numsamp = synthesize(isg, common, batch, otrue, args.synth_iti)
    if args.save_measurements:
        np.save(args.save_measurements, mes)
    S_leaf = numsamp[-1]
    flops_mv, pairs = near_flops_per_rhs(level, S_leaf)

    runs = run_paths(isg, args, wl, common, otrue, mes, paths, args.warmup, args.steps, timing)
    res, contrast, errors = runs["res"], runs["contrast"], runs["errors"]
    fast_leaf_fused, fast_split = runs["fused"], runs["split"]
    if not res:
        print(json.dumps({"metric": METRIC, "value": None, "error": errors}), flush=True)
        sys.exit(1)

    main_path = next(p for p in ("learned", "fast", "plain") if p in res)
    s_iter, stats, prof = res[main_path]
    mixed = main_path != "plain" and fast_prec == 1
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
    fused = main_path != "plain" and fast_leaf_fused
    m2l_gbs = (n_near * m2l_bytes_per_rhs(level, numsamp, cb, fused) / (m2l_ms * 1e-3) / 1e9
               if m2l_ms > 0 else 0.0)
    iters = sum(s["solve_iters"] for s in stats)
    split = fast_split if mixed else False
    traffic = measured_traffic(split) if (mixed and level == 8 and ntx == 64) else None
    peak = PEAK_TFLOPS["f16" if split else prec]
    geom = (("on a circle of radius %g lambda (line sources)" % args.radius) if args.radius > 0
            else "(plane waves, far-field receivers)")

    def path_desc(p):
        if p == "plain":
            return "1 RHS, FP64, zero start"
        return ("batched %d RHS + deflation, %s, %s total-field start, %s adjoint/derivative starts, "
                "%s derivative-source product"
                % (batch, ("FP32 operator / FP64 solver (FP32 solves to %g x tol) + FP64 defect correction"
                           % args.mixed_margin) if fast_prec == 1 else "FP64",
                   "learned (network + residual-minimising blend onto the first-order start)" if p == "learned"
                   else args.fast_start, args.fast_solve_start,
                   "FP32" if args.fast_products == "fp32" and fast_prec == 1 else "FP64"))

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
        "data": "synthetic (seeded ellipse phantom, measurements synthesized by the solver, every solve to 1e-6)",
        "config": {
            "workload": "DBIM %dx%d px lambda/10, level %d, %d Tx, %d Rx %s, %d ellipses seed %d, rel-res 1e-6, "
                        "steps cycle through iterations 1-%d from a cold start"
                        % (side, side, level, ntx, nrx, geom, wl["nell"], wl["seed"], CYCLE),
            "path": main_path + ": " + path_desc(main_path),
            "kernel_tuning": args.tuning,
        },
        "roofline": {
            "kernel": ("%s (P2P + L2P on v_mfma_f32_16x16x32_f16, FP32 operands split into FP16 hi + lo)"
                       % ("k_near_h16w" if split == 256 else "k_near_h16")
                       if split else "k_near_mfma (P2P + L2P on v_mfma_%s_16x16x4)" % prec),
            "bound": "mfma",
            "achieved": round(achieved, 2),
            "peak": peak,
            "unit": "TFLOP/s",
            "frac": round(achieved / peak, 4),
            "traffic": traffic,
            "traffic_launch": "64 RHS, level 8, mean over the near-field launches of a fast-path DBIM iteration "
                              "(rocprofv3 FETCH_SIZE x2 + WRITE_SIZE x1, copy-calibrated)"
                              if traffic else None,
            "flops_per_rhs_launch": flops_mv,
            "note": ("achieved counts 8 flops per complex multiply-add; the split kernel executes each real product "
                     "as three FP16 products (hi.hi + hi.lo + lo.hi) in the four-product complex form, 24 FP16 "
                     "flops per complex multiply-add, so its executed MFMA rate is 3 x achieved and its ceiling "
                     "for this arithmetic is peak / 3" if split else
                     "complex products in the three-product form: the matrix cores execute 6 of the 8 "
                     "algorithmic flops per complex multiply-add, so the executed MFMA rate is 0.75 x achieved"),
            "frac_of_split_ceiling": round(3 * achieved / peak, 4) if split else None,
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
        "misfit_by_iteration": misfit_by_iteration(stats),
        # the DBIM iteration whose contrast the contrast_rel_diff fields compare (the last step's)
        "contrast_after_iteration": (args.warmup + args.steps - 1) % CYCLE + 1,
        "synthesis": synth,
    }
    if timing:
        # GPU time of the live-timed kernel groups per step (HIP events on the solver's stream;
        # near/m2l are the FP64 operator, *_fp32 the FP32 one, p2m/interp either precision,
        # receivers the receiver operator calls)
        out["kernel_timers_ms_per_step"] = {k: round(v[0] / args.steps, 3) for k, v in prof.items() if v[1] > 0}
    if main_path == "learned":
        out["network_s_per_iter"] = round(sum(s["net_seconds"] for s in stats) / len(stats), 4)
        out["mean_abs_lambda"] = round(sum(s["lambda_abs"] for s in stats) / len(stats), 4)
    for p in ("fast", "plain"):
        if p in res and p != main_path:
            ps = res[p][1]
            out[p] = {"s_per_iter": round(res[p][0], 4), "path": path_desc(p),
                      "bicgstab_iterations": sum(s["solve_iters"] for s in ps),
                      "field_bicgstab_iterations": sum(s["field_iters"] for s in ps),
                      "misfit_by_iteration": misfit_by_iteration(ps)}
            if main_path in contrast and p in contrast:
                out[p]["contrast_rel_diff_vs_%s" % main_path] = float(
                    np.linalg.norm(contrast[main_path] - contrast[p]) / np.linalg.norm(contrast[p]))
    if "plain" in res and main_path != "plain":
        out["speedup_vs_plain"] = round(res["plain"][0] / s_iter, 3)
        if "fast" in res:
            out["fast"]["speedup_vs_plain"] = round(res["plain"][0] / res["fast"][0], 3)
    if errors:
        out["errors"] = errors

    if args.ref_geometry == "on" and args.radius > 0:
        # the reference's own geometry: plane-wave transmitters, far-field receivers
        rg_common = dict(common)
        rg_common["radius"] = 0.0
        rg_mes, rg_synth, _ = synthesize(isg, rg_common, batch, otrue, args.synth_iti)
        rg_paths = [p for p in ("learned", "plain") if p in paths]
        # one untimed cycle, then one timed cycle: the contrasts compared below are
        # then those after the cycle's last iteration (a cycle that ended on its
        # first iteration would compare the exact, solve-free first step)
        rg_runs = run_paths(isg, args, wl, rg_common, otrue, rg_mes, rg_paths, CYCLE, CYCLE, False)
        rres, rcon, rerr = rg_runs["res"], rg_runs["contrast"], rg_runs["errors"]
        rg = {"note": "plane waves and far-field receivers (the reference's geometry), one untimed and one "
                      "timed %d-iteration cycle per path, no live kernel timers" % CYCLE, "synthesis": rg_synth,
              "contrast_after_iteration": CYCLE}
        for p in rres:
            rg[p + "_s_per_iter"] = round(rres[p][0], 4)
            rg[p + "_misfit_by_iteration"] = misfit_by_iteration(rres[p][1])
        if "learned" in rres and "plain" in rres:
            rg["speedup_vs_plain"] = round(rres["plain"][0] / rres["learned"][0], 3)
            rg["contrast_rel_diff"] = float(np.linalg.norm(rcon["learned"] - rcon["plain"])
                                            / np.linalg.norm(rcon["plain"]))
        if rerr:
            rg["errors"] = rerr
        out["reference_geometry"] = rg

    if args.cpu_baseline == "on":
        try:
            o_rec = contrast[main_path]
            # transmitter 0's incident field from a context of the same geometry
            s0 = isg.Solver(max_rhs=1, **common)
            d0 = isg.Dbim(s0, otrue, field_start=0, measurements=mes)
            inc0 = d0.get("inc")[: side * side].copy()
            del d0
            s0.close()
            out["cpu_baseline"] = cpu_baselines(isg, level, args.radius, o_rec, inc0, 3 * ntx)
        except Exception as e:  # the CPU oracle needs its .so; report instead of failing the bench
            out["cpu_baseline"] = {"error": "%s: %s" % (type(e).__name__, e)}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

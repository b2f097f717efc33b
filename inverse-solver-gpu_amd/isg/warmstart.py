"""Train and evaluate the warm-start network on fields the solver generates.

    python inverse-solver-gpu_amd/isg/warmstart.py train --phantoms 200 --out W.bin
    python inverse-solver-gpu_amd/isg/warmstart.py eval --weights W.bin --phantoms 3 --level 8 --radius 60

Data: seeded random-ellipse phantoms (20 ellipses, eps_r - 1 in U(0.05, 0.8)
+ 0.02i) at 512x512 pixels (level 7), 64 transmitters, 256 receivers,
rel-res 1e-6, half of them with line sources and receivers on a 30-wavelength
circle (the benchmark's geometry scaled to this domain), half with plane waves
and far-field receivers (the reference's geometry).  For each phantom the
DBIM loop runs `iters` iterations on the GPU with the first-order total-field
start; for each iteration and a random subset of transmitters one sample is
input  (Re/Im new contrast, Re/Im incident field, Re/Im first-order start
        tot_prev - alpha dtot, Re/Im its step -alpha dtot), the field channels
        divided by the transmitter's incident-field rms,
target (new total field - first-order start) / the rms of the first-order
        step (the correction in units of the step it refines),
cropped to `crop` x `crop` pixels.  The network is fully convolutional, so
crops train it and whole images run through it.  Training augments with the 8
symmetries of the square (a rotation or reflection of the whole scene maps
solutions of the Helmholtz equation to solutions).  Everything numerical runs
in libisg.so; torch only holds the dataset in device memory.

`eval` reports, per total-field start (zero, previous iterate, first-order,
learned), the BiCGSTAB iterations and wall time of the total-field solves and
the whole DBIM iteration, on held-out phantoms.
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
import isg  # noqa: E402

TRAIN_SEED0, EVAL_SEED0 = 1000, 900000
STARTS = {"zero": 0, "previous": 1, "first-order": 2, "learned": 3}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def solver_for(level, ntx, nrx, radius=0.0, precision=1):
    return isg.Solver(level=level, numtx=ntx, numrx=nrx, toli=1e-6, iti=300, regular=float(np.sqrt(1e-3)),
                      max_rhs=ntx, precision=precision, radius=radius)


def phantom(solver, side, seed, nell=20):
    _, r2t = solver.geometry()
    return isg.raster_to_tree(isg.phantom_ellipses(side, nell, seed), r2t), r2t


def raster_fields(d, name, ntx, N, r2t, side):
    v = d.get(name)
    return v.reshape(ntx, N)[:, r2t].reshape(ntx, side, side)


def sample_inputs(o, inc, fo, step, rms):
    """Network input (..., 8) float32 of one transmitter: contrast, and the
    incident field, first-order start and first-order step divided by rms."""
    return np.stack([o.real, o.imag, inc.real / rms, inc.imag / rms, fo.real / rms, fo.imag / rms,
                     step.real / rms, step.imag / rms], axis=-1).astype(np.float32)


def gen_samples(level, ntx, nrx, seeds, radii, iters, crop, tx_per, crops_per, rng):
    """Solver-generated samples: X (S, crop, crop, 8), Y (S, crop, crop, 2)
    float32 and the index of each sample's phantom."""
    side = 8 << (level - 1)
    X, Y, P = [], [], []
    solvers = {}
    for i, seed in enumerate(seeds):
        radius = radii[i % len(radii)]
        if radius not in solvers:
            solvers[radius] = solver_for(level, ntx, nrx, radius)
        s = solvers[radius]
        N = s.N
        t0 = time.perf_counter()
        otrue, r2t = phantom(s, side, seed)
        d = isg.Dbim(s, otrue, field_start=2)
        inc = raster_fields(d, "inc", ntx, N, r2t, side)
        rms = np.sqrt(np.mean(np.abs(inc) ** 2, axis=(1, 2)))
        prev = raster_fields(d, "tot", ntx, N, r2t, side)
        for _ in range(iters):
            st = d.iterate()
            alpha = st["alpha"]
            o = d.get("o")[r2t].reshape(side, side)
            new = raster_fields(d, "tot", ntx, N, r2t, side)
            dtot = raster_fields(d, "dtot", ntx, N, r2t, side)
            for t in rng.choice(ntx, size=tx_per, replace=False):
                step = -alpha * dtot[t]
                fo = prev[t] + step
                xin = sample_inputs(o, inc[t], fo, step, rms[t])
                srms = float(np.sqrt(np.mean(np.abs(step) ** 2)))
                if not srms > 0:
                    continue
                dv = (new[t] - fo) / srms
                for _ in range(crops_per):
                    y0, x0 = rng.integers(0, side - crop + 1, 2)
                    sl = (slice(y0, y0 + crop), slice(x0, x0 + crop))
                    X.append(xin[sl])
                    Y.append(np.stack([dv.real[sl], dv.imag[sl]], axis=-1).astype(np.float32))
                    P.append(i)
            prev = new
        del d
        log("phantom %d (radius %g): %d iterations in %.1fs, %d samples"
            % (seed, radius, iters, time.perf_counter() - t0, len(X)))
    for s in solvers.values():
        s.close()
    return np.stack(X), np.stack(Y), np.array(P)


def fit_scales(X, Y):
    """Input scales: 1/std of the contrast channels and of the first-order
    step channels, 1 for the other (normalised) fields; output scale: rms of
    the target."""
    sc = np.ones(isg.WarmNet.NUM_SCALES, np.float32)
    so = float(np.sqrt(np.mean(X[..., 0:2] ** 2)))
    sc[0] = sc[1] = 1.0 / max(so, 1e-6)
    ss = float(np.sqrt(np.mean(X[..., 6:8] ** 2)))
    sc[6] = sc[7] = 1.0 / max(ss, 1e-6)
    sc[-1] = max(float(np.sqrt(np.mean(Y ** 2))), 1e-12)
    return sc


def augment(torch, x, y, k):
    """k in 0..7: rotation by 90 k degrees (k % 4), transposed when k >= 4."""
    if k >= 4:
        x, y = x.transpose(1, 2), y.transpose(1, 2)
    if k % 4:
        x, y = torch.rot90(x, k % 4, (1, 2)), torch.rot90(y, k % 4, (1, 2))
    return x.contiguous(), y.contiguous()


def fit(torch, Xg, Yg, tr, scales, args, rng, seed):
    """Adam with a cosine learning-rate schedule on the samples `tr`."""
    net = isg.WarmNet(max_batch=args.batch, max_pixels=args.crop * args.crop, seed=seed)
    if args.init:
        net.load(args.init)
    net.scales = scales
    trg = torch.from_numpy(tr).cuda()
    B = args.batch
    hist = []
    t0 = time.perf_counter()
    for step in range(args.steps):
        lr = args.lr_min + 0.5 * (args.lr - args.lr_min) * (1 + math.cos(math.pi * step / args.steps))
        idx = trg[torch.randint(0, len(tr), (B,), device="cuda")]
        xb, yb = augment(torch, Xg[idx], Yg[idx], int(rng.integers(0, 8)))
        torch.cuda.synchronize()
        loss = net.loss_grad_dev(xb.data_ptr(), yb.data_ptr(), B, args.crop, args.crop)
        net.adam(lr)
        hist.append(loss)
        if (step + 1) % max(1, args.steps // 10) == 0:
            log("step %d: loss %.4f (mean of last %d), lr %.2e, %.1fs"
                % (step + 1, np.mean(hist[-50:]), min(50, len(hist)), lr, time.perf_counter() - t0))
    return net, hist, time.perf_counter() - t0


def val_loss(torch, net, Xg, Yg, val, args):
    """Mean loss over the validation samples (1.0 = predicting no correction
    when the output scale is the targets' rms)."""
    vl, B = [], args.batch
    valg = torch.from_numpy(val).cuda()
    for i in range(0, len(val) - B + 1, B):
        xb, yb = Xg[valg[i:i + B]].contiguous(), Yg[valg[i:i + B]].contiguous()
        torch.cuda.synchronize()
        vl.append(net.loss_grad_dev(xb.data_ptr(), yb.data_ptr(), B, args.crop, args.crop))
    return float(np.mean(vl)) if vl else None


def train(args):
    import torch
    rng = np.random.default_rng(args.seed)
    seeds = [TRAIN_SEED0 + i for i in range(args.phantoms)]
    radii = [float(r) for r in args.radii.split(",")]
    t0 = time.perf_counter()
    X, Y, P = gen_samples(args.level, args.ntx, args.nrx, seeds, radii, args.iters, args.crop, args.tx, args.crops,
                          rng)
    t_data = time.perf_counter() - t0
    # validation: the last 5 % of the phantoms (held out from every fit)
    nvp = max(1, args.phantoms // 20)
    val = np.nonzero(P >= args.phantoms - nvp)[0]
    tr_all = np.nonzero(P < args.phantoms - nvp)[0]
    scales = fit_scales(X[tr_all], Y[tr_all])
    log("dataset: %d samples of %dx%d (%.1f GB), generated in %.1fs; scales %s"
        % (len(X), args.crop, args.crop, (X.nbytes + Y.nbytes) / 1e9, t_data, scales))
    Xg, Yg = torch.from_numpy(X).cuda(), torch.from_numpy(Y).cuda()
    del X, Y
    curve = []
    for k in sorted(int(c) for c in args.curve.split(",") if c):
        if k >= args.phantoms - nvp:
            continue
        tr = np.nonzero(P < k)[0]
        net_k, _, _ = fit(torch, Xg, Yg, tr, scales, args, rng, args.seed)
        curve.append({"phantoms": k, "samples": int(len(tr)), "val_loss": val_loss(torch, net_k, Xg, Yg, val, args)})
        log("learning curve: %s" % curve[-1])
        net_k.close()
    net, hist, t_train = fit(torch, Xg, Yg, tr_all, scales, args, rng, args.seed)
    vl = val_loss(torch, net, Xg, Yg, val, args)
    curve.append({"phantoms": args.phantoms - nvp, "samples": int(len(tr_all)), "val_loss": vl})
    net.save(args.out)
    res = {"phantoms": args.phantoms, "validation_phantoms": nvp, "radii": radii, "samples": int(len(tr_all)),
           "val_samples": int(len(val)), "steps": args.steps, "batch": args.batch,
           "crop": args.crop, "data_seconds": round(t_data, 1), "train_seconds": round(t_train, 1),
           "train_loss_last50": float(np.mean(hist[-50:])), "val_loss": vl, "learning_curve": curve,
           "scales": [float(v) for v in scales], "weights": args.out}
    print(json.dumps(res), flush=True)


def evaluate(args):
    """Zero / previous / first-order / learned total-field starts on held-out phantoms."""
    side = 8 << (args.level - 1)
    net = isg.WarmNet(max_batch=16, max_pixels=side * side, path=args.weights) if args.weights else None
    out = {k: {"field_iters": 0, "field_seconds": 0.0, "iter_seconds": 0.0, "net_seconds": 0.0,
               "solve_iters": 0, "lambda_abs": [], "start_ratio": []}
           for k in STARTS if k != "learned" or net is not None}
    misfit = {k: [] for k in out}
    for i in range(args.phantoms):
        seed = EVAL_SEED0 + i
        s = solver_for(args.level, args.ntx, args.nrx, args.radius)
        otrue, _ = phantom(s, side, seed)
        gen = isg.Dbim(s, otrue, field_start=2)
        mes = gen.get("mes")
        del gen
        for kind in out:
            d = isg.Dbim(s, otrue, field_start=STARTS[kind], measurements=mes)
            if kind == "learned":
                d.set_warmnet(net)
            d.iterate()  # first iteration (o = 0 before it) is common to all starts: excluded
            for _ in range(args.iters):
                st = d.iterate()
                r = out[kind]
                r["field_iters"] += st["field_iters"]
                r["field_seconds"] += st["field_seconds"]
                r["iter_seconds"] += st["seconds"]
                r["net_seconds"] += st["net_seconds"]
                r["solve_iters"] += st["solve_iters"]
                if kind == "learned":
                    r["lambda_abs"].append(st["lambda_abs"])
                    r["start_ratio"].append(st["start_ratio"])
            misfit[kind].append(st["misfit"])
            del d
        s.close()
        log("eval phantom %d done: %s" % (seed, {k: v["field_iters"] for k, v in out.items()}))
    n = args.phantoms * args.iters
    res = {"config": "%dx%d px, level %d, %d Tx, %d Rx, %s, 20 ellipses, rel-res 1e-6, held-out seeds %d..%d, "
                     "DBIM iterations 2..%d" % (side, side, args.level, args.ntx, args.nrx,
                                                ("line sources on a %g-lambda circle" % args.radius) if args.radius > 0
                                                else "plane waves", EVAL_SEED0,
                                                EVAL_SEED0 + args.phantoms - 1, args.iters + 1),
           "per_dbim_iteration": {}}
    for k, r in out.items():
        res["per_dbim_iteration"][k] = {
            "field_bicgstab_iterations": round(r["field_iters"] / n, 1),
            "field_seconds": round(r["field_seconds"] / n, 4),
            "iteration_seconds": round(r["iter_seconds"] / n, 4),
            "all_bicgstab_iterations": round(r["solve_iters"] / n, 1),
            "net_seconds": round(r["net_seconds"] / n, 4),
            "final_misfit_mean": float(np.mean(misfit[k])),
        }
        if r["lambda_abs"]:
            res["per_dbim_iteration"][k]["mean_abs_lambda"] = round(float(np.mean(r["lambda_abs"])), 3)
            # |initial residual of first-order start + lambda * network| / |of the first-order start|
            res["per_dbim_iteration"][k]["start_residual_ratio"] = round(float(np.mean(r["start_ratio"])), 3)
    print(json.dumps(res), flush=True)


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("mode", choices=["train", "eval"])
    ap.add_argument("--level", type=int, default=7)
    ap.add_argument("--ntx", type=int, default=64)
    ap.add_argument("--nrx", type=int, default=256)
    ap.add_argument("--phantoms", type=int, default=40)
    ap.add_argument("--iters", type=int, default=4)
    ap.add_argument("--crop", type=int, default=128)
    ap.add_argument("--crops", type=int, default=2, help="crops per sampled transmitter and iteration")
    ap.add_argument("--tx", type=int, default=16, help="transmitters sampled per phantom and iteration")
    ap.add_argument("--radii", default="30,0", help="training geometries, cycled over the phantoms (0: plane waves)")
    ap.add_argument("--radius", type=float, default=0.0, help="eval geometry (0: plane waves)")
    ap.add_argument("--curve", default="", help="also fit on the first k training phantoms, k in this comma "
                                                "list, and report the validation loss of each (learning curve)")
    ap.add_argument("--steps", type=int, default=3000)
    ap.add_argument("--batch", type=int, default=16)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--lr-min", type=float, default=1e-5)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--init", default=None, help="weights to continue from")
    ap.add_argument("--out", default="warmnet.bin")
    ap.add_argument("--weights", default=None)
    a = ap.parse_args()
    if a.mode == "train":
        train(a)
    else:
        evaluate(a)


if __name__ == "__main__":
    main()

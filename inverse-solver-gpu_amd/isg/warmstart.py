"""Train and evaluate the warm-start network on fields the solver generates.

    python inverse-solver-gpu_amd/isg/warmstart.py train --phantoms 40 --iters 4 --steps 3000 --out W.bin
    python inverse-solver-gpu_amd/isg/warmstart.py eval --weights W.bin --phantoms 3 --iters 4

Data: seeded random-ellipse phantoms (20 ellipses, eps_r - 1 in U(0.05, 0.8)
+ 0.02i) at 512x512 pixels (level 7), 64 plane-wave transmitters, 256
receivers, rel-res 1e-6.  For each phantom the DBIM loop runs `iters`
iterations on the GPU; every transmitter of every iteration gives one sample:
input (Re/Im new contrast, Re/Im incident field, Re/Im previous total field),
target = new total field - previous total field (the correction the network
predicts), cropped to `crop` x `crop` pixels.  The network is fully
convolutional, so crops train it and whole images run through it.
Training augments with the 8 symmetries of the square (a rotation or
reflection of the whole scene maps solutions of the Helmholtz equation to
solutions).  Everything numerical runs in libisg.so; torch only holds the
dataset in device memory.

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


def solver_for(level, ntx, nrx, precision=1):
    return isg.Solver(level=level, numtx=ntx, numrx=nrx, toli=1e-6, iti=300, regular=float(np.sqrt(1e-3)),
                      max_rhs=ntx, precision=precision)


def phantom(solver, side, seed, nell=20):
    _, r2t = solver.geometry()
    return isg.raster_to_tree(isg.phantom_ellipses(side, nell, seed), r2t), r2t


def raster_fields(d, name, ntx, N, r2t, side):
    v = d.get(name)
    return v.reshape(ntx, N)[:, r2t].reshape(ntx, side, side)


def gen_samples(level, ntx, nrx, seeds, iters, crop, crops_per, rng):
    """Solver-generated samples: X (S, crop, crop, 6), Y (S, crop, crop, 2) float32."""
    side = 8 << (level - 1)
    s = solver_for(level, ntx, nrx)
    N = s.N
    X, Y = [], []
    for seed in seeds:
        t0 = time.perf_counter()
        otrue, r2t = phantom(s, side, seed)
        d = isg.Dbim(s, otrue, field_start=2)
        inc = raster_fields(d, "inc", ntx, N, r2t, side)
        prev = raster_fields(d, "tot", ntx, N, r2t, side)
        for _ in range(iters):
            d.iterate()
            o = d.get("o")[r2t].reshape(side, side)
            new = raster_fields(d, "tot", ntx, N, r2t, side)
            for t in range(ntx):
                for _ in range(crops_per):
                    y0, x0 = rng.integers(0, side - crop + 1, 2)
                    sl = (slice(y0, y0 + crop), slice(x0, x0 + crop))
                    xs = np.stack([o.real[sl], o.imag[sl], inc[t].real[sl], inc[t].imag[sl],
                                   prev[t].real[sl], prev[t].imag[sl]], axis=-1)
                    dv = new[t][sl] - prev[t][sl]
                    X.append(xs.astype(np.float32))
                    Y.append(np.stack([dv.real, dv.imag], axis=-1).astype(np.float32))
            prev = new
        del d
        log("phantom %d: %d iterations in %.1fs, %d samples" % (seed, iters, time.perf_counter() - t0, len(X)))
    s.close()
    return np.stack(X), np.stack(Y)


def fit_scales(X, Y):
    """Input scales: 1/std of the contrast channels, 1 for the fields; output
    scale: rms of the target."""
    sc = np.ones(7, np.float32)
    so = float(np.sqrt(np.mean(X[..., 0:2] ** 2)))
    sc[0] = sc[1] = 1.0 / max(so, 1e-6)
    sc[6] = max(float(np.sqrt(np.mean(Y ** 2))), 1e-12)
    return sc


def augment(torch, x, y, k):
    """k in 0..7: rotation by 90 k degrees (k % 4), transposed when k >= 4."""
    if k >= 4:
        x, y = x.transpose(1, 2), y.transpose(1, 2)
    if k % 4:
        x, y = torch.rot90(x, k % 4, (1, 2)), torch.rot90(y, k % 4, (1, 2))
    return x.contiguous(), y.contiguous()


def train(args):
    import torch
    rng = np.random.default_rng(args.seed)
    seeds = [TRAIN_SEED0 + i for i in range(args.phantoms)]
    t0 = time.perf_counter()
    X, Y = gen_samples(args.level, args.ntx, args.nrx, seeds, args.iters, args.crop, args.crops, rng)
    t_data = time.perf_counter() - t0
    nval = max(1, len(X) // 20)
    perm = rng.permutation(len(X))
    val, tr = perm[:nval], perm[nval:]
    scales = fit_scales(X[tr], Y[tr])
    log("dataset: %d samples of %dx%d (%.1f GB), generated in %.1fs; scales %s"
        % (len(X), args.crop, args.crop, (X.nbytes + Y.nbytes) / 1e9, t_data, scales))
    Xg, Yg = torch.from_numpy(X).cuda(), torch.from_numpy(Y).cuda()
    del X, Y
    net = isg.WarmNet(max_batch=args.batch, max_pixels=args.crop * args.crop, seed=args.seed)
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
        if (step + 1) % max(1, args.steps // 20) == 0:
            log("step %d: loss %.4f (mean of last %d), lr %.2e, %.1fs"
                % (step + 1, np.mean(hist[-50:]), min(50, len(hist)), lr, time.perf_counter() - t0))
    t_train = time.perf_counter() - t0
    # validation loss (normalised: 1.0 = predicting no correction)
    vl = []
    valg = torch.from_numpy(val).cuda()
    for i in range(0, len(val) - B + 1, B):
        xb, yb = Xg[valg[i:i + B]].contiguous(), Yg[valg[i:i + B]].contiguous()
        torch.cuda.synchronize()
        vl.append(net.loss_grad_dev(xb.data_ptr(), yb.data_ptr(), B, args.crop, args.crop))
    net.save(args.out)
    res = {"samples": int(len(tr)), "val_samples": int(len(val)), "steps": args.steps, "batch": B,
           "crop": args.crop, "data_seconds": round(t_data, 1), "train_seconds": round(t_train, 1),
           "train_loss_last50": float(np.mean(hist[-50:])), "val_loss": float(np.mean(vl)) if vl else None,
           "scales": [float(v) for v in scales], "weights": args.out}
    print(json.dumps(res), flush=True)


def evaluate(args):
    """Zero / previous / first-order / learned total-field starts on held-out phantoms."""
    side = 8 << (args.level - 1)
    net = isg.WarmNet(max_batch=args.batch, max_pixels=side * side, path=args.weights) if args.weights else None
    out = {k: {"field_iters": 0, "field_seconds": 0.0, "iter_seconds": 0.0, "net_seconds": 0.0,
               "solve_iters": 0, "lambda_abs": []} for k in STARTS if k != "learned" or net is not None}
    misfit = {k: [] for k in out}
    for i in range(args.phantoms):
        seed = EVAL_SEED0 + i
        s = solver_for(args.level, args.ntx, args.nrx)
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
            misfit[kind].append(st["misfit"])
            del d
        s.close()
        log("eval phantom %d done: %s" % (seed, {k: v["field_iters"] for k, v in out.items()}))
    n = args.phantoms * args.iters
    res = {"config": "%dx%d px, level %d, %d Tx, %d Rx, 20 ellipses, rel-res 1e-6, held-out seeds %d..%d, "
                     "DBIM iterations 2..%d" % (side, side, args.level, args.ntx, args.nrx, EVAL_SEED0,
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
    ap.add_argument("--crops", type=int, default=1, help="crops per transmitter and iteration")
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

"""Learned troubled-cell indicator: training data, training on the GPU, evaluation.

Training data are element states of seeded runs at N = 7 on the curved mesh family on [0, 10]^2,
over a spread of mesh sizes (32^2 .. 512^2 elements) and run lengths, advanced by the solver with
the modal sensor and artificial viscosity (the sigma thresholds of the benchmark):
- dam breaks: step height ratio U(1.2, 5), step angle U(0, pi), a sharp step or one spread over
  0.5 .. 2 elements, noise U(0, 2e-3), b = 0.1 sin x cos y, eps0 grown with the step height;
- floods: seeded 16-mode Fourier bathymetry, a reservoir level U(0.5, 1) behind x = U(2, 5) over
  a dry or shallow plain (level U(-0.3, 0.1)), with the positivity limiter (dry tolerance 0.1,
  CFL 0.1, the flood benchmark's settings).
Every state is labelled by the modal sensor (troubled = positive viscosity, i.e. sigma above
sigma_min).  Per snapshot at most a fixed number of troubled elements is kept and smooth ones are
subsampled, so that no mesh size dominates.  The network (include/esdg_mlp.h) is trained with the
library's own forward/backward kernels and Adam on the device.  Weights are stored as JSON beside
this file.

Command line (runs on a GPU):  python -m esdgsem_mpiocca_amd.indicator --out weights.json
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

from .setup_api import Case, RunParams

WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "indicator_weights.json")


def av_run_params(**kw) -> RunParams:
    """The artificial-viscosity configuration of the curved dam-break benchmark (DFL 0.05: with 0.1 the
    512^2 run diverges once the viscous step limit governs dt)."""
    base = dict(g=9.81, CFL=0.5, DFL=0.05, num_flux=3, ES=1, flux_differencing=1, artificial_viscosity=1,
                positivity=0, pos_pres_tol=1e-8, eps0=0.005, sigma_min=-4.5, sigma_max=-2.5, rk_order=4,
                rk_ssp=0)
    base.update(kw)
    return RunParams(**base)


def dam_break_state(case: Case, ratio: float, angle: float, noise: float, rng, width: float = 0.0) -> np.ndarray:
    """h = ratio | 1 across the line through (5, 5) with normal angle `angle` (a tanh ramp of the
    given width when width > 0), plus U(-noise, noise), zero momentum, on the nodes of a curved
    dam-break mesh (b = 0.1 sin x cos y)."""
    x, y = case["x"], case["y"]
    side = (x - 5.0) * np.cos(angle) + (y - 5.0) * np.sin(angle)
    if width > 0.0:
        step = 1.0 + (ratio - 1.0) * 0.5 * (1.0 - np.tanh(side / width))
    else:
        step = np.where(side < 0.0, ratio, 1.0)
    h = step + noise * (2.0 * rng.uniform(size=x.shape) - 1.0)
    q = np.zeros((case.Nelem, 3, case.ngl ** 2))
    q[:, 0] = h.reshape(case.Nelem, -1)
    return q.ravel()


def load_weights(path: str = WEIGHTS):
    if not os.path.exists(path):
        return None, None
    with open(path) as f:
        d = json.load(f)
    return np.asarray(d["params"], dtype=np.float32), d.get("meta", {})


def save_weights(params, meta, path: str = WEIGHTS):
    with open(path, "w") as f:
        json.dump({"params": [float(v) for v in np.asarray(params, dtype=np.float32)], "meta": meta}, f)


class _Labeller:
    """Features (device) and modal-sensor labels of a solver's current state."""

    def __init__(self, dev, ctx, case, solver):
        self.dev, self.ctx, self.case, self.solver = dev, ctx, case, solver
        n = case.Nelem
        self.q = solver.view("q", n * 3 * 64)
        self.B = ctx.array(case["b"])
        self.V = ctx.array(case["VdmInv"])
        self.sizes = ctx.array(case["EleSizes"])
        self.feat = ctx.empty(n * 72, np.float32)
        self.visc = ctx.empty(n)
        self.vplot = ctx.empty(n)

    def __call__(self):
        n = self.case.Nelem
        self.dev.mlp_features(self.ctx, n, self.q, self.B, self.V, self.feat)
        self.ctx.shock_capturing(n, self.q, self.B, self.V, self.sizes, self.visc, self.vplot)
        return self.feat.download().reshape(n, 72), self.visc.download()


MESH_SIZES = (32, 48, 64, 96, 128, 192, 256, 384, 512)


def _scenario(rng):
    """One seeded training run: (case, run params, initial state or None, label)."""
    nel = int(rng.choice(MESH_SIZES))
    if rng.uniform() < 0.3:
        left, x_dam, right = rng.uniform(0.5, 1.0), rng.uniform(2.0, 5.0), rng.uniform(-0.3, 0.1)
        case = Case.flood(N=7, nel=nel, seed=int(rng.integers(1 << 30)), level_left=left, x_dam=x_dam,
                          level_right=right)
        rp = av_run_params(positivity=1, pos_pres_tol=0.1, CFL=0.1)
        return case, rp, None, f"flood nel {nel} levels {left:.2f}|{right:.2f} at x {x_dam:.2f}"
    ratio, angle = rng.uniform(1.2, 5.0), rng.uniform(0.0, np.pi)
    width = 0.0 if rng.uniform() < 0.5 else rng.uniform(0.5, 2.0) * 10.0 / nel
    noise = rng.uniform(0.0, 2e-3)
    case = Case.dam_break(N=7, nel=nel, noise=0.0, seed=0)
    # the viscosity amplitude grows with the step and the positivity limiter keeps the depth
    # non-negative, so that strong oblique steps stay stable; the labels depend on the sensor's
    # sigma only (sigma > sigma_min), not on eps0 or the limiter
    rp = av_run_params(eps0=0.005 * max(1.0, ratio - 1.0), positivity=1, pos_pres_tol=1e-3)
    q0 = dam_break_state(case, ratio, angle, noise, rng, width=width)
    return case, rp, q0, f"dam break nel {nel} ratio {ratio:.2f} angle {angle:.2f} width {width:.3f}"


def generate(dev, runs=64, snapshots=8, max_steps=400, max_troubled=4000, smooth_per_troubled=3.0,
             smooth_min=2000, seed=0, log=print):
    """Features and labels from seeded runs; returns (X float32 [n][72], y float32 [n])."""
    rng = np.random.default_rng(seed)
    Xs, ys = [], []
    for r in range(runs):
        case, rp, q0, what = _scenario(rng)
        total = int(rng.integers(40, max_steps + 1))
        every = max(1, total // snapshots)
        ctx = dev.Context(dev.make_params(case, rp))
        try:
            solver = dev.Solver(ctx, case, rp, fused=1)
            lab = _Labeller(dev, ctx, case, solver)
            if q0 is not None:
                solver.set_q(q0)
            dt0 = None
            for k in range(snapshots + 1):
                if k:
                    try:
                        solver.run(every)
                    except dev.EsdgDeviceError as e:   # a run that breaks down contributes what it has
                        log(f"# run {r} ({what}) stopped: {e}")
                        break
                    # a collapsing time step marks a run that is breaking down: its states are not kept
                    dt = solver.last_dt()
                    dt0 = dt if dt0 is None else dt0
                    if not (dt > 0.2 * dt0):
                        log(f"# run {r} ({what}): dt fell to {dt:.3g}, run dropped here")
                        break
                F, visc = lab()
                if not np.isfinite(F).all():
                    log(f"# run {r} ({what}): non-finite state, snapshot dropped")
                    break
                pos = np.nonzero(visc > 0)[0]
                neg = np.nonzero(visc <= 0)[0]
                if len(pos) > max_troubled:
                    pos = rng.choice(pos, max_troubled, replace=False)
                m = min(len(neg), max(smooth_min, int(smooth_per_troubled * len(pos))))
                take = np.concatenate([pos, rng.choice(neg, m, replace=False)])
                Xs.append(F[take])
                ys.append((visc[take] > 0).astype(np.float32))
            solver.close()
        finally:
            ctx.close()
        log(f"# run {r} ({what}, {every * snapshots} steps): {sum(len(v) for v in ys)} samples, "
            f"{int(sum(v.sum() for v in ys))} troubled")
    return np.concatenate(Xs).astype(np.float32), np.concatenate(ys).astype(np.float32)


def train(dev, ctx, X, y, epochs=8, batch=4096, lr=2e-3, pos_weight=None, seed=0, params=None, log=print):
    """Adam on the device; returns (params, history)."""
    rng = np.random.default_rng(seed)
    n = len(y)
    if pos_weight is None:
        pos_weight = float(min(20.0, max(1.0, (n - y.sum()) / max(y.sum(), 1.0))))
    net = dev.MLP(ctx, seed=seed, params=params)
    dX, dy = ctx.array(X), ctx.array(y)
    hist = []
    try:
        for ep in range(epochs):
            perm = rng.permutation(n).astype(np.int32)
            dperm = ctx.array(perm)
            lr_ep = lr * (0.5 ** (ep // 3))
            steps = n // batch
            tot = 0.0
            for s in range(steps):
                view = dev.DeviceView(ctx, dperm.ptr + 4 * s * batch, batch, np.int32)
                tot += net.train_step(batch, dX, dy, view, pos_weight, lr_ep, want_loss=(s % 64 == 0))
            dperm.free()
            dp = ctx.empty(n, np.float32)
            net.infer(dX, n, dp)
            p = dp.download()
            dp.free()
            acc = float(np.mean((p > 0.5) == (y > 0.5)))
            fn = float(np.sum((p <= 0.5) & (y > 0.5)) / max(y.sum(), 1))
            hist.append({"epoch": ep, "accuracy": acc, "false_negative_rate": fn, "lr": lr_ep})
            log(f"# epoch {ep}: accuracy {acc:.4f}, false negatives {fn:.4f}")
        params = net.get_params()
    finally:
        net.close()
        dX.free()
        dy.free()
    return params, {"pos_weight": pos_weight, "epochs": hist}


def main(argv=None):
    ap = argparse.ArgumentParser(description="train the learned troubled-cell indicator on the GPU")
    ap.add_argument("--out", default=WEIGHTS)
    ap.add_argument("--runs", type=int, default=64)
    ap.add_argument("--epochs", type=int, default=8)
    ap.add_argument("--seed", type=int, default=0)
    args = ap.parse_args(argv)
    import esdgsem_mpiocca_amd as E
    dev = E.device()
    log = lambda m: print(m, file=sys.stderr, flush=True)  # noqa: E731
    t0 = time.perf_counter()
    X, y = generate(dev, runs=args.runs, seed=args.seed, log=log)
    t_gen = time.perf_counter() - t0
    log(f"# dataset: {len(y)} samples, {int(y.sum())} troubled, {t_gen:.1f} s")
    # hold out 10% for validation
    rng = np.random.default_rng(args.seed + 1)
    perm = rng.permutation(len(y))
    nv = len(y) // 10
    Xv, yv, Xt, yt = X[perm[:nv]], y[perm[:nv]], X[perm[nv:]], y[perm[nv:]]
    case = Case.dam_break(N=7, nel=2, noise=0.0)
    ctx = dev.Context(dev.make_params(case, av_run_params()))
    try:
        t0 = time.perf_counter()
        params, hist = train(dev, ctx, Xt, yt, epochs=args.epochs, seed=args.seed, log=log)
        t_train = time.perf_counter() - t0
        net = dev.MLP(ctx, params=params)
        dX = ctx.array(Xv)
        dp = ctx.empty(len(yv), np.float32)
        net.infer(dX, len(yv), dp)
        p = dp.download()
        net.close()
    finally:
        ctx.close()
    val = {"accuracy": float(np.mean((p > 0.5) == (yv > 0.5))),
           "false_negative_rate_at_0.5": float(np.sum((p <= 0.5) & (yv > 0.5)) / max(yv.sum(), 1)),
           "false_negative_rate_at_0.1": float(np.sum((p <= 0.1) & (yv > 0.5)) / max(yv.sum(), 1)),
           "samples": int(len(yv))}
    meta = {"dataset": {"samples": int(len(y)), "troubled": int(y.sum()), "runs": args.runs,
                        "mesh_sizes": list(MESH_SIZES), "seconds": t_gen},
            "training": {**hist, "seconds": t_train}, "validation": val}
    save_weights(params, meta, args.out)
    print(json.dumps({"validation": val, "train_seconds": t_train, "samples": int(len(y))}), flush=True)


if __name__ == "__main__":
    main()

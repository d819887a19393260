#!/usr/bin/env python3
"""Benchmark of the per-stage SWE residual + time step on one MI355X.

Workload (default): the curved dam break -- 512 x 512 quads on [0,10]^2 with nodes displaced by
0.05 sin(pi x/5) sin(pi y/5) (isoparametric, N = 7), wall boundaries, h = 2 | 1 at x = 5 plus seeded
U(-1e-3, 1e-3) noise, b = 0.1 sin x cos y, modal shock sensor + artificial viscosity, rotated
entropy-stable surface flux, flux-differencing EC volume term, 5-stage 4th-order low-storage RK.
Synthetic data, FP64 throughout.

A step is one full time step (dt from the device reductions, then 5 RK stages); the metric is
DOF-stage updates per second = 3 (N+1)^2 Nelem stages / s over exactly --steps timed steps, with
the state resident in HBM.  The dominant kernel's roofline comes from HIP events recorded around
every launch on the solver's stream during the timed steps; its algorithmic bytes per node are the
reference decomposition's compulsory traffic.  The CPU baseline times the oracle (a single-core C
port of the reference kernels) on a bounded sample of the same workload.

Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0          # MI355X HBM3E
FP64_PEAK_TFLOPS = 78.6        # MI355X FP64 (vector and matrix)

# Algorithmic per-unit work of each kernel (reference decomposition; ngl = N+1).
#   volume flux-diff: 13 doubles/node (q 3, Jac, 4 metrics, Bx, By, Qt 3); ngl^2 (92 ngl + 10) flop/elem
def volume_bytes(nelem, ngl):
    return 13 * 8 * nelem * ngl * ngl


def volume_flops(nelem, ngl):
    return nelem * ngl * ngl * (92 * ngl + 10)


KERNELS = ["volume_kernel", "collect_edge_data", "calc_num_fluxes", "surface_kernel", "calc_rk",
           "update_kernel_ls", "shock_capturing", "calc_num_fluxes_gradient", "calc_gradient",
           "surface_kernel_gradient", "scale_gradient", "collect_edge_data_gradient",
           "volume_kernel_viscose", "calc_num_fluxes_viscose", "surface_kernel_visc",
           "find_lambda_max", "reduce_max", "pos_pres", "fused_face", "fused_volume", "fused_av",
           "fused_viscous_face", "fused_update", "fused_timestep"]


def c1_params(RunParams):
    return RunParams(g=9.81, CFL=0.5, DFL=0.1, num_flux=3, ES=1, flux_differencing=1,
                     artificial_viscosity=1, positivity=0, pos_pres_tol=1e-8, eps0=0.005,
                     sigma_min=-4.5, sigma_max=-2.5, rk_order=4, rk_ssp=0)


def cpu_baseline(nel, steps, N):
    """Time the oracle on a bounded sample of the workload (same mesh family, N, physics)."""
    from esdgsem_mpiocca_amd import Case, RunParams
    from oracle import binding as O
    case = Case.dam_break(N=N, nel=nel, noise=1e-3, seed=0)
    rp = c1_params(RunParams)
    O.timeloop(case, rp, 1)                      # warm-up (page-in, caches)
    t0 = time.perf_counter()
    _, n, _, _ = O.timeloop(case, rp, steps)
    dt = time.perf_counter() - t0
    dofs = 3 * case.ngl ** 2 * case.Nelem
    return {"value": dofs * 5 * n / dt, "unit": "DOF-stage updates/s", "cores": 1, "kind": "port",
            "sample": f"{nel}x{nel} curved dam break, N={N}, AV, {n} steps x 5 stages "
                      f"({dt:.1f} s, single thread, {platform.processor() or platform.machine()})"}


def load_traffic(name):
    path = os.path.join(ROOT, "profiles", f"{name}_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f)


def main():
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--nel", type=int, default=512)
    ap.add_argument("--N", type=int, default=7)
    ap.add_argument("--fused", type=int, default=1,
                    help="1: fused stage schedule (default); 0: the reference's kernel-by-kernel sequence")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-sample-nel", type=int, default=128)
    ap.add_argument("--cpu-sample-steps", type=int, default=15)
    ap.add_argument("--no-profile", action="store_true", help="time without per-launch events")
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        if int(os.environ.get("RANK", "0")) == 0:
            print("bench.py measures one GPU: this build has no multi-GPU path (the halo exchange "
                  "is out of scope); run with --gpus 1", file=sys.stderr, flush=True)
        sys.exit(2)

    import esdgsem_mpiocca_amd as E
    dev = E.device()

    t_setup = time.perf_counter()
    case = E.Case.dam_break(N=args.N, nel=args.nel, noise=1e-3, seed=0)
    rp = c1_params(E.RunParams)
    ctx = dev.Context(dev.make_params(case, rp), device=0)
    solver = dev.Solver(ctx, case, rp, fused=bool(args.fused))
    setup_s = time.perf_counter() - t_setup
    print(f"# setup {setup_s:.1f} s: {case.Nelem} elements, N={case.N}, {ctx.bytes_allocated()/2**30:.2f} GiB "
          "on device", file=sys.stderr, flush=True)

    if args.warmup:
        solver.run(args.warmup)
    ctx.finish()
    ctx.profile(not args.no_profile)
    ctx.profile_reset()
    t0 = time.perf_counter()
    n, t_sim = solver.run(args.steps)
    ctx.finish()
    wall = time.perf_counter() - t0
    if n != args.steps:
        raise SystemExit(f"solver stopped after {n} of {args.steps} steps")
    q = solver.get_q()
    import numpy as np
    if not np.isfinite(q).all():
        raise SystemExit("non-finite state after the timed steps")

    stages = 5
    dofs = 3 * case.ngl ** 2 * case.Nelem
    value = dofs * stages * n / wall

    per_kernel = {}
    for k in KERNELS:
        launches, ms = ctx.profile_query(k)
        if launches:
            per_kernel[k] = {"launches": launches, "avg_us": 1e3 * ms / launches, "total_ms": ms}
    ctx.profile(False)
    gpu_ms = sum(v["total_ms"] for v in per_kernel.values())
    for v in per_kernel.values():
        v["share"] = v["total_ms"] / gpu_ms if gpu_ms else None

    vol_name = "fused_volume" if args.fused else "volume_kernel"
    roofline = None
    if vol_name in per_kernel:
        avg_s = per_kernel[vol_name]["avg_us"] * 1e-6
        gbs = volume_bytes(case.Nelem, case.ngl) / avg_s / 1e9
        tfl = volume_flops(case.Nelem, case.ngl) / avg_s / 1e12
        tr = load_traffic(vol_name)
        roofline = {"kernel": vol_name, "bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS,
                    "unit": "GB/s", "frac": round(gbs / HBM_PEAK_GBS, 4),
                    "traffic": tr.get("hbm_bytes_per_launch") if tr else None,
                    "algorithmic_bytes_per_launch": volume_bytes(case.Nelem, case.ngl),
                    "avg_launch_us": round(per_kernel[vol_name]["avg_us"], 2),
                    "fp64_tflops_algorithmic": round(tfl, 2),
                    "fp64_frac_of_peak": round(tfl / FP64_PEAK_TFLOPS, 4)}

    cpu = None
    if not args.no_cpu_baseline:
        try:
            cpu = cpu_baseline(args.cpu_sample_nel, args.cpu_sample_steps, args.N)
        except Exception as e:  # the baseline is a reported figure; a failure must not hide the GPU line
            cpu = {"error": str(e)}

    out = {
        "metric": "FP64 DOF-stage updates/s at N=7 (curved dam break, AV)",
        "value": value,
        "unit": "DOF-stage updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * wall / n,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"curved dam break {args.nel}x{args.nel} quads, N={args.N}, walls, "
                               "modal sensor + artificial viscosity, rotated ES flux, EC flux "
                               "differencing, LSRK4(5)",
                   "Nelem": case.Nelem, "N": case.N, "dofs": dofs, "stages_per_step": stages,
                   "schedule": "fused" if args.fused else "reference kernel sequence",
                   "parallelism": "single GPU"},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "kernels": {k: {"avg_us": round(v["avg_us"], 2), "share": round(v["share"], 4)}
                    for k, v in sorted(per_kernel.items(), key=lambda kv: -kv[1]["total_ms"])},
        "sim_time": t_sim,
    }
    solver.close()
    ctx.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

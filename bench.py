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
reference decomposition's compulsory traffic.  The CPU baseline times the oracle (a C port of the
reference kernels, OpenMP over the element and face loops) on a bounded sample of the same workload.

Ensemble (also by default, reported under "ensemble"): 256 seeded tsunami scenarios over a 4000 m
ocean with a shelf (sizes {64,96,128,192,256}^2, N in {4,5,7}, elliptic Gaussian uplifts, open and
wall boundaries) advanced together by the batched single-kernel stage with the autotuned
schedule; scenario-steps/s, against per-scenario loops over the same scenarios (one solver per
distinct mesh, the state swapped in per scenario, only the stepping timed).

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


# fused volume pass (volume + surface lift [+ viscous terms] + LSRK in place): per node q 3,
# Yxi/Yeta/Xxi/Xeta 4, 1/J, Bx, By, gRK 3 read and gRK 3, q 3 written (19 doubles); per element
# side node F* 3 [+ viscous face flux 2] read.  (The viscous volume term, read only where the
# viscosity is non-zero, is not counted.)
def fused_volume_bytes(nelem, ngl, av, traces=True, sensor=False, stages=5, limiter=False):
    """Average bytes per launch.  Per node: q 3, Yxi/Yeta/Xxi/Xeta 4, 1/J, Bx, By, gRK 3 read, gRK 3
    and q 3 written; per element side node: F* 3 read, and with `traces` the side traces of the new
    state (3) written.  With `traces` the pass also evaluates, on the new state, the sensor when
    `sensor` (b read per node; the viscosity and its plot copy written per element) and, in the
    last of the step's `stages`, the element wave speed (one double per element).  With `limiter`
    the new state is limited before it is stored (the element size read per element; the GLL
    weights are 8 ngl bytes per launch).  The viscous terms (with `av`) are read only for the
    elements with viscosity in reach, a few percent of them, and are not counted."""
    per_elem = 19 * ngl * ngl + (12 + (12 if traces else 0)) * ngl
    if traces and sensor:
        per_elem += ngl * ngl + 2
    if traces:
        per_elem += 1.0 / stages
    if limiter:
        per_elem += 1
    return int(round(8 * nelem * per_elem))


KERNELS = ["volume_kernel", "collect_edge_data", "calc_num_fluxes", "surface_kernel", "calc_rk",
           "update_kernel_ls", "shock_capturing", "calc_num_fluxes_gradient", "calc_gradient",
           "surface_kernel_gradient", "scale_gradient", "collect_edge_data_gradient",
           "volume_kernel_viscose", "calc_num_fluxes_viscose", "surface_kernel_visc",
           "find_lambda_max", "reduce_max", "pos_pres", "fused_face", "fused_volume", "fused_av",
           "fused_viscous_face", "fused_clock", "mlp_indicator", "write_traces"]


def c1_params(RunParams):
    # DFL 0.05: with 0.1 the viscous step limit, which governs dt at 512^2 wherever the sensor fires,
    # lets the run diverge within 100 steps (the oracle too); 0.05 keeps it bounded over 500 steps
    return RunParams(g=9.81, CFL=0.5, DFL=0.05, num_flux=3, ES=1, flux_differencing=1,
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
    threads = int(O.lib().oracle_threads())
    return {"value": dofs * 5 * n / dt, "unit": "DOF-stage updates/s", "cores": threads, "kind": "port",
            "sample": f"{nel}x{nel} curved dam break, N={N}, AV, {n} steps x 5 stages "
                      f"({dt:.1f} s, {threads} OpenMP threads over element/face loops, "
                      f"{platform.processor() or platform.machine()}, nproc {os.cpu_count()})"}


def ensemble_bench(dev, args):
    """Batched ensemble vs per-scenario loops on the seeded scenario set."""
    import numpy as np
    from esdgsem_mpiocca_amd import scenarios
    t0 = time.perf_counter()
    params = scenarios.draw(args.ens_n, seed=0)
    members = scenarios.build(params)
    t_build = time.perf_counter() - t0
    n = len(members)
    K, W = args.ens_steps, args.ens_warmup
    nodes = sum(c.Nelem * c.ngl ** 2 for c, _, _ in members)
    case0, rp0, _ = members[0]
    ctx = dev.Context(dev.make_params(case0, rp0), device=0)
    t0 = time.perf_counter()
    ens = dev.Ensemble(ctx, members)
    t_create = time.perf_counter() - t0
    print(f"# ensemble: {n} scenarios, {nodes/1e6:.1f} M nodes, build {t_build:.1f} s, "
          f"create {t_create:.1f} s, {ens.nbytes()/2**30:.2f} GiB", file=sys.stderr, flush=True)
    tuned, tuned_ms = None, None
    if args.ens_schedule:
        ens.set_schedule(**json.loads(args.ens_schedule))
        tuned = ens.schedule().as_dict()
    elif args.ens_autotune:
        tuned, tuned_ms = ens.autotune(trial_steps=2)
        print(f"# ensemble autotune: {tuned} ({tuned_ms:.2f} ms/step)", file=sys.stderr, flush=True)
    ens.run(W)
    t0 = time.perf_counter()
    ens.run_async(K)
    ctx.finish()
    wall = time.perf_counter() - t0
    steps = [ens.status(i)[1] for i in range(n)]
    if any(s != W + K for s in steps):
        raise SystemExit("ensemble: a scenario stopped early")
    if not all(np.isfinite(ens.get_q(i)).all() for i in (0, n // 2, n - 1)):
        raise SystemExit("ensemble: non-finite state")
    batched = n * K / wall
    # the stage kernel's launch time (eager steps with per-launch events)
    ctx.profile(True)
    ctx.profile_reset()
    ens.run(2)
    launches, ms = ctx.profile_query("stage")
    ctx.profile(False)
    stage_us = 1e3 * ms / (2 * 5) if launches else None     # per RK stage (all launches of a stage)
    ens.close()
    # per-scenario loops: one context + solver per distinct mesh; only the stepping is timed
    loops = {}
    for fused, name in ((2, "single_kernel"), (0, "reference_sequence")) if args.ens_loops else ():
        solvers, tsum = {}, 0.0
        try:
            for case, rp, q in members:
                key = id(case)
                if key not in solvers:
                    c = dev.Context(dev.make_params(case, rp), device=0)
                    solvers[key] = (c, dev.Solver(c, case, rp, fused=fused))
                    solvers[key][1].run(2)
                c, s = solvers[key]
                s.set_q(q)
                c.finish()
                t1 = time.perf_counter()
                got, _ = s.run(K)
                tsum += time.perf_counter() - t1
                if got != K:
                    raise SystemExit("ensemble loop: a scenario stopped early")
        finally:
            for c, s in solvers.values():
                s.close()
                c.close()
        loops[name] = n * K / tsum
        print(f"# per-scenario loop ({name}): {loops[name]:.1f} scenario-steps/s", file=sys.stderr, flush=True)
    ctx.close()
    stage_bytes = 12 * 8 * nodes + 7 * 8 * sum(c.Nelem * c.ngl ** 2 for c in {id(m[0]): m[0] for m in members}.values())
    return {
        "metric": "ensemble scenario-steps/s",
        "value": batched,
        "unit": "scenario-steps/s",
        "ms_per_step": 1e3 * wall / K,
        "steps": K, "warmup": W,
        "scenarios": n, "nodes": nodes, "dofs": 3 * nodes,
        "dof_stage_updates_per_s": 3 * nodes * 5 * K / wall,
        "per_scenario_loop": loops,
        "speedup_vs_loop": {k: batched / v for k, v in loops.items()} or None,
        "schedule": tuned, "autotune_ms_per_step": tuned_ms,
        "stage_kernel": {"us_per_stage": stage_us, "launches_per_stage": launches / 10 if launches else None,
                         "compulsory_bytes_per_launch": stage_bytes,
                         "achieved_GBs": stage_bytes / (stage_us * 1e-6) / 1e9 if stage_us else None},
        "config": {"workload": "seeded ocean-uplift ensemble: sizes {64,96,128,192,256}^2 on [0,800 km]^2, "
                               "N in {4,5,7}, 4000 m ocean + 50 m shelf, open/wall boundaries, "
                               "rotated ES flux, no AV, LSRK4(5)", "seed": 0},
    }


def ensemble_av_bench(dev, args):
    """The seeded ensemble with artificial viscosity: batched with the learned indicator on the N=7
    members (the modal sensor on N=4, 5, where the network is not defined), against per-scenario
    loops of the plain method (modal sensor everywhere) on the fused (1), single-kernel (2) and
    reference (0) schedules, and against the batched plain method.  Every member runs W + K steps
    from its initial state (only the K steps timed); the L2 change of each member's final state
    against the plain loop's is reported (max over members)."""
    import numpy as np
    from esdgsem_mpiocca_amd import indicator, scenarios
    wparams, meta = indicator.load_weights()
    if wparams is None:
        raise SystemExit("no trained indicator weights (esdgsem_mpiocca_amd/indicator_weights.json)")
    # eps0 in m^2/s: the viscous step limit (DFL 0.05) stays above the CFL step on every mesh
    rp = scenarios.ocean_run_params(artificial_viscosity=1, eps0=1000.0, sigma_min=-4.5, sigma_max=-2.5, DFL=0.05)
    members = scenarios.build(scenarios.draw(args.ens_n, seed=0), rp)
    n = len(members)
    K, W = args.ens_av_steps, args.ens_warmup
    case0, rp0, _ = members[0]
    out = {}
    finals = {}

    def batched(learned):
        ctx = dev.Context(dev.make_params(case0, rp0), device=0)
        try:
            ens = dev.Ensemble(ctx, members)
            net = None
            if learned:
                net = dev.MLP(ctx, params=wparams)
                ens.set_indicator(net, args.ind_lo, args.ind_hi)
            tuned = None
            if args.ens_schedule:
                ens.set_schedule(**json.loads(args.ens_schedule))
                tuned = ens.schedule().as_dict()
            elif args.ens_autotune:
                tuned = ens.autotune(trial_steps=2)[0]
            ens.run(W)
            ctx.finish()
            t0 = time.perf_counter()
            ens.run_async(K)
            ctx.finish()
            wall = time.perf_counter() - t0
            if any(ens.status(i)[1] != W + K for i in range(n)):
                raise SystemExit("ensemble (AV): a scenario stopped early")
            counts = ens.indicator_counts() if learned else None
            finals["learned" if learned else "plain_batched"] = [ens.get_q(i) for i in range(n)]
            # per-kernel launch times of one more step, eager
            ctx.profile(True)
            ctx.profile_reset()
            ens.run(1)
            kern = {}
            for kn in ("stage", "ens_sensor", "ens_indicator", "ens_grad", "ens_viscous_face", "stage_dt", "stage_mark"):
                launches, ms = ctx.profile_query(kn)
                if launches:
                    kern[kn] = {"launches": launches, "ms_per_step": round(ms, 3)}
            ctx.profile(False)
            ens.close()
            if net:
                net.close()
        finally:
            ctx.close()
        r = {"scenario_steps_per_s": n * K / wall, "ms_per_step": 1e3 * wall / K, "schedule": tuned,
             "kernels_one_step": kern}
        if counts:
            es = max(counts["element_stages"], 1)
            r["indicator_counts"] = counts
            r["viscous_element_stage_fraction"] = counts["viscous"] / es
            r["sensor_evaluated_fraction"] = counts["not_smooth"] / es
        return r

    out["batched_learned"] = batched(True)
    print(f"# ensemble AV, learned: {out['batched_learned']['scenario_steps_per_s']:.1f} scenario-steps/s",
          file=sys.stderr, flush=True)
    out["batched_plain"] = batched(False)
    print(f"# ensemble AV, plain: {out['batched_plain']['scenario_steps_per_s']:.1f} scenario-steps/s",
          file=sys.stderr, flush=True)
    loops = {}
    for fused, name in ((1, "fused"), (2, "single_kernel"), (0, "reference_sequence")):
        solvers, tsum, qs = {}, 0.0, []
        try:
            for case, rpm, q in members:
                key = id(case)
                if key not in solvers:
                    c = dev.Context(dev.make_params(case, rpm), device=0)
                    solvers[key] = (c, dev.Solver(c, case, rpm, fused=fused))
                c, s = solvers[key]
                s.set_q(q)
                s.run(W)
                c.finish()
                t1 = time.perf_counter()
                got, _ = s.run(K)
                tsum += time.perf_counter() - t1
                if got != K:
                    raise SystemExit("ensemble AV loop: a scenario stopped early")
                if fused == 1:
                    qs.append(s.get_q())
        finally:
            for c, s in solvers.values():
                s.close()
                c.close()
        loops[name] = n * K / tsum
        if fused == 1:
            finals["plain_loop"] = qs
        print(f"# ensemble AV, per-scenario loop ({name}): {loops[name]:.1f} scenario-steps/s",
              file=sys.stderr, flush=True)

    def worst_change(a_list, b_list):
        worst = [0.0, 0.0, 0.0]
        for (case, _, _), a, b in zip(members, a_list, b_list):
            A, B = a.reshape(case.Nelem, 3, -1), b.reshape(case.Nelem, 3, -1)
            for v in range(3):
                nb = np.linalg.norm(B[:, v])
                d = np.linalg.norm(A[:, v] - B[:, v])
                worst[v] = max(worst[v], float(d / nb) if nb > 0 else float(d))
        return worst

    out["per_scenario_loop_plain"] = loops
    lv = out["batched_learned"]["scenario_steps_per_s"]
    out["speedup_learned_batched_vs_plain_loop"] = {k: lv / v for k, v in loops.items()}
    out["l2_change_learned_vs_plain_loop_max"] = worst_change(finals["learned"], finals["plain_loop"])
    out["plain_batched_equals_plain_loop"] = all(np.array_equal(a, b) for a, b in
                                                 zip(finals["plain_batched"], finals["plain_loop"]))
    out["finite"] = bool(all(np.isfinite(a).all() for a in finals["learned"]))
    out["metric"] = "ensemble scenario-steps/s (AV; learned indicator on N=7 members)"
    out["value"] = lv
    out["unit"] = "scenario-steps/s"
    out["steps"], out["warmup"] = K, W
    out["indicator"] = {"lo": args.ind_lo, "hi": args.ind_hi}
    out["config"] = {"workload": "the seeded ocean-uplift ensemble (256 scenarios, sizes {64..256}^2, N in {4,5,7}) "
                                 "with artificial viscosity (eps0 1000 m^2/s, sigma -4.5..-2.5, DFL 0.05): learned "
                                 "indicator on the N=7 members, modal sensor on the rest", "seed": 0}
    return out


def indicator_comparison(dev, args, spec, rp_kw, steps, modes):
    """Plain modal sensor against the learned indicator (modal-sensor fallback) on one case built on
    the device (esdg_device_case_create, handed to each solver device to device).  modes: "plain"
    (steps steps, timed), "learned" (timed) and "audit" (false negatives counted) to the plain
    run's end time, and optionally "control": the plain method with eps0 changed by 1e-12
    (relative), to the same time -- how far roundoff-sized differences travel in the flow, the scale
    to read the learned change against."""
    import numpy as np
    from esdgsem_mpiocca_amd import indicator
    params, meta = indicator.load_weights()
    if params is None:
        raise SystemExit("no trained indicator weights (esdgsem_mpiocca_amd/indicator_weights.json)")
    rp = indicator.av_run_params(**rp_kw)
    t0 = time.perf_counter()
    cctx = dev.Context(dev.make_params(spec, rp), device=0)
    dcase = dev.DeviceCase(cctx, spec)
    cctx.finish()
    t_case = time.perf_counter() - t0
    print(f"# {dcase.Nelem} elements, case built on the device in {t_case:.2f} s", file=sys.stderr, flush=True)
    out = {"case_build_s": round(t_case, 3)}
    qs = {}
    dofs = 3 * dcase.ngl ** 2 * dcase.Nelem
    t_end = None
    try:
        for mode in modes:
            eps0 = rp.eps0 * (1.0 + 1e-12) if mode == "control" else rp.eps0
            rpm = rp if mode == "plain" else indicator.av_run_params(**{**rp_kw, "T": t_end, "eps0": eps0})
            ctx = dev.Context(dev.make_params(spec, rpm), device=0)
            try:
                solver = dev.Solver(ctx, dcase, rpm, fused=1)
                net = None
                if mode in ("learned", "audit"):
                    net = dev.MLP(ctx, params=params)
                    solver.set_indicator(net, args.ind_lo, args.ind_hi)
                    solver.indicator_audit(mode == "audit")
                cap = steps if mode == "plain" else 2 * steps
                done, wall, t_sim = 0, 0.0, 0.0
                # the chunk timed with per-launch events (which make every step issue eagerly) is
                # left out of the throughput
                timed_steps, timed_wall = 0, 0.0
                # chunks of up to 50 steps, at least four per run, so that leaving the evented one
                # out still times three quarters of the steps
                chunk = max(1, min(50, steps // 4))
                per_kernel = {}
                while done < cap:
                    k = min(chunk, cap - done)
                    # per-kernel times from the second chunk (the first holds the run's first-step passes)
                    prof = mode in ("plain", "learned") and not per_kernel and (done > 0 or cap <= chunk)
                    if prof:
                        ctx.profile(True)
                        ctx.profile_reset()
                    t1 = time.perf_counter()
                    n, t_sim = solver.run(k)
                    ctx.finish()
                    dw = time.perf_counter() - t1
                    wall += dw
                    done += n
                    if not prof:
                        timed_steps += n
                        timed_wall += dw
                    if prof:
                        for kn in KERNELS:
                            launches, ms = ctx.profile_query(kn)
                            if launches:
                                per_kernel[kn] = {"avg_us": round(1e3 * ms / launches, 2),
                                                  "ms_per_step": round(ms / max(n, 1), 4)}
                        ctx.profile(False)
                    print(f"# {mode}: {done} steps, t = {t_sim:.6g}, {1e3 * wall / max(done, 1):.1f} ms/step",
                          file=sys.stderr, flush=True)
                    if n != k:
                        if mode != "plain" and t_sim >= t_end:
                            break
                        raise SystemExit(f"{mode}: solver stopped after {done} steps")
                if mode == "plain":
                    t_end = t_sim
                counts = solver.indicator_counts() if net else None
                if mode != "audit":
                    qs[mode] = solver.get_q()
                if mode in ("plain", "learned"):
                    if timed_steps == 0:   # a single chunk: it is the timed one
                        timed_steps, timed_wall = done, wall
                    out[mode] = {"dof_stage_updates_per_s": dofs * 5 * timed_steps / timed_wall,
                                 "ms_per_step": 1e3 * timed_wall / timed_steps, "steps": done,
                                 "timed_steps": timed_steps, "sim_time": t_sim, "kernels": per_kernel}
                    if counts:
                        es = max(counts["element_stages"], 1)
                        out[mode]["viscous_element_stage_fraction"] = counts["viscous"] / es
                        out[mode]["sensor_evaluated_fraction"] = counts["not_smooth"] / es
                elif mode == "audit":
                    out["audit"] = {"steps": done, **counts,
                                    "false_negative_rate": counts["false_negatives"] / max(counts["sensor_positives"], 1)}
                solver.close()
                if net:
                    net.close()
            finally:
                ctx.close()
    finally:
        dcase.close()
        cctx.close()

    def change(a, b):
        A, B = a.reshape(dcase.Nelem, 3, -1), b.reshape(dcase.Nelem, 3, -1)
        return [float(np.linalg.norm(A[:, v] - B[:, v]) / max(np.linalg.norm(B[:, v]), 1e-300)) for v in range(3)]

    out["l2_change_vs_plain"] = change(qs["learned"], qs["plain"])
    if "control" in qs:
        out["control_l2_change_eps0_1e-12"] = change(qs["control"], qs["plain"])
    A = qs["learned"].reshape(dcase.Nelem, 3, -1)
    B = qs["plain"].reshape(dcase.Nelem, 3, -1)
    out["finite"] = bool(np.isfinite(A).all() and np.isfinite(B).all())
    out["min_h"] = float(min(A[:, 0].min(), B[:, 0].min()))
    out["speedup_learned_vs_plain"] = out["learned"]["dof_stage_updates_per_s"] / out["plain"]["dof_stage_updates_per_s"]
    out["indicator"] = {"lo": args.ind_lo, "hi": args.ind_hi, "validation": meta.get("validation")}
    out["_dcase"] = (dcase.Nelem, dcase.ngl, dcase.N, dofs)
    return out


def c4_bench(dev, args, steps):
    """Wet/dry flood on the curved mesh family (2048^2, N=7), positivity limiter, artificial
    viscosity: the modal sensor everywhere against the learned indicator, with the rooflines of the
    plain run's two dominant kernels."""
    from esdgsem_mpiocca_amd import Case
    spec = Case.flood(N=7, nel=args.c4_nel, seed=0, build=False)
    # CFL 0.1: the dt formula's CFL/ngl stays below the smallest GLL weight fraction 1/(N(N+1)),
    # the step size under which the limiter keeps the element means positive.  Dry tolerance 0.1:
    # without the partial-dry treatment, thinner layers carry spurious velocities at the fronts
    # (tolerances 1e-2 and 5e-2 reach NaN or |u| of hundreds of m/s within 1000 steps at 2048^2).
    out = indicator_comparison(dev, args, spec, {"positivity": 1, "pos_pres_tol": 0.1, "CFL": 0.1}, steps,
                               ("plain", "learned", "audit", "control"))
    nelem, ngl, N, dofs = out.pop("_dcase")
    # roofline of the plain run's dominant kernel (HIP events around every launch of the profiled
    # chunk): the fused volume pass, which limits the new state before it stores it and writes the
    # side traces, the sensor and the wave speed of the limited state
    pk = out["plain"]["kernels"]
    roof = {}
    for name, alg in (("fused_volume", fused_volume_bytes(nelem, ngl, 1, traces=True, sensor=True, limiter=True)),):
        if name in pk:
            gbs = alg / (pk[name]["avg_us"] * 1e-6) / 1e9
            roof[name] = {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                          "frac": round(gbs / HBM_PEAK_GBS, 4), "algorithmic_bytes_per_launch": alg,
                          "avg_launch_us": pk[name]["avg_us"]}
    if "fused_volume" in roof:
        tfl = volume_flops(nelem, ngl) / (pk["fused_volume"]["avg_us"] * 1e-6) / 1e12
        roof["fused_volume"]["fp64_tflops_algorithmic"] = round(tfl, 2)
        roof["fused_volume"]["fp64_frac_of_peak"] = round(tfl / FP64_PEAK_TFLOPS, 4)
    out["roofline"] = roof
    out["metric"] = "FP64 DOF-stage updates/s at N=7 (curved wet/dry flood, limiter, AV)"
    out["value"] = out["plain"]["dof_stage_updates_per_s"]
    out["unit"] = "DOF-stage updates/s"
    out["config"] = {"workload": f"curved flood {args.c4_nel}x{args.c4_nel}, N=7, 16-mode Fourier bathymetry "
                                 "(amplitude 0.3), reservoir over dry land, positivity limiter, AV, rotated ES, "
                                 "LSRK4(5); learned indicator with the modal sensor as fallback beside it",
                     "Nelem": nelem, "N": N, "dofs": dofs}
    return out


def c2_bench(dev, args, steps):
    """The learned indicator evaluated on the 512^2 curved dam break at N=7 (the headline mesh)
    against the plain-sensor FP64 run: false-negative rate over the run and the L2 change of the
    state at equal simulated time (a flow without wet/dry fronts, where differences do not blow
    up), with its throughput beside the plain method's."""
    from esdgsem_mpiocca_amd import Case
    spec = Case.dam_break(N=7, nel=args.nel, noise=1e-3, seed=0, build=False)
    out = indicator_comparison(dev, args, spec, {}, steps, ("plain", "learned", "audit"))
    nelem, ngl, N, dofs = out.pop("_dcase")
    out["config"] = {"workload": f"curved dam break {args.nel}x{args.nel}, N=7, modal sensor + AV against the "
                                 "learned indicator (modal sensor as fallback)", "Nelem": nelem, "N": N, "steps": steps}
    return out


def load_profile_json(fname):
    path = os.path.join(ROOT, "profiles", fname)
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f)


def load_traffic(name):
    return load_profile_json(f"{name}_traffic.json")


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
    ap.add_argument("--cpu-sample-nel", type=int, default=512)
    ap.add_argument("--cpu-sample-steps", type=int, default=25)
    ap.add_argument("--no-profile", action="store_true", help="time without per-launch events")
    ap.add_argument("--c4-nel", type=int, default=2048)
    ap.add_argument("--c4-steps", type=int, default=None,
                    help="steps of the 2048^2 flood's plain run (default: 1000 with --workload c4, 100 beside the "
                         "headline line)")
    ap.add_argument("--c4", type=int, default=1, help="also run the 2048^2 flood beside the headline line")
    ap.add_argument("--c2", type=int, default=1,
                    help="also evaluate the learned indicator on the headline dam break (plain vs learned)")
    ap.add_argument("--c2-steps", type=int, default=200)

    # the learned indicator's operating point: p <= lo smooth (no viscosity), p >= hi troubled.  The
    # sensor is evaluated for every element either way, so lo trades false negatives only: at 0.002
    # the 512^2 dam break stays within 1e-3 relative L2 of the plain run
    ap.add_argument("--ind-lo", type=float, default=0.002)
    ap.add_argument("--ind-hi", type=float, default=0.5)
    ap.add_argument("--workload", choices=["c1", "ensemble", "ensemble_av", "c4"], default="c1",
                    help="c1: the curved dam break (headline line, plus the ensemble unless --ensemble 0); "
                         "ensemble: the ensemble only")
    ap.add_argument("--ensemble", type=int, default=1, help="also run the ensemble benchmark")
    ap.add_argument("--ens-n", type=int, default=256)
    ap.add_argument("--ens-steps", type=int, default=10)
    ap.add_argument("--ens-warmup", type=int, default=2)
    ap.add_argument("--ens-autotune", type=int, default=1)
    ap.add_argument("--ens-schedule", default="",
                    help='a fixed ensemble schedule instead of the autotuned one, e.g. \'{"order": 1, "tile": 2}\'')
    ap.add_argument("--ens-loops", type=int, default=1, help="also time the per-scenario loops (ensemble)")
    ap.add_argument("--ens-av", type=int, default=1,
                    help="also run the ensemble with AV: learned indicator batched vs the plain per-scenario loops")
    ap.add_argument("--ens-av-steps", type=int, default=10)
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        if int(os.environ.get("RANK", "0")) == 0:
            print("bench.py measures one GPU: this build has no multi-GPU path (the halo exchange "
                  "is out of scope); run with --gpus 1", file=sys.stderr, flush=True)
        sys.exit(2)

    import esdgsem_mpiocca_amd as E
    dev = E.device()

    if args.workload in ("ensemble", "ensemble_av", "c4"):
        if args.workload == "ensemble":
            out = ensemble_bench(dev, args)
        elif args.workload == "ensemble_av":
            out = ensemble_av_bench(dev, args)
        else:
            out = c4_bench(dev, args, args.c4_steps or 1000)
        out.update({"n_gpus": 1, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
                    "dtype": "f64", "data": "synthetic"})
        print(json.dumps(out), flush=True)
        return

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
    # the timed steps run as a caller runs them: from the second step of the call on, the fused
    # schedule replays a captured step graph
    ctx.profile(False)
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

    # the kernel times: the same number of steps again, with HIP events around every launch on
    # the solver's streams (per-launch events make every step issue eagerly, so this run is timed
    # on its own and does not enter `value`)
    per_kernel = {}
    wall_prof = None
    if not args.no_profile:
        ctx.profile(True)
        ctx.profile_reset()
        t0 = time.perf_counter()
        n_prof, _ = solver.run(args.steps)
        ctx.finish()
        wall_prof = time.perf_counter() - t0
        if n_prof != args.steps:
            raise SystemExit(f"solver stopped after {n_prof} of {args.steps} profiled steps")
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
        # this workload's viscosity is the modal sensor alone, which the fused schedule evaluates
        # inside the state update
        alg = (fused_volume_bytes(case.Nelem, case.ngl, rp.artificial_viscosity, not rp.positivity,
                                  sensor=bool(rp.artificial_viscosity), stages=5) if args.fused
               else volume_bytes(case.Nelem, case.ngl))
        gbs = alg / avg_s / 1e9
        tfl = volume_flops(case.Nelem, case.ngl) / avg_s / 1e12
        tr = load_traffic(vol_name)
        roofline = {"kernel": vol_name, "bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS,
                    "unit": "GB/s", "frac": round(gbs / HBM_PEAK_GBS, 4),
                    "traffic": tr.get("hbm_bytes_per_launch") if tr else None,
                    "algorithmic_bytes_per_launch": alg,
                    "avg_launch_us": round(per_kernel[vol_name]["avg_us"], 2),
                    "fp64_tflops_algorithmic": round(tfl, 2),
                    "fp64_frac_of_peak": round(tfl / FP64_PEAK_TFLOPS, 4)}
        # executed FP64 operations (PMC counts of the same kernel on the same workload), when
        # measured for this configuration
        fl = load_profile_json(f"{vol_name}_flops.json")
        if fl and args.fused and case.Nelem == 262144 and case.ngl == 8 and rp.artificial_viscosity:
            ex = fl["executed_fp64_flops_per_launch"] / avg_s / 1e12
            roofline["fp64_tflops_executed"] = round(ex, 2)
            roofline["fp64_executed_frac_of_peak"] = round(ex / FP64_PEAK_TFLOPS, 4)

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
        "ms_per_step_with_launch_events": 1e3 * wall_prof / args.steps if wall_prof else None,
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
    if args.ensemble:
        try:
            out["ensemble"] = ensemble_bench(dev, args)
        except Exception as e:  # reported beside the headline figure; must not hide it
            out["ensemble"] = {"error": str(e)}
    if args.ens_av:
        try:
            out["ensemble_av"] = ensemble_av_bench(dev, args)
        except (Exception, SystemExit) as e:  # reported beside the headline figure; must not hide it
            out["ensemble_av"] = {"error": str(e)}
    if args.c2:
        try:
            out["c2"] = c2_bench(dev, args, args.c2_steps)
        except (Exception, SystemExit) as e:  # reported beside the headline figure; must not hide it
            out["c2"] = {"error": str(e)}
    if args.c4:
        try:
            out["c4"] = c4_bench(dev, args, args.c4_steps or 100)
        except (Exception, SystemExit) as e:  # reported beside the headline figure; must not hide it
            out["c4"] = {"error": str(e)}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

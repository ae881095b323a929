"""Benchmark: converged Mn closures/s and s/timestep on the M4 lattice.

Workload (the north-star lattice case): M4 closure (15 moments), Q = 16
(256 hemisphere quadrature points), 280 x 280 cells on [-3.5, 3.5]^2 with the
7 x 7 block checkerboard of absorbers (sigma_a = 10) and a unit isotropic
source in the centre block, floor 1e-10, MOMOPT_TOL 1e-8, direct Hessian, the
robust line search (at TOL 1e-8 the reference's Armijo test cannot resolve the
last Newton decrease of near-isotropic floor cells and the run aborts with
"Max regularization" by step 34; both the GPU and the CPU leg run the same
robust rule),
Heun steps of dt = CFL 0.9 * 2/(2+theta) * dx dy / (dx + dy) up to t = 3.2
(569 steps).  Every step solves the closure twice in every cell including the
2 ghost layers: 2 * 284^2 = 161,312 closures per step.

A "step" is one Heun time step of the whole lattice.  W warm-up steps run first
(then the state is reset to t = 0); the timed region is the first K steps from
t = 0 (K = 569 is the whole run), with all state resident in HBM.

Prints one JSON line.  Single GPU only.
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "closures-2d_amd")

LATTICE = dict(num_cells_x=280, num_cells_y=280, a_x=-3.5, b_x=3.5, a_y=-3.5, b_y=3.5,
               gaussian_sigma=0.0, floor_value=1e-10, sigma=1.0, init_cond=1, moment_order=4,
               quad_order=16, cfl_factor=0.9, tol=1e-8, cond_h_max=1e10, max_iter=100,
               use_gaunt=0, theta=2.0, line_search=1)
T_FINAL = 3.2
FP64_PEAK_TFLOPS = 78.6     # MI355X dense FP64 (vector and matrix)


def _load(name):
    import importlib.util
    spec = importlib.util.spec_from_file_location("closures2d_amd_" + name, os.path.join(PKG, name + ".py"))
    mod = importlib.util.module_from_spec(spec)
    sys.modules[spec.name] = mod
    spec.loader.exec_module(mod)
    return mod


def flops_per_iter(n, nq):
    """Algorithmic FP64 flops of one Newton iteration (full f, g, H) and of one
    Armijo trial (f only): alpha.m (2n nq), lower-triangle M^T D M (n(n+1) nq),
    w e and f (2 nq), Cholesky (n^3/3), solve (2n^2), condition estimate (8n^2)."""
    f_iter = 2 * n * nq + n * (n + 1) * nq + 2 * nq + n ** 3 / 3.0 + 2 * n * n + 8 * n * n
    f_trial = 2 * n * nq + 2 * nq
    return f_iter, f_trial


def make_grid(mnc, gi, d):
    ctx = mnc.Context(d["moment_order"], d["quad_order"], use_gaunt=d["use_gaunt"])
    g = gi.initial_state(d["num_cells_x"], d["num_cells_y"], d["a_x"], d["b_x"], d["a_y"], d["b_y"],
                         d["init_cond"], d["gaussian_sigma"], d["floor_value"], d["sigma"], ctx.nm,
                         d["cfl_factor"], d["theta"])
    ctx.grid_create(g.nx, g.ny, g.dx, g.dy, g.init_x, g.init_y, d["init_cond"], d["theta"],
                    g.moments, g.alpha, g.sigma_t, g.sigma_s)
    return ctx, g


def run_form(mnc, gi, d, opt, steps, warmup):
    ctx, g = make_grid(mnc, gi, d)
    dts = gi.time_steps(g.max_dt, T_FINAL)
    for k in range(min(warmup, len(dts))):
        ctx.update(dts[k], opt)
    ctx.sync()
    ctx.grid_set(g.moments, g.alpha)        # back to t = 0
    ctx.reset_stats()
    ctx.timing_enable(True)
    ctx.sync()
    t0 = time.perf_counter()
    for k in range(steps):
        ctx.update(dts[k % len(dts)], opt)
    ctx.sync()
    wall = time.perf_counter() - t0
    st = ctx.stats()
    kn, kms = ctx.timing_get(mnc.KCLASS_NEWTON)
    fn, fms = ctx.timing_get(mnc.KCLASS_FLUX)
    on, oms = ctx.timing_get(mnc.KCLASS_OTHER)
    ctx.timing_enable(False)
    m, a = ctx.grid_get()
    ctx.close()
    closures = 2 * (g.nx + 4) * (g.ny + 4) * steps
    assert st.num_dual_solves == closures, (st.num_dual_solves, closures)
    return dict(wall=wall, stats=st, newton=(kn, kms), flux=(fn, fms), other=(on, oms), closures=closures,
                moments=m, dts=dts[:steps])


def cpu_baseline(d, seconds):
    """The CPU oracle (FP64 C, OpenMP schedule(dynamic,1) over cells, as the
    reference's solveOptimization) on the first steps of the same run, for
    about `seconds` of CPU time."""
    sys.path.insert(0, ROOT)
    import oracle
    nthreads = len(os.sched_getaffinity(0))
    s = oracle.Solver(d, nthreads=nthreads)
    import importlib
    gi = _load("grid_init")
    dts = gi.time_steps(s.max_dt, T_FINAL)
    t0 = time.perf_counter()
    k = 0
    while k < len(dts) and (time.perf_counter() - t0) < seconds:
        s.update(dts[k])
        k += 1
    wall = time.perf_counter() - t0
    st = s.stats()
    return dict(value=st.num_dual_solves / wall, steps=k, wall=wall, cores=nthreads,
                iters=st.iter_sum / max(st.num_dual_solves, 1))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=569)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build's hot path runs on a single MI355X "
              "(multi-GPU sharding is a design note only); run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    mnc, gi = _load("mnc"), _load("grid_init")
    d = dict(LATTICE)
    n = (d["moment_order"] + 1) * (d["moment_order"] + 2) // 2
    nq = d["quad_order"] ** 2
    plain = run_form(mnc, gi, d, mnc.options(max_iter=d["max_iter"], tol=d["tol"], cond_h_max=d["cond_h_max"],
                                             line_search=mnc.LS_ROBUST),
                     args.steps, args.warmup)
    best = plain
    form = "plain-fp64"
    st = best["stats"]
    value = best["closures"] / best["wall"]
    f_iter, f_trial = flops_per_iter(n, nq)
    kn, kms = best["newton"]
    alg_flops = st.iter_sum * f_iter + st.trial_sum * f_trial
    achieved = alg_flops / (kms * 1e-3) / 1e12 if kms > 0 else 0.0
    out = {
        "metric": "converged Mn closures/s (M4 lattice 280^2, t=0..3.2, TOL 1e-8)",
        "value": value,
        "unit": "closures/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * best["wall"] / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (the reference's lattice initial condition, generated in-process)",
        "config": {"workload": "M4 lattice 280x280, Q=16 (nq=256), 7x7 checkerboard on [-3.5,3.5]^2, "
                               "floor 1e-10, TOL 1e-8, direct Hessian, robust line search, Heun dt=%.6g" % plain["dts"][0],
                   "form": form, "closures_per_step": 2 * 284 * 284,
                   "s_per_timestep": best["wall"] / args.steps,
                   "plain_fp64_closures_per_s": plain["closures"] / plain["wall"],
                   "newton_iters_per_closure": st.iter_sum / st.num_dual_solves,
                   "trials_per_closure": st.trial_sum / st.num_dual_solves,
                   "max_iter": st.max_iter, "reg_hist": list(st.hist_reg),
                   "kernel_ms": {"newton": kms, "flux": best["flux"][1], "other": best["other"][1]},
                   "kernel_launches": {"newton": kn, "flux": best["flux"][0], "other": best["other"][0]}},
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / FP64_PEAK_TFLOPS, "traffic": None,
                     "kernel": "Newton solve (init + sweep/logic rounds + tail), M4 direct FP64",
                     "algorithmic_flops_per_launch": alg_flops / max(kn, 1),
                     "avg_launch_ms": kms / max(kn, 1)},
    }
    if not args.no_cpu:
        cb = cpu_baseline(d, args.cpu_seconds)
        out["cpu_baseline"] = {"value": cb["value"], "unit": "closures/s", "cores": cb["cores"], "kind": "port",
                               "sample": "oracle (FP64 C restatement, OpenMP dynamic,1) on the first %d steps "
                                         "of the same lattice run from t=0 (%.1f s, %.2f Newton iterations per "
                                         "closure)" % (cb["steps"], cb["wall"], cb["iters"])}
    print(json.dumps(out))


if __name__ == "__main__":
    main()

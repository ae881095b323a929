"""Benchmark: converged Mn closures/s and s/timestep on the M4 lattice.

Workload (the lattice case): M4 closure (15 moments), Q = 16
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

Three forms run on the same steps:
  plain-fp64    FP64 batched Newton from the previous step's alpha (the reference's method)
  mixed         mixed-precision lockstep rounds (MFMA Hessian), FP64 tail
  learned+mixed the same, started from the best of {MLP guess, isotropic alpha,
                previous alpha} by FP64 gradient norm; the MLP is first trained on
                GPU with the library's own kernels on (u, alpha) pairs harvested from
                plain-form lattice steps and an M4 Gaussian-pulse run (not timed,
                reported in config.mlp)
`value` is the learned+mixed form's rate (the metric's headline form); the
other forms are reported beside it.  config.isolated (the 10^6-vector M4
batch), config.m8 (M8 512^2) and config.ppn (the PPn closure on the same
lattice) are reported with their own rates.

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


def flops_per_iter_ppn(n, nq):
    """The same count for PPn: alpha.m (2n nq), H sums (n(n+1) nq), g sums
    (2n nq), the ansatz terms per point (~12 nq: G, its square root, log and
    the weights), Cholesky, solve and condition estimate; a trial is alpha.m
    and the f terms."""
    f_iter = 2 * n * nq + n * (n + 1) * nq + 2 * n * nq + 12 * nq + n ** 3 / 3.0 + 2 * n * n + 8 * n * n
    f_trial = 2 * n * nq + 8 * nq
    return f_iter, f_trial


def flops_per_iter(n, nq):
    """Algorithmic FP64 flops of one Newton iteration (full f, g, H) and of one
    Armijo trial (f only): alpha.m (2n nq), lower-triangle M^T D M (n(n+1) nq),
    w e and f (2 nq), Cholesky (n^3/3), solve (2n^2), condition estimate (8n^2)."""
    f_iter = 2 * n * nq + n * (n + 1) * nq + 2 * nq + n ** 3 / 3.0 + 2 * n * n + 8 * n * n
    f_trial = 2 * n * nq + 2 * nq
    return f_iter, f_trial


def make_grid(mnc, gi, d):
    mt = d.get("moment_type", 0)
    ctx = mnc.Context(d["moment_order"], d["quad_order"], use_gaunt=d["use_gaunt"], moment_type=mt)
    g = gi.initial_state(d["num_cells_x"], d["num_cells_y"], d["a_x"], d["b_x"], d["a_y"], d["b_y"],
                         d["init_cond"], d["gaussian_sigma"], d["floor_value"], d["sigma"], ctx.nm,
                         d["cfl_factor"], d["theta"], moment_type=mt, delta_ppn=d.get("delta_ppn", 1e-10))
    ctx.grid_create(g.nx, g.ny, g.dx, g.dy, g.init_x, g.init_y, d["init_cond"], d["theta"],
                    g.moments, g.alpha, g.sigma_t, g.sigma_s, delta_ppn=d.get("delta_ppn", 0.0))
    return ctx, g


def harvest(mnc, gi, d, steps, every):
    """(u/u0, scaled alpha) pairs of every cell after every `every`-th plain step."""
    ctx, g = make_grid(mnc, gi, d)
    dts = gi.time_steps(g.max_dt, T_FINAL)
    opt = mnc.options(max_iter=d["max_iter"], tol=d["tol"], cond_h_max=d["cond_h_max"], line_search=mnc.LS_ROBUST)
    us, al = [], []
    for k in range(steps):
        ctx.update(dts[k % len(dts)], opt)
        if (k + 1) % every == 0:
            m, a = ctx.grid_get()
            u1, a1 = mnc.scaled_pairs(m.reshape(-1, ctx.nm), a.reshape(-1, ctx.nm))
            us.append(u1)
            al.append(a1)
    ctx.close()
    return np.concatenate(us), np.concatenate(al)


def train_mlp(mnc, gi, d, pairs_target=2_000_000, train_steps=2000, seed=1):
    """Trains the warm-start network on pairs from plain runs: the lattice's
    first steps and an M4 Gaussian pulse (50 x 50 on [-1,1]^2, sigma 0.1, floor
    1e-4).  Returns (params, info)."""
    t0 = time.perf_counter()
    u1, a1 = harvest(mnc, gi, d, 48, 2)
    pulse = dict(d, num_cells_x=50, num_cells_y=50, a_x=-1.0, b_x=1.0, a_y=-1.0, b_y=1.0, gaussian_sigma=0.1,
                 floor_value=1e-4, init_cond=0)
    u2, a2 = harvest(mnc, gi, pulse, 100, 1)
    us, al = np.concatenate([u1, u2]), np.concatenate([a1, a2])
    # Most lattice cells are exactly isotropic (the floor), and a few pairs
    # near vacuum/absorber interfaces have |alpha_k| in the hundreds: keep the
    # anisotropic pairs with |alpha_k| <= 50 (k >= 1) and 1% of the isotropic
    # ones, so the standardised loss is not dominated by either.
    rng = np.random.default_rng(seed + 1)
    aniso = np.abs(us[:, 1:]).max(axis=1) > 1e-6
    tame = np.abs(al[:, 1:]).max(axis=1) <= 50.0
    keep = (aniso & tame) | (~aniso & (rng.random(len(us)) < 0.01))
    us, al = us[keep], al[keep]
    if len(us) > pairs_target:
        idx = np.random.default_rng(seed).choice(len(us), pairs_target, replace=False)
        us, al = us[idx], al[idx]
    t1 = time.perf_counter()
    progress("harvested %d pairs" % len(us))
    ctx = mnc.Context(d["moment_order"], d["quad_order"], use_gaunt=d["use_gaunt"])
    ctx.mlp_init(seed)
    ctx.timing_enable(True)
    loss = ctx.mlp_train(us, al, steps=train_steps, batch=4096, lr=1e-3, seed=seed)
    ctx.sync()
    n_mlp, ms_mlp = ctx.timing_get(mnc.KCLASS_MLP)
    params = ctx.mlp_params()
    ctx.close()
    info = {"pairs": int(len(us)), "harvest_s": t1 - t0, "train_steps": train_steps, "batch": 4096, "lr": 1e-3,
            "train_gpu_ms": ms_mlp, "loss_first": float(loss[:10].mean()), "loss_last": float(loss[-50:].mean())}
    return params, info


def stage_traffic(form):
    """HBM bytes of one Newton stage of `form` from the committed counter
    summary (profiles/newton_stage_traffic.md), or None."""
    try:
        for line in open(os.path.join(ROOT, "profiles", "newton_stage_traffic.md")):
            w = line.split()
            if len(w) == 3 and w[0] == "traffic_bytes_per_stage" and w[1] == form:
                return float(w[2])
    except OSError:
        pass
    return None


def progress(msg):
    print("[bench %.1fs] %s" % (time.perf_counter() - T0, msg), file=sys.stderr, flush=True)


T0 = time.perf_counter()


FORCE_TUNE = {}
ISOLATED_ROUNDS = (6, 12, 48, 96, 192)   # lockstep rounds tried on the isolated batch
TUNE_STEPS = 10           # steps per autotuning trial


def run_form(mnc, gi, d, opt, steps, warmup, mlp_params=None, kernel_steps=None, autotune=True, tune_step=None,
             full_run=True):
    """W warm-up steps, then (untimed) on to step `tune_step` (default: the
    middle of the run) and the library's autotuner on that state, then (from
    t = 0 again) K steps timed by wall clock with the tuned launch shape
    (hipGraph replay if the tuner picked it; per-launch timing off), then the
    same K steps (or the first kernel_steps of them) again from t = 0, eagerly,
    with HIP-event timing of every launch for the per-kernel-class split and
    the roofline; then (full_run) the whole run to t = 3.2 from t = 0 again,
    wall-timed with the tuned settings."""
    ctx, g = make_grid(mnc, gi, d)
    if mlp_params is not None:
        ctx.mlp_set_params(mlp_params)
    dts = gi.time_steps(g.max_dt, T_FINAL)
    if tune_step is None:
        tune_step = len(dts) // 2
    for k in range(min(max(warmup, tune_step if autotune else 0), len(dts))):
        ctx.update(dts[k], opt)
    ctx.sync()
    report = ctx.autotune(dts[0], opt, steps=TUNE_STEPS) if autotune else []
    progress("autotuned: %s" % "; ".join(report[-3:]))
    for k, v in FORCE_TUNE.items():
        ctx.tune(k, v)
    tuned = {k: ctx.get_tune(k) for k in ("graph", "rounds", "tail_mode", "tail_blocks", "mix_blocks", "flux_tile",
                                          "flux_groups", "flux_overlap",
                                          "mix_cond_log10", "mix_defer", "mix_split", "learned_gate_log10")}
    ctx.grid_set(g.moments, g.alpha)        # back to t = 0
    ctx.reset_stats()
    ctx.sync()
    t0 = time.perf_counter()
    for k in range(steps):
        ctx.update(dts[k % len(dts)], opt)
    ctx.sync()
    wall = time.perf_counter() - t0
    st = ctx.stats()
    m, a = ctx.grid_get()
    progress("timed %d steps in %.2f s" % (steps, wall))
    # eager pass with per-launch timing
    ks = steps if kernel_steps is None else min(steps, kernel_steps)
    ctx.grid_set(g.moments, g.alpha)
    ctx.reset_stats()
    ctx.timing_enable(True)
    for k in range(ks):
        ctx.update(dts[k % len(dts)], opt)
    ctx.sync()
    kst = ctx.stats()
    kn, kms = ctx.timing_get(mnc.KCLASS_NEWTON)
    tn, tms = ctx.timing_get(mnc.KCLASS_TAIL)
    fn, fms = ctx.timing_get(mnc.KCLASS_FLUX)
    on, oms = ctx.timing_get(mnc.KCLASS_OTHER)
    mn, mms = ctx.timing_get(mnc.KCLASS_MLP)
    ctx.timing_enable(False)
    full = None
    if full_run:
        # the whole run to t = 3.2 from t = 0 with the same tuned settings
        ctx.grid_set(g.moments, g.alpha)
        ctx.reset_stats()
        ctx.sync()
        t0 = time.perf_counter()
        for dt in dts:
            ctx.update(dt, opt)
        ctx.sync()
        fwall = time.perf_counter() - t0
        fst = ctx.stats()
        fcl = 2 * (g.nx + 4) * (g.ny + 4) * len(dts)
        assert fst.num_dual_solves == fcl, (fst.num_dual_solves, fcl)
        full = {"steps": len(dts), "t_final": T_FINAL, "wall_s": fwall, "closures": fcl,
                "closures_per_s": fcl / fwall, "s_per_timestep": fwall / len(dts),
                "newton_iters_per_closure": fst.iter_sum / fst.num_dual_solves,
                "trials_per_closure": fst.trial_sum / fst.num_dual_solves, "max_iter": fst.max_iter,
                "reg_hist": list(fst.hist_reg), "warm_pick_mlp_iso_prev": list(fst.warm_pick),
                "tail_cells_per_stage": fst.tail_cells / (2.0 * len(dts)),
                "tail_iters_per_tail_cell": fst.tail_iter_sum / max(fst.tail_cells, 1)}
        progress("full run: %d steps in %.2f s" % (len(dts), fwall))
    ctx.close()
    closures = 2 * (g.nx + 4) * (g.ny + 4) * steps
    assert st.num_dual_solves == closures, (st.num_dual_solves, closures)
    return dict(wall=wall, stats=st, kstats=kst, kernel_steps=ks, newton=(kn, kms), tail=(tn, tms), flux=(fn, fms),
                other=(on, oms), mlp=(mn, mms), closures=closures, moments=m, dts=dts[:steps], tuned=tuned,
                tune_report=report, full_run=full)


def form_summary(r, f_iter, f_trial):
    st, kst = r["stats"], r["kstats"]
    kms = r["newton"][1] + r["tail"][1]
    stages = 2 * r["kernel_steps"]
    return {"closures_per_s": r["closures"] / r["wall"], "s_per_timestep": r["wall"] / len(r["dts"]),
            "timed_steps": len(r["dts"]), "full_run": r["full_run"],
            "newton_iters_per_closure": st.iter_sum / st.num_dual_solves,
            "trials_per_closure": st.trial_sum / st.num_dual_solves, "max_iter": st.max_iter,
            "reg_hist": list(st.hist_reg), "warm_pick_mlp_iso_prev": list(st.warm_pick),
            "mixed_steps_redone_fp64": st.mixed_deferred,
            "tuned": r["tuned"],
            "autotune_trials": r["tune_report"],
            "kernel_timed_steps": r["kernel_steps"],
            "kernel_ms": {"newton_rounds": r["newton"][1], "tail": r["tail"][1], "flux": r["flux"][1],
                          "other": r["other"][1], "mlp": r["mlp"][1]},
            "tail_cells_per_stage": kst.tail_cells / max(stages, 1),
            "tail_iters_per_tail_cell": kst.tail_iter_sum / max(kst.tail_cells, 1),
            "tail_trials_per_tail_cell": kst.tail_trial_sum / max(kst.tail_cells, 1),
            "newton_tflops_fp64_equiv": (kst.iter_sum * f_iter + kst.trial_sum * f_trial) / (kms * 1e-3) / 1e12
            if kms > 0 else 0.0}


def bench_isolated(mnc, n, seed, oracle_sample, train_steps):
    """The isolated closure batch (n M4 vectors, Q = 16, 20% near-boundary
    lobes, isotropic start, TOL 1e-8, reference line search): converged
    closures/s of each form with u and alpha resident in HBM (mnc_batch_set
    before the clock starts), and each form's error against the FP64 CPU oracle
    on the first `oracle_sample` vectors.  Two variants beside it: the same
    batch with the robust line search (MNC_LS_ROBUST, the lattice decks'), and
    a smooth batch (no lobes) with the robust line search, where the network
    can learn every input."""
    sys.path.insert(0, ROOT)
    import oracle
    wl = _load("workloads")
    ctx = mnc.Context(4, 16, use_gaunt=False)
    w, p, xi, eta = ctx.tables()
    t0 = time.perf_counter()
    u, astar, lobe = wl.isolated_batch(w, p, xi, eta, 15, n, seed)
    a0 = wl.isotropic_start(u)
    gen_s = time.perf_counter() - t0
    progress("isolated batch generated (%.1f s)" % gen_s)
    tab = oracle.Tables(4, 16, False)
    m = oracle_sample
    # learned form: network trained on separate exact (u, alpha*) pairs of the same distribution
    ut, at, lt = wl.isolated_batch(w, p, xi, eta, 15, 2_000_000, seed + 1, near_boundary_frac=0.0)
    ust, ast_ = mnc.scaled_pairs(ut, at)
    ctx.mlp_init(seed)
    loss = ctx.mlp_train(ust, ast_, steps=train_steps, batch=4096, lr=1e-3, seed=seed)
    progress("isolated-batch network trained")

    def solve_forms(u, a0, ls):
        if m > 0:
            sto, uo, ao, ito, ro, _ = tab.opt_batch(u[:m], a0[:m], max_iter=100, tol=1e-8, cond_h_max=1e10,
                                                    nthreads=len(os.sched_getaffinity(0)), line_search=ls)
        res = {}
        for name, kw in (("plain-fp64", {}), ("mixed", {"precision": mnc.PREC_MIXED}),
                         ("learned+mixed", {"precision": mnc.PREC_MIXED, "warm_start": mnc.WARM_LEARNED})):
            opt = mnc.options(max_iter=100, tol=1e-8, cond_h_max=1e10, line_search=ls, **kw)

            def timed_solve():
                ctx.batch_set(u, a0)                   # inputs resident in HBM before the clock starts
                ctx.reset_stats()
                ctx.sync()
                t1 = time.perf_counter()
                ctx.batch_solve(opt)
                return time.perf_counter() - t1

            # the lockstep rounds before the persistent tail: a cold-started batch
            # needs ~20 sweeps per cell, so more rounds move work from the tail
            # (one wave per cell) to the lane-parallel rounds; keep the fastest
            # (and, mixed precision, whether steps redone in FP64 wait for the
            # tail or join FP64 lockstep rounds)
            ctx.tune("tail_mode", 0)                  # the context is shared by the forms
            timed_solve()                              # first run allocates
            tries = {}
            for defer in ((0, 1) if "precision" in kw else (0,)):
                for rounds in ISOLATED_ROUNDS:
                    ctx.tune("rounds", rounds)
                    if "precision" in kw:
                        ctx.tune("mix_defer", defer)
                    tries[(rounds, defer)] = min(timed_solve(), timed_solve())
            best_rounds, best_defer = min(tries, key=tries.get)
            ctx.tune("rounds", best_rounds)
            if "precision" in kw:
                ctx.tune("mix_defer", best_defer)
            # the tail at one or at two waves per SIMD (latency against cells in flight)
            tail_tries = {0: tries[(best_rounds, best_defer)]}
            ctx.tune("tail_mode", 3)
            tail_tries[3] = min(timed_solve(), timed_solve())
            best_tail = min(tail_tries, key=tail_tries.get)
            ctx.tune("tail_mode", best_tail)
            best = min(timed_solve() for _ in range(3))
            st = ctx.stats()
            ug, _, itg, rg = ctx.batch_get()
            progress("isolated %s (line search %d): %.3f s" % (name, ls, best))
            res[name] = {"closures_per_s": n / best, "s": best, "rounds": best_rounds, "mix_defer": best_defer,
                         "tail_mode": best_tail,
                         "tried_s": dict({"rounds=%d defer=%d" % k: v for k, v in tries.items()},
                                         **{"tail_mode=%d" % k: v for k, v in tail_tries.items()}),
                         "newton_iters_per_closure": st.iter_sum / st.num_dual_solves,
                         "trials_per_closure": st.trial_sum / st.num_dual_solves,
                         "reg_hist": list(st.hist_reg), "warm_pick_mlp_iso_prev": list(st.warm_pick),
                         "warm_fallback": st.warm_fallback, "tail_cells": st.tail_cells,
                         "tail_iters_per_tail_cell": st.tail_iter_sum / max(st.tail_cells, 1)}
            if m > 0:
                # per cell against the oracle: cells at the same regularization level
                # solved the same problem to the tolerance; a different level is a
                # discrete decision of the reference's algorithm taken the other way
                err = np.linalg.norm(ug[:m] - uo, axis=1) / np.linalg.norm(uo, axis=1)
                same_r = rg[:m] == ro
                res[name].update({
                    "oracle_same_r_frac": float(same_r.mean()),
                    "oracle_same_path_frac": float((same_r & (itg[:m] == ito)).mean()),
                    "oracle_same_r_max_rel_u_err": float(err[same_r].max()),
                    "oracle_median_rel_u_err": float(np.median(err)),
                    "oracle_max_rel_u_err": float(err.max())})
        return res

    res = solve_forms(u, a0, mnc.LS_REFERENCE)
    robust = solve_forms(u, a0, mnc.LS_ROBUST)
    us_, _, _ = wl.isolated_batch(w, p, xi, eta, 15, n, seed, near_boundary_frac=0.0)
    smooth = solve_forms(us_, wl.isotropic_start(us_), mnc.LS_ROBUST)
    ctx.close()

    def gain(r):
        return r["learned+mixed"]["closures_per_s"] / r["mixed"]["closures_per_s"]

    return {"workload": "isolated M4 batch, %d vectors (20%% near-boundary lobes), Q=16, isotropic start, "
                        "TOL 1e-8, direct Hessian, reference line search" % n,
            "generate_s": gen_s, "oracle_sample": m, "mlp_loss_first_last": [float(loss[:10].mean()),
                                                                           float(loss[-50:].mean())],
            "forms": res, "learned_over_mixed": gain(res),
            "robust_ls": {"workload": "the same batch, robust line search", "forms": robust,
                          "learned_over_mixed": gain(robust)},
            "smooth_robust_ls": {"workload": "%d smooth vectors (no lobes, the network's training "
                                             "distribution, fresh draws), robust line search" % n,
                                 "forms": smooth, "learned_over_mixed": gain(smooth)}}


def bench_optcb(mnc, n, seed, oracle_sample, bfgs=False):
    """The change-of-basis optimiser (optcb.cpp:90-415), or with bfgs its BFGS
    variant (optbfgs.cpp:92-545, 5 inner iterations, cond bound 20), on the
    isolated M4 batch of bench_isolated (n vectors, Q = 16, 20% near-boundary
    lobes, TOL 1e-8, reference line search) from the identity basis and
    isotropic alphaP: converged closures/s from the kernel's HIP-event time
    (the entry point takes host arrays; the wall time with the PCIe copies is
    reported beside it), and the FP64 CPU oracle's form of the same optimiser
    on the first `oracle_sample` vectors."""
    sys.path.insert(0, ROOT)
    import oracle
    wl = _load("workloads")
    ctx = mnc.Context(4, 16, use_gaunt=False)
    w, p, xi, eta = ctx.tables()
    u, astar, lobe = wl.isolated_batch(w, p, xi, eta, 15, n, seed)
    nm = 15
    P2M = np.tile(np.eye(nm).reshape(1, -1), (n, 1))
    aP = np.tile(oracle.isotropic_alpha(nm), (n, 1))
    if bfgs:
        opt = mnc.options(max_iter=100, tol=1e-8, cond_h_max=1e10, opt_type=mnc.OPT_BFGS, max_bfgs_iter=5,
                          cond_h_max_bfgs=20.0)
    else:
        opt = mnc.options(max_iter=100, tol=1e-8, cond_h_max=1e10)
    name = "optbfgs" if bfgs else "optcb"
    best = None
    for rep_ in range(2):
        ctx.timing_enable(True)
        t1 = time.perf_counter()
        st, ug, Bg, aPg, aMg, it, r, ng = ctx.optcb_batch_host(u, P2M, aP, opt, allow_status=(mnc.OK, mnc.ERR_MAXREG))
        wall = time.perf_counter() - t1
        kn, kms = ctx.timing_get(mnc.KCLASS_NEWTON)
        ctx.timing_enable(False)
        if best is None or kms < best[0]:
            best = (kms, wall)
    ctx.close()
    progress("%s batch: kernel %.3f s, with copies %.3f s" % (name, best[0] * 1e-3, best[1]))
    tab = oracle.Tables(4, 16, False)
    m = oracle_sample
    threads = len(os.sched_getaffinity(0))
    cpu_s, same = 0.0, np.zeros(0, dtype=bool)
    if m > 0:
        t1 = time.perf_counter()
        if bfgs:
            sto, uo, Bo, aPo, aMo, ito, ro, _ = tab.optbfgs_batch(u[:m], P2M[:m], aP[:m], nthreads=threads,
                                                                  max_bfgs_iter=5, cond_h_max_bfgs=20.0)
        else:
            sto, uo, Bo, aPo, aMo, ito, ro, _ = tab.optcb_batch(u[:m], P2M[:m], aP[:m], nthreads=threads)
        cpu_s = time.perf_counter() - t1
        same = (it[:m] == ito) & (r[:m] == ro)
    return {"workload": "isolated M4 batch, %d vectors (20%% near-boundary lobes), Q=16, identity basis, "
                        "isotropic alphaP, TOL 1e-8, direct Hessian, reference line search%s" %
                        (n, "; BFGS variant, 5 inner iterations, cond bound 20" if bfgs else ""),
            "status": int(st), "closures_per_s": n / (best[0] * 1e-3), "kernel_s": best[0] * 1e-3,
            "closures_per_s_with_pcie": n / best[1],
            ("outer_iters_per_closure" if bfgs else "newton_iters_per_closure"): float(it.mean()),
            "max_iter": int(it.max()),
            "reg_hist": [int((r == v).sum()) for v in (0.0, 1e-3, 1e-2, 1e-1, 0.5)],
            "cpu_baseline": {"value": m / cpu_s, "unit": "closures/s", "cores": threads, "kind": "port",
                             "sample": "oracle %s on the first %d vectors (%.1f s)" % (name, m, cpu_s)}
            if m > 0 else None,
            "oracle_sample_identical_path_frac": float(same.mean()) if m > 0 else None}


def bench_sn(mnc, n=512, Q=16, steps=10, cpu_n=128, cpu_steps=2):
    """The Sn discrete-ordinates solver (kinetic_update.cpp:120-235) on n x n
    periodic cells of [-1,1]^2 with Q*Q ordinates, sigma_t = sigma_s = 1, an
    isotropic density 1 + 0.5 sin(pi x) sin(pi y), dt at CFL 0.9
    (kinetic_init.cpp:73): cell-ordinate updates/s from the device time of
    `steps` Heun steps (the state already on the device), the algorithmic HBM
    rate (a step streams the state 5 times: read and write per stage, the
    step's start read once more) and the oracle's single-threaded step on a
    cpu_n x cpu_n sample."""
    sys.path.insert(0, ROOT)
    import oracle
    nq = Q * Q
    dx = dy = 2.0 / n
    x = -1.0 + (np.arange(n + 4) - 2 + 0.5) * dx
    rho = 1.0 + 0.5 * np.outer(np.sin(np.pi * x), np.sin(np.pi * x))
    f = np.empty((n + 4, n + 4, nq))
    f[:] = rho[:, :, None]
    sig = np.ones((n + 4, n + 4))
    dt = 0.9 * (2.0 / (2.0 + 2.0)) * (dx * dy) / (dx + dy)
    mnc.sn_run(f[:8 + 4, :8 + 4].copy(), sig[:12, :12].copy(), sig[:12, :12].copy(), Q, dx, dy, -1.0, -1.0, 2, dt, 1)
    out, ms = mnc.sn_run(f, sig, sig, Q, dx, dy, -1.0, -1.0, 2, dt, steps)
    w = oracle.Tables(1, Q, False).w
    drift = abs((out[2:-2, 2:-2] * w).sum() - (f[2:-2, 2:-2] * w).sum()) / (f[2:-2, 2:-2] * w).sum()
    upd = float(n) * n * nq * steps
    bytes_ = 5.0 * 8.0 * upd
    progress("Sn %dx%d Q=%d: %d steps in %.3f ms" % (n, n, Q, steps, ms))
    xi, eta = oracle.Tables(1, Q, False).xi, oracle.Tables(1, Q, False).eta
    m = cpu_n
    fc = np.ascontiguousarray(f[:m + 4, :m + 4])
    sc = np.ascontiguousarray(sig[:m + 4, :m + 4])
    t1 = time.perf_counter()
    oracle.sn_steps(fc, w, xi, eta, sc, sc, 2, -1.0, -1.0, dx, dy, dt, steps=cpu_steps)
    cpu_s = time.perf_counter() - t1
    return {"workload": "Sn %d^2 periodic cells, Q=%d (%d ordinates), sigma_t = sigma_s = 1, isotropic "
                        "1 + 0.5 sin(pi x) sin(pi y), CFL 0.9, %d Heun steps" % (n, Q, nq, steps),
            "cell_ordinate_updates_per_s": upd / (ms * 1e-3), "ms_per_step": ms / steps,
            "algorithmic_gb_per_s": bytes_ / (ms * 1e-3) / 1e9,
            "hbm_frac": bytes_ / (ms * 1e-3) / 8.0e12, "mass_drift": drift,
            "cpu_baseline": {"value": float(m) * m * nq * cpu_steps / cpu_s, "unit": "cell-ordinate updates/s",
                             "cores": 1, "kind": "port",
                             "sample": "oracle orc_sn_step, %d^2 cells, %d steps (%.1f s)" % (m, cpu_steps, cpu_s)}}


def bench_m8(mnc, gi, steps, n=512, sample=1000):
    """The M8 workload: 45 moments, Q = 32 (1024 points), n x n periodic cells on
    [-1,1]^2, sigma_s = 1, isotropic density 1 + 0.5 sin(pi x) sin(pi y), TOL
    1e-8 with the robust line search (the reference's Armijo test regularizes
    ~0.5% of these closures at 1e-8 on rounding), direct-Hessian algorithm
    (FP64 matrix-core path): converged closures/s
    and s/step over K steps, and the error of GPU closures against the FP64 CPU
    oracle on `sample` cells of the final state."""
    sys.path.insert(0, ROOT)
    import oracle
    wl = _load("workloads")
    g = gi.initial_state(n, n, -1.0, 1.0, -1.0, 1.0, 2, 0.0, 1e-4, 1.0, 45, 0.9, 2.0)
    wl.sine_state(g, 45)
    ctx = mnc.Context(8, 32, use_gaunt=False)
    ctx.grid_create(g.nx, g.ny, g.dx, g.dy, g.init_x, g.init_y, 2, 2.0, g.moments, g.alpha, g.sigma_t, g.sigma_s)
    opt = mnc.options(max_iter=100, tol=1e-8, cond_h_max=1e10, line_search=mnc.LS_ROBUST)
    ctx.update(g.max_dt, opt)                 # allocations
    ctx.grid_set(g.moments, g.alpha)
    ctx.reset_stats()
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        ctx.update(g.max_dt, opt)
    ctx.sync()
    wall = time.perf_counter() - t0
    st = ctx.stats()
    m, a = ctx.grid_get()
    # the same steps again from the start with HIP events around the Newton
    # launches (the FP64 matrix-core rounds), for the roofline
    ctx.grid_set(g.moments, g.alpha)
    ctx.reset_stats()
    ctx.timing_enable(True)
    for _ in range(steps):
        ctx.update(g.max_dt, opt)
    ctx.sync()
    kst = ctx.stats()
    kn, kms = ctx.timing_get(mnc.KCLASS_NEWTON)
    ctx.timing_enable(False)
    f_iter, f_trial = flops_per_iter(45, 32 * 32)
    m8_flops = kst.iter_sum * f_iter + kst.trial_sum * f_trial
    m8_achieved = m8_flops / (kms * 1e-3) / 1e12 if kms > 0 else 0.0
    # closures of `sample` cells of the final state, GPU vs oracle, from the stored alpha
    rng = np.random.default_rng(5)
    cells = rng.choice((n + 4) * (n + 4), sample, replace=False)
    u = m.reshape(-1, 45)[cells] * (1.0 + 1e-3 * rng.standard_normal((sample, 45)))   # a fresh solve
    u[:, 0] = m.reshape(-1, 45)[cells, 0]
    a0 = a.reshape(-1, 45)[cells]
    _, ug, _, itg, rg, _ = ctx.opt_batch_host(u, a0, opt)
    ctx.close()
    tab = oracle.Tables(8, 32, False)
    _, uo, _, ito, ro, _ = tab.opt_batch(u, a0, max_iter=100, tol=1e-8, cond_h_max=1e10,
                                         nthreads=len(os.sched_getaffinity(0)), line_search=1)
    err = np.linalg.norm(ug - uo, axis=1) / np.linalg.norm(uo, axis=1)
    closures = 2 * (n + 4) * (n + 4) * steps
    return {"workload": "M8 (n=45), Q=32 (nq=1024), %dx%d periodic on [-1,1]^2, sigma_s=1, density "
                        "1+0.5 sin(pi x) sin(pi y), TOL 1e-8, robust line search, Heun dt=%.6g" % (n, n, g.max_dt),
            "steps": steps, "closures_per_s": closures / wall, "s_per_timestep": wall / steps,
            "newton_iters_per_closure": st.iter_sum / st.num_dual_solves,
            "trials_per_closure": st.trial_sum / st.num_dual_solves, "reg_hist": list(st.hist_reg),
            "oracle_sample": sample, "oracle_max_rel_u_err": float(err.max()),
            "oracle_same_iterations": float(np.mean(itg == ito)),
            "roofline": {"bound": "mfma", "achieved": m8_achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                         "frac": m8_achieved / FP64_PEAK_TFLOPS, "traffic": None,
                         "kernel": "Newton solve per stage (sweep_g64_kernel + logic_g64_kernel rounds), HIP events "
                                   "over the %d steps run again; flops: iterations x %.3f Mflop + trials x %.1f kflop "
                                   "(the reference's direct-Hessian FP64 count at n = 45, nq = 1024)"
                                   % (steps, f_iter / 1e6, f_trial / 1e3),
                         "algorithmic_flops_per_launch": m8_flops / max(2 * steps, 1),
                         "avg_launch_ms": kms / max(2 * steps, 1)}}


def bench_ppn(mnc, gi, d, steps, warmup, cpu_seconds):
    """The PPn form of the same lattice workload (MOMOPT_MOMENT_TYPE 1,
    MOMOPT_DELTA_PPN 1e-10; FP64 lane kernels and cooperative tail, robust line
    search): converged closures/s and s/step over K steps from t = 0, the
    Newton roofline as for the Mn forms, and the CPU oracle on the first steps
    of the same run (cpu_seconds of wall clock, 0 skips it)."""
    dp = dict(d, moment_type=1, delta_ppn=1e-10)
    n = (dp["moment_order"] + 1) * (dp["moment_order"] + 2) // 2
    f_iter, f_trial = flops_per_iter_ppn(n, dp["quad_order"] ** 2)
    opt = mnc.options(max_iter=dp["max_iter"], tol=dp["tol"], cond_h_max=dp["cond_h_max"],
                      line_search=mnc.LS_ROBUST, delta_ppn=dp["delta_ppn"])
    r = run_form(mnc, gi, dp, opt, steps, warmup, full_run=False)
    st, kst = r["stats"], r["kstats"]
    kms = r["newton"][1] + r["tail"][1]
    achieved = (kst.iter_sum * f_iter + kst.trial_sum * f_trial) / (kms * 1e-3) / 1e12 if kms > 0 else 0.0
    out = {"closures_per_s": r["closures"] / r["wall"], "s_per_timestep": r["wall"] / len(r["dts"]),
           "steps": len(r["dts"]), "newton_iters_per_closure": st.iter_sum / st.num_dual_solves,
           "trials_per_closure": st.trial_sum / st.num_dual_solves, "max_iter": st.max_iter,
           "reg_hist": list(st.hist_reg), "tuned": r["tuned"],
           "kernel_ms": {"newton_rounds": r["newton"][1], "tail": r["tail"][1], "flux": r["flux"][1]},
           "roofline": {"bound": "mfma", "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                        "frac": achieved / FP64_PEAK_TFLOPS,
                        "flops_per_iteration": f_iter, "flops_per_trial": f_trial}}
    if cpu_seconds > 0:
        cb = cpu_baseline(dp, cpu_seconds)
        out["cpu_baseline"] = {"value": cb["value"], "unit": "closures/s", "cores": cb["cores"], "kind": "port",
                               "sample": "oracle PPn on the first %d steps of the same run (%.1f s)"
                                         % (cb["steps"], cb["wall"])}
    return out


def cpu_baseline(d, seconds, nthreads=None):
    """The CPU oracle (FP64 C, OpenMP schedule(dynamic,1) over cells, as the
    reference's solveOptimization) on the first steps of the same run, for
    about `seconds` of wall clock (at least one step), on nthreads threads
    (default: all the cores this process may run on)."""
    sys.path.insert(0, ROOT)
    import oracle
    if nthreads is None:
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


def stage_probe(mnc, gi, d, form, steps):
    """One closure solve of every lattice cell (mnc_solve_grid) on the state
    after `steps` plain steps, in the given form with the default launch shape:
    the unit whose counters the HBM-traffic figure of the roofline sums."""
    ctx, g = make_grid(mnc, gi, d)
    dts = gi.time_steps(g.max_dt, T_FINAL)
    plain = mnc.options(max_iter=d["max_iter"], tol=d["tol"], cond_h_max=d["cond_h_max"], line_search=mnc.LS_ROBUST)
    for k in range(steps):
        ctx.update(dts[k], plain)
    kw = {"plain-fp64": {}, "mixed": {"precision": mnc.PREC_MIXED},
          "learned+mixed": {"precision": mnc.PREC_MIXED, "warm_start": mnc.WARM_LEARNED}}[form]
    if form == "learned+mixed":
        params, _ = train_mlp(mnc, gi, d, train_steps=300)
        ctx.mlp_set_params(params)
    ctx.ghost_fill()
    ctx.sync()
    progress("probe solve (%s) after %d steps" % (form, steps))
    ctx.solve_grid(mnc.options(max_iter=d["max_iter"], tol=d["tol"], cond_h_max=d["cond_h_max"],
                               line_search=mnc.LS_ROBUST, **kw))
    ctx.sync()
    st = ctx.stats()
    ctx.close()
    progress("probe done: %d closures" % st.num_dual_solves)


def copy_probe(mib):
    """A device-to-device copy of `mib` MiB (the runtime's copy kernel): the
    calibration of the FETCH_SIZE / WRITE_SIZE counters."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    n = mib << 20
    a, b = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(a), C.c_size_t(n)) == 0 and hip.hipMalloc(C.byref(b), C.c_size_t(n)) == 0
    assert hip.hipMemset(a, 1, C.c_size_t(n)) == 0
    assert hip.hipDeviceSynchronize() == 0
    assert hip.hipMemcpy(b, a, C.c_size_t(n), 3) == 0      # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)
    progress("copied %d MiB device to device" % mib)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=569)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--plain-only", action="store_true", help="only the plain FP64 form")
    ap.add_argument("--only", choices=["plain-fp64", "mixed", "learned+mixed"], default=None,
                    help="run one lattice form (for profiling); value is that form's rate")
    ap.add_argument("--train-steps", type=int, default=2000)
    ap.add_argument("--isolated", type=int, default=1_000_000,
                    help="vectors of the isolated closure batch (0 skips it)")
    ap.add_argument("--m8-steps", type=int, default=20, help="steps of the M8 512^2 workload (0 skips it)")
    ap.add_argument("--optcb", type=int, default=200_000,
                    help="vectors of the change-of-basis optimiser's isolated batch (0 skips it)")
    ap.add_argument("--ppn-steps", type=int, default=569, help="steps of the PPn lattice form (0 skips it)")
    ap.add_argument("--ppn-only", action="store_true", help="only the PPn lattice form (for profiling)")
    ap.add_argument("--sn-only", action="store_true", help="only the Sn solver leg")
    ap.add_argument("--isolated-only", action="store_true",
                    help="only the isolated closure batch (the three forms, for profiling)")
    ap.add_argument("--m8-only", action="store_true", help="only the M8 512^2 leg (for profiling)")
    ap.add_argument("--optcb-only", action="store_true",
                    help="only the change-of-basis optimiser's isolated batch (for profiling)")
    ap.add_argument("--tune", action="append", default=[], metavar="KEY=VALUE",
                    help="force a launch/tuning setting after the autotuner (repeatable)")
    ap.add_argument("--stage-probe", choices=["plain-fp64", "mixed", "learned+mixed"], default=None,
                    help="only one closure solve of the lattice at --probe-step (for counter runs)")
    ap.add_argument("--probe-step", type=int, default=284)
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only a device-to-device copy of MIB MiB (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build's hot path runs on a single MI355X "
              "(multi-GPU sharding is a design note only); run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    for kv in args.tune:
        k, v = kv.split("=")
        FORCE_TUNE[k] = int(v)
    if args.copy_probe > 0:
        copy_probe(args.copy_probe)
        return
    mnc, gi = _load("mnc"), _load("grid_init")
    d = dict(LATTICE)
    if args.stage_probe:
        stage_probe(mnc, gi, d, args.stage_probe, args.probe_step)
        return
    if args.optcb_only:
        r = bench_optcb(mnc, args.optcb, 12345, 0 if args.no_cpu else 50000)
        rb = bench_optcb(mnc, args.optcb, 12345, 0 if args.no_cpu else 50000, bfgs=True)
        print(json.dumps({"metric": "converged optcb closures/s (isolated M4 batch, TOL 1e-8)",
                          "value": r["closures_per_s"], "unit": "closures/s", "n_gpus": 1,
                          "higher_is_better": True, "config": {"optcb": r, "optbfgs": rb}}))
        return
    if args.isolated_only:
        r = bench_isolated(mnc, args.isolated, 12345, 0 if args.no_cpu else 20000, args.train_steps)
        print(json.dumps({"metric": "converged closures/s (isolated M4 batch, TOL 1e-8)",
                          "value": r["forms"]["learned+mixed"]["closures_per_s"], "unit": "closures/s",
                          "n_gpus": 1, "higher_is_better": True, "config": {"isolated": r}}))
        return
    if args.m8_only:
        r = bench_m8(mnc, gi, args.m8_steps)
        print(json.dumps({"metric": "converged M8 closures/s (512^2 periodic, Q=32, TOL 1e-8)",
                          "value": r["closures_per_s"], "unit": "closures/s", "n_gpus": 1,
                          "higher_is_better": True, "config": {"m8": r}}))
        return
    if args.sn_only:
        r = bench_sn(mnc)
        print(json.dumps({"metric": "Sn cell-ordinate updates/s (512^2, Q=16)",
                          "value": r["cell_ordinate_updates_per_s"], "unit": "updates/s", "n_gpus": 1,
                          "higher_is_better": True, "config": {"sn": r}}))
        return
    if args.ppn_only:
        r = bench_ppn(mnc, gi, d, args.ppn_steps, args.warmup, 0.0 if args.no_cpu else min(args.cpu_seconds, 10.0))
        print(json.dumps({"metric": "converged PPn closures/s (M4 lattice 280^2, TOL 1e-8)",
                          "value": r["closures_per_s"], "unit": "closures/s", "n_gpus": 1,
                          "steps": r["steps"], "higher_is_better": True, "config": {"ppn": r}}))
        return
    n = (d["moment_order"] + 1) * (d["moment_order"] + 2) // 2
    nq = d["quad_order"] ** 2
    f_iter, f_trial = flops_per_iter(n, nq)

    def opts(**kw):
        return mnc.options(max_iter=d["max_iter"], tol=d["tol"], cond_h_max=d["cond_h_max"],
                           line_search=mnc.LS_ROBUST, **kw)

    progress("start")
    want = ["plain-fp64"] if args.plain_only else ([args.only] if args.only else
                                                    ["plain-fp64", "mixed", "learned+mixed"])
    forms = {}
    mlp_info = None
    if "plain-fp64" in want:
        forms["plain-fp64"] = run_form(mnc, gi, d, opts(), args.steps, args.warmup)
        progress("plain form done")
    if "mixed" in want:
        forms["mixed"] = run_form(mnc, gi, d, opts(precision=mnc.PREC_MIXED), args.steps, args.warmup)
        progress("mixed form done")
    if "learned+mixed" in want:
        params, mlp_info = train_mlp(mnc, gi, d, train_steps=args.train_steps)
        progress("network trained")
        forms["learned+mixed"] = run_form(mnc, gi, d, opts(precision=mnc.PREC_MIXED, warm_start=mnc.WARM_LEARNED),
                                          args.steps, args.warmup, mlp_params=params)
    form = want[-1]
    plain = forms.get("plain-fp64", forms[form])
    best = forms[form]
    value = best["closures"] / best["wall"]
    kst = best["kstats"]
    stages = 2 * best["kernel_steps"]                 # two closure solves per Heun step
    # the dominant kernel: the persistent tail (FP64, one wave per cell), its
    # Newton iterations and Armijo trials counted on the device
    tn, tms = best["tail"]
    tail_flops = kst.tail_iter_sum * f_iter + kst.tail_trial_sum * f_trial
    tail_achieved = tail_flops / (tms * 1e-3) / 1e12 if tms > 0 else 0.0
    # the whole Newton stage (prologue, lockstep rounds and tail)
    sms = best["newton"][1] + tms
    alg_flops = kst.iter_sum * f_iter + kst.trial_sum * f_trial
    achieved = alg_flops / (sms * 1e-3) / 1e12 if sms > 0 else 0.0
    # drift of each form's final state from the plain form's (density rel-L2)
    dens = {k: v["moments"][2:-2, 2:-2, 0] for k, v in forms.items()}
    ref = dens.get("plain-fp64", dens[form])
    drift = {k: float(np.sqrt(((dens[k] - ref) ** 2).sum() / (ref ** 2).sum())) for k in forms}
    t_end = float(sum(best["dts"]))
    net = ""
    if form == "learned+mixed":
        gate = best["tuned"].get("learned_gate_log10")
        picks = list(best["stats"].warm_pick)
        if gate == 5:
            net = " (network off: the autotuner's gate 5)"
        else:
            net = " (network gate 10^%d: %.2f%% of the closures start from its guess)" % (
                gate, 100.0 * picks[0] / max(sum(picks), 1))
    out = {
        "metric": "converged Mn closures/s, M4 lattice 280^2, steps 1-%d of 569 (t = 0 to %.4f), TOL 1e-8, "
                  "%s%s vs plain FP64" % (args.steps, t_end, form, net),
        "value": value,
        "unit": "closures/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * best["wall"] / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if form == "plain-fp64" else "f64+bf16",
        "data": "synthetic (the reference's lattice initial condition, generated in-process)",
        "config": {"workload": "M4 lattice 280x280, Q=16 (nq=256), 7x7 checkerboard on [-3.5,3.5]^2, "
                               "floor 1e-10, TOL 1e-8, direct Hessian, robust line search, Heun dt=%.6g; timed "
                               "window: steps 1-%d of the 569 to t = 3.2 (t = 0 to %.4f); config.forms[*].full_run "
                               "times all 569 steps" % (plain["dts"][0], args.steps, t_end),
                   "form": form, "closures_per_step": 2 * 284 * 284,
                   "s_per_timestep": best["wall"] / args.steps,
                   "plain_fp64_closures_per_s": plain["closures"] / plain["wall"],
                   "speedup_vs_plain_fp64": value / (plain["closures"] / plain["wall"]),
                   "full_run_speedup_vs_plain_fp64":
                       (best["full_run"]["closures_per_s"] / plain["full_run"]["closures_per_s"])
                       if best["full_run"] and plain["full_run"] else None,
                   "forms": {k: form_summary(v, f_iter, f_trial) for k, v in forms.items()},
                   "final_density_rel_l2_vs_plain": drift,
                   "mlp": mlp_info,
                   "roofline_newton_stage": {
                       "bound": "mfma", "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                       "frac": achieved / FP64_PEAK_TFLOPS, "traffic": stage_traffic(form),
                       "kernel": "Newton solve per stage (init + lockstep rounds + tail), %s form; flops are the "
                                 "reference algorithm's FP64 count (iterations x %.1f kflop + trials x %.1f kflop); "
                                 "traffic: HBM bytes per stage from calibrated FETCH_SIZE + WRITE_SIZE "
                                 "(profiles/newton_stage_traffic.md)" % (form, f_iter / 1e3, f_trial / 1e3),
                       "algorithmic_flops_per_launch": alg_flops / max(stages, 1),
                       "avg_launch_ms": sms / max(stages, 1)}},
        "roofline": {"bound": "mfma", "achieved": tail_achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": tail_achieved / FP64_PEAK_TFLOPS, "traffic": stage_traffic("tail-" + form),
                     "kernel": "tail_coop_kernel (the persistent FP64 tail, one wave per cell still iterating "
                               "after the lockstep rounds), %s form, HIP events on the context's stream over the %d "
                               "timed steps run again eagerly; flops: the Newton iterations (x %.1f kflop) and "
                               "Armijo trials (x %.1f kflop) done in the tail, counted on the device: %.1f cells, "
                               "%.0f iterations and %.0f trials per launch"
                               % (form, best["kernel_steps"], f_iter / 1e3, f_trial / 1e3,
                                  kst.tail_cells / max(stages, 1), kst.tail_iter_sum / max(stages, 1),
                                  kst.tail_trial_sum / max(stages, 1)),
                     "algorithmic_flops_per_launch": tail_flops / max(tn, 1),
                     "avg_launch_ms": tms / max(tn, 1)},
    }
    if args.isolated > 0 and not args.plain_only and not args.only:
        out["config"]["isolated"] = bench_isolated(mnc, args.isolated, 12345, 20000, args.train_steps)
        progress("isolated batch done")
        out["config"]["optcb"] = bench_optcb(mnc, args.optcb, 12345, 50000) if args.optcb > 0 else None
        out["config"]["optbfgs"] = bench_optcb(mnc, args.optcb, 12345, 50000, bfgs=True) if args.optcb > 0 else None
        progress("change-of-basis batches done")
        out["config"]["sn"] = bench_sn(mnc)
        progress("Sn done")
    if args.m8_steps > 0 and not args.plain_only and not args.only:
        out["config"]["m8"] = bench_m8(mnc, gi, args.m8_steps)
        progress("M8 done")
    if args.ppn_steps > 0 and not args.plain_only and not args.only:
        out["config"]["ppn"] = bench_ppn(mnc, gi, d, args.ppn_steps, args.warmup,
                                         0.0 if args.no_cpu else min(args.cpu_seconds, 10.0))
        progress("PPn done")
    if not args.no_cpu:
        progress("CPU baseline")
        cb = cpu_baseline(d, args.cpu_seconds)
        c1 = cpu_baseline(d, min(args.cpu_seconds, 8.0), nthreads=1)
        out["cpu_baseline"] = {"value": cb["value"], "unit": "closures/s", "cores": cb["cores"], "kind": "port",
                               "sample": "oracle (FP64 C restatement, OpenMP dynamic,1) on the first %d steps "
                                         "of the same lattice run from t=0 (%.1f s, %.2f Newton iterations per "
                                         "closure)" % (cb["steps"], cb["wall"], cb["iters"]),
                               "one_core": {"value": c1["value"], "unit": "closures/s", "cores": 1,
                                            "sample": "the same oracle on one thread, first %d step(s) (%.1f s)"
                                                      % (c1["steps"], c1["wall"])}}
    print(json.dumps(out))


if __name__ == "__main__":
    main()

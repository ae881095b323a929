"""hipGraph-captured time steps (mnc_tune("graph", 1)) against eager launches.

A step key (dt, options) runs eagerly the first time and is captured and
replayed from then on.  Every replay is compared with an eager run of the same
steps, with the program's own eager work (state reads, stats, a batch solve on
the same context) and synchronisations between replays.  The per-cell solves do
not depend on the order in which cells are listed, so the states agree
bitwise."""
import numpy as np
import pytest

from helpers import fixture_deck, gpu_grid_from_deck, load_pkg, mn_options


def _steps(mnc, gi, d, dts, opt, graph, between=None, tune=None):
    ctx, g = gpu_grid_from_deck(mnc, gi, d)
    ctx.tune("graph", graph)
    for key, v in (tune or {}).items():
        ctx.tune(key, v)
    states = []
    for k, dt in enumerate(dts):
        ctx.update(dt, opt)
        m, a = ctx.grid_get()
        states.append((m.copy(), a.copy()))
        if between is not None:
            between(ctx, k)
    st = ctx.stats()
    ctx.close()
    return states, st


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_graph_steps_equal_eager(precision):
    mnc, gi = load_pkg("mnc"), load_pkg("grid_init")
    print("case graph lattice %s" % precision, flush=True)
    d = dict(num_cells_x=80, num_cells_y=80, a_x=-3.5, b_x=3.5, a_y=-3.5, b_y=3.5, gaussian_sigma=0.0,
             floor_value=1e-10, sigma=1.0, init_cond=1, moment_order=4, quad_order=16, cfl_factor=0.9,
             tol=1e-8, cond_h_max=1e10, max_iter=100, use_gaunt=0, theta=2.0, line_search=1)
    prec = mnc.PREC_MIXED if precision == "mixed" else mnc.PREC_FP64
    opt = mn_options(mnc, d, precision=prec)
    g0 = gi.initial_state(80, 80, -3.5, 3.5, -3.5, 3.5, 1, 0.0, 1e-10, 1.0, 15, 0.9, 2.0)
    dts = [g0.max_dt] * 5 + [0.5 * g0.max_dt] * 3 + [g0.max_dt] * 2   # two step keys, revisited

    rng = np.random.default_rng(0)
    u = np.abs(rng.normal(1.0, 0.1, (64, 15)))
    u[:, 1:] *= 0.05
    a0 = np.zeros_like(u)

    def between(ctx, k):     # eager work on the same stream between replays
        ctx.stats()
        ctx.opt_batch_host(u, a0, mnc.options(tol=1e-8))

    eager, st_e = _steps(mnc, gi, d, dts, opt, 0, between)
    graph, st_g = _steps(mnc, gi, d, dts, opt, 1, between)
    for k, ((me, ae), (mg, ag)) in enumerate(zip(eager, graph)):
        assert np.array_equal(me, mg), "moments differ after step %d" % k
        assert np.array_equal(ae, ag), "alpha differ after step %d" % k
    assert st_e.num_dual_solves == st_g.num_dual_solves
    assert st_e.iter_sum == st_g.iter_sum and st_e.trial_sum == st_g.trial_sum


@pytest.mark.gpu
def test_graph_fixture_deck():
    """The Gaunt M3 fixture deck through captured steps equals eager steps; with
    per-launch timing on, graph mode runs eagerly and counts every launch."""
    mnc, gi = load_pkg("mnc"), load_pkg("grid_init")
    print("case graph fixture deck", flush=True)
    d = fixture_deck("lattice")
    opt = mn_options(mnc, d)
    res = {}
    for graph, timing in ((0, False), (1, False), (1, True)):
        ctx, g = gpu_grid_from_deck(mnc, gi, d)
        ctx.tune("graph", graph)
        ctx.timing_enable(timing)
        for _ in range(6):
            ctx.update(g.max_dt, opt)
        kn, _ = ctx.timing_get(mnc.KCLASS_NEWTON)
        res[(graph, timing)] = (ctx.grid_get()[0], kn)
        ctx.close()
    assert np.array_equal(res[(0, False)][0], res[(1, False)][0])
    assert np.array_equal(res[(0, False)][0], res[(1, True)][0])
    assert res[(1, True)][1] == 12


@pytest.mark.gpu
def test_autotune_restores_state_and_keeps_results():
    """mnc_autotune leaves the grid state and the statistics as they were, and
    the launch shape it picks solves the same steps to rounding level."""
    mnc, gi = load_pkg("mnc"), load_pkg("grid_init")
    print("case autotune", flush=True)
    d = dict(num_cells_x=80, num_cells_y=80, a_x=-3.5, b_x=3.5, a_y=-3.5, b_y=3.5, gaussian_sigma=0.0,
             floor_value=1e-10, sigma=1.0, init_cond=1, moment_order=4, quad_order=16, cfl_factor=0.9,
             tol=1e-8, cond_h_max=1e10, max_iter=100, use_gaunt=0, theta=2.0, line_search=1)
    opt = mn_options(mnc, d)
    ctx, g = gpu_grid_from_deck(mnc, gi, d)
    ctx.update(g.max_dt, opt)
    m0, a0 = ctx.grid_get()
    s0 = ctx.stats()
    report = ctx.autotune(g.max_dt, opt, steps=2)
    print("\n".join(report))
    m1, a1 = ctx.grid_get()
    s1 = ctx.stats()
    assert np.array_equal(m0, m1) and np.array_equal(a0, a1)
    assert s0.num_dual_solves == s1.num_dual_solves and s0.iter_sum == s1.iter_sum
    assert len(report) >= 10
    tuned = {k: ctx.get_tune(k) for k in ("graph", "rounds", "tail_mode", "tail_blocks")}
    print("picked", tuned)
    for _ in range(3):
        ctx.update(g.max_dt, opt)
    mt = ctx.grid_get()[0]
    ctx.close()
    ref, _ = gpu_grid_from_deck(mnc, gi, d)
    for _ in range(4):
        ref.update(g.max_dt, opt)
    mr = ref.grid_get()[0]
    ref.close()
    rel = np.abs(mt - mr).max() / np.abs(mr).max()
    print("tuned vs default after 4 steps: max rel %.3e" % rel)
    assert rel <= 1e-10


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_flux_overlap_equals_serial(precision):
    """flux_overlap 1 (the default: the flux of the tiles whose stencil region
    holds no cell left for the Newton tail runs on a second stream beside the
    tail, the flagged tiles after it) against 0 (tail, then every tile): the
    same kernels compute the same cells, so the states agree bitwise, eager
    and captured into a graph (the fork and join are graph edges)."""
    mnc, gi = load_pkg("mnc"), load_pkg("grid_init")
    print("case flux overlap %s" % precision, flush=True)
    d = dict(num_cells_x=100, num_cells_y=90, a_x=-3.5, b_x=3.5, a_y=-3.5, b_y=3.5, gaussian_sigma=0.0,
             floor_value=1e-10, sigma=1.0, init_cond=1, moment_order=4, quad_order=16, cfl_factor=0.9,
             tol=1e-8, cond_h_max=1e10, max_iter=100, use_gaunt=0, theta=2.0, line_search=1)
    prec = mnc.PREC_MIXED if precision == "mixed" else mnc.PREC_FP64
    opt = mn_options(mnc, d, precision=prec)
    g0 = gi.initial_state(100, 90, -3.5, 3.5, -3.5, 3.5, 1, 0.0, 1e-10, 1.0, 15, 0.9, 2.0)
    dts = [g0.max_dt] * 8
    serial, st_s = _steps(mnc, gi, d, dts, opt, 0, tune={"flux_overlap": 0, "rounds": 2})
    for graph in (0, 1):
        over, st_o = _steps(mnc, gi, d, dts, opt, graph, tune={"flux_overlap": 1, "rounds": 2})
        for k, ((ms, as_), (mo, ao)) in enumerate(zip(serial, over)):
            assert np.array_equal(ms, mo), "moments differ after step %d (graph %d)" % (k, graph)
            assert np.array_equal(as_, ao), "alpha differ after step %d (graph %d)" % (k, graph)
        assert st_s.iter_sum == st_o.iter_sum and st_s.tail_cells == st_o.tail_cells
    print("tail cells over 8 steps: %d" % st_s.tail_cells)
    assert st_s.tail_cells > 0

"""M8 (45 moments, Q = 32: 1024 points) on the FP64 matrix-core path against the
CPU oracle (the reference's direct-Hessian algorithm)."""
import math

import numpy as np
import pytest

from helpers import density, load_pkg, oracle, rel_l2


def _m8_inputs(tab, n, seed):
    rng = np.random.default_rng(seed)
    pm = tab.p[:, :tab.nm]
    a = np.zeros((n, tab.nm))
    a[:, 0] = rng.uniform(-2.0, 2.0, n)
    for k in range(1, tab.nm):
        a[:, k] = rng.normal(0.0, 1.0 / (1.0 + k), n)
    u = (np.exp(a @ pm.T) * tab.w[None, :]) @ pm
    return u, a


@pytest.mark.gpu
def test_m8_closure_batch_matches_oracle():
    mnc = load_pkg("mnc")
    o = oracle()
    print("case m8 batch", flush=True)
    tab = o.Tables(8, 32, False)
    u, astar = _m8_inputs(tab, 300, 8)
    a0 = np.zeros_like(u)
    a0[:, 0] = 2.0 * math.sqrt(math.pi) * (math.log(0.5 / math.sqrt(math.pi)) + np.log(u[:, 0]))
    ctx = mnc.Context(8, 32, use_gaunt=False)
    tol = 1e-8
    st, ug, ag, itg, rg, ngg = ctx.opt_batch_host(u, a0, mnc.options(max_iter=100, tol=tol, cond_h_max=1e10))
    ctx.close()
    sto, uo, ao, ito, ro, _ = tab.opt_batch(u, a0, max_iter=100, tol=tol, cond_h_max=1e10, nthreads=8)
    assert st == 0 and sto == 0
    assert np.array_equal(ug[:, 0], u[:, 0])
    same = (itg == ito) & (rg == ro)
    rs = rg == ro
    un = np.linalg.norm(ug[rs] - uo[rs], axis=1) / np.linalg.norm(uo[rs], axis=1)
    print("M8: identical (iter, r) %.4f, mean iters gpu %.2f oracle %.2f, max ||du||/||u|| %.3e"
          % (same.mean(), itg.mean(), ito.mean(), un.max()))
    assert rs.mean() >= 0.97
    assert same.mean() >= 0.9
    assert un.max() <= 4 * tol
    pm = tab.p[:, :tab.nm]
    recon = (np.exp(ag @ pm.T) * tab.w[None, :]) @ pm
    rel = np.abs(recon - ug).max(axis=1) / np.abs(ug).max(axis=1)
    assert rel.max() <= 1e-10


@pytest.mark.gpu
def test_m8_sine_grid_steps_match_oracle():
    """Two Heun steps of the smooth periodic M8 workload on a 20 x 20 grid, with
    the robust line search: at TOL 1e-8 the reference's Armijo test cannot
    certify the last Newton decrease of ~0.5% of these closures and regularizes
    them (r = 1e-3), a rounding-level decision that moves the state by ~1e-6;
    with the robust rule no cell regularizes and the oracle's own TOL 1e-8 and
    1e-10 runs agree to 1.2e-10 in density."""
    mnc, gi, wl = load_pkg("mnc"), load_pkg("grid_init"), load_pkg("workloads")
    print("case m8 grid", flush=True)
    d = dict(num_cells_x=20, num_cells_y=20, a_x=-1.0, b_x=1.0, a_y=-1.0, b_y=1.0, gaussian_sigma=0.0,
             floor_value=1e-4, sigma=1.0, init_cond=2, moment_order=8, quad_order=32, cfl_factor=0.9,
             tol=1e-8, cond_h_max=1e10, max_iter=100, use_gaunt=0, theta=2.0, line_search=1)
    g = gi.initial_state(20, 20, -1.0, 1.0, -1.0, 1.0, 2, 0.0, 1e-4, 1.0, 45, 0.9, 2.0)
    wl.sine_state(g, 45)
    ctx = mnc.Context(8, 32, use_gaunt=False)
    ctx.grid_create(g.nx, g.ny, g.dx, g.dy, g.init_x, g.init_y, 2, 2.0, g.moments, g.alpha, g.sigma_t, g.sigma_s)
    opt = mnc.options(max_iter=100, tol=1e-8, cond_h_max=1e10, line_search=mnc.LS_ROBUST)
    for _ in range(2):
        ctx.update(g.max_dt, opt)
    mg = ctx.grid_get()[0]
    ctx.close()
    s = oracle().Solver(d, nthreads=8)
    s.set_state(g.moments, g.alpha)
    for _ in range(2):
        s.update(g.max_dt)
    mo = s.state()[0]
    e = rel_l2(density(mg[2:-2, 2:-2]), density(mo[2:-2, 2:-2]))
    emax = np.abs(mg - mo).max() / np.abs(mo).max()
    print("M8 sine grid, 2 steps: density rel-L2 %.3e, max rel moment diff %.3e" % (e, emax))
    st = s.stats()
    assert list(st.hist_reg)[1:] == [0, 0, 0, 0]
    assert e <= 1e-9
    assert emax <= 1e-7


@pytest.mark.gpu
def test_m8_512_steps_conserve_mass_and_match_oracle_sample():
    """The M8 workload at full size (512 x 512 periodic cells, Q = 32, TOL
    1e-8, robust line search), 3 Heun steps on the GPU: the total mass of the
    periodic grid with sigma_t = sigma_s is conserved to rounding (a size-
    independent property of the scheme), and the closures of 400 sampled
    cells of the final state, solved afresh from the stored alpha (moments
    perturbed by 1e-3 so that they need Newton steps), agree with the
    oracle's per cell.  A full 512^2 oracle step would take about a minute
    per stage on the host, so the grid state itself is checked against the
    oracle on the 20 x 20 grid above."""
    import os
    mnc, gi, wl = load_pkg("mnc"), load_pkg("grid_init"), load_pkg("workloads")
    print("case m8 512^2", flush=True)
    n = 512
    g = gi.initial_state(n, n, -1.0, 1.0, -1.0, 1.0, 2, 0.0, 1e-4, 1.0, 45, 0.9, 2.0)
    wl.sine_state(g, 45)
    ctx = mnc.Context(8, 32, use_gaunt=False)
    ctx.grid_create(g.nx, g.ny, g.dx, g.dy, g.init_x, g.init_y, 2, 2.0, g.moments, g.alpha, g.sigma_t, g.sigma_s)
    opt = mnc.options(max_iter=100, tol=1e-8, cond_h_max=1e10, line_search=mnc.LS_ROBUST)
    m0 = density(g.moments[2:-2, 2:-2]).sum()
    for _ in range(3):
        ctx.update(g.max_dt, opt)
    m, a = ctx.grid_get()
    st = ctx.stats()
    dm = abs(density(m[2:-2, 2:-2]).sum() - m0) / m0
    print("M8 512^2: mass drift %.2e over 3 steps, %.3f iterations per closure" % (dm, st.iter_sum / st.num_dual_solves))
    assert dm < 1e-12
    assert st.num_dual_solves == 3 * 2 * (n + 4) ** 2 and st.hist_reg[1:] == [0, 0, 0, 0]
    rng = np.random.default_rng(5)
    cells = rng.choice((n + 4) * (n + 4), 400, replace=False)
    u = m.reshape(-1, 45)[cells] * (1.0 + 1e-3 * rng.standard_normal((400, 45)))
    u[:, 0] = m.reshape(-1, 45)[cells, 0]
    a0 = a.reshape(-1, 45)[cells]
    stg, ug, ag, itg, rg, _ = ctx.opt_batch_host(u, a0, opt)
    ctx.close()
    tab = oracle().Tables(8, 32, False)
    sto, uo, ao, ito, ro, _ = tab.opt_batch(u, a0, max_iter=100, tol=1e-8, cond_h_max=1e10, line_search=1,
                                            nthreads=min(16, len(os.sched_getaffinity(0))))
    assert stg == 0 and sto == 0
    err = np.linalg.norm(ug - uo, axis=1) / np.linalg.norm(uo, axis=1)
    print("M8 512^2 sample: same iterations %.3f, max ||du||/||u|| %.2e" % ((itg == ito).mean(), err.max()))
    assert (rg == ro).all() and (itg == ito).mean() >= 0.97
    assert err.max() <= 4e-8

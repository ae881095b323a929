"""The change-of-basis optimiser on the GPU (mnc_optcb_batch_host) against the
oracle's restatement (orc_optcb, itself pinned in tests/test_oracle_optcb.py).

The n x n algebra runs in the reference's order on both sides; the quadrature
sums differ in order (lane partials and wave sums on the GPU, sequential in the
oracle), so the two follow the same Newton path up to rounding: the checks are
identical (iterations, regularization) for nearly every cell, the multipliers
and bases to rounding on those cells, and every cell solved to the tolerance."""
import numpy as np
import pytest

import os
import subprocess

from helpers import (PKG, config2_batch, deck_text, density, fixture_deck, fixture_path, gpu_grid_from_deck, load_pkg,
                     oracle, read_pn, rel_l2)


def _moments(tab, alpha):
    pm = tab.p[:, :tab.nm]
    return (np.exp(alpha @ pm.T) * tab.w[None, :]) @ pm


def _identity_start(o, nm, n):
    P2M = np.tile(np.eye(nm).reshape(1, -1), (n, 1))
    aP = np.tile(o.isotropic_alpha(nm), (n, 1))
    return P2M, aP


def _run(order, Q, n, seed, nb_frac, P2M=None, aP=None, u=None, line_search=0, bfgs=None):
    """bfgs: None for optcb, else (max_bfgs_iter, cond_h_max_bfgs) for optbfgs."""
    mnc = load_pkg("mnc")
    o = oracle()
    print("case optcb order=%d Q=%d n=%d seed=%d lobes=%.2f ls=%d bfgs=%s" %
          (order, Q, n, seed, nb_frac, line_search, bfgs), flush=True)
    tab = o.Tables(order, Q, False)
    if u is None:
        rng = np.random.default_rng(seed)
        u, astar, lobe = config2_batch(tab.w, tab.p, tab.xi, tab.eta, tab.nm, n, rng, near_boundary_frac=nb_frac)
    if P2M is None:
        P2M, aP = _identity_start(o, tab.nm, len(u))
    ctx = mnc.Context(order, Q, use_gaunt=False)
    if bfgs is None:
        opt = mnc.options(max_iter=100, tol=1e-8, cond_h_max=1e10, line_search=line_search)
    else:
        opt = mnc.options(max_iter=100, tol=1e-8, cond_h_max=1e10, line_search=line_search,
                          opt_type=mnc.OPT_BFGS, max_bfgs_iter=bfgs[0], cond_h_max_bfgs=bfgs[1])
    g = ctx.optcb_batch_host(u, P2M, aP, opt)
    ctx.close()
    if bfgs is None:
        c = tab.optcb_batch(u, P2M, aP, tol=1e-8, nthreads=8, line_search=line_search)
    else:
        c = tab.optbfgs_batch(u, P2M, aP, tol=1e-8, nthreads=8, line_search=line_search, max_bfgs_iter=bfgs[0],
                              cond_h_max_bfgs=bfgs[1])
    return tab, u, g, c


def _compare(tab, u, g, c, same_frac):
    stg, ug, Bg, aPg, aMg, itg, rg, ngg = g
    sto, uo, Bo, aPo, aMo, ito, ro, ngo = c
    assert stg == 0 and sto == 0
    same = (itg == ito) & (rg == ro)
    print("identical (iterations, r): %.4f; mean iterations gpu %.3f oracle %.3f" % (same.mean(), itg.mean(), ito.mean()))
    assert same.mean() >= same_frac
    rel = np.abs(aMg - aMo).max(axis=1) / np.abs(aMo).max(axis=1)
    relB = np.abs(Bg - Bo).max(axis=1) / np.abs(Bo).max(axis=1)
    print("same-path cells: alphaM max rel %.3e median %.3e; P2M max rel %.3e" %
          (rel[same].max(), np.median(rel[same]), relB[same].max()))
    assert np.median(rel[same]) <= 1e-12
    # near the realizability boundary alpha is fixed only to ~cond(H) tol: the
    # same-path cells agree in the moments they reproduce
    mg, mo = _moments(tab, aMg[same]), _moments(tab, aMo[same])
    dm = np.linalg.norm(mg - mo, axis=1) / np.linalg.norm(mo, axis=1)
    print("same-path cells: ||moments(gpu) - moments(oracle)|| / ||u|| max %.3e" % dm.max())
    assert dm.max() <= 4e-8
    assert np.median(relB[same]) <= 1e-12
    ok = rg == 0.0
    res = np.linalg.norm(_moments(tab, aMg[ok]) - u[ok], axis=1) / np.linalg.norm(u[ok], axis=1)
    print("unregularized cells: ||moments(alphaM) - u|| / ||u|| max %.3e" % res.max())
    assert res.max() <= 2e-8
    np.testing.assert_allclose(ug[ok], u[ok], rtol=1e-15, atol=0)


@pytest.mark.gpu
@pytest.mark.parametrize("order,Q", [(1, 8), (2, 12), (4, 16)])
def test_optcb_matches_oracle(order, Q):
    tab, u, g, c = _run(order, Q, 600, 100 + order, 0.1)
    _compare(tab, u, g, c, 0.95)


@pytest.mark.gpu
@pytest.mark.parametrize("order,Q,ls", [(2, 12, 0), (4, 16, 0), (4, 16, 1)])
def test_optbfgs_matches_oracle(order, Q, ls):
    """The BFGS variant (MNC_OPT_BFGS, the decks' 5 inner iterations and cond
    bound 20) against the oracle's optbfgs: quasi-Newton paths amplify the
    quadrature sums' rounding, so fewer cells keep the exact path."""
    tab, u, g, c = _run(order, Q, 500, 200 + order, 0.1, line_search=ls, bfgs=(5, 20.0))
    _compare(tab, u, g, c, 0.9)


@pytest.mark.gpu
def test_optbfgs_one_inner_iteration_is_optcb_bitwise():
    """With one inner iteration and no cond bound the BFGS mode takes optcb's
    steps (H^-1 = I gives d = -g exactly), so its outputs are optcb's, bit
    for bit, on the GPU as in the oracle."""
    mnc = load_pkg("mnc")
    o = oracle()
    tab = o.Tables(3, 12, False)
    rng = np.random.default_rng(77)
    u, astar, lobe = config2_batch(tab.w, tab.p, tab.xi, tab.eta, tab.nm, 400, rng, near_boundary_frac=0.2)
    P2M, aP = _identity_start(o, tab.nm, len(u))
    ctx = mnc.Context(3, 12, use_gaunt=False)
    a = ctx.optcb_batch_host(u, P2M, aP, mnc.options())
    b = ctx.optcb_batch_host(u, P2M, aP, mnc.options(opt_type=mnc.OPT_BFGS, max_bfgs_iter=1, cond_h_max_bfgs=1e300))
    st = ctx.optcb_batch_host(u, P2M, aP, mnc.options(opt_type=mnc.OPT_BFGS, max_bfgs_iter=0),
                              allow_status=(mnc.OK, mnc.ERR_ARG))[0]
    assert st == mnc.ERR_ARG
    st = ctx.optcb_batch_host(u, P2M, aP, mnc.options(opt_type=mnc.OPT_BFGS, cond_h_max_bfgs=1.0),
                              allow_status=(mnc.OK, mnc.ERR_ARG))[0]
    assert st == mnc.ERR_ARG
    ctx.close()
    assert a[0] == 0 and b[0] == 0
    for x, y in zip(a[1:], b[1:]):
        assert np.array_equal(x, y)


@pytest.mark.gpu
@pytest.mark.parametrize("bfgs", [False, True])
def test_optcb_cell_results_do_not_depend_on_the_batch(bfgs):
    """The kernel's blocks take cells from a queue, as many blocks as fit the
    GPU at once. A cell's answer must not depend on which block took it, on
    what that block solved before (its workspace), or on the batch around it.
    With 5,000 cells, more than the resident blocks, some blocks take several
    cells. The whole batch, reversed, must equal it bit for bit, and so must
    a single cell and a slice of seven."""
    mnc = load_pkg("mnc")
    o = oracle()
    tab = o.Tables(4, 16, False)
    rng = np.random.default_rng(2024)
    u, astar, lobe = config2_batch(tab.w, tab.p, tab.xi, tab.eta, tab.nm, 5000, rng, near_boundary_frac=0.2)
    P2M, aP = _identity_start(o, tab.nm, len(u))
    opt = (mnc.options(max_iter=100, tol=1e-8, cond_h_max=1e10, opt_type=mnc.OPT_BFGS, max_bfgs_iter=5,
                       cond_h_max_bfgs=20.0) if bfgs else mnc.options(max_iter=100, tol=1e-8, cond_h_max=1e10))
    ok = (mnc.OK, mnc.ERR_MAXREG)
    print("case optcb batch independence bfgs=%d n=%d" % (bfgs, len(u)), flush=True)
    ctx = mnc.Context(4, 16, use_gaunt=False)
    full = ctx.optcb_batch_host(u, P2M, aP, opt, allow_status=ok)
    rev = ctx.optcb_batch_host(u[::-1], P2M[::-1], aP[::-1], opt, allow_status=ok)
    one = ctx.optcb_batch_host(u[4321:4322], P2M[4321:4322], aP[4321:4322], opt, allow_status=ok)
    seven = ctx.optcb_batch_host(u[1000:1007], P2M[1000:1007], aP[1000:1007], opt, allow_status=ok)
    ctx.close()
    assert np.all(full[5] > 0)                       # every cell was solved
    for k in range(1, 8):                            # u, P2M, alphaP, alphaM, iter, r, norm_g
        assert np.array_equal(full[k], rev[k][::-1])
        assert np.array_equal(full[k][4321:4322], one[k])
        assert np.array_equal(full[k][1000:1007], seven[k])


@pytest.mark.gpu
def test_optcb_robust_line_search_matches_oracle():
    tab, u, g, c = _run(4, 16, 400, 7, 0.1, line_search=1)
    _compare(tab, u, g, c, 0.95)


@pytest.mark.gpu
def test_optcb_reference_initial_basis_and_warm_restart():
    """The reference's grid initialisation (ones in P2M's first column) as
    input, then a second solve of nearby moments from the returned state."""
    o = oracle()
    mnc = load_pkg("mnc")
    tab = o.Tables(3, 12, False)
    nm = tab.nm
    rng = np.random.default_rng(55)
    u, astar, lobe = config2_batch(tab.w, tab.p, tab.xi, tab.eta, tab.nm, 300, rng, near_boundary_frac=0.0)
    P2M = np.zeros((len(u), nm * nm))
    P2M[:, :nm] = 1.0
    aP = np.tile(o.isotropic_alpha(nm), (len(u), 1))
    tab_, u_, g, c = _run(3, 12, 0, 0, 0.0, P2M=P2M, aP=aP, u=u)
    _compare(tab, u, g, c, 0.95)
    stg, ug, Bg, aPg, aMg, itg, rg, ngg = g
    un = u * (1.0 + 1e-3 * rng.standard_normal(u.shape))
    un[:, 0] = u[:, 0]
    tab_, u2, g2, c2 = _run(3, 12, 0, 0, 0.0, P2M=Bg, aP=aPg, u=un)
    _compare(tab, un, g2, c2, 0.95)
    print("warm restart: mean iterations %.2f (first solve %.2f)" % (g2[5].mean(), itg.mean()))
    assert g2[5].mean() <= 0.6 * itg.mean()


@pytest.mark.gpu
def test_optcb_rejects_unsupported_forms():
    mnc = load_pkg("mnc")
    o = oracle()
    nm = 6
    u = np.zeros((2, nm))
    u[:, 0] = 1.0
    P2M, aP = _identity_start(o, nm, 2)
    ctx = mnc.Context(2, 12, use_gaunt=True)
    st = ctx.optcb_batch_host(u, P2M, aP, mnc.options(), allow_status=(mnc.OK, mnc.ERR_ARG))[0]
    ctx.close()
    assert st == mnc.ERR_ARG
    ctx = mnc.Context(2, 12, use_gaunt=False)
    st = ctx.optcb_batch_host(u, P2M, aP, mnc.options(precision=mnc.PREC_MIXED),
                              allow_status=(mnc.OK, mnc.ERR_ARG))[0]
    assert st == mnc.ERR_ARG
    ctx.close()


@pytest.mark.gpu
def test_grid_solve_with_bfgs_equals_batch():
    """mnc_solve_grid with MNC_OPT_BFGS: the batch entry point's result from
    the reference's initial state, bit for bit."""
    mnc, gi = load_pkg("mnc"), load_pkg("grid_init")
    print("case optbfgs grid solve", flush=True)
    d = dict(fixture_deck("gaussian"), use_gaunt=0, tol=1e-8)
    ctx, g = gpu_grid_from_deck(mnc, gi, d)
    nm = ctx.nm
    m0, a0 = ctx.grid_get()
    opt = mnc.options(max_iter=d["max_iter"], tol=d["tol"], cond_h_max=d["cond_h_max"], opt_type=mnc.OPT_BFGS)
    ctx.reset_stats()
    ctx.solve_grid(opt)
    st = ctx.stats()
    mg, ag = ctx.grid_get()
    n = m0.reshape(-1, nm).shape[0]
    assert st.num_dual_solves == n
    B = np.zeros((n, nm * nm))
    B[:, :nm] = 1.0
    aP = np.tile(oracle().isotropic_alpha(nm), (n, 1))
    stb, ub, Bb, aPb, aMb, itb, rb, _ = ctx.optcb_batch_host(m0.reshape(-1, nm), B, aP, opt)
    ctx.close()
    assert stb == 0
    print("outer iterations per closure %.2f" % (st.iter_sum / n))
    assert np.array_equal(ag.reshape(-1, nm), aMb)
    assert np.array_equal(mg.reshape(-1, nm), ub)


@pytest.mark.gpu
def test_grid_solve_with_change_of_basis_equals_batch():
    """mnc_solve_grid with opt_type MNC_OPT_CHANGE_BASIS runs the same kernel
    on the grid's moments, alpha and the device-held basis state, initialised
    as momopt_init.cpp:144-158: bit for bit the batch entry point's result from
    that initial state, and the state carries over to the next solve."""
    mnc, gi = load_pkg("mnc"), load_pkg("grid_init")
    print("case optcb grid solve", flush=True)
    # the regression decks use the Gaunt form, which optcb cannot (it reads the
    # table with stride n): the same deck with the direct Hessian
    d = dict(fixture_deck("gaussian"), use_gaunt=0, tol=1e-8)
    ctx, g = gpu_grid_from_deck(mnc, gi, d)
    nm = ctx.nm
    m0, a0 = ctx.grid_get()
    opt = mnc.options(max_iter=d["max_iter"], tol=d["tol"], cond_h_max=d["cond_h_max"], opt_type=mnc.OPT_CHANGE_BASIS)
    ctx.reset_stats()
    ctx.solve_grid(opt)
    st = ctx.stats()
    mg, ag = ctx.grid_get()
    n = m0.reshape(-1, nm).shape[0]
    assert st.num_dual_solves == n
    B = np.zeros((n, nm * nm))
    B[:, :nm] = 1.0
    aP = np.tile(oracle().isotropic_alpha(nm), (n, 1))
    stb, ub, Bb, aPb, aMb, itb, rb, _ = ctx.optcb_batch_host(m0.reshape(-1, nm), B, aP, opt)
    assert stb == 0
    assert np.array_equal(ag.reshape(-1, nm), aMb)
    assert np.array_equal(mg.reshape(-1, nm), ub)
    # the same moments again: the kept basis solves them at once
    ctx.reset_stats()
    ctx.solve_grid(opt)
    st2 = ctx.stats()
    print("first solve %.2f iterations per closure, second %.2f" % (st.iter_sum / n, st2.iter_sum / n))
    assert st2.iter_sum / n <= 1.2
    # a new grid state restarts the basis state
    ctx.grid_set(m0, a0)
    ctx.solve_grid(opt)
    assert np.array_equal(ctx.grid_get()[1].reshape(-1, nm), aMb)
    ctx.close()
    # the batch entry points refuse the change-of-basis optimiser
    ctx = mnc.Context(2, 12, use_gaunt=False)
    u = np.zeros((1, 6))
    u[0, 0] = 1.0
    st = ctx.opt_batch_host(u, np.zeros((1, 6)), opt, allow_status=(mnc.OK, mnc.ERR_ARG))[0]
    ctx.close()
    assert st == mnc.ERR_ARG


@pytest.mark.gpu
def test_driver_change_of_basis_time_stepping(tmp_path):
    """MOMOPT_OPTIMIZATION_TYPE 1 and 2 through the deck driver: the smooth periodic
    deck (36 x 74, M3, Q = 30) with the direct Hessian and TOL 1e-8 to t =
    1.03, robust line search (at TOL 1e-8 the reference's Armijo test fails
    on f's rounding for ~1% of these closures and regularizes them).  Neither
    optimiser regularizes then (the .opt histograms), and both
    solve every closure to the tolerance, so the change-of-basis run lands on
    the Newton optimiser's run (itself checked against the oracle and the
    fixtures) at tolerance level.  (Near vacuum the two part ways by design:
    the Newton optimiser regularizes cells whose Hessian is ill-conditioned in
    the moment basis, the change of basis solves them unblended.)"""
    d = dict(fixture_deck("smooth"), use_gaunt=0, tol=1e-8)
    exe = os.path.join(PKG, "momopt_gpu")
    res = {}
    for ot in (0, 1, 2):
        run = tmp_path / ("opt%d" % ot)
        run.mkdir()
        (run / "input.deck").write_text(deck_text(d, 1.03, {"MOMOPT_GPU_LINE_SEARCH": 1}, opt_type=ot))
        r = subprocess.run([exe], cwd=run, capture_output=True, text=True, timeout=600)
        assert r.returncode == 0, r.stdout[-800:]
        res[ot] = read_pn(str(run / "out_1.030_0.pn"))[1]
        hist = (run / "out_1.030_0.opt").read_text()
        print("opt type %d:\n%s" % (ot, hist))
        regs = [int(line.split()[-1]) for line in hist.splitlines() if line.startswith("r = ")]
        assert len(regs) == 5 and sum(regs[1:]) == 0
    err = rel_l2(density(res[1]), density(res[0]))
    print("optcb vs Newton time stepping: density rel-L2 %.3e" % err)
    assert err < 1e-7
    err = rel_l2(density(res[2]), density(res[0]))
    print("optbfgs vs Newton time stepping: density rel-L2 %.3e" % err)
    assert err < 1e-7

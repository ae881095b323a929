"""Pins the oracle's PPn closure (fobj.cpp:137-236,355-428, opt.cpp:113-141,
isotropicAlpha opt.cpp:36-39, computeAnsatz momopt_update.cpp:54-57): the
reference's four PPn regression fixtures (tests/regression/momopt/<ic>/ppn.pn,
copied as data), derivatives by central differences, the Gaunt Hessian
against the direct one, the isotropic closed form and known answers.

The PPn ansatz is G(a) = a/2 + sqrt(a^2/4 + delta) of a = alpha.m: positive
for any alpha, and exp-free."""
import math
import os

import numpy as np
import pytest

from helpers import FIXTURE_T_FINAL, GOLDEN, density, fixture_deck, mass, oracle, read_pn, rel_l2

DELTA = 1e-10


def ppn_ansatz(a, delta):
    t = 0.5 * a
    s = np.sqrt(t * t + delta)
    return np.where(t < 0, -delta / (t - s), t + s)


@pytest.mark.parametrize("ic", ["gaussian", "smooth", "delta", "lattice"])
def test_ppn_fixture(ic):
    o = oracle()
    d = fixture_deck(ic, moment_type=o.MOMENT_PPN, delta_ppn=DELTA)
    s = o.Solver(d, nthreads=8)
    m0 = s.moments().copy()
    steps = s.run(FIXTURE_T_FINAL)
    m = s.moments()
    h, ref = read_pn(os.path.join(GOLDEN, "momopt", ic, "ppn.pn"))
    assert h[:2] == (36, 74) and h[6] == 10 and steps == 84
    err = rel_l2(density(m), density(ref))
    print("%s PPn rel L2 %.3e" % (ic, err))
    # with the supplementary stop test (tol_gamma 1.1, as the Mn fixtures):
    # every deck at the reference's criterion
    assert err < 1e-10
    dx, dy = (h[4] - h[2]) / h[0], (h[5] - h[3]) / h[1]
    dm = mass(density(m), dx, dy) - mass(density(m0), dx, dy)
    if ic == "smooth":
        assert abs(dm) < 1e-10
    if ic == "delta":
        assert abs(dm - (mass(density(ref), dx, dy) - mass(density(m0), dx, dy))) < 1e-10
    assert s.stats().num_dual_solves == 84 * 2 * 40 * 78


@pytest.mark.parametrize("gaunt", [False, True])
def test_ppn_derivatives_consistent(gaunt):
    """g is the gradient of f and H the Jacobian of g (central differences),
    and g is the direct quadrature of G m - u."""
    o = oracle()
    t = o.Tables(3, 16, gaunt)
    rng = np.random.default_rng(7)
    alpha = rng.normal(0, 0.5, t.nm)
    alpha[0] = 2.0
    u = rng.normal(0, 0.2, t.nm)
    u[0] = 1.0
    delta = 1e-3
    f, g, h = t.fobj_ppn(alpha, u, delta)
    G = ppn_ansatz(t.p[:, :t.nm] @ alpha, delta)
    np.testing.assert_allclose(g, (t.w * G) @ t.p[:, :t.nm] - u, atol=1e-13)
    eps = 1e-6
    for k in range(t.nm):
        e = np.zeros(t.nm)
        e[k] = eps
        fp, gp, _ = t.fobj_ppn(alpha + e, u, delta)
        fm, gm, _ = t.fobj_ppn(alpha - e, u, delta)
        assert abs((fp - fm) / (2 * eps) - g[k]) < 1e-7
        np.testing.assert_allclose((gp - gm) / (2 * eps), h[:, k], atol=1e-7)


def test_ppn_gaunt_and_direct_hessians_agree():
    o = oracle()
    for N, Q in ((2, 12), (3, 30), (4, 16)):
        t = o.Tables(N, Q, True)
        rng = np.random.default_rng(N)
        alpha = rng.normal(0, 1.0, t.nm)
        alpha[0] = 3.0
        u = np.zeros(t.nm)
        u[0] = 1
        _, _, hg = t.fobj_ppn(alpha, u, 1e-4, use_gaunt=True)
        _, _, hd = t.fobj_ppn(alpha, u, 1e-4, use_gaunt=False)
        np.testing.assert_allclose(hg, hd, rtol=0, atol=1e-13 * np.abs(hd).max())


def test_ppn_isotropic_input():
    """u = (u0, 0, ...) is solved by alpha = (u0 - 4 pi delta / u0, 0, ...)
    (the scaled isotropicAlpha 1 - 4 pi delta, opt.cpp:36-39, times u0)."""
    o = oracle()
    t = o.Tables(3, 30, True)
    for u0 in (1e-3, 1.0, 37.5):
        u = np.zeros(t.nm)
        u[0] = u0
        a_iso = np.zeros(t.nm)
        a_iso[0] = u0 - 4 * math.pi * DELTA / u0
        st, uo, ao, outs = t.opt(u, a_iso, tol=1e-12, moment_type=o.MOMENT_PPN, delta_ppn=DELTA)
        assert st == o.OK and outs.r == 0.0
        np.testing.assert_allclose(ao, a_iso, rtol=1e-12, atol=1e-14 * u0)
        np.testing.assert_array_equal(uo, u)       # PPn has no reconstruction step


@pytest.mark.parametrize("N,Q,gaunt", [(2, 12, False), (3, 30, True), (4, 16, False)])
def test_ppn_known_answer(N, Q, gaunt):
    """u = <m G(alpha*.m)> is solved back to alpha* from the isotropic start.
    At TOL 1e-12 the reference's Armijo test cannot resolve the last decreases
    of f (as for Mn at tight tolerance), so the robust line search runs."""
    o = oracle()
    t = o.Tables(N, Q, gaunt)
    rng = np.random.default_rng(200 + N)
    n = 200
    astar = rng.normal(0, 1.0, (n, t.nm)) / (1.0 + np.arange(t.nm))
    astar[:, 0] = rng.uniform(1.0, 4.0, n)
    P = t.p[:, :t.nm]
    u = (t.w[None, :] * ppn_ansatz(astar @ P.T, DELTA)) @ P
    a0 = np.zeros_like(u)
    a0[:, 0] = u[:, 0] - 4 * math.pi * DELTA / u[:, 0]
    st, uo, ao, it, r, ng = t.opt_batch(u, a0, tol=1e-12, nthreads=4, moment_type=o.MOMENT_PPN, delta_ppn=DELTA,
                                        line_search=1)
    assert st == o.OK
    ok = r == 0.0
    assert ok.all()
    err = np.abs(ao[ok] - astar[ok]).max(axis=1) / (np.abs(astar[ok]).max(axis=1) + 1)
    print("PPn known answer: median %.2e max %.2e, median iterations %d" % (np.median(err), err.max(), np.median(it)))
    assert np.median(err) < 1e-14 and err.max() < 1e-11
    assert np.median(it) <= 6

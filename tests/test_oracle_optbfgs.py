"""The oracle's BFGS variant of the change-of-basis optimiser (optbfgs.cpp:92-545,
MOMOPT_OPTIMIZATION_TYPE 2).  The reference holds no fixture for it, so the
restatement is pinned by what the mathematics fixes:

- the inverse update satisfies the secant equation H+^-1 y = s and equals
  the textbook product form (I - rho s y^T) H^-1 (I - rho y s^T) + rho s s^T
  (a different expression of the same matrix), which fails for a wrong
  sign, term or transpose;
- the reference updates H in two steps and forms the second from the
  first's result (optbfgs.cpp:447-456): H1 = H + y y^T / (s.y), then H1 less
  its rank-one part along H1 s.  So H+ s = 0 rather than y (H enters only
  the cond stop test); the restatement is pinned by that characterisation:
  H+ symmetric, H+ s = 0, and H+ v = H1 v for every v orthogonal to H1 s;
- with one inner iteration per basis the variant is optcb exactly (H = I
  gives d = -g, and the next outer iteration evaluates the full Hessian at
  the same alpha), bit for bit;
- with the decks' settings (5 inner iterations, cond bound 20) it solves the
  same dual problem as opt: the moments of its multipliers reproduce u to
  the stop tolerance, and it needs fewer full-Hessian evaluations than optcb;
- the settings that would take no step are rejected.
"""
import numpy as np
import pytest

from helpers import config2_batch, isotropic_start, oracle


def _moments(tab, alpha):
    pm = tab.p[:, :tab.nm]
    return (np.exp(alpha @ pm.T) * tab.w[None, :]) @ pm


def _start(o, tab, n):
    nm = tab.nm
    return np.tile(np.eye(nm).reshape(1, -1), (n, 1)), np.tile(o.isotropic_alpha(nm), (n, 1))


def test_bfgs_update_secant_and_product_form():
    o = oracle()
    rng = np.random.default_rng(3)
    for n in (1, 3, 6, 15):
        A = rng.standard_normal((n, n))
        H = A @ A.T + n * np.eye(n)
        Hinv = np.linalg.inv(H)
        s = rng.standard_normal(n)
        y = H @ s + 0.3 * rng.standard_normal(n)
        if s @ y <= 0:
            y = -y + 2 * (H @ s)
        assert s @ y > 0
        Hn, In = o.bfgs_update(H, Hinv, s, y)
        np.testing.assert_allclose(In @ y, s, rtol=1e-10, atol=1e-10 * np.abs(s).max())
        rho = 1.0 / (s @ y)
        E = np.eye(n)
        I_prod = (E - rho * np.outer(s, y)) @ Hinv @ (E - rho * np.outer(y, s)) + rho * np.outer(s, s)
        np.testing.assert_allclose(In, I_prod, rtol=1e-10, atol=1e-12 * np.abs(I_prod).max())
        scale = np.abs(H).max() + np.abs(y).max() ** 2 / abs(s @ y)
        np.testing.assert_allclose(Hn, Hn.T, rtol=0, atol=1e-13 * scale)
        np.testing.assert_allclose(Hn @ s, 0.0, atol=1e-12 * scale * np.abs(s).max())
        H1 = H + rho * np.outer(y, y)
        c = H1 @ s
        V = rng.standard_normal((n, 4))
        V -= np.outer(c, c @ V) / (c @ c)                   # (zero for n = 1)
        np.testing.assert_allclose(Hn @ V, H1 @ V, rtol=0, atol=1e-12 * scale)


@pytest.mark.parametrize("line_search", [0, 1])
def test_one_inner_iteration_is_optcb(line_search):
    o = oracle()
    tab = o.Tables(3, 12, False)
    rng = np.random.default_rng(21)
    u, astar, lobe = config2_batch(tab.w, tab.p, tab.xi, tab.eta, tab.nm, 120, rng, near_boundary_frac=0.2)
    P2M, aP = _start(o, tab, len(u))
    cb = tab.optcb_batch(u, P2M, aP, line_search=line_search, nthreads=8)
    bf = tab.optbfgs_batch(u, P2M, aP, line_search=line_search, max_bfgs_iter=1, cond_h_max_bfgs=1e300,
                           nthreads=8)
    assert cb[0] == 0 and bf[0] == 0
    for a, b in zip(cb[1:], bf[1:]):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("order,Q", [(2, 12), (4, 16)])
def test_optbfgs_solves_the_opt_problem(order, Q):
    o = oracle()
    tab = o.Tables(order, Q, False)
    rng = np.random.default_rng(61 + order)
    u, astar, lobe = config2_batch(tab.w, tab.p, tab.xi, tab.eta, tab.nm, 300, rng, near_boundary_frac=0.0)
    P2M, aP = _start(o, tab, len(u))
    tol = 1e-8
    st, uc, B, aPo, aM, it, r, ng = tab.optbfgs_batch(u, P2M, aP, tol=tol, nthreads=8)
    stc, _, _, _, aMc, itc, rc, _ = tab.optcb_batch(u, P2M, aP, tol=tol, nthreads=8)
    sto, uo, ao, ito, ro, _ = tab.opt_batch(u, isotropic_start(u), tol=tol, nthreads=8)
    assert st == 0 and stc == 0 and sto == 0
    ok = (r == 0.0) & (ro == 0.0)
    print("outer iterations: bfgs %.3f optcb %.3f opt %.3f; unregularized %.4f" %
          (it.mean(), itc.mean(), ito.mean(), ok.mean()))
    assert ok.mean() >= 0.98
    np.testing.assert_allclose(uc[ok], u[ok], rtol=1e-15, atol=0)
    res = np.linalg.norm(_moments(tab, aM[ok]) - u[ok], axis=1) / np.linalg.norm(u[ok], axis=1)
    print("||moments(alphaM) - u|| / ||u||: max %.3e" % res.max())
    assert res.max() <= 2 * tol
    rel = np.linalg.norm(aM - ao, axis=1)[ok] / np.linalg.norm(ao, axis=1)[ok]
    print("||alphaM - alpha_opt|| / ||alpha||: median %.3e max %.3e" % (np.median(rel), rel.max()))
    assert np.median(rel) <= 1e-8
    # the quasi-Newton steps replace full-Hessian evaluations
    assert it.mean() < itc.mean()


def test_optbfgs_rejects_settings_without_a_step():
    o = oracle()
    tab = o.Tables(2, 12, False)
    nm = tab.nm
    u = np.zeros((1, nm))
    u[0, 0] = 1.0
    P2M, aP = _start(o, tab, 1)
    assert tab.optbfgs_batch(u, P2M, aP, max_bfgs_iter=0)[0] == o.ERR_ARG
    assert tab.optbfgs_batch(u, P2M, aP, cond_h_max_bfgs=1.0)[0] == o.ERR_ARG
    assert o.Tables(2, 12, True).optbfgs_batch(u, P2M, aP)[0] == o.ERR_ARG
    assert tab.optbfgs_batch(u, P2M, aP)[0] == 0

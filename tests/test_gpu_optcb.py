"""The change-of-basis optimiser on the GPU (mnc_optcb_batch_host) against the
oracle's restatement (orc_optcb, itself pinned in tests/test_oracle_optcb.py).

The n x n algebra runs in the reference's order on both sides; the quadrature
sums differ in order (lane partials and wave sums on the GPU, sequential in the
oracle), so the two follow the same Newton path up to rounding: the checks are
identical (iterations, regularization) for nearly every cell, the multipliers
and bases to rounding on those cells, and every cell solved to the tolerance."""
import numpy as np
import pytest

from helpers import config2_batch, load_pkg, oracle


def _moments(tab, alpha):
    pm = tab.p[:, :tab.nm]
    return (np.exp(alpha @ pm.T) * tab.w[None, :]) @ pm


def _identity_start(o, nm, n):
    P2M = np.tile(np.eye(nm).reshape(1, -1), (n, 1))
    aP = np.tile(o.isotropic_alpha(nm), (n, 1))
    return P2M, aP


def _run(order, Q, n, seed, nb_frac, P2M=None, aP=None, u=None, line_search=0):
    mnc = load_pkg("mnc")
    o = oracle()
    print("case optcb order=%d Q=%d n=%d seed=%d lobes=%.2f ls=%d" % (order, Q, n, seed, nb_frac, line_search),
          flush=True)
    tab = o.Tables(order, Q, False)
    if u is None:
        rng = np.random.default_rng(seed)
        u, astar, lobe = config2_batch(tab.w, tab.p, tab.xi, tab.eta, tab.nm, n, rng, near_boundary_frac=nb_frac)
    if P2M is None:
        P2M, aP = _identity_start(o, tab.nm, len(u))
    ctx = mnc.Context(order, Q, use_gaunt=False)
    opt = mnc.options(max_iter=100, tol=1e-8, cond_h_max=1e10, line_search=line_search)
    g = ctx.optcb_batch_host(u, P2M, aP, opt)
    ctx.close()
    c = tab.optcb_batch(u, P2M, aP, tol=1e-8, nthreads=8, line_search=line_search)
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

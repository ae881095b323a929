"""Pins the oracle's objective, gradient and Hessian (fobj.cpp:37-129,255-335)."""
import math

import mpmath
import numpy as np
import pytest

from helpers import oracle


def test_isotropic_closed_form():
    """alpha = (a0, 0, ...): f = 4 pi e^{a0 m0} - a0 u0, g_k = 0 for k > 0,
    g_0 = 4 pi m0 e^{a0 m0} - u0."""
    o = oracle()
    t = o.Tables(3, 30, True)
    m0 = 0.5 / math.sqrt(math.pi)
    for a0 in (-3.0, 0.4, 2.0):
        alpha = np.zeros(t.nm)
        alpha[0] = a0
        u = np.zeros(t.nm)
        u[0] = 1.3
        f, g, h = t.fobj(alpha, u)
        e = math.exp(a0 * m0)
        assert abs(f - (4 * math.pi * e - a0 * 1.3)) < 1e-13 * abs(f)
        assert abs(g[0] - (4 * math.pi * m0 * e - 1.3)) < 1e-13
        assert np.abs(g[1:]).max() < 1e-14
        np.testing.assert_allclose(h, e * np.eye(t.nm), atol=1e-13 * e)


def test_brute_force_mpmath():
    """f, g, H against 50-digit sums over the same nodes."""
    o = oracle()
    t = o.Tables(2, 8, False)
    rng = np.random.default_rng(11)
    alpha = rng.normal(0, 0.7, t.nm)
    u = rng.normal(0, 0.3, t.nm)
    u[0] = 1.0
    mpmath.mp.dps = 50
    F = mpmath.mpf(0)
    G = [mpmath.mpf(0)] * t.nm
    H = [[mpmath.mpf(0)] * t.nm for _ in range(t.nm)]
    for q in range(t.nq):
        a = sum(mpmath.mpf(alpha[i]) * mpmath.mpf(t.p[q, i]) for i in range(t.nm))
        e = mpmath.mpf(t.w[q]) * mpmath.exp(a)
        F += e
        for i in range(t.nm):
            G[i] += e * mpmath.mpf(t.p[q, i])
            for j in range(t.nm):
                H[i][j] += e * mpmath.mpf(t.p[q, i]) * mpmath.mpf(t.p[q, j])
    F -= sum(mpmath.mpf(alpha[i]) * mpmath.mpf(u[i]) for i in range(t.nm))
    for direct_f in (True, False):
        f, g, h = t.fobj(alpha, u, want_h=not direct_f)
        assert abs(f - float(F)) < 1e-14 * abs(float(F))
        for i in range(t.nm):
            assert abs(g[i] - float(G[i] - mpmath.mpf(u[i]))) < 1e-14
    f, g, h = t.fobj(alpha, u)
    for i in range(t.nm):
        for j in range(t.nm):
            assert abs(h[i, j] - float(H[i][j])) < 1e-14 * float(H[0][0])


@pytest.mark.parametrize("gaunt", [False, True])
def test_derivatives_consistent(gaunt):
    """g is the gradient of f and H the Jacobian of g (central differences)."""
    o = oracle()
    t = o.Tables(3, 16, gaunt)
    rng = np.random.default_rng(5)
    alpha = rng.normal(0, 0.5, t.nm)
    u = rng.normal(0, 0.2, t.nm)
    u[0] = 1.0
    f, g, h = t.fobj(alpha, u)
    eps = 1e-6
    for k in range(t.nm):
        e = np.zeros(t.nm)
        e[k] = eps
        fp, gp, _ = t.fobj(alpha + e, u)
        fm, gm, _ = t.fobj(alpha - e, u)
        assert abs((fp - fm) / (2 * eps) - g[k]) < 1e-7
        np.testing.assert_allclose((gp - gm) / (2 * eps), h[:, k], atol=1e-7)


def test_gaunt_and_direct_hessians_agree():
    o = oracle()
    for N, Q in ((2, 12), (3, 30), (4, 16)):
        tg = o.Tables(N, Q, True)
        rng = np.random.default_rng(N)
        alpha = rng.normal(0, 1.0, tg.nm)
        u = np.zeros(tg.nm)
        u[0] = 1
        _, gg, hg = tg.fobj(alpha, u, use_gaunt=True)
        _, gd, hd = tg.fobj(alpha, u, use_gaunt=False)
        np.testing.assert_allclose(hg, hd, rtol=0, atol=1e-13 * np.abs(hd).max())
        np.testing.assert_allclose(gg, gd, atol=1e-13 * np.abs(hd).max())

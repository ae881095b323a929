"""Pins the oracle's quadrature, spherical-harmonic table and Gaunt map
(momopt_init.cpp:88-259, utils.cpp:55-60) against closed forms and
independent library routines."""
import math

import numpy as np
import pytest
import scipy.special

from helpers import oracle


@pytest.mark.parametrize("n", [2, 5, 12, 16, 30, 32, 64])
def test_gauss_legendre_vs_numpy(n):
    o = oracle()
    x, w = o.gauss_legendre(n)
    xr, wr = np.polynomial.legendre.leggauss(n)   # ascending nodes
    assert np.all(np.diff(x) > 0)
    np.testing.assert_allclose(x, xr, rtol=0, atol=4e-16)
    np.testing.assert_allclose(w, wr, rtol=2e-12, atol=0)   # numpy's weights carry ~1e-12 error
    # weights against 40-digit Newton iterates on P_n (mpmath)
    import mpmath
    mpmath.mp.dps = 40
    for i in range(n // 2):
        z = mpmath.mpf(-x[i])
        for _ in range(6):
            p, pm1 = mpmath.legendre(n, z), mpmath.legendre(n - 1, z)
            dp = n * (z * p - pm1) / (z * z - 1)
            z -= p / dp
        dp = n * (z * mpmath.legendre(n, z) - mpmath.legendre(n - 1, z)) / (z * z - 1)
        wex = float(2 / ((1 - z * z) * dp * dp))
        assert abs(float(z) + x[i]) <= 2e-16
        assert abs(w[i] - wex) <= 2e-13 * wex


@pytest.mark.parametrize("N,Q,gaunt", [(1, 8, False), (2, 12, True), (3, 30, True), (4, 16, True), (8, 32, True)])
def test_quadrature_integrates_sphere_and_orthonormal_basis(N, Q, gaunt):
    o = oracle()
    t = o.Tables(N, Q, gaunt)
    assert t.nq == Q * Q
    assert abs(t.w.sum() - 4 * math.pi) < 1e-13
    # the basis is orthonormal under the quadrature (exact for degree <= 2N' < 2Q)
    G = (t.p * t.w[:, None]).T @ t.p
    np.testing.assert_allclose(G, np.eye(t.np2), atol=5e-14)
    # directions are unit vectors in the mu < 0 hemisphere
    assert np.all(t.xi ** 2 + t.eta ** 2 < 1.0)


def test_low_order_harmonics_closed_form():
    """m0 = 1/(2 sqrt(pi)); m1 = sqrt(3/4pi) eta; m2 = -sqrt(3/4pi) xi (Condon-Shortley
    phase with sqrt2 sin(m phi) for m<0, momopt_init.cpp:200-209)."""
    o = oracle()
    t = o.Tables(2, 12, False)
    c = math.sqrt(3.0 / (4.0 * math.pi))
    np.testing.assert_allclose(t.p[:, 0], 0.5 / math.sqrt(math.pi), rtol=1e-15)
    np.testing.assert_allclose(t.p[:, 1], c * t.eta, atol=1e-15)
    np.testing.assert_allclose(t.p[:, 2], -c * t.xi, atol=1e-15)


def test_harmonics_vs_scipy():
    """Every table column equals the real combination of scipy's complex Y_l^m
    (which includes the Condon-Shortley phase), in the reference's order
    n = 0..N, m = -n..n, (n+m) even."""
    o = oracle()
    N, Q = 4, 16
    t = o.Tables(N, Q, False)
    mu = -np.sqrt(1.0 - t.xi ** 2 - t.eta ** 2)
    phi = np.mod(np.arctan2(t.eta, t.xi), 2 * np.pi)
    theta = np.arccos(mu)
    k = 0
    for n in range(N + 1):
        for m in range(-n, n + 1):
            if (m + n) % 2:
                continue
            am = abs(m)
            if hasattr(scipy.special, "sph_harm_y"):
                Y = scipy.special.sph_harm_y(n, am, theta, phi)
            else:
                Y = scipy.special.sph_harm(am, n, phi, theta)
            if m < 0:
                ref = -math.sqrt(2.0) * Y.imag      # sqrt2 sin(m phi) P = -sqrt2 Im(Y_n^|m|)
            elif m > 0:
                ref = math.sqrt(2.0) * Y.real
            else:
                ref = Y.real
            np.testing.assert_allclose(t.p[:, k], ref, atol=2e-14, err_msg="n=%d m=%d" % (n, m))
            k += 1
    assert k == t.nm


def test_sph_plm_special_values():
    o = oracle()
    # P_1^1 with CS phase: -sqrt(3/8pi) sqrt(1-x^2)
    for x in (-0.9, -0.3, 0.2, 0.7):
        assert abs(o.sph_plm(1, 1, x) + math.sqrt(3 / (8 * math.pi)) * math.sqrt(1 - x * x)) < 1e-15
        assert abs(o.sph_plm(2, 0, x) - math.sqrt(5 / (4 * math.pi)) * 0.5 * (3 * x * x - 1)) < 1e-15


@pytest.mark.parametrize("N,Q", [(2, 12), (3, 30), (4, 16)])
def test_gaunt_map_invariant(N, Q):
    """sum_k G_ijk m_k(q) = m_i(q) m_j(q) at every node; nnz counts of the reference's map
    (upper triangle M2 45, M3 167, M4 432)."""
    o = oracle()
    t = o.Tables(N, Q, True)
    nm = t.nm
    upper = 0
    for i in range(nm):
        for j in range(nm):
            idx = i * nm + j
            ks = t.gk[t.gptr[idx]:t.gptr[idx + 1]]
            vs = t.gv[t.gptr[idx]:t.gptr[idx + 1]]
            lhs = t.p[:, ks] @ vs
            np.testing.assert_allclose(lhs, t.p[:, i] * t.p[:, j], atol=1e-13)
            if j >= i:
                upper += len(ks)
    assert upper == {2: 45, 3: 167, 4: 432}[N]


def test_odd_quadrature_order_rejected():
    o = oracle()
    with pytest.raises(ValueError):
        o.Tables(2, 11, False)

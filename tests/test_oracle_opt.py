"""Pins the oracle's closure solve (opt.cpp:82-343, linesearch.cpp:21-99) by
known answers, invariants and each failure branch of the regularization ladder."""
import math

import numpy as np
import pytest

from helpers import config2_batch, isotropic_start, oracle

SQPI = math.sqrt(math.pi)


@pytest.mark.parametrize("N,Q,gaunt", [(2, 12, False), (3, 30, True), (4, 16, False)])
def test_known_answer(N, Q, gaunt):
    """u = <m exp(alpha*.m)> is solved back to alpha* (from the isotropic start)."""
    o = oracle()
    t = o.Tables(N, Q, gaunt)
    rng = np.random.default_rng(100 + N)
    u, astar, _ = config2_batch(t.w, t.p, t.xi, t.eta, t.nm, 200, rng, near_boundary_frac=0.0, scale_k=1.5)
    st, uo, ao, it, r, ng = t.opt_batch(u, isotropic_start(u), tol=1e-11, nthreads=4)
    assert st == o.OK
    ok = r == 0.0
    assert ok.mean() >= 0.95
    err = np.abs(ao[ok] - astar[ok]).max(axis=1) / (np.abs(astar[ok]).max(axis=1) + 1)
    assert np.median(err) < 1e-10 and err.max() < 1e-6
    # quadratic convergence: few iterations even from the isotropic start
    assert np.median(it) <= 12


def test_isotropic_input():
    """u = (u0, 0, ...) gives alpha0 = 2 sqrt(pi) ln(u0 / (2 sqrt(pi))) in one
    iteration (momopt_init.cpp:135-136) and u is returned unchanged to rounding."""
    o = oracle()
    t = o.Tables(3, 30, True)
    for u0 in (1e-4 * 2 * SQPI, 1.0, 37.5):
        u = np.zeros(t.nm)
        u[0] = u0
        a = isotropic_start(u[None, :])[0]
        st, uo, ao, outs = t.opt(u, a, tol=1e-4)
        assert st == o.OK and outs.iter == 1 and outs.r == 0.0
        assert abs(ao[0] - 2 * SQPI * math.log(u0 / (2 * SQPI))) < 1e-12 * (1 + abs(ao[0]))
        assert np.abs(ao[1:]).max() < 1e-14
        assert uo[0] == u0
        assert np.abs(uo[1:]).max() < 1e-15 * u0


def test_reconstruction_and_stop_test():
    """Output u is the ansatz's moments with u[0] preserved exactly; the scaled
    gradient met the stop test ||g|| <= tol ||u/u0||."""
    o = oracle()
    t = o.Tables(4, 16, False)
    rng = np.random.default_rng(3)
    u, _, lobe = config2_batch(t.w, t.p, t.xi, t.eta, t.nm, 300, rng)
    tol = 1e-8
    st, uo, ao, it, r, ng = t.opt_batch(u, isotropic_start(u), tol=tol, nthreads=4)
    assert st == o.OK
    assert np.array_equal(uo[:, 0], u[:, 0])
    pm = t.p[:, :t.nm]
    recon = (np.exp(ao @ pm.T) * t.w[None, :]) @ pm
    np.testing.assert_allclose(recon, uo, rtol=0, atol=1e-11 * np.abs(uo).max())
    ub = u / u[:, :1]
    blend = ub.copy()
    blend[:, 1:] *= (1.0 - r)[:, None]
    assert np.all(ng <= tol * np.linalg.norm(blend, axis=1) * (1 + 1e-12))


def test_nonpositive_density_is_an_error():
    o = oracle()
    t = o.Tables(2, 12, False)
    u = np.zeros(t.nm)
    a = np.zeros(t.nm)
    st, *_ = t.opt(u, a)
    assert st == o.ERR_U0
    u[0] = -1.0
    st, *_ = t.opt(u, a)
    assert st == o.ERR_U0


def test_max_iter_forces_the_whole_ladder():
    """MAX_ITER 1: every level regularizes at once until the last, which gets
    maxIter*10 (opt.cpp:244-273), so outs.r = 0.5 and u is the 50% blend."""
    o = oracle()
    t = o.Tables(2, 12, False)
    rng = np.random.default_rng(8)
    u, _, _ = config2_batch(t.w, t.p, t.xi, t.eta, t.nm, 50, rng, near_boundary_frac=0.0, scale_k=0.5)
    st, uo, ao, it, r, ng = t.opt_batch(u, isotropic_start(u), max_iter=1, tol=1e-6, nthreads=2)
    assert st == o.OK
    assert np.all(r == 0.5)
    assert np.all(it >= 5)           # 4 immediate regularizations + >= 1 iteration
    ub = u / u[:, :1]
    target = ub.copy()
    target[:, 1:] *= 0.5
    dn = np.linalg.norm(uo / uo[:, :1] - target, axis=1) / np.linalg.norm(target, axis=1)
    assert dn.max() < 3e-6


def test_condition_limit_exhausts_the_ladder():
    """COND_H_MAX below 1 regularizes at every iteration: "Max regularization"."""
    o = oracle()
    t = o.Tables(2, 12, False)
    u = np.zeros(t.nm)
    u[0] = 1.0
    u[1] = 0.3
    st, *_ = t.opt(u, np.zeros(t.nm), cond_h_max=0.5)
    assert st == o.ERR_MAXREG


def test_boundary_moments_need_regularization():
    """u = w_q m(Omega_q) (all mass on one node) lies on the boundary of the
    discrete realizable set: the Hessian degenerates as alpha grows, the solve
    fails (Cholesky / condition / line search / iteration limit) and the
    regularization ladder moves u inside (opt.cpp:244-273)."""
    o = oracle()
    t = o.Tables(2, 12, False)
    for q in (0, 17, 100):
        u = t.w[q] * t.p[q, :t.nm]
        st, uo, ao, outs = t.opt(u, isotropic_start(u[None, :])[0], tol=1e-6)
        assert st == o.OK
        assert outs.r in o.REGULARIZATIONS[1:]
        r = outs.r
        blend = u / u[0]
        blend[1:] *= 1.0 - r
        assert np.linalg.norm(uo / uo[0] - blend) <= 3e-6 * np.linalg.norm(blend)


def test_cholesky_and_rcond_kernels():
    o = oracle()
    rng = np.random.default_rng(0)
    for n in (1, 2, 6, 15, 45):
        A = rng.normal(size=(n, n))
        A = A @ A.T + n * np.eye(n)
        L = A.copy()
        assert o.lib().orc_cholesky(n, o._p(L)) == 0
        Lt = np.tril(L)
        np.testing.assert_allclose(Lt @ Lt.T, A, atol=1e-12 * np.abs(A).max())
        anorm = o.lib().orc_norm1(n, o._p(A))
        assert abs(anorm - np.abs(A).sum(axis=1).max()) < 1e-12 * anorm
        rc = o.lib().orc_rcond(n, o._p(L), anorm)
        true_rc = 1.0 / (np.abs(A).sum(axis=1).max() * np.abs(np.linalg.inv(A)).sum(axis=0).max())
        # Higham's estimator never overestimates ||A^-1||_1 and is usually within 3x
        assert true_rc * (1 - 1e-12) <= rc <= 3 * true_rc
    B = np.diag([1.0, -1.0])
    assert o.lib().orc_cholesky(2, o._p(B)) == 2


def test_rcond_estimate_within_exact_norm_when_ill_conditioned():
    """The tail skips the condition estimate when 2 ||A||_1 ||A^-1||_1 is
    within the limit, with ||A^-1||_1 computed from the factor: that relies on
    the estimate (1 / (rcond ||A||_1)) never exceeding ||A^-1||_1 by more than
    rounding, up to the regularization threshold kappa = 1e10.  Checked on
    SPD matrices of order 15 with spectra spread over 1e2 .. 1e10 and a
    spectrum with one isolated small eigenvalue, against the exact norm in
    long double from the eigendecomposition."""
    o = oracle()
    rng = np.random.default_rng(7)
    worst = 0.0
    n = 15
    for trial in range(300):
        kappa = 10.0 ** rng.uniform(2, 10)
        Q, _ = np.linalg.qr(rng.normal(size=(n, n)))
        if trial % 3 == 0:
            lam = np.ones(n)
            lam[rng.integers(n)] = 1.0 / kappa
        else:
            lam = np.exp(rng.uniform(0.0, np.log(kappa), n))
            lam[0], lam[1] = 1.0, kappa
        lam = lam * 10.0 ** rng.uniform(-3, 3)
        A = (Q * lam) @ Q.T
        A = 0.5 * (A + A.T)
        L = A.copy()
        if o.lib().orc_cholesky(n, o._p(L)) != 0:
            continue
        anorm = o.lib().orc_norm1(n, o._p(A))
        rc = o.lib().orc_rcond(n, o._p(L), anorm)
        Ql = Q.astype(np.longdouble)
        inv = (Ql / lam.astype(np.longdouble)) @ Ql.T
        exact = float(np.abs(inv).sum(axis=0).max())
        est = 1.0 / (rc * anorm)
        worst = max(worst, est / exact)
    print("largest estimate / exact ||A^-1||_1: %.6f" % worst)
    assert 0.0 < worst <= 1.0 + 1e-4


def test_cholesky_and_rcond_against_lapack():
    """orc_cholesky and orc_rcond against LAPACK's own dpotrf ('L') and dpocon
    (scipy's Fortran LAPACK; the reference calls them at opt.cpp:231-237):
    SPD matrices of orders 1-45 with condition numbers up to 3e10 (past the
    regularization threshold 1e10) and indefinite ones.  The factor agrees to
    rounding amplified by the conditioning, the reciprocal condition estimate
    to 1e-6 relative (the same sign vectors and iterations of Higham's
    estimator; only the rounding of the triangular solves differs), and
    failures report the same column."""
    from scipy.linalg import lapack
    o = oracle()
    rng = np.random.default_rng(11)
    worst_c = worst_r = 0.0
    checked = 0
    for trial in range(1800):
        n = (1, 3, 6, 10, 15, 45)[trial % 6]
        kappa = 10.0 ** rng.uniform(0, 10.5)
        Q, _ = np.linalg.qr(rng.normal(size=(n, n)))
        lam = np.exp(rng.uniform(0.0, np.log(kappa), n))
        lam[0] = 1.0
        if n > 1:
            lam[1] = kappa
        lam *= 10.0 ** rng.uniform(-3, 3)
        if trial % 10 == 9 and n > 1:
            lam[rng.integers(n)] *= -1.0          # indefinite
        A = (Q * lam) @ Q.T
        A = 0.5 * (A + A.T)
        L = A.copy()
        st = o.lib().orc_cholesky(n, o._p(L))
        c, info = lapack.dpotrf(A, lower=1, clean=1)
        if info != 0 or st != 0:
            assert st == info, (n, st, info)
            continue
        checked += 1
        worst_c = max(worst_c, np.abs(np.tril(L) - c).max() / np.abs(c).max())
        anorm = o.lib().orc_norm1(n, o._p(A))
        rc = o.lib().orc_rcond(n, o._p(L), anorm)
        rcl, info2 = lapack.dpocon(np.asfortranarray(c), anorm, uplo=b"L")
        assert info2 == 0
        worst_r = max(worst_r, abs(rc - rcl) / rcl)
    print("checked %d: factor max rel diff %.2e, rcond max rel diff %.2e" % (checked, worst_c, worst_r))
    assert checked > 1000
    assert worst_c < 1e-9
    assert worst_r < 1e-6

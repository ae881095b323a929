"""The oracle's Sn discrete-ordinates step (kinetic/kinetic_update.cpp:41-235,
kinetic_boundaries.cpp:190-223), the independent truth for the moment
closures.  Beside the reference's four kinetic fixtures
(tests/test_oracle_sn_fixtures.py), the restatement is pinned by what the
mathematics fixes:

- a uniform isotropic state with sigma_t = sigma_s on the periodic grid is
  steady (the ordinates' weights sum to 4 pi, the fluxes vanish);
- pure absorption of a uniform state follows Heun's factor
  1 - dt s + (dt s)^2 / 2 per step;
- on the periodic grid with sigma_t = sigma_s per cell, the total
  sum_cells sum_q w_q f_q is conserved (the face fluxes telescope), which
  fails for a face taken from the wrong cell;
- without collisions a smooth profile in one ordinate advects at its speed
  xi (or eta): compared with the exact translation, which fails for a wrong
  upwind side, sign or speed.
"""
import numpy as np

from helpers import oracle


def _quad(Q):
    t = oracle().Tables(1, Q, False)
    return t.w, t.xi, t.eta


def _grid(nx, ny, nq, val=1.0):
    return np.full((nx + 4, ny + 4, nq), val)


def test_sn_uniform_steady_state_and_absorption():
    o = oracle()
    w, xi, eta = _quad(8)
    nq = len(w)
    assert abs(w.sum() - 4 * np.pi) < 1e-13
    nx, ny = 6, 5
    f = _grid(nx, ny, nq, 0.7)
    s1 = np.ones((nx + 4, ny + 4))
    out = o.sn_steps(f, w, xi, eta, s1, s1, 2, -1.0, -1.0, 0.1, 0.1, 0.02, steps=5)
    np.testing.assert_allclose(out, 0.7, rtol=1e-13)
    sig, dt = 3.0, 0.01
    out = o.sn_steps(f, w, xi, eta, sig * s1, 0 * s1, 2, -1.0, -1.0, 0.1, 0.1, dt, steps=4)
    expect = 0.7 * (1 - dt * sig + 0.5 * (dt * sig) ** 2) ** 4
    np.testing.assert_allclose(out[2:-2, 2:-2], expect, rtol=1e-13)


def test_sn_periodic_mass_conservation():
    o = oracle()
    w, xi, eta = _quad(6)
    nq = len(w)
    nx, ny = 12, 10
    rng = np.random.default_rng(4)
    f = 1.0 + rng.random((nx + 4, ny + 4, nq))
    sig = 0.5 + rng.random((nx + 4, ny + 4))
    def mass(g):
        return (g[2:-2, 2:-2] * w).sum()
    m0 = mass(f)
    out = o.sn_steps(f, w, xi, eta, sig, sig, 2, -1.0, -1.0, 2.0 / nx, 2.0 / ny, 0.02, steps=6)
    print("mass drift %.3e" % (abs(mass(out) - m0) / m0))
    assert abs(mass(out) - m0) <= 1e-13 * m0
    assert not np.allclose(out, f)


def test_sn_advection_at_ordinate_speed():
    """sigma = 0, periodic, f_q(x) = 2 + sin(2 pi x / L) in every ordinate:
    after time T each ordinate is the profile moved by xi_q T (second-order
    scheme with a limiter: agreement to a few 1e-3 at 128 cells)."""
    o = oracle()
    w, xi, eta = _quad(4)
    nq = len(w)
    nx, ny, L = 128, 4, 2.0
    dx, dy = L / nx, L / ny
    x = -1.0 + (np.arange(nx + 4) - 2 + 0.5) * dx
    f = np.empty((nx + 4, ny + 4, nq))
    f[:] = (2.0 + np.sin(2 * np.pi * x / L))[:, None, None]
    z = np.zeros((nx + 4, ny + 4))
    dt, steps = 0.2 * dx, 40
    out = o.sn_steps(f, w, xi, eta, z, z, 2, -1.0, -1.0, dx, dy, dt, steps=steps)
    T = dt * steps
    xs = x[2:-2]
    err = 0.0
    for q in range(nq):
        exact = 2.0 + np.sin(2 * np.pi * (xs - xi[q] * T) / L)
        err = max(err, np.abs(out[2:-2, 2, q] - exact).max())
        # the y direction sees a uniform field: eta plays no part
        np.testing.assert_allclose(out[2:-2, 3, q], out[2:-2, 2, q], rtol=0, atol=1e-14)
    print("max |f - exact translation| = %.3e" % err)
    assert err < 5e-3
    # the same along y with the ordinates' eta
    g = np.transpose(f, (1, 0, 2)).copy()
    out = o.sn_steps(g, w, xi, eta, z.T.copy(), z.T.copy(), 2, -1.0, -1.0, dy, dx, dt, steps=steps)
    err = 0.0
    for q in range(nq):
        exact = 2.0 + np.sin(2 * np.pi * (xs - eta[q] * T) / L)
        err = max(err, np.abs(out[2, 2:-2, q] - exact).max())
    assert err < 5e-3

"""The Sn discrete-ordinates solver on the GPU (mnc_sn_run) against the oracle's
restatement (orc_sn_step, pinned in tests/test_oracle_sn.py).  The GPU sums
the scattering integral in a different order than the oracle's sequential
loop and contracts products into FMAs, so the states agree to rounding."""
import numpy as np
import pytest

from helpers import load_pkg, oracle


def _quad(Q):
    t = oracle().Tables(1, Q, False)
    return t.w, t.xi, t.eta


def _case(nx, ny, Q, seed, init_cond, steps, dx, dy, init_x, init_y, dt):
    mnc = load_pkg("mnc")
    print("case sn nx=%d ny=%d Q=%d seed=%d ic=%d steps=%d" % (nx, ny, Q, seed, init_cond, steps), flush=True)
    rng = np.random.default_rng(seed)
    nq = Q * Q
    f = 0.5 + rng.random((nx + 4, ny + 4, nq))
    ss = rng.random((nx + 4, ny + 4))
    st = ss + rng.random((nx + 4, ny + 4))
    w, xi, eta = _quad(Q)
    g, ms = mnc.sn_run(f, st, ss, Q, dx, dy, init_x, init_y, init_cond, dt, steps)
    o = oracle().sn_steps(f, w, xi, eta, st, ss, init_cond, init_x, init_y, dx, dy, dt, steps=steps)
    rel = np.abs(g - o).max() / np.abs(o).max()
    rin = np.abs(g - o)[2:-2, 2:-2].max() / np.abs(o).max()
    print("max |gpu - oracle| / max|f| = %.3e, interior %.3e (kernel %.3f ms)" % (rel, rin, ms))
    return f, g, o, rel


@pytest.mark.gpu
def test_sn_periodic_matches_oracle():
    f, g, o, rel = _case(20, 16, 8, 1, 2, 6, 0.1, 0.125, -1.0, -1.0, 0.02)
    assert rel <= 1e-13
    assert not np.allclose(g, f)


@pytest.mark.gpu
def test_sn_centre_source_fixed_ghosts_matches_oracle():
    """init_cond 1: the ghosts keep their values and the cells within
    |x|, |y| < 0.5 receive the unit source."""
    f, g, o, rel = _case(16, 18, 6, 2, 1, 5, 0.2, 0.2, -1.7, -1.9, 0.03)
    assert rel <= 1e-13
    np.testing.assert_array_equal(g[:2], f[:2])
    np.testing.assert_array_equal(g[:, -2:], f[:, -2:])


@pytest.mark.gpu
def test_sn_mass_conservation_at_size():
    """128 x 128 cells, Q = 16 (256 ordinates), sigma_t = sigma_s per cell,
    periodic: the total sum_q w_q f_q stays to rounding over 20 steps."""
    mnc = load_pkg("mnc")
    print("case sn mass 128x128 Q=16", flush=True)
    nx = ny = 128
    Q = 16
    rng = np.random.default_rng(9)
    f = 1.0 + rng.random((nx + 4, ny + 4, Q * Q))
    sig = 0.5 + rng.random((nx + 4, ny + 4))
    w = _quad(Q)[0]
    m0 = (f[2:-2, 2:-2] * w).sum()
    g, ms = mnc.sn_run(f, sig, sig, Q, 2.0 / nx, 2.0 / ny, -1.0, -1.0, 2, 0.004, 20)
    m1 = (g[2:-2, 2:-2] * w).sum()
    upd = 2.0 * 20 * nx * ny * Q * Q
    print("mass drift %.3e; %.3f ms for 20 steps, %.3e cell-ordinate stage updates/s" %
          (abs(m1 - m0) / m0, ms, upd / (ms * 1e-3)))
    assert abs(m1 - m0) <= 1e-12 * m0


@pytest.mark.gpu
def test_sn_rejects_bad_sizes():
    mnc = load_pkg("mnc")
    f = np.ones((8, 8, 25))
    s = np.ones((8, 8))
    with pytest.raises(mnc.MncError) as e:
        mnc.sn_run(f, s, s, 5, 0.1, 0.1, -1.0, -1.0, 2, 0.01, 1)     # odd Q
    assert e.value.code == mnc.ERR_ARG
    with pytest.raises(ValueError):
        mnc.sn_run(np.ones((8, 8, 16)), s, s, 5, 0.1, 0.1, -1.0, -1.0, 2, 0.01, 1)


@pytest.mark.gpu
def test_closures_converge_towards_sn_truth():
    """The use the Sn solver serves (a truth for convergence studies): the
    smooth periodic deck (36 x 74 cells, sigma 1.111, Q = 30) to t ~ 0.5 with
    the same dt, minmod MUSCL scheme and quadrature, so the runs differ only
    in the angular model.  The M3 closure's density lies much closer to
    S30's than the M1 closure's (measured: rel-L2 3.2e-2 against 2.4e-1)."""
    import math
    from helpers import density, fixture_deck, gpu_grid_from_deck, mn_options, rel_l2
    mnc, gi = load_pkg("mnc"), load_pkg("grid_init")
    print("case closures vs Sn truth, smooth deck", flush=True)
    dens = {}
    for order in (1, 3):
        d = fixture_deck("smooth", moment_order=order, use_gaunt=0, tol=1e-8, line_search=mnc.LS_ROBUST)
        ctx, g = gpu_grid_from_deck(mnc, gi, d)
        steps = int(round(0.5 / g.max_dt))
        opt = mn_options(mnc, d)
        for _ in range(steps):
            ctx.update(g.max_dt, opt)
        m, _ = ctx.grid_get()
        ctx.close()
        dens[order] = density(m)[2:-2, 2:-2]
        if order == 1:
            Q = d["quad_order"]
            f0 = np.repeat((g.moments[:, :, 0] / (2.0 * math.sqrt(math.pi)))[:, :, None], Q * Q, axis=2)
            f, ms = mnc.sn_run(f0, g.sigma_t, g.sigma_s, Q, g.dx, g.dy, g.init_x, g.init_y, d["init_cond"],
                               g.max_dt, steps)
            w = _quad(Q)[0]
            dens["sn"] = ((f * w).sum(axis=2) / (4.0 * math.pi))[2:-2, 2:-2]
            print("%d steps of dt %.4g; Sn %.2f ms" % (steps, g.max_dt, ms))
    e1, e3 = rel_l2(dens[1], dens["sn"]), rel_l2(dens[3], dens["sn"])
    print("density rel-L2 against S30: M1 %.3e, M3 %.3e" % (e1, e3))
    assert e3 < 0.25 * e1
    assert e3 < 5e-2


@pytest.mark.gpu
@pytest.mark.parametrize("ic", ["gaussian", "smooth", "delta", "lattice"])
def test_sn_fixture_regression(ic):
    """mnc_sn_run on the reference's kinetic regression deck (36 x 74, Q = 12,
    CFL 0.9, 84 steps including the final partial one, to t = 1.03) against
    the reference's own .sn fixture: density with the file's weights at the
    reference's 1e-10, and every ordinate."""
    from test_oracle_sn_fixtures import FIXTURE_IC, SN_Q, read_sn, sn_density, sn_fixture_case, sn_fixture_path
    from helpers import rel_l2
    mnc = load_pkg("mnc")
    print("case sn fixture ic=%s" % ic, flush=True)
    g, t, f, steps = sn_fixture_case(ic)
    full = [dt for dt in steps if dt == steps[0]]
    assert len(full) == len(steps) - 1
    f, _ = mnc.sn_run(f, g.sigma_t, g.sigma_s, SN_Q, g.dx, g.dy, g.init_x, g.init_y, FIXTURE_IC[ic][0], steps[0],
                      len(full))
    f, _ = mnc.sn_run(f, g.sigma_t, g.sigma_s, SN_Q, g.dx, g.dy, g.init_x, g.init_y, FIXTURE_IC[ic][0], steps[-1], 1)
    h, w, ref = read_sn(sn_fixture_path(ic))
    fin = f[2:-2, 2:-2]
    err = rel_l2(sn_density(fin, w), sn_density(ref, w))
    per = np.abs(fin - ref).max() / np.abs(ref).max()
    print("GPU Sn %s: density rel L2 %.3e, per-ordinate max rel %.3e" % (ic, err, per))
    assert err < 1e-10
    assert per < 1e-12

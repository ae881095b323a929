"""The PPn closure on the GPU (FP64 lane kernels and the cooperative tail;
fobj.cpp:137-236,355-428, opt.cpp:113-141) against the oracle: isolated
batches built from known multipliers, solved through the lockstep rounds and
through the tail alone, and the reference's four PPn regression fixtures
(tests/regression/momopt/<ic>/ppn.pn) through the resident grid, whose flux
uses the PPn ansatz (momopt_update.cpp:54-57)."""
import math
import os

import numpy as np
import pytest

from helpers import (GOLDEN, density, fixture_deck, gpu_grid_from_deck, load_pkg, mass, mn_options, oracle,
                     read_pn, rel_l2)

DELTA = 1e-10


def ppn_ansatz(a, delta):
    t = 0.5 * a
    s = np.sqrt(t * t + delta)
    return np.where(t < 0, -delta / (t - s), t + s)


def _batch(t, n, seed):
    rng = np.random.default_rng(seed)
    astar = rng.normal(0, 1.0, (n, t.nm)) / (1.0 + np.arange(t.nm))
    astar[:, 0] = rng.uniform(0.5, 4.0, n)
    P = t.p[:, :t.nm]
    u = (t.w[None, :] * ppn_ansatz(astar @ P.T, DELTA)) @ P
    a0 = np.zeros_like(u)
    a0[:, 0] = u[:, 0] - 4 * math.pi * DELTA / u[:, 0]      # the isotropic start of each u0
    return u, a0, astar


@pytest.mark.gpu
@pytest.mark.parametrize("N,Q,gaunt,rounds", [(2, 12, False, 6), (3, 30, True, 6), (4, 16, False, 6),
                                              (4, 16, False, 0), (3, 30, True, 0)])
def test_ppn_batch_matches_oracle(N, Q, gaunt, rounds):
    mnc = load_pkg("mnc")
    print("case PPn batch N=%d Q=%d gaunt=%s rounds=%d" % (N, Q, gaunt, rounds), flush=True)
    t = oracle().Tables(N, Q, gaunt)
    u, a0, astar = _batch(t, 1500, 300 + N)
    tol = 1e-8
    ctx = mnc.Context(N, Q, use_gaunt=gaunt, moment_type=mnc.MOMENT_PPN)
    ctx.tune("rounds", rounds)
    opt = mnc.options(max_iter=100, tol=tol, cond_h_max=1e10, line_search=mnc.LS_ROBUST, delta_ppn=DELTA)
    st, ug, ag, itg, rg, ngg = ctx.opt_batch_host(u, a0, opt)
    ctx.close()
    sto, uo, ao, ito, ro, ngo = t.opt_batch(u, a0, tol=tol, nthreads=8, line_search=1,
                                            moment_type=oracle().MOMENT_PPN, delta_ppn=DELTA)
    assert st == mnc.OK and sto == oracle().OK
    same_r = rg == ro
    same_path = same_r & (itg == ito)
    print("same r %.4f, same iterations %.4f, mean iterations %.3f (oracle %.3f)"
          % (same_r.mean(), same_path.mean(), itg.mean(), ito.mean()))
    assert same_r.mean() >= 0.99
    assert same_path.mean() >= 0.9
    ok = same_r & (rg == 0)
    # PPn returns the moments as (u / u0) * u0, the same operations on both sides
    np.testing.assert_array_equal(ug[ok], uo[ok])
    scale = np.abs(ao[ok]).max(axis=1) + 1.0
    err = np.abs(ag[ok] - ao[ok]).max(axis=1) / scale
    kerr = np.abs(ag[ok] - astar[ok]).max(axis=1) / scale
    print("vs oracle: median %.2e max %.2e; vs known alpha: median %.2e" % (np.median(err), err.max(), np.median(kerr)))
    assert np.median(err) < 1e-10
    assert err.max() < 1e-4
    assert np.median(kerr) < 1e-6


@pytest.mark.gpu
@pytest.mark.parametrize("ic", ["gaussian", "smooth", "delta", "lattice"])
def test_ppn_fixture_regression(ic):
    """The reference's PPn regression decks (M3 Q=30 Gaunt, delta 1e-10, 36x74
    to t = 1.03, tol_gamma 1.1) through mnc_update, against the oracle and the
    fixture at the reference's 1e-10, with the mass change of delta and smooth."""
    mnc, gi = load_pkg("mnc"), load_pkg("grid_init")
    print("case PPn fixture ic=%s" % ic, flush=True)
    d = fixture_deck(ic, moment_type=1, delta_ppn=DELTA)
    ctx, g = gpu_grid_from_deck(mnc, gi, d)
    opt = mn_options(mnc, d)
    m0 = density(ctx.grid_get()[0][2:-2, 2:-2])
    for dt in gi.time_steps(g.max_dt, 1.03):
        ctx.update(dt, opt)
    m = ctx.grid_get()[0][2:-2, 2:-2]
    st = ctx.stats()
    ctx.close()
    _, ref = read_pn(os.path.join(GOLDEN, "momopt", ic, "ppn.pn"))
    s = oracle().Solver(d, nthreads=8)
    s.run(1.03)
    mo = s.moments()
    vs_oracle = rel_l2(density(m), density(mo))
    vs_fixture = rel_l2(density(m), density(ref))
    print("PPn rel L2 vs oracle %.3e, vs fixture %.3e, max iterations %d" % (vs_oracle, vs_fixture, st.max_iter))
    # with the supplementary stop test the fixtures were written with
    # (tol_gamma 1.1) every deck is at rounding level, including the lattice,
    # whose no-FMA and FMA oracle builds otherwise differ by 1.3e-8
    # (test_oracle_sensitivity.py::test_ppn_lattice_fma_sensitivity)
    assert vs_oracle < 1e-10
    assert vs_fixture < 1e-10
    dmass = mass(density(m), g.dx, g.dy) - mass(m0, g.dx, g.dy)
    if ic == "smooth":
        assert abs(dmass) < 1e-10
    if ic == "delta":
        assert abs(dmass - (mass(density(ref), g.dx, g.dy) - mass(m0, g.dx, g.dy))) < 1e-10
    assert st.num_dual_solves == 84 * 2 * 40 * 78


@pytest.mark.gpu
def test_ppn_rejects_unsupported_forms():
    """PPn runs the FP64 form: mixed precision, the learned start and a
    non-positive delta are argument errors, not silent fallbacks."""
    mnc = load_pkg("mnc")
    ctx = mnc.Context(4, 16, moment_type=mnc.MOMENT_PPN)
    u = np.zeros((4, 15))
    u[:, 0] = 1.0
    a0 = np.zeros_like(u)
    a0[:, 0] = 1.0
    for opt in (mnc.options(delta_ppn=DELTA, precision=mnc.PREC_MIXED),
                mnc.options(delta_ppn=DELTA, warm_start=mnc.WARM_LEARNED),
                mnc.options(delta_ppn=0.0)):
        st = ctx.opt_batch_host(u, a0, opt, allow_status=(mnc.OK, mnc.ERR_ARG))[0]
        assert st == mnc.ERR_ARG
    st = ctx.opt_batch_host(u, a0, mnc.options(delta_ppn=DELTA))[0]
    assert st == mnc.OK
    ctx.close()

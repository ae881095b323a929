"""Mixed-precision closure (MNC_PREC_MIXED) on the GPU against the FP64 oracle.

The mixed path keeps f, g, the stop test, the Armijo test and the
reconstruction in FP64 and takes only the Newton direction from a Hessian built
on the matrix cores (bf16 triples with FP64 sums per 32 points, ~1e-7
relative; or, with mix_split 2, bf16 hi/lo pairs with FP32 sums, ~1e-5).  Its answers
are therefore solutions of the same problem to the same tolerance as the
reference's (||g|| <= tol ||u||), not bitwise equal to them: the checks are the
tolerance-level ones the FP64 path also meets against the oracle when rounding
sends a cell down a different Newton path."""
import numpy as np
import pytest

from helpers import config2_batch, isotropic_start, load_pkg, oracle

CASES = [(2, 12, False), (3, 16, True), (4, 16, False), (4, 16, True)]


def _run(order, Q, gaunt, n, tol, seed, nb_frac=0.2, tune=None, line_search=0):
    mnc = load_pkg("mnc")
    o = oracle()
    print("case mixed order=%d Q=%d gaunt=%d n=%d tol=%g seed=%d tune=%s" % (order, Q, gaunt, n, tol, seed, tune),
          flush=True)
    ctx = mnc.Context(order, Q, use_gaunt=gaunt)
    assert ctx.get_tune("has_mixed") == 1
    for k, v in (tune or {}).items():
        ctx.tune(k, v)
    tab = o.Tables(order, Q, gaunt)
    rng = np.random.default_rng(seed)
    u, astar, lobe = config2_batch(tab.w, tab.p, tab.xi, tab.eta, tab.nm, n, rng, near_boundary_frac=nb_frac)
    a0 = isotropic_start(u)
    ctx.reset_stats()
    opt = mnc.options(max_iter=100, tol=tol, cond_h_max=1e10, precision=mnc.PREC_MIXED, line_search=line_search)
    st, ug, ag, itg, rg, ngg = ctx.opt_batch_host(u, a0, opt)
    stats = ctx.stats()
    sto, uo, ao, ito, ro, ngo = tab.opt_batch(u, a0, max_iter=100, tol=tol, cond_h_max=1e10, nthreads=8,
                                              line_search=line_search)
    ctx.close()
    return dict(u=u, astar=astar, lobe=lobe, ug=ug, ag=ag, itg=itg, rg=rg, ngg=ngg, uo=uo, ao=ao, ito=ito,
                ro=ro, st=st, sto=sto, tab=tab, stats=stats)


def _check_solutions(R, tol):
    assert R["st"] == 0 and R["sto"] == 0
    # u[0] is preserved exactly (opt.cpp:296-303,107-109)
    assert np.array_equal(R["ug"][:, 0], R["u"][:, 0])
    rs = R["rg"] == R["ro"]
    print("same regularization level: %.4f; mean iterations mixed %.3f oracle %.3f"
          % (rs.mean(), R["itg"].mean(), R["ito"].mean()))
    assert rs.mean() >= 0.97
    un = np.linalg.norm(R["ug"][rs] - R["uo"][rs], axis=1) / np.linalg.norm(R["uo"][rs], axis=1)
    print("same-r cells: max ||u_mixed - u_oracle|| / ||u|| = %.3e, median %.3e (tol %.0e)"
          % (un.max(), np.median(un), tol))
    assert un.max() <= 4 * tol
    # the reconstruction is the moment vector of the returned multipliers
    tab = R["tab"]
    pm = tab.p[:, :tab.nm]
    recon = (np.exp(R["ag"] @ pm.T) * tab.w[None, :]) @ pm
    rel = np.abs(recon - R["ug"]).max(axis=1) / np.abs(R["ug"]).max(axis=1)
    print("reconstruction max rel %.3e" % rel.max())
    assert rel.max() <= 1e-10
    # unregularized cells: the FP64 gradient met the stop test of the scaled
    # problem at exit (opt.cpp:286-293)
    r0 = R["rg"] == 0.0
    unorm = np.linalg.norm(R["u"] / R["u"][:, :1], axis=1)
    assert np.all(R["ngg"][r0] <= tol * unorm[r0] * (1 + 1e-12))


@pytest.mark.gpu
@pytest.mark.parametrize("order,Q,gaunt", CASES)
def test_mixed_batch_matches_oracle(order, Q, gaunt):
    tol = 1e-6 if gaunt else 1e-8
    R = _run(order, Q, gaunt, n=3000, tol=tol, seed=4321 + order)
    _check_solutions(R, tol)
    # the approximate Hessian costs at most a few extra iterations on average
    assert R["itg"].mean() <= 1.5 * R["ito"].mean() + 0.5


@pytest.mark.gpu
@pytest.mark.parametrize("order,Q,gaunt", [(2, 12, False), (4, 16, False), (4, 16, True)])
def test_mixed_two_part_split_matches_oracle(order, Q, gaunt):
    """The two-part bf16 split (the default is three parts) with the cond bound
    at 10^5: the same solutions as the oracle, in about as many iterations."""
    tol = 1e-6 if gaunt else 1e-8
    R = _run(order, Q, gaunt, n=3000, tol=tol, seed=991 + order, tune={"mix_split": 2, "mix_cond_log10": 5})
    _check_solutions(R, tol)
    assert R["itg"].mean() <= 1.5 * R["ito"].mean() + 0.5


@pytest.mark.gpu
def test_mixed_all_deferred_equals_fp64_path():
    """cond bound 10^0: every mixed step is deferred, so each cell is redone from
    its first iterate by the FP64 tail and follows the oracle's path."""
    R = _run(4, 16, False, n=2000, tol=1e-8, seed=77, tune={"mix_cond_log10": 0})
    _check_solutions(R, 1e-8)
    same = (R["itg"] == R["ito"]) & (R["rg"] == R["ro"])
    print("identical (iter, r) with every step deferred: %.4f" % same.mean())
    assert same.mean() >= 0.95


@pytest.mark.gpu
@pytest.mark.parametrize("rounds", [0, 1, 12])
def test_mixed_round_counts(rounds):
    """The split between mixed lockstep rounds and the FP64 tail does not change
    what is solved."""
    R = _run(4, 16, False, n=1500, tol=1e-8, seed=5 + rounds, tune={"rounds": rounds})
    _check_solutions(R, 1e-8)


@pytest.mark.gpu
def test_mixed_lattice_floor_cells_robust():
    """The lattice floor cells (tests/golden/lattice_state) in mixed precision
    with the robust line search: all solved at r = 0, like the oracle."""
    import os
    from helpers import GOLDEN
    mnc = load_pkg("mnc")
    o = oracle()
    z = np.load(os.path.join(GOLDEN, "lattice_state", "floor_cells.npz"))
    u, a = z["u"], z["alpha"]
    ctx = mnc.Context(4, 16, use_gaunt=False)
    tab = o.Tables(4, 16, False)
    opt = mnc.options(max_iter=100, tol=1e-8, cond_h_max=1e10, line_search=mnc.LS_ROBUST, precision=mnc.PREC_MIXED)
    st, ug, ag, itg, rg, _ = ctx.opt_batch_host(u, a, opt)
    sto, uo, ao, ito, ro, _ = tab.opt_batch(u, a, tol=1e-8, nthreads=8, line_search=1)
    ctx.close()
    assert st == 0 and sto == 0
    assert np.all(rg == 0.0) and np.all(ro == 0.0)
    un = np.linalg.norm(ug - uo, axis=1) / np.linalg.norm(uo, axis=1)
    print("floor cells mixed vs oracle: max %.3e; iterations mixed %.3f oracle %.3f"
          % (un.max(), itg.mean(), ito.mean()))
    assert un.max() <= 4e-8


@pytest.mark.gpu
def test_mixed_lattice_steps_match_fp64():
    """Five Heun steps of an M4 lattice (80 x 80 on [-3.5, 3.5]^2, floor 1e-10,
    TOL 1e-8, robust line search): the mixed-precision run stays within
    tolerance level of the FP64 run and of the oracle."""
    mnc, gi = load_pkg("mnc"), load_pkg("grid_init")
    from helpers import density, gpu_grid_from_deck, rel_l2
    d = dict(num_cells_x=80, num_cells_y=80, a_x=-3.5, b_x=3.5, a_y=-3.5, b_y=3.5, gaussian_sigma=0.0,
             floor_value=1e-10, sigma=1.0, init_cond=1, moment_order=4, quad_order=16, cfl_factor=0.9,
             tol=1e-8, cond_h_max=1e10, max_iter=100, use_gaunt=0, theta=2.0, line_search=1)
    runs = {}
    for prec in (mnc.PREC_FP64, mnc.PREC_MIXED):
        ctx, g = gpu_grid_from_deck(mnc, gi, d)
        opt = mnc.options(max_iter=100, tol=1e-8, cond_h_max=1e10, line_search=mnc.LS_ROBUST, precision=prec)
        for _ in range(5):
            ctx.update(g.max_dt, opt)
        runs[prec] = ctx.grid_get()[0][2:-2, 2:-2]
        ctx.close()
    s = oracle().Solver(d, nthreads=8)
    for _ in range(5):
        s.update(s.max_dt)
    mo = s.state()[0][2:-2, 2:-2]
    e_mf = rel_l2(density(runs[mnc.PREC_MIXED]), density(runs[mnc.PREC_FP64]))
    e_mo = rel_l2(density(runs[mnc.PREC_MIXED]), density(mo))
    e_fo = rel_l2(density(runs[mnc.PREC_FP64]), density(mo))
    print("5 lattice steps, density rel-L2: mixed-fp64 %.3e mixed-oracle %.3e fp64-oracle %.3e" % (e_mf, e_mo, e_fo))
    assert e_fo <= 1e-12
    # each of the ten closure solves returns ||g|| <= tol ||u||: the FP64 path
    # usually ends far below tol (quadratic convergence), the mixed one just
    # below it (linear convergence with the approximate Hessian), so the runs
    # differ by a small multiple of tol -- the same 4 tol as the per-cell checks
    assert e_mf <= 4e-8 and e_mo <= 4e-8

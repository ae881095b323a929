"""The learned warm start on the GPU: the network's forward pass and gradient
against the numpy restatement (tests/mlp_ref.py, itself pinned by finite
differences), training, and Newton solves started from its guesses."""
import numpy as np
import pytest

import mlp_ref
from helpers import config2_batch, isotropic_start, load_pkg, oracle


def _ctx():
    mnc = load_pkg("mnc")
    return mnc, mnc.Context(4, 16, use_gaunt=False)


def _params_with_stats(ctx, rng):
    ctx.mlp_init(seed=11)
    p = ctx.mlp_params().astype(np.float64)
    # non-trivial standardisation
    p[mlp_ref.NPARAM:mlp_ref.NPARAM + 15] = rng.normal(0, 0.1, 15)
    p[mlp_ref.NPARAM + 15:mlp_ref.NPARAM + 30] = rng.uniform(0.2, 2.0, 15)
    p[mlp_ref.NPARAM + 30:mlp_ref.NPARAM + 45] = rng.normal(0, 1.0, 15)
    p[mlp_ref.NPARAM + 45:mlp_ref.NPARAM + 60] = rng.uniform(0.5, 3.0, 15)
    # non-zero biases
    layers_off = [15 * 128, 15 * 128 + 128 + 128 * 128, 15 * 128 + 2 * (128 + 128 * 128)]
    for o in layers_off:
        p[o:o + 128] = rng.normal(0, 0.1, 128)
    p[mlp_ref.NPARAM - 15:mlp_ref.NPARAM] = rng.normal(0, 0.1, 15)
    p = p.astype(np.float32)
    ctx.mlp_set_params(p)
    return p.astype(np.float64)


@pytest.mark.gpu
def test_mlp_forward_matches_reference():
    print("case mlp forward", flush=True)
    mnc, ctx = _ctx()
    rng = np.random.default_rng(1)
    p = _params_with_stats(ctx, rng)
    us = rng.normal(0, 0.3, (3000, 15))
    us[:, 0] = 1.0
    got = ctx.mlp_predict(us)
    x, _ = mlp_ref.standardise(p, us)
    _, st = mlp_ref.unpack(p)
    y = mlp_ref.forward(p, x, bf=True)[1][-1] * st[3] + st[2]
    y64 = mlp_ref.forward(p, x, bf=False)[1][-1] * st[3] + st[2]
    err = np.abs(got - y) / (np.abs(y).max() + 1e-30)
    err64 = np.abs(got - y64) / (np.abs(y64).max() + 1e-30)
    print("forward vs bf16 reference: max %.3e median %.3e; vs exact fp64: max %.3e" %
          (err.max(), np.median(err), err64.max()))
    assert np.median(err) <= 1e-5
    assert err.max() <= 1e-2
    assert err64.max() <= 5e-2
    ctx.close()


@pytest.mark.gpu
def test_mlp_gradient_matches_reference():
    print("case mlp gradient", flush=True)
    mnc, ctx = _ctx()
    rng = np.random.default_rng(2)
    p = _params_with_stats(ctx, rng)
    us = rng.normal(0, 0.3, (512, 15))
    us[:, 0] = 1.0
    al = rng.normal(0, 2.0, (512, 15))
    g, loss = ctx.mlp_gradient(us, al)
    x, t = mlp_ref.standardise(p, us, al)
    lref, gref = mlp_ref.loss_and_grad(p, x, t, bf=True)
    print("loss gpu %.8e ref %.8e" % (loss, lref))
    assert abs(loss - lref) <= 1e-4 * lref
    off = 0
    for i in range(4):
        a, b = mlp_ref.DIMS[i], mlp_ref.DIMS[i + 1]
        for name, n in (("W%d" % (i + 1), a * b), ("b%d" % (i + 1), b)):
            gg, gr = g[off:off + n], gref[off:off + n]
            rel = np.linalg.norm(gg - gr) / (np.linalg.norm(gr) + 1e-30)
            print("%s: rel err %.3e" % (name, rel))
            assert rel <= 1e-2, name
            off += n
    ctx.close()


def _smooth_pairs(mnc, tab, n, seed):
    rng = np.random.default_rng(seed)
    u, astar, lobe = config2_batch(tab.w, tab.p, tab.xi, tab.eta, tab.nm, n, rng, near_boundary_frac=0.0)
    return mnc.scaled_pairs(u, astar)


@pytest.mark.gpu
def test_mlp_training_reduces_loss():
    print("case mlp training", flush=True)
    mnc, ctx = _ctx()
    tab = oracle().Tables(4, 16, False)
    us, al = _smooth_pairs(mnc, tab, 100000, 3)
    ctx.mlp_init(seed=5)
    loss = ctx.mlp_train(us, al, steps=600, batch=4096, lr=1e-3, seed=9)
    first, last = loss[:10].mean(), loss[-20:].mean()
    print("training loss: first %.4e last %.4e" % (first, last))
    assert np.all(np.isfinite(loss))
    assert last <= 0.1 * first
    # held-out pairs: prediction error well below the spread of the targets
    ut, at = _smooth_pairs(mnc, tab, 5000, 4)
    pred = ctx.mlp_predict(ut)
    rel = np.linalg.norm(pred - at) / np.linalg.norm(at - at.mean(axis=0))
    print("held-out relative error %.3e" % rel)
    assert rel <= 0.3
    ctx.close()


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_learned_warm_start_solves_like_oracle(precision):
    """Train on exact smooth pairs, then solve fresh inputs (20% near-boundary
    lobes) from the learned start: every answer is a solution to the same
    tolerance as the oracle's from the isotropic start, and the learned start
    needs fewer Newton iterations than the isotropic one."""
    print("case learned warm start %s" % precision, flush=True)
    mnc, ctx = _ctx()
    o = oracle()
    tab = o.Tables(4, 16, False)
    us, al = _smooth_pairs(mnc, tab, 200000, 6)
    ctx.mlp_init(seed=7)
    ctx.mlp_train(us, al, steps=1500, batch=4096, lr=1e-3, seed=8)
    rng = np.random.default_rng(10)
    u, astar, lobe = config2_batch(tab.w, tab.p, tab.xi, tab.eta, tab.nm, 3000, rng, near_boundary_frac=0.2)
    a0 = isotropic_start(u)
    prec = mnc.PREC_MIXED if precision == "mixed" else mnc.PREC_FP64
    tol = 1e-8
    res = {}
    for warm in (mnc.WARM_ISOTROPIC, mnc.WARM_LEARNED):
        ctx.reset_stats()
        opt = mnc.options(max_iter=100, tol=tol, cond_h_max=1e10, precision=prec, warm_start=warm)
        res[warm] = ctx.opt_batch_host(u, a0, opt) + (ctx.stats(),)
    sto, uo, ao, ito, ro, _ = tab.opt_batch(u, a0, max_iter=100, tol=tol, cond_h_max=1e10, nthreads=8)
    ctx.close()
    st, ug, ag, itg, rg, ngg, stats = res[mnc.WARM_LEARNED]
    iso_iters = res[mnc.WARM_ISOTROPIC][3]
    picks = list(stats.warm_pick)
    print("iterations: isotropic %.3f learned %.3f (oracle %.3f); picks mlp/iso/prev %s"
          % (iso_iters.mean(), itg.mean(), ito.mean(), picks))
    assert st == 0 and sto == 0
    assert sum(picks) == len(u)
    assert picks[0] >= 0.5 * len(u)
    assert itg.mean() < iso_iters.mean()
    assert np.array_equal(ug[:, 0], u[:, 0])
    rs = rg == ro
    assert rs.mean() >= 0.97
    un = np.linalg.norm(ug[rs] - uo[rs], axis=1) / np.linalg.norm(uo[rs], axis=1)
    print("same-r cells: max ||u - u_oracle|| / ||u|| = %.3e" % un.max())
    assert un.max() <= 4 * tol


@pytest.mark.gpu
def test_learned_start_gate_keeps_a_good_previous_alpha():
    """When the alpha passed in already nearly solves the problem (the time-
    stepping case), the prologue keeps it and the network is not consulted:
    every cell picks 'previous' and converges in one or two iterations."""
    print("case learned gate", flush=True)
    mnc, ctx = _ctx()
    o = oracle()
    tab = o.Tables(4, 16, False)
    us, al = _smooth_pairs(mnc, tab, 50000, 16)
    ctx.mlp_init(seed=3)
    ctx.mlp_train(us, al, steps=200, batch=4096, lr=1e-3, seed=4)
    rng = np.random.default_rng(17)
    u, astar, lobe = config2_batch(tab.w, tab.p, tab.xi, tab.eta, tab.nm, 2000, rng, near_boundary_frac=0.0)
    a0 = isotropic_start(u)
    opt = mnc.options(max_iter=100, tol=1e-8, cond_h_max=1e10, warm_start=mnc.WARM_LEARNED)
    st, u1, a1, it1, r1, _ = ctx.opt_batch_host(u, a0, opt)
    assert st == 0
    ctx.reset_stats()
    st, u2, a2, it2, r2, _ = ctx.opt_batch_host(u1, a1, opt)       # already solved: the start is kept
    s = ctx.stats()
    ctx.close()
    print("second solve: picks mlp/iso/prev %s, max iterations %d" % (list(s.warm_pick), it2.max()))
    assert st == 0
    ok = r1 == 0
    assert s.warm_pick[2] >= ok.sum()
    assert it2[ok].max() <= 2


@pytest.mark.gpu
def test_learned_start_with_network_off_is_the_previous_start():
    """Gate value 5 switches the network off: without any network parameters
    the learned form then solves exactly as the previous-alpha start does (the
    same kernels, bit for bit), counts every finite start as 'previous', and a
    cell whose alpha is not finite starts from isotropicAlpha."""
    print("case learned gate off", flush=True)
    mnc = load_pkg("mnc")
    ctx = mnc.Context(4, 16, use_gaunt=False)
    tab = oracle().Tables(4, 16, False)
    rng = np.random.default_rng(23)
    u, astar, lobe = config2_batch(tab.w, tab.p, tab.xi, tab.eta, tab.nm, 2000, rng, near_boundary_frac=0.2)
    a0 = isotropic_start(u) + rng.normal(0, 0.05, u.shape)
    ctx.tune("learned_gate_log10", 5)
    assert ctx.get_tune("learned_gate_log10") == 5
    out = {}
    for prec in (mnc.PREC_FP64, mnc.PREC_MIXED):
        for warm in (mnc.WARM_PREVIOUS, mnc.WARM_LEARNED):
            ctx.reset_stats()
            opt = mnc.options(max_iter=100, tol=1e-8, cond_h_max=1e10, precision=prec, warm_start=warm)
            out[prec, warm] = ctx.opt_batch_host(u, a0, opt) + (list(ctx.stats().warm_pick),)
    for prec in (mnc.PREC_FP64, mnc.PREC_MIXED):
        p, l = out[prec, mnc.WARM_PREVIOUS], out[prec, mnc.WARM_LEARNED]
        assert p[0] == 0 and l[0] == 0
        for k in (1, 2, 3, 4):
            assert np.array_equal(p[k], l[k]), (prec, k)
        assert l[6] == [0, 0, len(u)]
    bad = a0.copy()
    bad[:5, 3] = np.nan
    ctx.reset_stats()
    opt = mnc.options(max_iter=100, tol=1e-8, cond_h_max=1e10, warm_start=mnc.WARM_LEARNED)
    st, ug, ag, itg, rg, _ = ctx.opt_batch_host(u, bad, opt)
    picks = list(ctx.stats().warm_pick)
    ctx.close()
    print("non-finite starts: picks %s" % picks)
    assert st == 0 and picks == [0, 5, len(u) - 5]
    assert np.all(np.isfinite(ag))


@pytest.mark.gpu
def test_fused_network_judge_matches_the_gemm_path():
    """The learned start's network and judge in one kernel (learned_fused 1,
    the default: weights and quadrature in LDS, activations on chip, the
    judge's gradient on the matrix cores in FP32) against the separate GEMM
    kernels and the FP64 comparison kernel (learned_fused 0): the same bf16
    operands and FP32 sums in the same order, so the network's guesses are
    the same; the two judges' gradient norms differ at rounding level (FP32
    against FP64), which can tip a near-tie between the guess and the parked
    start.  Solved from those starts the batches agree cell for cell."""
    print("case fused network judge", flush=True)
    mnc, ctx = _ctx()
    tab = oracle().Tables(4, 16, False)
    us, al = _smooth_pairs(mnc, tab, 100000, 31)
    ctx.mlp_init(seed=32)
    ctx.mlp_train(us, al, steps=600, batch=4096, lr=1e-3, seed=33)
    rng = np.random.default_rng(34)
    u, astar, lobe = config2_batch(tab.w, tab.p, tab.xi, tab.eta, tab.nm, 20000, rng, near_boundary_frac=0.2)
    a0 = isotropic_start(u)
    opt = mnc.options(max_iter=100, tol=1e-8, cond_h_max=1e10, precision=mnc.PREC_MIXED, warm_start=mnc.WARM_LEARNED)
    out = {}
    for fused in (1, 0):
        ctx.tune("learned_fused", fused)
        ctx.reset_stats()
        out[fused] = ctx.opt_batch_host(u, a0, opt) + (list(ctx.stats().warm_pick),)
    ctx.close()
    f, g = out[1], out[0]
    print("picks mlp/iso/prev: fused %s, gemm path %s" % (f[6], g[6]))
    assert f[0] == 0 and g[0] == 0
    assert sum(f[6]) == len(u) and f[6][0] >= 0.5 * len(u)
    assert abs(f[6][0] - g[6][0]) <= 0.001 * len(u)
    same = np.all(f[2] == g[2], axis=1) & np.all(f[1] == g[1], axis=1)
    print("bitwise identical cells: %.4f" % same.mean())
    assert same.mean() >= 0.99
    rs = f[4] == g[4]
    assert rs.mean() >= 0.99
    un = np.linalg.norm(f[1][rs] - g[1][rs], axis=1) / np.linalg.norm(g[1][rs], axis=1)
    assert un.max() <= 4e-8


@pytest.mark.gpu
def test_learned_lattice_steps_match_oracle():
    """The learned+mixed form time-stepping an M4 lattice (80 x 80 on
    [-3.5, 3.5]^2, floor 1e-10, TOL 1e-8, robust line search) against the
    oracle: a network trained on (u, alpha) pairs harvested from the plain
    form's first steps, consulted by every cell whose previous alpha leaves
    ||g|| > 1e-6 ||u|| (gate -6), keeps the state within tolerance level of
    the oracle's after 8 steps, as the mixed form does."""
    mnc, gi = load_pkg("mnc"), load_pkg("grid_init")
    from helpers import density, gpu_grid_from_deck, rel_l2
    print("case learned lattice", flush=True)
    d = dict(num_cells_x=80, num_cells_y=80, a_x=-3.5, b_x=3.5, a_y=-3.5, b_y=3.5, gaussian_sigma=0.0,
             floor_value=1e-10, sigma=1.0, init_cond=1, moment_order=4, quad_order=16, cfl_factor=0.9,
             tol=1e-8, cond_h_max=1e10, max_iter=100, use_gaunt=0, theta=2.0, line_search=1)
    steps = 8
    ctx, g = gpu_grid_from_deck(mnc, gi, d)
    plain = mnc.options(max_iter=100, tol=1e-8, cond_h_max=1e10, line_search=mnc.LS_ROBUST)
    us, al = [], []
    for _ in range(steps):
        ctx.update(g.max_dt, plain)
        m, a = ctx.grid_get()
        u1, a1 = mnc.scaled_pairs(m.reshape(-1, 15), a.reshape(-1, 15))
        us.append(u1)
        al.append(a1)
    us, al = np.concatenate(us), np.concatenate(al)
    keep = np.abs(al[:, 1:]).max(axis=1) <= 50.0
    ctx.mlp_init(seed=41)
    ctx.mlp_train(us[keep], al[keep], steps=600, batch=4096, lr=1e-3, seed=42)
    ctx.grid_set(g.moments, g.alpha)
    ctx.tune("learned_gate_log10", -6)
    ctx.reset_stats()
    learned = mnc.options(max_iter=100, tol=1e-8, cond_h_max=1e10, line_search=mnc.LS_ROBUST,
                          precision=mnc.PREC_MIXED, warm_start=mnc.WARM_LEARNED)
    for _ in range(steps):
        ctx.update(g.max_dt, learned)
    picks = list(ctx.stats().warm_pick)
    ml = ctx.grid_get()[0][2:-2, 2:-2]
    ctx.close()
    s = oracle().Solver(d, nthreads=8)
    for _ in range(steps):
        s.update(s.max_dt)
    mo = s.state()[0][2:-2, 2:-2]
    e = rel_l2(density(ml), density(mo))
    print("learned+mixed, %d lattice steps: picks mlp/iso/prev %s, density rel-L2 vs oracle %.3e" % (steps, picks, e))
    assert picks[0] > 0
    # the same bound as the mixed form's: every solve returns ||g|| <= tol ||u||
    assert e <= 4e-8


@pytest.mark.gpu
def test_learned_fallback_restores_the_isotropic_answer():
    """With the reference line search at TOL 1e-8 a start close to the solution
    can stall just above the tolerance (the Armijo test cannot resolve the
    decrease of f, t = 0) and the reference algorithm then regularizes: about
    1% of smooth cells do so from the network's guess though they converge
    unregularized from isotropicAlpha.  The fallback (learned_fallback_iters,
    default 10) restarts such a cell once from isotropicAlpha at level 0:
    the learned form then regularizes no more cells than the isotropic start
    does, and with the fallback off (0) it regularizes more."""
    print("case learned fallback", flush=True)
    mnc, ctx = _ctx()
    tab = oracle().Tables(4, 16, False)
    us, al = _smooth_pairs(mnc, tab, 200000, 51)
    ctx.mlp_init(seed=52)
    ctx.mlp_train(us, al, steps=1500, batch=4096, lr=1e-3, seed=53)
    rng = np.random.default_rng(54)
    u, astar, lobe = config2_batch(tab.w, tab.p, tab.xi, tab.eta, tab.nm, 100000, rng, near_boundary_frac=0.0)
    a0 = isotropic_start(u)
    res = {}
    for name, warm, fb in (("iso", mnc.WARM_ISOTROPIC, 10), ("learned", mnc.WARM_LEARNED, 10),
                           ("no-fallback", mnc.WARM_LEARNED, 0)):
        ctx.tune("learned_fallback_iters", fb)
        ctx.reset_stats()
        opt = mnc.options(max_iter=100, tol=1e-8, cond_h_max=1e10, precision=mnc.PREC_MIXED, warm_start=warm)
        st, uo, ao, it, r, ng = ctx.opt_batch_host(u, a0, opt)
        assert st == 0
        res[name] = (r, ctx.stats().warm_fallback)
    ctx.close()
    reg = {k: int((v[0] > 0).sum()) for k, v in res.items()}
    print("regularized cells: %s; fallbacks with the default %d" % (reg, res["learned"][1]))
    assert res["learned"][1] > 0 and res["no-fallback"][1] == 0 and res["iso"][1] == 0
    assert reg["learned"] <= reg["iso"]
    assert reg["no-fallback"] > reg["learned"]

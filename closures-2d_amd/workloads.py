"""Synthetic closure workloads (numpy, host side): the isolated batch of M4
moment vectors used to measure the closure on its own.

u = sum_q w_q m_q exp(alpha* . m_q) with alpha*_0 ~ U[-2, 2] and
alpha*_k ~ N(0, (3/(1+k))^2) (k = flat moment index), and a fraction of
near-boundary vectors u = sum_q w_q m_q f_q built from 1-3 narrow von
Mises-Fisher lobes f = sum_j c_j exp(kappa (Omega . Omega_j - 1)),
kappa = 400 (angular width 0.05), c_j ~ U[0.1, 1], Omega_j uniform on the
quadrature's hemisphere."""
import numpy as np


def isolated_batch(w, p, xi, eta, nm, n, seed, near_boundary_frac=0.2, kappa=400.0, chunk=65536):
    """Returns (u[n][nm], alpha_star[n][nm] with NaN rows for lobes, is_lobe[n])."""
    rng = np.random.default_rng(seed)
    pm = np.ascontiguousarray(p[:, :nm])
    wm = w[:, None] * pm
    alpha = np.empty((n, nm))
    alpha[:, 0] = rng.uniform(-2.0, 2.0, n)
    for k in range(1, nm):
        alpha[:, k] = rng.normal(0.0, 3.0 / (1.0 + k), n)
    u = np.empty((n, nm))
    for i in range(0, n, chunk):
        u[i:i + chunk] = np.exp(alpha[i:i + chunk] @ pm.T) @ wm
    nb = int(round(near_boundary_frac * n))
    lobe = np.zeros(n, dtype=bool)
    if nb:
        mu = -np.sqrt(np.maximum(0.0, 1.0 - xi * xi - eta * eta))
        om = np.stack([xi, eta, mu], axis=1)                    # [nq][3]
        idx = np.sort(rng.choice(n, nb, replace=False))
        J = rng.integers(1, 4, nb)
        c = rng.uniform(0.1, 1.0, (nb, 3)) * (np.arange(3)[None, :] < J[:, None])
        z = -rng.uniform(0.0, 1.0, (nb, 3))
        ph = rng.uniform(0.0, 2.0 * np.pi, (nb, 3))
        r = np.sqrt(1.0 - z * z)
        oj = np.stack([r * np.cos(ph), r * np.sin(ph), z], axis=2)   # [nb][3][3]
        for i in range(0, nb, chunk):
            dots = np.einsum("ljd,qd->ljq", oj[i:i + chunk], om)     # [l][3][nq]
            f = (c[i:i + chunk, :, None] * np.exp(kappa * (dots - 1.0))).sum(axis=1)
            u[idx[i:i + chunk]] = f @ wm
        alpha[idx] = np.nan
        lobe[idx] = True
    return u, alpha, lobe


def isotropic_start(u):
    """alpha whose scaled value is isotropicAlpha (opt.cpp:27-35) after opt()'s
    shift (opt.cpp:91-92)."""
    a = np.zeros_like(u)
    a[:, 0] = 2.0 * np.sqrt(np.pi) * (np.log(0.5 / np.sqrt(np.pi)) + np.log(u[:, 0]))
    return a


def sine_state(g, nm):
    """The smooth periodic M8 workload's initial state on a grid_init GridState
    (moments/alpha replaced in place): isotropic density 1 + 0.5 sin(pi x)
    sin(pi y) at the cell centres, ghosts included."""
    NX, NY = g.nx + 4, g.ny + 4
    i = np.arange(NX)[:, None] * np.ones((1, NY))
    j = np.ones((NX, 1)) * np.arange(NY)[None, :]
    x = g.init_x + (i - 2) * g.dx + g.dx / 2.0
    y = g.init_y + (j - 2) * g.dy + g.dy / 2.0
    dens = 1.0 + 0.5 * np.sin(np.pi * x) * np.sin(np.pi * y)
    g.moments = np.zeros((NX, NY, nm))
    g.alpha = np.zeros((NX, NY, nm))
    g.moments[:, :, 0] = dens * 2.0 * np.sqrt(np.pi)
    g.alpha[:, :, 0] = 2.0 * np.sqrt(np.pi) * np.log(dens)
    return g

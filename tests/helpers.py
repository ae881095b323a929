"""Shared helpers for the test suite: module loading, .pn files, decks."""
import importlib.util
import math
import os
import struct
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "closures-2d_amd")
GOLDEN = os.path.join(ROOT, "tests", "golden")


def load_pkg(name):
    """Import closures-2d_amd/<name>.py (the directory name is not a Python identifier)."""
    modname = "closures2d_amd_" + name
    if modname in sys.modules:
        return sys.modules[modname]
    spec = importlib.util.spec_from_file_location(modname, os.path.join(PKG, name + ".py"))
    mod = importlib.util.module_from_spec(spec)
    sys.modules[modname] = mod
    spec.loader.exec_module(mod)
    return mod


def oracle():
    if ROOT not in sys.path:
        sys.path.insert(0, ROOT)
    import oracle as o
    return o


PN_HEADER = struct.Struct("=2q4dq")   # util/formats.py:169


def read_pn(path):
    """(header, moments[sizeX][sizeY][n]) of a .pn file (util/formats.py:168-189)."""
    with open(path, "rb") as f:
        data = f.read()
    h = PN_HEADER.unpack(data[:PN_HEADER.size])
    m = np.frombuffer(data[PN_HEADER.size:], dtype=np.float64).reshape(h[0], h[1], h[6])
    return h, m


def density(m):
    """Density grid u0/(2 sqrt(pi)) (util/formats.py:184-187)."""
    return m[:, :, 0] / (2.0 * math.sqrt(math.pi))


def rel_l2(a, b):
    """abs(a - b) / abs(b) on density grids (site_scons/testing.py:127-132, formats.py:36-41)."""
    return float(np.sqrt(((a - b) ** 2).sum()) / np.sqrt((b ** 2).sum()))


def mass(dens, dx, dy):
    return float(dens.sum() * dx * dy)    # util/formats.py:88-93


# The supplementary stop test's bound the fixtures were written with
# (momopt_update.cpp:226, opt.cpp:288-293).
TOL_GAMMA = 1.1
# The regression decks of site_scons/testing.py:120-146 with util/decks.py defaults.
FIXTURE_BASE = dict(num_cells_x=36, num_cells_y=74, a_x=-1.55, b_x=1.45, a_y=-1.62, b_y=1.38,
                    floor_value=1e-4, sigma=1.111, moment_order=3, quad_order=30, cfl_factor=0.9,
                    tol=1e-4, cond_h_max=1e10, max_iter=100, use_gaunt=1, theta=2.0, tol_gamma=TOL_GAMMA)
FIXTURE_IC = {"delta": (0, 0.0), "gaussian": (0, 1.7), "lattice": (1, 1.7), "smooth": (2, 1.7)}
FIXTURE_T_FINAL = 1.03


def fixture_deck(ic, **over):
    d = dict(FIXTURE_BASE)
    d["init_cond"], d["gaussian_sigma"] = FIXTURE_IC[ic]
    d.update(over)
    return d


def fixture_path(ic):
    return os.path.join(GOLDEN, "momopt", ic, "mn.pn")


def deck_text(d, t_final, extra=None, out_dt=1000.1, opt_type=0):
    """An input.deck for the drivers (util/decks.py:26-104 key set, %.17f values)."""
    lines = ["SOLVER momopt",
             "NUM_CELLS_X %d" % d["num_cells_x"], "NUM_CELLS_Y %d" % d["num_cells_y"],
             "NUM_MPI_PARTITIONS_X 1", "NUM_MPI_PARTITIONS_Y 1",
             "A_X %.17f" % d["a_x"], "B_X %.17f" % d["b_x"], "A_Y %.17f" % d["a_y"], "B_Y %.17f" % d["b_y"],
             "T_FINAL %.17f" % t_final, "OUT_DELTA_T %.17f" % out_dt,
             "GAUSSIAN_SIGMA %.17f" % d["gaussian_sigma"], "FLOOR %.17f" % d["floor_value"],
             "INIT_COND %d" % d["init_cond"], "SIGMA %.17f" % d["sigma"],
             "MOMOPT_MOMENT_ORDER %d" % d["moment_order"], "MOMOPT_QUAD_ORDER %d" % d["quad_order"],
             "MOMOPT_CFL_FACTOR %.17f" % d["cfl_factor"], "MOMOPT_TOL %.17f" % d["tol"],
             "MOMOPT_COND_H_MAX %.17f" % d["cond_h_max"], "MOMOPT_COND_H_MAX_BFGS 20",
             "MOMOPT_MAX_ITER %d" % d["max_iter"], "MOMOPT_MAX_BFGS_ITER 5",
             "MOMOPT_USE_GAUNT %d" % d["use_gaunt"], "MOMOPT_THETA %.17f" % d["theta"],
             "MOMOPT_DELTA_PPN %.17f" % d.get("delta_ppn", 1e-10), "MOMOPT_MOMENT_TYPE %d" % d.get("moment_type", 0),
             "MOMOPT_OPTIMIZATION_TYPE %d" % opt_type]
    if d.get("tol_gamma", 0.0) > 0.0 and opt_type == 0:
        lines.append("MOMOPT_TOL_GAMMA %.17f" % d["tol_gamma"])
    for k, v in (extra or {}).items():
        lines.append("%s %s" % (k, v))
    return "\n".join(lines) + "\n"


def gpu_grid_from_deck(mnc, gi, d, device=0):
    """A mnc.Context holding the grid of deck d, initialised by grid_init."""
    mt = d.get("moment_type", 0)
    ctx = mnc.Context(d["moment_order"], d["quad_order"], use_gaunt=d["use_gaunt"], device=device, moment_type=mt)
    g = gi.initial_state(d["num_cells_x"], d["num_cells_y"], d["a_x"], d["b_x"], d["a_y"], d["b_y"],
                         d["init_cond"], d["gaussian_sigma"], d["floor_value"], d["sigma"], ctx.nm,
                         d["cfl_factor"], d["theta"], moment_type=mt, delta_ppn=d.get("delta_ppn", 1e-10))
    ctx.grid_create(g.nx, g.ny, g.dx, g.dy, g.init_x, g.init_y, d["init_cond"], d["theta"],
                    g.moments, g.alpha, g.sigma_t, g.sigma_s, delta_ppn=d.get("delta_ppn", 0.0))
    return ctx, g


def mn_options(mnc, d, **over):
    kw = dict(max_iter=d["max_iter"], tol=d["tol"], cond_h_max=d["cond_h_max"],
              line_search=d.get("line_search", 0), delta_ppn=d.get("delta_ppn", 0.0),
              tol_gamma=d.get("tol_gamma", 0.0))
    kw.update(over)
    return mnc.options(**kw)


def config2_batch(w, p, xi, eta, nm, n, rng, near_boundary_frac=0.2, kappa=400.0, scale_k=3.0):
    """Isolated-closure inputs in the style of the isolated M4 benchmark case:
    u = sum_q w_q m_q exp(alpha*.m_q) with alpha*_0 ~ U[-2,2] and
    alpha*_k ~ N(0, (scale_k/(1+k))^2) (k = flat moment index), and a fraction
    of near-boundary vectors u = sum_q w_q m_q f_q from 1-3 narrow von
    Mises-Fisher lobes f = sum_j c_j exp(kappa (Omega.Omega_j - 1)),
    c_j ~ U[0.1, 1], Omega_j uniform on the quadrature's hemisphere.
    Returns (u[n][nm], alpha_star[n][nm] (NaN rows for lobes), is_lobe[n])."""
    pm = p[:, :nm]
    alpha = np.zeros((n, nm))
    alpha[:, 0] = rng.uniform(-2.0, 2.0, n)
    for k in range(1, nm):
        alpha[:, k] = rng.normal(0.0, scale_k / (1.0 + k), n)
    u = (np.exp(alpha @ pm.T) * w[None, :]) @ pm
    nb = int(round(near_boundary_frac * n))
    lobe = np.zeros(n, dtype=bool)
    if nb:
        mu = -np.sqrt(np.maximum(0.0, 1.0 - xi * xi - eta * eta))
        om = np.stack([xi, eta, mu], axis=1)
        idx = rng.choice(n, nb, replace=False)
        for c in idx:
            J = rng.integers(1, 4)
            f = np.zeros(len(w))
            for _ in range(J):
                z = -rng.uniform(0.0, 1.0)
                ph = rng.uniform(0.0, 2.0 * np.pi)
                r = np.sqrt(1.0 - z * z)
                oj = np.array([r * np.cos(ph), r * np.sin(ph), z])
                f += rng.uniform(0.1, 1.0) * np.exp(kappa * (om @ oj - 1.0))
            u[c] = (w * f) @ pm
            alpha[c] = np.nan
        lobe[idx] = True
    return u, alpha, lobe


def isotropic_start(u):
    """alpha whose scaled value is isotropicAlpha (opt.cpp:27-35) after opt()'s shift (opt.cpp:91-92)."""
    a = np.zeros_like(u)
    a[:, 0] = 2.0 * math.sqrt(math.pi) * math.log(0.5 / math.sqrt(math.pi)) + \
        np.log(u[:, 0]) * 2.0 * math.sqrt(math.pi)
    return a

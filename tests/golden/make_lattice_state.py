"""Writes tests/golden/lattice_state/floor_cells.npz: (u, alpha) of 2 x 24 x 24
cells of the M4 lattice (280 x 280 on [-3.5, 3.5]^2, floor 1e-10, Q = 16,
TOL 1e-8) at the start of the first step in which the reference algorithm
(the CPU oracle, reference line search) aborts with "Max regularization",
centred on a cell whose solve climbs the regularization ladder to r >= 1e-2 and
on another regularized cell.

Run from the repository root: python tests/golden/make_lattice_state.py"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))
import oracle  # noqa: E402
from helpers import load_pkg  # noqa: E402

D = dict(num_cells_x=280, num_cells_y=280, a_x=-3.5, b_x=3.5, a_y=-3.5, b_y=3.5, gaussian_sigma=0.0,
         floor_value=1e-10, sigma=1.0, init_cond=1, moment_order=4, quad_order=16, cfl_factor=0.9, tol=1e-8,
         cond_h_max=1e10, max_iter=100, use_gaunt=0, theta=2.0, line_search=0)


def main():
    gi = load_pkg("grid_init")
    s = oracle.Solver(D, nthreads=len(os.sched_getaffinity(0)))
    dts = gi.time_steps(s.max_dt, 3.2)
    for k, dt in enumerate(dts):
        m, a = s.state()
        try:
            s.update(dt)
        except RuntimeError:
            break
    else:
        raise SystemExit("no step aborted")
    print("first aborting step:", k)
    nm = m.shape[-1]
    tab = oracle.Tables(4, 16, False)
    st, uo, ao, it, r, _ = tab.opt_batch(m.reshape(-1, nm), a.reshape(-1, nm), max_iter=100, tol=1e-8,
                                         cond_h_max=1e10, nthreads=len(os.sched_getaffinity(0)))
    r = r.reshape(m.shape[:2])
    bad = np.argwhere((r >= 1e-2) | (r < 0))
    reg = np.argwhere(r != 0.0)
    print("cells climbing to r >= 1e-2 or aborting:", len(bad), "; regularized:", len(reg))
    picks = [bad[0]]
    for b in reg:
        if np.abs(b - picks[0]).max() > 24:
            picks.append(b)
            break
    us, al = [], []
    for ci, cj in picks[:2]:
        i0 = int(np.clip(ci - 12, 0, m.shape[0] - 24))
        j0 = int(np.clip(cj - 12, 0, m.shape[1] - 24))
        us.append(m[i0:i0 + 24, j0:j0 + 24].reshape(-1, nm))
        al.append(a[i0:i0 + 24, j0:j0 + 24].reshape(-1, nm))
        print("patch around", (int(ci), int(cj)))
    out = os.path.join(ROOT, "tests", "golden", "lattice_state", "floor_cells.npz")
    np.savez_compressed(out, u=np.concatenate(us), alpha=np.concatenate(al))
    print("wrote", out)


if __name__ == "__main__":
    main()

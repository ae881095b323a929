"""Host-side grid set-up for the Mn solver: initial density, cross sections and
the initial (u, alpha) of every cell, ghost cells included.

Restates Solver::initializeGrid (src/main.cpp:257-372) and the initial-state
part of MomOptSolver::init (src/momopt/momopt_init.cpp:72,120-143) for one rank
(c_initX = A_X, c_initY = A_Y).  Arrays use the reference's I3D layout
[i][j][k] with 2 ghost cells per side.
"""
import math

import numpy as np

LATTICE_CENTRES = ((2.0, 2.0), (2.0, 0.0), (2.0, -2.0), (1.0, 1.0), (1.0, -1.0), (0.0, -2.0),
                   (-1.0, 1.0), (-1.0, -1.0), (-2.0, 2.0), (-2.0, 0.0), (-2.0, -2.0))


class GridState:
    pass


def initial_state(nx, ny, a_x, b_x, a_y, b_y, init_cond, gaussian_sigma, floor_value, sigma, nm,
                  cfl_factor=0.9, theta=2.0, moment_type=0, delta_ppn=1e-10):
    g = GridState()
    g.nx, g.ny = nx, ny
    g.dx = (b_x - a_x) / nx
    g.dy = (b_y - a_y) / ny
    g.init_x, g.init_y = a_x, a_y
    NX, NY = nx + 4, ny + 4
    i = np.arange(NX)[:, None] * np.ones((1, NY))
    j = np.ones((NX, 1)) * np.arange(NY)[None, :]
    x = g.init_x + (i - 2) * g.dx + g.dx / 2.0
    y = g.init_y + (j - 2) * g.dy + g.dy / 2.0
    ghost = (i < 2) | (i > nx + 1) | (j < 2) | (j > ny + 1)
    init = np.full((NX, NY), float(floor_value))
    st = np.full((NX, NY), float(sigma))
    ss = np.full((NX, NY), float(sigma))
    inner = ~ghost
    if init_cond == 0 and gaussian_sigma == 0.0:
        init[(nx + 2) // 2, (ny + 2) // 2] = 1.0 / (g.dx * g.dy)
    elif init_cond == 0:
        s2 = gaussian_sigma * gaussian_sigma
        gf = 1.0 / (2.0 * math.pi * s2) * np.exp(-(x * x + y * y) / (2.0 * s2))
        init[inner] = np.maximum(gf, floor_value)[inner]
    elif init_cond == 1:
        st[inner] = 1.0
        ss[inner] = 1.0
        blk = np.zeros((NX, NY), dtype=bool)
        for cx, cy in LATTICE_CENTRES:
            blk |= (np.abs(x - cx) < 0.5) & (np.abs(y - cy) < 0.5)
        blk &= inner
        st[blk] = 10.0
        ss[blk] = 0.0
    elif init_cond == 2:
        init[inner] = (1 + np.sin(2 * math.pi * x) * np.cos(2 * math.pi * y))[inner]
    g.init = init
    g.sigma_t = st
    g.sigma_s = ss
    moments = np.zeros((NX, NY, nm))
    alpha = np.zeros((NX, NY, nm))
    moments[:, :, 0] = init * 2.0 * math.sqrt(math.pi)
    if moment_type == 1:      # PPn (momopt_init.cpp:138-141)
        alpha[:, :, 0] = moments[:, :, 0] - 4.0 * delta_ppn * math.pi / moments[:, :, 0]
    else:
        alpha[:, :, 0] = 2.0 * math.sqrt(math.pi) * np.log(moments[:, :, 0] / (2.0 * math.sqrt(math.pi)))
    g.moments = moments
    g.alpha = alpha
    g.max_dt = cfl_factor * (2.0 / (2.0 + theta)) * (g.dx * g.dy) / (g.dx + g.dy)
    return g


def time_steps(max_dt, t_final, out_delta_t=1000.1):
    """The dt sequence of the reference time loop (src/main.cpp:204-230)."""
    t = 0.0
    steps = []
    t_out = min(t + out_delta_t, t_final)
    while t_out <= t_final:
        dt = max_dt
        while t < t_out:
            t += dt
            if t > t_out:
                dt = t_out - (t - dt)
                t = t_out
            steps.append(dt)
        if t_out > t_final - 1e-12 * t_final:
            break
        t_out = min(t_out + out_delta_t, t_final)
    return steps

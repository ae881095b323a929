"""The oracle's Sn step against the reference's own kinetic regression
fixtures (tests/regression/kinetic/<ic>.sn, copied as data under
tests/golden/kinetic/).  Procedure of site_scons/testing.py:109-180,225-240:
the regression deck (36 x 74 cells on [-1.55, 1.45] x [-1.62, 1.38],
sigma 1.111, t = 1.03) with util/decks.py:1-5's kinetic deck (Q = 12,
CFL 0.9); the initial state is every ordinate equal to the cell's initial
density (kinetic_init.cpp:76-86), dt = cfl (2 / 4) dx dy / (dx + dy)
(kinetic_init.cpp:49) and the main loop's last partial step
(main.cpp:204-230).  Density compared with the file's own weights
(util/formats.py:130-153) at the reference's 1e-10, plus every ordinate and
the weights stored in the header against the oracle's quadrature."""
import math
import os

import numpy as np
import pytest

from helpers import FIXTURE_BASE, FIXTURE_IC, FIXTURE_T_FINAL, GOLDEN, load_pkg, oracle, rel_l2
import struct

SN_HEADER = struct.Struct("=2q4dq")   # util/formats.py:131
SN_Q = 12
SN_CFL = 0.9


def read_sn(path):
    """(header, weights[nq], f[sizeX][sizeY][nq]) of a .sn file (util/formats.py:130-153)."""
    with open(path, "rb") as f:
        data = f.read()
    h = SN_HEADER.unpack(data[:SN_HEADER.size])
    nq = h[6]
    w = np.frombuffer(data[SN_HEADER.size:SN_HEADER.size + 8 * nq], dtype=np.float64)
    m = np.frombuffer(data[SN_HEADER.size + 8 * nq:], dtype=np.float64).reshape(h[0], h[1], nq)
    return h, w, m


def sn_density(f, w):
    return (f @ w) / (4.0 * math.pi)


def sn_fixture_case(ic):
    """Initial state, cross sections, quadrature and dt sequence of the
    kinetic regression deck for initial condition ic."""
    gi = load_pkg("grid_init")
    d = dict(FIXTURE_BASE)
    d["init_cond"], d["gaussian_sigma"] = FIXTURE_IC[ic]
    g = gi.initial_state(d["num_cells_x"], d["num_cells_y"], d["a_x"], d["b_x"], d["a_y"], d["b_y"],
                         d["init_cond"], d["gaussian_sigma"], d["floor_value"], d["sigma"], 1)
    t = oracle().Tables(1, SN_Q, False)
    f0 = np.repeat(g.init[:, :, None], len(t.w), axis=2)
    max_dt = SN_CFL * (2.0 / (2.0 + 2.0)) * (g.dx * g.dy) / (g.dx + g.dy)
    return g, t, f0, gi.time_steps(max_dt, FIXTURE_T_FINAL)


def sn_fixture_path(ic):
    return os.path.join(GOLDEN, "kinetic", "%s.sn" % ic)


@pytest.mark.parametrize("ic", ["gaussian", "smooth", "delta", "lattice"])
def test_sn_fixture(ic):
    g, t, f, steps = sn_fixture_case(ic)
    assert len(steps) == 84
    o = oracle()
    for dt in steps:
        f = o.sn_steps(f, t.w, t.xi, t.eta, g.sigma_t, g.sigma_s, FIXTURE_IC[ic][0], g.init_x, g.init_y,
                       g.dx, g.dy, dt)
    h, w, ref = read_sn(sn_fixture_path(ic))
    assert h[:2] == (36, 74) and h[6] == SN_Q * SN_Q
    assert np.abs(w - t.w).max() <= 1e-15 * np.abs(w).max() * 4
    fin = f[2:-2, 2:-2]
    err = rel_l2(sn_density(fin, w), sn_density(ref, w))
    per = np.abs(fin - ref).max() / np.abs(ref).max()
    print("Sn %s: density rel L2 %.3e, per-ordinate max rel %.3e" % (ic, err, per))
    assert err < 1e-10
    assert per < 1e-12
    if ic == "smooth":    # periodic, sigma_t = sigma_s: testing.py:174-178
        dm = (sn_density(fin, w).sum() - sn_density(g.init[2:-2, 2:-2, None] * np.ones(len(w)), w).sum()) * g.dx * g.dy
        assert abs(dm) < 1e-10

"""The oracle against the reference's own regression fixtures
(tests/regression/momopt/<ic>/mn.pn; procedure of site_scons/testing.py:109-180).

gaussian and smooth reproduce to rounding.  delta and lattice differ by a small,
mirror-symmetric amount that no rounding-level change moves (see the module
docstring of oracle/mn_oracle.c): those two are held to the measured level so
that any regression shows."""
import numpy as np
import pytest

from helpers import FIXTURE_T_FINAL, density, fixture_deck, fixture_path, mass, oracle, read_pn, rel_l2

LIMIT = {"gaussian": 1e-10, "smooth": 1e-10, "delta": 1.0e-5, "lattice": 1.0e-6}


@pytest.mark.parametrize("ic", ["gaussian", "smooth", "delta", "lattice"])
def test_fixture(ic):
    o = oracle()
    d = fixture_deck(ic)
    s = o.Solver(d, nthreads=8)
    m0 = s.moments().copy()
    steps = s.run(FIXTURE_T_FINAL)
    m = s.moments()
    h, ref = read_pn(fixture_path(ic))
    assert h[:2] == (36, 74) and h[6] == 10 and steps == 84
    dx, dy = (h[4] - h[2]) / h[0], (h[5] - h[3]) / h[1]
    err = rel_l2(density(m), density(ref))
    print("%s rel L2 %.3e" % (ic, err))
    assert err < LIMIT[ic]
    dm = mass(density(m), dx, dy) - mass(density(m0), dx, dy)
    dm_ref = mass(density(ref), dx, dy) - mass(density(m0), dx, dy)
    if ic == "smooth":
        assert abs(dm) < 1e-10
    if ic == "delta":
        # outflow through the boundary: the same as the fixture's own
        assert abs(dm - dm_ref) < 1e-11
    st = s.stats()
    assert st.num_dual_solves == 84 * 2 * 40 * 78

"""The C-ABI library on a machine without a GPU: it loads, exports every
function include/mnclosure.h declares, and a host-only context (device < 0)
builds the same tables as the oracle by an independent method."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

from helpers import PKG, deck_text, fixture_deck, load_pkg, oracle


def test_library_exports_header_symbols():
    mnc = load_pkg("mnc")
    syms = mnc.header_symbols()
    assert len(syms) >= 20
    L = ctypes.CDLL(mnc.LIB_PATH)
    missing = [s for s in syms if not hasattr(L, s)]
    assert not missing, missing


@pytest.mark.parametrize("N,Q,gaunt", [(2, 12, False), (3, 30, True), (4, 16, True), (4, 16, False)])
def test_host_tables_match_oracle(N, Q, gaunt):
    mnc = load_pkg("mnc")
    o = oracle()
    ctx = mnc.Context(N, Q, use_gaunt=gaunt, device=-1)
    w, p, xi, eta = ctx.tables()
    t = o.Tables(N, Q, gaunt)
    np.testing.assert_allclose(w, t.w, rtol=1e-14, atol=0)
    np.testing.assert_allclose(p, t.p, rtol=0, atol=2e-14)
    np.testing.assert_allclose(xi, t.xi, atol=1e-15)
    np.testing.assert_allclose(eta, t.eta, atol=1e-15)
    if gaunt:
        ptr, k, v = ctx.gaunt()
        assert np.array_equal(ptr, t.gptr) and np.array_equal(k, t.gk[:len(k)])
        np.testing.assert_allclose(v, t.gv[:len(v)], atol=1e-13)
    ctx.close()


def test_host_only_context_rejects_device_calls():
    mnc = load_pkg("mnc")
    ctx = mnc.Context(2, 12, device=-1)
    with pytest.raises(mnc.MncError) as e:
        ctx.opt_batch_host(np.ones((4, ctx.nm)), np.zeros((4, ctx.nm)))
    assert e.value.code == mnc.ERR_STATE
    ctx.close()


def test_bad_setup_rejected():
    mnc = load_pkg("mnc")
    with pytest.raises(mnc.MncError):
        mnc.Context(2, 11, device=-1)     # odd quadrature order


def test_driver_reports_missing_deck_key(tmp_path):
    """A missing key aborts with the reference's message (main.cpp:33-39)."""
    text = "\n".join(l for l in deck_text(fixture_deck("gaussian"), 1.03).splitlines()
                     if not l.startswith("SIGMA ")) + "\n"
    (tmp_path / "input.deck").write_text(text)
    r = subprocess.run([os.path.join(PKG, "momopt_gpu")], cwd=tmp_path, capture_output=True, text=True,
                       timeout=60)
    assert r.returncode != 0
    assert "main.cpp:input error at line 83" in r.stdout

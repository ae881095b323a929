"""The no-allocation entry point (mnc_opt_batch_noalloc) against the entry
point it varies (mnc_opt_batch), on the same device arrays: the same kernels
on the same inputs, so the outputs agree bit for bit.  Two calls in a row, the
second with a smaller batch, so what the first call left in the library's
fixed scratch array cannot leak into the second; FP64 and mixed precision, Mn
and PPn; and the limits it rejects."""
import ctypes as C

import numpy as np
import pytest

from helpers import config2_batch, isotropic_start, load_pkg, oracle

_hip = None


def hip():
    global _hip
    if _hip is None:
        _hip = C.CDLL("libamdhip64.so")
        _hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        _hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        _hip.hipFree.argtypes = [C.c_void_p]
    return _hip


class DevArray:
    """A device copy of a host array (the caller's buffers; the entry point
    under test allocates nothing)."""

    def __init__(self, host):
        self.nbytes = host.nbytes
        self.ptr = C.c_void_p()
        assert hip().hipMalloc(C.byref(self.ptr), max(self.nbytes, 8)) == 0
        assert hip().hipMemcpy(self.ptr, host.ctypes.data, self.nbytes, 1) == 0

    def get(self, like):
        out = np.empty_like(like)
        assert hip().hipMemcpy(out.ctypes.data, self.ptr, self.nbytes, 2) == 0
        return out

    def free(self):
        hip().hipFree(self.ptr)


def _solve(mnc, ctx, fn, u, a, opt):
    outs_host = np.zeros(len(u) * 3)           # mnc_outs: r, iter (+ padding), norm_g
    du, da, do = DevArray(u), DevArray(a), DevArray(outs_host)
    st = getattr(mnc.lib(), fn)(ctx.h, len(u), du.ptr, da.ptr, C.byref(opt), do.ptr)
    res = (st, du.get(u), da.get(a), do.get(outs_host))
    for d in (du, da, do):
        d.free()
    return res


def _inputs(order, Q, n, seed, lobes=0.2):
    t = oracle().Tables(order, Q, False)
    u, _, _ = config2_batch(t.w, t.p, t.xi, t.eta, t.nm, n, np.random.default_rng(seed), near_boundary_frac=lobes)
    return u, isotropic_start(u)


@pytest.mark.gpu
@pytest.mark.parametrize("order,Q,prec,ppn", [(4, 16, 0, False), (4, 16, 1, False), (2, 12, 0, True),
                                              (3, 12, 1, False)])
def test_noalloc_matches_opt_batch(order, Q, prec, ppn):
    mnc = load_pkg("mnc")
    ctx = mnc.Context(order, Q, use_gaunt=False, moment_type=mnc.MOMENT_PPN if ppn else mnc.MOMENT_MN)
    opt = mnc.options(tol=1e-8, precision=prec, line_search=mnc.LS_ROBUST, delta_ppn=1e-8 if ppn else 0.0)
    for call, (n, seed) in enumerate(((12000, 5), (3000, 6))):
        print("case noalloc order=%d Q=%d prec=%d ppn=%d call=%d n=%d" % (order, Q, prec, ppn, call, n), flush=True)
        u, a = _inputs(order, Q, n, seed)
        if ppn:
            a = np.zeros_like(u)
            a[:, 0] = u[:, 0]
        ref = _solve(mnc, ctx, "mnc_opt_batch", u, a, opt)
        got = _solve(mnc, ctx, "mnc_opt_batch_noalloc", u, a, opt)
        assert got[0] == ref[0] and (ppn or ref[0] == mnc.OK)
        for x, y in zip(ref[1:], got[1:]):
            np.testing.assert_array_equal(x, y)
        assert not np.array_equal(got[2], a)          # something was solved
    ctx.close()


@pytest.mark.gpu
def test_noalloc_limits():
    mnc = load_pkg("mnc")
    ctx = mnc.Context(4, 16, use_gaunt=False)
    u, a = _inputs(4, 16, mnc.NOALLOC_MAX_CELLS + 1, 7, lobes=0.0)
    st = _solve(mnc, ctx, "mnc_opt_batch_noalloc", u, a, mnc.options(tol=1e-8))[0]
    assert st == mnc.ERR_ARG
    st = _solve(mnc, ctx, "mnc_opt_batch_noalloc", u[:100], a[:100],
                mnc.options(tol=1e-8, warm_start=mnc.WARM_LEARNED))[0]
    assert st == mnc.ERR_ARG
    st, uo, ao, _ = _solve(mnc, ctx, "mnc_opt_batch_noalloc", u[:mnc.NOALLOC_MAX_CELLS], a[:mnc.NOALLOC_MAX_CELLS],
                           mnc.options(tol=1e-8))
    assert st == mnc.OK
    ctx.close()
    ctx = mnc.Context(8, 32, use_gaunt=False)
    u8 = np.zeros((4, 45))
    u8[:, 0] = 1.0
    st = _solve(mnc, ctx, "mnc_opt_batch_noalloc", u8, np.zeros_like(u8), mnc.options(tol=1e-8))[0]
    assert st == mnc.ERR_ARG
    ctx.close()

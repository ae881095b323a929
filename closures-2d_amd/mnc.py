"""Python view of the C-ABI in include/mnclosure.h (libmnclosure.so).

The library is the product: hand-written gfx950 HIP kernels behind a plain C
interface.  This module only marshals numpy arrays through ctypes; it has no
compute fallback and raises if the library is missing or a call fails.
"""
import ctypes as C
import os
import re

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmnclosure.so")
HEADER = os.path.join(HERE, os.pardir, "include", "mnclosure.h")

OK, ERR_ARG, ERR_U0, ERR_MAXREG, ERR_HIP, ERR_STATE = 0, 1, 2, 3, 4, 5
PREC_FP64, PREC_MIXED = 0, 1
LS_REFERENCE, LS_ROBUST = 0, 1
WARM_PREVIOUS, WARM_ISOTROPIC, WARM_LEARNED = 0, 1, 2
NOALLOC_MAX_CELLS = 16384   # MNC_NOALLOC_MAX_CELLS
OPT_NEWTON, OPT_CHANGE_BASIS, OPT_BFGS = 0, 1, 2
MOMENT_MN, MOMENT_PPN = 0, 1
KCLASS_NEWTON, KCLASS_FLUX, KCLASS_OTHER, KCLASS_MLP, KCLASS_TAIL = 0, 1, 2, 3, 4

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


class Setup(C.Structure):
    _fields_ = [("moment_order", C.c_int), ("quad_order", C.c_int), ("use_gaunt", C.c_int),
                ("moment_type", C.c_int), ("device", C.c_int)]


class Options(C.Structure):
    _fields_ = [("max_iter", C.c_int), ("tol_rel", C.c_double), ("cond_h_max", C.c_double),
                ("precision", C.c_int), ("warm_start", C.c_int), ("line_search", C.c_int),
                ("delta_ppn", C.c_double), ("opt_type", C.c_int), ("max_bfgs_iter", C.c_int),
                ("cond_h_max_bfgs", C.c_double), ("tol_gamma", C.c_double)]


class Outs(C.Structure):
    _fields_ = [("r", C.c_double), ("iter", C.c_int), ("norm_g", C.c_double)]


class GridDesc(C.Structure):
    _fields_ = [("num_cells_x", C.c_int), ("num_cells_y", C.c_int), ("dx", C.c_double),
                ("dy", C.c_double), ("init_x", C.c_double), ("init_y", C.c_double),
                ("init_cond", C.c_int), ("theta", C.c_double), ("delta_ppn", C.c_double)]


class Stats(C.Structure):
    _fields_ = [("max_iter", C.c_int), ("num_dual_solves", C.c_longlong),
                ("iter_sum", C.c_longlong), ("hist_reg", C.c_longlong * 5),
                ("trial_sum", C.c_longlong), ("warm_pick", C.c_longlong * 3),
                ("mixed_deferred", C.c_longlong), ("tail_cells", C.c_longlong),
                ("tail_iter_sum", C.c_longlong), ("tail_trial_sum", C.c_longlong),
                ("warm_fallback", C.c_longlong)]


class MncError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("%s (status %d)" % (msg, code))
        self.code = code


def header_symbols():
    """Function names declared in include/mnclosure.h."""
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(mnc_\w+)\s*\(", text, re.M)))


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise OSError("libmnclosure.so is not built (run __graft_entry__.build()); "
                          "there is no fallback path")
        L = C.CDLL(LIB_PATH)
        vp = C.c_void_p
        L.mnc_create.argtypes = [C.POINTER(Setup), C.POINTER(vp)]
        L.mnc_destroy.argtypes = [vp]
        L.mnc_destroy.restype = None
        for f in ("mnc_num_moments", "mnc_num_quad", "mnc_table_stride", "mnc_gaunt_nnz"):
            getattr(L, f).argtypes = [vp]
        L.mnc_get_tables.argtypes = [vp, _dp, _dp, _dp, _dp]
        L.mnc_get_gaunt.argtypes = [vp, _ip, _ip, _dp]
        L.mnc_opt_batch.argtypes = [vp, C.c_int, vp, vp, C.POINTER(Options), vp]
        L.mnc_opt_batch_noalloc.argtypes = [vp, C.c_int, vp, vp, C.POINTER(Options), vp]
        L.mnc_opt_batch_host.argtypes = [vp, C.c_int, _dp, _dp, C.POINTER(Options), C.POINTER(Outs)]
        L.mnc_optcb_batch_host.argtypes = [vp, C.c_int, _dp, _dp, _dp, _dp, C.POINTER(Options), C.POINTER(Outs)]
        L.mnc_batch_set.argtypes = [vp, C.c_int, _dp, _dp]
        L.mnc_batch_solve.argtypes = [vp, C.POINTER(Options)]
        L.mnc_batch_get.argtypes = [vp, _dp, _dp, C.POINTER(Outs)]
        L.mnc_grid_create.argtypes = [vp, C.POINTER(GridDesc), _dp, _dp, _dp, _dp]
        L.mnc_grid_get.argtypes = [vp, _dp, _dp]
        L.mnc_grid_set.argtypes = [vp, _dp, _dp]
        L.mnc_grid_get_basis.argtypes = [vp, _dp, _dp]
        L.mnc_grid_set_basis.argtypes = [vp, _dp, _dp]
        L.mnc_ghost_fill.argtypes = [vp]
        L.mnc_solve_grid.argtypes = [vp, C.POINTER(Options)]
        L.mnc_flux.argtypes = [vp, _dp]
        L.mnc_update.argtypes = [vp, C.c_double, C.POINTER(Options)]
        L.mnc_get_stats.argtypes = [vp, C.POINTER(Stats)]
        L.mnc_reset_stats.argtypes = [vp]
        L.mnc_sync.argtypes = [vp]
        L.mnc_timing_enable.argtypes = [vp, C.c_int]
        L.mnc_timing_get.argtypes = [vp, C.c_int, C.POINTER(C.c_longlong), C.POINTER(C.c_double)]
        L.mnc_tune.argtypes = [vp, C.c_char_p, C.c_int]
        L.mnc_autotune.argtypes = [vp, C.c_double, C.POINTER(Options), C.c_int, C.c_char_p, C.c_int]
        L.mnc_get_tune.argtypes = [vp, C.c_char_p, C.POINTER(C.c_int)]
        L.mnc_sn_run.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double, C.c_int,
                                 _dp, _dp, C.c_double, C.c_int, _dp, C.c_int, C.POINTER(C.c_double)]
        L.mnc_last_error.argtypes = [vp]
        L.mnc_last_error.restype = C.c_char_p
        _fp = C.POINTER(C.c_float)
        L.mnc_mlp_num_params.argtypes = []
        L.mnc_mlp_init.argtypes = [vp, C.c_ulonglong]
        L.mnc_mlp_train.argtypes = [vp, C.c_longlong, _dp, _dp, C.c_int, C.c_int, C.c_double, C.c_ulonglong, _fp]
        L.mnc_mlp_predict.argtypes = [vp, C.c_int, _dp, _dp]
        L.mnc_mlp_get_params.argtypes = [vp, _fp]
        L.mnc_mlp_set_params.argtypes = [vp, _fp]
        L.mnc_mlp_gradient.argtypes = [vp, C.c_int, _dp, _dp, _fp, _fp]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(_dp)


def options(max_iter=100, tol=1e-8, cond_h_max=1e10, precision=PREC_FP64, warm_start=WARM_PREVIOUS,
            line_search=LS_REFERENCE, delta_ppn=0.0, opt_type=0, max_bfgs_iter=5, cond_h_max_bfgs=20.0,
            tol_gamma=0.0):
    return Options(max_iter, tol, cond_h_max, precision, warm_start, line_search, delta_ppn, opt_type, max_bfgs_iter,
                   cond_h_max_bfgs, tol_gamma)


class Context:
    """One mnc_ctx: tables, a HIP stream and (optionally) one grid."""

    def __init__(self, moment_order, quad_order, use_gaunt=False, device=0, moment_type=MOMENT_MN):
        h = C.c_void_p()
        st = lib().mnc_create(C.byref(Setup(moment_order, quad_order, int(bool(use_gaunt)), moment_type, device)),
                              C.byref(h))
        if st != OK:
            raise MncError(st, "mnc_create failed")
        self.h = h
        L = lib()
        self.nm = L.mnc_num_moments(h)
        self.nq = L.mnc_num_quad(h)
        self.np2 = L.mnc_table_stride(h)
        self.use_gaunt = bool(use_gaunt)
        self.moment_type = moment_type
        self.grid = None

    def close(self):
        if self.h:
            lib().mnc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, st, what):
        if st != OK:
            raise MncError(st, "%s: %s" % (what, lib().mnc_last_error(self.h).decode()))

    def tables(self):
        w = np.zeros(self.nq)
        p = np.zeros((self.nq, self.np2))
        xi = np.zeros(self.nq)
        eta = np.zeros(self.nq)
        self._check(lib().mnc_get_tables(self.h, _p(w), _p(p), _p(xi), _p(eta)), "mnc_get_tables")
        return w, p, xi, eta

    def gaunt(self):
        nnz = lib().mnc_gaunt_nnz(self.h)
        ptr = np.zeros(self.nm * self.nm + 1, dtype=np.int32)
        k = np.zeros(max(nnz, 1), dtype=np.int32)
        v = np.zeros(max(nnz, 1))
        self._check(lib().mnc_get_gaunt(self.h, ptr.ctypes.data_as(_ip), k.ctypes.data_as(_ip), _p(v)),
                    "mnc_get_gaunt")
        return ptr, k[:nnz], v[:nnz]

    def batch_set(self, u, alpha):
        """Stages a batch in device memory (mnc_batch_set)."""
        u = np.ascontiguousarray(u, dtype=np.float64).reshape(-1, self.nm)
        a = np.ascontiguousarray(alpha, dtype=np.float64).reshape(-1, self.nm)
        self._check(lib().mnc_batch_set(self.h, u.shape[0], _p(u), _p(a)), "mnc_batch_set")
        self._batch_n = u.shape[0]

    def batch_solve(self, opt):
        self._check(lib().mnc_batch_solve(self.h, C.byref(opt)), "mnc_batch_solve")

    def batch_get(self):
        n = self._batch_n
        u = np.zeros((n, self.nm))
        a = np.zeros((n, self.nm))
        outs = (Outs * n)()
        self._check(lib().mnc_batch_get(self.h, _p(u), _p(a), outs), "mnc_batch_get")
        it = np.array([o.iter for o in outs], dtype=np.int32)
        r = np.array([o.r for o in outs])
        return u, a, it, r

    def opt_batch_host(self, u, alpha, opt=None, allow_status=(OK,)):
        """Batched opt() on host arrays; returns (status, u, alpha, iter, r, normG)."""
        opt = opt or options()
        u = np.array(u, dtype=np.float64, order="C").reshape(-1, self.nm)
        alpha = np.array(alpha, dtype=np.float64, order="C").reshape(-1, self.nm)
        n = u.shape[0]
        outs = (Outs * max(n, 1))()
        st = lib().mnc_opt_batch_host(self.h, n, _p(u), _p(alpha), C.byref(opt), outs)
        if st not in allow_status:
            self._check(st, "mnc_opt_batch_host")
        it = np.array([o.iter for o in outs[:n]], dtype=np.int32)
        r = np.array([o.r for o in outs[:n]])
        ng = np.array([o.norm_g for o in outs[:n]])
        return st, u, alpha, it, r, ng

    def optcb_batch_host(self, u, P2M, alphaP, opt=None, allow_status=(OK,)):
        """The change-of-basis optimiser (optcb.cpp:90-415) on host arrays: P2M
        [n][nm*nm] (column-major bases) and alphaP [n][nm] are the per-cell
        state, updated; returns (status, u, P2M, alphaP, alphaM, iter, r,
        normG)."""
        opt = opt or options()
        u = np.array(u, dtype=np.float64, order="C").reshape(-1, self.nm)
        n = u.shape[0]
        P2M = np.array(P2M, dtype=np.float64, order="C").reshape(n, self.nm * self.nm)
        alphaP = np.array(alphaP, dtype=np.float64, order="C").reshape(n, self.nm)
        alphaM = np.zeros_like(alphaP)
        outs = (Outs * max(n, 1))()
        st = lib().mnc_optcb_batch_host(self.h, n, _p(u), _p(P2M), _p(alphaP), _p(alphaM), C.byref(opt), outs)
        if st not in allow_status:
            self._check(st, "mnc_optcb_batch_host")
        it = np.array([o.iter for o in outs[:n]], dtype=np.int32)
        r = np.array([o.r for o in outs[:n]])
        ng = np.array([o.norm_g for o in outs[:n]])
        return st, u, P2M, alphaP, alphaM, it, r, ng

    # ---- grid ----
    def grid_create(self, nx, ny, dx, dy, init_x, init_y, init_cond, theta, moments, alpha, sigma_t, sigma_s,
                    delta_ppn=0.0):
        gd = GridDesc(nx, ny, dx, dy, init_x, init_y, init_cond, theta, delta_ppn)
        m = np.ascontiguousarray(moments, dtype=np.float64).ravel()
        a = np.ascontiguousarray(alpha, dtype=np.float64).ravel()
        st_ = np.ascontiguousarray(sigma_t, dtype=np.float64).ravel()
        ss_ = np.ascontiguousarray(sigma_s, dtype=np.float64).ravel()
        self._check(lib().mnc_grid_create(self.h, C.byref(gd), _p(m), _p(a), _p(st_), _p(ss_)),
                    "mnc_grid_create")
        self.grid = (nx, ny)

    def grid_get(self):
        nx, ny = self.grid
        n = (nx + 4) * (ny + 4) * self.nm
        m = np.zeros(n)
        a = np.zeros(n)
        self._check(lib().mnc_grid_get(self.h, _p(m), _p(a)), "mnc_grid_get")
        shp = (nx + 4, ny + 4, self.nm)
        return m.reshape(shp), a.reshape(shp)

    def grid_set(self, moments, alpha):
        m = np.ascontiguousarray(moments, dtype=np.float64).ravel()
        a = np.ascontiguousarray(alpha, dtype=np.float64).ravel()
        self._check(lib().mnc_grid_set(self.h, _p(m), _p(a)), "mnc_grid_set")

    def grid_get_basis(self):
        """The change-of-basis state (P2M [cells][nm*nm] column-major, alphaP
        [cells][nm]) of every cell including ghosts (mnc_grid_get_basis)."""
        nx, ny = self.grid
        nc = (nx + 4) * (ny + 4)
        B = np.zeros((nc, self.nm * self.nm))
        aP = np.zeros((nc, self.nm))
        self._check(lib().mnc_grid_get_basis(self.h, _p(B), _p(aP)), "mnc_grid_get_basis")
        return B, aP

    def grid_set_basis(self, P2M, alphaP):
        B = np.ascontiguousarray(P2M, dtype=np.float64).ravel()
        aP = np.ascontiguousarray(alphaP, dtype=np.float64).ravel()
        self._check(lib().mnc_grid_set_basis(self.h, _p(B), _p(aP)), "mnc_grid_set_basis")

    def ghost_fill(self):
        self._check(lib().mnc_ghost_fill(self.h), "mnc_ghost_fill")

    def solve_grid(self, opt):
        self._check(lib().mnc_solve_grid(self.h, C.byref(opt)), "mnc_solve_grid")

    def flux(self):
        nx, ny = self.grid
        out = np.zeros((nx + 4) * (ny + 4) * self.nm)
        self._check(lib().mnc_flux(self.h, _p(out)), "mnc_flux")
        return out.reshape(nx + 4, ny + 4, self.nm)

    def update(self, dt, opt):
        self._check(lib().mnc_update(self.h, dt, C.byref(opt)), "mnc_update")

    def stats(self):
        s = Stats()
        self._check(lib().mnc_get_stats(self.h, C.byref(s)), "mnc_get_stats")
        return s

    def reset_stats(self):
        self._check(lib().mnc_reset_stats(self.h), "mnc_reset_stats")

    def sync(self):
        self._check(lib().mnc_sync(self.h), "mnc_sync")

    def tune(self, key, value):
        self._check(lib().mnc_tune(self.h, key.encode(), int(value)), "mnc_tune")

    def autotune(self, dt, opt, steps=3):
        """Picks the fastest launch shape for steps of dt on the current grid
        (state restored afterwards); returns the trial report lines."""
        buf = C.create_string_buffer(8192)
        self._check(lib().mnc_autotune(self.h, dt, C.byref(opt), steps, buf, len(buf)), "mnc_autotune")
        return buf.value.decode().splitlines()

    def get_tune(self, key):
        v = C.c_int(0)
        self._check(lib().mnc_get_tune(self.h, key.encode(), C.byref(v)), "mnc_get_tune")
        return v.value

    def timing_enable(self, on=True):
        self._check(lib().mnc_timing_enable(self.h, int(on)), "mnc_timing_enable")

    # ---- learned warm start ----
    def mlp_init(self, seed=1):
        self._check(lib().mnc_mlp_init(self.h, seed), "mnc_mlp_init")

    def mlp_train(self, u_scaled, alpha_scaled, steps, batch=4096, lr=1e-3, seed=1):
        """Adam steps on (u/u0, scaled alpha) pairs; returns the per-step losses."""
        us = np.ascontiguousarray(u_scaled, dtype=np.float64).reshape(-1, 15)
        a = np.ascontiguousarray(alpha_scaled, dtype=np.float64).reshape(-1, 15)
        loss = np.zeros(max(steps, 1), dtype=np.float32)
        self._check(lib().mnc_mlp_train(self.h, us.shape[0], _p(us), _p(a), steps, batch, lr, seed,
                                        loss.ctypes.data_as(C.POINTER(C.c_float))), "mnc_mlp_train")
        return loss[:steps]

    def mlp_predict(self, u_scaled):
        us = np.ascontiguousarray(u_scaled, dtype=np.float64).reshape(-1, 15)
        out = np.zeros_like(us)
        self._check(lib().mnc_mlp_predict(self.h, us.shape[0], _p(us), _p(out)), "mnc_mlp_predict")
        return out

    def mlp_params(self):
        out = np.zeros(lib().mnc_mlp_num_params(), dtype=np.float32)
        self._check(lib().mnc_mlp_get_params(self.h, out.ctypes.data_as(C.POINTER(C.c_float))), "mnc_mlp_get_params")
        return out

    def mlp_set_params(self, params):
        p = np.ascontiguousarray(params, dtype=np.float32)
        assert p.size == lib().mnc_mlp_num_params()
        self._check(lib().mnc_mlp_set_params(self.h, p.ctypes.data_as(C.POINTER(C.c_float))), "mnc_mlp_set_params")

    def mlp_gradient(self, u_scaled, alpha_scaled):
        us = np.ascontiguousarray(u_scaled, dtype=np.float64).reshape(-1, 15)
        a = np.ascontiguousarray(alpha_scaled, dtype=np.float64).reshape(-1, 15)
        g = np.zeros(37007, dtype=np.float32)
        loss = C.c_float(0.0)
        self._check(lib().mnc_mlp_gradient(self.h, us.shape[0], _p(us), _p(a), g.ctypes.data_as(C.POINTER(C.c_float)),
                                           C.byref(loss)), "mnc_mlp_gradient")
        return g, float(loss.value)

    def timing_get(self, kclass):
        n = C.c_longlong(0)
        ms = C.c_double(0.0)
        self._check(lib().mnc_timing_get(self.h, kclass, C.byref(n), C.byref(ms)), "mnc_timing_get")
        return n.value, ms.value


def sn_run(f, sigma_t, sigma_s, quad_order, dx, dy, init_x, init_y, init_cond, dt, steps, device=0):
    """`steps` Heun steps of the Sn solver on the GPU (mnc_sn_run): f is
    [nx+4][ny+4][Q*Q] with ghosts, sigma_t/sigma_s [nx+4][ny+4].  Returns
    (f after the steps, kernel milliseconds)."""
    f = np.array(f, dtype=np.float64, order="C")
    st = np.ascontiguousarray(sigma_t, dtype=np.float64)
    ss = np.ascontiguousarray(sigma_s, dtype=np.float64)
    nx, ny = f.shape[0] - 4, f.shape[1] - 4
    if f.shape[2] != quad_order * quad_order or st.shape != f.shape[:2] or ss.shape != f.shape[:2]:
        raise ValueError("Sn state must be [nx+4][ny+4][Q*Q], sigma [nx+4][ny+4]")
    ms = C.c_double(0.0)
    rc = lib().mnc_sn_run(nx, ny, quad_order, dx, dy, init_x, init_y, init_cond, _p(st), _p(ss), dt, steps, _p(f),
                          device, C.byref(ms))
    if rc != OK:
        raise MncError(rc, "mnc_sn_run failed")
    return f, ms.value


def scaled_pairs(u, alpha):
    """(u/u0, alpha of the u0 = 1 problem) from unscaled moments and multipliers:
    the shift of opt.cpp:91-92 (alpha_0 - 2 sqrt(pi) ln u0)."""
    u = np.asarray(u, dtype=np.float64).reshape(-1, u.shape[-1])
    a = np.array(alpha, dtype=np.float64).reshape(-1, u.shape[-1])
    us = u / u[:, :1]
    a[:, 0] -= 2.0 * np.sqrt(np.pi) * np.log(u[:, 0])
    return us, a

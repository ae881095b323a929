"""ctypes binding of the CPU FP64 oracle (oracle/libmnoracle.so).

TEST INFRASTRUCTURE ONLY: the checker for the HIP path.  Only tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg import this module;
the product (closures-2d_amd/) never does.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("MN_ORACLE_LIB") or os.path.join(HERE, "libmnoracle.so")
XP_LIB_PATH = os.path.join(HERE, "libmnoracle_xp.so")   # exp in long double (sensitivity probe)
FMA_LIB_PATH = os.path.join(HERE, "libmnoracle_fma.so")  # FMA contraction (sensitivity probe)

OK, ERR_ARG, ERR_U0, ERR_MAXREG, ERR_ALLOC = 0, 1, 2, 3, 4
REGULARIZATIONS = (0.0, 1e-3, 1e-2, 1e-1, 5e-1)  # opt.h:39-40

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


MOMENT_MN, MOMENT_PPN = 0, 1


class Options(C.Structure):
    """OPTIONS (opt.h:11-22), the fields the opt path reads."""
    _fields_ = [("max_iter", C.c_int), ("tol_rel", C.c_double),
                ("cond_h_max", C.c_double), ("use_gaunt", C.c_int), ("line_search", C.c_int),
                ("moment_type", C.c_int), ("delta_ppn", C.c_double),
                ("max_bfgs_iter", C.c_int), ("cond_h_max_bfgs", C.c_double),
                ("tol_gamma", C.c_double), ("trace", _dp), ("trace_cap", C.c_int),
                ("trace_len", _ip)]


class Outs(C.Structure):
    """OUTS (opt.h:25-32)."""
    _fields_ = [("r", C.c_double), ("iter", C.c_int), ("norm_g", C.c_double)]


class Deck(C.Structure):
    _fields_ = [("num_cells_x", C.c_int), ("num_cells_y", C.c_int),
                ("a_x", C.c_double), ("b_x", C.c_double),
                ("a_y", C.c_double), ("b_y", C.c_double),
                ("gaussian_sigma", C.c_double), ("floor_value", C.c_double),
                ("sigma", C.c_double), ("init_cond", C.c_int),
                ("moment_order", C.c_int), ("quad_order", C.c_int),
                ("cfl_factor", C.c_double), ("tol", C.c_double),
                ("cond_h_max", C.c_double), ("max_iter", C.c_int),
                ("use_gaunt", C.c_int), ("theta", C.c_double), ("line_search", C.c_int),
                ("moment_type", C.c_int), ("delta_ppn", C.c_double), ("tol_gamma", C.c_double)]


class Stats(C.Structure):
    _fields_ = [("max_iter", C.c_int), ("num_dual_solves", C.c_longlong),
                ("iter_sum", C.c_longlong), ("hist_reg", C.c_longlong * 5)]


def build():
    subprocess.check_call(["make", "-s", "-C", HERE])


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.orc_gauss_legendre.argtypes = [C.c_int, _dp, _dp]
        L.orc_sph_plm.argtypes = [C.c_int, C.c_int, C.c_double]
        L.orc_sph_plm.restype = C.c_double
        L.orc_tables_create.argtypes = [C.c_int, C.c_int, C.c_int, C.POINTER(C.c_void_p)]
        L.orc_tables_destroy.argtypes = [C.c_void_p]
        for f in ("orc_tables_nm", "orc_tables_np2", "orc_tables_nq", "orc_tables_gaunt_nnz"):
            getattr(L, f).argtypes = [C.c_void_p]
        L.orc_tables_get.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp]
        L.orc_tables_gaunt_get.argtypes = [C.c_void_p, _ip, _ip, _dp]
        L.orc_fobj.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp, C.c_int]
        L.orc_fobj_ppn.argtypes = [C.c_void_p, _dp, _dp, C.c_double, _dp, _dp, C.c_int]
        L.orc_fobj_ppn.restype = C.c_double
        L.orc_fobj.restype = C.c_double
        L.orc_cholesky.argtypes = [C.c_int, _dp]
        L.orc_rcond.argtypes = [C.c_int, _dp, C.c_double]
        L.orc_rcond.restype = C.c_double
        L.orc_norm1.argtypes = [C.c_int, _dp]
        L.orc_norm1.restype = C.c_double
        L.orc_opt.argtypes = [C.c_void_p, _dp, _dp, C.POINTER(Options), C.POINTER(Outs)]
        L.orc_opt_batch.argtypes = [C.c_void_p, C.c_int, _dp, _dp, C.POINTER(Options),
                                    C.POINTER(Outs), C.c_int]
        L.orc_optcb_batch.argtypes = [C.c_void_p, C.c_int, _dp, _dp, _dp, _dp, C.POINTER(Options),
                                      C.POINTER(Outs), C.c_int]
        L.orc_optbfgs_batch.argtypes = [C.c_void_p, C.c_int, _dp, _dp, _dp, _dp, C.POINTER(Options),
                                        C.POINTER(Outs), C.c_int]
        L.orc_bfgs_update.argtypes = [C.c_int, _dp, _dp, _dp, _dp]
        L.orc_sn_step.argtypes = [C.c_int, C.c_int, C.c_int, _dp, _dp, _dp, _dp, _dp, C.c_int, C.c_double,
                                  C.c_double, C.c_double, C.c_double, C.c_double, _dp]
        L.orc_solver_create.argtypes = [C.POINTER(Deck), C.c_int, C.POINTER(C.c_void_p)]
        L.orc_solver_destroy.argtypes = [C.c_void_p]
        L.orc_solver_max_dt.argtypes = [C.c_void_p]
        L.orc_solver_max_dt.restype = C.c_double
        L.orc_solver_nm.argtypes = [C.c_void_p]
        L.orc_solver_update.argtypes = [C.c_void_p, C.c_double]
        L.orc_solver_run.argtypes = [C.c_void_p, C.c_double, C.c_double, _ip]
        L.orc_solver_get_state.argtypes = [C.c_void_p, _dp, _dp]
        L.orc_solver_set_state.argtypes = [C.c_void_p, _dp, _dp]
        L.orc_solver_get_moments.argtypes = [C.c_void_p, _dp]
        L.orc_solver_get_stats.argtypes = [C.c_void_p, C.POINTER(Stats)]
        L.orc_solver_solve_all.argtypes = [C.c_void_p]
        L.orc_solver_flux.argtypes = [C.c_void_p, _dp]
        L.orc_solver_ghost_fill.argtypes = [C.c_void_p]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(_dp)


def gauss_legendre(n):
    x = np.zeros(n)
    w = np.zeros(n)
    lib().orc_gauss_legendre(n, _p(x), _p(w))
    return x, w


def sph_plm(l, m, x):
    return lib().orc_sph_plm(l, m, float(x))


class Tables:
    """Quadrature, real-SH table and Gaunt map (momopt_init.cpp:88-259)."""

    def __init__(self, moment_order, quad_order, use_gaunt=False):
        h = C.c_void_p()
        st = lib().orc_tables_create(moment_order, quad_order, int(bool(use_gaunt)), C.byref(h))
        if st != OK:
            raise ValueError("orc_tables_create failed: %d" % st)
        self.h = h
        L = lib()
        self.nm = L.orc_tables_nm(h)
        self.np2 = L.orc_tables_np2(h)
        self.nq = L.orc_tables_nq(h)
        self.use_gaunt = bool(use_gaunt)
        self.w = np.zeros(self.nq)
        self.p = np.zeros((self.nq, self.np2))
        self.xi = np.zeros(self.nq)
        self.eta = np.zeros(self.nq)
        L.orc_tables_get(h, _p(self.w), _p(self.p), _p(self.xi), _p(self.eta))
        if self.use_gaunt:
            nnz = L.orc_tables_gaunt_nnz(h)
            self.gptr = np.zeros(self.nm * self.nm + 1, dtype=np.int32)
            self.gk = np.zeros(max(nnz, 1), dtype=np.int32)
            self.gv = np.zeros(max(nnz, 1))
            L.orc_tables_gaunt_get(h, self.gptr.ctypes.data_as(_ip),
                                   self.gk.ctypes.data_as(_ip), _p(self.gv))

    def __del__(self):
        try:
            lib().orc_tables_destroy(self.h)
        except Exception:
            pass

    def fobj(self, alpha, u, want_g=True, want_h=True, use_gaunt=None):
        alpha = np.ascontiguousarray(alpha, dtype=np.float64)
        u = np.ascontiguousarray(u, dtype=np.float64)
        g = np.zeros(self.nm) if (want_g or want_h) else None
        h = np.zeros((self.nm, self.nm)) if want_h else None
        ug = self.use_gaunt if use_gaunt is None else bool(use_gaunt)
        f = lib().orc_fobj(self.h, _p(alpha), _p(u), _p(g) if g is not None else None,
                           _p(h) if h is not None else None, int(ug))
        return f, g, h

    def fobj_ppn(self, alpha, u, delta, want_g=True, want_h=True, use_gaunt=None):
        alpha = np.ascontiguousarray(alpha, dtype=np.float64)
        u = np.ascontiguousarray(u, dtype=np.float64)
        g = np.zeros(self.nm) if want_g else None
        h = np.zeros((self.nm, self.nm)) if want_h else None
        ug = self.use_gaunt if use_gaunt is None else bool(use_gaunt)
        f = lib().orc_fobj_ppn(self.h, _p(alpha), _p(u), delta, _p(g) if g is not None else None,
                               _p(h) if h is not None else None, int(ug))
        return f, g, h

    def opt(self, u, alpha, max_iter=100, tol=1e-8, cond_h_max=1e10, use_gaunt=None, line_search=0,
            moment_type=MOMENT_MN, delta_ppn=0.0, tol_gamma=0.0):
        """One opt() call (opt.cpp:82-148); returns (status, u_out, alpha_out, outs).
        tol_gamma > 0 adds the supplementary stop test (opt.cpp:52-76,288-293)."""
        u = np.array(u, dtype=np.float64)
        alpha = np.array(alpha, dtype=np.float64)
        o = Options(max_iter, tol, cond_h_max, int(self.use_gaunt if use_gaunt is None else use_gaunt),
                    line_search, moment_type, delta_ppn, 0, 0.0, tol_gamma)
        outs = Outs()
        st = lib().orc_opt(self.h, _p(u), _p(alpha), C.byref(o), C.byref(outs))
        return st, u, alpha, outs

    def opt_trace(self, u, alpha, max_iter=100, tol=1e-8, cond_h_max=1e10, use_gaunt=None, line_search=0,
                  moment_type=MOMENT_MN, delta_ppn=0.0, tol_gamma=0.0, cap=4096):
        """One opt() call with optScaled's decisions recorded: returns (status,
        u_out, alpha_out, outs, records[k][7]) with each record (total
        iteration, regularization index, decision, condition estimate, |g|,
        rtol, gamma); decision 0 none, 1 iteration limit, 2 Cholesky failed,
        3 line search t = 0, 4 condition estimate above cond_h_max, 5 converged."""
        u = np.array(u, dtype=np.float64)
        alpha = np.array(alpha, dtype=np.float64)
        buf = np.zeros((cap, 7))
        n = C.c_int(0)
        o = Options(max_iter, tol, cond_h_max, int(self.use_gaunt if use_gaunt is None else use_gaunt),
                    line_search, moment_type, delta_ppn, 0, 0.0, tol_gamma, _p(buf), cap, C.pointer(n))
        outs = Outs()
        st = lib().orc_opt(self.h, _p(u), _p(alpha), C.byref(o), C.byref(outs))
        return st, u, alpha, outs, buf[:n.value].copy()

    def opt_batch(self, u, alpha, max_iter=100, tol=1e-8, cond_h_max=1e10, use_gaunt=None,
                  nthreads=0, line_search=0, moment_type=MOMENT_MN, delta_ppn=0.0, tol_gamma=0.0):
        u = np.array(u, dtype=np.float64, order="C")
        alpha = np.array(alpha, dtype=np.float64, order="C")
        n = u.shape[0]
        o = Options(max_iter, tol, cond_h_max, int(self.use_gaunt if use_gaunt is None else use_gaunt),
                    line_search, moment_type, delta_ppn, 0, 0.0, tol_gamma)
        outs = (Outs * n)()
        st = lib().orc_opt_batch(self.h, n, _p(u), _p(alpha), C.byref(o), outs, nthreads)
        it = np.array([o_.iter for o_ in outs], dtype=np.int32)
        r = np.array([o_.r for o_ in outs])
        ng = np.array([o_.norm_g for o_ in outs])
        return st, u, alpha, it, r, ng


    def optcb_batch(self, u, P2M, alphaP, max_iter=100, tol=1e-8, cond_h_max=1e10, nthreads=0,
                    line_search=0, moment_type=MOMENT_MN, delta_ppn=0.0):
        """The change-of-basis optimiser (optcb.cpp:90-415) per cell.  P2M is
        [n][n*n] (each cell's basis, column-major), alphaP [n][nm]; both are
        updated in place as the reference's per-cell state.  Returns (status,
        u, P2M, alphaP, alphaM, iterations, r, |g|)."""
        u = np.array(u, dtype=np.float64, order="C")
        n = u.shape[0]
        P2M = np.array(P2M, dtype=np.float64, order="C").reshape(n, self.nm * self.nm)
        alphaP = np.array(alphaP, dtype=np.float64, order="C")
        alphaM = np.zeros_like(alphaP)
        o = Options(max_iter, tol, cond_h_max, 0, line_search, moment_type, delta_ppn)
        outs = (Outs * n)()
        st = lib().orc_optcb_batch(self.h, n, _p(u), _p(P2M), _p(alphaP), _p(alphaM), C.byref(o), outs, nthreads)
        it = np.array([o_.iter for o_ in outs], dtype=np.int32)
        r = np.array([o_.r for o_ in outs])
        ng = np.array([o_.norm_g for o_ in outs])
        return st, u, P2M, alphaP, alphaM, it, r, ng

    def optbfgs_batch(self, u, P2M, alphaP, max_iter=100, tol=1e-8, cond_h_max=1e10, nthreads=0,
                      line_search=0, moment_type=MOMENT_MN, delta_ppn=0.0, max_bfgs_iter=5, cond_h_max_bfgs=20.0):
        """The BFGS variant (optbfgs.cpp:92-513) per cell, with optcb_batch's
        arrays and returns; iterations count the outer (full-Hessian) ones."""
        u = np.array(u, dtype=np.float64, order="C")
        n = u.shape[0]
        P2M = np.array(P2M, dtype=np.float64, order="C").reshape(n, self.nm * self.nm)
        alphaP = np.array(alphaP, dtype=np.float64, order="C")
        alphaM = np.zeros_like(alphaP)
        o = Options(max_iter, tol, cond_h_max, 0, line_search, moment_type, delta_ppn, max_bfgs_iter,
                    cond_h_max_bfgs)
        outs = (Outs * n)()
        st = lib().orc_optbfgs_batch(self.h, n, _p(u), _p(P2M), _p(alphaP), _p(alphaM), C.byref(o), outs, nthreads)
        it = np.array([o_.iter for o_ in outs], dtype=np.int32)
        r = np.array([o_.r for o_ in outs])
        ng = np.array([o_.norm_g for o_ in outs])
        return st, u, P2M, alphaP, alphaM, it, r, ng


def bfgs_update(H, Hinv, s, y):
    """One BFGS update of H and H^-1 (optbfgs.cpp:440-477); returns new copies."""
    n = len(s)
    Hc = np.array(H, dtype=np.float64).reshape(n, n).T.copy()      # column-major storage
    Ic = np.array(Hinv, dtype=np.float64).reshape(n, n).T.copy()
    s = np.ascontiguousarray(s, dtype=np.float64)
    y = np.ascontiguousarray(y, dtype=np.float64)
    lib().orc_bfgs_update(n, _p(Hc), _p(Ic), _p(s), _p(y))
    return Hc.T.copy(), Ic.T.copy()


def sn_steps(f, w, xi, eta, sigma_t, sigma_s, init_cond, init_x, init_y, dx, dy, dt, steps=1):
    """`steps` Heun steps of the Sn solver (kinetic_update.cpp:120-235) on a
    copy of f, shaped [nx+4][ny+4][nq]; sigma_t/sigma_s [nx+4][ny+4]."""
    f = np.array(f, dtype=np.float64, order="C")
    nx, ny, nq = f.shape[0] - 4, f.shape[1] - 4, f.shape[2]
    args = [np.ascontiguousarray(a, dtype=np.float64) for a in (w, xi, eta, sigma_t, sigma_s)]
    for _ in range(steps):
        st = lib().orc_sn_step(nx, ny, nq, *[_p(a) for a in args], int(init_cond), float(init_x), float(init_y),
                               float(dx), float(dy), float(dt), _p(f))
        if st != 0:
            raise RuntimeError("orc_sn_step status %d" % st)
    return f


def isotropic_alpha(nm):
    """opt.cpp:27-35 (Mn)."""
    a = np.zeros(nm)
    a[0] = 2.0 * np.sqrt(np.pi) * np.log(0.5 / np.sqrt(np.pi))
    return a


class Solver:
    """MomOptSolver on one rank (main.cpp + momopt_*.cpp)."""

    def __init__(self, deck, nthreads=0):
        if isinstance(deck, dict):
            deck = deck_from_dict(deck)
        self.deck = deck
        h = C.c_void_p()
        st = lib().orc_solver_create(C.byref(deck), nthreads, C.byref(h))
        if st != OK:
            raise ValueError("orc_solver_create failed: %d" % st)
        self.h = h
        self.nm = lib().orc_solver_nm(h)
        self.nx, self.ny = deck.num_cells_x, deck.num_cells_y

    def __del__(self):
        try:
            lib().orc_solver_destroy(self.h)
        except Exception:
            pass

    @property
    def max_dt(self):
        return lib().orc_solver_max_dt(self.h)

    def update(self, dt):
        st = lib().orc_solver_update(self.h, dt)
        if st != OK:
            raise RuntimeError("orc_solver_update: %d" % st)

    def run(self, t_final, out_delta_t=1000.1):
        n = C.c_int(0)
        st = lib().orc_solver_run(self.h, t_final, out_delta_t, C.byref(n))
        if st != OK:
            raise RuntimeError("orc_solver_run: %d" % st)
        return n.value

    def state(self):
        sz = (self.nx + 4) * (self.ny + 4) * self.nm
        m = np.zeros(sz)
        a = np.zeros(sz)
        lib().orc_solver_get_state(self.h, _p(m), _p(a))
        shp = (self.nx + 4, self.ny + 4, self.nm)
        return m.reshape(shp), a.reshape(shp)

    def set_state(self, moments, alpha):
        m = np.ascontiguousarray(moments, dtype=np.float64).ravel()
        a = np.ascontiguousarray(alpha, dtype=np.float64).ravel()
        lib().orc_solver_set_state(self.h, _p(m), _p(a))

    def moments(self):
        out = np.zeros(self.nx * self.ny * self.nm)
        lib().orc_solver_get_moments(self.h, _p(out))
        return out.reshape(self.nx, self.ny, self.nm)

    def stats(self):
        s = Stats()
        lib().orc_solver_get_stats(self.h, C.byref(s))
        return s

    def solve_all(self):
        st = lib().orc_solver_solve_all(self.h)
        if st != OK:
            raise RuntimeError("orc_solver_solve_all: %d" % st)

    def flux(self):
        out = np.zeros((self.nx + 4) * (self.ny + 4) * self.nm)
        lib().orc_solver_flux(self.h, _p(out))
        return out.reshape(self.nx + 4, self.ny + 4, self.nm)

    def ghost_fill(self):
        lib().orc_solver_ghost_fill(self.h)


DECK_DEFAULTS = dict(num_cells_x=50, num_cells_y=50, a_x=-1.5, b_x=1.5, a_y=-1.5, b_y=1.5,
                     gaussian_sigma=0.4, floor_value=1e-4, sigma=1.0, init_cond=0,
                     moment_order=3, quad_order=30, cfl_factor=0.9, tol=1e-4,
                     cond_h_max=1e10, max_iter=100, use_gaunt=1, theta=2.0, line_search=0,
                     moment_type=MOMENT_MN, delta_ppn=1e-10, tol_gamma=0.0)


def deck_from_dict(d):
    v = dict(DECK_DEFAULTS)
    v.update(d)
    return Deck(**{k: v[k] for k, _ in Deck._fields_})

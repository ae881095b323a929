/*
 * mn_oracle.h -- CPU FP64 restatement of the closures-2d Mn ("momopt") path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.  The
 * product path (closures-2d_amd/) never links or calls it.
 *
 * Every function cites the reference file:line it restates.  Arithmetic is
 * plain IEEE double, evaluated in the reference's order, built with
 * -ffp-contract=off (the reference was built by g++ -O3 for x86-64 without
 * FMA: SConstruct:41, examples/config.py:11).
 */
#ifndef MN_ORACLE_H
#define MN_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes (the reference calls utils_abort() instead: utils.cpp:26-33). */
enum {
    ORC_OK = 0,
    ORC_ERR_ARG = 1,          /* bad argument (odd Q, N<0, ...)                 */
    ORC_ERR_U0 = 2,           /* u[0] <= 0            (opt.cpp:97-101)          */
    ORC_ERR_MAXREG = 3,       /* max regularization   (opt.cpp:257-260)         */
    ORC_ERR_ALLOC = 4
};

/* Mirrors OPTIONS (opt.h:11-22) for the fields the opt path reads. */
typedef struct {
    int max_iter;
    double tol_rel;
    double cond_h_max;
    int use_gaunt;
    int line_search;   /* 0: linesearch.cpp exactly; 1: robust (see orc_opt) */
    int moment_type;   /* ORC_MOMENT_MN or ORC_MOMENT_PPN (opt.h:35-37) */
    double delta_ppn;  /* PPn regularization delta (opt.h:21) */
    int max_bfgs_iter;       /* optbfgs only: inner iterations per basis (opt.h:15) */
    double cond_h_max_bfgs;  /* optbfgs only: bound on cond(H) cond(H^-1) (opt.h:20) */
    double tol_gamma;  /* opt only: the supplementary stop test gamma = modcfl(d) <= tol_gamma
                          (opt.cpp:52-76,288-293, commented out in the reference's current
                          source; momopt_update.cpp:226 sets 1.1); <= 0 switches it off */
    /* opt only, for tests: when trace is set, every iteration of optScaled
     * appends one record of ORC_TRACE_WORDS doubles (total iteration,
     * regularization index, decision, condition estimate, |g|, rtol, gamma)
     * while *trace_len < trace_cap.  Decision: 0 none, 1 the iteration limit,
     * 2 the Cholesky failed, 3 the line search returned t = 0, 4 the
     * condition estimate exceeded cond_h_max, 5 converged. */
    double *trace;
    int trace_cap;
    int *trace_len;
} orc_options;
#define ORC_TRACE_WORDS 7

enum { ORC_MOMENT_MN = 0, ORC_MOMENT_PPN = 1 };

/* Mirrors OUTS (opt.h:25-32). */
typedef struct {
    double r;
    int iter;
    double norm_g;
} orc_outs;

typedef struct orc_tables orc_tables;

/* Gauss-Legendre rule on [-1,1], ascending nodes (utils.cpp:55-60). */
int orc_gauss_legendre(int n, double *x, double *w);

/* Normalised associated Legendre function with Condon-Shortley phase,
 * sqrt((2l+1)/(4pi) (l-m)!/(l+m)!) P_l^m(x)  (the role GSL sphPlm plays at
 * momopt_init.cpp:203-209). */
double orc_sph_plm(int l, int m, double x);

/* Quadrature, real spherical-harmonic table and Gaunt map
 * (momopt_init.cpp:88-93,164-259). */
int orc_tables_create(int moment_order, int quad_order, int use_gaunt, orc_tables **out);
void orc_tables_destroy(orc_tables *t);
int orc_tables_nm(const orc_tables *t);
int orc_tables_np2(const orc_tables *t);
int orc_tables_nq(const orc_tables *t);
/* copies out w[nq], p[nq*np2], xi[nq], eta[nq] (any may be NULL) */
void orc_tables_get(const orc_tables *t, double *w, double *p, double *xi, double *eta);
/* Gaunt map as CSR over the nm*nm (i,j) pairs: ptr[nm*nm+1], k[nnz], v[nnz] */
int orc_tables_gaunt_nnz(const orc_tables *t);
void orc_tables_gaunt_get(const orc_tables *t, int *ptr, int *k, double *v);

/* f, g, H of the Mn dual objective (fobj.cpp:37-129,255-335).
 * h != NULL: full evaluation through the Hessian (g must be non-NULL).
 * h == NULL: f (and g if non-NULL) by direct quadrature. */
double orc_fobj(const orc_tables *t, const double *alpha, const double *u,
                double *g, double *h, int use_gaunt);

/* Unblocked lower Cholesky (role of LAPACK dpotrf 'L'); returns 0 or j+1. */
int orc_cholesky(int n, double *a);
/* Reciprocal 1-norm condition estimate from the Cholesky factor (role of
 * LAPACK dpocon): Higham's estimator (ACM TOMS 14(4), 1988). */
double orc_rcond(int n, const double *l, double anorm);
/* max absolute row sum (utils.cpp:36-52) */
double orc_norm1(int n, const double *a);

/* f, g, H of the PPn dual objective (fobj.cpp:137-236,355-428): the ansatz
 * G(a) = a/2 + sqrt(a^2/4 + delta) of a = alpha.m, f = sum w (t G - delta/2 +
 * delta log G) - alpha.u with t = a/2, g = sum w G m - u, H = sum w G/(2 s) m m^T
 * with s = sqrt(t^2 + delta).  h may be NULL; g may be NULL. */
double orc_fobj_ppn(const orc_tables *t, const double *alpha, const double *u, double delta,
                    double *g, double *h, int use_gaunt);

/* One closure solve: opt() for MOMENT_TYPE_MN and MOMENT_TYPE_PPN
 * (opt.cpp:82-148,154-343).
 * u and alpha are in/out exactly as in the reference.  With
 * opt->line_search == 1 the Armijo loop additionally accepts the full Newton
 * step when its predicted decrease -g.d is at most 1e3 * DBL_EPSILON * fscale,
 * fscale = |H00/m0^2| + sum |alpha_i u_i| (the rounding resolution of f): an
 * extension for tight tolerances, not in the reference. */
int orc_opt(const orc_tables *t, double *u, double *alpha,
            const orc_options *opt, orc_outs *outs);

/* Batched opt over ncells cells, [ncells][nm] layout, OpenMP schedule(dynamic,1)
 * as MomOptSolver::solveOptimization (momopt_update.cpp:208-277).
 * nthreads <= 0: OpenMP default.  Returns the first error status seen. */
int orc_opt_batch(const orc_tables *t, int ncells, double *u, double *alpha,
                  const orc_options *opt, orc_outs *outs, int nthreads);

/* One closure solve by the change-of-basis optimiser optcb (optcb.cpp:90-415,
 * MOMOPT_OPTIMIZATION_TYPE 1).  Per cell it carries the basis P2M (n x n,
 * lower triangular, column-major) and alphaP, the multipliers in that basis;
 * both are in/out state for the next call.  Each Newton step factors the
 * Hessian in the current basis, H_P = L L^T, and moves the basis by L
 * (P2M <- P2M L, alphaP <- L^T alphaP, g_P, u_P and the table <- L^{-1} .), so
 * the direction is -g_P; the stop test measures g in the moment basis.  On
 * success alphaM = P2M^{-T} alphaP (rescaled as opt does) and u is the
 * (regularized) input, not a reconstruction.  Only the direct Hessian: the
 * reference copies the table with stride n (optcb.cpp:183), so use_gaunt is
 * ORC_ERR_ARG here. */
int orc_optcb(const orc_tables *t, double *u, double *P2M, double *alphaP, double *alphaM,
              const orc_options *opt, orc_outs *outs);
/* Batched optcb, [ncells][n], [ncells][n*n] layouts, OpenMP dynamic,1. */
int orc_optcb_batch(const orc_tables *t, int ncells, double *u, double *P2M, double *alphaP, double *alphaM,
                    const orc_options *opt, orc_outs *outs, int nthreads);

/* The BFGS variant optbfgs (optbfgs.cpp:92-513, MOMOPT_OPTIMIZATION_TYPE 2):
 * optcb's outer loop (full Hessian, factor, basis move by L), then, from
 * H = H^-1 = I in the new basis, quasi-Newton steps d = -H^-1 g with the
 * gradient alone and rank-two updates of H and H^-1, while
 * ||H||_inf ||H^-1||_inf < cond_h_max_bfgs and at most max_bfgs_iter of them
 * (optbfgs.cpp:376-478).  outs.iter counts the outer iterations, as the
 * reference's totalIter.  cond_h_max_bfgs <= 1 or max_bfgs_iter < 1 would
 * take no step (the reference then loops to its max-regularization abort):
 * ORC_ERR_ARG here. */
int orc_optbfgs(const orc_tables *t, double *u, double *P2M, double *alphaP, double *alphaM,
                const orc_options *opt, orc_outs *outs);
int orc_optbfgs_batch(const orc_tables *t, int ncells, double *u, double *P2M, double *alphaP, double *alphaM,
                      const orc_options *opt, orc_outs *outs, int nthreads);
/* One BFGS update (optbfgs.cpp:440-477) of H and H^-1 (n x n, column-major)
 * for the step s = alpha - alpha_old and y = g - g_old. */
void orc_bfgs_update(int n, double *H, double *Hinv, const double *s, const double *y);

/* The Sn discrete-ordinates solver's Heun step (KineticSolver::update,
 * kinetic/kinetic_update.cpp:120-235) on one rank: f is [(nx+4)(ny+4)][nq]
 * with two ghost layers (I3D order), the ordinates (w, xi, eta) those of the
 * moment tables for the same Q (kinetic_init.cpp:76-110 builds the same
 * quadrature), sigma_t/sigma_s [(nx+4)(ny+4)].  Periodic ghosts for
 * init_cond 2, the centre source for init_cond 1, as the reference. */
int orc_sn_step(int nx, int ny, int nq, const double *w, const double *xi, const double *eta,
                const double *sigma_t, const double *sigma_s, int init_cond, double init_x, double init_y,
                double dx, double dy, double dt, double *f);

/* ---------------- grid solver (main.cpp, momopt_*.cpp) ---------------- */

/* Deck values the momopt path reads (main.cpp:69-83, momopt_init.cpp:56-68). */
typedef struct {
    int num_cells_x, num_cells_y;
    double a_x, b_x, a_y, b_y;
    double gaussian_sigma, floor_value, sigma;
    int init_cond;
    int moment_order, quad_order;
    double cfl_factor, tol, cond_h_max;
    int max_iter, use_gaunt;
    double theta;
    int line_search;
    int moment_type;    /* MOMOPT_MOMENT_TYPE (momopt_init.cpp:67) */
    double delta_ppn;   /* MOMOPT_DELTA_PPN (momopt_init.cpp:66) */
    double tol_gamma;   /* orc_options.tol_gamma (momopt_update.cpp:226); <= 0: off */
} orc_deck;

/* Mirrors OPTIMIZATION_STATS (momopt_solver.h:33-41). */
typedef struct {
    int max_iter;
    long long num_dual_solves;
    long long iter_sum;
    long long hist_reg[5];
} orc_stats;

typedef struct orc_solver orc_solver;

int orc_solver_create(const orc_deck *deck, int nthreads, orc_solver **out);
void orc_solver_destroy(orc_solver *s);
double orc_solver_max_dt(const orc_solver *s);
int orc_solver_nm(const orc_solver *s);
/* One Heun step (MomOptSolver::update, momopt_update.cpp:283-382). */
int orc_solver_update(orc_solver *s, double dt);
/* The time loop of main.cpp:204-230 up to t_final; returns steps taken in *steps. */
int orc_solver_run(orc_solver *s, double t_final, double out_delta_t, int *steps);
/* Ghost-inclusive state [(nx+4)*(ny+4)*nm] (I3D layout, solver.h:13). */
void orc_solver_get_state(const orc_solver *s, double *moments, double *alpha);
void orc_solver_set_state(orc_solver *s, const double *moments, const double *alpha);
/* Interior moments in .pn order [nx][ny][nm] (momopt_output.cpp:41-47). */
void orc_solver_get_moments(const orc_solver *s, double *out);
void orc_solver_get_stats(const orc_solver *s, orc_stats *st);
/* Solve stages separately (for stage-level parity tests). */
int orc_solver_solve_all(orc_solver *s);                /* solveOptimization */
void orc_solver_flux(orc_solver *s, double *flux);       /* solveFlux -> [cells][nm] */
void orc_solver_ghost_fill(orc_solver *s);               /* communicateBoundaries (serial) */

#ifdef __cplusplus
}
#endif
#endif

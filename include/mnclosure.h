/*
 * mnclosure.h -- C-ABI of the MI355X Mn (and PPn) closure hot path.
 *
 * The reference (ckrisgarrett/closures-2d) runs the Mn closure as one host
 * call per cell, opt() (src/momopt/opt/opt.h:43-44), from an OpenMP loop in
 * MomOptSolver::solveOptimization (src/momopt/momopt_update.cpp:208-277), and
 * consumes the ansatz in MomOptSolver::solveFlux / update
 * (src/momopt/momopt_update.cpp:96-202,283-382).  This header is the thin
 * boundary those calls cross into hand-written gfx950 HIP kernels.
 *
 * Moment types (mnc_setup.moment_type, MOMOPT_MOMENT_TYPE): 0 = Mn, the
 * entropy closure with the exp ansatz; 1 = PPn, the positive Pn closure with
 * the ansatz a/2 + sqrt(a^2/4 + delta) of a = alpha.m (fobj.cpp:137-236,
 * 355-428; opt.cpp:113-141; momopt_update.cpp:54-57).  PPn runs the FP64 form
 * for moment orders 1-4 (options.delta_ppn > 0; the grid description carries
 * the same delta for the flux); mixed precision, the learned warm start and
 * M8 are Mn only and return MNC_ERR_ARG for PPn.
 *
 * Conventions
 *   - Plain C types only; every array is row-major double.
 *   - A moment/multiplier array is [ncells][n] with n = (N+1)(N+2)/2, i.e.
 *     the reference's I3D layout (src/solver.h:13) when it covers a grid.
 *   - Functions return an MNC_* status; nothing in the library aborts the
 *     process.  The reference's fatal paths (utils_abort at opt.cpp:97-101 and
 *     opt.cpp:257-260) come back as MNC_ERR_U0 / MNC_ERR_MAXREG and the caller
 *     decides (the reference driver aborts).
 *   - One context = one device, one HIP stream (default priority), one grid.
 *     Calls on one context are stream-ordered and must come from one host
 *     thread at a time.
 */
#ifndef MNCLOSURE_H
#define MNCLOSURE_H

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---- */
#define MNC_OK 0
#define MNC_ERR_ARG 1          /* bad argument or unsupported size            */
#define MNC_ERR_U0 2           /* some cell had u[0] <= 0   (opt.cpp:97-101)  */
#define MNC_ERR_MAXREG 3       /* some cell passed the last regularization
                                  level                       (opt.cpp:257-260) */
#define MNC_ERR_HIP 4          /* a HIP runtime call failed                    */
#define MNC_ERR_STATE 5        /* call out of order (e.g. no grid yet)         */

/* ---- option values ---- */
#define MNC_PREC_FP64 0        /* plain FP64 Newton (reproduces the reference) */
#define MNC_PREC_MIXED 1       /* BF16/FP32-MFMA Hessian, FP64 g/f/Cholesky/stop */

#define MNC_WARM_PREVIOUS 0    /* start from the alpha passed in (the reference) */
#define MNC_WARM_ISOTROPIC 1   /* start from isotropicAlpha (opt.cpp:27-35)      */
#define MNC_WARM_LEARNED 2     /* MLP prediction, judged against the candidates
                                  above by the FP64 gradient norm; a cell started
                                  from the prediction that would regularize at
                                  level 0 restarts once from isotropicAlpha at
                                  level 0 (mnc_stats.warm_fallback)             */

#define MNC_LS_REFERENCE 0     /* Armijo on f exactly as linesearch.cpp:21-99     */
#define MNC_LS_ROBUST 1        /* the same, but a full Newton step whose predicted
                                  decrease is below the rounding resolution of f
                                  is accepted (needed for MOMOPT_TOL ~1e-8, where
                                  the reference aborts with "Max regularization") */

#define MNC_HESS_DIRECT 0      /* H = sum_q w e m m^T          (fobj.cpp:85-93)  */
#define MNC_HESS_GAUNT 1       /* H from order-2N moments + Gaunt map (fobj.cpp:76-82,99-114) */

typedef struct mnc_ctx mnc_ctx;

/* Setup keys of MomOptSolver::init (src/momopt/momopt_init.cpp:56-68). */
typedef struct {
    int moment_order;          /* MOMOPT_MOMENT_ORDER  N (n = (N+1)(N+2)/2)    */
    int quad_order;            /* MOMOPT_QUAD_ORDER    Q (even; nq = Q*Q)      */
    int use_gaunt;             /* MOMOPT_USE_GAUNT     0/1                     */
    int moment_type;           /* MOMOPT_MOMENT_TYPE: 0 = Mn, 1 = PPn           */
    int device;                /* HIP device ordinal; < 0 = host-only tables    */
} mnc_setup;

/* OPTIONS (src/momopt/opt/opt.h:11-22), plus the GPU-only choices. */
typedef struct {
    int max_iter;              /* MOMOPT_MAX_ITER                               */
    double tol_rel;            /* MOMOPT_TOL                                    */
    double cond_h_max;         /* MOMOPT_COND_H_MAX                             */
    int precision;             /* MNC_PREC_*                                    */
    int warm_start;            /* MNC_WARM_*                                    */
    int line_search;           /* MNC_LS_*                                      */
    double delta_ppn;          /* MOMOPT_DELTA_PPN (contexts of moment type 1)  */
    int opt_type;              /* MOMOPT_OPTIMIZATION_TYPE: MNC_OPT_* (0 when the
                                  struct is zero-initialised past delta_ppn)    */
    int max_bfgs_iter;         /* MOMOPT_MAX_BFGS_ITER (MNC_OPT_BFGS only, >= 1) */
    double cond_h_max_bfgs;    /* MOMOPT_COND_H_MAX_BFGS (MNC_OPT_BFGS only, > 1) */
    double tol_gamma;          /* optScaled's supplementary stop test (opt.cpp:52-76,
                                  288-293; tolGamma, momopt_update.cpp:226): a cell whose
                                  gradient passes tol_rel stops only when
                                  gamma = exp(2 |d|_1 m_inf) (Mn; 1 + 2 |d|_1 m_inf for
                                  PPn) <= tol_gamma, d the Newton direction and
                                  m_inf = sqrt((2 (floor(sqrt(n)) + 1) + 1) / (2 pi)).
                                  <= 0 (zero-initialised): off, as the reference's current
                                  source; its regression fixtures were written with 1.1.
                                  MNC_OPT_NEWTON only (MNC_ERR_ARG otherwise)         */
} mnc_options;

/* Optimisers (opt.h:42-46).  MNC_OPT_CHANGE_BASIS in the grid solves
 * (mnc_solve_grid, mnc_update) runs optcb (optcb.cpp:90-415) with the
 * per-cell basis and multipliers kept on the device, initialised as
 * momopt_init.cpp:144-158 does at mnc_grid_create and mnc_grid_set: Mn,
 * direct Hessian, FP64, warm start MNC_WARM_PREVIOUS.  The batch entry points
 * take MNC_OPT_NEWTON only (the change-of-basis batch is
 * mnc_optcb_batch_host). */
#define MNC_OPT_NEWTON 0
#define MNC_OPT_CHANGE_BASIS 1
/* The BFGS variant (optbfgs.cpp:92-545): optcb's outer loop, then from
 * H = H^-1 = I in the new basis up to max_bfgs_iter quasi-Newton steps with
 * the gradient alone while ||H||_inf ||H^-1||_inf < cond_h_max_bfgs.  Same
 * state, forms and entry points as MNC_OPT_CHANGE_BASIS (mnc_optcb_batch_host
 * runs it when opt_type is MNC_OPT_BFGS); outs.iter counts the outer
 * (full-Hessian) iterations. */
#define MNC_OPT_BFGS 2

/* OUTS (src/momopt/opt/opt.h:25-32). */
typedef struct {
    double r;                  /* regularization level used                    */
    int iter;                  /* total Newton iterations incl. restarts       */
    double norm_g;             /* ||g||_2 at exit (scaled problem)             */
} mnc_outs;

/* Grid description: the fields of Solver/MomOptSolver the update reads
 * (src/solver.h:411-424, src/main.cpp:140-155). */
typedef struct {
    int num_cells_x;           /* c_sizeX (interior)                            */
    int num_cells_y;           /* c_sizeY                                       */
    double dx, dy;
    double init_x, init_y;     /* c_initX, c_initY (lower-left domain corner)   */
    int init_cond;             /* 1 = lattice source term, 2 = periodic ghosts  */
    double theta;              /* MOMOPT_THETA (minmod slope factor)            */
    double delta_ppn;          /* MOMOPT_DELTA_PPN: the PPn ansatz of the flux  */
} mnc_grid_desc;

/* Per-solve statistics (OPTIMIZATION_STATS, src/momopt/momopt_solver.h:33-41). */
typedef struct {
    int max_iter;
    long long num_dual_solves;
    long long iter_sum;          /* Newton iterations (full f, g, H evaluations)  */
    long long hist_reg[5];
    long long trial_sum;         /* Armijo trial evaluations (f only)             */
    long long warm_pick[3];      /* MNC_WARM_LEARNED starts taken from: [0] the
                                    MLP, [1] isotropicAlpha, [2] the alpha passed in */
    long long mixed_deferred;    /* MNC_PREC_MIXED Newton steps redone in FP64       */
    long long tail_cells;        /* cells the persistent tail took (those still
                                    iterating after the lockstep rounds)            */
    long long tail_iter_sum;     /* Newton iterations done in the tail              */
    long long tail_trial_sum;    /* Armijo trials done in the tail                  */
    long long warm_fallback;     /* MNC_WARM_LEARNED cells started from the MLP's
                                    guess that met a regularization criterion at
                                    level 0 and restarted from isotropicAlpha at
                                    level 0 (once per cell)                          */
} mnc_stats;

/* Creates a context: builds the quadrature, the real spherical-harmonic table
 * and (use_gaunt) the Gaunt map on the host and uploads them once.  With
 * setup->device < 0 the context is host-only (tables, no HIP calls; every
 * device entry point then returns MNC_ERR_STATE).
 * Replaces MomOptSolver::init's table setup (momopt_init.cpp:88-259). */
int mnc_create(const mnc_setup *setup, mnc_ctx **out);
void mnc_destroy(mnc_ctx *ctx);

int mnc_num_moments(const mnc_ctx *ctx);       /* n  (c_numMoments)            */
int mnc_num_quad(const mnc_ctx *ctx);          /* nq (c_numQuadPoints)         */
int mnc_table_stride(const mnc_ctx *ctx);      /* np2 (c_numManyMoments)       */
/* Host copies of w[nq], p[nq*np2], xi[nq], eta[nq] (momopt_init.cpp:176-217);
 * any pointer may be NULL. */
int mnc_get_tables(const mnc_ctx *ctx, double *w, double *p, double *xi, double *eta);
/* Gaunt map as CSR over the n*n (i,j) pairs (g_hStruct, fobj.h:12-22). */
int mnc_gaunt_nnz(const mnc_ctx *ctx);
int mnc_get_gaunt(const mnc_ctx *ctx, int *ptr, int *k, double *v);

/* Batched closure on DEVICE arrays u[ncells][n], alpha[ncells][n] (in/out,
 * exactly opt()'s in/out meaning: u becomes the reconstructed ansatz moments,
 * alpha the multipliers); outs is a device array of ncells or NULL.
 * Replaces the per-cell opt() loop of solveOptimization
 * (momopt_update.cpp:214-238 -> opt.cpp:82-148). */
int mnc_opt_batch(mnc_ctx *ctx, int ncells, double *u, double *alpha,
                  const mnc_options *opt, mnc_outs *outs);
/* The same on HOST arrays (copies in and out): the drop-in for the
 * reference's c_moments / c_alpha host buffers. */
int mnc_opt_batch_host(mnc_ctx *ctx, int ncells, double *u, double *alpha,
                       const mnc_options *opt, mnc_outs *outs);
/* No counterpart in the reference: mnc_opt_batch for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  Same device arrays, options, results and
 * errors as mnc_opt_batch, but the per-cell Newton state, the active lists and
 * the control words live in one fixed device array of the library (about
 * 27 MB, one copy per GPU, held until the process ends) instead of the
 * context's buffers, and the call makes no hipMalloc or hipFree.  Limits:
 * ncells <= MNC_NOALLOC_MAX_CELLS, moment order <= 4, warm start previous or
 * isotropic (the learned start allocates the network's buffers), no per-launch
 * timing; otherwise MNC_ERR_ARG.  Every call on one GPU, from any context,
 * shares that array and the call returns after its kernels finish: calls that
 * overlap on one GPU (from several host threads) race, so callers serialise
 * them.  Statistics accumulate in the context as for mnc_opt_batch. */
#define MNC_NOALLOC_MAX_CELLS 16384
int mnc_opt_batch_noalloc(mnc_ctx *ctx, int ncells, double *u, double *alpha,
                          const mnc_options *opt, mnc_outs *outs);
/* The change-of-basis optimiser on a batch, HOST arrays (copies in and out):
 * replaces the per-cell optcb() loop that solveOptimization runs for
 * MOMOPT_OPTIMIZATION_TYPE 1 (momopt_update.cpp:240-243 -> optcb.cpp:90-415).
 * Per cell: u[nm] (in; out: the input, blended by the regularization used),
 * P2M[nm*nm] (the basis, lower triangular, column-major; in/out state),
 * alphaP[nm] (the multipliers in that basis; in/out state) and alphaM[nm]
 * (out: the multipliers in the moment basis, as c_alpha).  The reference's
 * grid initialisation (ones in the first column of P2M, isotropic alphaP;
 * momopt_init.cpp:144-158) is a valid first input: its singular basis takes
 * the first-evaluation reset to the identity (optcb.cpp:264-281).  Mn
 * contexts with the direct Hessian only (the reference copies the table with
 * stride nm, optcb.cpp:183), FP64, warm_start MNC_WARM_PREVIOUS; otherwise
 * MNC_ERR_ARG.  Errors as mnc_opt_batch. */
int mnc_optcb_batch_host(mnc_ctx *ctx, int ncells, double *u, double *P2M, double *alphaP, double *alphaM,
                         const mnc_options *opt, mnc_outs *outs);

/* The Sn discrete-ordinates solver (SOLVER kinetic), the independent truth
 * for the closures: `steps` Heun steps of KineticSolver::update
 * (src/kinetic/kinetic_update.cpp:120-235) on the GPU, the main loop's call
 * repeated (src/main.cpp:204-230 with a fixed dt).  kinetic is the host state
 * [(nx+4)*(ny+4)][Q*Q] with two ghost layers (I3D order), in and out;
 * sigma_t/sigma_s [(nx+4)*(ny+4)].  The ordinates are the moment tables'
 * (kinetic_init.cpp:76-110 builds the same quadrature).  Periodic ghosts
 * for init_cond 2 (kinetic_boundaries.cpp:190-223), the centre source for
 * init_cond 1.  kernel_ms (may be null) receives the device time of the
 * steps.  MNC_ERR_ARG for bad sizes, MNC_ERR_HIP on a runtime error. */
int mnc_sn_run(int nx, int ny, int quad_order, double dx, double dy, double init_x, double init_y, int init_cond,
               const double *sigma_t, const double *sigma_s, double dt, int steps, double *kinetic, int device,
               double *kernel_ms);

/* A batch kept in device memory across calls (the staging arrays of
 * mnc_opt_batch_host): upload it, solve it in place any number of times (each
 * solve is mnc_opt_batch on those arrays, outs kept on the device), read it
 * back.  No counterpart in the reference; it serves callers that reuse one
 * batch and timing of the closure without the PCIe copies. */
int mnc_batch_set(mnc_ctx *ctx, int ncells, const double *u, const double *alpha);
int mnc_batch_solve(mnc_ctx *ctx, const mnc_options *opt);
int mnc_batch_get(mnc_ctx *ctx, double *u, double *alpha, mnc_outs *outs);

/* Grid state (ghost-inclusive, [(nx+4)*(ny+4)][n] moments and alpha;
 * [(nx+4)*(ny+4)] sigma_t, sigma_s), uploaded from host arrays prepared as
 * Solver::initializeGrid + MomOptSolver::init do (main.cpp:257-372,
 * momopt_init.cpp:120-143). */
int mnc_grid_create(mnc_ctx *ctx, const mnc_grid_desc *grid, const double *moments,
                    const double *alpha, const double *sigma_t, const double *sigma_s);
int mnc_grid_get(mnc_ctx *ctx, double *moments, double *alpha);
int mnc_grid_set(mnc_ctx *ctx, const double *moments, const double *alpha);

/* The change-of-basis optimisers' per-cell state of the grid, the
 * reference's c_P2M ([cells][n*n], each basis lower triangular and
 * column-major) and c_alphaP ([cells][n]) (momopt_init.cpp:115-158), ghost
 * cells included, read or written for restart files: with
 * MOMOPT_OPTIMIZATION_TYPE 1 or 2 the next solve starts from this state, so a
 * restart that continues bit for bit needs it beside the moments and alpha.
 * Before the grid's first change-of-basis solve it reads the initial state
 * (ones in each basis's first column, isotropic alphaP).  mnc_grid_create and
 * mnc_grid_set reset it to that initial state. */
int mnc_grid_get_basis(mnc_ctx *ctx, double *P2M, double *alphaP);
int mnc_grid_set_basis(mnc_ctx *ctx, const double *P2M, const double *alphaP);

/* Serial communicateBoundaries: periodic ghost copy of moments only
 * (comm.cpp:148-155 -> momopt_boundaries.cpp:190-223); no-op unless
 * init_cond == 2. */
int mnc_ghost_fill(mnc_ctx *ctx);
/* solveOptimization over every cell including ghosts (momopt_update.cpp:208-277). */
int mnc_solve_grid(mnc_ctx *ctx, const mnc_options *opt);
/* solveFlux (momopt_update.cpp:96-202) into a host array [(nx+4)*(ny+4)][n]
 * (interior rows filled, ghost rows zero); flux may be NULL to keep it on device. */
int mnc_flux(mnc_ctx *ctx, double *flux_host);
/* MomOptSolver::update (momopt_update.cpp:283-382): one Heun step of dt. */
int mnc_update(mnc_ctx *ctx, double dt, const mnc_options *opt);
/* Accumulated solve statistics since the last reset. */
int mnc_get_stats(mnc_ctx *ctx, mnc_stats *st);
int mnc_reset_stats(mnc_ctx *ctx);

/* Tuning knobs of the batched solve (set by the autotuner or by hand):
 *   "rounds"       lockstep sweep/logic rounds before the persistent tail (0-512)
 *   "tail_blocks"  workgroups of the persistent tail kernel (1-65536)
 *   "tail_mode"    0: one wave per remaining cell, dense algebra shared in LDS;
 *                  1: one lane per cell; 2: one wave per cell, algebra per lane;
 *                  3: as 0 at two waves per SIMD (more cells in flight, each
 *                  a little slower: for tails of many cells; direct Mn
 *                  Hessian, elsewhere 0)
 *   "mix_cond_log10"  MNC_PREC_MIXED: a Newton step uses the mixed-precision
 *                  Hessian only while its condition estimate is <= 10^value
 *                  (0-10, default 9); otherwise the step is redone in FP64
 *   "mix_defer"    MNC_PREC_MIXED: 0: a step to redo in FP64 waits for the
 *                  FP64 tail (default); 1: it joins FP64 lockstep rounds
 *                  that run beside the mixed-precision ones
 *   "mix_split"    MNC_PREC_MIXED: bf16 parts of the quadrature values and of
 *                  the table in the Hessian products: 2 (three products,
 *                  FP32 sums, ~1e-5 relative) or 3 (default; six products,
 *                  FP64 sums per 32 points, ~1e-7 relative)
 *   "mix_blocks"   workgroups of the mixed-precision sweep (1-65536)
 *   "learned_gate_log10"  MNC_WARM_LEARNED: the network is consulted for a
 *                  cell only when the gradient norm at its start exceeds
 *                  10^value ||u|| (scaled problem; -16..4, default -2);
 *                  5 switches the network off: every cell starts from the
 *                  alpha passed in (isotropicAlpha if it is not finite) and
 *                  neither the network nor the kernel that compares its guess
 *                  with the start is launched
 *   "learned_fused"  MNC_WARM_LEARNED: 1 (default): the network and the
 *                  comparison of its guess with the parked start run in one
 *                  kernel (weights and quadrature in LDS, activations never
 *                  leave the chip, the guess's gradient norm from FP32
 *                  matrix-core sums; quadratures up to 512 points, larger ones
 *                  take the 0 path); 0: the separate GEMM kernels of
 *                  mnc_mlp_predict, then the FP64 comparison kernel
 *   "learned_fallback_iters"  MNC_WARM_LEARNED: a cell started from the
 *                  network's guess that meets a regularization criterion at
 *                  level 0 within this many Newton iterations restarts once
 *                  from isotropicAlpha at level 0 (0..100000, default 10;
 *                  0 never)
 *   "flux_tile"    0: ansatz written to HBM, then the flux; 1: ansatz fused
 *                  into the flux with 8x8-cell LDS tiles; 2: 16x16 tiles
 *                  (n <= 15); these three give identical results.  3 (the
 *                  default at n <= 15; above it 0 is the default and 3 runs
 *                  as 1): 16x16 tiles with the directions split
 *                  into "flux_groups" groups summed in order and reciprocals
 *                  of dx, dy (rounding-level differences from 0-2)
 *   "flux_groups"  direction groups of flux_tile 3 (1-8, default 6)
 *   "flux_overlap" flux_tile 3 in mnc_update: 1 (default): after a stage's
 *                  lockstep rounds, the flux of the tiles whose stencil region
 *                  holds no cell left for the Newton tail runs on a second
 *                  stream (default priority) beside the tail, the other tiles'
 *                  after it; 0: the tail, then every tile.  The same kernels on
 *                  the same cells: identical results.  Per-launch timing runs 0.
 *   "graph"        1: mnc_update steps are captured into hipGraphs (one per
 *                  distinct dt and options, captured on the second step with
 *                  that key) and replayed; 0: eager launches (default).  While
 *                  per-launch timing (mnc_timing_enable) is on, steps run eagerly.
 * mnc_get_tune also reads "has_mixed" (1 if this context's order and
 * quadrature have the mixed-precision kernels: N <= 4, Q <= 16).
 * Returns MNC_ERR_ARG for an unknown key or an out-of-range value. */
int mnc_tune(mnc_ctx *ctx, const char *key, int value);
int mnc_get_tune(const mnc_ctx *ctx, const char *key, int *value);

/* Autotuner (no counterpart in the reference): times `steps` mnc_update steps
 * of dt from the current grid state for each candidate setting, one
 * coordinate-descent pass over, in this order, the knobs that change the work
 * -- (learned warm start) "learned_gate_log10" (5, the network off, among
 * the candidates), (mixed precision)
 * "mix_cond_log10", "mix_split", "mix_defer" -- then the launch shapes
 * "graph", "flux_tile", "flux_groups", "flux_overlap", "rounds", "tail_mode", "tail_blocks" and (mixed
 * precision) "mix_blocks", and keeps the fastest (a gate that consults the
 * network only when it beats the network off by more than 1%).  Every candidate converges
 * to the same tolerance; "mix_cond_log10" trades mixed-precision steps
 * against FP64 ones.  The grid state
 * and the statistics are restored after every trial, so the call leaves the
 * solution where it was.  report (may be NULL) receives one "key=value ms/step"
 * line per trial.  The tail kernels sum in different orders, so the choice can
 * move results at rounding level. */
int mnc_autotune(mnc_ctx *ctx, double dt, const mnc_options *opt, int steps, char *report, int report_len);

/* Synchronises the context's stream; returns the first error seen. */
int mnc_sync(mnc_ctx *ctx);
/* Kernel timing with HIP events on the context's stream.
 * mnc_timing_enable(ctx, 1) starts recording every launch; mnc_timing_get
 * synchronises and returns, for kernel class k (MNC_KCLASS_*), the number of
 * launches and their summed duration in milliseconds. */
#define MNC_KCLASS_NEWTON 0
#define MNC_KCLASS_FLUX 1
#define MNC_KCLASS_OTHER 2
#define MNC_KCLASS_MLP 3
#define MNC_KCLASS_TAIL 4       /* the persistent tail of the Newton solve (timed apart
                                   from MNC_KCLASS_NEWTON, its prologue and rounds) */
#define MNC_NUM_KCLASS 5
int mnc_timing_enable(mnc_ctx *ctx, int on);
int mnc_timing_get(mnc_ctx *ctx, int kclass, long long *launches, double *ms);
/* Last error text for the most recent failing call on this context. */
const char *mnc_last_error(const mnc_ctx *ctx);

/* ---- learned warm start (no counterpart in the reference, which starts every
 * solve from the previous step's alpha: momopt_update.cpp:236-238) ----
 *
 * An MLP 15 -> 128 -> 128 -> 128 -> 15 (GELU) maps the scaled moments u/u0 of
 * an M4 closure to the multipliers of the scaled (u0 = 1) problem.  It is
 * trained and evaluated on the context's device (bf16 matrix-core products,
 * FP32 sums and Adam).  With MNC_WARM_LEARNED the Newton prologue evaluates
 * the network for every cell and starts from whichever of {MLP guess,
 * isotropicAlpha, the alpha passed in} has the smallest FP64 gradient norm.
 * The parameter block is mnc_mlp_num_params() floats: the 37,007 weights and
 * biases (W1[15][128] b1 W2[128][128] b2 W3[128][128] b3 W4[128][15] b4)
 * followed by the input/target standardisation (in_mean, in_std, out_mean,
 * out_std; 15 each).  Only moment order 4 (n = 15) has a network. */
int mnc_mlp_num_params(void);
/* Random initial weights (uniform, fan-in scaled; SplitMix64 stream `seed`). */
int mnc_mlp_init(mnc_ctx *ctx, unsigned long long seed);
/* Adam steps on minibatches drawn with replacement from npairs host pairs
 * (u_scaled[npairs][15], alpha_scaled[npairs][15]).  The first call after
 * mnc_mlp_init fixes the standardisation from these pairs.  loss_out[steps]
 * (may be NULL) receives each step's mean squared error in standardised units. */
int mnc_mlp_train(mnc_ctx *ctx, long long npairs, const double *u_scaled, const double *alpha_scaled, int steps,
                  int batch, double lr, unsigned long long seed, float *loss_out);
/* Network output for n host rows of scaled moments -> scaled multipliers. */
int mnc_mlp_predict(mnc_ctx *ctx, int n, const double *u_scaled, double *alpha_scaled);
int mnc_mlp_get_params(mnc_ctx *ctx, float *params);
int mnc_mlp_set_params(mnc_ctx *ctx, const float *params);
/* One training step's gradient (no update) on the given rows, for checking:
 * grad[37007], loss = mean squared error in standardised units. */
int mnc_mlp_gradient(mnc_ctx *ctx, int rows, const double *u_scaled, const double *alpha_scaled, float *grad,
                     float *loss);

#ifdef __cplusplus
}
#endif
#endif

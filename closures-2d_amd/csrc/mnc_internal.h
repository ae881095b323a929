// mnc_internal.h -- shared declarations of the MI355X Mn-closure library.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/mnclosure.h"

namespace mnc {

// Quadrature + real spherical harmonics + Gaunt map, built on the host
// (momopt_init.cpp:88-259).  p is [nq][np2] like c_spHarm.
struct Tables {
    int N = 0, Q = 0, nm = 0, nmm = 0, np2 = 0, nq = 0, use_gaunt = 0;
    std::vector<double> w, p, xi, eta;
    std::vector<int> gptr, gk;      // CSR over nm*nm pairs
    std::vector<double> gv;
};

// Builds the tables; returns MNC_OK or MNC_ERR_ARG.
int build_tables(int moment_order, int quad_order, int use_gaunt, Tables &t);

// Device-side kernel classes for timing.
enum KClass { KC_NEWTON = MNC_KCLASS_NEWTON, KC_FLUX = MNC_KCLASS_FLUX, KC_OTHER = MNC_KCLASS_OTHER,
              KC_MLP = MNC_KCLASS_MLP };

// Arguments of the batched Newton kernels.
struct NewtonArgs {
    int ncells;
    double *u;            // [ncells][nm] in/out
    double *alpha;        // [ncells][nm] in/out
    mnc_outs *outs;       // [ncells] or nullptr
    const double *w;      // [nq]
    const double *p;      // [nq][np2]
    const double *pT;     // [np2][nq], the same table transposed
    int nq;
    const int *gptr;      // Gaunt CSR (nm*nm+1)
    const int *gk;
    const double *gv;
    int max_iter;
    double tol;
    double cond_h_max;
    double tol_gamma;     // the supplementary stop test's bound (opt.cpp:288-293); <= 0: off
    double gamma_minf;    // its m_inf = sqrt((2 (floor(sqrt(n)) + 1) + 1) / (2 pi)) (opt.cpp:61-62)
    int warm;             // MNC_WARM_*
    long long *stats;     // [0] max iter, [1] iter sum, [2] solves, [3..7] hist, [8] trials,
                          // [9..11] warm starts taken: learned, isotropic, previous,
                          // [12] mixed-precision steps redone in FP64, [13] cells the
                          // persistent tail took, [14], [15] Newton iterations and
                          // Armijo trials done in the tail
    int ls_mode;          // MNC_LS_*
    const double *alpha_guess; // [ncells][nm] scaled-problem guess (learned warm start) or nullptr
    double gate;          // learned warm start: the network is consulted when the best of isotropic
                          // and previous alpha has ||g|| > gate * ||u|| (scaled problem)
    int divert;           // learned warm start: this launch's logic holds the gate (first round)
    const float *mlp_w;   // learned warm start: the network's parameters and standardisation (FP32)
    int fallback_iters;   // learned warm start: a start from the network's guess that meets a
                          // regularization criterion at level 0 within this many iterations
                          // restarts from isotropicAlpha at level 0
    int ppn;              // MOMENT_TYPE_PPN: the PPn closure (opt.cpp:113-141, fobj.cpp:137-236,355-428)
    double delta;         // its delta (MOMOPT_DELTA_PPN), unscaled
    // per-cell solver state and the active lists
    double *st_d;         // [dfields][cap]
    int *st_i;            // [ifields][cap]
    int *lists;           // [4][cap]: the lockstep pair 0/1, the FP64 pair 2/3 of a mixed solve
    int *ctl;             // [0],[1] counts of lists 0/1, [2] tail queue, [3] error bits, [4],[5] lists 2/3
    size_t cap;
    // mixed precision (MixTables, device copies)
    const double *mz;     // [nt][4][64]  alpha.m MFMA operand
    const double *mg;     // [nt][4][64]  w m MFMA operand (column 15: w)
    const unsigned short *mh; // [nt/2][nnt][2][64][8] bf16 hi/lo of w m2N
    int mnt;              // quadrature tiles of 16 points (even)
    const int *mgptr, *mgk;   // Gaunt map of the order-2N table
    const double *mgv;
    const float *mpv;         // the same map in compile-time pattern order, or nullptr
    double cmix;          // condition bound above which a mixed-precision step is redone in FP64
    // FP64 matrix-core sweep of the high orders (G64Tables, device copies)
    const double *gz, *gg, *gh;
    int gnt;              // quadrature tiles of 16 points
    int gnmm;             // order-2N moments
};

// Control words of a batched solve: [0], [1] counts of lists 0/1 (the
// lockstep pair), [2] the tail's queue head (the change-of-basis kernel's
// cell queue when that runs instead), [3] error bits, [4], [5] counts of
// lists 2/3 (cells redone in FP64 by a mixed-precision solve), [6] count of
// list 4 (cells waiting for the learned warm start's network).
constexpr int kCtlWords = 8;
constexpr int kNumLists = 5;
constexpr int kQueueWord = 2, kErrWord = 3;
constexpr int kMlpListIndex = 4, kMlpListWord = 6;
constexpr int kStatWords = 17;   // ... [13] tail cells, [14] tail iterations, [15] tail trials,
                                 // [16] learned starts that fell back to isotropicAlpha

// Result layout of v_mfma_f64_16x16x4f64 on gfx950 (measured with a one-wave
// probe): result element r of lane l is row 4r + l/16, column l%16.  (The
// inputs are lane l: A[l%16][l/16], B[l/16][l%16]; the f32/bf16 16x16 forms
// return row 4(l/16) + r instead.)
constexpr int mfma_f64_row(int r, int g) { return 4 * r + g; }

// Operand tables of the mixed-precision sweep (built on the host from Tables):
// the quadrature is padded to a multiple of 32 points (zero weight) and the
// order-2N harmonics to a multiple of 16 columns.
struct MixTables {
    int nt = 0, nnt = 0, nmm = 0;
    std::vector<double> z, g;
    std::vector<unsigned short> h;
    std::vector<int> gptr, gk;
    std::vector<double> gv;
    std::vector<float> pv;   // the Gaunt map in compile-time pattern order (empty: pattern does not cover it)
};
// The fitted Gaunt map (CSR over nm*nm pairs) in the order of the compile-time
// pattern of moment order n (mnc_gaunt_pattern.h), 0 where the fit has no
// entry.  False when the map has an entry outside the pattern.
bool gaunt_pattern_values(int n, const std::vector<int> &gptr, const std::vector<int> &gk,
                          const std::vector<double> &gv, std::vector<float> &out);
// Returns MNC_OK, or MNC_ERR_ARG when the order/quadrature has no mixed kernel.
int build_mix_tables(int moment_order, int quad_order, MixTables &m);

// Operand tables of the FP64 matrix-core sweep (the high orders, where the
// direct Hessian does not fit a lane): quadrature padded to 16-point tiles
// (zero weight), moments to kz = ceil(n/4) groups of 4, the sums to g tiles
// of 16 columns (the last column is w), the order-2N moments to h tiles.
struct G64Tables {
    int nt = 0, kz = 0, ngt = 0, nht = 0, nmm = 0;
    std::vector<double> z, g, h;
    std::vector<int> gptr, gk;
    std::vector<double> gv;
};
int build_g64_tables(int moment_order, int quad_order, G64Tables &m);

// Launchers (mnc_newton.hip).  Return false if (nm, np2, gaunt) has no instance.
// One call enqueues (no host synchronisation) the prologue if `init`, the
// lockstep sweep/logic rounds r0..r1-1 with compaction (round r reads list
// r & 1) and, if `tail`, the persistent tail.  tail_mode 0: one wave per
// remaining cell (cooperative sweeps); 1: one lane per cell; 2: one wave per
// cell, algebra per lane; 3: mode 0 held to two waves per SIMD (direct Mn
// Hessian; mode 0 elsewhere).
struct LaunchPlan {
    int r0 = 0, r1 = 0;
    bool init = true, tail = true;
    int tail_blocks = 0, tail_mode = 0;
    int mix_blocks = 0, defer_rounds = 0;   // mixed precision only
    int mix_split = 3;                       // bf16 parts in the mixed sweep's Hessian products (2 or 3)
};
bool launch_newton_fp64(const NewtonArgs &a, int nm, int np2, bool gaunt, const LaunchPlan &lp, hipStream_t s);
// The control word holding list l's count.
int newton_list_word(int l);
// The same with the lockstep rounds in mixed precision (MFMA sweeps); the tail
// is FP64.
bool launch_newton_mixed(const NewtonArgs &a, int nm, int np2, bool gaunt, const LaunchPlan &lp, hipStream_t s);
// The learned warm start (M4): with NewtonArgs::divert set, the first round's
// logic parks the cells whose start fails the gate on the network's list; the
// judge compares the network's guesses for them and moves them to list 1.
bool launch_newton_judge(const NewtonArgs &a, int nm, int np2, bool gaunt, hipStream_t s);
// The same second half in one kernel: the network (bf16 matrix cores, weights
// in LDS) and the FP64 judge for the cells of the network's list (at most
// max_rows), sixteen cells per wave, `blocks` persistent workgroups.
// The network and judge kernel holds the quadrature in LDS: up to this many points.
constexpr int kNetMaxQ = 512;
bool launch_newton_net_judge(const NewtonArgs &a, int nm, int np2, bool gaunt, int max_rows, int blocks,
                             hipStream_t s);
// The FP64 matrix-core path (M8: n = 45).  rounds: lockstep rounds per launch
// batch; the caller repeats batches until control word `lin` reads zero.
// Returns false for an order without an instance.
bool launch_newton_g64_init(const NewtonArgs &a, int nm, int np2, hipStream_t s);
bool launch_newton_g64_rounds(const NewtonArgs &a, int nm, int np2, int &lin, int rounds, int blocks, hipStream_t s);
// Scratch of the no-allocation entry point (mnc_opt_batch_noalloc): a
// file-scope device array in mnc_newton.hip with the Newton state, lists and
// control words of up to kNoAllocCells cells of moment order <= 4.
constexpr int kNoAllocCells = MNC_NOALLOC_MAX_CELLS;
struct NoAllocState {
    double *st_d = nullptr;
    int *st_i = nullptr, *lists = nullptr, *ctl = nullptr;
    size_t cap = 0;
};
size_t noalloc_scratch_bytes();
// The array's address on `device`, which must be current: resolved with
// hipGetSymbolAddress on the first call for that device and kept per device
// (each GPU holds its own copy of the array at its own address).
hipError_t noalloc_scratch_address(int device, void **ptr);
// The regions of the array at device address base for order (nm, np2); false
// when that order's state does not fit.
bool noalloc_layout(int nm, int np2, void *base, NoAllocState &out);
constexpr int kMixMaxTiles = 16;   // quadrature tiles of 16 points held in LDS (nq <= 256)
constexpr int kMixParts = 3;       // bf16 parts per table value in the mixed sweep's Hessian table
int mix_tiles(int nm);             // 16-column tiles of the order-2N moments the mixed kernel expects
int newton_block();
// Zeroes the solve-control words (a kernel, so the step can be captured).
void launch_ctl_reset(int *ctl, hipStream_t s);
void newton_store_fields(int nm, int np2, int *dfields, int *ifields);

// The learned warm start (mnc_mlp.hip).  w: the 37,007 parameters followed by
// in_mean, in_std, out_mean, out_std (15 each); grad, m, v: like the
// parameters; scratch: mlp_scratch_floats(rows); loss: one float (accumulated).
struct MlpBuffers {
    float *w, *grad, *m, *v, *scratch, *loss;
};
constexpr int kMlpStats = 60;
int mlp_param_count();
size_t mlp_scratch_floats(int rows);
// Minibatch (seed != 0: rows drawn with replacement; seed == 0: rows i0.. in
// order; as == nullptr: inputs only) into the scratch's input/target slots.
void mlp_gather(const MlpBuffers &b, const double *us, const double *as, long long npairs, int rows,
                unsigned long long seed, long long i0, hipStream_t s);
// mdev: if set, the row count is *mdev (at most rows), read on the device.
void mlp_forward(const MlpBuffers &b, int rows, hipStream_t s, const int *mdev = nullptr);
void mlp_backward(const MlpBuffers &b, int rows, hipStream_t s);
void mlp_adam(const MlpBuffers &b, float lr, int step, hipStream_t s);
// Scaled multipliers of the last forward pass into alpha[rows][15] (FP64).
void mlp_unscale(const MlpBuffers &b, int rows, double *alpha, hipStream_t s);
void mlp_scale_moments(const double *u, int ncells, double *us, hipStream_t s);
// The network for the cells listed in list[0 .. *count) (a device count; at
// most max_rows): scaled moments gathered from u (unscaled), guesses scattered
// into alpha[cell][15].
void mlp_predict_listed(const MlpBuffers &b, const double *u, const int *list, const int *count, int max_rows,
                        double *alpha, hipStream_t s);
float *mlp_output(const MlpBuffers &b, int rows);

// Arguments of the grid kernels (mnc_grid.hip).
struct GridArgs {
    int nx, ny, NX, NY, nm, np2, nq;
    double dx, dy, theta, init_x, init_y;
    int init_cond;
    const double *w, *p, *xi, *eta;
    double *moments, *alpha, *moments_old, *sigma_t, *sigma_s, *flux, *ansatz;
    double *fpart;        // [groups][nx*ny][nm]: per-group partial fluxes (flux_tile 3)
    int groups;           // direction groups of flux_tile 3 (1..kMaxFluxGroups)
    int ppn;              // the PPn ansatz (computeAnsatz, momopt_update.cpp:54-57)
    double delta;         // its delta
    // flux_tile 3 on part of the grid: 0 every tile; 1 the tiles whose flag is
    // 0 (no unfinished closure in their stencil region); 2 the flagged ones
    const int *tile_flag; // [tiles_x][tiles_y] of kSplitTile x kSplitTile cells
    int tile_mode;
};
constexpr int kMaxFluxGroups = 8;
constexpr int kSplitTile = 16;   // flux_tile 3's cells per tile side
// flags[tx][ty] = 1 for the flux_tile 3 tiles whose stencil region (the tile
// and two cells around it) holds a cell on list0 (count *cnt0) or list1 (*cnt1);
// 0 for the others.  Grid cell c (ghosts included) is (c / NY, c % NY).
void launch_tile_flags(const int *list0, const int *cnt0, const int *list1, const int *cnt1, int ncells, int nx,
                       int ny, int *flags, hipStream_t s);

// The change-of-basis optimiser (optcb.cpp:90-415) on a batch: one block per
// cell at a time, ws holds each block's table seen through its cell's basis.
struct CbArgs {
    int ncells, nm, nq;
    const double *w, *p;              // quadrature weights, moment table [nq][nm]
    double *u, *P2M, *alphaP, *alphaM;  // [ncells][nm], [ncells][nm*nm] (column-major), [ncells][nm] x2
    mnc_outs *outs;                   // may be null
    double *ws;                       // [blocks][nq][nm]
    int max_iter, ls_mode;
    double tol, cond_h_max;
    int *err;                         // error bits (1 << MNC_ERR_*)
    long long *stats;                 // [kStatWords] as the Newton kernels, or null
    int bfgs, max_bfgs_iter;          // the BFGS variant (optbfgs.cpp:169-513) and its inner count
    double cond_h_max_bfgs;           // its bound on cond(H) cond(H^-1)
    int *queue;                       // the next cell to take, zero at launch
};
// Launches on min(the resident blocks, max_blocks, ncells) blocks (max_blocks:
// the workspace's size); returns that count, or -1 for an order with no
// kernel instance.
int launch_optcb(const CbArgs &a, int max_blocks, int num_cu, hipStream_t s);

void launch_ghost_fill(const GridArgs &g, hipStream_t s);
void launch_ansatz(const GridArgs &g, int q0, int nqc, hipStream_t s);
// stage: 0 = flux only (writes g.flux), 1 = Euler stage, 2 = Euler + Heun average
bool launch_flux_euler(const GridArgs &g, int q0, int nqc, int stage, double dt, bool first_chunk,
                       bool last_chunk, hipStream_t s);
// The ansatz fused into the flux/Euler stage (E kept in LDS per tile):
// tile 1 = 8 x 8 cells, 2 = 16 x 16, 3 = 16 x 16 with the directions split
// into g.groups launches' worth of blocks (needs fpart).
bool launch_flux_fused(const GridArgs &g, int tile, int stage, double dt, hipStream_t s);

}  // namespace mnc

#define MNC_HIP_CHECK(expr)                                                  \
    do {                                                                     \
        hipError_t _e = (expr);                                              \
        if (_e != hipSuccess) return ctx_fail(ctx, MNC_ERR_HIP, #expr, _e);  \
    } while (0)

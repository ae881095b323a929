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
    const float *pf;      // [nq][np2] fp32 copy (mixed precision)
    const unsigned short *pbf; // bf16 MFMA operand table (mixed precision)
    int nq;
    const int *gptr;      // Gaunt CSR (nm*nm+1)
    const int *gk;
    const double *gv;
    int max_iter;
    double tol;
    double cond_h_max;
    int warm;             // MNC_WARM_*
    long long *stats;     // [0] max iter, [1] iter sum, [2] solves, [3..7] hist, [8] trials
    int ls_mode;          // MNC_LS_*
    const double *alpha_guess; // [ncells][nm] scaled-problem guess (learned warm start) or nullptr
    // per-cell solver state and the active lists
    double *st_d;         // [dfields][cap]
    int *st_i;            // [ifields][cap]
    int *lists;           // [2][cap]
    int *ctl;             // [0],[1] list counts, [2] tail queue, [3] error bits
    size_t cap;
};

// Launchers (mnc_newton.hip).  Return false if (nm, np2, gaunt) has no instance.
// Enqueues the whole batched solve: prologue, `rounds` lockstep sweep/logic
// rounds with compaction, then the persistent tail.  No host synchronisation.
// tail_mode 0: one wave per remaining cell (cooperative sweeps); 1: one lane per cell.
bool launch_newton_fp64(const NewtonArgs &a, int nm, int np2, bool gaunt, int rounds, int tail_blocks,
                        int tail_mode, hipStream_t s);
int newton_block();
// Zeroes the four solve-control words (a kernel, so the step can be captured).
void launch_ctl_reset(int *ctl, hipStream_t s);
void newton_store_fields(int nm, int np2, int *dfields, int *ifields);

// Arguments of the grid kernels (mnc_grid.hip).
struct GridArgs {
    int nx, ny, NX, NY, nm, np2, nq;
    double dx, dy, theta, init_x, init_y;
    int init_cond;
    const double *w, *p, *xi, *eta;
    double *moments, *alpha, *moments_old, *sigma_t, *sigma_s, *flux, *ansatz;
};

void launch_ghost_fill(const GridArgs &g, hipStream_t s);
void launch_ansatz(const GridArgs &g, int q0, int nqc, hipStream_t s);
// stage: 0 = flux only (writes g.flux), 1 = Euler stage, 2 = Euler + Heun average
bool launch_flux_euler(const GridArgs &g, int q0, int nqc, int stage, double dt, bool first_chunk,
                       bool last_chunk, hipStream_t s);

}  // namespace mnc

#define MNC_HIP_CHECK(expr)                                                  \
    do {                                                                     \
        hipError_t _e = (expr);                                              \
        if (_e != hipSuccess) return ctx_fail(ctx, MNC_ERR_HIP, #expr, _e);  \
    } while (0)

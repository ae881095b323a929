// mnc_coop_la.h -- dense kernels for one 64-lane wave on a matrix in LDS
// (rows L[i][*], lower factor), used by the wave-per-cell kernels: the FP64
// tail's M8 logic and the change-of-basis optimiser.  The arithmetic of each
// element is LAPACK dpotf2's / dtrsv's, the lanes share the independent
// updates of a step.
#pragma once

#include <hip/hip_runtime.h>

namespace mnc {
namespace {

// v (a double) from lane l (wave-uniform l)
__device__ __forceinline__ double readlane_d(double v, int l) {
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b & 0xffffffffu), l);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b >> 32), l);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// x <- L^{-1} x then L^{-T} x, lane i owning x_i (i < NM); division as dtrsv.
template <int NM>
__device__ __forceinline__ void coop_apply_inv(double (*L)[NM], double *x, int lane) {
    for (int j = 0; j < NM; ++j) {
        if (lane == j) x[j] = x[j] / L[j][j];
        __builtin_amdgcn_wave_barrier();
        const double xj = x[j];
        if (lane > j && lane < NM) x[lane] = fma(-xj, L[lane][j], x[lane]);
        __builtin_amdgcn_wave_barrier();
    }
    for (int j = NM - 1; j >= 0; --j) {
        if (lane == j) x[j] = x[j] / L[j][j];
        __builtin_amdgcn_wave_barrier();
        const double xj = x[j];
        if (lane < j) x[lane] = fma(-L[j][lane], xj, x[lane]);
        __builtin_amdgcn_wave_barrier();
    }
}

// coop_apply_inv with x in registers (x_i on lane i): x_j is read with
// v_readlane instead of a store, a barrier and a load per step; the same
// operations on each element.
template <int NM>
__device__ __forceinline__ void coop_apply_inv_reg(double (*L)[NM], double &x, int lane) {
    for (int j = 0; j < NM; ++j) {
        const double xj = readlane_d(x, j) / L[j][j];
        if (lane == j) x = xj;
        else if (lane > j && lane < NM) x = fma(-xj, L[lane][j], x);
    }
    for (int j = NM - 1; j >= 0; --j) {
        const double xj = readlane_d(x, j) / L[j][j];
        if (lane == j) x = xj;
        else if (lane < j) x = fma(-L[j][lane], xj, x);
    }
}

// Cholesky in place on the lower triangle (dpotf2 order); uniform result.
template <int NM>
__device__ __forceinline__ bool coop_cholesky(double (*L)[NM], int lane) {
    bool ok = true;
    for (int j = 0; j < NM; ++j) {
        // the inner loops unrolled so that their LDS reads are issued ahead of
        // the dependent FMA chains (the same FMAs in the same order)
        double dot = 0.0;
#pragma unroll 8
        for (int k = 0; k < j; ++k) dot = fma(L[j][k], L[j][k], dot);
        double ajj = L[j][j] - dot;
        ok = ok && (ajj > 0.0);
        ajj = sqrt(ajj > 0.0 ? ajj : 1.0);
        double y = 0.0;
        if (lane > j && lane < NM) {
            y = L[lane][j];
#pragma unroll 8
            for (int k = 0; k < j; ++k) y = fma(-L[j][k], L[lane][k], y);
        }
        __builtin_amdgcn_wave_barrier();
        if (lane == j) L[j][j] = ajj;
        const double r = 1.0 / ajj;
        if (lane > j && lane < NM) L[lane][j] = y * r;
        __builtin_amdgcn_wave_barrier();
    }
    return ok;
}

// Higham's estimate of ||A^{-1}||_1 (dpocon's role); every lane ends with it.
template <int NM>
__device__ __forceinline__ double coop_rcond(double (*L)[NM], double *x, double anorm, int lane) {
    if (anorm == 0.0) return 0.0;
    unsigned isgn = 0;
    double est = 0.0;
    int jump = 1, j = 0, iter = 0;
    if (lane < NM) x[lane] = 1.0 / NM;
    __builtin_amdgcn_wave_barrier();
    for (int guard = 0; guard < 16; ++guard) {
        coop_apply_inv<NM>(L, x, lane);
        double s = 0.0, m = -1.0;
        unsigned sg = 0;
        int im = 0;
        for (int i = 0; i < NM; ++i) {
            const double v = x[i];
            s += fabs(v);
            if (fabs(v) > m) { m = fabs(v); im = i; }
            sg |= !(v >= 0.0) ? (1u << i) : 0u;
        }
        const double xj_old = x[j];
        __builtin_amdgcn_wave_barrier();
        double nx = 0.0;
        bool done = false;
        if (jump == 1) {
            if (NM == 1) { est = fabs(x[0]); done = true; }
            else { est = s; isgn = sg; nx = (lane < NM && (sg >> lane & 1u)) ? -1.0 : 1.0; jump = 2; }
        } else if (jump == 2) {
            j = im; iter = 2; nx = (lane == j) ? 1.0 : 0.0; jump = 3;
        } else if (jump == 3) {
            const double estold = est;
            est = s;
            if (sg == isgn || est <= estold) {
                nx = ((lane & 1) ? -1.0 : 1.0) * (1.0 + double(lane) / double(NM - 1));
                jump = 5;
            } else {
                nx = (lane < NM && (sg >> lane & 1u)) ? -1.0 : 1.0;
                isgn = sg;
                jump = 4;
            }
        } else if (jump == 4) {
            const int jlast = j;
            j = im;
            (void)jlast;   // xj_old is x[jlast], read above before j moved
            if (xj_old != m && iter < 5) {
                ++iter; nx = (lane == j) ? 1.0 : 0.0; jump = 3;
            } else {
                nx = ((lane & 1) ? -1.0 : 1.0) * (1.0 + double(lane) / double(NM - 1));
                jump = 5;
            }
        } else {
            const double temp = 2.0 * (s / double(3 * NM));
            if (temp > est) est = temp;
            done = true;
        }
        if (done) break;
        if (lane < NM) x[lane] = nx;
        __builtin_amdgcn_wave_barrier();
    }
    if (!isfinite(est) || est == 0.0) return 0.0;
    return (1.0 / est) / anorm;
}

__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o));
    return v;
}

// True when the condition test `1/rcond > limit` (opt.cpp:231-244) is false
// whatever the estimate: Higham's estimate never exceeds ||A^-1||_1 <=
// sqrt(n) / lambda_min, and Gershgorin gives lambda_min >= gersh =
// min_i (2 a_ii - sum_j |a_ij|).  With a factor 2 of margin for rounding, the
// estimate (the condition number's main cost after the factorisation) is
// skipped when anorm sqrt(n) / gersh cannot reach the limit; the decision is
// the reference's either way.
__device__ __forceinline__ bool cond_proven_below(double anorm, double gersh, int n, double limit) {
    return gersh > 0.0 && 2.0 * anorm * sqrt(double(n)) <= limit * gersh;
}

__device__ __forceinline__ double wave_min(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o));
    return v;
}

// v from the lane a DPP control names within each 16-lane row (CTRL: a quad
// permutation or a row rotation, so that every lane has a source)
template <int CTRL>
__device__ __forceinline__ double dpp_row_d(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_mov_dpp((int)b, CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), CTRL, 0xF, 0xF, false);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// The largest (smallest) value of lanes 0..15, wave-uniform: swaps within the
// quads (lane ^ 1, lane ^ 2), then rotations of the row by 4 and 8 lanes.
// For vectors of at most 16 elements on lanes 0..15 with neutral values
// elsewhere this is wave_max (wave_min) exactly -- max and min do not round
// and ignore a NaN beside a number in any order -- in four DPP steps instead
// of six LDS permutes.
__device__ __forceinline__ double row16_max(double v) {
    v = fmax(v, dpp_row_d<0xB1>(v));
    v = fmax(v, dpp_row_d<0x4E>(v));
    v = fmax(v, dpp_row_d<0x124>(v));
    v = fmax(v, dpp_row_d<0x128>(v));
    return readlane_d(v, 0);
}
__device__ __forceinline__ double row16_min(double v) {
    v = fmin(v, dpp_row_d<0xB1>(v));
    v = fmin(v, dpp_row_d<0x4E>(v));
    v = fmin(v, dpp_row_d<0x124>(v));
    v = fmin(v, dpp_row_d<0x128>(v));
    return readlane_d(v, 0);
}

// ---- wave-parallel dense algebra with the rows in registers ----------------
//
// Lane i (< NM) holds row i of the symmetric matrix (row[k] = A[i][k]); after
// reg_cholesky, row[k] = L[i][k] for k <= i, and col[k] = L[k][i] (k > i) is
// filled from it for the back substitution.  Entries of other lanes are read
// with v_readlane (wave-uniform lane index), so each dependent step of the
// factorisation and of the triangular solves is a register operation rather
// than an LDS round trip.  The arithmetic and its order are those of
// cholesky<N>, apply_inv<N> and rcond<N>.

template <int NM>
__device__ __forceinline__ bool reg_cholesky(double (&row)[NM], int lane) {
    bool ok = true;
#pragma unroll
    for (int j = 0; j < NM; ++j) {
        double dot = 0.0;
#pragma unroll
        for (int k = 0; k < j; ++k) dot = fma(row[k], row[k], dot);
        double ajj = readlane_d(row[j] - dot, j);
        ok = ok && (ajj > 0.0);
        ajj = sqrt(ajj > 0.0 ? ajj : 1.0);
        double y = row[j];
#pragma unroll
        for (int k = 0; k < j; ++k) y = fma(-readlane_d(row[k], j), row[k], y);
        const double r = 1.0 / ajj;
        if (lane == j) row[j] = ajj;
        else if (lane > j && lane < NM) row[j] = y * r;
    }
    return ok;
}

// x <- (L L^T)^{-1} x, x_i on lane i, with the reciprocals of the diagonal
// (wave-uniform rd[j]) in place of divisions: the tail's Newton direction and
// the condition estimate's repeated solves.
template <int NM>
__device__ __forceinline__ void reg_apply_inv_rcp(const double (&row)[NM], const double (&col)[NM],
                                                  const double (&rd)[NM], double &x, int lane) {
#pragma unroll
    for (int j = 0; j < NM; ++j) {
        const double xj = readlane_d(x, j) * rd[j];
        if (lane == j) x = xj;
        else if (lane > j && lane < NM) x = fma(-xj, row[j], x);
    }
#pragma unroll
    for (int j = NM - 1; j >= 0; --j) {
        const double xj = readlane_d(x, j) * rd[j];
        if (lane == j) x = xj;
        else if (lane < j) x = fma(-col[j], xj, x);
    }
}

// Reciprocal 1-norm condition estimate (dpocon's role; Higham 1988), x_i on lane i.
template <int NM>
__device__ __forceinline__ double reg_rcond(const double (&row)[NM], const double (&col)[NM], double anorm, int lane) {
    if (anorm == 0.0) return 0.0;
    unsigned isgn = 0;
    double est = 0.0;
    int jump = 1, j = 0, iter = 0;
    double x = lane < NM ? 1.0 / NM : 0.0;
    double rd[NM];
#pragma unroll
    for (int k = 0; k < NM; ++k) rd[k] = 1.0 / readlane_d(row[k], k);
#pragma unroll 1
    for (int guard = 0; guard < 16; ++guard) {
        reg_apply_inv_rcp<NM>(row, col, rd, x, lane);
        double s = 0.0, m = -1.0;
        int im = 0;
#pragma unroll
        for (int i = 0; i < NM; ++i) {                 // dasum / idamax in the serial order
            const double v = fabs(readlane_d(x, i));
            s += v;
            if (v > m) { m = v; im = i; }
        }
        const unsigned sg = (unsigned)(__ballot(lane < NM && !(x >= 0.0)) & ((NM >= 32) ? 0xffffffffull : ((1ull << NM) - 1)));
        const double xj_old = readlane_d(x, j);
        double nx = 0.0;
        bool done = false;
        if (jump == 1) {
            if (NM == 1) { est = fabs(readlane_d(x, 0)); done = true; }
            else { est = s; isgn = sg; nx = (lane < NM && (sg >> lane & 1u)) ? -1.0 : 1.0; jump = 2; }
        } else if (jump == 2) {
            j = im; iter = 2; nx = (lane == j) ? 1.0 : 0.0; jump = 3;
        } else if (jump == 3) {
            const double estold = est;
            est = s;
            if (sg == isgn || est <= estold) {
                nx = ((lane & 1) ? -1.0 : 1.0) * (1.0 + double(lane) / double(NM - 1));
                jump = 5;
            } else {
                nx = (lane < NM && (sg >> lane & 1u)) ? -1.0 : 1.0;
                isgn = sg;
                jump = 4;
            }
        } else if (jump == 4) {
            j = im;
            if (xj_old != m && iter < 5) {
                ++iter; nx = (lane == j) ? 1.0 : 0.0; jump = 3;
            } else {
                nx = ((lane & 1) ? -1.0 : 1.0) * (1.0 + double(lane) / double(NM - 1));
                jump = 5;
            }
        } else {
            const double temp = 2.0 * (s / double(3 * NM));
            if (temp > est) est = temp;
            done = true;
        }
        if (done) break;
        x = lane < NM ? nx : 0.0;
    }
    if (!isfinite(est) || est == 0.0) return 0.0;
    return (1.0 / est) / anorm;
}

// ||A^{-1}||_1 for A = L L^T from the factor (rows in registers as after
// reg_cholesky: lane i holds L[i][k], k <= i; rd[k] = 1 / L[k][k]).  Lane j
// forms column j of Z = L^{-1} by forward substitution (every column at once,
// the products L[i][k] z_k issued as soon as z_k is final), then column j of
// A^{-1} = Z^T Z as dot products of Z's columns, and the wave takes the largest
// absolute column sum: about 250 register reads and 240 FMAs, against the tens
// of dependent triangular-solve steps of the estimate.  Higham's estimate never
// exceeds ||A^{-1}||_1, so where anorm times this (with a factor 2 of margin for
// rounding) is within the limit the estimate cannot fail the condition test.
template <int NM>
__device__ __forceinline__ double reg_inv_norm1(const double (&row)[NM], const double (&rd)[NM], int lane) {
    double z[NM];
#pragma unroll
    for (int i = 0; i < NM; ++i) z[i] = (i == lane) ? 1.0 : 0.0;
#pragma unroll
    for (int k = 0; k < NM; ++k) {
        z[k] = z[k] * rd[k];
#pragma unroll
        for (int i = k + 1; i < NM; ++i) z[i] = fma(-readlane_d(row[k], i), z[k], z[i]);
    }
    double c = 0.0;
#pragma unroll
    for (int i = 0; i < NM; ++i) {
        double d = 0.0;
#pragma unroll
        for (int k = i; k < NM; ++k) d = fma(readlane_d(z[k], i), z[k], d);
        c += fabs(d);
    }
    if constexpr (NM <= 16) return row16_max(lane < NM ? c : 0.0);
    else return wave_max(lane < NM ? c : 0.0);
}

// reg_inv_norm1 that also solves x <- (L L^T)^{-1} x (x_i on lane i) for the
// cost of one more lane: lane NM carries x through the forward substitution
// that forms Z, and the back substitution then runs beside the Z^T Z
// products.  Lane NM's updates fma(-L[i][k], z_k, z_i) are
// reg_apply_inv_rcp's fma(-x_k, L[i][k], x_i) (the same exact product), in
// the same order, and the back substitution is its second loop, so x comes
// out bit for bit as reg_apply_inv_rcp leaves it.
template <int NM>
__device__ __forceinline__ double reg_inv_norm1_solve(const double (&row)[NM], const double (&col)[NM],
                                                      const double (&rd)[NM], double &x, int lane) {
    static_assert(NM < 64, "lane NM carries the right-hand side");
    double z[NM];
#pragma unroll
    for (int i = 0; i < NM; ++i) {
        const double xi = readlane_d(x, i);
        z[i] = lane == NM ? xi : ((i == lane) ? 1.0 : 0.0);
    }
#pragma unroll
    for (int k = 0; k < NM; ++k) {
        z[k] = z[k] * rd[k];
#pragma unroll
        for (int i = k + 1; i < NM; ++i) z[i] = fma(-readlane_d(row[k], i), z[k], z[i]);
    }
    // y = L^{-1} x from lane NM, y_j on lane j (where reg_apply_inv_rcp's
    // forward loop leaves it), then its back substitution
    double y = 0.0;
#pragma unroll
    for (int j = 0; j < NM; ++j) {
        const double v = readlane_d(z[j], NM);
        y = lane == j ? v : y;
    }
#pragma unroll
    for (int j = NM - 1; j >= 0; --j) {
        const double xj = readlane_d(y, j) * rd[j];
        if (lane == j) y = xj;
        else if (lane < j) y = fma(-col[j], xj, y);
    }
    x = lane < NM ? y : x;
    double c = 0.0;
#pragma unroll
    for (int i = 0; i < NM; ++i) {
        double d = 0.0;
#pragma unroll
        for (int k = i; k < NM; ++k) d = fma(readlane_d(z[k], i), z[k], d);
        c += fabs(d);
    }
    if constexpr (NM <= 16) return row16_max(lane < NM ? c : 0.0);
    else return wave_max(lane < NM ? c : 0.0);
}

}  // namespace
}  // namespace mnc

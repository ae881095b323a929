// mnc_coop_la.h -- dense kernels for one 64-lane wave on a matrix in LDS
// (rows L[i][*], lower factor), used by the wave-per-cell kernels: the FP64
// tail's M8 logic and the change-of-basis optimiser.  The arithmetic of each
// element is LAPACK dpotf2's / dtrsv's, the lanes share the independent
// updates of a step.
#pragma once

#include <hip/hip_runtime.h>

namespace mnc {
namespace {

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

// Cholesky in place on the lower triangle (dpotf2 order); uniform result.
template <int NM>
__device__ __forceinline__ bool coop_cholesky(double (*L)[NM], int lane) {
    bool ok = true;
    for (int j = 0; j < NM; ++j) {
        double dot = 0.0;
        for (int k = 0; k < j; ++k) dot = fma(L[j][k], L[j][k], dot);
        double ajj = L[j][j] - dot;
        ok = ok && (ajj > 0.0);
        ajj = sqrt(ajj > 0.0 ? ajj : 1.0);
        double y = 0.0;
        if (lane > j && lane < NM) {
            y = L[lane][j];
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

}  // namespace
}  // namespace mnc

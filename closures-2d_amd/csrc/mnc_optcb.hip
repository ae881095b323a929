// mnc_optcb.hip -- the change-of-basis optimiser (src/momopt/opt/optcb.cpp:90-415,
// MOMOPT_OPTIMIZATION_TYPE 1) for batches of Mn closures on gfx950.
//
// One 64-lane block per cell.  The cell's basis P2M (n x n lower triangular,
// column-major as the reference's BLAS calls see it), its multipliers alphaP
// in that basis, and the quadrature table seen through the basis, pP[q][k] =
// (P2M^{-1} p_q)_k, live in HBM (pP in a workspace of nq*n doubles per cell).
// The quadrature work -- the table's triangular solves, f, g and the Hessian
// sums, the line search's f -- is spread over the lanes (lane l takes the
// points q = l, l + 64, ...; wave sums by butterfly).  The n x n algebra
// (Cholesky, condition estimate, the basis products and solves) runs on lane 0
// in the reference's order, on LDS copies, between block barriers.
#include "mnc_internal.h"

#include <cfloat>

namespace mnc {
namespace {

__constant__ double kCbReg[5] = {0.0, 1e-3, 1e-2, 1e-1, 5e-1};   // opt.h:39-40
#define kCbSqrtPi sqrt(M_PI)

enum : int { CB_STEP = 0, CB_RESET = 1, CB_REG = 2, CB_FAIL = 3 };

template <int N>
struct CbShared {
    double P2M[N * N], P2Min[N * N];   // column-major (element (i,j) at [i + j*N])
    double H[N * N], L[N * N];         // row-major; L lower
    double u[N], uP[N], aP[N], aP0[N], gP[N], dP[N], gM[N], aold[N], aM[N];
    double x[N], v[N];
    int isgn[N];
    double f, fscale, rtol, fnew, t, err;
    int action, done, converged;
};

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

// ---- n x n kernels (lane 0), the Netlib reference BLAS / LAPACK loop orders ----
// dtrsv('L','N','N') on a column-major A: x <- A^{-1} x
template <int N>
__device__ void cb_trsv_ln(const double *a, double *x) {
    for (int j = 0; j < N; j++) {
        if (x[j] != 0.0) {
            x[j] = x[j] / a[j + j * N];
            const double temp = x[j];
            for (int i = j + 1; i < N; i++) x[i] = x[i] - temp * a[i + j * N];
        }
    }
}
// dtrsv('L','T','N'): x <- A^{-T} x
template <int N>
__device__ void cb_trsv_lt(const double *a, double *x) {
    for (int j = N - 1; j >= 0; j--) {
        double temp = x[j];
        for (int i = N - 1; i > j; i--) temp = temp - a[i + j * N] * x[i];
        x[j] = temp / a[j + j * N];
    }
}
// dtrmv('L','N','N'): x <- A x
template <int N>
__device__ void cb_trmv_ln(const double *a, double *x) {
    for (int j = N - 1; j >= 0; j--) {
        if (x[j] != 0.0) {
            const double temp = x[j];
            for (int i = N - 1; i > j; i--) x[i] = x[i] + temp * a[i + j * N];
            x[j] = x[j] * a[j + j * N];
        }
    }
}
// dtrmv('L','T','N'): x <- A^T x
template <int N>
__device__ void cb_trmv_lt(const double *a, double *x) {
    for (int j = 0; j < N; j++) {
        double temp = x[j] * a[j + j * N];
        for (int i = j + 1; i < N; i++) temp = temp + a[i + j * N] * x[i];
        x[j] = temp;
    }
}
// dtrmm('R','L','N','N', N, N, 1, A, B): B <- B A (both column-major)
template <int N>
__device__ void cb_trmm_rln(const double *a, double *b) {
    for (int j = 0; j < N; j++) {
        double temp = a[j + j * N];
        for (int i = 0; i < N; i++) b[i + j * N] = temp * b[i + j * N];
        for (int k = j + 1; k < N; k++) {
            if (a[k + j * N] != 0.0) {
                temp = a[k + j * N];
                for (int i = 0; i < N; i++) b[i + j * N] = b[i + j * N] + temp * b[i + k * N];
            }
        }
    }
}
template <int N>
__device__ void cb_identity(double *a) {
    for (int i = 0; i < N * N; i++) a[i] = 0.0;
    for (int i = 0; i < N; i++) a[i + i * N] = 1.0;
}
__device__ double cb_nrm2(int n, const double *x) {
    double s = 0.0;
    for (int i = 0; i < n; i++) s += x[i] * x[i];
    return sqrt(s);
}
// dpotf2 on a row-major copy (lower); 0 or the failing column + 1
template <int N>
__device__ int cb_cholesky(double *a) {
    for (int j = 0; j < N; j++) {
        double dot = 0.0;
        for (int k = 0; k < j; k++) dot += a[j * N + k] * a[j * N + k];
        double ajj = a[j * N + j] - dot;
        if (ajj <= 0.0 || isnan(ajj)) {
            a[j * N + j] = ajj;
            return j + 1;
        }
        ajj = sqrt(ajj);
        a[j * N + j] = ajj;
        for (int i = j + 1; i < N; i++) {
            double y = a[i * N + j];
            for (int k = 0; k < j; k++) y += (-a[j * N + k]) * a[i * N + k];
            a[i * N + j] = y;
        }
        const double r = 1.0 / ajj;
        for (int i = j + 1; i < N; i++) a[i * N + j] *= r;
    }
    return 0;
}
// x <- (L L^T)^{-1} x with L row-major lower
template <int N>
__device__ void cb_apply_inv_rm(const double *l, double *x) {
    for (int j = 0; j < N; j++) {
        if (x[j] != 0.0) {
            x[j] = x[j] / l[j * N + j];
            const double temp = x[j];
            for (int i = j + 1; i < N; i++) x[i] = x[i] - temp * l[i * N + j];
        }
    }
    for (int j = N - 1; j >= 0; j--) {
        double temp = x[j];
        for (int i = N - 1; i > j; i--) temp = temp - l[i * N + j] * x[i];
        x[j] = temp / l[j * N + j];
    }
}
template <int N>
__device__ double cb_dasum(const double *x) {
    double s = 0.0;
    for (int i = 0; i < N; i++) s += fabs(x[i]);
    return s;
}
template <int N>
__device__ int cb_idamax(const double *x) {
    int idx = 0;
    double m = fabs(x[0]);
    for (int i = 1; i < N; i++)
        if (fabs(x[i]) > m) { m = fabs(x[i]); idx = i; }
    return idx;
}
// dpocon's role: Higham's estimate of ||A^{-1}||_1 (ACM TOMS 14(4), 1988), then
// the reciprocal condition number
template <int N>
__device__ double cb_rcond(const double *l, double anorm, double *x, double *v, int *isgn) {
    if (anorm == 0.0) return 0.0;
    double est;
    for (int i = 0; i < N; i++) x[i] = 1.0 / N;
    cb_apply_inv_rm<N>(l, x);
    if (N == 1) {
        est = fabs(x[0]);
    } else {
        est = cb_dasum<N>(x);
        for (int i = 0; i < N; i++) {
            x[i] = (x[i] >= 0.0) ? 1.0 : -1.0;
            isgn[i] = (int)x[i];
        }
        cb_apply_inv_rm<N>(l, x);
        int j = cb_idamax<N>(x);
        int iter = 2;
        for (;;) {
            for (int i = 0; i < N; i++) x[i] = 0.0;
            x[j] = 1.0;
            cb_apply_inv_rm<N>(l, x);
            for (int i = 0; i < N; i++) v[i] = x[i];
            const double estold = est;
            est = cb_dasum<N>(v);
            bool repeated = true;
            for (int i = 0; i < N; i++)
                if (((x[i] >= 0.0) ? 1 : -1) != isgn[i]) { repeated = false; break; }
            if (repeated || est <= estold) break;
            for (int i = 0; i < N; i++) {
                x[i] = (x[i] >= 0.0) ? 1.0 : -1.0;
                isgn[i] = (int)x[i];
            }
            cb_apply_inv_rm<N>(l, x);
            const int jlast = j;
            j = cb_idamax<N>(x);
            if (x[jlast] != fabs(x[j]) && iter < 5) {
                ++iter;
                continue;
            }
            break;
        }
        double altsgn = 1.0;
        for (int i = 0; i < N; i++) {
            x[i] = altsgn * (1.0 + (double)i / (double)(N - 1));
            altsgn = -altsgn;
        }
        cb_apply_inv_rm<N>(l, x);
        const double temp = 2.0 * (cb_dasum<N>(x) / (double)(3 * N));
        if (temp > est) est = temp;
    }
    if (!isfinite(est) || est == 0.0) return 0.0;
    return (1.0 / est) / anorm;
}

// ---- the table seen through a basis (all lanes, their points) ----
// pP_q <- p_q, then (if basis) pP_q <- B^{-1} pP_q with B column-major
template <int N>
__device__ void cb_table_reset(const CbArgs &A, double *pP, const double *basis, int lane) {
    for (int q = lane; q < A.nq; q += 64) {
        double x[N];
#pragma unroll
        for (int k = 0; k < N; k++) x[k] = A.p[size_t(q) * N + k];
        if (basis) cb_trsv_ln<N>(basis, x);
#pragma unroll
        for (int k = 0; k < N; k++) pP[size_t(q) * N + k] = x[k];
    }
}
// pP_q <- L^{-1} pP_q with L row-major lower (dtrsm('L','L','N','N') by columns)
template <int N>
__device__ void cb_table_solve(const CbArgs &A, double *pP, const double *l, int lane) {
    for (int q = lane; q < A.nq; q += 64) {
        double x[N];
#pragma unroll
        for (int k = 0; k < N; k++) x[k] = pP[size_t(q) * N + k];
        for (int j = 0; j < N; j++) {
            if (x[j] != 0.0) {
                x[j] = x[j] / l[j * N + j];
                const double temp = x[j];
                for (int i = j + 1; i < N; i++) x[i] = x[i] - temp * l[i * N + j];
            }
        }
#pragma unroll
        for (int k = 0; k < N; k++) pP[size_t(q) * N + k] = x[k];
    }
}

// f, g and H in the current basis (fobj.cpp:37-129,255-287 on pP): the lanes'
// partial sums, then wave sums; lane 0 stores.
template <int N>
__device__ void cb_fobj_full(const CbArgs &A, const double *pP, CbShared<N> &sh, int lane) {
    constexpr int TRI = N * (N + 1) / 2;
    double part[TRI];
#pragma unroll
    for (int k = 0; k < TRI; k++) part[k] = 0.0;
    for (int q = lane; q < A.nq; q += 64) {
        double pq[N];
#pragma unroll
        for (int k = 0; k < N; k++) pq[k] = pP[size_t(q) * N + k];
        double temp = 0.0;
#pragma unroll
        for (int i = 0; i < N; i++) temp = temp + sh.aP[i] * pq[i];
        temp = exp(temp);
        const double wq = A.w[q];
#pragma unroll
        for (int i = 0; i < N; i++)
#pragma unroll
            for (int j = 0; j <= i; j++) part[i * (i + 1) / 2 + j] += wq * pq[i] * pq[j] * temp;
    }
#pragma unroll
    for (int i = 0; i < N; i++)
#pragma unroll
        for (int j = 0; j <= i; j++) {
            const double s = wave_sum(part[i * (i + 1) / 2 + j]);
            if (lane == 0) sh.H[i * N + j] = sh.H[j * N + i] = s;
        }
    if (lane == 0) {
        const double p0 = pP[0];
        double f = sh.H[0] / (p0 * p0), fscale = fabs(f);
        for (int i = 0; i < N; i++) {
            sh.gP[i] = sh.H[i] / p0 - sh.uP[i];
            f = f - sh.aP[i] * sh.uP[i];
            fscale += fabs(sh.aP[i] * sh.uP[i]);
        }
        sh.f = f;
        sh.fscale = fscale;
    }
}

// f alone at alpha a (fobj.cpp:291-321 on pP)
template <int N>
__device__ double cb_fobj_f(const CbArgs &A, const double *pP, const double *a, const double *uP, int lane) {
    double s = 0.0;
    for (int q = lane; q < A.nq; q += 64) {
        double temp = 0.0;
#pragma unroll
        for (int i = 0; i < N; i++) temp = temp + a[i] * pP[size_t(q) * N + i];
        s = s + A.w[q] * exp(temp);
    }
    double f = __shfl(wave_sum(s), 0);                   // lane 0's sum on every lane
    for (int i = 0; i < N; i++) f = f - a[i] * uP[i];
    return f;
}

template <int N>
__device__ void cb_isotropic(double *a) {     // isotropicAlpha, Mn (opt.cpp:27-35)
    a[0] = 2.0 * kCbSqrtPi * log(0.5 / kCbSqrtPi);
    for (int i = 1; i < N; i++) a[i] = 0.0;
}

template <int N>
__global__ __launch_bounds__(64) void optcb_kernel(CbArgs A) {
    __shared__ CbShared<N> sh;
    const int lane = threadIdx.x;
    for (int c = blockIdx.x; c < A.ncells; c += gridDim.x) {
        double *pP = A.ws + size_t(blockIdx.x) * A.nq * N;
        double *ug = A.u + size_t(c) * N;
        double *Bg = A.P2M + size_t(c) * N * N;
        double *aPg = A.alphaP + size_t(c) * N;
        const double u0 = ug[0];
        if (!(u0 > 0.0)) {                                 // optcb.cpp:104-108
            if (lane == 0) {
                atomicOr(A.err, 1 << MNC_ERR_U0);
                if (A.outs) { A.outs[c].iter = -1; A.outs[c].r = 0.0; A.outs[c].norm_g = 0.0; }
            }
            continue;
        }
        if (lane == 0) {
            for (int k = 0; k < N; k++) sh.u[k] = ug[k] / u0;
            for (int k = 0; k < N; k++) sh.aP[k] = sh.aP0[k] = aPg[k];
            for (int k = 0; k < N * N; k++) sh.P2M[k] = sh.P2Min[k] = Bg[k];
            for (int k = 0; k < N; k++) sh.uP[k] = sh.u[k];
            cb_trsv_ln<N>(sh.P2M, sh.uP);                   // :198-200
            sh.rtol = cb_nrm2(N, sh.u) * A.tol;
            sh.done = 0;
            sh.converged = 0;
        }
        __syncthreads();
        cb_table_reset<N>(A, pP, sh.P2Min, lane);           // :185-187
        __syncthreads();
        int iter = 0, total = 0, maxit = A.max_iter, reg = 0, hitmax = 0;
        double t = 1.0, r = 0.0;
        for (;;) {
            ++iter;
            ++total;
            cb_fobj_full<N>(A, pP, sh, lane);               // :234-242
            __syncthreads();
            if (lane == 0) {                                // :247-259
                for (int k = 0; k < N * N; k++) sh.L[k] = sh.H[k];
                const int error = cb_cholesky<N>(sh.L);
                double norm = 0.0;                          // utils_norm1
                for (int i = 0; i < N; i++) {
                    double s = 0.0;
                    for (int j = 0; j < N; j++) s += fabs(sh.H[i * N + j]);
                    if (s > norm) norm = s;
                }
                double condest = 0.0;
                if (error == 0) condest = 1.0 / cb_rcond<N>(sh.L, norm, sh.x, sh.v, sh.isgn);
                if (total == 1 && (error != 0 || condest > A.cond_h_max)) sh.action = CB_RESET;
                else if (iter >= maxit || error != 0 || t == 0.0 || condest > A.cond_h_max) sh.action = CB_REG;
                else sh.action = CB_STEP;
            }
            __syncthreads();
            const int action = sh.action;
            if (action == CB_RESET) {                       // :264-281
                if (lane == 0) {
                    cb_isotropic<N>(sh.aP0);
                    cb_identity<N>(sh.P2M);
                    cb_identity<N>(sh.P2Min);
                    for (int k = 0; k < N; k++) sh.aP[k] = sh.aP0[k];
                    for (int k = 0; k < N; k++) sh.uP[k] = sh.u[k];
                    sh.rtol = cb_nrm2(N, sh.u) * A.tol;
                }
                __syncthreads();
                cb_table_reset<N>(A, pP, nullptr, lane);
                __syncthreads();
                iter = 0;
                t = 1.0;
                continue;
            }
            if (action == CB_REG) {                         // :285-329
                if (reg < 4) {
                    ++reg;
                    r = kCbReg[reg];
                    if (lane == 0) {
                        sh.uP[0] = 1.0;
                        for (int i = 1; i < N; i++) sh.uP[i] = sh.u[i] * (1.0 - r);
                    }
                } else {
                    if (hitmax) {                           // "Max regularization achieved."
                        if (lane == 0) {
                            atomicOr(A.err, 1 << MNC_ERR_MAXREG);
                            if (A.outs) { A.outs[c].iter = total; A.outs[c].r = -1.0; A.outs[c].norm_g = sh.err; }
                        }
                        __syncthreads();
                        break;
                    }
                    if (lane == 0) {
                        cb_isotropic<N>(sh.aP0);
                        cb_identity<N>(sh.P2Min);
                        for (int k = 0; k < N; k++) sh.uP[k] = sh.u[k];
                    }
                    reg = 0;
                    hitmax = 1;
                    maxit = A.max_iter;
                }
                if (lane == 0) {
                    sh.rtol = cb_nrm2(N, sh.uP) * A.tol;
                    for (int k = 0; k < N; k++) sh.aP[k] = sh.aP0[k];
                    for (int k = 0; k < N * N; k++) sh.P2M[k] = sh.P2Min[k];
                    cb_trsv_ln<N>(sh.P2M, sh.uP);
                }
                __syncthreads();
                cb_table_reset<N>(A, pP, sh.P2Min, lane);
                __syncthreads();
                iter = 0;
                t = 1.0;
                if (reg == 4) maxit = 10 * maxit;
                continue;
            }
            // CB_STEP: the basis moves by L (:333-343)
            if (lane == 0) {
                double Lc[N * N];                           // column-major view of the factor
                for (int i = 0; i < N; i++)
                    for (int j = 0; j < N; j++) Lc[i + j * N] = (i >= j) ? sh.L[i * N + j] : 0.0;
                cb_trsv_ln<N>(Lc, sh.gP);
                cb_trmm_rln<N>(Lc, sh.P2M);
                cb_trmv_lt<N>(Lc, sh.aP);
                cb_trsv_ln<N>(Lc, sh.uP);
                for (int i = 0; i < N; i++) sh.dP[i] = -sh.gP[i];   // :347-348
                for (int i = 0; i < N; i++) sh.gM[i] = sh.gP[i];    // :352-354
                cb_trmv_ln<N>(sh.P2M, sh.gM);
                sh.err = cb_nrm2(N, sh.gM);
                sh.converged = sh.err <= sh.rtol;
                if (sh.converged) {                         // :358-375
                    for (int i = 0; i < N; i++) sh.aM[i] = sh.aP[i];
                    cb_trsv_lt<N>(sh.P2M, sh.aM);
                }
            }
            cb_table_solve<N>(A, pP, sh.L, lane);
            __syncthreads();
            if (sh.converged) break;
            // line search in the current basis (linesearch.cpp:21-99 on pP)
            double maxf = 0.0;
            for (int i = 0; i < N; i++) {
                const double tmp = fabs(sh.dP[i]) / (fabs(sh.aP[i]) + 1e-100);
                if (tmp > maxf) maxf = tmp;
            }
            double gd = 0.0;
            for (int i = 0; i < N; i++) gd += sh.gP[i] * sh.dP[i];
            __syncthreads();
            if (lane == 0)
                for (int i = 0; i < N; i++) sh.aold[i] = sh.aP[i];
            __syncthreads();
            double tt = 1.0;
            bool accepted = false;
            while (tt * maxf > DBL_EPSILON / 2) {
                if (lane == 0)
                    for (int i = 0; i < N; i++) sh.aP[i] = sh.aold[i] + tt * sh.dP[i];
                __syncthreads();
                const double fnew = cb_fobj_f<N>(A, pP, sh.aP, sh.uP, lane);
                bool acc = fnew <= sh.f + 1e-3 * tt * gd;
                if (!acc && A.ls_mode == MNC_LS_ROBUST && tt == 1.0 && -gd <= 1e3 * DBL_EPSILON * sh.fscale)
                    acc = true;
                __syncthreads();
                if (acc) { accepted = true; break; }
                tt = 0.5 * tt;
            }
            if (!accepted) {
                if (lane == 0)
                    for (int i = 0; i < N; i++) sh.aP[i] = sh.aold[i];
                tt = 0.0;
            }
            t = tt;
            __syncthreads();
        }
        if (sh.converged) {
            if (lane == 0) {                                // outputs (:367-369, :111-118)
                sh.aM[0] = sh.aM[0] + log(u0) * 2.0 * kCbSqrtPi;
                for (int k = 0; k < N; k++) A.alphaM[size_t(c) * N + k] = sh.aM[k];
                ug[0] = 1.0 * u0;
                for (int i = 1; i < N; i++) ug[i] = (sh.u[i] * (1.0 - r)) * u0;
                for (int k = 0; k < N; k++) aPg[k] = sh.aP[k];
                for (int k = 0; k < N * N; k++) Bg[k] = sh.P2M[k];
                if (A.outs) { A.outs[c].iter = total; A.outs[c].r = r; A.outs[c].norm_g = sh.err; }
                if (A.stats) {                              // c_optStats (momopt_update.cpp:254-273)
                    atomicMax((unsigned long long *)&A.stats[0], (unsigned long long)total);
                    atomicAdd((unsigned long long *)&A.stats[1], (unsigned long long)total);
                    atomicAdd((unsigned long long *)&A.stats[2], 1ull);
                    atomicAdd((unsigned long long *)&A.stats[3 + reg], 1ull);
                }
            }
        }
        __syncthreads();
    }
}

}  // namespace

bool launch_optcb(const CbArgs &a, int blocks, hipStream_t s) {
    switch (a.nm) {
    case 1: hipLaunchKernelGGL(optcb_kernel<1>, dim3(blocks), dim3(64), 0, s, a); return true;
    case 3: hipLaunchKernelGGL(optcb_kernel<3>, dim3(blocks), dim3(64), 0, s, a); return true;
    case 6: hipLaunchKernelGGL(optcb_kernel<6>, dim3(blocks), dim3(64), 0, s, a); return true;
    case 10: hipLaunchKernelGGL(optcb_kernel<10>, dim3(blocks), dim3(64), 0, s, a); return true;
    case 15: hipLaunchKernelGGL(optcb_kernel<15>, dim3(blocks), dim3(64), 0, s, a); return true;
    default: return false;
    }
}

}  // namespace mnc

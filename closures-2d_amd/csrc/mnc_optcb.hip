// mnc_optcb.hip -- the change-of-basis optimiser (src/momopt/opt/optcb.cpp:90-415,
// MOMOPT_OPTIMIZATION_TYPE 1) for batches of Mn closures on gfx950.
//
// One 64-lane block per cell at a time, the blocks taking cells from a queue
// (a counter in device memory) so that the few slow cells (near-boundary
// lobes, many iterations) do not leave a block's statically assigned share
// late.  The cell's basis P2M (n x n lower triangular,
// column-major as the reference's BLAS calls see it), its multipliers alphaP
// in that basis, and the quadrature table seen through the basis, pP[q][k] =
// (P2M^{-1} p_q)_k, live in HBM (pP in a workspace of nq*n doubles per block).
// The quadrature work -- the table's triangular solves, f, g and the Hessian
// sums, the line search's f -- is spread over the lanes (lane l takes the
// points q = l, l + 64, ...; wave sums by butterfly).  The n x n algebra
// (Cholesky, condition estimate, the basis products and solves) runs in the
// reference's order for each element, the independent operations of a step
// side by side across the lanes, on LDS copies or registers, between block
// barriers.
//
// The block is one wave and the kernel latency-bound, so waves per SIMD count
// more than anything inside one wave: the kernel is compiled for three (168
// VGPRs), the grid sized to the occupancy the runtime reports.  Two things
// keep the register count there without spills in the iteration: the lane
// index passes through an empty asm at the top of each iteration (the
// compiler would otherwise hoist dozens of lane masks and lane-dependent
// constants out of the loop and keep them live across it), and the table
// solve re-reads the factor from LDS for each point instead of holding all of
// it in registers.
#include "mnc_internal.h"
#include "mnc_coop_la.h"

#include <cfloat>

namespace mnc {
namespace {

__constant__ double kCbReg[5] = {0.0, 1e-3, 1e-2, 1e-1, 5e-1};   // opt.h:39-40
#define kCbSqrtPi sqrt(M_PI)

enum : int { CB_STEP = 0, CB_RESET = 1, CB_REG = 2, CB_FAIL = 3 };

template <int N>
struct CbShared {
    double P2M[N * N];                 // column-major (element (i,j) at [i + j*N])
    double L[N * N];                   // row-major: H, then its factor in the lower triangle
    double hB[N * N], hI[N * N];       // BFGS's H and H^-1 in the current basis (column-major)
    double gOld[N], sD[N], yD[N], tmp[N];
    double u[N], uP[N], aP[N], aP0[N], gP[N], dP[N], gM[N], aold[N], aM[N];
    double x[N];
    double e[64];                      // a 64-point chunk's exponentials (cb_fobj_full)
    double f, fscale, rtol, fnew, t, err;
    int action, done, converged;
    int p2min_id;                      // the restart basis: 1 the identity, 0 the cell's input basis
};

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

// ---- n x n kernels (lane 0), the Netlib reference BLAS / LAPACK loop orders ----
// element (i,j) of an n x n matrix, column-major or (RM) row-major
template <int N, bool RM>
__device__ __forceinline__ int cb_ix(int i, int j) { return RM ? i * N + j : i + j * N; }
// dtrsv('L','N','N') on a column-major A (RM: row-major, lower triangle read
// only): x <- A^{-1} x
template <int N, bool RM = false>
__device__ void cb_trsv_ln(const double *a, double *x) {
    for (int j = 0; j < N; j++) {
        if (x[j] != 0.0) {
            x[j] = x[j] / a[cb_ix<N, RM>(j, j)];
            const double temp = x[j];
            for (int i = j + 1; i < N; i++) x[i] = x[i] - temp * a[cb_ix<N, RM>(i, j)];
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
// dtrmv('L','T','N'): x <- A^T x (RM as cb_trsv_ln)
template <int N, bool RM = false>
__device__ void cb_trmv_lt(const double *a, double *x) {
    for (int j = 0; j < N; j++) {
        double temp = x[j] * a[cb_ix<N, RM>(j, j)];
        for (int i = j + 1; i < N; i++) temp = temp + a[cb_ix<N, RM>(i, j)] * x[i];
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
// The same four triangular kernels shared by the wave, lane i owning x_i (the
// vector in LDS, read at entry and written at exit); each element sees
// Netlib's operations in Netlib's order, only the independent ones of a step
// run side by side.  Row-major (rm) factors are the Hessian's L; column-major
// (cm) ones the basis P2M.
// dtrsv('L','N','N'), row-major: column j divides x_j (if nonzero), then x_i -= x_j l_ij
template <int N>
__device__ void cb_par_trsv_ln_rm(const double *l, double *x, int lane) {
    double xi = lane < N ? x[lane] : 0.0;
    for (int j = 0; j < N; j++) {
        double xj = readlane_d(xi, j);
        if (xj != 0.0) {
            xj = xj / l[j * N + j];
            if (lane == j) xi = xj;
            else if (lane > j && lane < N) xi = xi - xj * l[lane * N + j];
        }
    }
    __syncthreads();
    if (lane < N) x[lane] = xi;
    __syncthreads();
}
// dtrmv('L','T','N'), row-major: x_j <- x_j l_jj + sum_{i > j} l_ij x_i (old x)
template <int N>
__device__ void cb_par_trmv_lt_rm(const double *l, double *x, int lane) {
    const double xo = lane < N ? x[lane] : 0.0;
    double temp = lane < N ? xo * l[lane * N + lane] : 0.0;
    for (int i = 1; i < N; i++) {
        const double xi = readlane_d(xo, i);
        if (lane < i) temp = temp + l[i * N + lane] * xi;
    }
    __syncthreads();
    if (lane < N) x[lane] = temp;
    __syncthreads();
}
// dtrmv('L','N','N'), column-major: for j = n-1..0 with x_j != 0, x_i += x_j a_ij
// (i > j) and x_j *= a_jj; x_j is still the input when column j is applied
template <int N>
__device__ void cb_par_trmv_ln_cm(const double *a, double *x, int lane) {
    const double xo = lane < N ? x[lane] : 0.0;
    double xi = xo;
    for (int j = N - 1; j >= 0; j--) {
        const double xj = readlane_d(xo, j);
        if (xj != 0.0) {
            if (lane == j) xi = xi * a[j + j * N];
            else if (lane > j && lane < N) xi = xi + xj * a[lane + j * N];
        }
    }
    __syncthreads();
    if (lane < N) x[lane] = xi;
    __syncthreads();
}
// dtrsv('L','T','N'), column-major: for i = n-1..0, x_i = temp_i / a_ii, then
// temp_j -= a_ij x_i for j < i (the dot products' order)
template <int N>
__device__ void cb_par_trsv_lt_cm(const double *a, double *x, int lane) {
    double tj = lane < N ? x[lane] : 0.0;
    for (int i = N - 1; i >= 0; i--) {
        const double xi = readlane_d(tj, i) / a[i + i * N];
        if (lane == i) tj = xi;
        else if (lane < i) tj = tj - a[i + lane * N] * xi;
    }
    __syncthreads();
    if (lane < N) x[lane] = tj;
    __syncthreads();
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
        __asm__ volatile("" ::: "memory");          // the factor's reads stay in the loop
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

typedef double cb_f64x4 __attribute__((ext_vector_type(4)));

// f, g and H in the current basis (fobj.cpp:37-129,255-287 on pP).  H is an
// FP64 matrix-core sum, 16 x 16 over 4 points per instruction: lane l supplies
// point q0 + 4s + l/16, component l%16 (A = w p_c e, B = p_c).  Each point's
// exponential (the sequential dot over its row) is formed once, by lane
// q - q0 of its 64-point chunk, and read from LDS by the sixteen lanes that
// use it.  Result element r of lane l is row 4r + l/16, column l%16; the lower
// triangle is stored and mirrored into L.  Lane 0 forms g and f from the first
// column.
template <int N>
__device__ void cb_fobj_full(const CbArgs &A, const double *pP, CbShared<N> &sh, int lane) {
    static_assert(N <= 16, "one 16 x 16 tile");
    cb_f64x4 acc0 = {0.0, 0.0, 0.0, 0.0}, acc1 = {0.0, 0.0, 0.0, 0.0};
    const int c = lane & 15, k = lane >> 4;
    for (int q0 = 0; q0 < A.nq; q0 += 64) {
        {
            const int q = q0 + lane;
            double e = 0.0;
            if (q < A.nq) {
                const double *row = pP + size_t(q) * N;
                double temp = 0.0;
#pragma unroll
                for (int i = 0; i < N; i++) temp = temp + sh.aP[i] * row[i];
                e = exp(temp);
            }
            sh.e[lane] = e;
        }
        __syncthreads();
#pragma unroll 4
        for (int st = 0; st < 16; st++) {
            const int q = q0 + 4 * st + k;
            double av = 0.0, bv = 0.0;
            if (q < A.nq && c < N) {
                bv = pP[size_t(q) * N + c];
                av = A.w[q] * bv * sh.e[4 * st + k];
            }
            if ((st & 1) == 0) acc0 = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, acc0, 0, 0, 0);
            else acc1 = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, acc1, 0, 0, 0);
        }
        __syncthreads();
    }
#pragma unroll
    for (int r = 0; r < 4; r++) {
        const int i = 4 * r + k;
        if (i < N && c <= i) {
            const double v = acc0[r] + acc1[r];
            sh.L[i * N + c] = v;
            sh.L[c * N + i] = v;
        }
    }
    __syncthreads();
    if (lane == 0) {
        const double p0 = pP[0];
        double f = sh.L[0] / (p0 * p0), fscale = fabs(f);
        for (int i = 0; i < N; i++) {
            sh.gP[i] = sh.L[i * N] / p0 - sh.uP[i];
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

// f and g alone at alphaP (fobj.cpp:291-321 on pP), and f's resolution for the
// robust line search (the mass f + alpha.u plus sum |alpha_i u_i|); lane 0 stores
template <int N>
__device__ void cb_fobj_g(const CbArgs &A, const double *pP, CbShared<N> &sh, int lane) {
    double part[N + 1];
#pragma unroll
    for (int k = 0; k <= N; k++) part[k] = 0.0;
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
        for (int i = 0; i < N; i++) part[i] += wq * pq[i] * temp;
        part[N] += wq * temp;
    }
#pragma unroll
    for (int k = 0; k <= N; k++) part[k] = wave_sum(part[k]);
    if (lane == 0) {
        double f = part[N], s = 0.0, au = 0.0;
        for (int i = 0; i < N; i++) {
            sh.gP[i] = part[i] - sh.uP[i];
            f = f - sh.aP[i] * sh.uP[i];
            s += fabs(sh.aP[i] * sh.uP[i]);
            au += sh.aP[i] * sh.uP[i];
        }
        sh.f = f;
        sh.fscale = s + fabs(f + au);
    }
}

// cond (optbfgs.cpp:517-545): ||.|| over h[i*n+j] rows (j inner) of H and H^-1,
// lane i a row each, the largest by a butterfly max (NaN sums drop out, as the
// reference's '>' test); every lane ends with the product
template <int N>
__device__ double cb_bfgs_cond(const CbShared<N> &sh, int lane) {
    double n1 = 0.0, n2 = 0.0;
    if (lane < N) {
        for (int j = 0; j < N; j++) n1 += fabs(sh.hB[lane * N + j]);
        for (int j = 0; j < N; j++) n2 += fabs(sh.hI[lane * N + j]);
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        n1 = fmax(n1, __shfl_xor(n1, o));
        n2 = fmax(n2, __shfl_xor(n2, o));
    }
    return n1 * n2;
}

// Netlib dger's element update for row i: a(i,j) += x_i (alpha y_j), y_j != 0
template <int N>
__device__ __forceinline__ void cb_ger_row(double *a, double alpha, double xi, const double *y, int i) {
    for (int j = 0; j < N; j++)
        if (y[j] != 0.0) a[i + j * N] = a[i + j * N] + xi * (alpha * y[j]);
}
// Netlib dgemv('N'), beta = 0, row i: sum_j (alpha x_j) a(i,j) in column order
template <int N>
__device__ __forceinline__ double cb_gemv_row(const double *a, double alpha, const double *x, int i) {
    double y = 0.0;
    for (int j = 0; j < N; j++) y = y + (alpha * x[j]) * a[i + j * N];
    return y;
}
template <int N>
__device__ __forceinline__ double cb_dot(const double *x, const double *y) {
    double s = 0.0;
    for (int i = 0; i < N; i++) s = s + x[i] * y[i];
    return s;
}

// The rank-two updates of H and H^-1 (optbfgs.cpp:440-477), lane i owning row
// i; each element takes the reference's sequence of dger updates.  sD and yD
// hold the step and the gradient change.
template <int N>
__device__ void cb_bfgs_update(CbShared<N> &sh, int lane) {
    const bool own = lane < N;
    double dot = cb_dot<N>(sh.sD, sh.yD);
    double coef = 1.0 / dot;
    if (own) cb_ger_row<N>(sh.hB, coef, sh.yD[lane], sh.yD, lane);     // H += y y^T / (s.y)
    __syncthreads();
    if (own) sh.tmp[lane] = cb_gemv_row<N>(sh.hB, 1.0, sh.sD, lane);   // H s, from the updated H
    __syncthreads();
    dot = cb_dot<N>(sh.sD, sh.tmp);
    coef = -1.0 / dot;
    if (own) cb_ger_row<N>(sh.hB, coef, sh.tmp[lane], sh.tmp, lane);   // H -= (H s)(H s)^T / (s.H s)
    double t2 = 0.0;
    if (own) t2 = cb_gemv_row<N>(sh.hI, 1.0, sh.yD, lane);             // H^-1 y
    __syncthreads();
    if (own) sh.tmp[lane] = t2;
    __syncthreads();
    coef = cb_dot<N>(sh.yD, sh.tmp);
    dot = cb_dot<N>(sh.sD, sh.yD);
    coef = (coef + dot) / (dot * dot);
    if (own) {
        cb_ger_row<N>(sh.hI, coef, sh.sD[lane], sh.sD, lane);          // + coef s s^T
        coef = -1.0 / dot;
        cb_ger_row<N>(sh.hI, coef, sh.tmp[lane], sh.sD, lane);         // - (H^-1 y) s^T / (s.y)
        cb_ger_row<N>(sh.hI, coef, sh.sD[lane], sh.tmp, lane);         // - s (H^-1 y)^T / (s.y)
    }
    __syncthreads();
}

template <int N>
__device__ void cb_isotropic(double *a) {     // isotropicAlpha, Mn (opt.cpp:27-35)
    a[0] = 2.0 * kCbSqrtPi * log(0.5 / kCbSqrtPi);
    for (int i = 1; i < N; i++) a[i] = 0.0;
}

constexpr int kCbWavesPerSimd = 3;

template <int N>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(kCbWavesPerSimd))) void optcb_kernel(CbArgs A) {
    __shared__ CbShared<N> sh;
    __shared__ int next_cell;
    int lane = threadIdx.x;
    for (;;) {
        if (lane == 0) next_cell = atomicAdd(A.queue, 1);
        __syncthreads();
        const int c = next_cell;
        __syncthreads();
        if (c >= A.ncells) break;
        // the table seen through the basis: in an HBM workspace per block (an
        // LDS copy, 61 KB at n = 15 and 512 points, would leave room for two
        // cells per CU instead of twelve, and measured slower)
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
        for (int k = lane; k < N * N; k += 64) sh.P2M[k] = Bg[k];
        if (lane < N) {
            const double uk = ug[lane] / u0, ak = aPg[lane];
            sh.u[lane] = uk;
            sh.uP[lane] = uk;
            sh.aP[lane] = ak;
            sh.aP0[lane] = ak;
        }
        __syncthreads();
        if (lane == 0) {
            sh.p2min_id = 0;
            cb_trsv_ln<N>(sh.P2M, sh.uP);                   // :198-200
            sh.rtol = cb_nrm2(N, sh.u) * A.tol;
            sh.done = 0;
            sh.converged = 0;
        }
        __syncthreads();
        cb_table_reset<N>(A, pP, Bg, lane);                 // :185-187
        __syncthreads();
        int iter = 0, total = 0, maxit = A.max_iter, reg = 0, hitmax = 0;
        double t = 1.0, r = 0.0;
        for (;;) {
            ++iter;
            ++total;
            __asm__ volatile("" : "+v"(lane));              // no lane-derived values live across iterations
            cb_fobj_full<N>(A, pP, sh, lane);               // :234-242
            __syncthreads();
            // :247-259: the factor, the 1-norm and the condition estimate, the
            // wave sharing each step's independent updates
            // the factor with the rows in registers (lane i: row i), then back
            // to LDS for the basis update and the table solve
            double row[N];
            double norm = 0.0, dg = 0.0;                    // utils_norm1: row sums in order, their max
#pragma unroll
            for (int j = 0; j < N; j++) {
                row[j] = lane < N ? sh.L[lane * N + j] : 0.0;
                norm += fabs(row[j]);
                dg = (j == lane) ? row[j] : dg;
            }
            const double gersh = row16_min(lane < N ? 2.0 * dg - norm : DBL_MAX);
            norm = row16_max(lane < N ? norm : 0.0);
            const bool chol_ok = reg_cholesky<N>(row, lane);
            __syncthreads();
            if (lane < N)
                for (int j = 0; j < N; j++) sh.L[lane * N + j] = row[j];
            __syncthreads();
            double condest = 0.0;
            // the estimate only where the bounds cannot decide the test (as in
            // the Newton tail)
            if (chol_ok && !cond_proven_below(norm, gersh, N, A.cond_h_max)) {
                double rd[N];
#pragma unroll
                for (int k = 0; k < N; k++) rd[k] = 1.0 / readlane_d(row[k], k);
                if (!(2.0 * norm * reg_inv_norm1<N>(row, rd, lane) <= A.cond_h_max)) {
                    double col[N];
#pragma unroll
                    for (int k = 0; k < N; k++) col[k] = (lane < N && k > lane) ? sh.L[k * N + lane] : 0.0;
                    condest = 1.0 / reg_rcond<N>(row, col, norm, lane);
                }
            }
            if (lane == 0) {
                if (total == 1 && (!chol_ok || condest > A.cond_h_max)) sh.action = CB_RESET;
                else if (iter >= maxit || !chol_ok || t == 0.0 || condest > A.cond_h_max) sh.action = CB_REG;
                else sh.action = CB_STEP;
            }
            __syncthreads();
            const int action = sh.action;
            if (action == CB_RESET) {                       // :264-281
                if (lane == 0) {
                    cb_isotropic<N>(sh.aP0);
                    cb_identity<N>(sh.P2M);
                    sh.p2min_id = 1;
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
                        sh.p2min_id = 1;
                        for (int k = 0; k < N; k++) sh.uP[k] = sh.u[k];
                    }
                    reg = 0;
                    hitmax = 1;
                    maxit = A.max_iter;
                }
                if (lane == 0) {
                    sh.rtol = cb_nrm2(N, sh.uP) * A.tol;
                    for (int k = 0; k < N; k++) sh.aP[k] = sh.aP0[k];
                    if (sh.p2min_id) cb_identity<N>(sh.P2M);
                    else
                        for (int k = 0; k < N * N; k++) sh.P2M[k] = Bg[k];
                    cb_trsv_ln<N>(sh.P2M, sh.uP);
                }
                __syncthreads();
                cb_table_reset<N>(A, pP, sh.p2min_id ? nullptr : Bg, lane);
                __syncthreads();
                iter = 0;
                t = 1.0;
                if (reg == 4) maxit = 10 * maxit;
                continue;
            }
            // CB_STEP: the basis moves by L (:333-343).  P2M <- P2M L with
            // dtrmm('R','L','N','N')'s order for each element: element e of
            // the column-major result (row e % n, column e / n) on lane e % 64,
            // formed from the old P2M, stored after the barrier.
            constexpr int kEl = (N * N + 63) / 64;
            double bel[kEl];
#pragma unroll
            for (int r = 0; r < kEl; r++) {
                const int e = lane + 64 * r, i = e % N, j = e / N;
                double b = 0.0;
                if (e < N * N) {
                    b = sh.L[j * N + j] * sh.P2M[i + j * N];
                    for (int k = j + 1; k < N; k++) {
                        const double lkj = sh.L[k * N + j];
                        if (lkj != 0.0) b = b + lkj * sh.P2M[i + k * N];
                    }
                }
                bel[r] = b;
            }
            __syncthreads();
#pragma unroll
            for (int r = 0; r < kEl; r++)
                if (lane + 64 * r < N * N) sh.P2M[lane + 64 * r] = bel[r];
            __syncthreads();
            cb_par_trsv_ln_rm<N>(sh.L, sh.gP, lane);
            cb_par_trmv_lt_rm<N>(sh.L, sh.aP, lane);
            cb_par_trsv_ln_rm<N>(sh.L, sh.uP, lane);
            cb_table_solve<N>(A, pP, sh.L, lane);
            if (A.bfgs)                                     // optbfgs.cpp:371-373: H = H^-1 = I
                for (int k = lane; k < N * N; k += 64) sh.hB[k] = sh.hI[k] = (k % N == k / N) ? 1.0 : 0.0;
            __syncthreads();
            // optcb takes one step along -g_P (:347-388); optbfgs takes steps
            // along -H^-1 g_P while cond(H) cond(H^-1) and the count allow
            // (optbfgs.cpp:376-478)
            for (int ib = 0;; ++ib) {
                if (A.bfgs) {
                    const double cnd = cb_bfgs_cond<N>(sh, lane);
                    if (!(cnd < A.cond_h_max_bfgs && ib < A.max_bfgs_iter)) break;
                    double d = 0.0;
                    if (lane < N) d = cb_gemv_row<N>(sh.hI, -1.0, sh.gP, lane);
                    if (lane < N) sh.dP[lane] = d;
                } else if (lane < N) {
                    sh.dP[lane] = -sh.gP[lane];
                }
                __syncthreads();
                if (lane < N) sh.gM[lane] = sh.gP[lane];       // :352-354
                __syncthreads();
                cb_par_trmv_ln_cm<N>(sh.P2M, sh.gM, lane);
                if (lane == 0) {
                    sh.err = cb_nrm2(N, sh.gM);
                    sh.converged = sh.err <= sh.rtol;
                }
                __syncthreads();
                if (sh.converged) {                         // :358-375
                    if (lane < N) sh.aM[lane] = sh.aP[lane];
                    __syncthreads();
                    cb_par_trsv_lt_cm<N>(sh.P2M, sh.aM, lane);
                }
                __syncthreads();
                if (sh.converged) break;
                if (A.bfgs && lane < N) sh.gOld[lane] = sh.gP[lane];
                // line search in the current basis (linesearch.cpp:21-99 on pP)
                double maxf = 0.0;
                for (int i = 0; i < N; i++) {
                    const double tmp = fabs(sh.dP[i]) / (fabs(sh.aP[i]) + 1e-100);
                    if (tmp > maxf) maxf = tmp;
                }
                double gd = 0.0;
                for (int i = 0; i < N; i++) gd += sh.gP[i] * sh.dP[i];
                __syncthreads();
                if (lane < N) sh.aold[lane] = sh.aP[lane];
                __syncthreads();
                double tt = 1.0;
                bool accepted = false;
                while (tt * maxf > DBL_EPSILON / 2) {
                    if (lane < N) sh.aP[lane] = sh.aold[lane] + tt * sh.dP[lane];
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
                    if (lane < N) sh.aP[lane] = sh.aold[lane];
                    tt = 0.0;
                }
                t = tt;
                __syncthreads();
                if (!A.bfgs) break;
                cb_fobj_g<N>(A, pP, sh, lane);              // optbfgs.cpp:429-436
                __syncthreads();
                if (lane < N) {                             // :440-445
                    sh.yD[lane] = sh.gP[lane] - sh.gOld[lane];
                    sh.sD[lane] = sh.aP[lane] - sh.aold[lane];
                }
                __syncthreads();
                cb_bfgs_update<N>(sh, lane);
            }
            if (sh.converged) break;
        }
        if (sh.converged) {                                 // outputs (:367-369, :111-118)
            for (int k = lane; k < N * N; k += 64) Bg[k] = sh.P2M[k];
            if (lane < N) {
                A.alphaM[size_t(c) * N + lane] = lane == 0 ? sh.aM[0] + log(u0) * 2.0 * kCbSqrtPi : sh.aM[lane];
                ug[lane] = lane == 0 ? 1.0 * u0 : (sh.u[lane] * (1.0 - r)) * u0;
                aPg[lane] = sh.aP[lane];
            }
            if (lane == 0) {
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

// Blocks: the resident count (the runtime's occupancy for this instance, a
// block per wave slot) times the CUs, capped by max_blocks (the workspace's
// size) and the cell count; blocks beyond the resident ones would only wait.
template <int N>
int optcb_blocks_n(int max_blocks, int num_cu, int ncells) {
    static int per_cu = 0;
    if (per_cu == 0) {
        int nb = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, optcb_kernel<N>, 64, 0) != hipSuccess || nb < 1)
            nb = 4 * kCbWavesPerSimd;
        per_cu = nb;
    }
    const int b = per_cu * num_cu;
    return b < max_blocks ? (b < ncells ? b : ncells) : (max_blocks < ncells ? max_blocks : ncells);
}
template <int N>
int launch_optcb_n(const CbArgs &a, int max_blocks, int num_cu, hipStream_t s) {
    const int blocks = optcb_blocks_n<N>(max_blocks, num_cu, a.ncells);
    if (blocks > 0) hipLaunchKernelGGL(optcb_kernel<N>, dim3(blocks), dim3(64), 0, s, a);
    return blocks;
}

int launch_optcb(const CbArgs &a, int max_blocks, int num_cu, hipStream_t s) {
    switch (a.nm) {
    case 1: return launch_optcb_n<1>(a, max_blocks, num_cu, s);
    case 3: return launch_optcb_n<3>(a, max_blocks, num_cu, s);
    case 6: return launch_optcb_n<6>(a, max_blocks, num_cu, s);
    case 10: return launch_optcb_n<10>(a, max_blocks, num_cu, s);
    case 15: return launch_optcb_n<15>(a, max_blocks, num_cu, s);
    default: return -1;
    }
}

}  // namespace mnc

// mnc_newton.hip -- batched Newton solve of the Mn dual problem on gfx950.
//
// Lockstep rounds with compaction: every round is one quadrature sweep for all
// active cells (sweep_kernel: one lane per cell, wave-uniform table rows as
// scalar loads; or sweep_mix_kernel: sixteen cells per wave on the matrix
// cores) followed by one step of the per-cell Newton/Armijo state machine
// (logic_kernel), which compacts the cells still iterating onto the next list.
// One sweep yields both a trial's f and the full Hessian at the trial point, so
// an accepted trial goes straight to its Newton step with no second sweep (the
// reference re-evaluates at the same alpha, giving the same numbers).  After the
// rounds, a persistent tail (tail_coop_kernel: one wave per remaining cell)
// takes the few slow cells near the realizability boundary from a queue and
// iterates each to completion.
//
// The per-cell algorithm restates src/momopt/opt/opt.cpp:82-343 (opt,
// optScaled), src/momopt/opt/linesearch.cpp:21-99 and fobj.cpp:37-129,255-335,
// with the LAPACK roles (dpotrf 'L', dpocon, dpotrs, dnrm2) done in registers
// or LDS: unblocked column Cholesky, Higham's 1-norm estimator (ACM TOMS 14(4),
// 1988), forward/back substitution.
#include <algorithm>
#include <cfloat>
#include <mutex>
#include <type_traits>

#include "mnc_internal.h"
#include "mnc_coop_la.h"
#include "mnc_gaunt_pattern.h"

namespace mnc {
namespace {

constexpr int kBlock = 256;
#define kSqrtPi sqrt(M_PI)
#define kPi M_PI
__constant__ double kReg[5] = {0.0, 1e-3, 1e-2, 1e-1, 5e-1};   // opt.h:39-40

template <int N> struct Tri {
    static constexpr int size = N * (N + 1) / 2;
    __device__ static constexpr int at(int i, int j) { return i * (i + 1) / 2 + j; }   // i >= j
};

// ---- dense kernels on a packed lower triangle in registers ----------------

// The dense kernels are generic in the element type T: FP64 everywhere except
// the mixed-precision logic, whose Hessian is only ~1e-5 accurate and is
// factored in FP32 (the gradient, f and the stop test stay FP64).

// dpotf2-order Cholesky in place; returns false on a non-positive (or NaN) pivot.
template <int N, class T = double>
__device__ __forceinline__ bool cholesky(T (&a)[Tri<N>::size]) {
    bool ok = true;
#pragma unroll
    for (int j = 0; j < N; ++j) {
        T dot = T(0);
#pragma unroll
        for (int k = 0; k < j; ++k) dot = fma(a[Tri<N>::at(j, k)], a[Tri<N>::at(j, k)], dot);
        T ajj = a[Tri<N>::at(j, j)] - dot;
        ok = ok && (ajj > T(0));
        ajj = sqrt(ajj > T(0) ? ajj : T(1));
        a[Tri<N>::at(j, j)] = ajj;
#pragma unroll
        for (int i = j + 1; i < N; ++i) {
            T y = a[Tri<N>::at(i, j)];
#pragma unroll
            for (int k = 0; k < j; ++k) y = fma(-a[Tri<N>::at(j, k)], a[Tri<N>::at(i, k)], y);
            a[Tri<N>::at(i, j)] = y;
        }
        const T r = T(1) / ajj;
#pragma unroll
        for (int i = j + 1; i < N; ++i) a[Tri<N>::at(i, j)] *= r;
    }
    return ok;
}

// x <- (L L^T)^{-1} x
template <int N, class T = double>
__device__ __forceinline__ void apply_inv(const T (&l)[Tri<N>::size], T (&x)[N]) {
#pragma unroll
    for (int j = 0; j < N; ++j) {
        x[j] = x[j] / l[Tri<N>::at(j, j)];
#pragma unroll
        for (int i = j + 1; i < N; ++i) x[i] = fma(-x[j], l[Tri<N>::at(i, j)], x[i]);
    }
#pragma unroll
    for (int j = N - 1; j >= 0; --j) {
        T t = x[j];
#pragma unroll
        for (int i = N - 1; i > j; --i) t = fma(-l[Tri<N>::at(i, j)], x[i], t);
        x[j] = t / l[Tri<N>::at(j, j)];
    }
}

// apply_inv with the reciprocals of the diagonal precomputed (rinv): the
// condition estimate's repeated solves, where a division per step would sit on
// the dependency chain; the estimate moves only at rounding level.
template <int N, class T = double>
__device__ __forceinline__ void apply_inv_rcp(const T (&l)[Tri<N>::size], const T (&rinv)[N], T (&x)[N]) {
#pragma unroll
    for (int j = 0; j < N; ++j) {
        x[j] = x[j] * rinv[j];
#pragma unroll
        for (int i = j + 1; i < N; ++i) x[i] = fma(-x[j], l[Tri<N>::at(i, j)], x[i]);
    }
#pragma unroll
    for (int j = N - 1; j >= 0; --j) {
        T t = x[j];
#pragma unroll
        for (int i = N - 1; i > j; --i) t = fma(-l[Tri<N>::at(i, j)], x[i], t);
        x[j] = t * rinv[j];
    }
}

template <int N, class T>
__device__ __forceinline__ T dasum(const T (&x)[N]) {
    T s = T(0);
#pragma unroll
    for (int i = 0; i < N; ++i) s += fabs(x[i]);
    return s;
}

template <int N, class T>
__device__ __forceinline__ int idamax(const T (&x)[N]) {
    int idx = 0;
    T m = fabs(x[0]);
#pragma unroll
    for (int i = 1; i < N; ++i) {
        const T v = fabs(x[i]);
        if (v > m) { m = v; idx = i; }
    }
    return idx;
}

// Reciprocal condition estimate of A = L L^T in the 1-norm (the role of
// dpocon): Higham's estimate of ||A^{-1}||_1, with a single call site of the
// A^{-1} product inside a small state machine to keep the unrolled code short.
template <int N, class T = double>
__device__ __forceinline__ double rcond(const T (&l)[Tri<N>::size], double anorm) {
    if (anorm == 0.0) return 0.0;
    T x[N];
    unsigned isgn = 0;   // bit i set <=> isgn_i = -1
    T est = T(0);
    int jump = 1, j = 0, iter = 0;
    T rinv[N];
#pragma unroll
    for (int i = 0; i < N; ++i) {
        x[i] = T(1) / T(N);
        rinv[i] = T(1) / l[Tri<N>::at(i, i)];
    }
#pragma unroll 1
    for (int guard = 0; guard < 16; ++guard) {
        apply_inv_rcp<N, T>(l, rinv, x);
        if (jump == 1) {                       // first product
            if (N == 1) { est = fabs(x[0]); break; }
            est = dasum<N, T>(x);
            isgn = 0;
#pragma unroll
            for (int i = 0; i < N; ++i) {
                const bool neg = !(x[i] >= T(0));
                x[i] = neg ? T(-1) : T(1);
                isgn |= neg ? (1u << i) : 0u;
            }
            jump = 2;
        } else if (jump == 2) {                // after the transposed product
            j = idamax<N, T>(x);
            iter = 2;
#pragma unroll
            for (int i = 0; i < N; ++i) x[i] = (i == j) ? T(1) : T(0);
            jump = 3;
        } else if (jump == 3) {                // x = A^{-1} e_j
            const T estold = est;
            est = dasum<N, T>(x);
            unsigned s = 0;
#pragma unroll
            for (int i = 0; i < N; ++i) s |= !(x[i] >= T(0)) ? (1u << i) : 0u;
            bool done = (s == isgn) || (est <= estold);
            if (!done) {
#pragma unroll
                for (int i = 0; i < N; ++i) x[i] = (s >> i & 1u) ? T(-1) : T(1);
                isgn = s;
                jump = 4;
            } else {
                T alt = T(1);
#pragma unroll
                for (int i = 0; i < N; ++i) { x[i] = alt * (T(1) + T(i) / T(N - 1)); alt = -alt; }
                jump = 5;
            }
        } else if (jump == 4) {                // after the transposed product
            const int jlast = j;
            j = idamax<N, T>(x);
            T xjl = T(0), xj = T(0);
#pragma unroll
            for (int i = 0; i < N; ++i) { xjl = (i == jlast) ? x[i] : xjl; xj = (i == j) ? x[i] : xj; }
            if (xjl != fabs(xj) && iter < 5) {
                ++iter;
#pragma unroll
                for (int i = 0; i < N; ++i) x[i] = (i == j) ? T(1) : T(0);
                jump = 3;
            } else {
                T alt = T(1);
#pragma unroll
                for (int i = 0; i < N; ++i) { x[i] = alt * (T(1) + T(i) / T(N - 1)); alt = -alt; }
                jump = 5;
            }
        } else {                               // jump == 5: final alternating vector
            const T temp = T(2) * (dasum<N, T>(x) / T(3 * N));
            if (temp > est) est = temp;
            break;
        }
    }
    if (!isfinite(est) || est == T(0)) return 0.0;
    return (1.0 / double(est)) / anorm;
}

template <int N>
__device__ __forceinline__ double nrm2(const double (&x)[N]) {
    double s = 0.0;
#pragma unroll
    for (int i = 0; i < N; ++i) s = fma(x[i], x[i], s);
    return sqrt(s);
}

__device__ __forceinline__ unsigned long long lane_mask_lt() {
    const int lane = threadIdx.x & 63;
    return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

enum : int { ST_NEWTON = 0, ST_TRIAL = 1 };
constexpr int kHalvings = 2;   // line-search halvings evaluated per wave sweep in the tail (8 measured
                               // 2-3% slower: the unused candidates' exponentials and registers)
constexpr int kTailCap = 4;    // quadrature points per lane whose w e the tail keeps between passes
constexpr int kTailPts = 2;    // points per lane whose chains interleave (2 of the 4 at a time:
                               // with 4, the rows and the partial sums outgrow the VGPRs and the
                               // sums move through AGPRs, 4% slower on the M4 lattice)
constexpr int kTailWaves = 4;  // waves (cells in flight) per block of the cooperative tail
constexpr int kFirstHalvings = 2;   // candidates in the first batch of halvings after a failed trial
enum : int { LOGIC_CONTINUE = 0, LOGIC_DONE = 1, LOGIC_DEFER = 2, LOGIC_MLP = 3 };
constexpr int kMlpList = kMlpListIndex;   // cells waiting for the learned warm start's network
static_assert(kMlpListWord == kMlpListIndex + 2, "list word of cw()");

// The table is read-only for the whole kernel; reading it through the constant
// address space lets wave-uniform row reads become scalar loads (SGPR operands).
typedef const __attribute__((address_space(4))) double cdouble;

// One quadrature sweep: acc += sum_q w_q e_q (outer products of rows [R0,R1) of
// the table, lower triangle) or, for the Gaunt form, acc[k] += sum_q w_q e_q p_qk.
// Returns sum_q w_q e_q (the line-search objective's integral, fobj.cpp:297-317).
template <int NM, int NP2, bool GAUNT, int R0, int R1, int NA>
__device__ __forceinline__ double sweep(const double *Pg, const double *Wg, int nq, const double (&al)[NM],
                                        double (&acc)[NA]) {
    cdouble *P = (cdouble *)Pg;
    cdouble *W = (cdouble *)Wg;
    double sf = 0.0;
#pragma unroll
    for (int k = 0; k < NA; ++k) acc[k] = 0.0;
#pragma unroll 2
    for (int q = 0; q < nq; ++q) {
        cdouble *row = P + size_t(q) * NP2;
        double am = 0.0;
#pragma unroll
        for (int k = 0; k < NM; ++k) am = fma(al[k], row[k], am);
        const double we = W[q] * exp(am);
        sf += we;
        if constexpr (GAUNT) {
#pragma unroll
            for (int k = 0; k < NP2; ++k) acc[k] = fma(we, row[k], acc[k]);
        } else {
            constexpr int off = Tri<NM>::at(R0, 0);
#pragma unroll
            for (int i = R0; i < R1; ++i) {
                const double v = we * row[i];
#pragma unroll
                for (int j = 0; j <= i; ++j) acc[Tri<NM>::at(i, j) - off] = fma(v, row[j], acc[Tri<NM>::at(i, j) - off]);
            }
        }
    }
    return sf;
}

// The same sweep shared by a whole wave for one cell: lane l takes the
// quadrature points q = l, l+64, ..., and the partial sums are reduced across
// the wave (butterfly), so every lane ends with the full sums.
// The PPn ansatz at t = (alpha.m)/2 (fobj.cpp:166-172): G = t + s with
// s = sqrt(t^2 + delta), taken as -delta / (t - s) for t < 0.
__device__ __forceinline__ double ppn_g(double t, double delta, double &s) {
    s = sqrt(t * t + delta);
    return t < 0.0 ? -delta / (t - s) : t + s;
}

// The PPn sweep (fobj.cpp:137-236,355-428): f = sum w (t G - delta/2 + delta log G)
// is returned, the Hessian sums use the weight w G / (2 s) and, with NG = NM,
// gs collects the gradient sums sum w G m_k (G's weight differs from H's, so g
// is not H's first column as it is for Mn).
template <int NM, int NP2, bool GAUNT, int R0, int R1, int NA, int NG>
__device__ __forceinline__ double sweep_ppn(const double *Pg, const double *Wg, int nq, const double (&al)[NM],
                                            double delta, double (&acc)[NA], double (&gs)[NG > 0 ? NG : 1]) {
    cdouble *P = (cdouble *)Pg;
    cdouble *W = (cdouble *)Wg;
    double sf = 0.0;
#pragma unroll
    for (int k = 0; k < NA; ++k) acc[k] = 0.0;
#pragma unroll
    for (int k = 0; k < NG; ++k) gs[k] = 0.0;
    for (int q = 0; q < nq; ++q) {
        cdouble *row = P + size_t(q) * NP2;
        double am = 0.0;
#pragma unroll
        for (int k = 0; k < NM; ++k) am = fma(al[k], row[k], am);
        const double t = 0.5 * am;
        double sq;
        const double G = ppn_g(t, delta, sq);
        sf = sf + W[q] * (t * G - 0.5 * delta + delta * log(G));
        const double wg = W[q] * G, wh = 0.5 * W[q] * (G / sq);
#pragma unroll
        for (int k = 0; k < NG; ++k) gs[k] = fma(row[k], wg, gs[k]);
        if constexpr (GAUNT) {
#pragma unroll
            for (int k = 0; k < NP2; ++k) acc[k] = fma(row[k], wh, acc[k]);
        } else {
            constexpr int off = Tri<NM>::at(R0, 0);
#pragma unroll
            for (int i = R0; i < R1; ++i) {
                const double v = row[i] * wh;
#pragma unroll
                for (int j = 0; j <= i; ++j) acc[Tri<NM>::at(i, j) - off] = fma(v, row[j], acc[Tri<NM>::at(i, j) - off]);
            }
        }
    }
    return sf;
}

template <int NM, int NP2, bool GAUNT, int R0, int R1, int NA>
__device__ __forceinline__ double sweep_wave(const double *P, const double *W, int nq, const double (&al)[NM],
                                             double (&acc)[NA]) {
    const int lane = threadIdx.x & 63;
    double sf = 0.0;
#pragma unroll
    for (int k = 0; k < NA; ++k) acc[k] = 0.0;
    for (int q = lane; q < nq; q += 64) {
        const double *row = P + size_t(q) * NP2;
        double am = 0.0;
#pragma unroll
        for (int k = 0; k < NM; ++k) am = fma(al[k], row[k], am);
        const double we = W[q] * exp(am);
        sf += we;
        if constexpr (GAUNT) {
#pragma unroll
            for (int k = 0; k < NP2; ++k) acc[k] = fma(we, row[k], acc[k]);
        } else {
            constexpr int off = Tri<NM>::at(R0, 0);
#pragma unroll
            for (int i = R0; i < R1; ++i) {
                const double v = we * row[i];
#pragma unroll
                for (int j = 0; j <= i; ++j) acc[Tri<NM>::at(i, j) - off] = fma(v, row[j], acc[Tri<NM>::at(i, j) - off]);
            }
        }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        sf += __shfl_xor(sf, off);
#pragma unroll
        for (int k = 0; k < NA; ++k) acc[k] += __shfl_xor(acc[k], off);
    }
    return sf;
}

// ---- per-cell solver state in HBM ------------------------------------------
// Structure of arrays [field][cap]: threads handling neighbouring list entries
// (mostly ascending cell indices) touch neighbouring words.
// Moment order N of n = (N+1)(N+2)/2 moments, and the order-2N count padded to
// the 16 columns of an MFMA tile (the mixed sweep's Hessian moments).
constexpr int order_of(int nm) { int N = 0; while ((N + 1) * (N + 2) / 2 < nm) ++N; return N; }
constexpr int mix_cols(int nm) { return ((2 * order_of(nm) + 1) * (2 * order_of(nm) + 2) / 2 + 15) / 16 * 16; }
constexpr int cmax(int a, int b) { return a > b ? a : b; }

template <int NM, int NP2>
struct Store {
    static constexpr int nacc = cmax(cmax(NP2, NM * (NM + 1) / 2), mix_cols(NM));
    enum : int {
        F_AL = 0, F_U = NM, F_AOLD = 2 * NM, F_D = 3 * NM, F_T = 4 * NM, F_RTOL, F_U0, F_ASCALE, F_FCUR,
        F_GD, F_MAXF, F_FSCALE, F_DELTA, F_SF, F_GS, F_ACC = F_GS + NM
    };
    static constexpr int dfields = F_ACC + nacc;
    enum : int { I_ITER = 0, I_TOTAL, I_REG, I_MAXIT, I_MODE, I_WARM, ifields };
    double *D;
    int *I;
    size_t cap;
    __device__ double &d(int f, int c) const { return D[size_t(f) * cap + c]; }
    __device__ int &i(int f, int c) const { return I[size_t(f) * cap + c]; }
};

// Per-lane statistics, reduced per block and flushed with one atomic each.
// One lane's counts within one launch (32-bit: a lane counts far fewer than
// 2^31 events in a launch; the block sums are 64-bit).
struct LaneStats {
    int max_iter = 0, err_bits = 0;
    int sum = 0, n = 0, trials = 0, hist[5] = {0, 0, 0, 0, 0}, warm[3] = {0, 0, 0}, deferred = 0;
    // persistent tail only: cells taken, Newton iterations and Armijo trials done in it
    int tcells = 0, titer = 0, ttrials = 0;
    int fallback = 0;   // learned starts abandoned for isotropicAlpha (net_fallback)
};

// Counting into the arrays with a run-time index would place the whole
// LaneStats in scratch memory; these select with compares instead.
__device__ __forceinline__ void count_warm(LaneStats &s, int k) {
#pragma unroll
    for (int i = 0; i < 3; ++i) s.warm[i] += (k == i);
}
__device__ __forceinline__ void count_reg(LaneStats &s, int r) {
#pragma unroll
    for (int i = 0; i < 5; ++i) s.hist[i] += (r == i);
}

template <int WAVES = kBlock / 64>
__device__ void flush_stats(const NewtonArgs &A, LaneStats &s) {
    __shared__ long long red[WAVES][18];
    for (int off = 32; off > 0; off >>= 1) {
        s.max_iter = max(s.max_iter, __shfl_xor(s.max_iter, off));
        s.err_bits |= __shfl_xor(s.err_bits, off);
        s.sum += __shfl_xor(s.sum, off);
        s.n += __shfl_xor(s.n, off);
        s.trials += __shfl_xor(s.trials, off);
#pragma unroll
        for (int r = 0; r < 5; ++r) s.hist[r] += __shfl_xor(s.hist[r], off);
#pragma unroll
        for (int r = 0; r < 3; ++r) s.warm[r] += __shfl_xor(s.warm[r], off);
        s.deferred += __shfl_xor(s.deferred, off);
        s.tcells += __shfl_xor(s.tcells, off);
        s.titer += __shfl_xor(s.titer, off);
        s.ttrials += __shfl_xor(s.ttrials, off);
        s.fallback += __shfl_xor(s.fallback, off);
    }
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (lane == 0) {
        red[w][0] = s.max_iter; red[w][1] = s.err_bits; red[w][2] = s.sum; red[w][3] = s.n; red[w][4] = s.trials;
        for (int r = 0; r < 5; ++r) red[w][5 + r] = s.hist[r];
        for (int r = 0; r < 3; ++r) red[w][10 + r] = s.warm[r];
        red[w][13] = s.deferred;
        red[w][14] = s.tcells;
        red[w][15] = s.titer;
        red[w][16] = s.ttrials;
        red[w][17] = s.fallback;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        long long v[18] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
        for (int ww = 0; ww < WAVES; ++ww) {
            v[0] = max(v[0], red[ww][0]);
            v[1] |= red[ww][1];
            for (int k = 2; k < 18; ++k) v[k] += red[ww][k];
        }
        if (v[3] || v[4] || v[1] || v[10] || v[11] || v[12] || v[13] || v[14] || v[17]) {
            atomicMax((unsigned long long *)&A.stats[0], (unsigned long long)v[0]);
            atomicAdd((unsigned long long *)&A.stats[1], (unsigned long long)v[2]);
            atomicAdd((unsigned long long *)&A.stats[2], (unsigned long long)v[3]);
            for (int r = 0; r < 5; ++r) atomicAdd((unsigned long long *)&A.stats[3 + r], (unsigned long long)v[5 + r]);
            atomicAdd((unsigned long long *)&A.stats[8], (unsigned long long)v[4]);
            for (int r = 0; r < 3; ++r) atomicAdd((unsigned long long *)&A.stats[9 + r], (unsigned long long)v[10 + r]);
            atomicAdd((unsigned long long *)&A.stats[12], (unsigned long long)v[13]);
            if (v[14]) {
                atomicAdd((unsigned long long *)&A.stats[13], (unsigned long long)v[14]);
                atomicAdd((unsigned long long *)&A.stats[14], (unsigned long long)v[15]);
                atomicAdd((unsigned long long *)&A.stats[15], (unsigned long long)v[16]);
            }
            if (v[17]) atomicAdd((unsigned long long *)&A.stats[16], (unsigned long long)v[17]);
            if (v[1]) atomicOr(&A.ctl[3], int(v[1]));
        }
    }
}

// Squared FP64 gradient norm of the scaled dual problem at al:
// || sum_q w_q e^{al.m_q} m_q - us ||^2 (fobj.cpp:291-334 without f).
template <int NM, int NP2>
__device__ double grad_norm2(const NewtonArgs &A, const double (&al)[NM], const double (&us)[NM]) {
    cdouble *P = (cdouble *)A.p;
    cdouble *W = (cdouble *)A.w;
    double g[NM];
#pragma unroll
    for (int k = 0; k < NM; ++k) g[k] = 0.0;
#pragma unroll 2
    for (int q = 0; q < A.nq; ++q) {
        cdouble *row = P + size_t(q) * NP2;
        double am = 0.0;
#pragma unroll
        for (int k = 0; k < NM; ++k) am = fma(al[k], row[k], am);
        const double we = W[q] * exp(am);
#pragma unroll
        for (int k = 0; k < NM; ++k) g[k] = fma(we, row[k], g[k]);
    }
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < NM; ++k) {
        const double d = g[k] - us[k];
        s = fma(d, d, s);
    }
    return s;
}

// The supplementary stop test (opt.cpp:52-76,288-293): with tol_gamma > 0 a
// cell whose gradient passes the tolerance stops only when
// gamma = exp(2 |d|_1 m_inf) (Mn) or 1 + 2 |d|_1 m_inf (PPn) is <= tol_gamma,
// d the Newton direction and m_inf = sqrt((2 P + 1) / (2 pi)) with
// P = floor(sqrt(n)) + 1 (computed on the host).  da1 is |d|_1 summed in
// index order.
__device__ __forceinline__ bool gamma_ok(const NewtonArgs &A, double da1, bool ppn) {
    const double x = 2.0 * da1 * A.gamma_minf;
    return (ppn ? 1.0 + x : exp(x)) <= A.tol_gamma;
}

// Control word holding the count of list l: lists 0 and 1 (the lockstep
// ping-pong pair) use words 0 and 1, lists 2 and 3 (the FP64 pair of a
// mixed-precision solve) words 4 and 5, list 4 (the learned warm start's
// network rows) word 6.  Word 2 is the tail's queue head, word 3 the error bits.
__device__ __forceinline__ int cw(int l) { return l < 2 ? l : l + 2; }

// Appends the calling lanes with `want` set to list `dst`, whose count is
// control word `cw` (wave-aggregated).
__device__ __forceinline__ void list_append(const NewtonArgs &A, int dst, int cw, bool want, int c) {
    const unsigned long long m = __ballot(want);
    if (!m) return;
    const int lane = threadIdx.x & 63;
    const int leader = __ffsll((long long)m) - 1;
    int base = 0;
    if (lane == leader) base = atomicAdd(&A.ctl[cw], __popcll(m));
    base = __shfl(base, leader);
    if (want) A.lists[size_t(dst) * A.cap + base + __popcll(m & lane_mask_lt())] = c;
}

enum : int { INIT_INVALID = 0, INIT_READY = 1 };

// opt() prologue (opt.cpp:87-101) and optScaled set-up (:170-205) for cell c.
// Returns INIT_INVALID if u[0] <= 0 (the reference aborts: opt.cpp:97-101).
// MNC_WARM_LEARNED starts from the alpha passed in (isotropicAlpha if it is not
// finite; I_WARM records which); the first round's logic then decides whether
// the network is consulted (cell_logic, divert).  With the network switched off
// (no divert at the prologue) the start is counted here.
template <int NM, int NP2>
__device__ int cell_init(const NewtonArgs &A, const Store<NM, NP2> &S, int c, LaneStats &st) {
    using St = Store<NM, NP2>;
    const double *uin = A.u + size_t(c) * NM;
    const double *ain = A.alpha + size_t(c) * NM;
    const double u0 = uin[0];
    const double ascale = log(u0) * 2.0 * kSqrtPi;
    double ss = 0.0;
#pragma unroll
    for (int k = 0; k < NM; ++k) {
        const double uk = uin[k] / u0;
        S.d(St::F_U + k, c) = uk;
        ss = fma(uk, uk, ss);
    }
    double al[NM];
    int warm = 2;
    bool iso = A.warm == MNC_WARM_ISOTROPIC;
    if (A.warm == MNC_WARM_LEARNED) {
        bool finite = true;
#pragma unroll
        for (int k = 0; k < NM; ++k) finite = finite && isfinite(ain[k]);
        iso = !finite;
        warm = finite ? 2 : 1;
    }
    // PPn (opt.cpp:113-141): alpha / u0, u / u0 and delta / u0^2, no shift
    const double delta = A.ppn ? A.delta / (u0 * u0) : 0.0;
    if (iso) {                                                     // isotropicAlpha (opt.cpp:27-46)
        al[0] = A.ppn ? 1.0 - 4.0 * delta * kPi : 2.0 * kSqrtPi * log(0.5 / kSqrtPi);
#pragma unroll
        for (int k = 1; k < NM; ++k) al[k] = 0.0;
    } else if (A.ppn) {
#pragma unroll
        for (int k = 0; k < NM; ++k) al[k] = ain[k] / u0;
    } else {
#pragma unroll
        for (int k = 0; k < NM; ++k) al[k] = ain[k];
        al[0] = al[0] - ascale;
    }
#pragma unroll
    for (int k = 0; k < NM; ++k) S.d(St::F_AL + k, c) = al[k];
    if (!(u0 > 0.0)) {
        st.err_bits |= 1 << MNC_ERR_U0;
        if (A.outs) { A.outs[c].iter = -1; A.outs[c].r = 0.0; A.outs[c].norm_g = 0.0; }
        return INIT_INVALID;
    }
    S.d(St::F_U0, c) = u0;
    S.d(St::F_ASCALE, c) = ascale;
    S.d(St::F_DELTA, c) = delta;
    S.d(St::F_RTOL, c) = sqrt(ss) * A.tol;
    S.d(St::F_T, c) = 1.0;
    S.i(St::I_ITER, c) = 0;
    S.i(St::I_TOTAL, c) = 0;
    S.i(St::I_REG, c) = 0;
    S.i(St::I_MAXIT, c) = A.max_iter;
    S.i(St::I_MODE, c) = ST_NEWTON;
    S.i(St::I_WARM, c) = warm;
    if (A.warm == MNC_WARM_LEARNED && !A.divert) count_warm(st, warm);
    return INIT_READY;
}

// The learned warm start's fallback: a cell started from the network's guess
// (I_WARM 0) that meets a regularization criterion at level 0 -- the iteration
// limit, a failed factorisation, an exhausted line search, a condition estimate
// above cond_h_max -- within A.fallback_iters Newton iterations restarts once
// from isotropicAlpha at level 0 instead of climbing the ladder: the guess, not
// the problem, may be what failed, and the start from isotropicAlpha is the
// reference's own (momopt_update.cpp:236-238 with an isotropic previous alpha).
// A late failure is the problem's (cells beyond what the quadrature resolves
// fail at level 0 from any start) and climbs the ladder.  I_WARM becomes
// kWarmFellBack, so a second failure climbs the ladder as the reference does.
// Only Mn, where the network exists.
constexpr int kWarmFellBack = 3;
__device__ __forceinline__ bool net_fallback(const NewtonArgs &A, int reg, int warm, int iter) {
    return reg == 0 && warm == 0 && iter <= A.fallback_iters;
}

// One step of optScaled's loop for cell c after a sweep at its evaluation
// point (the sweep's results are in S: F_SF and F_ACC).  Mirrors
// opt.cpp:209-318 and linesearch.cpp:38-98, with the sweep shared between an
// accepted Armijo trial and the Newton iteration that follows it.
//
// MIX: the sweep was the mixed-precision one: f from the FP64 sum of w e, g from
// the FP64 sums of w e m (F_GS), and H from order-2N moments accumulated on the
// matrix cores in bf16 hi/lo pairs with FP32 sums (relative error ~1e-5).  The
// stop test, the Armijo test and the reconstruction are FP64 as in the
// reference; only the Newton direction uses the approximate H.  Whenever the
// reference would act on H itself (a failed factorisation, a condition estimate
// above min(cmix, cond_h_max), an exhausted line search, the iteration limit),
// the iteration is not taken: the cell returns LOGIC_DEFER at the point where
// that Newton iteration starts, to be redone by an FP64 sweep.
template <int NM, int NP2, bool GAUNT, bool MIX = false, bool PPN = false>
__device__ int cell_logic(const NewtonArgs &A, const Store<NM, NP2> &S, int c, LaneStats &st,
                          float *gstage = nullptr) {
    using St = Store<NM, NP2>;
    const double p0 = A.p[0];
    int mode = S.i(St::I_MODE, c), iter = S.i(St::I_ITER, c), total = S.i(St::I_TOTAL, c);
    int reg = S.i(St::I_REG, c), maxit = S.i(St::I_MAXIT, c);
    double t = S.d(St::F_T, c);
    double al[NM];
#pragma unroll
    for (int k = 0; k < NM; ++k) al[k] = S.d(St::F_AL + k, c);
    bool newton = (mode == ST_NEWTON);
    bool do_reg = false, converged = false, defer = false, to_mlp = false;
    double err = 0.0;
    if (mode == ST_TRIAL) {                                        // linesearch.cpp:45-70
        ++st.trials;
        double fnew = S.d(St::F_SF, c);
#pragma unroll
        for (int k = 0; k < NM; ++k) fnew = fnew - al[k] * S.d(St::F_U + k, c);
        const double fcur = S.d(St::F_FCUR, c), gd = S.d(St::F_GD, c);
        bool accept = fnew <= fcur + 1e-3 * t * gd;
        // Robust mode: a full Newton step whose predicted decrease -g.d lies
        // below the rounding resolution of f cannot be judged by Armijo; in
        // that (quadratic) regime the step is taken.
        if (!accept && A.ls_mode == MNC_LS_ROBUST && t == 1.0 &&
            -gd <= 1e3 * DBL_EPSILON * S.d(St::F_FSCALE, c))
            accept = true;
        if (accept) {
            newton = true;
        } else {
            t = 0.5 * t;
            if (t * S.d(St::F_MAXF, c) > DBL_EPSILON / 2) {
#pragma unroll
                for (int k = 0; k < NM; ++k) al[k] = fma(t, S.d(St::F_D + k, c), S.d(St::F_AOLD + k, c));
            } else {                                               // exhausted: t = 0 (:74-98)
#pragma unroll
                for (int k = 0; k < NM; ++k) al[k] = S.d(St::F_AOLD + k, c);
                if constexpr (MIX) {        // redo the iteration that produced d in FP64
                    t = 1.0;
                    --iter; --total;
                    defer = true;
                } else {
                    t = 0.0;
                    ++iter; ++total;
                    do_reg = true;                                 // opt.cpp:244 (t == 0)
                }
            }
        }
    }
    double g0[NM];   // the gradient at convergence (reconstruction)
    if (newton) {
        ++iter; ++total;                                           // opt.cpp:212-213
#pragma unroll
        for (int k = 0; k < NM; ++k) S.d(St::F_AL + k, c) = al[k];
        using HT = typename std::conditional<MIX, float, double>::type;
        HT h[Tri<NM>::size];
        if constexpr (MIX) {                                       // order-2N moments -> H
            // The moments are staged in this thread's LDS column (stride
            // kBlock), so the Gaunt terms, whose indices are the same for every
            // cell (scalar loads), read LDS instead of the store.
            constexpr int NMM = mix_cols(NM);
#pragma unroll
            for (int k = 0; k < NMM; ++k) gstage[k * kBlock] = float(S.d(St::F_ACC + k, c));
            if (A.mpv != nullptr) {
                // the map's pattern fixed at compile time: every index folds,
                // the coefficients come in as batched scalar loads; the same
                // products in the same order as the CSR walk below
                constexpr GauntPattern<order_of(NM)> P = gaunt_pattern<order_of(NM)>();
#pragma unroll
                for (int k = 0; k < Tri<NM>::size; ++k) h[k] = 0.0f;
#pragma unroll
                for (int e = 0; e < P.count; ++e) {
                    const int hk = Tri<NM>::at(P.j[e], P.i[e]);
                    h[hk] = fmaf(A.mpv[e], gstage[P.k[e] * kBlock], h[hk]);
                }
            } else {
#pragma unroll
                for (int i = 0; i < NM; ++i) {
#pragma unroll
                    for (int j = i; j < NM; ++j) {
                        float hij = 0.0f;
                        const int pr = i * NM + j;
                        for (int cc = A.mgptr[pr]; cc < A.mgptr[pr + 1]; ++cc)
                            hij = fmaf(float(A.mgv[cc]), gstage[A.mgk[cc] * kBlock], hij);
                        h[Tri<NM>::at(j, i)] = HT(hij);
                    }
                }
            }
        } else if constexpr (GAUNT) {                              // fobj.cpp:99-113
#pragma unroll
            for (int i = 0; i < NM; ++i) {
#pragma unroll
                for (int j = i; j < NM; ++j) {
                    double hij = 0.0;
                    const int pr = i * NM + j;
                    for (int cc = A.gptr[pr]; cc < A.gptr[pr + 1]; ++cc)
                        hij = fma(A.gv[cc], S.d(St::F_ACC + A.gk[cc], c), hij);
                    h[Tri<NM>::at(j, i)] = hij;
                }
            }
        } else {
#pragma unroll
            for (int k = 0; k < Tri<NM>::size; ++k) h[k] = S.d(St::F_ACC + k, c);
        }
        // Phase 1: g, f and the 1-norm from H (fobj.cpp:270-286, utils.cpp:36-52).
        // g goes to the gradient slot of the state (F_D doubles as it until the
        // direction is formed), so only the factor stays live below.
        {
            double fcur = (MIX || PPN) ? S.d(St::F_SF, c) : h[0] / (p0 * p0);
            double fscale = fabs(fcur), gg = 0.0;
#pragma unroll
            for (int i = 0; i < NM; ++i) {
                const double ui = S.d(St::F_U + i, c);
                const double gi = ((MIX || PPN) ? S.d(St::F_GS + i, c) : h[Tri<NM>::at(i, 0)] / p0) - ui;
                S.d(St::F_D + i, c) = gi;
                gg = fma(gi, gi, gg);
                fcur = fcur - al[i] * ui;
                fscale += fabs(al[i] * ui);
            }
            S.d(St::F_FCUR, c) = fcur;
            S.d(St::F_FSCALE, c) = fscale;
            err = sqrt(gg);                                        // dnrm2 (:286)
        }
        bool diverted = false;
        if (A.divert) {
            // The learned warm start's gate at the first evaluation: the start
            // (the alpha passed in, or isotropicAlpha if that was not finite)
            // is kept while ||g|| <= gate * ||u||.  Otherwise the better of it
            // and isotropicAlpha -- whose gradient norm is sum_{k>0} u_k^2
            // without a sweep: there the integrals are (1, 0, ...) up to
            // rounding -- is parked (norm^2 in F_FCUR, label in I_WARM) for
            // the network's guess to be judged against (judge_kernel), and
            // the evaluation does not count as an iteration.
            const int warm = S.i(St::I_WARM, c);
            double iso2 = 0.0;
#pragma unroll
            for (int k = 1; k < NM; ++k) {
                const double uk = S.d(St::F_U + k, c);
                iso2 = fma(uk, uk, iso2);
            }
            const double g2 = err * err;
            if (g2 > A.gate * A.gate * (1.0 + iso2)) {
                int pick = warm;
                double best = g2;
                if (warm == 2 && iso2 < g2) {
                    pick = 1;
                    best = iso2;
                    al[0] = 2.0 * kSqrtPi * log(0.5 / kSqrtPi);
#pragma unroll
                    for (int k = 1; k < NM; ++k) al[k] = 0.0;
                }
                S.d(St::F_FCUR, c) = best;
                S.i(St::I_WARM, c) = pick;
                --iter;
                --total;
                diverted = true;
                to_mlp = true;
            } else {
                count_warm(st, warm);
            }
        }
        if (!diverted) {
            double anorm = 0.0, gersh = DBL_MAX;
#pragma unroll
            for (int i = 0; i < NM; ++i) {
                double s = 0.0;
#pragma unroll
                for (int j = 0; j < NM; ++j) s += fabs(h[i >= j ? Tri<NM>::at(i, j) : Tri<NM>::at(j, i)]);
                anorm = s > anorm ? s : anorm;
                gersh = fmin(gersh, 2.0 * double(h[Tri<NM>::at(i, i)]) - s);
            }
            // Phase 2: factor, condition estimate, direction (dpotrf, dpocon, dpotrs).
            const double climit = MIX ? fmin(A.cmix, A.cond_h_max) : A.cond_h_max;
            const bool chol_ok = cholesky<NM, HT>(h);
            double condest = 0.0;
            if (chol_ok && !cond_proven_below(anorm, gersh, NM, climit)) condest = 1.0 / rcond<NM, HT>(h, anorm);
            if (iter >= maxit || !chol_ok || t == 0.0 || condest > climit) {
                if constexpr (MIX) {
                    --iter; --total;
                    defer = true;
                } else {
                    do_reg = true;                                     // opt.cpp:244
                }
            } else {
                double d[NM];
                {
                    HT dh[NM];
#pragma unroll
                    for (int i = 0; i < NM; ++i) dh[i] = HT(S.d(St::F_D + i, c));
                    apply_inv<NM, HT>(h, dh);                          // dpotrs (:278-282)
#pragma unroll
                    for (int i = 0; i < NM; ++i) d[i] = double(dh[i]);
                }
                // Phase 3: stop test, or the line search set-up (linesearch.cpp:28-43)
                bool stop = err <= S.d(St::F_RTOL, c);
                if (stop && A.tol_gamma > 0.0) {
                    double da1 = 0.0;
#pragma unroll
                    for (int i = 0; i < NM; ++i) da1 = da1 + fabs(d[i]);
                    stop = gamma_ok(A, da1, PPN);
                }
                if (stop) {
                    converged = true;                                  // :293-306
#pragma unroll
                    for (int i = 0; i < NM; ++i) g0[i] = S.d(St::F_D + i, c);
                } else {
                    double maxf = 0.0, gd = 0.0;
#pragma unroll
                    for (int i = 0; i < NM; ++i) {
                        const double ai = S.d(St::F_AL + i, c);
                        const double gi = S.d(St::F_D + i, c);
                        const double di = -d[i];
                        const double tmp = fabs(di) / (fabs(ai) + 1e-100);
                        maxf = tmp > maxf ? tmp : maxf;
                        gd = fma(gi, di, gd);
                        S.d(St::F_AOLD + i, c) = ai;
                        S.d(St::F_D + i, c) = di;
                        al[i] = ai;
                        d[i] = di;
                    }
                    S.d(St::F_MAXF, c) = maxf;
                    S.d(St::F_GD, c) = gd;
                    t = 1.0;
                    if (t * maxf > DBL_EPSILON / 2) {
#pragma unroll
                        for (int i = 0; i < NM; ++i) al[i] = al[i] + t * d[i];
                        mode = ST_TRIAL;
                    } else if constexpr (MIX) {
                        --iter; --total;
                        defer = true;
                    } else {
                        // the Armijo loop cannot start: linesearch returns 0 and the
                        // next iteration regularizes (opt.cpp:212-213,244)
                        t = 0.0;
                        ++iter; ++total;
                        do_reg = true;
                    }
                }
            }
            if (!converged && mode != ST_TRIAL) {
#pragma unroll
                for (int i = 0; i < NM; ++i) al[i] = S.d(St::F_AL + i, c);
            }
        }
    }
    int result = LOGIC_CONTINUE;
    if (to_mlp) {
        result = LOGIC_MLP;
    } else if (defer) {
        mode = ST_NEWTON;
        result = LOGIC_DEFER;
    } else if (do_reg && !PPN && net_fallback(A, reg, S.i(St::I_WARM, c), iter)) {
        // the network's guess failed: plain Newton from isotropicAlpha, still at
        // level 0 (u and the tolerance are those of level 0)
        S.i(St::I_WARM, c) = kWarmFellBack;
        ++st.fallback;
        al[0] = 2.0 * kSqrtPi * log(0.5 / kSqrtPi);
#pragma unroll
        for (int k = 1; k < NM; ++k) al[k] = 0.0;
        iter = 0;
        t = 1.0;
        mode = ST_NEWTON;
    } else if (do_reg) {                                           // opt.cpp:246-273
        if (reg < 4) {
            ++reg;
            const double r = kReg[reg];
            const double *uin = A.u + size_t(c) * NM;
            const double u0 = S.d(St::F_U0, c);
            double ss = 1.0;
            S.d(St::F_U + 0, c) = 1.0;
#pragma unroll
            for (int k = 1; k < NM; ++k) {
                const double uk = (uin[k] / u0) * (1.0 - r);
                S.d(St::F_U + k, c) = uk;
                ss = fma(uk, uk, ss);
            }
            // isotropicAlpha (opt.cpp:27-46)
            al[0] = PPN ? 1.0 - 4.0 * S.d(St::F_DELTA, c) * kPi : 2.0 * kSqrtPi * log(0.5 / kSqrtPi);
#pragma unroll
            for (int k = 1; k < NM; ++k) al[k] = 0.0;
            S.d(St::F_RTOL, c) = sqrt(ss) * A.tol;
            iter = 0;
            t = 1.0;
            if (reg == 4) maxit = 10 * maxit;
            mode = ST_NEWTON;
        } else {                                                   // "Max regularization achieved."
            st.err_bits |= 1 << MNC_ERR_MAXREG;
            if (A.outs) { A.outs[c].iter = total; A.outs[c].r = -1.0; A.outs[c].norm_g = err; }
            result = LOGIC_DONE;
        }
    } else if (converged) {
        const double u0 = S.d(St::F_U0, c);
#pragma unroll
        for (int k = 0; k < NM; ++k) al[k] = S.d(St::F_AL + k, c);
        double *uo = A.u + size_t(c) * NM;
        double *ao = A.alpha + size_t(c) * NM;
        if constexpr (PPN) {                                   // no reconstruction; opt.cpp:131-136
#pragma unroll
            for (int k = 0; k < NM; ++k) {
                ao[k] = al[k] * u0;
                uo[k] = S.d(St::F_U + k, c) * u0;
            }
        } else {
            const double u0p = S.d(St::F_U + 0, c) + g0[0];
#pragma unroll
            for (int k = 0; k < NM; ++k) uo[k] = ((S.d(St::F_U + k, c) + g0[k]) / u0p) * u0;   // :298-301, opt.cpp:107-108
            al[0] -= log(u0p) * 2.0 * kSqrtPi;                                               // :302
            al[0] = al[0] + S.d(St::F_ASCALE, c);                                            // opt.cpp:109
#pragma unroll
            for (int k = 0; k < NM; ++k) ao[k] = al[k];
        }
        if (A.outs) { A.outs[c].iter = total; A.outs[c].r = kReg[reg]; A.outs[c].norm_g = err; }
        st.max_iter = total > st.max_iter ? total : st.max_iter;
        st.sum += total;
        ++st.n;
        count_reg(st, reg);
        result = LOGIC_DONE;
    }
    if (result != LOGIC_DONE) {
#pragma unroll
        for (int k = 0; k < NM; ++k) S.d(St::F_AL + k, c) = al[k];
        S.d(St::F_T, c) = t;
        S.i(St::I_ITER, c) = iter;
        S.i(St::I_TOTAL, c) = total;
        S.i(St::I_REG, c) = reg;
        S.i(St::I_MAXIT, c) = maxit;
        S.i(St::I_MODE, c) = mode;
    }
    return result;
}

// One evaluation at the cell's current point: S.F_SF and S.F_ACC (and, for
// PPn, the gradient sums S.F_GS).
template <int NM, int NP2, bool GAUNT, int RS, bool PPN = false>
__device__ __forceinline__ void cell_sweep(const NewtonArgs &A, const Store<NM, NP2> &S, int c) {
    using St = Store<NM, NP2>;
    constexpr int NA = GAUNT ? NP2 : Tri<NM>::at(RS, 0);
    constexpr int NB = (GAUNT || RS >= NM) ? 1 : Tri<NM>::size - Tri<NM>::at(RS, 0);
    static_assert(NA <= St::nacc, "accumulators exceed the store");
    double al[NM];
#pragma unroll
    for (int k = 0; k < NM; ++k) al[k] = S.d(St::F_AL + k, c);
    if constexpr (PPN) {
        const double delta = S.d(St::F_DELTA, c);
        {
            double acc[NA], gs[NM];
            const double sf = sweep_ppn<NM, NP2, GAUNT, 0, (GAUNT ? NM : RS), NA, NM>(A.p, A.w, A.nq, al, delta,
                                                                                       acc, gs);
            S.d(St::F_SF, c) = sf;
#pragma unroll
            for (int k = 0; k < NM; ++k) S.d(St::F_GS + k, c) = gs[k];
#pragma unroll
            for (int k = 0; k < NA; ++k) S.d(St::F_ACC + k, c) = acc[k];
        }
        if constexpr (!GAUNT && RS < NM) {
            double accb[NB], none[1];
            sweep_ppn<NM, NP2, false, RS, NM, NB, 0>(A.p, A.w, A.nq, al, delta, accb, none);
#pragma unroll
            for (int k = 0; k < NB; ++k) S.d(St::F_ACC + NA + k, c) = accb[k];
        }
        return;
    }
    {
        double acc[NA];
        const double sf = sweep<NM, NP2, GAUNT, 0, (GAUNT ? NM : RS), NA>(A.p, A.w, A.nq, al, acc);
        S.d(St::F_SF, c) = sf;
#pragma unroll
        for (int k = 0; k < NA; ++k) S.d(St::F_ACC + k, c) = acc[k];
    }
    if constexpr (!GAUNT && RS < NM) {
        double accb[NB];
        sweep<NM, NP2, false, RS, NM, NB>(A.p, A.w, A.nq, al, accb);
#pragma unroll
        for (int k = 0; k < NB; ++k) S.d(St::F_ACC + NA + k, c) = accb[k];
    }
}

// ---- kernels ------------------------------------------------------------------

// Prologue of every cell; the cells go on list 0.
template <int NM, int NP2>
__global__ __launch_bounds__(kBlock) void init_kernel(NewtonArgs A) {
    const Store<NM, NP2> S{A.st_d, A.st_i, A.cap};
    const int c = blockIdx.x * kBlock + threadIdx.x;
    LaneStats st;
    int res = INIT_INVALID;
    if (c < A.ncells) res = cell_init<NM, NP2>(A, S, c, st);
    list_append(A, 0, 0, res == INIT_READY, c);
    flush_stats(A, st);
}

// The learned warm start's second half: for the cells the first round's
// logic parked on list kMlpList, the network's guess (alpha_guess, scaled
// problem) replaces the parked start if its FP64 gradient norm is smaller; the
// cells join list 1, the second round's input.
template <int NM, int NP2>
__global__ __launch_bounds__(kBlock) void judge_kernel(NewtonArgs A) {
    using St = Store<NM, NP2>;
    const St S{A.st_d, A.st_i, A.cap};
    const int slot = blockIdx.x * kBlock + threadIdx.x;
    LaneStats st;
    bool ok = false;
    int c = -1;
    if (slot < A.ctl[cw(kMlpList)]) {
        c = A.lists[size_t(kMlpList) * A.cap + slot];
        double us[NM], cand[NM];
        bool finite = true;
#pragma unroll
        for (int k = 0; k < NM; ++k) {
            us[k] = S.d(St::F_U + k, c);
            cand[k] = A.alpha_guess[size_t(c) * NM + k];
            finite = finite && isfinite(cand[k]);
        }
        int pick = S.i(St::I_WARM, c);
        if (finite && grad_norm2<NM, NP2>(A, cand, us) < S.d(St::F_FCUR, c)) {
            pick = 0;
            S.i(St::I_WARM, c) = 0;
#pragma unroll
            for (int k = 0; k < NM; ++k) S.d(St::F_AL + k, c) = cand[k];
        }
        count_warm(st, pick);
        ok = true;
    }
    list_append(A, 1, cw(1), ok, c);
    flush_stats(A, st);
}

// ---- the learned warm start in one kernel: network and judge ---------------
//
// For the cells the first round parked on list kMlpList: the MLP
// 15 -> 128 -> 128 -> 128 -> 15 (GELU) on the matrix cores, sixteen cells per
// wave (v_mfma_f32_16x16x32_bf16: bf16 operands, FP32 sums, bias and GELU in
// FP32, the arithmetic of the unfused GEMM kernels in mnc_mlp.hip, so the
// guesses are the same), then the judge: the gradient norm of the scaled
// problem at the guess, also on the matrix cores (v_mfma_f32_16x16x4_f32), which
// replaces the parked start when it is smaller; the cells join list 1.
//
// The judge per wave, sixteen cells c and 16-point tiles of the quadrature:
//   a[q][c] = sum_k P[q][k] alpha[c][k]               (MFMA, FP32)
//   S[k][c] += sum_q P[q][k] w_q exp(a[q][c] - m_c)   (MFMA, FP32)
// with m_c the running maximum of a over the points seen (the sums rescaled by
// exp(m_old - m_new) when it grows, so no exponential overflows), and
// g_k = exp(m_c) S[k][c] - u_k, ||g||^2 in FP64.  FP32 resolves ||g|| to about
// 1e-6 ||u||: enough to rank two starts (the judge only picks the start; the
// Newton solve that follows is the usual one).
//
// LDS: the weights in the bf16 operand layout (lane l of (k-step, tile) holds
// W[32 ks + 8 (l/16) + t][16 tile + l%16], t = 0..7), the quadrature as FP32
// (P[q][0..15] with row stride 17, w_q), and per wave one activation buffer
// (bf16), which the guesses (FP64) reuse after the last layer.
typedef __bf16 nbf16x8 __attribute__((ext_vector_type(8)));
typedef unsigned short nu16x8 __attribute__((ext_vector_type(8)));
typedef float nf32x4 __attribute__((ext_vector_type(4)));

constexpr int kNetH = 128, kNetPar = 37007;   // hidden width; the parameter count (mnc_mlp.hip)
constexpr int kNetOffW1 = 0, kNetOffB1 = 15 * 128, kNetOffW2 = kNetOffB1 + 128, kNetOffB2 = kNetOffW2 + 128 * 128,
              kNetOffW3 = kNetOffB2 + 128, kNetOffB3 = kNetOffW3 + 128 * 128, kNetOffW4 = kNetOffB3 + 128,
              kNetOffB4 = kNetOffW4 + 128 * 15;
static_assert(kNetOffB4 + 15 == kNetPar, "parameter layout");
constexpr int kNetWaves = 8;                  // waves per block (one block per CU: the weights fill LDS)
constexpr int kNetActStride = kNetH + 8;      // bf16 activation row (16-byte aligned rows, banks offset)
constexpr int kNetPStride = 17;               // FP32 quadrature row

__device__ __forceinline__ unsigned short net_bf16(float x) {   // round to nearest even, inf/NaN kept
    const unsigned u = __float_as_uint(x);
    if ((u & 0x7f800000u) == 0x7f800000u) return (unsigned short)(u >> 16);
    return (unsigned short)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}
__device__ __forceinline__ float net_gelu(float x) { return 0.5f * x * (1.0f + erff(x * 0.70710678118654752f)); }

struct NetShared {
    nu16x8 w1[8][64];          // k-step 0 (K = 15, padded to 32) x 8 output tiles
    nu16x8 w2[4][8][64];
    nu16x8 w3[4][8][64];
    nu16x8 w4[4][64];          // one output tile (15 outputs, padded to 16)
    float b[3][kNetH], b4[16], st[60];
    float pq[kNetMaxQ][kNetPStride];   // P[q][k] (k < 15; column 15 zero) in FP32
    float wq[kNetMaxQ];
    // per wave: the activations of its sixteen rows (bf16), then their guesses (FP64)
    union alignas(16) Act {
        unsigned short a[16][kNetActStride];
        double guess[16][16];
    } act[kNetWaves];
};
static_assert(sizeof(NetShared) <= 160 * 1024, "LDS");

// Weight element of the operand layout: W[k][j] of a layer with K inputs and N outputs.
__device__ __forceinline__ nu16x8 net_operand(const float *W, int K, int N, int ks, int tile, int lane) {
    nu16x8 v;
    const int j = 16 * tile + (lane & 15);
#pragma unroll
    for (int t = 0; t < 8; ++t) {
        const int k = 32 * ks + 8 * (lane >> 4) + t;
        v[t] = (k < K && j < N) ? net_bf16(W[size_t(k) * N + j]) : (unsigned short)0;
    }
    return v;
}

// One hidden layer for the wave's sixteen rows, in place: act = gelu(act W + b)
// (K = 128).  Every operand is read before the first result is written.
__device__ __forceinline__ void net_layer(const nu16x8 (&w)[4][8][64], const float *bias,
                                          unsigned short (*act)[kNetActStride], int lane) {
    nu16x8 a[4];
#pragma unroll
    for (int ks = 0; ks < 4; ++ks) a[ks] = *reinterpret_cast<const nu16x8 *>(&act[lane & 15][32 * ks + 8 * (lane >> 4)]);
    nf32x4 acc[8];
#pragma unroll
    for (int tile = 0; tile < 8; ++tile) {
        acc[tile] = nf32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int ks = 0; ks < 4; ++ks)
            acc[tile] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(nbf16x8, a[ks]),
                                                                __builtin_bit_cast(nbf16x8, w[ks][tile][lane]), acc[tile], 0, 0, 0);
    }
    __builtin_amdgcn_wave_barrier();
#pragma unroll
    for (int tile = 0; tile < 8; ++tile) {
        const int j = 16 * tile + (lane & 15);
#pragma unroll
        for (int r = 0; r < 4; ++r) act[4 * (lane >> 4) + r][j] = net_bf16(net_gelu(acc[tile][r] + bias[j]));
    }
}

template <int NM, int NP2>
__global__ __launch_bounds__(64 * kNetWaves) void net_judge_kernel(NewtonArgs A) {
    static_assert(NM == 15, "the network is M4's");
    using St = Store<NM, NP2>;
    const St S{A.st_d, A.st_i, A.cap};
    __shared__ NetShared sh;
    const float *W = A.mlp_w;
    for (int i = threadIdx.x; i < 8 * 64; i += blockDim.x) sh.w1[i / 64][i % 64] = net_operand(W + kNetOffW1, 15, kNetH, 0, i / 64, i % 64);
    for (int i = threadIdx.x; i < 4 * 8 * 64; i += blockDim.x) {
        const int ks = i / 512, tile = (i / 64) % 8, l = i % 64;
        sh.w2[ks][tile][l] = net_operand(W + kNetOffW2, kNetH, kNetH, ks, tile, l);
        sh.w3[ks][tile][l] = net_operand(W + kNetOffW3, kNetH, kNetH, ks, tile, l);
    }
    for (int i = threadIdx.x; i < 4 * 64; i += blockDim.x) sh.w4[i / 64][i % 64] = net_operand(W + kNetOffW4, kNetH, 15, i / 64, 0, i % 64);
    for (int i = threadIdx.x; i < kNetH; i += blockDim.x) {
        sh.b[0][i] = W[kNetOffB1 + i];
        sh.b[1][i] = W[kNetOffB2 + i];
        sh.b[2][i] = W[kNetOffB3 + i];
    }
    for (int i = threadIdx.x; i < 16; i += blockDim.x) sh.b4[i] = i < 15 ? W[kNetOffB4 + i] : 0.f;
    for (int i = threadIdx.x; i < 60; i += blockDim.x) sh.st[i] = W[kNetPar + i];
    const int nq = A.nq;                     // the host launches this kernel only for nq <= kNetMaxQ
    const int nt = (nq + 15) / 16;
    for (int i = threadIdx.x; i < 16 * nt * 16; i += blockDim.x) {
        const int q = i / 16, k = i % 16;
        sh.pq[q][k] = (q < nq && k < NM) ? float(A.p[size_t(q) * NP2 + k]) : 0.f;
        if (k == 0) sh.wq[q] = q < nq ? float(A.w[q]) : 0.f;
    }
    __syncthreads();     // the only block-wide barrier: from here on the waves are independent
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int row = lane & 15, part = lane >> 4;
    const int count = A.ctl[cw(kMlpList)];
    const int *list = A.lists + size_t(kMlpList) * A.cap;
    unsigned short (*act)[kNetActStride] = sh.act[wv].a;
    double (*guess)[16] = sh.act[wv].guess;
    LaneStats st;
    for (int base = (blockIdx.x * kNetWaves + wv) * 16; base < count; base += gridDim.x * kNetWaves * 16) {
        const int c = base + row < count ? list[base + row] : -1;
        // layer 1: the standardised scaled moments u/u0 (k < 15; lane group
        // `part` supplies k = 8 part + t)
        nu16x8 x;
#pragma unroll
        for (int t = 0; t < 8; ++t) {
            const int k = 8 * part + t;
            float v = 0.f;
            if (c >= 0 && k < 15) v = float((S.d(St::F_U + k, c) - double(sh.st[k])) / double(sh.st[15 + k]));
            x[t] = net_bf16(v);
        }
#pragma unroll
        for (int tile = 0; tile < 8; ++tile) {
            nf32x4 acc = {0.f, 0.f, 0.f, 0.f};
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(nbf16x8, x), __builtin_bit_cast(nbf16x8, sh.w1[tile][lane]),
                                                          acc, 0, 0, 0);
            const int j = 16 * tile + row;
#pragma unroll
            for (int r = 0; r < 4; ++r) act[4 * part + r][j] = net_bf16(net_gelu(acc[r] + sh.b[0][j]));
        }
        __builtin_amdgcn_wave_barrier();
        net_layer(sh.w2, sh.b[1], act, lane);
        __builtin_amdgcn_wave_barrier();
        net_layer(sh.w3, sh.b[2], act, lane);
        __builtin_amdgcn_wave_barrier();
        {
            nf32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int ks = 0; ks < 4; ++ks) {
                const nu16x8 a = *reinterpret_cast<const nu16x8 *>(&act[row][32 * ks + 8 * part]);
                acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(nbf16x8, a),
                                                              __builtin_bit_cast(nbf16x8, sh.w4[ks][lane]), acc, 0, 0, 0);
            }
            // the guesses overwrite the activations just read: keep the
            // compiler from moving a store above those loads (the hardware
            // keeps one wave's LDS accesses in order)
            __asm__ volatile("" ::: "memory");
            __builtin_amdgcn_wave_barrier();
            if (row < 15) {
#pragma unroll
                for (int r = 0; r < 4; ++r)   // the scaled multipliers: y out_std + out_mean
                    guess[4 * part + r][row] = double(acc[r] + sh.b4[row]) * double(sh.st[45 + row]) + double(sh.st[30 + row]);
            }
        }
        __asm__ volatile("" ::: "memory");
        __builtin_amdgcn_wave_barrier();
        // ---- the judge (this lane: cell `row`, moments part + 4 ks of its
        // operand, moments 4 part + r of its sums) ----
        float alf[4];
        bool finite = true;
#pragma unroll
        for (int ks = 0; ks < 4; ++ks) {
            const int k = part + 4 * ks;
            const double v = k < NM ? guess[row][k] : 0.0;
            finite = finite && isfinite(v);
            alf[ks] = float(v);
        }
        finite = finite && __shfl_xor(int(finite), 16) && __shfl_xor(int(finite), 32);
        if (!finite || c < 0) {
#pragma unroll
            for (int ks = 0; ks < 4; ++ks) alf[ks] = 0.f;
        }
        nf32x4 sum = {0.f, 0.f, 0.f, 0.f};
        float m = -FLT_MAX;
        for (int t = 0; t < nt; ++t) {
            nf32x4 a = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int ks = 0; ks < 4; ++ks)       // a[16t + 4 part + r][row]
                a = __builtin_amdgcn_mfma_f32_16x16x4f32(sh.pq[16 * t + row][part + 4 * ks], alf[ks], a, 0, 0, 0);
            const int q0 = 16 * t + 4 * part;
            float tm = -FLT_MAX;
#pragma unroll
            for (int r = 0; r < 4; ++r) tm = q0 + r < nq ? fmaxf(tm, a[r]) : tm;
            tm = fmaxf(tm, __shfl_xor(tm, 16));
            tm = fmaxf(tm, __shfl_xor(tm, 32));
            const float mn = fmaxf(m, tm);
            const float sc = __expf(m - mn);
            m = mn;
            float e[4];
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                sum[r] *= sc;
                e[r] = sh.wq[q0 + r] * __expf(a[r] - mn);   // w = 0 past nq
            }
#pragma unroll
            for (int ks = 0; ks < 4; ++ks)       // S[4 part + r][row] += P[q][.] w e over q = 16t + 4 (l/16) + ks
                sum = __builtin_amdgcn_mfma_f32_16x16x4f32(sh.pq[16 * t + 4 * part + ks][row], e[ks], sum, 0, 0, 0);
        }
        const double em = exp(double(m));
        double n2 = 0.0;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int k = 4 * part + r;
            if (c >= 0 && k < NM) {
                const double gk = em * double(sum[r]) - S.d(St::F_U + k, c);
                n2 = fma(gk, gk, n2);
            }
        }
        n2 += __shfl_xor(n2, 16);
        n2 += __shfl_xor(n2, 32);
        const bool take = c >= 0 && finite && n2 < S.d(St::F_FCUR, c);
        if (take) {
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int k = 4 * part + r;
                if (k < NM) S.d(St::F_AL + k, c) = guess[row][k];
            }
        }
        bool ok = false;
        if (c >= 0 && part == 0) {
            int pick = S.i(St::I_WARM, c);
            if (take) {
                pick = 0;
                S.i(St::I_WARM, c) = 0;
            }
            count_warm(st, pick);
            ok = true;
        }
        list_append(A, 1, cw(1), ok, c);
        __asm__ volatile("" ::: "memory");
        __builtin_amdgcn_wave_barrier();
    }
    flush_stats<kNetWaves>(A, st);
}

// One sweep for every cell on list `lin`; also clears the count of the list
// the following logic kernel fills.
template <int NM, int NP2, bool GAUNT, int RS, bool PPN = false>
__global__ __launch_bounds__(kBlock) void sweep_kernel(NewtonArgs A, int lin) {
    if (blockIdx.x == 0 && threadIdx.x == 0) A.ctl[cw(lin ^ 1)] = 0;
    const int slot = blockIdx.x * kBlock + threadIdx.x;
    if (slot >= A.ctl[cw(lin)]) return;
    const int c = A.lists[size_t(lin) * A.cap + slot];
    const Store<NM, NP2> S{A.st_d, A.st_i, A.cap};
    cell_sweep<NM, NP2, GAUNT, RS, PPN>(A, S, c);
}

// Newton / Armijo logic for every cell on list `lin`; cells needing another
// sweep are compacted onto list lin^1.
// MIX: after a mixed-precision sweep; deferred cells go to list defer_to
// (the FP64 pair's next list).
template <int NM, int NP2, bool GAUNT, bool MIX = false, bool PPN = false>
__global__ __launch_bounds__(kBlock) void logic_kernel(NewtonArgs A, int lin, int defer_to) {
    const int slot = blockIdx.x * kBlock + threadIdx.x;
    const Store<NM, NP2> S{A.st_d, A.st_i, A.cap};
    __shared__ float gstage[MIX ? mix_cols(NM) : 1][kBlock];
    LaneStats st;
    int res = LOGIC_DONE;
    int c = -1;
    if (slot < A.ctl[cw(lin)]) {
        c = A.lists[size_t(lin) * A.cap + slot];
        res = cell_logic<NM, NP2, GAUNT, MIX, PPN>(A, S, c, st, &gstage[0][threadIdx.x]);
        if (res == LOGIC_DEFER) ++st.deferred;
    }
    list_append(A, lin ^ 1, cw(lin ^ 1), res == LOGIC_CONTINUE, c);
    if constexpr (MIX) list_append(A, defer_to, cw(defer_to), res == LOGIC_DEFER, c);
    if (A.divert) list_append(A, kMlpList, cw(kMlpList), res == LOGIC_MLP, c);
    flush_stats(A, st);
}

// Entry s of the tail's work: the cells left on list `lin`, then those left
// on list `lin2` (the FP64 list of a mixed-precision solve).  -1 past the end.
__device__ __forceinline__ int tail_cell(const NewtonArgs &A, int lin, int count, int lin2, int count2, int s) {
    if (s < count) return A.lists[size_t(lin) * A.cap + s];
    if (s < count + count2) return A.lists[size_t(lin2) * A.cap + (s - count)];
    return -1;
}

// Persistent tail: lanes pull the remaining cells of list `lin` from a queue
// and iterate each to completion, so a few cells needing tens of iterations
// near the realizability boundary cost one launch, not one per iteration.
template <int NM, int NP2, bool GAUNT, int RS>
__global__ __launch_bounds__(kBlock) void tail_kernel(NewtonArgs A, int lin, int lin2) {
    const Store<NM, NP2> S{A.st_d, A.st_i, A.cap};
    const int lane = threadIdx.x & 63;
    const int count = A.ctl[cw(lin)], dcount = A.ctl[cw(lin2)];
    LaneStats st;
    int c = -1;          // -1 wants work, -2 finished
    for (;;) {
        const bool need = (c == -1);
        const unsigned long long m = __ballot(need);
        if (m) {
            const int leader = __ffsll((long long)m) - 1;
            int base = 0;
            if (lane == leader) base = atomicAdd(&A.ctl[2], __popcll(m));
            base = __shfl(base, leader);
            if (need) {
                const int s = base + __popcll(m & lane_mask_lt());
                c = tail_cell(A, lin, count, lin2, dcount, s);
                if (c < 0) c = -2;
                if (c >= 0) {
                    ++st.tcells;
                    st.titer -= S.i(Store<NM, NP2>::I_TOTAL, c);
                }
            }
        }
        if (__ballot(c >= 0) == 0) break;
        if (c >= 0) {
            cell_sweep<NM, NP2, GAUNT, RS>(A, S, c);
            if (cell_logic<NM, NP2, GAUNT>(A, S, c, st) == LOGIC_DONE) c = -1;
        }
    }
    st.titer += st.sum;           // the converged cells' final totals
    st.ttrials = st.trials;
    flush_stats(A, st);
}

// Cooperative tail: each wave takes one remaining cell at a time and iterates
// it to completion, the 64 lanes sharing every quadrature sweep.  The slowest
// near-boundary cells (tens of iterations) then cost microseconds per
// iteration instead of a single lane's full sweep.  All lanes run the per-cell
// logic on identical data (uniform control flow; stores write equal values).
template <int NM, int NP2, bool GAUNT, int RS>
__global__ __launch_bounds__(kBlock) void tail_wave_kernel(NewtonArgs A, int lin, int lin2) {
    using St = Store<NM, NP2>;
    const Store<NM, NP2> S{A.st_d, A.st_i, A.cap};
    constexpr int NA = GAUNT ? NP2 : Tri<NM>::at(RS, 0);
    constexpr int NB = (GAUNT || RS >= NM) ? 1 : Tri<NM>::size - Tri<NM>::at(RS, 0);
    const int lane = threadIdx.x & 63;
    const int count = A.ctl[cw(lin)], dcount = A.ctl[cw(lin2)];
    LaneStats st;
    for (;;) {
        int s = 0;
        if (lane == 0) s = atomicAdd(&A.ctl[2], 1);
        s = __shfl(s, 0);
        const int c = tail_cell(A, lin, count, lin2, dcount, s);
        if (c < 0) break;
        if (lane == 0) {
            ++st.tcells;
            st.titer -= S.i(St::I_TOTAL, c);
        }
        for (;;) {
            double al[NM];
#pragma unroll
            for (int k = 0; k < NM; ++k) al[k] = S.d(St::F_AL + k, c);
            {
                double acc[NA];
                const double sf = sweep_wave<NM, NP2, GAUNT, 0, (GAUNT ? NM : RS), NA>(A.p, A.w, A.nq, al, acc);
                S.d(St::F_SF, c) = sf;
#pragma unroll
                for (int k = 0; k < NA; ++k) S.d(St::F_ACC + k, c) = acc[k];
            }
            if constexpr (!GAUNT && RS < NM) {
                double accb[NB];
                sweep_wave<NM, NP2, false, RS, NM, NB>(A.p, A.w, A.nq, al, accb);
#pragma unroll
                for (int k = 0; k < NB; ++k) S.d(St::F_ACC + NA + k, c) = accb[k];
            }
            LaneStats one;
            const int res = cell_logic<NM, NP2, GAUNT>(A, S, c, one);
            if (lane == 0) {
                st.max_iter = max(st.max_iter, one.max_iter);
                st.err_bits |= one.err_bits;
                st.sum += one.sum;
                st.n += one.n;
                st.trials += one.trials;
#pragma unroll
                for (int r = 0; r < 5; ++r) st.hist[r] += one.hist[r];
            }
            if (res == LOGIC_DONE) break;
        }
    }
    st.titer += st.sum;
    st.ttrials = st.trials;
    flush_stats(A, st);
}

// ---- cooperative tail: one wave per cell, dense algebra in LDS -------------
//
// For the few cells still iterating after the lockstep rounds, a wave owns one
// cell: lane l sweeps the quadrature points q = l, l+64, ...; the partial sums
// are reduced through LDS; the Hessian and its factor live in LDS and lane i
// owns row i of the Cholesky factorisation and of the triangular solves.  The
// order of every sum is the serial one (the column-oriented back substitution
// subtracts x_{n-1}, x_{n-2}, ... in the order of the dot-product form), so a
// cell takes the same path here as in the lockstep kernels up to FMA rounding.
// Per-wave state of the cooperative tail.  The direct Hessian's TRI sums are
// accumulated in NPASS passes over the quadrature (a flat split of the packed
// triangle: KH sums, then the rest), so that a pass's partial sums fit the
// registers; the first pass also carries the PPn gradient sums.
template <int NM, int NP2, bool GAUNT, int RS, bool PPN = false>
struct CoopShared {
    static constexpr int NPASS = (!GAUNT && RS < NM) ? 2 : 1;
    static constexpr int TRI = Tri<NM>::size;
    static constexpr int KH = GAUNT ? NP2 : (NPASS == 2 ? (TRI + 1) / 2 : TRI);
    static constexpr int NG = PPN ? NM : 0;
    static constexpr int NV0 = KH + NG, NV1 = NPASS == 2 ? TRI - KH : 0;
    static constexpr int NV = NV0 > NV1 ? NV0 : NV1;      // partial sums per lane and pass
    static constexpr int H1 = (NV + 1) / 2, H2 = (H1 + 1) / 2;   // values per lane after the two exchanges
    static constexpr int NACC = Store<NM, NP2>::nacc;
    double part[H2][65];      // the lane quads' partials, padded
    double accs[NACC];        // reduced sweep sums
    double L[NM][NM];         // H, then its Cholesky factor (lower)
    double x[NM], gvec[NM];
    double gsum[NM];          // PPn: sum w G m
};

// The quadrature table (transposed, pT[k][q]) and weights in LDS, shared by the
// block's waves, when nq <= 64 kTailCap and the rows are the moments (direct
// Hessian): LDS reads instead of L2 round trips in every sweep.
template <int NM, bool ON>
struct TailTable {
    double t[ON ? NM : 1][ON ? 64 * kTailCap : 1];
    double w[ON ? 64 * kTailCap : 1];
};

// v from lane l ^ 1 and from lane l ^ 2 (DPP quad permutations)
__device__ __forceinline__ double dpp_xor1(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_mov_dpp((int)b, 0xB1, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), 0xB1, 0xF, 0xF, false);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

__device__ __forceinline__ double dpp_xor2(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_mov_dpp((int)b, 0x4E, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), 0x4E, 0xF, 0xF, false);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}


// WPE: waves per SIMD the registers are held to.  At 1 the kernel takes up to
// 512 registers, for the latency of few cells (the lattice's tail: 2.65
// against 2.88 ms/step at 2); at 2 it holds 256, with a few spill slots,
// for throughput when many cells enter the tail (the isolated lobe batch:
// 10.9 against 9.2 M closures/s).  At 2 the lane index passes through an
// empty asm each pass and the state's base offsets are formed per cell, so
// that the compiler does not keep lane masks and field addresses live
// across the loops.
template <int NM, int NP2, bool GAUNT, int RS, bool PPN = false, int WPE = 1>
__global__ __launch_bounds__(64 * kTailWaves) __attribute__((amdgpu_waves_per_eu(WPE)))
void tail_coop_kernel(NewtonArgs A, int lin, int lin2) {
    // points per lane interleaved: at two waves per SIMD one at a time (fewer
    // live registers; 4-5% faster there), otherwise kTailPts
    constexpr int PTS = WPE > 1 ? 1 : kTailPts;
    static_assert(NM <= 16, "the max/min reductions over the row lanes take one 16-lane DPP row");
    using St = Store<NM, NP2>;
    using CS = CoopShared<NM, NP2, GAUNT, RS, PPN>;
    constexpr int NG = CS::NG;
    __shared__ CS shw[kTailWaves];
    __shared__ TailTable<NM, !GAUNT> tab;
    CS &sh = shw[threadIdx.x >> 6];
    int lane = threadIdx.x & 63;
    const bool tabled = !GAUNT && A.nq <= 64 * kTailCap;
    if constexpr (!GAUNT) {
        if (tabled) {
            for (int i = threadIdx.x; i < NM * A.nq; i += blockDim.x) tab.t[i / A.nq][i % A.nq] = A.pT[i];
            for (int i = threadIdx.x; i < A.nq; i += blockDim.x) tab.w[i] = A.w[i];
        }
    }
    __syncthreads();      // the only block-wide barrier: from here on the waves are independent
    const int count = A.ctl[cw(lin)], dcount = A.ctl[cw(lin2)];
    const double p0 = A.p[0];
    LaneStats st;
    for (;;) {
        int s = 0;
        if (lane == 0) s = atomicAdd(&A.ctl[2], 1);
        s = __shfl(s, 0);
        // wave-uniform: state loads become scalar loads
        const int c = __builtin_amdgcn_readfirstlane(tail_cell(A, lin, count, lin2, dcount, s));
        if (c < 0) break;
        size_t capc = A.cap;
        if constexpr (WPE > 1) __asm__ volatile("" : "+s"(capc));
        const Store<NM, NP2> S{A.st_d, A.st_i, capc};
        // The cell's state comes into registers once: the tail runs it to the
        // end, so nothing is stored back until the outputs.  Vectors are
        // lane-distributed (element k on lane k), scalars wave-uniform.
        double al[NM];
#pragma unroll
        for (int k = 0; k < NM; ++k) al[k] = S.d(St::F_AL + k, c);
        int mode = S.i(St::I_MODE, c), iter = S.i(St::I_ITER, c), total = S.i(St::I_TOTAL, c);
        int reg = S.i(St::I_REG, c), maxit = S.i(St::I_MAXIT, c), warm = S.i(St::I_WARM, c);
        double t = S.d(St::F_T, c), rtol = S.d(St::F_RTOL, c);
        if (lane == 0) {
            ++st.tcells;
            st.titer -= total;
        }
        const double u0 = S.d(St::F_U0, c), ascale = S.d(St::F_ASCALE, c);
        double fcur_s = S.d(St::F_FCUR, c), gd_s = S.d(St::F_GD, c), fscale_s = S.d(St::F_FSCALE, c);
        double maxf_s = S.d(St::F_MAXF, c);
        double ul = lane < NM ? S.d(St::F_U + lane, c) : 0.0;
        const double delta = PPN ? S.d(St::F_DELTA, c) : 0.0;
        double aold_l = lane < NM ? S.d(St::F_AOLD + lane, c) : 0.0;
        double d_l = lane < NM ? S.d(St::F_D + lane, c) : 0.0;
        for (;;) {
            if constexpr (WPE > 1) __asm__ volatile("" : "+v"(lane));
            // ---- sweep(s): partial sums per lane, reduced through DPP and LDS ----
            double sf = 0.0;
            double wes[kTailCap];
#pragma unroll
            for (int pass = 0; pass < CS::NPASS; ++pass) {
                // flat range [K0, K1) of the sums of this pass (GAUNT: the
                // order-2N moments), then the PPn gradient sums
                const int K0 = pass == 0 ? 0 : CS::KH, K1 = pass == 0 ? CS::KH : CS::TRI;
                const int nv = pass == 0 ? CS::NV0 : CS::NV1;
                double acc[CS::NV];
#pragma unroll
                for (int k = 0; k < CS::NV; ++k) acc[k] = 0.0;
                // Lane l takes the points q = l, l + 64, ...; the moment values
                // come from the LDS table or from the transposed table in global
                // memory (for each k the wave reads 64 consecutive doubles).
                // With at most kTailCap points per lane, w e of the first pass
                // is kept for the second.  The points run PTS at a time,
                // a scheduling barrier between the groups so that the next
                // group's rows are not loaded while this group's are live
                // (each lane's sums still take its points in ascending order).
                const bool cached = A.nq <= 64 * kTailCap;
                for (int q0 = 0; q0 < A.nq; q0 += 64 * kTailCap) {
#pragma unroll
                    for (int qi = 0; qi < kTailCap; ++qi) {
                        if (qi > 0 && qi % PTS == 0) {
                            if (q0 + 64 * qi >= A.nq) break;
                            __builtin_amdgcn_sched_barrier(0);
                        }
                        // points past nq read point 0 and add exact zeros, so
                        // the PTS points' chains can interleave (the
                        // order-2N rows of the Gaunt form are too many registers
                        // for that: there the points stay sequential)
                        const int q = q0 + 64 * qi + lane;
                        if (GAUNT && q >= A.nq) break;
                        const bool valid = q < A.nq;
                        const int qq = valid ? q : 0;
                        double row[GAUNT ? NP2 : NM];
                        double wq;
                        if constexpr (!GAUNT) {
                            if (tabled) {
#pragma unroll
                                for (int k = 0; k < NM; ++k) row[k] = tab.t[k][qq];
                                wq = tab.w[qq];
                            } else {
#pragma unroll
                                for (int k = 0; k < NM; ++k) row[k] = A.pT[size_t(k) * A.nq + qq];
                                wq = A.w[qq];
                            }
                        } else {
#pragma unroll
                            for (int k = 0; k < NP2; ++k) row[k] = A.pT[size_t(k) * A.nq + qq];
                            wq = A.w[qq];
                        }
                        double we;                    // the Hessian's weight
                        if (pass > 0 && cached) {
                            we = wes[qi];
                        } else {
                            double am = 0.0;
#pragma unroll
                            for (int k = 0; k < NM; ++k) am = fma(al[k], row[k], am);
                            if constexpr (PPN) {
                                const double tq = 0.5 * am;
                                double sq;
                                const double G = ppn_g(tq, delta, sq);
                                we = valid ? 0.5 * wq * (G / sq) : 0.0;
                                if (pass == 0) {
                                    sf = sf + (valid ? wq * (tq * G - 0.5 * delta + delta * log(G)) : 0.0);
                                    const double wg = valid ? wq * G : 0.0;
#pragma unroll
                                    for (int k = 0; k < NG; ++k) acc[CS::KH + k] = fma(row[k], wg, acc[CS::KH + k]);
                                }
                            } else {
                                we = valid ? wq * exp(am) : 0.0;
                            }
                            wes[qi] = we;
                        }
                        if (!PPN && pass == 0) sf += we;
                        if constexpr (GAUNT) {
#pragma unroll
                            for (int k = 0; k < NP2; ++k) acc[k] = fma(we, row[k], acc[k]);
                        } else {
#pragma unroll
                            for (int i = 0; i < NM; ++i) {
                                if (Tri<NM>::at(i, i) < K0 || Tri<NM>::at(i, 0) >= K1) continue;
                                const double v = we * row[i];
#pragma unroll
                                for (int jj = 0; jj <= i; ++jj) {
                                    const int k = Tri<NM>::at(i, jj) - K0;
                                    if (k >= 0 && k < K1 - K0) acc[k] = fma(v, row[jj], acc[k]);
                                }
                            }
                        }
                    }
                }
                // Two exchanges between lanes (l ^ 1, then l ^ 2): each halves
                // the values a lane holds and sums them over the pair, so that
                // lane l ends with H2 values summed over its quad -- value
                // (l & 1 ? H1 : 0) + (l & 2 ? H2 : 0) + r in slot r -- and the
                // sixteen quads' partials of each value are then added through
                // LDS by one lane, four chains in quad order.
                double a1[CS::H1];
#pragma unroll
                for (int r = 0; r < CS::H1; ++r) {
                    const double lo = acc[r], hi = (CS::H1 + r < CS::NV) ? acc[CS::H1 + r] : 0.0;
                    const bool b = lane & 1;
                    a1[r] = (b ? hi : lo) + dpp_xor1(b ? lo : hi);
                }
#pragma unroll
                for (int r = 0; r < CS::H2; ++r) {
                    const double lo = a1[r], hi = (CS::H2 + r < CS::H1) ? a1[CS::H2 + r] : 0.0;
                    const bool b = lane & 2;
                    sh.part[r][lane] = (b ? hi : lo) + dpp_xor2(b ? lo : hi);
                }
                __builtin_amdgcn_wave_barrier();
                for (int k = lane; k < nv; k += 64) {
                    const int hb = k >= CS::H1 ? 1 : 0, rel = k - hb * CS::H1;
                    const int qb = rel >= CS::H2 ? 1 : 0, r = rel - qb * CS::H2;
                    const int cls = hb | (qb << 1);
                    double t4[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
                    for (int m = 0; m < 16; m += 4) {
#pragma unroll
                        for (int u = 0; u < 4; ++u) t4[u] += sh.part[r][4 * (m + u) + cls];
                    }
                    const double tot = (t4[0] + t4[1]) + (t4[2] + t4[3]);
                    if (PPN && pass == 0 && k >= CS::KH) sh.gsum[k - CS::KH] = tot;
                    else sh.accs[K0 + k] = tot;
                }
                __builtin_amdgcn_wave_barrier();
            }
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) sf += __shfl_xor(sf, o);
            __builtin_amdgcn_wave_barrier();
            // ---- per-cell logic (cell_logic's steps, cooperative) ----
            bool newton = (mode == ST_NEWTON), do_reg = false, converged = false;
            double err = 0.0;
            if (mode == ST_TRIAL) {
                if (lane == 0) ++st.trials;
                double fnew = sf;
#pragma unroll
                for (int k = 0; k < NM; ++k) fnew = fnew - al[k] * readlane_d(ul, k);
                const double fcur = fcur_s, gd = gd_s;
                bool accept = fnew <= fcur + 1e-3 * t * gd;
                if (!accept && A.ls_mode == MNC_LS_ROBUST && t == 1.0 && -gd <= 1e3 * DBL_EPSILON * fscale_s)
                    accept = true;
                if (accept) {
                    newton = true;
                } else {
                    // The reference halves t and evaluates f again until
                    // Armijo holds or t * maxf <= eps/2 (linesearch.cpp:45-98).
                    // The next candidates (kFirstHalvings in the first batch,
                    // where most searches end, then kHalvings per batch) are
                    // evaluated in one wave sweep (f only, the same sums in the
                    // same order as a trial sweep) and the first accepted one is
                    // taken: the reference's decisions without one pass per
                    // halving.
                    const double maxf = maxf_s;
                    if (lane < NM) {
                        sh.x[lane] = aold_l;
                        sh.gvec[lane] = d_l;
                    }
                    __builtin_amdgcn_wave_barrier();
                    int state = 0;                 // 0 searching, 1 accepted, 2 exhausted
                    int cap = kFirstHalvings;      // candidates in this batch
                    while (state == 0) {
                        int nval = 0;
#pragma unroll
                        for (int jj = 0; jj < kHalvings; ++jj)
                            if (jj < cap && nval == jj && ldexp(t, -(jj + 1)) * maxf > DBL_EPSILON / 2) nval = jj + 1;
                        if (nval == 0) { state = 2; break; }
                        double fj[kHalvings];
#pragma unroll
                        for (int jj = 0; jj < kHalvings; ++jj) fj[jj] = 0.0;
                        // candidates outside, the lane's points inside (their
                        // chains interleave); each fj sums the lane's points in
                        // order, as the trial sweep's f does
                        for (int q0 = 0; q0 < A.nq; q0 += 64 * PTS) {
#pragma unroll
                            for (int jj = 0; jj < kHalvings; ++jj) {
                                if (jj >= nval) break;                  // wave-uniform
                                const double tj = ldexp(t, -(jj + 1));
                                double aj[NM];
#pragma unroll
                                for (int k = 0; k < NM; ++k) aj[k] = fma(tj, sh.gvec[k], sh.x[k]);
#pragma unroll
                                for (int qi = 0; qi < PTS; ++qi) {
                                    const int q = q0 + 64 * qi + lane;
                                    if (GAUNT && q >= A.nq) break;      // Gaunt form: sequential points
                                    const bool valid = q < A.nq;
                                    const int qq = valid ? q : 0;
                                    double row[NM], wq;
                                    if constexpr (!GAUNT) {
                                        if (tabled) {
#pragma unroll
                                            for (int k = 0; k < NM; ++k) row[k] = tab.t[k][qq];
                                            wq = tab.w[qq];
                                        } else {
#pragma unroll
                                            for (int k = 0; k < NM; ++k) row[k] = A.pT[size_t(k) * A.nq + qq];
                                            wq = A.w[qq];
                                        }
                                    } else {
#pragma unroll
                                        for (int k = 0; k < NM; ++k) row[k] = A.pT[size_t(k) * A.nq + qq];
                                        wq = A.w[qq];
                                    }
                                    double am = 0.0;
#pragma unroll
                                    for (int k = 0; k < NM; ++k) am = fma(aj[k], row[k], am);
                                    if constexpr (PPN) {
                                        const double tq = 0.5 * am;
                                        double sq;
                                        const double G = ppn_g(tq, delta, sq);
                                        fj[jj] += valid ? wq * (tq * G - 0.5 * delta + delta * log(G)) : 0.0;
                                    } else {
                                        fj[jj] += valid ? wq * exp(am) : 0.0;
                                    }
                                }
                            }
                        }
                        int acc = -1;
                        for (int jj = 0; jj < kHalvings; ++jj) {
                            if (jj >= nval) break;
                            double fv = fj[jj];
#pragma unroll
                            for (int o = 32; o > 0; o >>= 1) fv += __shfl_xor(fv, o);
                            const double tj = ldexp(t, -(jj + 1));
                            for (int k = 0; k < NM; ++k)
                                fv = fv - fma(tj, sh.gvec[k], sh.x[k]) * readlane_d(ul, k);
                            if (acc < 0 && fv <= fcur + 1e-3 * tj * gd) acc = jj;
                        }
                        if (lane == 0) st.trials += (acc >= 0) ? acc + 1 : nval;
                        if (acc >= 0) {
                            t = ldexp(t, -(acc + 1));
                            state = 1;
                        } else {
                            t = ldexp(t, -nval);
                            if (nval < cap) state = 2;           // the next halving fails the bound
                        }
                        cap = kHalvings;
                    }
                    if (state == 1) {                  // a full sweep at the accepted point next
#pragma unroll
                        for (int k = 0; k < NM; ++k) al[k] = fma(t, readlane_d(d_l, k), readlane_d(aold_l, k));
                        mode = ST_NEWTON;
                    } else {
#pragma unroll
                        for (int k = 0; k < NM; ++k) al[k] = readlane_d(aold_l, k);
                        t = 0.0;
                        ++iter; ++total;
                        do_reg = true;
                    }
                    __builtin_amdgcn_wave_barrier();
                }
            }
            if (newton) {
                ++iter; ++total;
                // H: row i in the registers of lane i (full symmetric row)
                double row[NM];
#pragma unroll
                for (int jj = 0; jj < NM; ++jj) {
                    const int i0 = lane > jj ? lane : jj, j0 = lane > jj ? jj : lane;
                    double hij = 0.0;
                    if (lane < NM) {
                        if constexpr (GAUNT) {
                            const int pr = j0 * NM + i0;
                            for (int cc = A.gptr[pr]; cc < A.gptr[pr + 1]; ++cc)
                                hij = fma(A.gv[cc], sh.accs[A.gk[cc]], hij);
                        } else {
                            hij = sh.accs[Tri<NM>::at(i0, j0)];
                        }
                    }
                    row[jj] = hij;
                }
                // g_i = H_i0 / m0 - u_i on lane i; f, its scale and ||g|| in the serial order
                const double gv = lane < NM ? (PPN ? sh.gsum[lane] : row[0] / p0) - ul : 0.0;
                if (lane < NM) sh.gvec[lane] = gv;
                double fcur = PPN ? sf : readlane_d(row[0], 0) / (p0 * p0), fscale = fabs(fcur), gg = 0.0;
#pragma unroll
                for (int i = 0; i < NM; ++i) {
                    const double ui = readlane_d(ul, i);
                    const double gi = readlane_d(gv, i);
                    gg = fma(gi, gi, gg);
                    fcur = fcur - al[i] * ui;
                    fscale += fabs(al[i] * ui);
                }
                double rs = 0.0, dg = 0.0;             // row sums (utils.cpp:36-52), diagonal
#pragma unroll
                for (int jj = 0; jj < NM; ++jj) {
                    rs += fabs(row[jj]);
                    dg = (jj == lane) ? row[jj] : dg;
                }
                const double anorm = row16_max(lane < NM ? rs : 0.0);
                const double gersh = row16_min(lane < NM ? 2.0 * dg - rs : DBL_MAX);
                err = sqrt(gg);
                const bool chol_ok = reg_cholesky<NM>(row, lane);
                // the columns of the factor (lane i: L[k][i], k > i) for the back substitution
                double col[NM];
                if (lane < NM) {
#pragma unroll
                    for (int k = 0; k < NM; ++k) sh.L[lane][k] = row[k];
                }
                __builtin_amdgcn_wave_barrier();
#pragma unroll
                for (int k = 0; k < NM; ++k) col[k] = (lane < NM && k > lane) ? sh.L[k][lane] : 0.0;
                __builtin_amdgcn_wave_barrier();
                // reciprocals of the factor's diagonal, independent of each
                // other: the solves below multiply by them (as blocked trsm
                // kernels do) instead of dividing on their dependent chains
                double rd[NM];
#pragma unroll
                for (int k = 0; k < NM; ++k) rd[k] = 1.0 / readlane_d(row[k], k);
                double condest = 0.0;
                // where the bounds cannot decide the test, the exact ||A^-1||_1
                // bound and the direction come from one pass over the factor
                double xv = gv;
                bool solved = false;
                if (chol_ok && !cond_proven_below(anorm, gersh, NM, A.cond_h_max)) {
                    const double inv1 = reg_inv_norm1_solve<NM>(row, col, rd, xv, lane);
                    solved = true;
                    if (!(2.0 * anorm * inv1 <= A.cond_h_max))
                        condest = 1.0 / reg_rcond<NM>(row, col, anorm, lane);
                }
                if (iter >= maxit || !chol_ok || t == 0.0 || condest > A.cond_h_max) {
                    do_reg = true;
                } else {
                    if (!solved) reg_apply_inv_rcp<NM>(row, col, rd, xv, lane);
                    bool stop = err <= rtol;
                    if (stop && A.tol_gamma > 0.0) {
                        double da1 = 0.0;
#pragma unroll
                        for (int i = 0; i < NM; ++i) da1 = da1 + fabs(readlane_d(xv, i));
                        stop = gamma_ok(A, da1, PPN);
                    }
                    if (stop) {
                        converged = true;
                    } else {
                        const double di = -xv;
                        double al_l = 0.0;
#pragma unroll
                        for (int k = 0; k < NM; ++k) al_l = (k == lane) ? al[k] : al_l;
                        const double maxf = row16_max(lane < NM ? fabs(di) / (fabs(al_l) + 1e-100) : 0.0);
                        double gd = 0.0;
#pragma unroll
                        for (int i = 0; i < NM; ++i) gd = fma(readlane_d(gv, i), readlane_d(di, i), gd);
                        aold_l = lane < NM ? al_l : 0.0;
                        d_l = lane < NM ? di : 0.0;
                        maxf_s = maxf;
                        gd_s = gd;
                        fcur_s = fcur;
                        fscale_s = fscale;
                        t = 1.0;
                        if (t * maxf > DBL_EPSILON / 2) {
#pragma unroll
                            for (int i = 0; i < NM; ++i) al[i] = al[i] + t * readlane_d(di, i);
                            mode = ST_TRIAL;
                        } else {
                            t = 0.0;
                            ++iter; ++total;
                            do_reg = true;
                        }
                    }
                }
            }
            __builtin_amdgcn_wave_barrier();
            bool done = false;
            if (do_reg && !PPN && net_fallback(A, reg, warm, iter)) {   // the network's guess failed
                warm = kWarmFellBack;
                if (lane == 0) {
                    S.i(St::I_WARM, c) = kWarmFellBack;
                    ++st.fallback;
                }
                al[0] = 2.0 * kSqrtPi * log(0.5 / kSqrtPi);
#pragma unroll
                for (int k = 1; k < NM; ++k) al[k] = 0.0;
                iter = 0;
                t = 1.0;
                mode = ST_NEWTON;
            } else if (do_reg) {
                if (reg < 4) {
                    ++reg;
                    const double r = kReg[reg];
                    const double *uin = A.u + size_t(c) * NM;
                    ul = lane == 0 ? 1.0 : (lane < NM ? (uin[lane] / u0) * (1.0 - r) : 0.0);
                    double ss = 1.0;
                    for (int k = 1; k < NM; ++k) {
                        const double uk = readlane_d(ul, k);
                        ss = fma(uk, uk, ss);
                    }
                    rtol = sqrt(ss) * A.tol;
                    al[0] = PPN ? 1.0 - 4.0 * delta * kPi : 2.0 * kSqrtPi * log(0.5 / kSqrtPi);
#pragma unroll
                    for (int k = 1; k < NM; ++k) al[k] = 0.0;
                    iter = 0;
                    t = 1.0;
                    if (reg == 4) maxit = 10 * maxit;
                    mode = ST_NEWTON;
                } else {
                    if (lane == 0) {
                        st.err_bits |= 1 << MNC_ERR_MAXREG;
                        if (A.outs) { A.outs[c].iter = total; A.outs[c].r = -1.0; A.outs[c].norm_g = err; }
                    }
                    done = true;
                }
            } else if (converged) {
                if (PPN) {                                    // no reconstruction; opt.cpp:131-136
                    if (lane < NM) {
                        double ak = 0.0;
#pragma unroll
                        for (int kk = 0; kk < NM; ++kk) ak = (kk == lane) ? al[kk] : ak;
                        A.u[size_t(c) * NM + lane] = ul * u0;
                        A.alpha[size_t(c) * NM + lane] = ak * u0;
                    }
                } else {
                    const double u0p = readlane_d(ul, 0) + sh.gvec[0];
                    double a0 = al[0];
                    a0 -= log(u0p) * 2.0 * kSqrtPi;
                    a0 = a0 + ascale;
                    if (lane < NM) {
                        A.u[size_t(c) * NM + lane] = ((ul + sh.gvec[lane]) / u0p) * u0;
                        double ak = 0.0;
#pragma unroll
                        for (int kk = 0; kk < NM; ++kk) ak = (kk == lane) ? al[kk] : ak;
                        A.alpha[size_t(c) * NM + lane] = lane == 0 ? a0 : ak;
                    }
                }
                if (lane == 0) {
                    if (A.outs) { A.outs[c].iter = total; A.outs[c].r = kReg[reg]; A.outs[c].norm_g = err; }
                    st.max_iter = total > st.max_iter ? total : st.max_iter;
                    st.sum += total;
                    ++st.n;
                    count_reg(st, reg);
                }
                done = true;
            }
            __builtin_amdgcn_wave_barrier();
            if (done) break;
        }
    }
    // statistics: only lane 0 of each wave counted
    if (lane == 0 && (st.n || st.trials || st.err_bits || st.tcells)) {
        atomicAdd((unsigned long long *)&A.stats[13], (unsigned long long)st.tcells);
        atomicAdd((unsigned long long *)&A.stats[14], (unsigned long long)(st.titer + st.sum));
        atomicAdd((unsigned long long *)&A.stats[15], (unsigned long long)st.trials);
        atomicMax((unsigned long long *)&A.stats[0], (unsigned long long)st.max_iter);
        atomicAdd((unsigned long long *)&A.stats[1], (unsigned long long)st.sum);
        atomicAdd((unsigned long long *)&A.stats[2], (unsigned long long)st.n);
        for (int r = 0; r < 5; ++r) atomicAdd((unsigned long long *)&A.stats[3 + r], (unsigned long long)st.hist[r]);
        atomicAdd((unsigned long long *)&A.stats[8], (unsigned long long)st.trials);
        if (st.fallback) atomicAdd((unsigned long long *)&A.stats[16], (unsigned long long)st.fallback);
        if (st.err_bits) atomicOr(&A.ctl[3], st.err_bits);
    }
}

// ---- mixed-precision sweep on the matrix cores ------------------------------
//
// A wave evaluates sixteen cells of the active list at once.  Lane l works for
// cell l%16 on the quadrature points q = 16T + mfma_f64_row(r, l/16) (r = 0..3)
// of every 16-point tile T, which is where the first product leaves them, so no
// value crosses lanes:
//   z = alpha.m       v_mfma_f64_16x16x4f64, (16 points x 16 moments) x (16 moments x 16 cells)
//   e = exp(z)        FP64, per lane
//   G += e (w m, w)   v_mfma_f64_16x16x4f64: FP64 sums of w e m_i and w e (columns 0..14, 15)
//   Hm += e (w m2N)   v_mfma_f32_16x16x32_bf16 on bf16 parts of e and of the
//                     table: the order-2N moments from which the logic forms H
//                     through the Gaunt map.  NS = 2: hi/lo pairs, three products
//                     (hi hi, hi lo, lo hi; relative error ~2^-16), FP32 sums over
//                     the whole quadrature.  NS = 3: hi/mid/lo triples, the six
//                     products down to order 2^-16 (error ~2^-24), FP32 sums per
//                     32 points added in FP64.
// The operand tables (built by build_mix_tables) sit in LDS for the whole
// launch; blocks loop over tiles.
typedef double mf64x4 __attribute__((ext_vector_type(4)));
typedef float mf32x4 __attribute__((ext_vector_type(4)));
typedef __bf16 mbf16x8 __attribute__((ext_vector_type(8)));
typedef unsigned short mu16x8 __attribute__((ext_vector_type(8)));

constexpr int kMixWaves = 16;

template <int NNT, int NS>
struct MixShared {
    double z[kMixMaxTiles][4][64];
    double g[kMixMaxTiles][4][64];
    mu16x8 h[kMixMaxTiles / 2][NNT][NS][64];
};

// x = hi + lo up to 2^-17 |x|, both rounded to nearest even (non-finite x
// gives a non-finite pair, which fails the factorisation downstream).
__device__ __forceinline__ void split_bf16(float x, unsigned short &hi, unsigned short &lo) {
    const unsigned u = __float_as_uint(x);
    const unsigned r = u + 0x7fffu + ((u >> 16) & 1u);
    hi = (unsigned short)(r >> 16);
    const unsigned v = __float_as_uint(x - __uint_as_float(r & 0xffff0000u));
    lo = (unsigned short)((v + 0x7fffu + ((v >> 16) & 1u)) >> 16);
}

// x = hi + mid + lo up to 2^-25 |x| (x - hi and x - hi - mid are exact in FP32)
__device__ __forceinline__ unsigned short bf16_rne_bits(float x) {
    const unsigned u = __float_as_uint(x);
    return (unsigned short)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}

__device__ __forceinline__ float bf16_bits_val(unsigned short h) { return __uint_as_float(unsigned(h) << 16); }

__device__ __forceinline__ void split3_bf16(float x, unsigned short &hi, unsigned short &mid, unsigned short &lo) {
    hi = bf16_rne_bits(x);
    const float r = x - bf16_bits_val(hi);
    mid = bf16_rne_bits(r);
    lo = bf16_rne_bits(r - bf16_bits_val(mid));
}

template <int NM, int NP2, int NNT, int NS>
__global__ __launch_bounds__(64 * kMixWaves) void sweep_mix_kernel(NewtonArgs A, int lin) {
    using St = Store<NM, NP2>;
    static_assert(NM <= 15, "moment 15 of the sum tile holds sum w e");
    static_assert(16 * NNT <= St::nacc, "Hessian moments exceed the store");
    static_assert(NS == 2 || NS == 3, "two or three bf16 parts");
    static_assert(NS <= kMixParts, "the table holds kMixParts parts");
    __shared__ MixShared<NNT, NS> sh;
    if (blockIdx.x == 0 && threadIdx.x == 0) A.ctl[lin ^ 1] = 0;
    const int count = A.ctl[lin];
    const int ntiles = (count + 15) >> 4;
    if (blockIdx.x * kMixWaves >= ntiles) return;          // uniform per block
    const int nt = A.mnt, nkb = nt >> 1;
    {
        double *sz = &sh.z[0][0][0], *sg = &sh.g[0][0][0];
        for (int i = threadIdx.x; i < nt * 256; i += blockDim.x) {
            sz[i] = A.mz[i];
            sg[i] = A.mg[i];
        }
        // table blocks (kb, n) hold kMixParts parts of 64 lanes; the first NS are used
        const mu16x8 *gh = reinterpret_cast<const mu16x8 *>(A.mh);
        mu16x8 *hh = &sh.h[0][0][0][0];
        for (int i = threadIdx.x; i < nkb * NNT * NS * 64; i += blockDim.x) {
            const int l = i & 63, rest = i >> 6;
            hh[i] = gh[((rest / NS) * kMixParts + rest % NS) * 64 + l];
        }
    }
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int ci = lane & 15, g = lane >> 4;
    const St S{A.st_d, A.st_i, A.cap};
    for (int tile = blockIdx.x * kMixWaves + wave; tile < ntiles; tile += gridDim.x * kMixWaves) {
        const int slot = tile * 16 + ci;
        const int c = slot < count ? A.lists[size_t(lin) * A.cap + slot] : -1;
        double ab[4];
#pragma unroll
        for (int kk = 0; kk < 4; ++kk) {
            const int k = 4 * kk + g;
            ab[kk] = (c >= 0 && k < NM) ? S.d(St::F_AL + k, c) : 0.0;
        }
        mf64x4 gacc = {0.0, 0.0, 0.0, 0.0};
        mf32x4 hacc[NNT];
        mf64x4 hsum[NS == 3 ? NNT : 1];
#pragma unroll
        for (int n = 0; n < NNT; ++n) hacc[n] = mf32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int n = 0; n < (NS == 3 ? NNT : 1); ++n) hsum[n] = mf64x4{0.0, 0.0, 0.0, 0.0};
        for (int kb = 0; kb < nkb; ++kb) {
            mu16x8 ahi, alo, amid;
#pragma unroll
            for (int hh = 0; hh < 2; ++hh) {
                const int T = 2 * kb + hh;
                mf64x4 z = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
                for (int kk = 0; kk < 4; ++kk)
                    z = __builtin_amdgcn_mfma_f64_16x16x4f64(sh.z[T][kk][lane], ab[kk], z, 0, 0, 0);
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const double e = exp(z[r]);
                    gacc = __builtin_amdgcn_mfma_f64_16x16x4f64(e, sh.g[T][r][lane], gacc, 0, 0, 0);
                    unsigned short hi, mid, lo;
                    if constexpr (NS == 3) {
                        split3_bf16((float)e, hi, mid, lo);
                        amid[4 * hh + r] = mid;
                    } else {
                        split_bf16((float)e, hi, lo);
                    }
                    ahi[4 * hh + r] = hi;
                    alo[4 * hh + r] = lo;
                }
            }
            const mbf16x8 Ahi = __builtin_bit_cast(mbf16x8, ahi), Alo = __builtin_bit_cast(mbf16x8, alo);
            if constexpr (NS == 3) {
                const mbf16x8 Amid = __builtin_bit_cast(mbf16x8, amid);
#pragma unroll
                for (int n = 0; n < NNT; ++n) {
                    const mbf16x8 Bhi = __builtin_bit_cast(mbf16x8, sh.h[kb][n][0][lane]);
                    const mbf16x8 Bmid = __builtin_bit_cast(mbf16x8, sh.h[kb][n][1][lane]);
                    const mbf16x8 Blo = __builtin_bit_cast(mbf16x8, sh.h[kb][n][NS - 1][lane]);
                    // smallest terms first
                    mf32x4 t = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Ahi, Blo, mf32x4{0.f, 0.f, 0.f, 0.f}, 0, 0, 0);
                    t = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Amid, Bmid, t, 0, 0, 0);
                    t = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Alo, Bhi, t, 0, 0, 0);
                    t = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Ahi, Bmid, t, 0, 0, 0);
                    t = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Amid, Bhi, t, 0, 0, 0);
                    t = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Ahi, Bhi, t, 0, 0, 0);
#pragma unroll
                    for (int r = 0; r < 4; ++r) hsum[n][r] += (double)t[r];
                }
            } else {
#pragma unroll
                for (int n = 0; n < NNT; ++n) {
                    const mbf16x8 Bhi = __builtin_bit_cast(mbf16x8, sh.h[kb][n][0][lane]);
                    const mbf16x8 Blo = __builtin_bit_cast(mbf16x8, sh.h[kb][n][1][lane]);
                    hacc[n] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Ahi, Bhi, hacc[n], 0, 0, 0);
                    hacc[n] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Ahi, Blo, hacc[n], 0, 0, 0);
                    hacc[n] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Alo, Bhi, hacc[n], 0, 0, 0);
                }
            }
        }
        // Result rows are cells: the cell of slot 16 tile + i is held by lane i.
        // Element r of lane l is row mfma_f64_row(r, l/16) of the FP64 sums and
        // row 4(l/16) + r of the bf16 products; column l%16 of both.
        const int j = lane & 15;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int cg = __shfl(c, mfma_f64_row(r, g));
            const int ch = __shfl(c, 4 * g + r);
            if (cg >= 0) {
                if (j < NM) S.d(St::F_GS + j, cg) = gacc[r];
                if (j == 15) S.d(St::F_SF, cg) = gacc[r];
            }
            if (ch >= 0) {
#pragma unroll
                for (int n = 0; n < NNT; ++n)
                    S.d(St::F_ACC + 16 * n + j, ch) = NS == 3 ? hsum[n < (NS == 3 ? NNT : 1) ? n : 0][r] : (double)hacc[n][r];
            }
        }
    }
}

// ---- FP64 matrix-core path for the high orders (M8: n = 45, nq = 1024) -------
//
// The direct Hessian of 45 moments (1035 sums) does not fit a lane, so the
// sweep runs on the matrix cores with the lane scheme of the mixed sweep, all in
// FP64, sixteen cells per wave:
//   z = alpha.m        KZ products per 16-point tile
//   G = sums w e m_i   NGT column tiles; the last column is sum w e (f)
//   Hm = sums w e p2N  NHT column tiles: the order-2N moments, from which the
//                      logic forms H through the Gaunt map
// f, g and H are then the reference's full evaluation (fobj.cpp:266-287) to
// rounding: f = sum w e - alpha.u, g_i = sum w e m_i - u_i, and H differs from
// the direct sum only by the Gaunt least-squares fit (~1e-14).  The operand
// tables (build_g64_tables, 2 MB at M8/Q32) are read from global memory.
template <int NM, int NP2, int KZ, int NGT, int NHT>
__global__ __launch_bounds__(256) void sweep_g64_kernel(NewtonArgs A, int lin) {
    using St = Store<NM, NP2>;
    static_assert(16 * NGT - 1 >= NM, "the last sum column holds sum w e");
    static_assert(16 * NHT <= St::nacc, "order-2N moments exceed the store");
    if (blockIdx.x == 0 && threadIdx.x == 0) A.ctl[lin ^ 1] = 0;
    const int count = A.ctl[lin];
    const int ntiles = (count + 15) >> 4;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int ci = lane & 15, g = lane >> 4;
    const St S{A.st_d, A.st_i, A.cap};
    const int nt = A.gnt, nmm = A.gnmm;
    for (int tile = blockIdx.x * 4 + wave; tile < ntiles; tile += gridDim.x * 4) {
        const int slot = tile * 16 + ci;
        const int c = slot < count ? A.lists[size_t(lin) * A.cap + slot] : -1;
        double ab[KZ];
#pragma unroll
        for (int kk = 0; kk < KZ; ++kk) {
            const int k = 4 * kk + g;
            ab[kk] = (c >= 0 && k < NM) ? S.d(St::F_AL + k, c) : 0.0;
        }
        mf64x4 gacc[NGT], hacc[NHT];
#pragma unroll
        for (int n = 0; n < NGT; ++n) gacc[n] = mf64x4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int n = 0; n < NHT; ++n) hacc[n] = mf64x4{0.0, 0.0, 0.0, 0.0};
        for (int T = 0; T < nt; ++T) {
            const double *zt = A.gz + size_t(T) * KZ * 64 + lane;
            const double *gt = A.gg + size_t(T) * 4 * NGT * 64 + lane;
            const double *ht = A.gh + size_t(T) * 4 * NHT * 64 + lane;
            mf64x4 z = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int kk = 0; kk < KZ; ++kk) z = __builtin_amdgcn_mfma_f64_16x16x4f64(zt[kk * 64], ab[kk], z, 0, 0, 0);
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const double e = exp(z[r]);
#pragma unroll
                for (int n = 0; n < NGT; ++n)
                    gacc[n] = __builtin_amdgcn_mfma_f64_16x16x4f64(e, gt[(r * NGT + n) * 64], gacc[n], 0, 0, 0);
#pragma unroll
                for (int n = 0; n < NHT; ++n)
                    hacc[n] = __builtin_amdgcn_mfma_f64_16x16x4f64(e, ht[(r * NHT + n) * 64], hacc[n], 0, 0, 0);
            }
        }
        const int j = lane & 15;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int cg = __shfl(c, mfma_f64_row(r, g));
            if (cg < 0) continue;
#pragma unroll
            for (int n = 0; n < NGT; ++n) {
                const int col = 16 * n + j;
                if (col < NM) S.d(St::F_GS + col, cg) = gacc[n][r];
                if (col == 16 * NGT - 1) S.d(St::F_SF, cg) = gacc[n][r];
            }
#pragma unroll
            for (int n = 0; n < NHT; ++n) {
                const int col = 16 * n + j;
                if (col < nmm) S.d(St::F_ACC + col, cg) = hacc[n][r];
            }
        }
    }
}

template <int NM, int NACC>
struct G64Shared {
    double L[NM][NM];   // H, then its Cholesky factor (lower)
    double x[NM], gvec[NM], al[NM];
    double mom[NACC];   // the sweep's order-2N moments of the cell
};

// One step of optScaled's logic per cell after sweep_g64_kernel, one wave per
// cell (lane i owns row i of H and of its factor; the iterate in LDS).  The
// steps are those of cell_logic (opt.cpp:209-318, linesearch.cpp:38-98);
// cells needing another sweep go onto list lin^1.
template <int NM, int NP2>
__global__ __launch_bounds__(64) void logic_g64_kernel(NewtonArgs A, int lin) {
    using St = Store<NM, NP2>;
    constexpr int NK = gp_nm(2 * order_of(NM));    // the order-2N moments (153 at n = 45)
    __shared__ G64Shared<NM, NK> sh;
    const St S{A.st_d, A.st_i, A.cap};
    const int lane = threadIdx.x;
    const int count = A.ctl[lin];
    LaneStats st;
    for (int slot = blockIdx.x; slot < count; slot += gridDim.x) {
        const int c = A.lists[size_t(lin) * A.cap + slot];
        int mode = S.i(St::I_MODE, c), iter = S.i(St::I_ITER, c), total = S.i(St::I_TOTAL, c);
        int reg = S.i(St::I_REG, c), maxit = S.i(St::I_MAXIT, c);
        double t = S.d(St::F_T, c);
        const double ul = lane < NM ? S.d(St::F_U + lane, c) : 0.0;
        if (lane < NM) sh.al[lane] = S.d(St::F_AL + lane, c);
        __builtin_amdgcn_wave_barrier();
        bool newton = (mode == ST_NEWTON), do_reg = false, converged = false;
        double err = 0.0;
        if (mode == ST_TRIAL) {                                    // linesearch.cpp:45-70
            if (lane == 0) ++st.trials;
            const double fnew = S.d(St::F_SF, c) - wave_sum(lane < NM ? sh.al[lane] * ul : 0.0);
            const double fcur = S.d(St::F_FCUR, c), gd = S.d(St::F_GD, c);
            bool accept = fnew <= fcur + 1e-3 * t * gd;
            if (!accept && A.ls_mode == MNC_LS_ROBUST && t == 1.0 && -gd <= 1e3 * DBL_EPSILON * S.d(St::F_FSCALE, c))
                accept = true;
            __builtin_amdgcn_wave_barrier();
            if (accept) {
                newton = true;
            } else {
                t = 0.5 * t;
                if (t * S.d(St::F_MAXF, c) > DBL_EPSILON / 2) {
                    if (lane < NM) sh.al[lane] = fma(t, S.d(St::F_D + lane, c), S.d(St::F_AOLD + lane, c));
                } else {                                            // exhausted (:74-98)
                    if (lane < NM) sh.al[lane] = S.d(St::F_AOLD + lane, c);
                    t = 0.0;
                    ++iter; ++total;
                    do_reg = true;
                }
            }
            __builtin_amdgcn_wave_barrier();
        }
        if (newton) {
            ++iter; ++total;                                        // opt.cpp:212-213
            // the cell's moments into LDS first: each Gaunt term then reads its
            // moment from LDS, not through a load that waits on the term's index
            for (int k = lane; k < NK; k += 64) sh.mom[k] = S.d(St::F_ACC + k, c);
            __builtin_amdgcn_wave_barrier();
            double rs = 0.0, gi = 0.0, dg = 0.0;
            if (lane < NM) {                                        // row `lane` of H (Gaunt map)
                for (int jj = 0; jj < NM; ++jj) {
                    const int pr = (lane < jj) ? lane * NM + jj : jj * NM + lane;
                    double hij = 0.0;
                    // the same sum in the same order; four terms' loads in flight
#pragma unroll 4
                    for (int cc = A.mgptr[pr]; cc < A.mgptr[pr + 1]; ++cc)
                        hij = fma(A.mgv[cc], sh.mom[A.mgk[cc]], hij);
                    sh.L[lane][jj] = hij;
                    rs += fabs(hij);
                    if (jj == lane) dg = hij;
                }
                gi = S.d(St::F_GS + lane, c) - ul;
                sh.gvec[lane] = gi;
            }
            const double au = wave_sum(lane < NM ? sh.al[lane] * ul : 0.0);
            const double aus = wave_sum(lane < NM ? fabs(sh.al[lane] * ul) : 0.0);
            const double fcur = S.d(St::F_SF, c) - au;
            const double fscale = fabs(S.d(St::F_SF, c)) + aus;
            err = sqrt(wave_sum(gi * gi));                          // dnrm2 (:286)
            const double anorm = wave_max(rs);                     // utils.cpp:36-52
            const double gersh = wave_min(lane < NM ? 2.0 * dg - rs : DBL_MAX);
            __builtin_amdgcn_wave_barrier();
            const bool chol_ok = coop_cholesky<NM>(sh.L, lane);
            double condest = 0.0;
            if (chol_ok && !cond_proven_below(anorm, gersh, NM, A.cond_h_max))
                condest = 1.0 / coop_rcond<NM>(sh.L, sh.x, anorm, lane);
            if (iter >= maxit || !chol_ok || t == 0.0 || condest > A.cond_h_max) {
                do_reg = true;                                      // opt.cpp:244
            } else {
                double xr = lane < NM ? sh.gvec[lane] : 0.0;
                coop_apply_inv_reg<NM>(sh.L, xr, lane);             // dpotrs (:278-282)
                if (lane < NM) sh.x[lane] = xr;
                __builtin_amdgcn_wave_barrier();
                bool stop = err <= S.d(St::F_RTOL, c);
                if (stop && A.tol_gamma > 0.0) {
                    double da1 = 0.0;
                    for (int i = 0; i < NM; ++i) da1 = da1 + fabs(sh.x[i]);
                    stop = gamma_ok(A, da1, false);
                }
                if (stop) {
                    converged = true;                               // :293-306
                } else {                                            // linesearch.cpp:28-43
                    const double di = lane < NM ? -sh.x[lane] : 0.0;
                    const double ai = lane < NM ? sh.al[lane] : 0.0;
                    const double maxf = wave_max(lane < NM ? fabs(di) / (fabs(ai) + 1e-100) : 0.0);
                    const double gd = wave_sum(lane < NM ? sh.gvec[lane] * di : 0.0);
                    if (lane < NM) {
                        S.d(St::F_AOLD + lane, c) = ai;
                        S.d(St::F_D + lane, c) = di;
                    }
                    S.d(St::F_MAXF, c) = maxf;
                    S.d(St::F_GD, c) = gd;
                    S.d(St::F_FCUR, c) = fcur;
                    S.d(St::F_FSCALE, c) = fscale;
                    t = 1.0;
                    __builtin_amdgcn_wave_barrier();
                    if (t * maxf > DBL_EPSILON / 2) {
                        if (lane < NM) sh.al[lane] = ai + t * di;
                        mode = ST_TRIAL;
                    } else {
                        t = 0.0;
                        ++iter; ++total;
                        do_reg = true;
                    }
                }
            }
            __builtin_amdgcn_wave_barrier();
        }
        bool done = false;
        if (do_reg) {                                               // opt.cpp:246-273
            if (reg < 4) {
                ++reg;
                const double r = kReg[reg];
                const double u0 = S.d(St::F_U0, c);
                const double uk = (lane > 0 && lane < NM) ? (A.u[size_t(c) * NM + lane] / u0) * (1.0 - r) : 0.0;
                const double ss = 1.0 + wave_sum(uk * uk);
                if (lane < NM) {
                    S.d(St::F_U + lane, c) = lane == 0 ? 1.0 : uk;
                    sh.al[lane] = lane == 0 ? 2.0 * kSqrtPi * log(0.5 / kSqrtPi) : 0.0;
                }
                S.d(St::F_RTOL, c) = sqrt(ss) * A.tol;
                iter = 0;
                t = 1.0;
                if (reg == 4) maxit = 10 * maxit;
                mode = ST_NEWTON;
            } else {                                                // "Max regularization achieved."
                if (lane == 0) {
                    st.err_bits |= 1 << MNC_ERR_MAXREG;
                    if (A.outs) { A.outs[c].iter = total; A.outs[c].r = -1.0; A.outs[c].norm_g = err; }
                }
                done = true;
            }
        } else if (converged) {                                     // in-place reconstruction
            const double u0 = S.d(St::F_U0, c);
            const double u0p = S.d(St::F_U, c) + sh.gvec[0];
            if (lane < NM) {
                A.u[size_t(c) * NM + lane] = ((ul + sh.gvec[lane]) / u0p) * u0;
                const double ak = sh.al[lane];
                A.alpha[size_t(c) * NM + lane] =
                    lane == 0 ? (ak - log(u0p) * 2.0 * kSqrtPi) + S.d(St::F_ASCALE, c) : ak;
            }
            if (lane == 0) {
                if (A.outs) { A.outs[c].iter = total; A.outs[c].r = kReg[reg]; A.outs[c].norm_g = err; }
                st.max_iter = total > st.max_iter ? total : st.max_iter;
                st.sum += total;
                ++st.n;
                count_reg(st, reg);
            }
            done = true;
        }
        __builtin_amdgcn_wave_barrier();
        if (!done) {
            if (lane < NM) S.d(St::F_AL + lane, c) = sh.al[lane];
            if (lane == 0) {
                S.d(St::F_T, c) = t;
                S.i(St::I_ITER, c) = iter;
                S.i(St::I_TOTAL, c) = total;
                S.i(St::I_REG, c) = reg;
                S.i(St::I_MAXIT, c) = maxit;
                S.i(St::I_MODE, c) = mode;
                const int pos = atomicAdd(&A.ctl[lin ^ 1], 1);
                A.lists[size_t(lin ^ 1) * A.cap + pos] = c;
            }
        }
        __builtin_amdgcn_wave_barrier();
    }
    if (lane == 0 && (st.n || st.trials || st.err_bits)) {
        atomicMax((unsigned long long *)&A.stats[0], (unsigned long long)st.max_iter);
        atomicAdd((unsigned long long *)&A.stats[1], (unsigned long long)st.sum);
        atomicAdd((unsigned long long *)&A.stats[2], (unsigned long long)st.n);
        for (int r = 0; r < 5; ++r) atomicAdd((unsigned long long *)&A.stats[3 + r], (unsigned long long)st.hist[r]);
        atomicAdd((unsigned long long *)&A.stats[8], (unsigned long long)st.trials);
        if (st.err_bits) atomicOr(&A.ctl[3], st.err_bits);
    }
}

template <int NM, int NP2, bool G, int RS, bool PPN = false>
void launch_tail(const NewtonArgs &a, int lin, int lin2, int tail_blocks, int tail_mode, hipStream_t s) {
    if constexpr (!G && !PPN) {   // the cooperative tail at two waves per SIMD (direct Mn Hessian)
        if (tail_mode == 3) {
            hipLaunchKernelGGL((tail_coop_kernel<NM, NP2, G, RS, PPN, 2>), dim3(tail_blocks * 4 / kTailWaves),
                               dim3(64 * kTailWaves), 0, s, a, lin, lin2);
            return;
        }
    }
    if (tail_mode == 0 || tail_mode == 3 || PPN)   // PPn: the cooperative tail only
        hipLaunchKernelGGL((tail_coop_kernel<NM, NP2, G, RS, PPN>), dim3(tail_blocks * 4 / kTailWaves),
                           dim3(64 * kTailWaves), 0, s, a, lin, lin2);
    else if (tail_mode == 2)
        hipLaunchKernelGGL((tail_wave_kernel<NM, NP2, G, RS>), dim3(tail_blocks), dim3(kBlock), 0, s, a, lin, lin2);
    else
        hipLaunchKernelGGL((tail_kernel<NM, NP2, G, RS>), dim3(tail_blocks), dim3(kBlock), 0, s, a, lin, lin2);
}

// Rounds r0..r1-1 of the lockstep solve (round r reads list r & 1), with the
// prologue first if `init` and the persistent tail last if `tail`.
template <int NM, int NP2, bool G, int RS, bool PPN = false>
void launch_all(const NewtonArgs &a, const LaunchPlan &lp, hipStream_t s) {
    const int nb = (a.ncells + kBlock - 1) / kBlock;
    if (lp.init) hipLaunchKernelGGL((init_kernel<NM, NP2>), dim3(nb), dim3(kBlock), 0, s, a);
    int lin = lp.r0 & 1;
    for (int r = lp.r0; r < lp.r1; ++r) {
        hipLaunchKernelGGL((sweep_kernel<NM, NP2, G, RS, PPN>), dim3(nb), dim3(kBlock), 0, s, a, lin);
        hipLaunchKernelGGL((logic_kernel<NM, NP2, G, false, PPN>), dim3(nb), dim3(kBlock), 0, s, a, lin, 2);
        lin ^= 1;
    }
    if (lp.tail) launch_tail<NM, NP2, G, RS, PPN>(a, lin, 2, lp.tail_blocks, lp.tail_mode, s);   // list 2 empty
}

template <int NM, int NP2, bool G, int RS>
void launch_mixed_all(const NewtonArgs &a, const LaunchPlan &lp, hipStream_t s) {
    constexpr int NNT = mix_cols(NM) / 16;
    const int nb = (a.ncells + kBlock - 1) / kBlock;
    const int mix_blocks = lp.mix_blocks, defer_rounds = lp.defer_rounds;
    if (lp.init) hipLaunchKernelGGL((init_kernel<NM, NP2>), dim3(nb), dim3(kBlock), 0, s, a);
    // Mixed-precision cells on lists 0/1.  A cell whose step must be redone in
    // FP64 either waits on list 2 for the FP64 tail (defer_rounds 0) or moves
    // to the FP64 pair (lists 2/3), which the same rounds sweep with the FP64
    // lane kernels (defer_rounds 1); what is left after the rounds goes to the tail.
    int lin = lp.r0 & 1, flin = 2 + (defer_rounds ? (lp.r0 & 1) : 0);
    for (int r = lp.r0; r < lp.r1; ++r) {
        if (lp.mix_split == 3)
            hipLaunchKernelGGL((sweep_mix_kernel<NM, NP2, NNT, 3>), dim3(mix_blocks), dim3(64 * kMixWaves), 0, s, a, lin);
        else
            hipLaunchKernelGGL((sweep_mix_kernel<NM, NP2, NNT, 2>), dim3(mix_blocks), dim3(64 * kMixWaves), 0, s, a, lin);
        if (defer_rounds) {
            hipLaunchKernelGGL((sweep_kernel<NM, NP2, G, RS>), dim3(nb), dim3(kBlock), 0, s, a, flin);
            hipLaunchKernelGGL((logic_kernel<NM, NP2, G, true>), dim3(nb), dim3(kBlock), 0, s, a, lin, flin ^ 1);
            hipLaunchKernelGGL((logic_kernel<NM, NP2, G>), dim3(nb), dim3(kBlock), 0, s, a, flin, 2);
            flin ^= 1;
        } else {
            hipLaunchKernelGGL((logic_kernel<NM, NP2, G, true>), dim3(nb), dim3(kBlock), 0, s, a, lin, 2);
        }
        lin ^= 1;
    }
    if (lp.tail) launch_tail<NM, NP2, G, RS>(a, lin, flin, lp.tail_blocks, lp.tail_mode, s);
}

}  // namespace

int newton_block() { return kBlock; }
int newton_list_word(int l) { return l < 2 ? l : l + 2; }   // cw() on the host

__global__ void ctl_reset_kernel(int *ctl) {
    if (threadIdx.x < kCtlWords) ctl[threadIdx.x] = 0;
}

void launch_ctl_reset(int *ctl, hipStream_t s) { hipLaunchKernelGGL(ctl_reset_kernel, dim3(1), dim3(64), 0, s, ctl); }

void newton_store_fields(int nm, int np2, int *dfields, int *ifields) {
    const int tri = nm * (nm + 1) / 2;
    *dfields = 5 * nm + 10 + cmax(cmax(np2, tri), mix_cols(nm));
    *ifields = 6;
}

bool launch_newton_fp64(const NewtonArgs &a, int nm, int np2, bool gaunt, const LaunchPlan &lp, hipStream_t s) {
#define MNC_CASE(NMV, NP2D, NP2G, RSPLIT)                                                             \
    if (nm == NMV) {                                                                                   \
        static_assert(Store<NMV, NP2G>::dfields == 5 * NMV + 10 + Store<NMV, NP2G>::nacc, "store");    \
        if (gaunt && np2 == NP2G) {                                                                    \
            if (a.ppn) launch_all<NMV, NP2G, true, NMV, true>(a, lp, s);                               \
            else launch_all<NMV, NP2G, true, NMV>(a, lp, s);                                           \
            return true;                                                                               \
        }                                                                                              \
        if (!gaunt && np2 == NP2D) {                                                                   \
            if (a.ppn) launch_all<NMV, NP2D, false, RSPLIT, true>(a, lp, s);                           \
            else launch_all<NMV, NP2D, false, RSPLIT>(a, lp, s);                                       \
            return true;                                                                               \
        }                                                                                              \
        return false;                                                                                  \
    }
    MNC_CASE(1, 1, 1, 1)      // M0
    MNC_CASE(3, 3, 6, 3)      // M1
    MNC_CASE(6, 6, 15, 6)     // M2
    MNC_CASE(10, 10, 28, 10)  // M3
    MNC_CASE(15, 15, 45, 10)  // M4: direct Hessian in two sweeps (55 + 65 entries)
#undef MNC_CASE
    return false;
}

bool launch_newton_mixed(const NewtonArgs &a, int nm, int np2, bool gaunt, const LaunchPlan &lp, hipStream_t s) {
#define MNC_CASE(NMV, NP2D, NP2G, RSPLIT)                                                                \
    if (nm == NMV) {                                                                                      \
        if (a.mnt < 2 || a.mnt > kMixMaxTiles || (a.mnt & 1)) return false;                              \
        if (gaunt && np2 == NP2G) { launch_mixed_all<NMV, NP2G, true, NMV>(a, lp, s); return true; }         \
        if (!gaunt && np2 == NP2D) { launch_mixed_all<NMV, NP2D, false, RSPLIT>(a, lp, s); return true; }    \
        return false;                                                                                     \
    }
    MNC_CASE(1, 1, 1, 1)
    MNC_CASE(3, 3, 6, 3)
    MNC_CASE(6, 6, 15, 6)
    MNC_CASE(10, 10, 28, 10)
    MNC_CASE(15, 15, 45, 10)
#undef MNC_CASE
    return false;
}

int mix_tiles(int nm) { return mix_cols(nm) / 16; }

bool launch_newton_g64_init(const NewtonArgs &a, int nm, int np2, hipStream_t s) {
    const int nb = (a.ncells + kBlock - 1) / kBlock;
    if (nm == 45 && np2 == 45) {
        hipLaunchKernelGGL((init_kernel<45, 45>), dim3(nb), dim3(kBlock), 0, s, a);
        return true;
    }
    return false;
}

bool launch_newton_g64_rounds(const NewtonArgs &a, int nm, int np2, int &lin, int rounds, int blocks, hipStream_t s) {
    if (!(nm == 45 && np2 == 45)) return false;
    constexpr int KZ = 12, NGT = 3, NHT = 10;   // 48 moment rows, 46 sum columns, 153 order-16 moments
    if (a.gnmm > 16 * NHT || a.gnmm != gp_nm(2 * order_of(45))) return false;   // logic_g64's LDS moments
    for (int r = 0; r < rounds; ++r) {
        hipLaunchKernelGGL((sweep_g64_kernel<45, 45, KZ, NGT, NHT>), dim3(blocks), dim3(256), 0, s, a, lin);
        hipLaunchKernelGGL((logic_g64_kernel<45, 45>), dim3(4 * blocks), dim3(64), 0, s, a, lin);
        lin ^= 1;
    }
    return true;
}

// ---- the no-allocation entry point's scratch ------------------------------------
//
// One file-scope device array per GPU holds, for up to kNoAllocCells cells,
// what mnc_opt_batch keeps in the context's lazily grown buffers: the per-cell
// Newton state (Store, column stride kNoAllocCells), the kNumLists active lists
// and the control words.  Sized for the largest order it serves (M4, both
// Hessian forms have the same store).  Host code reaches it only through the
// address hipGetSymbolAddress returns on each device; the array's name on the
// host is not a device address.
namespace {
constexpr size_t align256(size_t b) { return (b + 255) / 256 * 256; }
constexpr size_t noalloc_bytes(int dfields, int ifields) {
    return align256(size_t(dfields) * kNoAllocCells * sizeof(double)) +
           align256(size_t(ifields) * kNoAllocCells * sizeof(int)) +
           align256(size_t(kNumLists) * kNoAllocCells * sizeof(int)) + align256(kCtlWords * sizeof(int));
}
static_assert(Store<15, 15>::dfields == Store<15, 45>::dfields, "M4 stores");
constexpr size_t kNoAllocBytes = noalloc_bytes(Store<15, 45>::dfields, Store<15, 45>::ifields);
}  // namespace

static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

size_t noalloc_scratch_bytes() { return kNoAllocBytes; }

bool noalloc_layout(int nm, int np2, void *base, NoAllocState &out) {
    int df = 0, ifl = 0;
    newton_store_fields(nm, np2, &df, &ifl);
    if (noalloc_bytes(df, ifl) > kNoAllocBytes) return false;
    unsigned char *p = static_cast<unsigned char *>(base);
    out.st_d = reinterpret_cast<double *>(p);
    p += align256(size_t(df) * kNoAllocCells * sizeof(double));
    out.st_i = reinterpret_cast<int *>(p);
    p += align256(size_t(ifl) * kNoAllocCells * sizeof(int));
    out.lists = reinterpret_cast<int *>(p);
    p += align256(size_t(kNumLists) * kNoAllocCells * sizeof(int));
    out.ctl = reinterpret_cast<int *>(p);
    out.cap = kNoAllocCells;
    return true;
}

hipError_t noalloc_scratch_address(int device, void **ptr) {
    constexpr int kMaxDevices = 64;
    static std::mutex mu;
    static void *addr[kMaxDevices] = {};
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(mu);
    if (!addr[device]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        addr[device] = p;
    }
    *ptr = addr[device];
    return hipSuccess;
}

bool launch_newton_net_judge(const NewtonArgs &a, int nm, int np2, bool gaunt, int max_rows, int blocks,
                             hipStream_t s) {
    if (a.nq > kNetMaxQ) return false;
    const int b = std::max(1, std::min(blocks, (max_rows + 16 * kNetWaves - 1) / (16 * kNetWaves)));
    const dim3 blk(64 * kNetWaves);
    if (nm == 15 && !gaunt && np2 == 15) hipLaunchKernelGGL((net_judge_kernel<15, 15>), dim3(b), blk, 0, s, a);
    else if (nm == 15 && gaunt && np2 == 45) hipLaunchKernelGGL((net_judge_kernel<15, 45>), dim3(b), blk, 0, s, a);
    else return false;
    return true;
}

bool launch_newton_judge(const NewtonArgs &a, int nm, int np2, bool gaunt, hipStream_t s) {
    const int nb = (a.ncells + kBlock - 1) / kBlock;
    if (nm == 15 && !gaunt && np2 == 15) hipLaunchKernelGGL((judge_kernel<15, 15>), dim3(nb), dim3(kBlock), 0, s, a);
    else if (nm == 15 && gaunt && np2 == 45) hipLaunchKernelGGL((judge_kernel<15, 45>), dim3(nb), dim3(kBlock), 0, s, a);
    else return false;
    return true;
}

}  // namespace mnc

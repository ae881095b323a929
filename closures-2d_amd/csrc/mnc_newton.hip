// mnc_newton.hip -- batched FP64 Newton solve of the Mn dual problem on gfx950.
//
// One lane owns one cell at a time (thread-per-cell).  Each wave is a
// persistent worker: lanes whose cell has converged pull the next cell index
// from an on-device work queue (one atomic per wave, wave-aggregated), so cells
// needing 30+ iterations near the realizability boundary do not hold back the
// ~2-iteration interior cells that share their wave.  All lanes of a wave sweep
// the quadrature together (uniform q loop: the table rows are wave-uniform loads),
// each evaluating its own cell at its own point -- the Newton iterate for lanes
// about to take a Newton step, or the line-search trial point for lanes inside
// the Armijo loop.  One sweep yields both the trial's f and the full Hessian at
// the trial point, so an accepted trial goes straight to its Newton step with
// no second sweep (the reference re-evaluates at the same alpha, giving the same
// numbers).
//
// The per-cell algorithm restates src/momopt/opt/opt.cpp:82-343 (opt,
// optScaled), src/momopt/opt/linesearch.cpp:21-99 and fobj.cpp:37-129,255-335,
// with the LAPACK roles (dpotrf 'L', dpocon, dpotrs, dnrm2) done in registers:
// unblocked column Cholesky, Higham's 1-norm estimator (ACM TOMS 14(4), 1988),
// forward/back substitution.
#include <cfloat>

#include "mnc_internal.h"

namespace mnc {
namespace {

constexpr int kBlock = 256;
#define kSqrtPi sqrt(M_PI)
__constant__ double kReg[5] = {0.0, 1e-3, 1e-2, 1e-1, 5e-1};   // opt.h:39-40

template <int N> struct Tri {
    static constexpr int size = N * (N + 1) / 2;
    __device__ static constexpr int at(int i, int j) { return i * (i + 1) / 2 + j; }   // i >= j
};

// ---- dense kernels on a packed lower triangle in registers ----------------

// dpotf2-order Cholesky in place; returns false on a non-positive (or NaN) pivot.
template <int N>
__device__ __forceinline__ bool cholesky(double (&a)[Tri<N>::size]) {
    bool ok = true;
#pragma unroll
    for (int j = 0; j < N; ++j) {
        double dot = 0.0;
#pragma unroll
        for (int k = 0; k < j; ++k) dot = fma(a[Tri<N>::at(j, k)], a[Tri<N>::at(j, k)], dot);
        double ajj = a[Tri<N>::at(j, j)] - dot;
        ok = ok && (ajj > 0.0);
        ajj = sqrt(ajj > 0.0 ? ajj : 1.0);
        a[Tri<N>::at(j, j)] = ajj;
#pragma unroll
        for (int i = j + 1; i < N; ++i) {
            double y = a[Tri<N>::at(i, j)];
#pragma unroll
            for (int k = 0; k < j; ++k) y = fma(-a[Tri<N>::at(j, k)], a[Tri<N>::at(i, k)], y);
            a[Tri<N>::at(i, j)] = y;
        }
        const double r = 1.0 / ajj;
#pragma unroll
        for (int i = j + 1; i < N; ++i) a[Tri<N>::at(i, j)] *= r;
    }
    return ok;
}

// x <- (L L^T)^{-1} x
template <int N>
__device__ __forceinline__ void apply_inv(const double (&l)[Tri<N>::size], double (&x)[N]) {
#pragma unroll
    for (int j = 0; j < N; ++j) {
        x[j] = x[j] / l[Tri<N>::at(j, j)];
#pragma unroll
        for (int i = j + 1; i < N; ++i) x[i] = fma(-x[j], l[Tri<N>::at(i, j)], x[i]);
    }
#pragma unroll
    for (int j = N - 1; j >= 0; --j) {
        double t = x[j];
#pragma unroll
        for (int i = N - 1; i > j; --i) t = fma(-l[Tri<N>::at(i, j)], x[i], t);
        x[j] = t / l[Tri<N>::at(j, j)];
    }
}

template <int N>
__device__ __forceinline__ double dasum(const double (&x)[N]) {
    double s = 0.0;
#pragma unroll
    for (int i = 0; i < N; ++i) s += fabs(x[i]);
    return s;
}

template <int N>
__device__ __forceinline__ int idamax(const double (&x)[N]) {
    int idx = 0;
    double m = fabs(x[0]);
#pragma unroll
    for (int i = 1; i < N; ++i) {
        const double v = fabs(x[i]);
        if (v > m) { m = v; idx = i; }
    }
    return idx;
}

// Reciprocal condition estimate of A = L L^T in the 1-norm (the role of
// dpocon): Higham's estimate of ||A^{-1}||_1, with a single call site of the
// A^{-1} product inside a small state machine to keep the unrolled code short.
template <int N>
__device__ double rcond(const double (&l)[Tri<N>::size], double anorm) {
    if (anorm == 0.0) return 0.0;
    double x[N];
    unsigned isgn = 0;   // bit i set <=> isgn_i = -1
    double est = 0.0;
    int jump = 1, j = 0, iter = 0;
#pragma unroll
    for (int i = 0; i < N; ++i) x[i] = 1.0 / N;
    for (int guard = 0; guard < 16; ++guard) {
        apply_inv<N>(l, x);
        if (jump == 1) {                       // first product
            if (N == 1) { est = fabs(x[0]); break; }
            est = dasum<N>(x);
            isgn = 0;
#pragma unroll
            for (int i = 0; i < N; ++i) {
                const bool neg = !(x[i] >= 0.0);
                x[i] = neg ? -1.0 : 1.0;
                isgn |= neg ? (1u << i) : 0u;
            }
            jump = 2;
        } else if (jump == 2) {                // after the transposed product
            j = idamax<N>(x);
            iter = 2;
#pragma unroll
            for (int i = 0; i < N; ++i) x[i] = (i == j) ? 1.0 : 0.0;
            jump = 3;
        } else if (jump == 3) {                // x = A^{-1} e_j
            const double estold = est;
            est = dasum<N>(x);
            unsigned s = 0;
#pragma unroll
            for (int i = 0; i < N; ++i) s |= !(x[i] >= 0.0) ? (1u << i) : 0u;
            bool done = (s == isgn) || (est <= estold);
            if (!done) {
#pragma unroll
                for (int i = 0; i < N; ++i) x[i] = (s >> i & 1u) ? -1.0 : 1.0;
                isgn = s;
                jump = 4;
            } else {
                double alt = 1.0;
#pragma unroll
                for (int i = 0; i < N; ++i) { x[i] = alt * (1.0 + double(i) / double(N - 1)); alt = -alt; }
                jump = 5;
            }
        } else if (jump == 4) {                // after the transposed product
            const int jlast = j;
            j = idamax<N>(x);
            double xjl = 0.0, xj = 0.0;
#pragma unroll
            for (int i = 0; i < N; ++i) { xjl = (i == jlast) ? x[i] : xjl; xj = (i == j) ? x[i] : xj; }
            if (xjl != fabs(xj) && iter < 5) {
                ++iter;
#pragma unroll
                for (int i = 0; i < N; ++i) x[i] = (i == j) ? 1.0 : 0.0;
                jump = 3;
            } else {
                double alt = 1.0;
#pragma unroll
                for (int i = 0; i < N; ++i) { x[i] = alt * (1.0 + double(i) / double(N - 1)); alt = -alt; }
                jump = 5;
            }
        } else {                               // jump == 5: final alternating vector
            const double temp = 2.0 * (dasum<N>(x) / double(3 * N));
            if (temp > est) est = temp;
            break;
        }
    }
    if (!isfinite(est) || est == 0.0) return 0.0;
    return (1.0 / est) / anorm;
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

enum : int { ST_NEWTON = 0, ST_TRIAL = 1 };

// Per-slot scratch fields.  Layout [wave][field][64 lanes]: a wave's accesses
// to one field are 512 contiguous bytes, and the address is a wave-uniform base
// (SGPRs) plus the lane offset, so no per-field address registers stay live.
template <int NM, int NP2>
struct Scratch {
    // u, aold, d, g, then the accumulators: the Gaunt moments (NP2) or the
    // packed lower triangle of the direct Hessian (NM(NM+1)/2), whichever is larger
    static constexpr int nacc = NP2 > NM * (NM + 1) / 2 ? NP2 : NM * (NM + 1) / 2;
    static constexpr int fields = 4 * NM + nacc;
    double *wb;      // wave base (uniform)
    int lane;
    __device__ double &at(int f) const { return wb[f * 64 + lane]; }
    __device__ double &u(int k) const { return at(k); }
    __device__ double &aold(int k) const { return at(NM + k); }
    __device__ double &d(int k) const { return at(2 * NM + k); }
    __device__ double &g(int k) const { return at(3 * NM + k); }
    __device__ double &acc(int k) const { return at(4 * NM + k); }
};

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

// ---- the persistent Newton kernel ----------------------------------------
// RS: for the direct form, rows [0,RS) are accumulated in the first sweep and
// rows [RS,NM) in a second one (RS == NM: a single sweep).  Splitting keeps the
// M4 accumulators (120 doubles) out of spill space.
template <int NM, int NP2, bool GAUNT, int RS>
__global__ __launch_bounds__(kBlock) void newton_fp64_kernel(NewtonArgs A) {
    constexpr int NA = GAUNT ? NP2 : Tri<NM>::at(RS, 0);            // first-sweep accumulators
    constexpr int NB = (GAUNT || RS >= NM) ? 1 : Tri<NM>::size - Tri<NM>::at(RS, 0);
    constexpr bool SPLIT = !GAUNT && RS < NM;
    static_assert(NA <= Scratch<NM, NP2>::nacc, "accumulators exceed the scratch field set");
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6));
    const Scratch<NM, NP2> S{A.scratch + size_t(wave) * Scratch<NM, NP2>::fields * 64, lane};
    const double p0 = A.p[0];

    int cell = -1;           // -1 wants work, -2 finished, >= 0 working
    int state = ST_NEWTON, iter = 0, total = 0, reg = 0, max_iter = A.max_iter;
    double t = 1.0, rtol = 0.0, u0 = 1.0, ascale = 0.0, fcur = 0.0, gd = 0.0, maxf = 0.0;
    double al[NM];           // evaluation point (Newton iterate or line-search trial)
    int st_max = 0;
    long long st_sum = 0, st_n = 0, st_hist[5] = {0, 0, 0, 0, 0}, st_trials = 0;
    int err_bits = 0;

    for (;;) {
        // -------- refill lanes from the work queue --------
        const bool need = (cell == -1);
        const unsigned long long m = __ballot(need);
        if (m) {
            const int leader = __ffsll((long long)m) - 1;
            int base = 0;
            if (lane == leader) base = atomicAdd(&A.queue[0], __popcll(m));
            base = __shfl(base, leader);
            if (need) {
                const int c = base + __popcll(m & lane_mask_lt());
                cell = (c < A.ncells) ? c : -2;
                if (cell >= 0) {
                    // opt() prologue (opt.cpp:87-101) + optScaled set-up (:170-205)
                    const double *uin = A.u + size_t(cell) * NM;
                    const double *ain = A.alpha + size_t(cell) * NM;
                    u0 = uin[0];
                    ascale = log(u0) * 2.0 * kSqrtPi;
                    double ss = 0.0;
#pragma unroll
                    for (int k = 0; k < NM; ++k) {
                        const double uk = uin[k] / u0;
                        S.u(k) = uk;
                        ss = fma(uk, uk, ss);
                        al[k] = ain[k];
                    }
                    if (A.warm == MNC_WARM_ISOTROPIC) {
                        al[0] = 2.0 * kSqrtPi * log(0.5 / kSqrtPi);
#pragma unroll
                        for (int k = 1; k < NM; ++k) al[k] = 0.0;
                    } else if (A.warm == MNC_WARM_LEARNED && A.alpha_guess) {
                        const double *ag = A.alpha_guess + size_t(cell) * NM;
#pragma unroll
                        for (int k = 0; k < NM; ++k) al[k] = ag[k];
                    } else {
                        al[0] = al[0] - ascale;
                    }
                    if (!(u0 > 0.0)) {
                        err_bits |= 1 << MNC_ERR_U0;
                        if (A.outs) { A.outs[cell].iter = -1; A.outs[cell].r = 0.0; A.outs[cell].norm_g = 0.0; }
                        cell = -1;
                    } else {
                        rtol = sqrt(ss) * A.tol;
                        iter = 0; total = 0; reg = 0; max_iter = A.max_iter; t = 1.0;
                        state = ST_NEWTON;
                    }
                }
            }
            continue;   // re-ballot: lanes that failed the prologue want work again
        }
        if (__ballot(cell >= 0) == 0) break;   // every lane finished

        // -------- first quadrature sweep at each lane's evaluation point --------
        double sf = 0.0;
        {
            double acc[NA];
            if (cell >= 0) {
                sf = sweep<NM, NP2, GAUNT, 0, (GAUNT ? NM : RS), NA>(A.p, A.w, A.nq, al, acc);
#pragma unroll
                for (int k = 0; k < NA; ++k) S.acc(k) = acc[k];
            }
        }
        if (cell < 0) continue;

        // -------- Armijo test for lanes inside the line search (linesearch.cpp:45-70) --------
        bool newton = (state == ST_NEWTON);
        bool do_reg = false;
        if (state == ST_TRIAL) {
            ++st_trials;
            double fnew = sf;
#pragma unroll
            for (int k = 0; k < NM; ++k) fnew = fnew - al[k] * S.u(k);
            if (fnew <= fcur + 1e-3 * t * gd) {
                newton = true;                                     // accepted: al is the new iterate
            } else {
                t = 0.5 * t;
                if (t * maxf > DBL_EPSILON / 2) {
#pragma unroll
                    for (int k = 0; k < NM; ++k) al[k] = S.aold(k) + t * S.d(k);
                } else {                                           // exhausted: t = 0 (:74-98)
#pragma unroll
                    for (int k = 0; k < NM; ++k) al[k] = S.aold(k);
                    t = 0.0;
                    ++iter; ++total;
                    do_reg = true;                                 // opt.cpp:244 (t == 0)
                }
            }
        }

        // -------- second sweep (split direct Hessian) where a Newton step follows --------
        double accb[NB];
        if constexpr (SPLIT) {
            if (__ballot(newton) != 0) {
                if (newton) sweep<NM, NP2, false, RS, NM, NB>(A.p, A.w, A.nq, al, accb);
            }
        }
        double err = 0.0;
        bool converged = false;
        if (newton) {
            ++iter; ++total;                                       // opt.cpp:212-213
            double h[Tri<NM>::size];
            if constexpr (GAUNT) {                                 // fobj.cpp:99-113
#pragma unroll
                for (int i = 0; i < NM; ++i) {
#pragma unroll
                    for (int j = i; j < NM; ++j) {
                        double hij = 0.0;
                        const int pr = i * NM + j;
                        for (int c = A.gptr[pr]; c < A.gptr[pr + 1]; ++c) hij = fma(A.gv[c], S.acc(A.gk[c]), hij);
                        h[Tri<NM>::at(j, i)] = hij;
                    }
                }
            } else {
#pragma unroll
                for (int k = 0; k < NA; ++k) h[k] = S.acc(k);
                if constexpr (SPLIT) {
#pragma unroll
                    for (int k = 0; k < NB; ++k) h[NA + k] = accb[k];
                }
            }
            // g and f from H (fobj.cpp:270-286)
            fcur = h[0] / (p0 * p0);
            double gg = 0.0;
#pragma unroll
            for (int i = 0; i < NM; ++i) {
                const double ui = S.u(i);
                const double gi = h[Tri<NM>::at(i, 0)] / p0 - ui;
                S.g(i) = gi;
                gg = fma(gi, gi, gg);
                fcur = fcur - al[i] * ui;
            }
            // 1-norm = max absolute row sum of the symmetric H (utils.cpp:36-52)
            double anorm = 0.0;
#pragma unroll
            for (int i = 0; i < NM; ++i) {
                double s = 0.0;
#pragma unroll
                for (int j = 0; j < NM; ++j) s += fabs(h[i >= j ? Tri<NM>::at(i, j) : Tri<NM>::at(j, i)]);
                anorm = s > anorm ? s : anorm;
            }
            const bool chol_ok = cholesky<NM>(h);
            double condest = 0.0;
            if (chol_ok) condest = 1.0 / rcond<NM>(h, anorm);
            if (iter >= max_iter || !chol_ok || t == 0.0 || condest > A.cond_h_max) {
                do_reg = true;                                      // opt.cpp:244
            } else {
                double d[NM];
#pragma unroll
                for (int i = 0; i < NM; ++i) d[i] = S.g(i);
                apply_inv<NM>(h, d);                                // dpotrs (:278-282)
                err = sqrt(gg);                                     // dnrm2 (:286)
                if (err <= rtol) {
                    converged = true;                               // :293-306
                } else {                                            // set up the line search
                    maxf = 0.0;
                    gd = 0.0;
#pragma unroll
                    for (int i = 0; i < NM; ++i) {
                        const double di = -d[i];
                        const double tmp = fabs(di) / (fabs(al[i]) + 1e-100);
                        maxf = tmp > maxf ? tmp : maxf;
                        gd = fma(S.g(i), di, gd);
                        S.aold(i) = al[i];
                        S.d(i) = di;
                    }
                    t = 1.0;
                    if (t * maxf > DBL_EPSILON / 2) {
#pragma unroll
                        for (int i = 0; i < NM; ++i) al[i] = al[i] + t * S.d(i);
                        state = ST_TRIAL;
                    } else {
                        // the Armijo loop cannot start: linesearch returns 0 and
                        // the next iteration regularizes (opt.cpp:212-213,244)
                        t = 0.0;
                        ++iter; ++total;
                        do_reg = true;
                    }
                }
            }
        }
        if (do_reg) {                                               // opt.cpp:246-273
            if (reg < 4) {
                ++reg;
                const double r = kReg[reg];
                const double *uin = A.u + size_t(cell) * NM;
                double ss = 1.0;
                S.u(0) = 1.0;
#pragma unroll
                for (int k = 1; k < NM; ++k) {
                    const double uk = (uin[k] / u0) * (1.0 - r);
                    S.u(k) = uk;
                    ss = fma(uk, uk, ss);
                }
                al[0] = 2.0 * kSqrtPi * log(0.5 / kSqrtPi);
#pragma unroll
                for (int k = 1; k < NM; ++k) al[k] = 0.0;
                rtol = sqrt(ss) * A.tol;
                iter = 0;
                t = 1.0;
                if (reg == 4) max_iter = 10 * max_iter;
                state = ST_NEWTON;
            } else {                                                // "Max regularization achieved."
                err_bits |= 1 << MNC_ERR_MAXREG;
                if (A.outs) { A.outs[cell].iter = total; A.outs[cell].r = -1.0; A.outs[cell].norm_g = err; }
                cell = -1;
            }
        } else if (converged) {
            const double u0p = S.u(0) + S.g(0);
            double *uo = A.u + size_t(cell) * NM;
            double *ao = A.alpha + size_t(cell) * NM;
#pragma unroll
            for (int k = 0; k < NM; ++k) uo[k] = ((S.u(k) + S.g(k)) / u0p) * u0;   // :298-301, opt.cpp:107-108
            al[0] -= log(u0p) * 2.0 * kSqrtPi;                                      // :302
            al[0] = al[0] + ascale;                                                 // opt.cpp:109
#pragma unroll
            for (int k = 0; k < NM; ++k) ao[k] = al[k];
            if (A.outs) { A.outs[cell].iter = total; A.outs[cell].r = kReg[reg]; A.outs[cell].norm_g = err; }
            st_max = total > st_max ? total : st_max;
            st_sum += total;
            ++st_n;
            ++st_hist[reg];
            cell = -1;
        }
    }

    // -------- flush statistics: one atomic per wave --------
    for (int off = 32; off > 0; off >>= 1) {
        st_max = max(st_max, __shfl_xor(st_max, off));
        st_sum += __shfl_xor(st_sum, off);
        st_n += __shfl_xor(st_n, off);
        st_trials += __shfl_xor(st_trials, off);
        err_bits |= __shfl_xor(err_bits, off);
#pragma unroll
        for (int r = 0; r < 5; ++r) st_hist[r] += __shfl_xor(st_hist[r], off);
    }
    if (lane == 0) {
        atomicMax((unsigned long long *)&A.stats[0], (unsigned long long)st_max);
        atomicAdd((unsigned long long *)&A.stats[1], (unsigned long long)st_sum);
        atomicAdd((unsigned long long *)&A.stats[2], (unsigned long long)st_n);
#pragma unroll
        for (int r = 0; r < 5; ++r) atomicAdd((unsigned long long *)&A.stats[3 + r], (unsigned long long)st_hist[r]);
        atomicAdd((unsigned long long *)&A.stats[8], (unsigned long long)st_trials);
        if (err_bits) atomicOr(&A.queue[1], err_bits);
    }
}

template <int NM, int NP2, bool G, int RS>
void launch(const NewtonArgs &a, int nblocks, hipStream_t s) {
    hipLaunchKernelGGL((newton_fp64_kernel<NM, NP2, G, RS>), dim3(nblocks), dim3(kBlock), 0, s, a);
}

}  // namespace

int newton_slots_per_block() { return kBlock; }

int newton_scratch_fields(int nm, int np2) {
    const int tri = nm * (nm + 1) / 2;
    return 4 * nm + (np2 > tri ? np2 : tri);
}

bool launch_newton_fp64(const NewtonArgs &a, int nm, int np2, bool gaunt, int nblocks, hipStream_t s) {
#define MNC_CASE(NMV, NP2D, NP2G, RSPLIT)                                                    \
    if (nm == NMV) {                                                                          \
        if (gaunt && np2 == NP2G) { launch<NMV, NP2G, true, NMV>(a, nblocks, s); return true; }   \
        if (!gaunt && np2 == NP2D) { launch<NMV, NP2D, false, RSPLIT>(a, nblocks, s); return true; } \
        return false;                                                                         \
    }
    MNC_CASE(1, 1, 1, 1)      // M0
    MNC_CASE(3, 3, 6, 3)      // M1
    MNC_CASE(6, 6, 15, 6)     // M2
    MNC_CASE(10, 10, 28, 10)  // M3
    MNC_CASE(15, 15, 45, 10)  // M4: direct Hessian in two sweeps (55 + 65 entries)
#undef MNC_CASE
    return false;
}

}  // namespace mnc

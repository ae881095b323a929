// mnc_grid.hip -- ansatz, kinetic flux, Euler/Heun stage and ghost fill of the
// Mn solver on gfx950 (src/momopt/momopt_update.cpp:38-202,283-382 and
// src/momopt/momopt_boundaries.cpp:190-223).
//
// Data layout in HBM: grid arrays are the reference's I3D layout
// [(i*(ny+4)+j)][k] (x-major, 2 ghost cells per side); the ansatz is stored
// [q][cell] so that, for one quadrature direction, neighbouring threads (cells)
// read neighbouring words.
#include "mnc_internal.h"

namespace mnc {
namespace {

// The left edge x0 + k h of a cell for the lattice source test |x| < 0.5
// (momopt_update.cpp:328-334), rounded as a product and a sum, as the
// reference's x86-64 build without FMA evaluates it: on grids whose edges
// fall on +-0.5 exactly (280 cells on [-3.5, 3.5]: h = 0.025) a contracted
// fma lands inside and adds a row and a column of source cells.
// (__dadd_rn / __dmul_rn alone do not stop the contraction: the pragma does.)
__device__ __forceinline__ double edge_coord(double x0, int k, double h) {
#pragma clang fp contract(off)
    return x0 + double(k) * h;
}

constexpr int kBlock = 256;

// momopt_update.cpp:77-90 (SIGN/MIN/MAX of utils.h:28-30)
__device__ __forceinline__ double sgn1(double x) { return (x >= 0.0) ? 1.0 : -1.0; }
__device__ __forceinline__ double minmod(double x, double y) {
    const double a = fabs(x), b = y * sgn1(x);
    const double mn = (a < b) ? a : b;
    return sgn1(x) * ((0.0 > mn) ? 0.0 : mn);
}
__device__ __forceinline__ double slopefit(double l, double c, double r, double theta) {
    return minmod(theta * (r - c), minmod(0.5 * (r - l), theta * (c - l)));
}

// The Euler stage of interior cell (i, j) = c with flux F (:316-336 / :344-364),
// the lattice source (:328-334) and, on the second stage, the Heun average
// (:368-377); stage 0 stores the flux only.
template <int NM>
__device__ __forceinline__ void euler_epilogue(const GridArgs &g, int i, int j, size_t c, const double (&F)[NM],
                                               int stage, double dt) {
    if (stage == 0) {
#pragma unroll
        for (int k = 0; k < NM; ++k) g.flux[c * NM + k] = F[k];
        return;
    }
    double *u = g.moments + c * NM;
    const double st = g.sigma_t[c], ss = g.sigma_s[c];
    double v[NM];
    v[0] = u[0] * (1.0 + dt * ss - dt * st) - dt * F[0];
#pragma unroll
    for (int k = 1; k < NM; ++k) v[k] = u[k] * (1.0 - dt * st) - dt * F[k];
    if (g.init_cond == 1) {
        const double x_i = edge_coord(g.init_x, i - 2, g.dx);
        const double y_j = edge_coord(g.init_y, j - 2, g.dy);
        if (fabs(x_i) < 0.5 && fabs(y_j) < 0.5) v[0] = v[0] + dt * 2.0 * sqrt(M_PI);
    }
    if (stage == 2) {
        const double *uo = g.moments_old + c * NM;
#pragma unroll
        for (int k = 0; k < NM; ++k) v[k] = 0.5 * (v[k] + uo[k]);
    }
#pragma unroll
    for (int k = 0; k < NM; ++k) u[k] = v[k];
}

// The ansatz at kinetic = alpha.m_q (computeAnsatz, :38-58): exp for Mn; for
// PPn kinetic/2 + sqrt(kinetic^2 + 4 delta)/2, written for kinetic <= 0 in
// the form without cancellation.
__device__ __forceinline__ double ansatz_value(const GridArgs &g, double kinetic) {
    if (g.ppn) {
        if (kinetic > 0) return 0.5 * kinetic + 0.5 * sqrt(kinetic * kinetic + 4 * g.delta);
        return -g.delta / (0.5 * kinetic - 0.5 * sqrt(kinetic * kinetic + 4 * g.delta));
    }
    return exp(kinetic);
}

// E[q][cell] = ansatz(alpha_cell . m_q) for q in [q0, q0+nqc)  (computeAnsatz, :38-58)
template <int NM>
__global__ __launch_bounds__(kBlock) void ansatz_kernel(GridArgs g, int q0, int nqc) {
    const int ncell = g.NX * g.NY;
    const int c = blockIdx.x * kBlock + threadIdx.x;
    if (c >= ncell) return;
    double a[NM];
#pragma unroll
    for (int k = 0; k < NM; ++k) a[k] = g.alpha[size_t(c) * NM + k];
    for (int qq = 0; qq < nqc; ++qq) {
        const double *row = g.p + size_t(q0 + qq) * g.np2;
        double s = 0.0;
#pragma unroll
        for (int k = 0; k < NM; ++k) s = fma(a[k], row[k], s);
        g.ansatz[size_t(qq) * ncell + c] = ansatz_value(g, s);
    }
}

// Interior flux (solveFlux :128-196) accumulated over q in [q0, q0+nqc) in
// ascending q, then (last chunk) the Euler stage (:316-336 / :344-364) and, on
// the second stage, the Heun average (:368-377).
template <int NM>
__global__ __launch_bounds__(kBlock) void flux_euler_kernel(GridArgs g, int q0, int nqc, int stage, double dt,
                                                             int first_chunk, int last_chunk) {
    const int idx = blockIdx.x * kBlock + threadIdx.x;
    if (idx >= g.nx * g.ny) return;
    const int i = 2 + idx / g.ny;
    const int j = 2 + idx % g.ny;
    const int NY = g.NY;
    const int ncell = g.NX * g.NY;
    const size_t c = size_t(i) * NY + j;
    double F[NM];
    if (first_chunk) {
#pragma unroll
        for (int k = 0; k < NM; ++k) F[k] = 0.0;
    } else {
#pragma unroll
        for (int k = 0; k < NM; ++k) F[k] = g.flux[c * NM + k];
    }
    const double theta = g.theta, dx = g.dx, dy = g.dy;
    for (int qq = 0; qq < nqc; ++qq) {
        const int q = q0 + qq;
        const double *E = g.ansatz + size_t(qq) * ncell;
        const double xi = g.xi[q], eta = g.eta[q], w = g.w[q];
        double f1 = 0.0, f2 = 0.0, f3 = 0.0, f4 = 0.0;
        if (xi > 0) {
            const double k1 = E[(i - 2) * NY + j], k2 = E[(i - 1) * NY + j];
            const double k3 = E[i * NY + j], k4 = E[(i + 1) * NY + j];
            f1 = k3 + 0.5 * slopefit(k2, k3, k4, theta) - k2 - 0.5 * slopefit(k1, k2, k3, theta);
            f1 = f1 * xi * w;
        }
        if (xi < 0) {
            const double k1 = E[(i - 1) * NY + j], k2 = E[i * NY + j];
            const double k3 = E[(i + 1) * NY + j], k4 = E[(i + 2) * NY + j];
            f2 = k3 - 0.5 * slopefit(k2, k3, k4, theta) - k2 + 0.5 * slopefit(k1, k2, k3, theta);
            f2 = f2 * xi * w;
        }
        if (eta > 0) {
            const double k1 = E[i * NY + (j - 2)], k2 = E[i * NY + (j - 1)];
            const double k3 = E[i * NY + j], k4 = E[i * NY + (j + 1)];
            f3 = k3 + 0.5 * slopefit(k2, k3, k4, theta) - k2 - 0.5 * slopefit(k1, k2, k3, theta);
            f3 = f3 * eta * w;
        }
        if (eta < 0) {
            const double k1 = E[i * NY + (j - 1)], k2 = E[i * NY + j];
            const double k3 = E[i * NY + (j + 1)], k4 = E[i * NY + (j + 2)];
            f4 = k3 - 0.5 * slopefit(k2, k3, k4, theta) - k2 + 0.5 * slopefit(k1, k2, k3, theta);
            f4 = f4 * eta * w;
        }
        const double comb = (f1 + f2) / dx + (f3 + f4) / dy;
        const double *row = g.p + size_t(q) * g.np2;
#pragma unroll
        for (int k = 0; k < NM; ++k) F[k] = fma(row[k], comb, F[k]);
    }
    if (!last_chunk || stage == 0) {
#pragma unroll
        for (int k = 0; k < NM; ++k) g.flux[c * NM + k] = F[k];
        return;
    }
    double *u = g.moments + c * NM;
    const double st = g.sigma_t[c], ss = g.sigma_s[c];
    double v[NM];
    v[0] = u[0] * (1.0 + dt * ss - dt * st) - dt * F[0];
#pragma unroll
    for (int k = 1; k < NM; ++k) v[k] = u[k] * (1.0 - dt * st) - dt * F[k];
    if (g.init_cond == 1) {                                   // lattice source (:328-334)
        const double x_i = edge_coord(g.init_x, i - 2, dx);
        const double y_j = edge_coord(g.init_y, j - 2, dy);
        if (fabs(x_i) < 0.5 && fabs(y_j) < 0.5) v[0] = v[0] + dt * 2.0 * sqrt(M_PI);
    }
    if (stage == 2) {
        const double *uo = g.moments_old + c * NM;
#pragma unroll
        for (int k = 0; k < NM; ++k) v[k] = 0.5 * (v[k] + uo[k]);
    }
#pragma unroll
    for (int k = 0; k < NM; ++k) u[k] = v[k];
}

// The ansatz fused into the flux: a block owns a TX x TY tile of interior cells,
// keeps alpha of the tile and its 2-cell halo in LDS, and for each chunk of QC
// directions evaluates the ansatz E of the halo'd tile into LDS (the halo
// is evaluated by both neighbouring tiles) before the stencil reads it.  E never
// goes to HBM.  The arithmetic of every cell is that of ansatz_kernel +
// flux_euler_kernel (same summation order over k and q), so the results are
// identical.
template <int NM, int TX, int TY, int QC>
__global__ __launch_bounds__(TX * TY) void flux_fused_kernel(GridArgs g, int stage, double dt) {
    constexpr int HX = TX + 4, HY = TY + 4, NH = HX * HY, NT = TX * TY;
    __shared__ double al_s[NH][NM];
    __shared__ double E_s[QC][NH];
    const int tid = threadIdx.x;
    const int tx0 = blockIdx.x * TX, ty0 = blockIdx.y * TY;   // halo cell (hi, hj) is grid cell (tx0+hi, ty0+hj)
    const int NX = g.NX, NY = g.NY;
    for (int idx = tid; idx < NH * NM; idx += NT) {
        const int h = idx / NM, k = idx % NM;
        const int gi = tx0 + h / HY, gj = ty0 + h % HY;
        al_s[h][k] = (gi < NX && gj < NY) ? g.alpha[(size_t(gi) * NY + gj) * NM + k] : 0.0;
    }
    const int li = tid / TY, lj = tid % TY;
    const int i = 2 + tx0 + li, j = 2 + ty0 + lj;
    const bool active = (tx0 + li < g.nx) && (ty0 + lj < g.ny);
    const int hc = (li + 2) * HY + (lj + 2);                  // this cell's halo index
    const double theta = g.theta, dx = g.dx, dy = g.dy;
    double F[NM];
#pragma unroll
    for (int k = 0; k < NM; ++k) F[k] = 0.0;
    for (int q0 = 0; q0 < g.nq; q0 += QC) {
        const int nqc = min(QC, g.nq - q0);
        __syncthreads();
        for (int qq = 0; qq < nqc; ++qq) {
            const double *row = g.p + size_t(q0 + qq) * g.np2;
            for (int h = tid; h < NH; h += NT) {
                double sa = 0.0;
#pragma unroll
                for (int k = 0; k < NM; ++k) sa = fma(al_s[h][k], row[k], sa);
                E_s[qq][h] = ansatz_value(g, sa);
            }
        }
        __syncthreads();
        if (!active) continue;
        for (int qq = 0; qq < nqc; ++qq) {
            const int q = q0 + qq;
            const double *E = E_s[qq];
            const double xi = g.xi[q], eta = g.eta[q], w = g.w[q];
            double f1 = 0.0, f2 = 0.0, f3 = 0.0, f4 = 0.0;
            if (xi > 0) {
                const double k1 = E[hc - 2 * HY], k2 = E[hc - HY], k3 = E[hc], k4 = E[hc + HY];
                f1 = k3 + 0.5 * slopefit(k2, k3, k4, theta) - k2 - 0.5 * slopefit(k1, k2, k3, theta);
                f1 = f1 * xi * w;
            }
            if (xi < 0) {
                const double k1 = E[hc - HY], k2 = E[hc], k3 = E[hc + HY], k4 = E[hc + 2 * HY];
                f2 = k3 - 0.5 * slopefit(k2, k3, k4, theta) - k2 + 0.5 * slopefit(k1, k2, k3, theta);
                f2 = f2 * xi * w;
            }
            if (eta > 0) {
                const double k1 = E[hc - 2], k2 = E[hc - 1], k3 = E[hc], k4 = E[hc + 1];
                f3 = k3 + 0.5 * slopefit(k2, k3, k4, theta) - k2 - 0.5 * slopefit(k1, k2, k3, theta);
                f3 = f3 * eta * w;
            }
            if (eta < 0) {
                const double k1 = E[hc - 1], k2 = E[hc], k3 = E[hc + 1], k4 = E[hc + 2];
                f4 = k3 - 0.5 * slopefit(k2, k3, k4, theta) - k2 + 0.5 * slopefit(k1, k2, k3, theta);
                f4 = f4 * eta * w;
            }
            const double comb = (f1 + f2) / dx + (f3 + f4) / dy;
            const double *row = g.p + size_t(q) * g.np2;
#pragma unroll
            for (int k = 0; k < NM; ++k) F[k] = fma(row[k], comb, F[k]);
        }
    }
    if (!active) return;
    const size_t c = size_t(i) * NY + j;
    if (stage == 0) {
#pragma unroll
        for (int k = 0; k < NM; ++k) g.flux[c * NM + k] = F[k];
        return;
    }
    double *u = g.moments + c * NM;
    const double st = g.sigma_t[c], ss = g.sigma_s[c];
    double v[NM];
    v[0] = u[0] * (1.0 + dt * ss - dt * st) - dt * F[0];
#pragma unroll
    for (int k = 1; k < NM; ++k) v[k] = u[k] * (1.0 - dt * st) - dt * F[k];
    if (g.init_cond == 1) {                                   // lattice source (:328-334)
        const double x_i = edge_coord(g.init_x, i - 2, dx);
        const double y_j = edge_coord(g.init_y, j - 2, dy);
        if (fabs(x_i) < 0.5 && fabs(y_j) < 0.5) v[0] = v[0] + dt * 2.0 * sqrt(M_PI);
    }
    if (stage == 2) {
        const double *uo = g.moments_old + c * NM;
#pragma unroll
        for (int k = 0; k < NM; ++k) v[k] = 0.5 * (v[k] + uo[k]);
    }
#pragma unroll
    for (int k = 0; k < NM; ++k) u[k] = v[k];
}

// duplicateBoundaries (momopt_boundaries.cpp:190-223): x strips, then y strips.
__global__ void ghost_x_kernel(GridArgs g) {
    const int idx = blockIdx.x * kBlock + threadIdx.x;   // over 2 * ny * nm
    const int total = 2 * g.ny * g.nm;
    if (idx >= total) return;
    const int k = idx % g.nm;
    const int j = 2 + (idx / g.nm) % g.ny;
    const int i = idx / (g.nm * g.ny);
    const int x0 = 0, x1 = 2, x2 = 2 + g.nx - 2, x3 = 2 + g.nx;
    const int NY = g.NY, nm = g.nm;
    g.moments[(size_t(x0 + i) * NY + j) * nm + k] = g.moments[(size_t(x2 + i) * NY + j) * nm + k];
    g.moments[(size_t(x3 + i) * NY + j) * nm + k] = g.moments[(size_t(x1 + i) * NY + j) * nm + k];
}

__global__ void ghost_y_kernel(GridArgs g) {
    const int idx = blockIdx.x * kBlock + threadIdx.x;   // over nx * 2 * nm
    const int total = g.nx * 2 * g.nm;
    if (idx >= total) return;
    const int k = idx % g.nm;
    const int jj = (idx / g.nm) % 2;
    const int i = 2 + idx / (g.nm * 2);
    const int y0 = 0, y1 = 2, y2 = 2 + g.ny - 2, y3 = 2 + g.ny;
    const int NY = g.NY, nm = g.nm;
    g.moments[(size_t(i) * NY + y0 + jj) * nm + k] = g.moments[(size_t(i) * NY + y2 + jj) * nm + k];
    g.moments[(size_t(i) * NY + y3 + jj) * nm + k] = g.moments[(size_t(i) * NY + y1 + jj) * nm + k];
}

template <int NM>
void launch_flux_t(const GridArgs &g, int q0, int nqc, int stage, double dt, bool fc, bool lc, hipStream_t s) {
    const int n = g.nx * g.ny;
    hipLaunchKernelGGL(flux_euler_kernel<NM>, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, s, g, q0, nqc,
                       stage, dt, int(fc), int(lc));
}

// The fused ansatz + flux with the directions split into gridDim.z groups of
// consecutive q: block (tile, z) sums the flux of its TX x TY cells over group
// z's directions into fpart[z] (interior cells, [cell][k]), and
// flux_sum_kernel adds the groups in order and applies the Euler stage.  With
// 80,656 cells a launch of one group per tile is about one wave per SIMD and
// waits on latency; the groups multiply the waves in flight.  alpha of the
// halo cells a thread evaluates stays in its registers, so LDS holds only E.
// Per cell and direction the arithmetic is flux_fused_kernel's except that
// (f1 + f2) / dx + (f3 + f4) / dy multiplies by reciprocals, and the q sum is
// split into the groups: the result moves at rounding level.
template <int NM, int TX, int TY, int QC>
__global__ __launch_bounds__(TX * TY) void flux_split_kernel(GridArgs g) {
    static_assert(TX == kSplitTile && TY == kSplitTile, "the tile flags' geometry");
    if (g.tile_mode != 0) {   // this launch's share of the tiles (uniform over the block)
        const bool flagged = g.tile_flag[blockIdx.x * gridDim.y + blockIdx.y] != 0;
        if (flagged != (g.tile_mode == 2)) return;
    }
    constexpr int HX = TX + 4, HY = TY + 4, NH = HX * HY, NT = TX * TY;
    constexpr int HPT = (NH + NT - 1) / NT;                  // halo cells evaluated per thread
    __shared__ double E_s[QC][NH];
    const int tid = threadIdx.x;
    const int tx0 = blockIdx.x * TX, ty0 = blockIdx.y * TY;
    const int NX = g.NX, NY = g.NY;
    double al[HPT][NM];
#pragma unroll
    for (int hh = 0; hh < HPT; ++hh) {
        const int h = tid + hh * NT;
        const int gi = tx0 + h / HY, gj = ty0 + h % HY;
        const bool in = h < NH && gi < NX && gj < NY;
#pragma unroll
        for (int k = 0; k < NM; ++k) al[hh][k] = in ? g.alpha[(size_t(gi) * NY + gj) * NM + k] : 0.0;
    }
    const int li = tid / TY, lj = tid % TY;
    const int i = 2 + tx0 + li, j = 2 + ty0 + lj;
    const bool active = (tx0 + li < g.nx) && (ty0 + lj < g.ny);
    const int hc = (li + 2) * HY + (lj + 2);
    const double theta = g.theta, rdx = 1.0 / g.dx, rdy = 1.0 / g.dy;
    const int per = (g.nq + gridDim.z - 1) / gridDim.z;
    const int qa = blockIdx.z * per, qb = min(g.nq, qa + per);
    double F[NM];
#pragma unroll
    for (int k = 0; k < NM; ++k) F[k] = 0.0;
    for (int q0 = qa; q0 < qb; q0 += QC) {
        const int nqc = min(QC, qb - q0);
        __syncthreads();
        for (int qq = 0; qq < nqc; ++qq) {
            const double *row = g.p + size_t(q0 + qq) * g.np2;
#pragma unroll
            for (int hh = 0; hh < HPT; ++hh) {
                const int h = tid + hh * NT;
                if (h < NH) {
                    double sa = 0.0;
#pragma unroll
                    for (int k = 0; k < NM; ++k) sa = fma(al[hh][k], row[k], sa);
                    E_s[qq][h] = ansatz_value(g, sa);
                }
            }
        }
        __syncthreads();
        if (!active) continue;
        for (int qq = 0; qq < nqc; ++qq) {
            const int q = q0 + qq;
            const double *E = E_s[qq];
            const double xi = g.xi[q], eta = g.eta[q], w = g.w[q];
            double f1 = 0.0, f2 = 0.0, f3 = 0.0, f4 = 0.0;
            if (xi > 0) {
                const double k1 = E[hc - 2 * HY], k2 = E[hc - HY], k3 = E[hc], k4 = E[hc + HY];
                f1 = k3 + 0.5 * slopefit(k2, k3, k4, theta) - k2 - 0.5 * slopefit(k1, k2, k3, theta);
                f1 = f1 * xi * w;
            }
            if (xi < 0) {
                const double k1 = E[hc - HY], k2 = E[hc], k3 = E[hc + HY], k4 = E[hc + 2 * HY];
                f2 = k3 - 0.5 * slopefit(k2, k3, k4, theta) - k2 + 0.5 * slopefit(k1, k2, k3, theta);
                f2 = f2 * xi * w;
            }
            if (eta > 0) {
                const double k1 = E[hc - 2], k2 = E[hc - 1], k3 = E[hc], k4 = E[hc + 1];
                f3 = k3 + 0.5 * slopefit(k2, k3, k4, theta) - k2 - 0.5 * slopefit(k1, k2, k3, theta);
                f3 = f3 * eta * w;
            }
            if (eta < 0) {
                const double k1 = E[hc - 1], k2 = E[hc], k3 = E[hc + 1], k4 = E[hc + 2];
                f4 = k3 - 0.5 * slopefit(k2, k3, k4, theta) - k2 + 0.5 * slopefit(k1, k2, k3, theta);
                f4 = f4 * eta * w;
            }
            const double comb = (f1 + f2) * rdx + (f3 + f4) * rdy;
            const double *row = g.p + size_t(q) * g.np2;
#pragma unroll
            for (int k = 0; k < NM; ++k) F[k] = fma(row[k], comb, F[k]);
        }
    }
    if (!active) return;
    double *o = g.fpart + (size_t(blockIdx.z) * g.nx * g.ny + size_t(i - 2) * g.ny + (j - 2)) * NM;
#pragma unroll
    for (int k = 0; k < NM; ++k) o[k] = F[k];
}

// F = fpart[0] + fpart[1] + ... (in group order), then the Euler stage.
template <int NM>
__global__ __launch_bounds__(kBlock) void flux_sum_kernel(GridArgs g, int groups, int stage, double dt) {
    const int idx = blockIdx.x * kBlock + threadIdx.x;
    if (idx >= g.nx * g.ny) return;
    const int i = 2 + idx / g.ny, j = 2 + idx % g.ny;
    if (g.tile_mode != 0) {
        const int tiles_y = (g.ny + kSplitTile - 1) / kSplitTile;
        const bool flagged = g.tile_flag[((i - 2) / kSplitTile) * tiles_y + (j - 2) / kSplitTile] != 0;
        if (flagged != (g.tile_mode == 2)) return;
    }
    double F[NM];
#pragma unroll
    for (int k = 0; k < NM; ++k) F[k] = g.fpart[size_t(idx) * NM + k];
    for (int z = 1; z < groups; ++z) {
        const double *p = g.fpart + (size_t(z) * g.nx * g.ny + idx) * NM;
#pragma unroll
        for (int k = 0; k < NM; ++k) F[k] += p[k];
    }
    euler_epilogue<NM>(g, i, j, size_t(i) * g.NY + j, F, stage, dt);
}

template <int NM>
void launch_fused_t(const GridArgs &g, int tile, int stage, double dt, hipStream_t s) {
    if constexpr (NM <= 15) {   // alpha of two halo cells per thread in registers
        if (tile == 3) {
            const dim3 grid((g.nx + 15) / 16, (g.ny + 15) / 16, g.groups);
            hipLaunchKernelGGL((flux_split_kernel<NM, 16, 16, 8>), grid, dim3(256), 0, s, g);
            hipLaunchKernelGGL((flux_sum_kernel<NM>), dim3((g.nx * g.ny + kBlock - 1) / kBlock), dim3(kBlock), 0, s,
                               g, g.groups, stage, dt);
            return;
        }
    }
    if constexpr (NM <= 15) {   // alpha of a 20 x 20 halo'd tile fits LDS up to n = 15
        if (tile == 2) {
            hipLaunchKernelGGL((flux_fused_kernel<NM, 16, 16, 8>), dim3((g.nx + 15) / 16, (g.ny + 15) / 16), dim3(256),
                               0, s, g, stage, dt);
            return;
        }
    }
    hipLaunchKernelGGL((flux_fused_kernel<NM, 8, 8, 8>), dim3((g.nx + 7) / 8, (g.ny + 7) / 8), dim3(64), 0, s, g,
                       stage, dt);
}

template <int NM>
void launch_ansatz_t(const GridArgs &g, int q0, int nqc, hipStream_t s) {
    const int n = g.NX * g.NY;
    hipLaunchKernelGGL(ansatz_kernel<NM>, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, s, g, q0, nqc);
}

__global__ void tile_flag_clear_kernel(int *flags, int n) {
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < n; t += gridDim.x * blockDim.x) flags[t] = 0;
}

// One thread per list entry: the tiles whose stencil region [16 t, 16 t + 20)
// (grid coordinates, ghosts included) holds the entry's cell.
__global__ void tile_flag_mark_kernel(const int *list0, const int *cnt0, const int *list1, const int *cnt1, int ncells,
                                      int NY, int tiles_x, int tiles_y, int *flags) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= ncells) return;
    const int n0 = *cnt0, n1 = *cnt1;
    int c = -1;
    if (s < n0) c = list0[s];
    else if (s - n0 < n1) c = list1[s - n0];
    if (c < 0) return;
    const int ci = c / NY, cj = c % NY;
    constexpr int T = kSplitTile, R = kSplitTile + 4;
    const int tx0 = ci >= R ? (ci - R + T) / T : 0, tx1 = min(tiles_x - 1, ci / T);
    const int ty0 = cj >= R ? (cj - R + T) / T : 0, ty1 = min(tiles_y - 1, cj / T);
    for (int tx = tx0; tx <= tx1; ++tx)
        for (int ty = ty0; ty <= ty1; ++ty) flags[tx * tiles_y + ty] = 1;
}

}  // namespace

void launch_tile_flags(const int *list0, const int *cnt0, const int *list1, const int *cnt1, int ncells, int nx,
                       int ny, int *flags, hipStream_t s) {
    const int tx = (nx + kSplitTile - 1) / kSplitTile, ty = (ny + kSplitTile - 1) / kSplitTile;
    hipLaunchKernelGGL(tile_flag_clear_kernel, dim3((tx * ty + 255) / 256), dim3(256), 0, s, flags, tx * ty);
    hipLaunchKernelGGL(tile_flag_mark_kernel, dim3((ncells + 255) / 256), dim3(256), 0, s, list0, cnt0, list1, cnt1,
                       ncells, ny + 4, tx, ty, flags);
}

void launch_ghost_fill(const GridArgs &g, hipStream_t s) {
    if (g.init_cond != 2) return;
    const int nxs = 2 * g.ny * g.nm, nys = g.nx * 2 * g.nm;
    hipLaunchKernelGGL(ghost_x_kernel, dim3((nxs + kBlock - 1) / kBlock), dim3(kBlock), 0, s, g);
    hipLaunchKernelGGL(ghost_y_kernel, dim3((nys + kBlock - 1) / kBlock), dim3(kBlock), 0, s, g);
}

#define MNC_NM_SWITCH(NMV, CALL) \
    switch (NMV) {               \
    case 1: CALL(1); break;      \
    case 3: CALL(3); break;      \
    case 6: CALL(6); break;      \
    case 10: CALL(10); break;    \
    case 15: CALL(15); break;    \
    case 21: CALL(21); break;    \
    case 28: CALL(28); break;    \
    case 36: CALL(36); break;    \
    case 45: CALL(45); break;    \
    default: return false;       \
    }

void launch_ansatz(const GridArgs &g, int q0, int nqc, hipStream_t s) {
    auto f = [&]() -> bool {
#define CALLA(N) launch_ansatz_t<N>(g, q0, nqc, s)
        MNC_NM_SWITCH(g.nm, CALLA)
#undef CALLA
        return true;
    };
    f();
}

bool launch_flux_fused(const GridArgs &g, int tile, int stage, double dt, hipStream_t s) {
#define CALLX(N) launch_fused_t<N>(g, tile, stage, dt, s)
    MNC_NM_SWITCH(g.nm, CALLX)
#undef CALLX
    return true;
}

bool launch_flux_euler(const GridArgs &g, int q0, int nqc, int stage, double dt, bool first_chunk, bool last_chunk,
                       hipStream_t s) {
#define CALLF(N) launch_flux_t<N>(g, q0, nqc, stage, dt, first_chunk, last_chunk, s)
    MNC_NM_SWITCH(g.nm, CALLF)
#undef CALLF
    return true;
}

}  // namespace mnc

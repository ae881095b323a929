// mnc_sn.hip -- the Sn discrete-ordinates solver (src/kinetic/) on the GPU:
// the independent truth for the moment closures.  One Heun step
// (KineticSolver::update, kinetic_update.cpp:120-235) is two fused stage
// launches: each wave takes one interior cell, its lanes the ordinates;
// the flux of every ordinate (kinetic_update.cpp:41-114, MUSCL with the
// minmod slope fit, theta = 2), the isotropic scattering integral (a wave
// sum), the init_cond 1 source and the Euler update are one pass from the
// stage's input grid into its output grid.  The second stage folds in the
// Heun average with the step's initial state.  Periodic ghosts
// (kinetic_boundaries.cpp:190-223) are a separate small launch.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <climits>
#include <cmath>
#include <vector>

#include "mnc_internal.h"

namespace {

// x0 + k h rounded as a product and a sum: the contraction into one fma is
// switched off here, as the reference's x86-64 build (no FMA) rounds it.
__device__ __forceinline__ double edge_coord(double x0, int k, double h) {
#pragma clang fp contract(off)
    return x0 + double(k) * h;
}

constexpr int kSnThreads = 256;

struct SnArgs {
    int nx, ny, nq, init_cond;
    double dx, dy, dt, init_x, init_y;
    const double *w, *xi, *eta;       // [nq]
    const double *sigma_t, *sigma_s;  // [(nx+4)(ny+4)]
};

// The reference's selects as v_min / v_max (kinetic_update.cpp:22-25):
// fabs(x) < y s ? fabs(x) : y s is fmin(fabs(x), y s) and m > 0 ? m : 0 is
// fmax(m, 0) -- equal values, NaN included; only the sign of an exact zero
// slope may differ, which a sum with any nonzero term does not see.
__device__ __forceinline__ double sn_minmod(double x, double y) {
    const double s = (x >= 0.0) ? 1.0 : -1.0;
    const double m = fmin(fabs(x), y * s);
    return s * fmax(m, 0.0);
}
__device__ __forceinline__ double sn_slopefit(double l, double c, double r) {   // :31-35, theta = 2
    return sn_minmod(2.0 * (r - c), sn_minmod(0.5 * (r - l), 2.0 * (c - l)));
}

// ghosts of f (and of f2 when not null) <- the opposite interior layers of f
// (x: interior rows; y: interior columns)
__global__ void sn_ghost_kernel(double *f, double *f2, int nx, int ny, int nq) {
    const int NY = ny + 4;
    const long long nxw = 2LL * ny * nq, nyw = 2LL * nx * nq;
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < nxw + nyw;
         t += (long long)gridDim.x * blockDim.x) {
        if (t < nxw) {
            const int q = int(t % nq), j = 2 + int((t / nq) % ny), g = int(t / (nq * (long long)ny));
            const double lo = f[(size_t(nx + g) * NY + j) * nq + q], hi = f[(size_t(2 + g) * NY + j) * nq + q];
            f[(size_t(g) * NY + j) * nq + q] = lo;
            f[(size_t(nx + 2 + g) * NY + j) * nq + q] = hi;
            if (f2) {
                f2[(size_t(g) * NY + j) * nq + q] = lo;
                f2[(size_t(nx + 2 + g) * NY + j) * nq + q] = hi;
            }
        } else {
            const long long u = t - nxw;
            const int q = int(u % nq), i = 2 + int((u / nq) % nx), g = int(u / (nq * (long long)nx));
            const double lo = f[(size_t(i) * NY + ny + g) * nq + q], hi = f[(size_t(i) * NY + 2 + g) * nq + q];
            f[(size_t(i) * NY + g) * nq + q] = lo;
            f[(size_t(i) * NY + ny + 2 + g) * nq + q] = hi;
            if (f2) {
                f2[(size_t(i) * NY + g) * nq + q] = lo;
                f2[(size_t(i) * NY + ny + 2 + g) * nq + q] = hi;
            }
        }
    }
}

// out <- Euler stage of in for every interior cell; with old, the Heun average
// 0.5 (stage + old) (old may alias out: each element is read before it is written
// by the same thread).  One wave per cell, four consecutive cells per block.
// XCD-aware and column-tiled: workgroups go to the 8 XCDs round robin, so
// block b runs on XCD b % 8, which owns one contiguous range of rows; it
// sweeps that range in strips of kSnStrip columns, strip by strip, row by row
// within a strip.  The stencil's x neighbours (a row, NY nq 8 bytes, apart)
// then span five rows of one strip, kSnStrip nq 8 bytes each (640 KB at
// 256 ordinates), which stay in the XCD's 4 MB L2 until they are reused.
constexpr int kSnStrip = 64;
constexpr int kSnWaves = kSnThreads / 64;

__host__ __device__ __forceinline__ int sn_rows_per_xcd(int nx) { return (nx + 7) / 8; }

__global__ __launch_bounds__(kSnThreads) void sn_stage_kernel(SnArgs a, const double *__restrict__ in, double *out,
                                                              const double *old) {
    const int NY = a.ny + 4, nq = a.nq;
    const int lane = threadIdx.x & 63;
    {
        const int rows_per = sn_rows_per_xcd(a.nx);
        const int r0 = (blockIdx.x & 7) * rows_per;
        const int rows = min(a.nx, r0 + rows_per) - r0;
        const int L = (blockIdx.x >> 3) * kSnWaves + (threadIdx.x >> 6);   // the XCD's cell, strip order
        if (rows <= 0 || L >= rows * a.ny) return;   // uniform over the wave; no block barrier follows
        const int full = rows * kSnStrip;             // cells of a full strip
        const int sidx = L / full, j0 = sidx * kSnStrip;
        const int ws = min(kSnStrip, a.ny - j0), rem = L - sidx * full;
        const int i = 2 + r0 + rem / ws, j = 2 + j0 + rem % ws;
        const size_t c = size_t(i) * NY + j;
        double part = 0.0;
        for (int q = lane; q < nq; q += 64) part += a.w[q] * in[c * nq + q];
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) part += __shfl_xor(part, o);
        double integral = part * a.sigma_s[c] / (4 * M_PI);
        const double st = a.sigma_t[c];
        // rounded as a product and a sum (no fma), as the reference's build
        // evaluates the source test |x| < 0.5 (kinetic_update.cpp:173-175)
        const double x_i = edge_coord(a.init_x, i, a.dx), y_j = edge_coord(a.init_y, j, a.dy);
        const bool src = a.init_cond == 1 && fabs(x_i) < 0.5 && fabs(y_j) < 0.5;
        const size_t sx = size_t(NY) * nq;     // x neighbour stride; y neighbours are nq apart
        for (int q = lane; q < nq; q += 64) {
            // The upwind side differs between the lanes' ordinates, so both
            // sides' stencils are selected, not branched to (a branch would
            // run both slope fits on every wave).  s + h * slope with h = +-0.5
            // is the reference's s +- 0.5 * slope exactly (0.5 * slope is exact).
            const double *f = in + c * nq + q;
            const double c0 = f[0];
            const double ym2 = f[-2 * nq], ym1 = f[-nq], yp1 = f[nq], yp2 = f[2 * nq];
            const double xm2 = f[-2 * sx], xm1 = f[-sx], xp1 = f[sx], xp2 = f[2 * sx];
            const double xiq = a.xi[q], etq = a.eta[q];
            const bool ep = etq > 0, xp = xiq > 0;
            const double hy = ep ? 0.5 : -0.5, hx = xp ? 0.5 : -0.5;
            const double y1 = ep ? ym2 : ym1, y2 = ep ? ym1 : c0, y3 = ep ? c0 : yp1, y4 = ep ? yp1 : yp2;
            const double x1 = xp ? xm2 : xm1, x2 = xp ? xm1 : c0, x3 = xp ? c0 : xp1, x4 = xp ? xp1 : xp2;
            const double nF = y3 + hy * sn_slopefit(y2, y3, y4);
            const double sF = y2 + hy * sn_slopefit(y1, y2, y3);
            const double eF = x3 + hx * sn_slopefit(x2, x3, x4);
            const double wF = x2 + hx * sn_slopefit(x1, x2, x3);
            const double flux = (eF - wF) * xiq / a.dx + (nF - sF) * etq / a.dy;
            double v = c0;
            v = v - a.dt * st * v + a.dt * integral;
            if (src) v = v + a.dt;
            v = v - a.dt * flux;
            if (old) v = 0.5 * (v + old[c * nq + q]);
            out[c * nq + q] = v;
        }
    }
}

}  // namespace

extern "C" int mnc_sn_run(int nx, int ny, int quad_order, double dx, double dy, double init_x, double init_y,
                          int init_cond, const double *sigma_t, const double *sigma_s, double dt, int steps,
                          double *kinetic, int device, double *kernel_ms) {
    if (nx < 2 || ny < 2 || quad_order < 2 || (quad_order & 1) || steps < 0 || !sigma_t || !sigma_s || !kinetic ||
        !(dx > 0.0) || !(dy > 0.0) || !std::isfinite(dt))
        return MNC_ERR_ARG;
    if (size_t(nx + 4) * (ny + 4) * size_t(quad_order) * quad_order > (size_t(1) << 33)) return MNC_ERR_ARG;
    if (size_t(nx) * ny > size_t(INT_MAX) - 8) return MNC_ERR_ARG;   // one block per cell
    mnc::Tables t;
    if (mnc::build_tables(1, quad_order, 0, t) != MNC_OK) return MNC_ERR_ARG;   // the same ordinates
    const int nq = t.nq;
    const size_t ncell = size_t(nx + 4) * (ny + 4), n = ncell * nq;
    if (hipSetDevice(device) != hipSuccess) return MNC_ERR_HIP;
    double *d = nullptr;
    if (hipMalloc((void **)&d, (2 * n + 2 * ncell + 3 * size_t(nq)) * sizeof(double)) != hipSuccess) return MNC_ERR_HIP;
    double *f = d, *f1 = d + n, *st = f1 + n, *ss = st + ncell, *w = ss + ncell, *xi = w + nq, *eta = xi + nq;
    hipStream_t s = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = MNC_OK;
    auto ok = [&](hipError_t e) { if (e != hipSuccess && rc == MNC_OK) rc = MNC_ERR_HIP; return rc == MNC_OK; };
    if (ok(hipStreamCreate(&s)) && ok(hipEventCreate(&e0)) && ok(hipEventCreate(&e1)) &&
        ok(hipMemcpyAsync(f, kinetic, n * sizeof(double), hipMemcpyHostToDevice, s)) &&
        ok(hipMemcpyAsync(f1, f, n * sizeof(double), hipMemcpyDeviceToDevice, s)) &&   // ghosts of the stage grid
        ok(hipMemcpyAsync(st, sigma_t, ncell * sizeof(double), hipMemcpyHostToDevice, s)) &&
        ok(hipMemcpyAsync(ss, sigma_s, ncell * sizeof(double), hipMemcpyHostToDevice, s)) &&
        ok(hipMemcpyAsync(w, t.w.data(), nq * sizeof(double), hipMemcpyHostToDevice, s)) &&
        ok(hipMemcpyAsync(xi, t.xi.data(), nq * sizeof(double), hipMemcpyHostToDevice, s)) &&
        ok(hipMemcpyAsync(eta, t.eta.data(), nq * sizeof(double), hipMemcpyHostToDevice, s)) &&
        ok(hipEventRecord(e0, s))) {
        SnArgs a{nx, ny, nq, init_cond, dx, dy, dt, init_x, init_y, w, xi, eta, st, ss};
        // four cells each, one row range per XCD
        const int blocks = 8 * ((sn_rows_per_xcd(nx) * ny + kSnWaves - 1) / kSnWaves);
        const long long gw = 2LL * (nx + ny) * nq;
        const int gblocks = int(std::min<long long>((gw + 255) / 256, 4096));
        for (int k = 0; k < steps && rc == MNC_OK; ++k) {
            // the reference fills the ghosts of its one array before each stage,
            // so after the step they hold the first stage's layers: the second
            // fill writes them into the state too
            if (init_cond == 2)
                hipLaunchKernelGGL(sn_ghost_kernel, dim3(gblocks), dim3(256), 0, s, f, (double *)nullptr, nx, ny, nq);
            hipLaunchKernelGGL(sn_stage_kernel, dim3(blocks), dim3(kSnThreads), 0, s, a, f, f1, nullptr);
            if (init_cond == 2) hipLaunchKernelGGL(sn_ghost_kernel, dim3(gblocks), dim3(256), 0, s, f1, f, nx, ny, nq);
            hipLaunchKernelGGL(sn_stage_kernel, dim3(blocks), dim3(kSnThreads), 0, s, a, f1, f, f);
            ok(hipGetLastError());
        }
        if (ok(hipEventRecord(e1, s)) && ok(hipMemcpyAsync(kinetic, f, n * sizeof(double), hipMemcpyDeviceToHost, s)) &&
            ok(hipStreamSynchronize(s)) && kernel_ms) {
            float ms = 0.0f;
            if (ok(hipEventElapsedTime(&ms, e0, e1))) *kernel_ms = ms;
        }
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (s) (void)hipStreamDestroy(s);
    (void)hipFree(d);
    return rc;
}

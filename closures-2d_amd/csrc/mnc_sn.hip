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

__device__ __forceinline__ double sn_minmod(double x, double y) {       // kinetic_update.cpp:22-25
    const double s = (x >= 0.0) ? 1.0 : -1.0;
    const double m = fabs(x) < y * s ? fabs(x) : y * s;
    return s * (m > 0.0 ? m : 0.0);
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
// by the same thread).  One wave per cell, four consecutive cells per block,
// XCD-aware: workgroups go to the 8 XCDs round robin, so block b takes group
// (b % 8) * per + b / 8 and each XCD sweeps one contiguous range of rows; the
// stencil's x neighbours (a row, ~NY nq 8 bytes, apart) are then still in
// that XCD's L2 when they are reused.
__global__ __launch_bounds__(kSnThreads) void sn_stage_kernel(SnArgs a, const double *__restrict__ in, double *out,
                                                              const double *old) {
    constexpr int kWaves = kSnThreads / 64;
    const int NY = a.ny + 4, nq = a.nq;
    const int ncell = a.nx * a.ny, ngroup = (ncell + kWaves - 1) / kWaves, per = (ngroup + 7) / 8;
    const int lane = threadIdx.x & 63;
    {
        const int grp = (blockIdx.x & 7) * per + (blockIdx.x >> 3);
        const int cid = grp * kWaves + (threadIdx.x >> 6);
        if (grp >= ngroup || cid >= ncell) return;   // uniform over the wave; no block barrier follows
        const int i = 2 + cid / a.ny, j = 2 + cid % a.ny;
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
        for (int q = lane; q < nq; q += 64) {
            auto F = [&](int di, int dj) { return in[(size_t(i + di) * NY + (j + dj)) * nq + q]; };
            double nF, sF, eF, wF;
            if (a.eta[q] > 0) {
                const double s1 = F(0, -2), s2 = F(0, -1), s3 = F(0, 0), s4 = F(0, 1);
                nF = s3 + 0.5 * sn_slopefit(s2, s3, s4);
                sF = s2 + 0.5 * sn_slopefit(s1, s2, s3);
            } else {
                const double s1 = F(0, -1), s2 = F(0, 0), s3 = F(0, 1), s4 = F(0, 2);
                nF = s3 - 0.5 * sn_slopefit(s2, s3, s4);
                sF = s2 - 0.5 * sn_slopefit(s1, s2, s3);
            }
            if (a.xi[q] > 0) {
                const double s1 = F(-2, 0), s2 = F(-1, 0), s3 = F(0, 0), s4 = F(1, 0);
                eF = s3 + 0.5 * sn_slopefit(s2, s3, s4);
                wF = s2 + 0.5 * sn_slopefit(s1, s2, s3);
            } else {
                const double s1 = F(-1, 0), s2 = F(0, 0), s3 = F(1, 0), s4 = F(2, 0);
                eF = s3 - 0.5 * sn_slopefit(s2, s3, s4);
                wF = s2 - 0.5 * sn_slopefit(s1, s2, s3);
            }
            const double flux = (eF - wF) * a.xi[q] / a.dx + (nF - sF) * a.eta[q] / a.dy;
            double v = F(0, 0);
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
        const int blocks = 8 * (((nx * ny + 3) / 4 + 7) / 8);   // four cells each, in XCD order
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

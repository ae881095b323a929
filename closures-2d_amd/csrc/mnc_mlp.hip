// mnc_mlp.hip -- the learned warm start: an MLP 15 -> 128 -> 128 -> 128 -> 15
// with GELU, mapping the scaled moments u/u0 of an M4 closure to its scaled
// multipliers (alpha of the u0 = 1 problem), trained and evaluated with
// hand-written gfx950 kernels.  The reference has no counterpart (it starts
// every Newton solve from the previous step's alpha, momopt_update.cpp:236-238).
//
// Arithmetic: GEMM operands are rounded to bf16 (weights and activations) and
// multiplied on the matrix cores (v_mfma_f32_16x16x32_bf16) with FP32 sums; the
// master weights, biases, Adam moments and every elementwise step are FP32.
// Inputs and targets are standardised per component with statistics kept
// beside the weights.
//
// Layouts (row-major FP32): activations [rows][cols]; a layer's weight is
// W[in][out] (y = x W + b).  The parameter vector is
//   W1[15][128] b1[128] W2[128][128] b2[128] W3[128][128] b3[128] W4[128][15] b4[15]
// (37,007 values), followed by in_mean[15] in_std[15] out_mean[15] out_std[15].
#include <algorithm>

#include "mnc_internal.h"

namespace mnc {
namespace {

typedef float mf32x4 __attribute__((ext_vector_type(4)));
typedef __bf16 mbf16x8 __attribute__((ext_vector_type(8)));
typedef unsigned short mu16x8 __attribute__((ext_vector_type(8)));

__device__ __forceinline__ unsigned short to_bf16(float x) {
    const unsigned u = __float_as_uint(x);
    if ((u & 0x7f800000u) == 0x7f800000u) return (unsigned short)(u >> 16);   // inf/NaN kept
    return (unsigned short)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}

// GELU(x) = x Phi(x) and its derivative Phi(x) + x phi(x) (the exact, erf form).
__device__ __forceinline__ float gelu(float x) { return 0.5f * x * (1.0f + erff(x * 0.70710678118654752f)); }
__device__ __forceinline__ float gelu_grad(float x) {
    const float cdf = 0.5f * (1.0f + erff(x * 0.70710678118654752f));
    const float pdf = 0.39894228040143268f * expf(-0.5f * x * x);
    return cdf + x * pdf;
}

enum : int { EPI_STORE = 0, EPI_GELU = 1, EPI_DGELU = 2 };

// C[M][N] = A (M x K) * B (K x N) (+ bias), with general strides so one kernel
// serves x W, dY W^T and X^T dY.  One wave per 16 x 16 tile of C; K in steps of
// 32 (zero padded).  Lane l supplies A[i0 + l%16][k0 + 8(l/16) + t] and
// B[k0 + 8(l/16) + t][j0 + l%16] (t = 0..7) and receives
// C[i0 + 4(l/16) + r][j0 + l%16] (r = 0..3).
// Epilogues: EPI_STORE: C = AB + bias;  EPI_GELU: Z = AB + bias -> C, gelu(Z) -> C2;
// EPI_DGELU: C = AB * gelu'(Zaux).
struct GemmArgs {
    const float *A, *B, *bias, *aux;
    float *C, *C2;
    int M, N, K;
    const int *mdev;      // if set, M = min(M, *mdev) (row count known only on the device)
    long long sai, sak, sbk, sbj;
    int ldc;
    int epi;
};

__device__ __forceinline__ void gemm_tile(const GemmArgs &g, int tile, int tiles_n, int lane) {
    const int i0 = (tile / tiles_n) * 16, j0 = (tile % tiles_n) * 16;
    const int ia = i0 + (lane & 15), jb = j0 + (lane & 15), kq = 8 * (lane >> 4);
    mf32x4 acc = {0.f, 0.f, 0.f, 0.f};
    for (int k0 = 0; k0 < g.K; k0 += 32) {
        mu16x8 a, b;
#pragma unroll
        for (int t = 0; t < 8; ++t) {
            const int k = k0 + kq + t;
            const float av = (ia < g.M && k < g.K) ? g.A[ia * g.sai + k * g.sak] : 0.f;
            const float bv = (jb < g.N && k < g.K) ? g.B[k * g.sbk + jb * g.sbj] : 0.f;
            a[t] = to_bf16(av);
            b[t] = to_bf16(bv);
        }
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(mbf16x8, a), __builtin_bit_cast(mbf16x8, b),
                                                      acc, 0, 0, 0);
    }
    const int jc = j0 + (lane & 15);
    if (jc >= g.N) return;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int ic = i0 + 4 * (lane >> 4) + r;
        if (ic >= g.M) continue;
        float v = acc[r];
        const size_t o = size_t(ic) * g.ldc + jc;
        if (g.epi == EPI_DGELU) {
            g.C[o] = v * gelu_grad(g.aux[o]);
        } else {
            if (g.bias) v += g.bias[jc];
            g.C[o] = v;
            if (g.epi == EPI_GELU) g.C2[o] = gelu(v);
        }
    }
}

__global__ __launch_bounds__(256) void gemm_kernel(GemmArgs g) {
    const int lane = threadIdx.x & 63;
    const int tiles_n = (g.N + 15) >> 4;
    if (g.mdev) g.M = min(g.M, *g.mdev);
    const int tiles = ((g.M + 15) >> 4) * tiles_n;
    // grid-stride over the tiles: with the row count known only on the device
    // the grid is sized for the maximum, so blocks past the work leave at once
    for (int tile = blockIdx.x * 4 + (threadIdx.x >> 6); tile < tiles; tile += gridDim.x * 4)
        gemm_tile(g, tile, tiles_n, lane);
}

// db[j] = sum_i dZ[i][j]  (one block per column, tree sum)
__global__ __launch_bounds__(256) void colsum_kernel(const float *dz, int rows, int cols, float *db) {
    __shared__ float red[256];
    const int j = blockIdx.x;
    float s = 0.f;
    for (int i = threadIdx.x; i < rows; i += 256) s += dz[size_t(i) * cols + j];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) db[j] = red[0];
}

// SplitMix64 (Steele, Lea & Flood 2014): the sample index of row i at step s.
__device__ __forceinline__ unsigned long long splitmix64(unsigned long long x) {
    x += 0x9e3779b97f4a7c15ull;
    x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
    x = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
    return x ^ (x >> 31);
}

// Gathers a minibatch: X[i] = (u_s[idx] - in_mean) / in_std, T[i] likewise for
// the targets; idx drawn with replacement (or rows i0.. in order when seed = 0).
__global__ void gather_kernel(const double *us, const double *as, long long npairs, int rows,
                              unsigned long long seed, long long i0, const float *stats, float *X, float *T) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= rows) return;
    long long idx = seed ? (long long)(splitmix64(seed ^ (unsigned long long)i * 0xd1b54a32d192ed03ull) %
                                       (unsigned long long)npairs)
                         : (i0 + i);
    if (idx >= npairs) idx = npairs - 1;
#pragma unroll
    for (int k = 0; k < 15; ++k) {
        X[i * 15 + k] = float((us[idx * 15 + k] - stats[k]) / stats[15 + k]);
        if (T) T[i * 15 + k] = float((as[idx * 15 + k] - stats[30 + k]) / stats[45 + k]);
    }
}

// dY = 2 (Y - T) / (rows * 15); loss += sum (Y - T)^2 / (rows * 15)
__global__ __launch_bounds__(256) void mse_kernel(const float *Y, const float *T, int rows, float *dY, float *loss) {
    __shared__ float red[256];
    const int n = rows * 15;
    const float inv = 1.0f / float(n);
    float s = 0.f;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
        const float d = Y[i] - T[i];
        dY[i] = 2.0f * d * inv;
        s += d * d * inv;
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicAdd(loss, red[0]);
}

// Adam (Kingma & Ba 2015) on the flat parameter vector.
__global__ void adam_kernel(float *w, const float *grad, float *m, float *v, int n, float lr, float b1, float b2,
                            float eps, float c1, float c2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = grad[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    w[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// Listed rows: X[i] = standardised u[cell]/u[cell][0], cell = list[i], i < *count.
__global__ void gather_list_kernel(const double *u, const int *list, const int *count, const float *stats,
                                   float *X) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= *count) return;
    const double *uc = u + size_t(list[i]) * 15;
    const double r = uc[0] > 0.0 ? 1.0 / uc[0] : 0.0;
#pragma unroll
    for (int k = 0; k < 15; ++k) X[i * 15 + k] = float((uc[k] * r - stats[k]) / stats[15 + k]);
}

// Listed rows: alpha[cell] = Y[i] * out_std + out_mean, cell = list[i].
__global__ void unscale_list_kernel(const float *Y, const int *list, const int *count, const float *stats,
                                    double *alpha) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= *count * 15) return;
    const int row = i / 15, k = i % 15;
    alpha[size_t(list[row]) * 15 + k] = double(Y[i]) * double(stats[45 + k]) + double(stats[30 + k]);
}

// Scaled multipliers from the network output: alpha_s = Y * out_std + out_mean.
__global__ void unscale_kernel(const float *Y, int rows, const float *stats, double *alpha) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= rows * 15) return;
    const int k = i % 15;
    alpha[i] = double(Y[i]) * double(stats[45 + k]) + double(stats[30 + k]);
}

// Scaled inputs of the Newton prologue: u_s = u / u0 for every cell.
__global__ void scale_moments_kernel(const double *u, int ncells, double *us) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ncells) return;
    const double u0 = u[size_t(i) * 15];
    const double r = u0 > 0.0 ? 1.0 / u0 : 0.0;
    for (int k = 0; k < 15; ++k) us[size_t(i) * 15 + k] = u[size_t(i) * 15 + k] * r;
}

}  // namespace

// ---- host side ---------------------------------------------------------------

namespace {
constexpr int kIn = 15, kH = 128, kOut = 15;
constexpr int OFF_W1 = 0, OFF_B1 = OFF_W1 + kIn * kH, OFF_W2 = OFF_B1 + kH, OFF_B2 = OFF_W2 + kH * kH,
              OFF_W3 = OFF_B2 + kH, OFF_B3 = OFF_W3 + kH * kH, OFF_W4 = OFF_B3 + kH, OFF_B4 = OFF_W4 + kH * kOut,
              NPARAM = OFF_B4 + kOut;
static_assert(NPARAM == 37007, "parameter count");

void gemm(hipStream_t s, const float *A, long long sai, long long sak, const float *B, long long sbk, long long sbj,
          float *C, int ldc, int M, int N, int K, const float *bias, int epi, float *C2, const float *aux,
          const int *mdev = nullptr) {
    GemmArgs g{A, B, bias, aux, C, C2, M, N, K, mdev, sai, sak, sbk, sbj, ldc, epi};
    const int tiles = ((M + 15) / 16) * ((N + 15) / 16);
    const int blocks = std::min((tiles + 3) / 4, 2048);     // 8 blocks per CU at 256 CUs; tiles beyond loop
    hipLaunchKernelGGL(gemm_kernel, dim3(blocks), dim3(256), 0, s, g);
}
}  // namespace

int mlp_param_count() { return NPARAM; }

// Scratch: for `rows` rows, X[rows*15], T[rows*15], Z1..Z3, A1..A3 [rows*128],
// Y, dY [rows*15], dA/dZ [2][rows*128].
size_t mlp_scratch_floats(int rows) { return size_t(rows) * (15 * 4 + 128 * 8); }

// rows: the scratch layout's row count (and the GEMMs' M unless mdev gives fewer).
void mlp_forward(const MlpBuffers &b, int rows, hipStream_t s, const int *mdev) {
    float *X = b.scratch, *Z1 = X + size_t(rows) * 30, *A1 = Z1 + size_t(rows) * kH, *Z2 = A1 + size_t(rows) * kH,
          *A2 = Z2 + size_t(rows) * kH, *Z3 = A2 + size_t(rows) * kH, *A3 = Z3 + size_t(rows) * kH,
          *Y = A3 + size_t(rows) * kH;
    const float *W = b.w;
    gemm(s, X, kIn, 1, W + OFF_W1, kH, 1, Z1, kH, rows, kH, kIn, W + OFF_B1, EPI_GELU, A1, nullptr, mdev);
    gemm(s, A1, kH, 1, W + OFF_W2, kH, 1, Z2, kH, rows, kH, kH, W + OFF_B2, EPI_GELU, A2, nullptr, mdev);
    gemm(s, A2, kH, 1, W + OFF_W3, kH, 1, Z3, kH, rows, kH, kH, W + OFF_B3, EPI_GELU, A3, nullptr, mdev);
    gemm(s, A3, kH, 1, W + OFF_W4, kOut, 1, Y, kOut, rows, kOut, kH, W + OFF_B4, EPI_STORE, nullptr, nullptr, mdev);
}

void mlp_predict_listed(const MlpBuffers &b, const double *u, const int *list, const int *count, int max_rows,
                        double *alpha, hipStream_t s) {
    hipLaunchKernelGGL(gather_list_kernel, dim3((max_rows + 255) / 256), dim3(256), 0, s, u, list, count,
                       b.w + NPARAM, b.scratch);
    mlp_forward(b, max_rows, s, count);
    hipLaunchKernelGGL(unscale_list_kernel, dim3((max_rows * 15 + 255) / 256), dim3(256), 0, s,
                       mlp_output(b, max_rows), list, count, b.w + NPARAM, alpha);
}

float *mlp_output(const MlpBuffers &b, int rows) { return b.scratch + size_t(rows) * (30 + 6 * kH); }
float *mlp_input(const MlpBuffers &b) { return b.scratch; }

// Gradient of the MSE loss on the last forward pass (targets in T); gradients
// of all parameters into b.grad; loss accumulated into *b.loss.
void mlp_backward(const MlpBuffers &b, int rows, hipStream_t s) {
    const size_t R = size_t(rows);
    float *X = b.scratch, *T = X + R * kIn, *Z1 = X + R * 30, *A1 = Z1 + R * kH, *Z2 = A1 + R * kH,
          *A2 = Z2 + R * kH, *Z3 = A2 + R * kH, *A3 = Z3 + R * kH, *Y = A3 + R * kH, *dY = Y + R * kOut,
          *dA = dY + R * kOut, *dZ = dA + R * kH;
    const float *W = b.w;
    float *G = b.grad;
    hipLaunchKernelGGL(mse_kernel, dim3(64), dim3(256), 0, s, Y, T, rows, dY, b.loss);
    // layer 4: dW4 = A3^T dY, db4, dA3 = dY W4^T -> dZ3 = dA3 * gelu'(Z3)
    gemm(s, A3, 1, kH, dY, kOut, 1, G + OFF_W4, kOut, kH, kOut, rows, nullptr, EPI_STORE, nullptr, nullptr);
    hipLaunchKernelGGL(colsum_kernel, dim3(kOut), dim3(256), 0, s, dY, rows, kOut, G + OFF_B4);
    gemm(s, dY, kOut, 1, W + OFF_W4, 1, kOut, dZ, kH, rows, kH, kOut, nullptr, EPI_DGELU, nullptr, Z3);
    // layer 3
    gemm(s, A2, 1, kH, dZ, kH, 1, G + OFF_W3, kH, kH, kH, rows, nullptr, EPI_STORE, nullptr, nullptr);
    hipLaunchKernelGGL(colsum_kernel, dim3(kH), dim3(256), 0, s, dZ, rows, kH, G + OFF_B3);
    gemm(s, dZ, kH, 1, W + OFF_W3, 1, kH, dA, kH, rows, kH, kH, nullptr, EPI_DGELU, nullptr, Z2);
    // layer 2
    gemm(s, A1, 1, kH, dA, kH, 1, G + OFF_W2, kH, kH, kH, rows, nullptr, EPI_STORE, nullptr, nullptr);
    hipLaunchKernelGGL(colsum_kernel, dim3(kH), dim3(256), 0, s, dA, rows, kH, G + OFF_B2);
    gemm(s, dA, kH, 1, W + OFF_W2, 1, kH, dZ, kH, rows, kH, kH, nullptr, EPI_DGELU, nullptr, Z1);
    // layer 1
    gemm(s, X, 1, kIn, dZ, kH, 1, G + OFF_W1, kH, kIn, kH, rows, nullptr, EPI_STORE, nullptr, nullptr);
    hipLaunchKernelGGL(colsum_kernel, dim3(kH), dim3(256), 0, s, dZ, rows, kH, G + OFF_B1);
}

void mlp_gather(const MlpBuffers &b, const double *us, const double *as, long long npairs, int rows,
                unsigned long long seed, long long i0, hipStream_t s) {
    float *X = b.scratch, *T = as ? X + size_t(rows) * kIn : nullptr;
    hipLaunchKernelGGL(gather_kernel, dim3((rows + 255) / 256), dim3(256), 0, s, us, as, npairs, rows, seed, i0,
                       b.w + NPARAM, X, T);
}

void mlp_adam(const MlpBuffers &b, float lr, int step, hipStream_t s) {
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const float c1 = 1.0f - powf(b1, float(step)), c2 = 1.0f - powf(b2, float(step));
    hipLaunchKernelGGL(adam_kernel, dim3((NPARAM + 255) / 256), dim3(256), 0, s, b.w, b.grad, b.m, b.v, NPARAM, lr,
                       b1, b2, eps, c1, c2);
}

void mlp_unscale(const MlpBuffers &b, int rows, double *alpha, hipStream_t s) {
    hipLaunchKernelGGL(unscale_kernel, dim3((rows * 15 + 255) / 256), dim3(256), 0, s, mlp_output(b, rows), rows,
                       b.w + NPARAM, alpha);
}

void mlp_scale_moments(const double *u, int ncells, double *us, hipStream_t s) {
    hipLaunchKernelGGL(scale_moments_kernel, dim3((ncells + 255) / 256), dim3(256), 0, s, u, ncells, us);
}

}  // namespace mnc

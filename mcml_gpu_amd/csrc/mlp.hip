/*
 * mlp.hip - the reflectance surrogate: a fully connected network
 *     x[14] -> (Linear 256 + GELU) x 4 -> Linear 3 (per-SDS log reflectance)
 * trained and evaluated with the project's own kernels.  No counterpart in the reference.
 *
 * Every matrix product runs on the matrix cores: v_mfma_f32_16x16x16_bf16 takes bf16
 * operands (weights and activations rounded to bf16 as they are loaded) and accumulates
 * in FP32.  A wavefront computes a 32x32 block of C as 2x2 tiles of 16x16; a workgroup
 * of four wavefronts covers 64x64.  Operands are read straight from global memory (L1/L2
 * resident at these sizes) through strides, so one kernel serves
 *     forward     Z = X W^T + b,  A = GELU(Z)
 *     backward    dA_prev = dZ W,  dZ = dA * GELU'(Z)    (fused epilogue)
 *     weight grad dW = dZ^T A_prev  (split over the batch, partials reduced in a fixed order)
 * Master weights and Adam moments are FP32.
 */
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "mlp.h"

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

namespace {

/* fp32 -> bf16, round to nearest even (NaN kept quiet). */
__device__ __forceinline__ short to_bf16(float x)
{
    uint32_t u = __float_as_uint(x);
    if ((u & 0x7F800000u) == 0x7F800000u && (u & 0x007FFFFFu)) return (short)((u >> 16) | 0x40u);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (short)(u >> 16);
}

__device__ __forceinline__ float gelu(float z) { return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f)); }

__device__ __forceinline__ float gelu_grad(float z)
{
    const float cdf = 0.5f * (1.0f + erff(z * 0.70710678118654752f));
    const float pdf = 0.39894228040143268f * expf(-0.5f * z * z);
    return cdf + z * pdf;
}

} // namespace

/* C = A B with A(m,k) = A[m*sam + k*sak], B(k,n) = B[k*sbk + n*sbn] over k in
 * [k_lo, k_hi), then the epilogue.  grid = (ceil(N/64), ceil(M/64), splits). */
extern "C" __global__ void __launch_bounds__(256) mcml_mlp_gemm(MlpGemmArgs g)
{
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int m0 = blockIdx.y * 64 + (wave >> 1) * 32;
    const int n0 = blockIdx.x * 64 + (wave & 1) * 32;
    /* split-K: this workgroup's share of the reduction */
    const int kper = (g.K + (int)gridDim.z - 1) / (int)gridDim.z;
    const int kbeg = (int)blockIdx.z * kper;
    const int kend = min(g.K, kbeg + kper);
    const int r = lane & 15;   /* row of A / column of B this lane feeds */
    const int kq = lane >> 4;  /* which 4-wide slice of the 16-deep k step */
    f32x4 acc[2][2];
    for (int i = 0; i < 2; i++)
        for (int j = 0; j < 2; j++) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
    for (int k0 = kbeg; k0 < kend; k0 += 16) {
        bf16x4 a[2], b[2];
        for (int t = 0; t < 2; t++) {
            const int m = m0 + 16 * t + r;
            const int n = n0 + 16 * t + r;
            for (int j = 0; j < 4; j++) {
                const int k = k0 + 4 * kq + j;
                const bool kin = k < kend;
                const float av = (kin && m < g.M) ? g.A[(int64_t)m * g.sam + (int64_t)k * g.sak] : 0.0f;
                const float bv = (kin && n < g.N) ? g.B[(int64_t)k * g.sbk + (int64_t)n * g.sbn] : 0.0f;
                a[t][j] = to_bf16(av);
                b[t][j] = to_bf16(bv);
            }
        }
        for (int i = 0; i < 2; i++)
            for (int j = 0; j < 2; j++) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[i], b[j], acc[i][j], 0, 0, 0);
    }
    /* lane holds C[4*kq + e][r] of each 16x16 tile */
    for (int i = 0; i < 2; i++)
        for (int j = 0; j < 2; j++)
            for (int e = 0; e < 4; e++) {
                const int m = m0 + 16 * i + 4 * kq + e;
                const int n = n0 + 16 * j + r;
                if (m >= g.M || n >= g.N) continue;
                float v = acc[i][j][e];
                const int64_t o = (int64_t)m * g.ldc + n;
                switch (g.epilogue) {
                case MLP_EPI_STORE: g.C[o] = v; break;
                case MLP_EPI_SPLIT: g.C[(int64_t)blockIdx.z * g.split_stride + o] = v; break;
                case MLP_EPI_BIAS: g.C[o] = v + g.bias[n]; break;
                case MLP_EPI_BIAS_GELU: {
                    const float z = v + g.bias[n];
                    if (g.Z != nullptr) g.Z[o] = z;
                    g.C[o] = gelu(z);
                    break;
                }
                case MLP_EPI_DGELU: g.C[o] = v * gelu_grad(g.Z[o]); break;
                default: break;
                }
            }
}

/* out[i] = sum over s of part[s*stride + i] (fixed order), i < n. */
extern "C" __global__ void __launch_bounds__(256) mcml_mlp_reduce(const float* part, int splits, int64_t stride,
                                                                  int64_t n, float* out)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        float s = 0.0f;
        for (int k = 0; k < splits; k++) s += part[(int64_t)k * stride + i];
        out[i] = s;
    }
}

/* Column sums of dZ [M x N] (bias gradient): one thread per column, rows in order. */
extern "C" __global__ void __launch_bounds__(256) mcml_mlp_colsum(const float* dz, int M, int N, float* out)
{
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    float s = 0.0f;
    for (int m = 0; m < M; m++) s += dz[(int64_t)m * N + n];
    out[n] = s;
}

/* Weighted squared error with one weight per target: dY = 2 w (Y - T) / (B * outs), and the
 * row's mean of w (Y - T)^2 into rowloss. */
extern "C" __global__ void __launch_bounds__(256) mcml_mlp_loss(const float* y, const float* t, const float* w, int B,
                                                                int outs, float* dy, float* rowloss)
{
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= B) return;
    const float scale = 2.0f / (float)(B * outs);
    float l = 0.0f;
    for (int o = 0; o < outs; o++) {
        const int64_t i = (int64_t)m * outs + o;
        const float wm = w ? w[i] : 1.0f;
        const float d = y[i] - t[i];
        dy[i] = (scale * wm) * d;
        l += wm * d * d;
    }
    rowloss[m] = l / (float)outs;
}

/* Adam (Kingma & Ba 2015) with bias correction folded into the step size. */
extern "C" __global__ void __launch_bounds__(256) mcml_mlp_adam(float* p, const float* g, float* m1, float* m2, int64_t n,
                                                                float lr_t, float b1, float b2, float eps, float wd)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const float gi = g[i] + wd * p[i];
        const float a = b1 * m1[i] + (1.0f - b1) * gi;
        const float b = b2 * m2[i] + (1.0f - b2) * gi * gi;
        m1[i] = a;
        m2[i] = b;
        p[i] -= lr_t * a / (sqrtf(b) + eps);
    }
}

/* dst[i][:] = src[idx[i]][:] (mini-batch assembly). */
extern "C" __global__ void __launch_bounds__(256) mcml_mlp_gather(const float* src, const int32_t* idx, int rows,
                                                                  int cols, float* dst)
{
    const int64_t total = (int64_t)rows * cols;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
        const int64_t r = t / cols, c = t - r * cols;
        dst[t] = src[(int64_t)idx[r] * cols + c];
    }
}

/* ------------------------------------------------------------------ launchers */
hipError_t mcml_mlp_launch_gather(const float* src, const int32_t* idx, int rows, int cols, float* dst, hipStream_t st)
{
    const int64_t blocks = ((int64_t)rows * cols + 255) / 256;
    hipLaunchKernelGGL(mcml_mlp_gather, dim3((unsigned)(blocks < 4096 ? blocks : 4096)), dim3(256), 0, st, src, idx,
                       rows, cols, dst);
    return hipGetLastError();
}

hipError_t mcml_mlp_launch_gemm(const MlpGemmArgs& g, int splits, hipStream_t st)
{
    dim3 grid((g.N + 63) / 64, (g.M + 63) / 64, splits);
    hipLaunchKernelGGL(mcml_mlp_gemm, grid, dim3(256), 0, st, g);
    return hipGetLastError();
}

hipError_t mcml_mlp_launch_reduce(const float* part, int splits, int64_t stride, int64_t n, float* out, hipStream_t st)
{
    const int64_t blocks = (n + 255) / 256;
    hipLaunchKernelGGL(mcml_mlp_reduce, dim3((unsigned)(blocks < 4096 ? blocks : 4096)), dim3(256), 0, st, part, splits,
                       stride, n, out);
    return hipGetLastError();
}

hipError_t mcml_mlp_launch_colsum(const float* dz, int M, int N, float* out, hipStream_t st)
{
    hipLaunchKernelGGL(mcml_mlp_colsum, dim3((N + 255) / 256), dim3(256), 0, st, dz, M, N, out);
    return hipGetLastError();
}

hipError_t mcml_mlp_launch_loss(const float* y, const float* t, const float* w, int B, int outs, float* dy,
                                float* rowloss, hipStream_t st)
{
    hipLaunchKernelGGL(mcml_mlp_loss, dim3((B + 255) / 256), dim3(256), 0, st, y, t, w, B, outs, dy, rowloss);
    return hipGetLastError();
}

hipError_t mcml_mlp_launch_adam(float* p, const float* g, float* m1, float* m2, int64_t n, float lr_t, float b1,
                                float b2, float eps, float wd, hipStream_t st)
{
    const int64_t blocks = (n + 255) / 256;
    hipLaunchKernelGGL(mcml_mlp_adam, dim3((unsigned)(blocks < 4096 ? blocks : 4096)), dim3(256), 0, st, p, g, m1, m2,
                       n, lr_t, b1, b2, eps, wd);
    return hipGetLastError();
}

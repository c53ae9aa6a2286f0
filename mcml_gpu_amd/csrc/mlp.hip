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

/* GELU in its tanh form (Hendrycks & Gimpel, "Gaussian Error Linear Units", 2016; PyTorch's
 * gelu(approximate="tanh")): 0.5 z (1 + tanh(k (z + c z^3))) = z sigmoid(a),
 * a = 2k (z + c z^3), k = sqrt(2/pi), c = 0.044715.  The sigmoid is 1 / (1 + 2^(-a log2 e))
 * with the hardware exp2 and reciprocal (a few ulp, far below the bf16 rounding the
 * activations then get), log2(e) folded into the constants: 2^(z (GELU_A + GELU_B z^2)). */
constexpr float GELU_A = -2.3022082f;   /* -2k log2(e) */
constexpr float GELU_B = -0.10294324f;  /* -2kc log2(e) */
constexpr float GELU_K2 = 1.5957691f;   /* 2k: a'(z) = 2k + 6kc z^2 */
constexpr float GELU_K6C = 0.21406445f; /* 6kc */

__device__ __forceinline__ float gelu(float z)
{
    const float e = __builtin_amdgcn_exp2f(z * fmaf(z * z, GELU_B, GELU_A));
    return z * __builtin_amdgcn_rcpf(1.0f + e);
}

/* d/dz z s(a) = s + z s (1 - s) a'(z), s = sigmoid(a) */
__device__ __forceinline__ float gelu_grad(float z)
{
    const float z2 = z * z;
    const float s = __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(z * fmaf(z2, GELU_B, GELU_A)));
    return fmaf(z * s * (1.0f - s), fmaf(z2, GELU_K6C, GELU_K2), s);
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

/* Partial column sums of dZ [M x N] (bias gradient) over row range blockIdx.y of `rows`
 * rows: 64 columns per workgroup, four interleaved row slices per column summed with four
 * independent accumulators, combined in a fixed order.  part[blockIdx.y * N + n]. */
extern "C" __global__ void __launch_bounds__(256) mcml_mlp_colsum(const float* dz, int M, int N, int rows, float* part)
{
    __shared__ float slice_sum[4][64];
    const int c = threadIdx.x & 63, slice = threadIdx.x >> 6;
    const int n = blockIdx.x * 64 + c;
    const int beg = blockIdx.y * rows, end = min(M, beg + rows);
    float acc[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    if (n < N) {
        int m = beg + slice;
        for (; m + 12 < end; m += 16)
            for (int u = 0; u < 4; u++) acc[u] += dz[(int64_t)(m + 4 * u) * N + n];
        for (int u = 0; m < end; m += 4, u++) acc[u & 3] += dz[(int64_t)m * N + n];
    }
    slice_sum[slice][c] = (acc[0] + acc[1]) + (acc[2] + acc[3]);
    __syncthreads();
    if (slice == 0 && n < N)
        part[(int64_t)blockIdx.y * N + n] = (slice_sum[0][c] + slice_sum[1][c]) + (slice_sum[2][c] + slice_sum[3][c]);
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

/* ------------------------------------------------------------------ fused inference
 * Transposed formulation: each layer computes Z^T = W X^T, so a 16x16 output tile held by
 * a lane as rows n = 4q + e (q = lane / 16, e = 0..3) of column m = lane % 16 is already
 * the B operand the next layer's v_mfma_f32_16x16x32_bf16 wants for its k = n: two adjacent
 * output tiles make one 32-deep k step, with k order {4q + e, 16 + 4q + e} that the packed
 * weights follow.  Activations therefore stay in registers from the first layer to the
 * last; each layer's bf16 weights are staged once per workgroup into LDS (128 KB for a
 * 256 x 256 layer) and every A-operand read there feeds two row blocks.  A workgroup of
 * eight wavefronts covers 256 rows, 32 per wavefront. */
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

namespace {

constexpr int FW_KT = MLP_FUSED_H / 32; /* 32-deep k steps of a hidden layer */
constexpr int FW_MB = 2;                /* 16-row blocks per wavefront */

__device__ __forceinline__ bf16x8 as_bf16x8(uint4 v) { return __builtin_bit_cast(bf16x8, v); }

/* Copy N 16-byte words of packed weights into LDS with direct global-to-LDS loads
 * (global_load_lds_dwordx4: lane i of a wavefront writes the 16 bytes it loads to the
 * wavefront's LDS base + 16 i).  No VGPRs hold the data, so all N / T loads of a lane are
 * in flight together; the barrier that follows waits for them.  (Through registers, with
 * the activations of the layer occupying most of them, the compiler issued one load,
 * waited for it and stored it, sixteen times in a row per hidden layer.) */
template <int N>
__device__ __forceinline__ void stage(uint4* lds, const uint4* g)
{
    constexpr int T = MLP_FUSED_WAVES * 64, PER = N / T;
    static_assert(N % T == 0, "stage size must be a multiple of the workgroup");
    const int wave_base = threadIdx.x & ~63;
#pragma unroll
    for (int i = 0; i < PER; i++)
        __builtin_amdgcn_global_load_lds(static_cast<const void*>(g + threadIdx.x + i * T),
                                         (__attribute__((address_space(3))) void*)(lds + i * T + wave_base), 16, 0,
                                         0);
}

/* One GELU layer from KT_IN k steps of input to MLP_FUSED_H outputs. */
template <int KT_IN>
__device__ __forceinline__ void fused_hidden(const uint4* w, const float* __restrict__ bias,
                                             const bf16x8 (&in)[KT_IN][FW_MB], bf16x8 (&out)[FW_KT][FW_MB], int r, int q)
{
#pragma unroll
    for (int c = 0; c < FW_KT; c++) {
        f32x4 acc[2][FW_MB];
#pragma unroll
        for (int mb = 0; mb < FW_MB; mb++) acc[0][mb] = acc[1][mb] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int t = 0; t < KT_IN; t++) {
            const bf16x8 a0 = as_bf16x8(w[(t * 4 + q) * MLP_FUSED_H + 32 * c + r]);
            const bf16x8 a1 = as_bf16x8(w[(t * 4 + q) * MLP_FUSED_H + 32 * c + 16 + r]);
#pragma unroll
            for (int mb = 0; mb < FW_MB; mb++) {
                acc[0][mb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a0, in[t][mb], acc[0][mb], 0, 0, 0);
                acc[1][mb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1, in[t][mb], acc[1][mb], 0, 0, 0);
            }
        }
        const float4 b0 = *reinterpret_cast<const float4*>(bias + 32 * c + 4 * q);
        const float4 b1 = *reinterpret_cast<const float4*>(bias + 32 * c + 16 + 4 * q);
        const float bb0[4] = {b0.x, b0.y, b0.z, b0.w}, bb1[4] = {b1.x, b1.y, b1.z, b1.w};
#pragma unroll
        for (int mb = 0; mb < FW_MB; mb++) {
            bf16x8 v;
#pragma unroll
            for (int e = 0; e < 4; e++) {
                v[e] = (__bf16)gelu(acc[0][mb][e] + bb0[e]);
                v[4 + e] = (__bf16)gelu(acc[1][mb][e] + bb1[e]);
            }
            out[c][mb] = v;
        }
    }
}

} // namespace

extern "C" __global__ void __launch_bounds__(MLP_FUSED_WAVES * 64) mcml_mlp_fused(MlpFusedArgs a)
{
    __shared__ uint4 wlds[FW_KT * 4 * MLP_FUSED_H];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int r = lane & 15, q = lane >> 4;
    const int row0 = blockIdx.x * MLP_FUSED_ROWS + wave * (FW_MB * 16);
    const uint4* pk = reinterpret_cast<const uint4*>(a.packed);

    /* first layer's B operand straight from x: k = 4q + j (j < 4), 12 + 4q + j (j >= 4).
     * Addresses are clamped into the array and out-of-range values zeroed afterwards, so
     * all sixteen loads are in flight together (B >= 1 and in >= 1). */
    float xv[FW_MB][8];
#pragma unroll
    for (int mb = 0; mb < FW_MB; mb++) {
        const int m = min(row0 + 16 * mb + r, a.B - 1);
#pragma unroll
        for (int j = 0; j < 8; j++) {
            const int k = min(4 * q + (j < 4 ? j : 12 + j), a.in - 1);
            xv[mb][j] = a.x[(int64_t)m * a.in + k];
        }
    }
    bf16x8 x0[1][FW_MB];
#pragma unroll
    for (int mb = 0; mb < FW_MB; mb++) {
        const bool row_in = row0 + 16 * mb + r < a.B;
#pragma unroll
        for (int j = 0; j < 8; j++) {
            const bool k_in = 4 * q + (j < 4 ? j : 12 + j) < a.in;
            x0[0][mb][j] = (__bf16)((row_in && k_in) ? xv[mb][j] : 0.0f);
        }
    }
    stage<4 * MLP_FUSED_H>(wlds, pk + a.w_off[0] / 8);
    __syncthreads();
    bf16x8 act[FW_KT][FW_MB];
    fused_hidden<1>(wlds, a.bias + a.b_off[0], x0, act, r, q);
    for (int l = 1; l < a.hidden; l++) {
        __syncthreads();
        stage<FW_KT * 4 * MLP_FUSED_H>(wlds, pk + a.w_off[l] / 8);
        __syncthreads();
        bf16x8 nxt[FW_KT][FW_MB];
        fused_hidden<FW_KT>(wlds, a.bias + a.b_off[l], act, nxt, r, q);
#pragma unroll
        for (int t = 0; t < FW_KT; t++)
#pragma unroll
            for (int mb = 0; mb < FW_MB; mb++) act[t][mb] = nxt[t][mb];
    }
    /* linear output layer, one 16-row block of outputs */
    __syncthreads();
    stage<FW_KT * 4 * 16>(wlds, pk + a.w_off[a.hidden] / 8);
    __syncthreads();
    const float* bl = a.bias + a.b_off[a.hidden];
    const float4 bq = *reinterpret_cast<const float4*>(bl + 4 * q); /* padded to 16 outputs */
    f32x4 acc[FW_MB];
#pragma unroll
    for (int mb = 0; mb < FW_MB; mb++) acc[mb] = f32x4{bq.x, bq.y, bq.z, bq.w};
#pragma unroll
    for (int t = 0; t < FW_KT; t++) {
        const bf16x8 w = as_bf16x8(wlds[(t * 4 + q) * 16 + r]);
#pragma unroll
        for (int mb = 0; mb < FW_MB; mb++) acc[mb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w, act[t][mb], acc[mb], 0, 0, 0);
    }
#pragma unroll
    for (int mb = 0; mb < FW_MB; mb++) {
        const int m = row0 + 16 * mb + r;
#pragma unroll
        for (int e = 0; e < 4; e++) {
            const int n = 4 * q + e;
            if (m < a.B && n < a.outs) a.y[(int64_t)m * a.outs + n] = acc[mb][e];
        }
    }
}

/* Pack one layer (W [out x in] fp32, b [out]) into the fused kernel's operand order. */
extern "C" __global__ void __launch_bounds__(256) mcml_mlp_pack(const float* W, const float* b, int in, int out,
                                                                int Kp, int Np, uint16_t* dst, float* dst_bias)
{
    const int64_t total = (int64_t)Kp * Np;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
        const int j = (int)(e & 7);
        const int n = (int)((e >> 3) % Np);
        const int64_t tq = (e >> 3) / Np;
        const int qq = (int)(tq & 3);
        const int t = (int)(tq >> 2);
        const int k = 32 * t + 4 * qq + (j < 4 ? j : 12 + j);
        const float v = (n < out && k < in) ? W[(int64_t)n * in + k] : 0.0f;
        dst[e] = (uint16_t)to_bf16(v);
        if (e < Np) dst_bias[e] = e < out ? b[e] : 0.0f;
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

hipError_t mcml_mlp_launch_colsum(const float* dz, int M, int N, int splits, float* part, hipStream_t st)
{
    const int rows = (M + splits - 1) / splits;
    hipLaunchKernelGGL(mcml_mlp_colsum, dim3((N + 63) / 64, splits), dim3(256), 0, st, dz, M, N, rows, part);
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

hipError_t mcml_mlp_launch_pack(const float* W, const float* b, int in, int out, int Kp, int Np, uint16_t* dst,
                                float* dst_bias, hipStream_t st)
{
    const int64_t blocks = ((int64_t)Kp * Np + 255) / 256;
    hipLaunchKernelGGL(mcml_mlp_pack, dim3((unsigned)(blocks < 4096 ? blocks : 4096)), dim3(256), 0, st, W, b, in, out,
                       Kp, Np, dst, dst_bias);
    return hipGetLastError();
}

hipError_t mcml_mlp_launch_fused(const MlpFusedArgs& a, hipStream_t st)
{
    const int blocks = (a.B + MLP_FUSED_ROWS - 1) / MLP_FUSED_ROWS;
    hipLaunchKernelGGL(mcml_mlp_fused, dim3(blocks), dim3(MLP_FUSED_WAVES * 64), 0, st, a);
    return hipGetLastError();
}

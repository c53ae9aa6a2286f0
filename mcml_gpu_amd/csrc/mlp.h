/* Surrogate network kernels (internal to libmcml). */
#ifndef MCML_MLP_H
#define MCML_MLP_H
#include <hip/hip_runtime.h>
#include <stdint.h>

enum { MLP_EPI_STORE = 0, MLP_EPI_SPLIT = 1, MLP_EPI_BIAS = 2, MLP_EPI_BIAS_GELU = 3, MLP_EPI_DGELU = 4 };

struct MlpGemmArgs {
    const float* A;
    const float* B;
    float* C;
    const float* bias;   /* [N] (BIAS, BIAS_GELU) */
    float* Z;            /* pre-activation out (BIAS_GELU, may be NULL) or in (DGELU) */
    int64_t sam, sak, sbk, sbn, ldc;
    int64_t split_stride; /* elements between split-K partial planes (SPLIT) */
    int32_t M, N, K, epilogue;
};

hipError_t mcml_mlp_launch_gemm(const MlpGemmArgs& g, int splits, hipStream_t st);

/* Fused inference of a network with in <= 32 inputs, 1..MLP_FUSED_MAX_HIDDEN GELU layers of
 * width MLP_FUSED_H and outs <= 16 linear outputs, all in one kernel.  Weights are packed
 * once per parameter change by mcml_mlp_launch_pack into bf16 operand order:
 *   layer l, padded to Kp (32 for the first layer, MLP_FUSED_H after) x Np (MLP_FUSED_H,
 *   16 for the last), element [t][q][n][j] = W[n][32t + 4q + (j < 4 ? j : 12 + j)]
 *   (zero outside the real in x out), biases zero-padded to Np. */
#define MLP_FUSED_H 256
#define MLP_FUSED_MAX_HIDDEN 8
#define MLP_FUSED_WAVES 8
#define MLP_FUSED_ROWS (MLP_FUSED_WAVES * 32)

struct MlpFusedArgs {
    const float* x;                               /* [B x in] */
    float* y;                                     /* [B x outs] */
    const uint16_t* packed;                       /* packed bf16 weights */
    const float* bias;                            /* packed biases */
    int64_t w_off[MLP_FUSED_MAX_HIDDEN + 1];      /* per layer, in bf16 elements */
    int64_t b_off[MLP_FUSED_MAX_HIDDEN + 1];      /* per layer, in floats */
    int32_t B, in, outs, hidden;
};

hipError_t mcml_mlp_launch_pack(const float* W, const float* b, int in, int out, int Kp, int Np, uint16_t* dst,
                                float* dst_bias, hipStream_t st);
hipError_t mcml_mlp_launch_fused(const MlpFusedArgs& a, hipStream_t st);
hipError_t mcml_mlp_launch_reduce(const float* part, int splits, int64_t stride, int64_t n, float* out, hipStream_t st);
/* part[s * N + n] = sum of dz[m][n] over the s-th of `splits` row ranges */
hipError_t mcml_mlp_launch_colsum(const float* dz, int M, int N, int splits, float* part, hipStream_t st);
hipError_t mcml_mlp_launch_loss(const float* y, const float* t, const float* w, int B, int outs, float* dy,
                                float* rowloss, hipStream_t st);
hipError_t mcml_mlp_launch_adam(float* p, const float* g, float* m1, float* m2, int64_t n, float lr_t, float b1,
                                float b2, float eps, float wd, hipStream_t st);
hipError_t mcml_mlp_launch_gather(const float* src, const int32_t* idx, int rows, int cols, float* dst, hipStream_t st);

#endif

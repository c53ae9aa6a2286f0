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
hipError_t mcml_mlp_launch_reduce(const float* part, int splits, int64_t stride, int64_t n, float* out, hipStream_t st);
hipError_t mcml_mlp_launch_colsum(const float* dz, int M, int N, float* out, hipStream_t st);
hipError_t mcml_mlp_launch_loss(const float* y, const float* t, const float* w, int B, int outs, float* dy,
                                float* rowloss, hipStream_t st);
hipError_t mcml_mlp_launch_adam(float* p, const float* g, float* m1, float* m2, int64_t n, float lr_t, float b1,
                                float b2, float eps, float wd, hipStream_t st);
hipError_t mcml_mlp_launch_gather(const float* src, const int32_t* idx, int rows, int cols, float* dst, hipStream_t st);

#endif

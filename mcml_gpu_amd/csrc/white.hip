/*
 * white.hip - White Monte Carlo: detected weight at new absorption coefficients from the
 * recorded per-layer pathlengths, without new transport.
 *
 * For a detected photon with weight w and pathlength L_l in layer l, simulated with
 * absorption mua_l, the weight it would carry at mua'_l is w * exp(-sum_l (mua'_l - mua_l) L_l).
 * This is the estimator the reference uses for its average pathlength (transport.cu:
 * 285-292), and it is exact in expectation for the reference's transport: with mus fixed,
 * the step-length density contributes (mut'/mut)^n and the per-interaction drops
 * (mus/mut')^n / (mus/mut)^n = (mut/mut')^n over n interactions, so only the Beer-Lambert
 * factor remains; roulette is unbiased for any weight history.
 *
 * One workgroup per (set, SDS, mua' vector) sums its group's records in FP64 in a fixed
 * order, so the result does not depend on scheduling.  The sum of squared rescaled weights
 * comes out beside it: the label's variance estimate (importance weights far from 1 mean
 * few effective photons).
 */
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "white.h"

#define WHITE_BLOCK 256

extern "C" __global__ void __launch_bounds__(WHITE_BLOCK) mcml_white_kernel(WhiteArgs a)
{
    __shared__ double part[WHITE_BLOCK];
    __shared__ double part_sq[WHITE_BLOCK];
    const int m = blockIdx.x;  /* mua' vector of the set */
    const int g = blockIdx.y;  /* group = set * S + sds - 1 */
    const int set = g / a.num_sds;
    const int L = a.num_layers;
    float dmua[16];
    for (int l = 0; l < L; l++)
        dmua[l] = a.mua_new[((uint64_t)set * a.m + m) * L + l] - a.mua_old[(uint64_t)set * L + l];
    const uint64_t b = a.offsets[g], e = a.offsets[g + 1];
    const int stride = L + 2;
    double acc = 0.0, acc_sq = 0.0;
    for (uint64_t i = b + threadIdx.x; i < e; i += WHITE_BLOCK) {
        const float* row = a.rows + i * (uint64_t)stride;
        float x = 0.0f;
        for (int l = 0; l < L; l++) x = fmaf(dmua[l], row[1 + l], x);
        const double w = (double)row[0] * exp(-(double)x);
        acc += w;
        acc_sq += w * w;
    }
    part[threadIdx.x] = acc;
    part_sq[threadIdx.x] = acc_sq;
    __syncthreads();
    for (int off = WHITE_BLOCK / 2; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) {
            part[threadIdx.x] += part[threadIdx.x + off];
            part_sq[threadIdx.x] += part_sq[threadIdx.x + off];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const uint64_t o = ((uint64_t)set * a.m + m) * a.num_sds + (g - set * a.num_sds);
        a.out[o] = part[0];
        if (a.out_sq) a.out_sq[o] = part_sq[0];
    }
}

hipError_t mcml_launch_white(const WhiteArgs& a, int nsets, hipStream_t st)
{
    hipLaunchKernelGGL(mcml_white_kernel, dim3(a.m, nsets * a.num_sds), dim3(WHITE_BLOCK), 0, st, a);
    return hipGetLastError();
}

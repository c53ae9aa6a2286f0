/* White Monte Carlo rescaling kernel (internal to libmcml). */
#ifndef MCML_WHITE_H
#define MCML_WHITE_H
#include <hip/hip_runtime.h>
#include <stdint.h>

struct WhiteArgs {
    const float* rows;        /* [n * (L+2)] pathlength records, grouped by (set, SDS) */
    const uint64_t* offsets;  /* [nsets*S + 1] group starts */
    const float* mua_old;     /* [nsets * L] absorption the records were simulated with */
    const float* mua_new;     /* [nsets * m * L] absorption to rescale to */
    double* out;              /* [nsets * m * S] detected weight at mua_new, W-units */
    double* out_sq;           /* same shape: sum of squared rescaled weights (or NULL) */
    int32_t m, num_layers, num_sds, pad;
};

hipError_t mcml_launch_white(const WhiteArgs& a, int nsets, hipStream_t st);

#endif

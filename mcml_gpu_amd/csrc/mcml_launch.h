/* Host-side launchers of the kernels in transport.hip (internal to libmcml). */
#ifndef MCML_LAUNCH_H
#define MCML_LAUNCH_H
#include <hip/hip_runtime.h>
#include "mcml_device.h"

/* Cached HIP graph of one first pass (see mcml_graph_first_pass). */
struct FirstPassGraph {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipGraphNode_t reset = nullptr, kernel = nullptr, readback = nullptr;
    int grid = 0, block = 0, nctrl = 0;
    size_t lds = 0;
    unsigned long long* ctrl = nullptr;
    void* host_ctrl = nullptr;
    void* func = nullptr;
};

hipError_t mcml_launch_reset(unsigned long long* ctrl, int n, hipStream_t st);
hipError_t mcml_graph_first_pass(FirstPassGraph& G, const FirstPassArgs& a, int grid, int block,
                                 unsigned long long* ctrl, int nctrl, void* host_ctrl, hipStream_t st);
void mcml_graph_release(FirstPassGraph& G);
hipError_t mcml_launch_first_pass(const FirstPassArgs& a, int grid, int block, hipStream_t st);

/* Allocation-free first pass (sums sink) and its device scratch (limits in mcml.h). */
#define MCML_NOALLOC_OFF_CTRL 0
#define MCML_NOALLOC_OFF_SETS 256
#define MCML_NOALLOC_OFF_BEGIN (MCML_NOALLOC_OFF_SETS + MCML_NOALLOC_MAX_SETS * 2048)
#define MCML_NOALLOC_OFF_ACC (MCML_NOALLOC_OFF_BEGIN + ((MCML_NOALLOC_MAX_SETS + 1) * 8 + 255) / 256 * 256)
#define MCML_NOALLOC_SCRATCH_BYTES \
    (MCML_NOALLOC_OFF_ACC + MCML_NOALLOC_MAX_GROUPS * MCML_ACC_SLOTS * MCML_ACC_WORDS * 8)
hipError_t mcml_scratch_address(void** p);
hipError_t mcml_launch_selftest(unsigned long long* bad, hipStream_t st);
hipError_t mcml_launch_first_pass_sums(const FirstPassArgs& a, int grid, int block, hipStream_t st);
hipError_t mcml_occupancy_first_pass_sums(const FirstPassArgs& a, int block, int* blocks_per_cu);
hipError_t mcml_launch_clear(unsigned long long* p, uint64_t n, hipStream_t st);
hipError_t mcml_launch_sums_finalize(const unsigned long long* accum, int groups, double* reflectance,
                                     uint64_t* detected, hipStream_t st);
/* the replay, then (when a.stage is set) the reduction of its staging copies into the grids */
hipError_t mcml_launch_replay(const ReplayArgs& a, int grid, hipStream_t st);
/* add k staging copies [k][ngroups][mcml_replay_stage_words()] into the grids */
hipError_t mcml_launch_stage_reduce(const unsigned long long* stage, int k, int ngroups, unsigned long long* A_rz,
                                    unsigned long long* A0_z, int nz, int nr, hipStream_t st);
hipError_t mcml_launch_dump(const GatherArgs& a, int grid, hipStream_t st);
hipError_t mcml_launch_groups(const uint64_t* keys, uint64_t n, int nsets, int nsds, uint64_t* offsets,
                              const float* w, double* sums, hipStream_t st);
hipError_t mcml_occupancy_first_pass(const FirstPassArgs& a, int block, int* blocks_per_cu);
hipError_t mcml_sort_keys(void* temp, size_t& temp_bytes, const uint64_t* keys_in, uint64_t* keys_out,
                          uint32_t* idx_out, uint64_t n, hipStream_t st);
hipError_t mcml_occupancy_replay(int* blocks_per_cu);
int mcml_replay_block();
size_t mcml_replay_stage_words(); /* words of one group's tile corner (A_rz corner + A0_z rows) */
/* reference launch shape: bytes of one slot's saved state; init and one step-budget launch */
size_t mcml_refshape_slot_bytes();
hipError_t mcml_launch_refshape_init(const RefShapeArgs& a, int block, hipStream_t st);
hipError_t mcml_launch_refshape_pass(const RefShapeArgs& a, int block, hipStream_t st);

#endif

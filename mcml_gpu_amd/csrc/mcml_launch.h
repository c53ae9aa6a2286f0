/* Host-side launchers of the kernels in transport.hip (internal to libmcml). */
#ifndef MCML_LAUNCH_H
#define MCML_LAUNCH_H
#include <hip/hip_runtime.h>
#include "mcml_device.h"

hipError_t mcml_launch_first_pass(const FirstPassArgs& a, int grid, int block, hipStream_t st);
hipError_t mcml_launch_replay(const ReplayArgs& a, int grid, hipStream_t st);
hipError_t mcml_launch_dump(const DumpArgs& a, int grid, hipStream_t st);
hipError_t mcml_occupancy_first_pass(int block, size_t lds_bytes, int* blocks_per_cu);
hipError_t mcml_occupancy_replay(int* blocks_per_cu);
int mcml_replay_block();

#endif

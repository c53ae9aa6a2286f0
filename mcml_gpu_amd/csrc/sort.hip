/*
 * sort.hip - device radix sort of detection-record keys (rocPRIM), so records never go
 * through a host sort: 64-bit keys (set | SDS | photon) with their source indices.
 */
#include <hip/hip_runtime.h>

#include <cstring> /* rocPRIM headers use memset without including it */

#include <rocprim/rocprim.hpp>

#include "mcml_launch.h"

/* Sort keys ascending, carrying each key's source index (0..n-1) along.
 * temp == nullptr: only report the temporary storage size in temp_bytes. */
hipError_t mcml_sort_keys(void* temp, size_t& temp_bytes, const uint64_t* keys_in, uint64_t* keys_out,
                          uint32_t* idx_out, uint64_t n, hipStream_t st)
{
    rocprim::counting_iterator<uint32_t> idx_in(0u);
    return rocprim::radix_sort_pairs(temp, temp_bytes, keys_in, keys_out, idx_in, idx_out, (size_t)n, 0u, 64u, st);
}

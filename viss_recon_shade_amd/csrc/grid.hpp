// uvw gridding / de-duplication / counts (grid.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace vrs {
size_t grid_scratch_bytes(int64_t n_keys);
int grid_tile();
hipError_t grid_build_keys(const double4 *pos, int n_t, const int2 *pairs, int n_pairs, double sc,
                           uint64_t *keys, int *bad, hipStream_t st);
// Sorts keys[n] (tmp: [n] scratch), writes the unique keys and their first
// positions (iscratch: [2n] ints; small: [n / grid_tile() + 2] ints).
hipError_t grid_sort_unique(uint64_t *keys, uint64_t *tmp, int64_t n, int *iscratch, int *small,
                            uint64_t *ukeys, int32_t *counts, int *n_unique_dev, hipStream_t st);
hipError_t grid_counts(const int *start, int n_unique, int64_t n, int32_t *counts, hipStream_t st);
hipError_t grid_decode(const uint64_t *ukeys, int n_unique, float *uvw, hipStream_t st);
hipError_t grid_vis_weights(const uint64_t *vkeys, int64_t n_vis, const uint64_t *ukeys,
                            const int32_t *counts, int n_unique, float *wgt, int32_t *vcount,
                            hipStream_t st);
}  // namespace vrs

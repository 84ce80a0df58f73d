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
// Stable sort of keys by bits [bit_lo, bit_hi) (tmp: [n]; small: [n / grid_tile() + 2] ints).
hipError_t grid_sort_keys(uint64_t *keys, uint64_t *tmp, int64_t n, int *small, int bit_lo,
                          int bit_hi, hipStream_t st);
// Fine-cell sky: pack cells, sort (parent pixel, cell) keys, emit idx (cells grouped by
// RING parent pixel, ascending cell index within a pixel) and per chunk-order pixel
// start / count.  *bad |= 1: a parent out of range; 2: a pixel without cells;
// 4: a cell direction that is not a unit vector.
hipError_t fine_build_index(const double *lmn, const float *ff, const int32_t *nx,
                            int32_t nx_base, int64_t n, int64_t n_pix, const int32_t *ring,
                            int64_t n_pad, float4 *c32, double4 *c64, uint64_t *keys,
                            uint64_t *tmp, int *small, int *bad, int32_t *idx, int32_t *start,
                            int32_t *count, hipStream_t st);
// out[ring[pi]] = count[pi] for the non-padding pixels
hipError_t fine_count_ring(const int32_t *count, const int32_t *ring, int64_t n_pad, int32_t *out,
                           hipStream_t st);
}  // namespace vrs

// uvw gridding / de-duplication / counts (grid.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace vrs {
hipError_t grid_build_keys(const double4 *pos, int n_t, const int2 *pairs, int n_pairs, double sc,
                           uint64_t *keys, int *bad, hipStream_t st);
// Stable LSD radix sort of keys[n] over the 8-bit digits [d_lo, d_hi) (bits
// 8 d_lo .. 8 d_hi - 1); tmp: [n] keys; scratch: radix_scratch_bytes(n) bytes.
// Digits that every key shares are skipped: one read of the digit histograms
// synchronizes the stream.  passes_run (nullable) = digit passes performed.
size_t radix_scratch_bytes(int64_t n);
hipError_t radix_sort(uint64_t *keys, uint64_t *tmp, int64_t n, int d_lo, int d_hi, void *scratch,
                      hipStream_t st, int *passes_run);
// Runs of sorted keys: sorted_unique_count fills scratch (unique_scratch_bytes(n))
// and returns the number of runs on the host (synchronizes); sorted_unique_runs
// then writes each run's key, first position (start: [n_unique]) and length.
size_t unique_scratch_bytes(int64_t n);
hipError_t sorted_unique_count(const uint64_t *keys, int64_t n, void *scratch, int64_t *n_unique,
                               hipStream_t st);
hipError_t sorted_unique_runs(const uint64_t *keys, int64_t n, const void *scratch, int64_t n_unique,
                              uint64_t *ukeys, int64_t *start, int32_t *counts, hipStream_t st);
hipError_t grid_decode(const uint64_t *ukeys, int64_t n_unique, float *uvw, hipStream_t st);
hipError_t grid_vis_weights(const uint64_t *vkeys, int64_t n_vis, const uint64_t *ukeys,
                            const int32_t *counts, int64_t n_unique, float *wgt, int32_t *vcount,
                            hipStream_t st);
// Fine-cell sky: pack cells, sort (parent pixel, cell) keys (scratch:
// radix_scratch_bytes(n)), emit idx (cells grouped by
// RING parent pixel, ascending cell index within a pixel) and per chunk-order pixel
// start / count.  *bad |= 1: a parent out of range; 2: a pixel without cells;
// 4: a cell direction that is not a unit vector.
hipError_t fine_build_index(const double *lmn, const float *ff, const int32_t *nx,
                            int32_t nx_base, int64_t n, int64_t n_pix, const int32_t *ring,
                            int64_t n_pad, float4 *c32, double4 *c64, uint64_t *keys,
                            uint64_t *tmp, void *scratch, int *bad, int32_t *idx, int32_t *start,
                            int32_t *count, hipStream_t st);
// out[ring[pi]] = count[pi] for the non-padding pixels
hipError_t fine_count_ring(const int32_t *count, const int32_t *ring, int64_t n_pad, int32_t *out,
                           hipStream_t st);
}  // namespace vrs

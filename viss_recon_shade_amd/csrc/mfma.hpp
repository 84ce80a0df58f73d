// Launch interface of the split-FP16 matrix-core path (mfma.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "kernels.hpp"

namespace vrs {

int mfma_vis_pad(int64_t n_vis, int64_t *n_vis_pad);
size_t mfma_a16_bytes_visgen(int32_t n_maps, int64_t n_pad);
size_t mfma_a16_bytes_recon(int32_t n_maps, int64_t n_vis_pad);

// I: [n_maps][n_pix] float (RING); amax/inv_scale: [n_maps]; a16: mfma_a16_bytes_visgen;
// V: [n_maps][n_vis] complex64.
hipError_t launch_mfma_visgen(const GeomDev &g, const PhaseKnobs &k, int32_t n_maps, int64_t n_pix,
                              const float *I, unsigned int *amax, float *inv_scale, void *a16,
                              float2 *V, int grid, hipStream_t st);
// part: [n_maps][n_seg][n_pad] partial images, summed by launch_recon_finalize.
hipError_t launch_mfma_recon(const GeomDev &g, const PhaseKnobs &k, int32_t n_maps,
                             const float2 *V, const float *w, float kappa, int64_t n_vis_pad,
                             unsigned int *amax, float *inv_scale, void *a16, int32_t n_seg,
                             float *part, int grid, hipStream_t st);

// Per-satellite phase factors (orbit geometries whose u v w are satellite
// position differences): steps per vis-gen block for this array, 0 if the
// satellite kernels do not apply (n_sat outside [2, 16]).
int mfma_sat_tb(int32_t n_sat, int32_t n_base);
// 1 if the array's blocks can hold star-ordered baselines (16 satellites, every
// pair; the pair list must be the lexicographic one): each wave's visibilities
// then share one satellite, and a wave skips the K-steps on which that
// satellite is occulted at both steps of the block.
int mfma_visgen_sat_star(int32_t n_sat, int32_t n_base);
hipError_t launch_mfma_visgen_sat(const GeomDev &g, const PhaseKnobs &k, double inv_lambda_km,
                                  int32_t n_maps, int64_t n_pix, const float *I,
                                  unsigned int *amax, float *inv_scale, void *a16, float2 *V,
                                  bool star, int grid, hipStream_t st);
const void *mfma_visgen_sat_kernel_ptr(int tb);

// Recon on per-satellite factors: the satellite counts it is built for (16, 8;
// 0 otherwise; the pair list must be the lexicographic one), the operand's slot
// count, and the launch (spos: n_sat * n_t double4 of workspace).
int mfma_recon_sat_ns(int32_t n_sat);
int64_t mfma_recon_sat_slots(int32_t n_sat, int32_t n_t);
hipError_t launch_mfma_recon_sat(const GeomDev &g, const PhaseKnobs &k, double inv_lambda_km,
                                 int32_t n_maps, const float2 *V, const float *w, float kappa,
                                 int64_t n_slots, unsigned int *amax, float *inv_scale, void *a16,
                                 void *spos, int32_t n_seg, float *part, int grid, hipStream_t st);
const void *mfma_recon_sat_kernel_ptr(int ns);
// One-map vis-gen on per-satellite factors (at most 16 satellites, any pair
// list): pixel segments, the Gram-block workspace, and the launch (rows:
// quad_recon_rows_bytes; amax: one word).
int64_t gram_visgen_segments(int64_t n_pad, int32_t n_t);
size_t gram_visgen_part_bytes(int32_t n_t, int64_t n_seg);
hipError_t launch_gram_visgen(const GeomDev &g, const PhaseKnobs &k, double inv_lambda_km, int64_t n_pix,
                              const float *I, unsigned int *amax, void *rows, int32_t n_seg, float *part,
                              float2 *V, hipStream_t st);
// One-map recon on per-satellite factors (at most 16 satellites, any pair list):
// step segments, workspace sizes (W operand, satellite rows) and the launch;
// part: [n_seg][n_pad] partial images, summed by launch_recon_finalize.
int64_t quad_recon_segments(int64_t n_pad, int32_t n_t);
size_t quad_recon_wop_bytes(int32_t n_t);
size_t quad_recon_rows_bytes(int32_t n_t);
hipError_t launch_quad_recon(const GeomDev &g, const PhaseKnobs &k, double inv_lambda_km,
                             const float2 *V, const float *w, float kappa, unsigned int *amax,
                             float *inv_scale, void *wop, void *rows, int32_t n_seg, float *part,
                             hipStream_t st);
int mfma_recon_sat_threads();
int mfma_recon_sat_pix_per_block();

const void *mfma_visgen_kernel_ptr();
const void *mfma_recon_kernel_ptr();
int mfma_block_threads();         // vis-gen block
int mfma_recon_block_threads();   // recon block
int mfma_vis_per_block();
int mfma_pix_per_block();

}  // namespace vrs

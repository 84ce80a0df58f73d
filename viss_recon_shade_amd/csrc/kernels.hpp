// Launch interface of the gfx950 kernels (kernels.hip).  Host code in capi.hip
// validates every argument before calling these.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "vrs_internal.hpp"

namespace vrs {

// A fine-cell sky attached to a geometry (sub-pixel shading): cells grouped by
// their parent pixel, pixel (chunk-order) -> [start, start + count) in idx.
struct FineDev {
    const int32_t *start;   // [n_pad]
    const int32_t *count;   // [n_pad] (0 for padding and empty pixels)
    const int32_t *idx;     // [n_cells] cell indices, grouped by pixel, ascending within a pixel
    const float4 *cell32;   // [n_cells] (l, m, n, FF)
    const double4 *cell64;  // [n_cells] (l, m, n, 0)
    const float2 *pixcap;   // [n_pad] (cos, sin) of the angle that holds the pixel's cells
};

// Device view of one geometry; passed to kernels by value.
struct GeomDev {
    int32_t n_sat, n_t, n_base, n_chunk;
    // > 0: antipodal layout, chunk k + n_half holds the antipodes of chunk k
    // slot by slot (n_chunk = 2 n_half); 0: no such pairing
    int32_t n_half;
    int64_t n_vis, n_pad;
    const double4 *pos;    // [n_sat * n_t]  (x, y, z, |r|^2)
    const CapRec *cap;     // [n_sat * n_t]
    const double4 *uvw64;  // [n_vis]
    const int2 *vis_cap;   // [n_vis] cap rows (sat_a * n_t + t, sat_b * n_t + t)
    const int2 *pairs;     // [n_base]
    const float4 *pix32;   // [n_pad]
    const double4 *pix64;  // [n_pad]
    const int32_t *ring;   // [n_pad]
    const ChunkRec *chunk; // [n_chunk]
    const ChunkRec *sub;   // [n_pad / kSubChunk] bounding caps of 64-pixel groups (8 x 8 blocks)
    LimbConsts limb;
    // Ragged batch ("list mode"); all null for a single geometry.
    const int2 *vis_sat;       // [n_vis] (sat_a * n_t + t, sat_b * n_t + t); x < 0: padding
    const int32_t *tile_map;   // [n_vis / list_pad()] map (channel) of each padded group
    const double *inv_lambda;  // [n_maps] 1e3 / lambda_m: km -> wavelengths
    const int4 *slots;         // [n_slots] (map, v_begin, v_end, 0) recon visibility ranges
    const int32_t *map_slot0;  // [n_maps + 1] first slot of each map
    const float *kappa_map;    // [n_maps] recon normalisation of each map
    int64_t n_slots;
    FineDev fine;              // all null unless a fine-cell vis-gen is running
};

struct OrbitArgs {
    int32_t n_sat, n_t;
    double dt, a, cos_i, sin_i, n_mot, r_moon, inv_lambda_km;
    const double *raan, *m0;  // device [n_sat]
    // device [n_t] or null: step j is taken at time t_idx[j] * dt; a negative
    // entry is a zero-filled sample (position 0, a cap that shades nothing)
    const int64_t *t_idx;
    int32_t ceil_uvw;         // 1: uvw = ceil(dx * 1e3 / lambda), the reference's gridding
};

struct PhaseKnobs {
    double s;       // phase scale
    int32_t wfac;   // 1: use n (trailing exp(-2 pi i s w) folded in), 0: n - 1
    float t_occ;
    int32_t pair;   // 1: antipodal pixel pairs share a phase (host-side kernel choice)
};

// Stage 1: positions, caps, uvw.
hipError_t launch_orbit(const OrbitArgs &a, const GeomDev &g, double4 *pos, CapRec *cap,
                        double4 *uvw64, int2 *vis_cap, hipStream_t st);

// Stage 2: per-visibility shade statistics.  bits may be null; tot = [shaded, near, fp64].
hipError_t launch_shade(const GeomDev &g, int32_t *count, uint32_t *hash, uint32_t *bits,
                        int64_t n_words, unsigned long long *tot, int grid, hipStream_t st);

// Gather RING-ordered maps into chunk order (zero padding).
hipError_t launch_to_chunk_order(const GeomDev &g, int32_t n_maps, int64_t n_pix,
                                 const float *I, float *Ich, hipStream_t st);

// Stage 3.  Ich: [n_maps][n_pad]; V: [n_maps][n_vis] complex64.
// vt = visibilities per tile (8, 16 or 32); pair: one phase serves a pixel and
// its antipode (used when g.n_half > 0).
hipError_t launch_visgen(const GeomDev &g, const PhaseKnobs &k, bool f64_phase, int vt, bool pair,
                         int32_t n_maps, const float *Ich, float2 *V, int grid, hipStream_t st);

// Stage 4.  V: [n_maps][n_vis] complex64; w: [n_vis] or null; part: [n_maps][n_seg][n_pad].
// queue: one device counter, cleared by the launch (a kernel, no memset node).
hipError_t launch_recon(const GeomDev &g, const PhaseKnobs &k, bool f64_phase, int32_t n_maps,
                        const float2 *V, const float *w, float kappa, int32_t n_seg,
                        float *part, unsigned long long *queue, int grid, hipStream_t st);
// F[map][ring[p]] = sum_seg part[map][seg][p].
hipError_t launch_recon_finalize(const GeomDev &g, int32_t n_maps, int32_t n_seg, int64_t n_pix,
                                 const float *part, float *F, hipStream_t st);

// Fine-cell sky set-up: per chunk-order pixel, cells' mean FF (cbar) and the
// angle holding its cells (pixcap); per chunk, bounds widened to hold every
// cell (fchunk).  g.fine.idx/start/count/cell* come from the caller.
hipError_t launch_fine_pixels(const GeomDev &g, float *cbar, float2 *pixcap, ChunkRec *fchunk,
                              hipStream_t st);

// One map, chunk order -> RING order (padding dropped).
hipError_t launch_from_chunk_order(const GeomDev &g, int64_t n_pix, const float *in, float *out,
                                   hipStream_t st);

// Vis-gen with the sub-pixel (fine-cell) shade mask; Ich = per-pixel mean FF (chunk order);
// g.chunk must be the widened fine chunk bounds.
// Bare-term ceiling kernel: `blocks` x 256 threads, each writes out[thread] once.
hipError_t launch_term_ceiling(const float4 *uvw, int32_t iters, int32_t blocks, float *out,
                               hipStream_t st);
int64_t term_ceiling_terms(int32_t iters, int32_t blocks);
hipError_t launch_visgen_fine(const GeomDev &g, const PhaseKnobs &k, bool f64, const float *Ich,
                              float2 *V, int grid, hipStream_t st);
const void *visgen_fine_kernel_ptr(bool f64);

int visgen_tile_vis();           // default visibilities per vis-gen tile
int list_pad();                  // list-mode padding granule (entries)
int recon_stage_vis();           // visibilities staged per recon LDS tile
const void *visgen_kernel_ptr(bool f64_phase, int vt, bool pair);
const void *recon_kernel_ptr(bool f64_phase);
const void *shade_kernel_ptr();

}  // namespace vrs

// Internal layouts shared by the host side and the gfx950 kernels.
//
// HBM layout of one geometry (all arrays owned by vrs_geom):
//   pos    double4 [n_sat][n_t]      satellite position (km) and |r|^2, FP64
//   cap    CapRec  [n_sat][n_t]      shade cap of that satellite at that step (FP32)
//   uvw64  double4 [n_vis]           (u, v, w, 0) wavelengths, v = b * n_t + t
//   vis_cap int2   [n_vis]           cap rows of the pair, (sat_a * n_t + t, sat_b * n_t + t)
//   pairs  int2    [n_base]          satellite indices of each baseline
//   pix32  float4  [n_pad]           (l, m, n, 0) in chunk order, FP32
//   pix64  double4 [n_pad]           (l, m, n, 0) in chunk order, FP64
//   ring   int32   [n_pad]           RING index of each chunk-ordered pixel, -1 for padding
//   chunk  ChunkRec[n_chunk]         bounding cap of each chunk of kChunk pixels
//   sub    ChunkRec[n_pad / 64]      the same for each group of kSubChunk pixels
// Chunk order is HEALPix NESTED order, so a chunk of 256 pixels is a 16 x 16
// patch of one base face whenever nside >= 16.  From nside 16 on the second
// half is the antipodal image of the first (position n/2 + q holds the
// antipode of position q; base faces 0-5 face faces 6-11), so chunk
// k + n_chunk/2 is chunk k mirrored through the centre, slot by slot.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace vrs {

constexpr int kWave = 64;
constexpr int kPixPerLane = 4;
constexpr int kChunk = kWave * kPixPerLane;  // pixels per chunk
constexpr int kWavesPerBlock = 4;
constexpr int kBlock = kWave * kWavesPerBlock;
constexpr int kSubChunk = 64;                // pixels of a sub-chunk (one wave's lane group)

// Shade cap of one satellite at one step: pixel s is shaded iff
// dot(s, dir) >= c where dir = -r/|r| and c = sqrt(1 - R^2/|r|^2).
// sa = sin(alpha) = R/|r|.  g = half-width of the FP32 guard band; lo = c - g
// and hi = c + g in FP32: the FP32 pixel test shades when dot >= hi, clears
// when dot < lo, and leaves pixels with lo <= dot < hi to the FP64 predicate.
struct alignas(32) CapRec {
    float dx, dy, dz, c;
    float sa, g, lo, hi;
};

// Bounding cap of a chunk: every pixel centre lies within angle rho of (gx,gy,gz).
struct alignas(32) ChunkRec {
    float gx, gy, gz, crho;
    float srho;
    int32_t flags;      // kChunkAlwaysStraddle: classify pixel by pixel
    int32_t n_valid;    // pixels that are not padding
    uint32_t hash_all;  // XOR of pix_hash over the chunk's RING indices
};
constexpr int32_t kChunkAlwaysStraddle = 1;

enum ChunkClass : int { kOut = 0, kIn = 1, kStraddle = 2 };

// Margin (in cosine units) added to the cap guard band when a whole chunk is
// classified; covers FP32 rounding of the bounding-cap arithmetic.
constexpr float kChunkMargin = 3e-5f;
// Minimum FP32 guard band; the FP32 dot of two unit vectors rounded to FP32
// is within 4e-7 of the exact value.
constexpr float kGuardMin = 4e-6f;
// Near-limb band: limb distance within 1e-6 R.
constexpr double kNearLimbRel = 1e-6;

struct LimbConsts {
    double R2, lo2, hi2;
};

__host__ __device__ inline uint32_t pix_hash_u64(uint64_t p) {
    uint64_t z = p + 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    z ^= z >> 31;
    return (uint32_t)(z ^ (z >> 32));
}

}  // namespace vrs

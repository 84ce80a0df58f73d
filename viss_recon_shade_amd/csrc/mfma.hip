// Multi-map vis-gen and recon as real contractions on the gfx950 matrix cores.
//
// With n_maps sky maps the direct DFT is a GEMM whose middle operand, the
// phase-factor matrix E[v][p] = exp(-2 pi i s phi_vp) with the shade mask
// folded in, is generated on the fly:
//   vis-gen  V[m][v] = sum_p I[m][p] E'[v][p] + t_occ sum_{p shaded} E[v][p]
//            (E' = E with shaded entries zeroed; the second term is
//             map-independent and accumulated on the VALU side)
//   recon    F[m][p] = sum_v (Vr w kappa)[m][v] Re E* - (Vi w kappa)[m][v] Im E*
//            with shaded (v, p) entries zeroed.
// Each lane generates the E entries its MFMA B operand needs (FP64 phase and
// fract, FP32 sin/cos), splits them into FP16 hi + lo and issues three
// products per tile (hi*hi + hi*lo + lo*hi) into FP32 accumulators.  The map
// operand (I, or the weighted V) is pre-split the same way, scaled per map by
// a power of two into FP16 range, and staged through LDS for all waves of the
// block.  Instructions: vis-gen v_mfma_f32_16x16x32_f16 (M = 16 maps, N = 16
// visibilities, K = 32 pixels), recon v_mfma_f32_32x32x16_f16 (M = 32 maps,
// N = 32 pixels, K = 16 visibilities).
//
// On orbit geometries the second family of kernels (the "per-satellite
// factor" kernels further down) forms E from factors of single satellites:
// u v w = (r_y - r_x) / lambda, so E_xy(p) = E_y(p) conj(E_x(p)) with E_s(p) =
// exp(-2 pi i s (r_s / lambda) . (l, m, n - 1)), one FP64 phase and one sin/cos
// per (satellite, step, pixel) instead of per (baseline, step, pixel), and the
// occultation of a (visibility, pixel) pair is that of either satellite, so
// zeroing each satellite's occulted factors zeroes exactly the shaded entries.
#include <hip/hip_runtime.h>

#include "kernels.hpp"
#include "mfma.hpp"

namespace vrs {

namespace {

typedef _Float16 half8 __attribute__((ext_vector_type(8)));
typedef float floatx16 __attribute__((ext_vector_type(16)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

constexpr int kMT = 32;            // rows of one M tile (maps)
constexpr int kMaps = 64;          // maps per pass
constexpr int kNT = 32;            // columns per wave (visibilities or pixels)
constexpr int kKS = 16;            // K per MFMA
constexpr int kStage = 64;         // K per LDS stage
// vis-gen: v_mfma_f32_16x16x32_f16, 16 maps x 16 visibilities x 32 pixels per
// instruction; 16 waves x 16 visibilities per block, four waves per SIMD.
constexpr int kMTV = 16;          // rows of one vis-gen M tile (maps)
constexpr int kNTV = 16;          // visibilities per vis-gen wave
constexpr int kKSV = 32;          // K (pixels) per vis-gen MFMA
constexpr int kWaves = 16;        // vis-gen block: one per CU
constexpr int kThreads = kWaves * 64;
constexpr int kWavesR = 16;       // recon block: two chunks, one block per CU, four waves per SIMD
constexpr int kThreadsR = kWavesR * 64;

static_assert(kChunk % kStage == 0, "stages must tile chunks");
static_assert(kStage == kSubChunk, "vis-gen stages are classified by their sub-chunk record");
static_assert(kWavesR * kNT == 2 * kChunk, "a recon block covers two chunks");

__device__ __forceinline__ floatx16 mfma(half8 a, half8 b, floatx16 c) {
    return __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ floatx4 mfma16(half8 a, half8 b, floatx4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
}

// x = hi + lo with hi, lo in FP16 (hi rounded toward zero, lo the remainder).
__device__ __forceinline__ void split2(float x0, float x1, uint32_t &hi, uint32_t &lo) {
    const auto h = __builtin_amdgcn_cvt_pkrtz(x0, x1);  // 2 x f16
    const float r0 = x0 - (float)h[0];
    const float r1 = x1 - (float)h[1];
    const auto l = __builtin_amdgcn_cvt_pkrtz(r0, r1);
    hi = __builtin_bit_cast(uint32_t, h);
    lo = __builtin_bit_cast(uint32_t, l);
}

union H8 {
    half8 v;
    uint32_t u[4];
};

__device__ __forceinline__ bool occ32m(float sx, float sy, float sz, const CapRec &a, bool &unc) {
    const float x = fmaf(sz, a.dz, fmaf(sy, a.dy, sx * a.dx));
    unc = unc || (fabsf(x - a.c) <= a.g);
    return x >= a.c;
}

__device__ __forceinline__ bool occ64m(const double4 r, double sx, double sy, double sz,
                                       const LimbConsts &L) {
    const double d = fma(r.z, sz, fma(r.y, sy, r.x * sx));
    if (!(d < 0.0)) return false;
    const double q = fma(-d, d, r.w);
    return q <= L.R2;
}

__device__ __forceinline__ int classify_cap_m(const ChunkRec &ch, const CapRec &a) {
    if (ch.flags & kChunkAlwaysStraddle) return kStraddle;
    const float x = fmaf(ch.gz, a.dz, fmaf(ch.gy, a.dy, ch.gx * a.dx));
    const float mg = a.g + kChunkMargin;
    if ((x >= fmaf(a.c, ch.crho, a.sa * ch.srho) + mg) && (a.c < ch.crho)) return kIn;
    if (x < fmaf(a.c, ch.crho, -a.sa * ch.srho) - mg) return kOut;
    return kStraddle;
}

__device__ __forceinline__ int classify_m(const ChunkRec &ch, const CapRec &a, const CapRec &b) {
    const int ca = classify_cap_m(ch, a), cb = classify_cap_m(ch, b);
    return (ca == kIn || cb == kIn) ? kIn : ((ca == kOut && cb == kOut) ? kOut : kStraddle);
}

// Shade decision for one (visibility, pixel) pair inside a straddling chunk.
__device__ __forceinline__ bool pair_shaded(const GeomDev &g, const double4 s64,
                                            const CapRec &A, const CapRec &B, int ia, int ib) {
    const float sx = (float)s64.x, sy = (float)s64.y, sz = (float)s64.z;
    bool unc = false;
    bool sh = occ32m(sx, sy, sz, A, unc) | occ32m(sx, sy, sz, B, unc);
    if (unc)
        sh = occ64m(g.pos[ia], s64.x, s64.y, s64.z, g.limb) |
             occ64m(g.pos[ib], s64.x, s64.y, s64.z, g.limb);
    return sh;
}

// Shade bits of 8 staged pixels sp[S t], t = 0..7, for one pair (bit t: pixel S t):
// FP32 cap tests, the exact FP64 predicate for pixels in the guard band.
template <int S>
__device__ __forceinline__ uint32_t straddle_mask8(const GeomDev &g, const double4 *sp,
                                                const CapRec &A, const CapRec &B, int ia, int ib) {
    uint32_t m = 0, um = 0;
#pragma unroll
    for (int t = 0; t < 8; ++t) {
        const double4 s = sp[S * t];
        const float sx = (float)s.x, sy = (float)s.y, sz = (float)s.z;
        bool unc = false;
        const bool sh = occ32m(sx, sy, sz, A, unc) | occ32m(sx, sy, sz, B, unc);
        m |= (uint32_t)sh << t;
        um |= (uint32_t)unc << t;
    }
    while (um) {
        const int t = __ffs(um) - 1;
        um &= um - 1;
        const double4 s = sp[S * t];
        const bool ex = occ64m(g.pos[ia], s.x, s.y, s.z, g.limb) |
                        occ64m(g.pos[ib], s.x, s.y, s.z, g.limb);
        m = (m & ~(1u << t)) | ((uint32_t)ex << t);
    }
    return m;
}

// ------------------------------------------------------------------ prep
__global__ void zero_u32_kernel(unsigned int *__restrict__ x, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] = 0u;
}

// Per-row scale exponents: e[m] such that max |x[m][:] - shift| * 2^e[m] <= 2^14
// (shift applies to real rows only).
__global__ void __launch_bounds__(256) row_absmax_kernel(const float *__restrict__ x, int64_t n,
                                                         int64_t stride_row, int complex_in,
                                                         const float *__restrict__ w,
                                                         unsigned int *__restrict__ amax,
                                                         float shift) {
    const int64_t row = blockIdx.y;
    const float *xr = x + row * stride_row;
    float m = 0.0f;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const float ww = w ? fabsf(w[i]) : 1.0f;
        if (complex_in) {
            m = fmaxf(m, fabsf(xr[2 * i]) * ww);
            m = fmaxf(m, fabsf(xr[2 * i + 1]) * ww);
        } else {
            m = fmaxf(m, fabsf(xr[i] - shift) * ww);
        }
    }
    for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o, 64));
    if ((threadIdx.x & 63) == 0) atomicMax(&amax[row], __float_as_uint(m));
}

__device__ __forceinline__ float row_scale(unsigned int amax_bits, float extra) {
    const float m = __uint_as_float(amax_bits) * extra;
    if (!(m > 0.0f) || !isfinite(m)) return 1.0f;
    int e;
    frexpf(m, &e);              // m = f 2^e, f in [0.5, 1)
    return ldexpf(1.0f, 14 - e);  // m * scale < 2^14
}

// K order of the vis-gen map operand inside each K-step of 32 pixels: position
// 8 g + t holds pixel 4 t + g.  Lane group g of a wave (lanes 16 g .. 16 g + 15)
// supplies K positions 8 g .. 8 g + 7, so its E entry t is pixel 4 t + g: the
// four groups read four consecutive 32-byte pixel slots, which lie on distinct
// LDS banks (8 pixels apart they would share the same banks).
__device__ __forceinline__ int64_t kstep_pos(int64_t p) {
    return (p & ~int64_t(31)) | ((p & 3) << 3) | ((p >> 2) & 7);
}

// I - shift [n_maps][n_pix] RING -> hi/lo FP16 [group][2][kMaps][n_pad] chunk
// order (permuted by kstep_pos inside each K-step when `permute`), scaled;
// padding pixels are zero.
__global__ void __launch_bounds__(256) split_maps_kernel(GeomDev g, int32_t n_maps, int64_t n_pix,
                                                         const float *__restrict__ I,
                                                         const unsigned int *__restrict__ amax,
                                                         _Float16 *__restrict__ out,
                                                         float *__restrict__ inv_scale,
                                                         int32_t permute, float shift) {
    const int64_t n_groups = (n_maps + kMaps - 1) / kMaps;
    const int64_t total = n_groups * kMaps * g.n_pad;
    for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
         idx += (int64_t)gridDim.x * blockDim.x) {
        const int64_t pi = idx % g.n_pad;
        const int64_t mrow = idx / g.n_pad;  // group * kMaps + row
        const int64_t grp = mrow / kMaps, row = mrow % kMaps;
        const int64_t m = grp * kMaps + row;
        float x = 0.0f, sc = 1.0f;
        if (m < n_maps) {
            sc = row_scale(amax[m], 1.0f);
            const int r = g.ring[pi];
            if (r >= 0) x = (I[m * n_pix + r] - shift) * sc;
            if (pi == 0) inv_scale[m] = 1.0f / sc;
        }
        const _Float16 h = (_Float16)x;  // nearest
        const _Float16 l = (_Float16)(x - (float)h);
        const int64_t kp = permute ? kstep_pos(pi) : pi;
        out[((grp * 2 + 0) * kMaps + row) * g.n_pad + kp] = h;
        out[((grp * 2 + 1) * kMaps + row) * g.n_pad + kp] = l;
    }
}

// V [n_maps][n_vis] complex -> hi/lo FP16 [group][re,-im][hi,lo][kMaps][n_vis_pad],
// weighted by w * kappa and scaled per map.  The imaginary rows hold -Im V so
// that Re(V e^{+i theta}) = Re V cos + (-Im V) sin is a plain sum of products.
__global__ void __launch_bounds__(256) split_vis_kernel(int64_t n_vis, int64_t n_vis_pad,
                                                        int32_t n_maps,
                                                        const float2 *__restrict__ V,
                                                        const float *__restrict__ w, float kappa,
                                                        const unsigned int *__restrict__ amax,
                                                        _Float16 *__restrict__ out,
                                                        float *__restrict__ inv_scale) {
    const int64_t n_groups = (n_maps + kMaps - 1) / kMaps;
    const int64_t total = n_groups * kMaps * n_vis_pad;
    for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
         idx += (int64_t)gridDim.x * blockDim.x) {
        const int64_t v = idx % n_vis_pad;
        const int64_t mrow = idx / n_vis_pad;
        const int64_t grp = mrow / kMaps, row = mrow % kMaps;
        const int64_t m = grp * kMaps + row;
        float xr = 0.0f, xi = 0.0f;
        if (m < n_maps) {
            const float sc = row_scale(amax[m], fabsf(kappa));
            if (v < n_vis) {
                const float2 q = V[m * n_vis + v];
                const float ww = (w ? w[v] : 1.0f) * kappa * sc;
                xr = q.x * ww;
                xi = -(q.y * ww);
            }
            if (v == 0) inv_scale[m] = 1.0f / sc;
        }
        const _Float16 hr = (_Float16)xr, hi_ = (_Float16)xi;
        const _Float16 lr = (_Float16)(xr - (float)hr), li = (_Float16)(xi - (float)hi_);
        const int64_t base = grp * 4 * kMaps;
        out[((base + 0 * kMaps) + row) * n_vis_pad + v] = hr;  // re hi
        out[((base + 1 * kMaps) + row) * n_vis_pad + v] = lr;  // re lo
        out[((base + 2 * kMaps) + row) * n_vis_pad + v] = hi_; // -im hi
        out[((base + 3 * kMaps) + row) * n_vis_pad + v] = li;  // -im lo
    }
}

// Output row of accumulator element r for this lane in a 32x32 tile.
__device__ __forceinline__ int acc_row(int r, int lane) {
    return 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
}

// ------------------------------------------------------------------ vis-gen
// Block: 16 waves x 16 visibilities; all waves share the staged map tile (each
// staged tile serves 256 visibilities).  The 16x16x32 instruction keeps the
// complex accumulators of 64 maps in 32 registers, so four waves fit a SIMD.
// Stages of kStage pixels are double-buffered: the next stage's map tile and
// pixel vectors go straight from global memory into the other LDS buffer
// while the current stage computes (one barrier per stage).
// LDS tiles of the map operand are [NH][kMaps rows][8 segments of 8 halves],
// unpadded (128-byte rows).  Segment s of row r sits in slot s ^ ((r >> 1) & 7):
// with 64 four-byte banks, even and odd rows start on opposite bank halves and
// the slot swizzle spreads the 8 rows of each half over its 8 16-byte groups,
// so the 16 rows read per LDS cycle by an A-operand load hit 64 distinct banks.
__device__ __forceinline__ int a_slot(int row, int seg) { return seg ^ ((row >> 1) & 7); }

typedef __attribute__((address_space(3))) void *lds_ptr_t;

// Direct global -> LDS load of stage st (gfx950 global_load_lds_dwordx4):
// every wave moves whole 1 KiB blocks, lane l writing bytes [16 l, 16 l + 16)
// of its block.  NH halves-planes of the map operand, plus kStage FP64 pixel
// vectors when pix64 is given.  Completion: s_waitcnt vmcnt(0) + barrier.
template <int NH, int NW>
__device__ __forceinline__ void stage_load_lds(const _Float16 *__restrict__ Ag, int64_t ld,
                                               const double4 *__restrict__ pix64, int64_t st,
                                               _Float16 *sA, double4 *sP) {
    constexpr int kBlocks = NH * kMaps * (kStage / 8) / 64;   // 1 KiB blocks of the map tile
    static_assert(kBlocks % NW == 0, "map tile splits evenly over the waves");
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int64_t p0 = st * kStage;
#pragma unroll
    for (int r = 0; r < kBlocks / NW; ++r) {
        const int blk = wave * (kBlocks / NW) + r;
        const int i = blk * 64 + lane;                       // 16-byte slot index in the tile
        const int hl = i / (kMaps * (kStage / 8));
        const int rem = i % (kMaps * (kStage / 8));
        const int row = rem / (kStage / 8), slot = rem % (kStage / 8);
        const int seg = a_slot(row, slot);                   // the swizzle is an involution
        const _Float16 *src = Ag + ((int64_t)hl * kMaps + row) * ld + p0 + seg * 8;
        __builtin_amdgcn_global_load_lds(src, (lds_ptr_t)(sA + blk * 512), 16, 0, 0);
    }
    if (pix64) {
        // kStage double4 = kStage / 32 blocks of 1 KiB, one per wave
        static_assert(kStage / 32 <= NW, "pixel blocks fit the waves");
        if (wave < kStage / 32) {
            const char *src = reinterpret_cast<const char *>(pix64 + p0) + wave * 1024 + lane * 16;
            __builtin_amdgcn_global_load_lds(src, (lds_ptr_t)(reinterpret_cast<char *>(sP) + wave * 1024),
                                             16, 0, 0);
        }
    }
}

__device__ __forceinline__ void wait_vm() { asm volatile("s_waitcnt vmcnt(0)" ::: "memory"); }

__global__ void __launch_bounds__(kThreads) mfma_visgen_kernel(
    GeomDev g, PhaseKnobs k, int32_t n_maps, const _Float16 *__restrict__ A16,
    const float *__restrict__ inv_scale, float2 *__restrict__ V) {
    __shared__ _Float16 sA[2][2][kMaps][kStage];  // [buffer][hi/lo][map][pixel slot]
    __shared__ double4 sP[2][kStage];             // [buffer] pixel (l, m, n) FP64
    __shared__ CapRec sCap[kWaves * kNTV][2];     // caps of the block's visibilities
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // wave-uniform
    const int col = lane & 15;                 // this lane's visibility within the wave tile
    const int kg = lane >> 4;                  // which 8 of the 32 K values this lane supplies
    const int64_t n_vtiles = (g.n_vis + kWaves * kNTV - 1) / (kWaves * kNTV);
    const int64_t n_groups = (n_maps + kMaps - 1) / kMaps;
    const int64_t n_tiles = n_vtiles * n_groups;
    const int64_t n_stages = g.n_pad / kStage;
    const float t_occ = k.t_occ;

    for (int64_t tile = blockIdx.x; tile < n_tiles; tile += gridDim.x) {
        const int64_t grp = tile / n_vtiles;
        const int64_t v = (tile % n_vtiles) * (kWaves * kNTV) + wave * kNTV + col;
        const bool vok = v < g.n_vis;
        // this lane's visibility
        double u = 0.0, vv = 0.0, w = 0.0;
        int ia = 0, ib = 0;
        const int vc = wave * kNTV + col;       // this lane's visibility within the block
        if (vok) {
            const int2 vcap = g.vis_cap[v];
            ia = vcap.x;
            ib = vcap.y;
            const double4 q = g.uvw64[v];
            u = q.x * k.s; vv = q.y * k.s; w = q.z * k.s;
        }
        // the caps are read once per chunk and in straddling K-steps: keep them in LDS
        // (sCap is free: every wave passed the previous tile's last barrier)
        if (kg == 0) {
            const CapRec none = {0.f, 0.f, 0.f, 2.0f, 0.f, 0.f, 2.0f, 2.0f};
            sCap[vc][0] = vok ? g.cap[ia] : none;
            sCap[vc][1] = vok ? g.cap[ib] : none;
        }
        const double wo = k.wfac ? 0.0 : -w;
        floatx4 acc[kMaps / kMTV][2];  // [m tile][re, im]
#pragma unroll
        for (int a = 0; a < kMaps / kMTV; ++a)
#pragma unroll
            for (int c = 0; c < 2; ++c) acc[a][c] = floatx4{};
        float gre = 0.0f, gim = 0.0f;  // sum of (cos, sin) over shaded pixels
        int cls = kOut, ccls = kOut;  // class of this lane's pair: stage, chunk
        bool cmixed = false;  // the wave's chunk classes mix or some lane straddles
        bool mixed = false;   // some lane's pair shades part of the stage, or lanes differ
        bool all_in = false;  // every visibility of the wave shades the whole chunk: E' = 0

        const _Float16 *Ag = A16 + grp * 2 * kMaps * g.n_pad;
        // Both buffers are free here: every wave passed the last barrier of
        // the previous tile after its final reads.
        stage_load_lds<2, kWaves>(Ag, g.n_pad, g.pix64, 0, &sA[0][0][0][0], &sP[0][0]);
        wait_vm();
        __syncthreads();
        for (int64_t st = 0; st < n_stages; ++st) {
            const int buf = (int)(st & 1);
            const int64_t p0 = st * kStage;
            // Classes per chunk; where the wave's chunk is mixed, lanes whose chunk
            // straddles re-classify each stage against the 64-pixel group's own
            // bounding cap (records first, so that waiting for them does not wait
            // for the prefetch).
            if ((p0 % kChunk) == 0) {
                const ChunkRec ch = g.chunk[p0 / kChunk];
                ccls = classify_m(ch, sCap[vc][0], sCap[vc][1]);
                cmixed = __ballot(ccls != kOut) != 0 && __ballot(ccls == kIn) != ~0ull;
                cls = ccls;
                all_in = __ballot(ccls == kIn) == ~0ull;
                mixed = cmixed;
            }
            if (cmixed) {  // wave-uniform
                cls = ccls;
                if (ccls == kStraddle) cls = classify_m(g.sub[p0 / kSubChunk], sCap[vc][0], sCap[vc][1]);
                all_in = __ballot(cls == kIn) == ~0ull;
                mixed = !all_in && __ballot(cls != kOut) != 0;
            }
            // the other buffer was last read before the previous barrier
            if (st + 1 < n_stages)
                stage_load_lds<2, kWaves>(Ag, g.n_pad, g.pix64, st + 1, &sA[buf ^ 1][0][0][0],
                                          &sP[buf ^ 1][0]);
            for (int ks = 0; ks < kStage / kKSV; ++ks) {
                // E for this lane's 8 pixels: straight-line phase and sin/cos
                // first, then the shade mask, so that the 8 chains interleave.
                // entry t of this lane is staged pixel ks * 32 + 4 t + kg (see kstep_pos)
                const int kb = ks * kKSV + kg;
                float cs[8], sn[8];
#pragma unroll
                for (int t = 0; t < 8; ++t) {
                    const double4 sp = sP[buf][kb + 4 * t];
                    const double ph = fma(u, sp.x, fma(vv, sp.y, fma(w, sp.z, wo)));
                    const float r = (float)__builtin_amdgcn_fract(ph);
                    cs[t] = __builtin_amdgcn_cosf(r);
                    sn[t] = __builtin_amdgcn_sinf(r);
                }
                // The B operand holds (cos, sin) of E = cos - i sin; the sign of the
                // imaginary part is applied once, at the output.
                H8 bre_h, bre_l, bim_h, bim_l;
                if (mixed) {  // wave-uniform: shaded and clear pixels both occur
                    uint32_t shm = (cls == kIn) ? 0xffu : 0u;
                    if (cls == kStraddle)
                        shm = straddle_mask8<4>(g, &sP[buf][kb], sCap[vc][0], sCap[vc][1], ia, ib);
#pragma unroll
                    for (int t = 0; t < 8; ++t) {
                        const bool sh = (shm >> t) & 1u;
                        gre += sh ? cs[t] : 0.0f;
                        gim += sh ? sn[t] : 0.0f;
                        cs[t] = sh ? 0.0f : cs[t];
                        sn[t] = sh ? 0.0f : sn[t];
                    }
                } else if (all_in) {
                    // every pixel is shaded for every lane: E' = 0, only the shaded sums
                    float sr = cs[0], si = sn[0];
#pragma unroll
                    for (int t = 1; t < 8; ++t) {
                        sr += cs[t];
                        si += sn[t];
                    }
                    gre += sr;
                    gim += si;
                }
                if (!all_in) {  // wave-uniform
#pragma unroll
                    for (int t2 = 0; t2 < 4; ++t2) {
                        split2(cs[2 * t2], cs[2 * t2 + 1], bre_h.u[t2], bre_l.u[t2]);
                        split2(sn[2 * t2], sn[2 * t2 + 1], bim_h.u[t2], bim_l.u[t2]);
                    }
#pragma unroll
                    for (int mt = 0; mt < kMaps / kMTV; ++mt) {
                        const int row = mt * kMTV + col, so = a_slot(row, ks * 4 + kg) * 8;
                        const half8 ah = *reinterpret_cast<const half8 *>(&sA[buf][0][row][so]);
                        const half8 al = *reinterpret_cast<const half8 *>(&sA[buf][1][row][so]);
                        acc[mt][0] = mfma16(ah, bre_h.v, acc[mt][0]);
                        acc[mt][0] = mfma16(ah, bre_l.v, acc[mt][0]);
                        acc[mt][0] = mfma16(al, bre_h.v, acc[mt][0]);
                        acc[mt][1] = mfma16(ah, bim_h.v, acc[mt][1]);
                        acc[mt][1] = mfma16(ah, bim_l.v, acc[mt][1]);
                        acc[mt][1] = mfma16(al, bim_h.v, acc[mt][1]);
                    }
                }
            }
            wait_vm();
            __syncthreads();
        }
        // shaded-pixel term: the four K groups of this column
        gre += __shfl_xor(gre, 16, 64);
        gim += __shfl_xor(gim, 16, 64);
        gre += __shfl_xor(gre, 32, 64);
        gim += __shfl_xor(gim, 32, 64);
        if (vok) {
#pragma unroll
            for (int mt = 0; mt < kMaps / kMTV; ++mt)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    // D of 16x16x32: lane holds rows 4 * (lane / 16) + r of column lane % 16
                    const int64_t m = grp * kMaps + mt * kMTV + 4 * kg + r;
                    if (m < n_maps) {
                        const float is = inv_scale[m];
                        // V = sum I E' + T_occ sum_shaded E, E = cos - i sin
                        V[m * g.n_vis + v] = make_float2(acc[mt][0][r] * is + t_occ * gre,
                                                         -(acc[mt][1][r] * is) - t_occ * gim);
                    }
                }
        }
    }
}

// ------------------------------------------------------------------ recon
// Block: 16 waves x 32 pixels (two chunks, waves 0-7 and 8-15); K runs over
// visibilities, staged 64 at a time; each staged tile of weighted V serves the
// block's 512 pixels.
// The weighted-V tile of the next stage moves straight into the other LDS
// buffer while the current stage computes; wave 0 reads the next stage's
// visibility geometry into registers at the top of a stage and classifies
// and stores it at the bottom.
struct CapLite {  // the cap fields a pixel test reads (CapRec without its 32-byte alignment)
    float dx, dy, dz, c, g;
};
// 88 bytes: lanes 0-31 and 32-63 read entries 8 apart, 704 bytes = 176 dwords,
// so their 16-byte reads land on disjoint banks.
struct VisS {
    double u, v, w, wo;
    int32_t cls, ia, ib, pad0;  // cls: class against the block's chunk h in bits 2h, 2h+1
    CapLite a, b;
};
static_assert(sizeof(VisS) == 88, "staged visibility record layout");

__device__ __forceinline__ CapLite cap_lite(const CapRec &r) { return {r.dx, r.dy, r.dz, r.c, r.g}; }
__device__ __forceinline__ CapRec cap_full(const CapLite &l) {
    return {l.dx, l.dy, l.dz, l.c, 0.f, l.g, 0.f, 0.f};
}

struct VisRaw {  // what wave 0 holds between the two halves of staging
    int2 cap;
    bool ok;
    double4 q;
};

__device__ __forceinline__ void vis_raw_load(const GeomDev &g, int64_t vv, VisRaw &r) {
    r.ok = vv < g.n_vis;
    if (r.ok) {
        r.cap = g.vis_cap[vv];
        r.q = g.uvw64[vv];
    }
}

__device__ __forceinline__ void vis_stage_store(const GeomDev &g, const PhaseKnobs &k,
                                                const ChunkRec &ch0, const ChunkRec &ch1,
                                                const VisRaw &r, VisS &e) {
    if (r.ok) {
        e.ia = r.cap.x;
        e.ib = r.cap.y;
        const CapRec a = g.cap[e.ia], b = g.cap[e.ib];
        e.a = cap_lite(a);
        e.b = cap_lite(b);
        e.u = r.q.x * k.s; e.v = r.q.y * k.s; e.w = r.q.z * k.s;
        e.wo = k.wfac ? 0.0 : -e.w;
        e.cls = classify_m(ch0, a, b) | (classify_m(ch1, a, b) << 2);
    } else {
        e.ia = e.ib = 0;
        const CapLite none = {0.f, 0.f, 0.f, 2.0f, 0.f};
        e.a = none;
        e.b = none;
        e.u = e.v = e.w = e.wo = 0.0;
        e.cls = kIn | (kIn << 2);  // padding: contributes nothing
    }
}

// A wave-uniform 64-bit value kept in scalar registers.
__device__ __forceinline__ uint64_t uniform64(uint64_t x) {
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)x);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(x >> 32));
    return ((uint64_t)hi << 32) | lo;
}

// Wave 0, after staging: bit masks over the 64 staged visibilities of which
// ones shade all of the block's chunk h, and which straddle it (h = 0, 1).
__device__ __forceinline__ void stage_masks(int cls, uint64_t (&m)[2][2]) {
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const int c = (cls >> (2 * h)) & 3;
        const uint64_t in_m = __ballot(c == kIn), st_m = __ballot(c == kStraddle);
        if ((threadIdx.x & 63) == 0) {
            m[h][0] = in_m;
            m[h][1] = st_m;
        }
    }
}

__global__ void __launch_bounds__(kThreadsR, 4) mfma_recon_kernel(
    GeomDev g, PhaseKnobs k, int32_t n_maps, int64_t n_vis_pad, const _Float16 *__restrict__ A16,
    const float *__restrict__ inv_scale, int32_t n_seg, float *__restrict__ part) {
    __shared__ _Float16 sA[2][4][kMaps][kStage];  // [buffer][re hi, re lo, im hi, im lo][map][vis slot]
    __shared__ VisS sV[2][kStage];
    // [buffer][chunk h][0: shades all of chunk h, 1: straddles it], one bit per staged visibility
    __shared__ uint64_t sMask[2][2][2];
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // wave-uniform
    const int col = lane & 31;
    const int kg = lane >> 5;
    const int half = wave / (kWavesR / 2);  // which of the block's two chunks
    const int64_t n_ptiles = (g.n_pad + kWavesR * kNT - 1) / (kWavesR * kNT);
    const int64_t n_groups = (n_maps + kMaps - 1) / kMaps;
    const int64_t per_seg = (n_vis_pad + n_seg - 1) / n_seg;
    const int64_t seg_len = (per_seg + kStage - 1) / kStage * kStage;
    const int64_t n_items = n_ptiles * n_groups * n_seg;

    for (int64_t item = blockIdx.x; item < n_items; item += gridDim.x) {
        const int64_t seg = item % n_seg;
        const int64_t rest = item / n_seg;
        const int64_t pt = rest % n_ptiles;
        const int64_t grp = rest / n_ptiles;
        const int64_t pi = pt * (kWavesR * kNT) + wave * kNT + col;  // chunk-order pixel
        // an odd chunk count leaves the last block's second chunk empty
        const int64_t c0 = (pt * (kWavesR * kNT)) / kChunk;
        const bool live = c0 + half < g.n_chunk;  // wave-uniform
        const double4 s = live ? g.pix64[pi] : make_double4(0.0, 0.0, 0.0, 0.0);
        const ChunkRec ch0 = g.chunk[c0];
        const ChunkRec ch1 = g.chunk[c0 + 1 < g.n_chunk ? c0 + 1 : c0];
        floatx16 acc[2];
        acc[0] = floatx16{};
        acc[1] = floatx16{};
        const _Float16 *Ag = A16 + grp * 4 * kMaps * n_vis_pad;
        const int64_t vb = seg * seg_len;
        const int64_t ve = (vb + seg_len < n_vis_pad) ? vb + seg_len : n_vis_pad;
        const int64_t n_st = ve > vb ? (ve - vb) / kStage : 0;   // both multiples of kStage
        VisRaw raw;
        if (n_st > 0) {
            // Both buffers are free: every wave passed the previous item's last barrier.
            stage_load_lds<4, kWavesR>(Ag, n_vis_pad, nullptr, vb / kStage, &sA[0][0][0][0], nullptr);
            if (threadIdx.x < kStage) {  // wave 0
                vis_raw_load(g, vb + threadIdx.x, raw);
                vis_stage_store(g, k, ch0, ch1, raw, sV[0][threadIdx.x]);
                stage_masks(sV[0][threadIdx.x].cls, sMask[0]);
            }
            wait_vm();
            __syncthreads();
        }
        for (int64_t st = 0; st < n_st; ++st) {
            const int buf = (int)(st & 1);
            const bool more = st + 1 < n_st;
            const int64_t v1 = vb + (st + 1) * kStage;
            if (more) {
                if (threadIdx.x < kStage) vis_raw_load(g, v1 + threadIdx.x, raw);
                stage_load_lds<4, kWavesR>(Ag, n_vis_pad, nullptr, v1 / kStage, &sA[buf ^ 1][0][0][0],
                                  nullptr);
            }
            const uint64_t in_m = uniform64(sMask[buf][half][0]), st_m = uniform64(sMask[buf][half][1]);
            // a stage whose visibilities all shade the wave's chunk adds nothing: skip it
            if (live && in_m != ~0ull) {
#pragma unroll
                for (int ks = 0; ks < kStage / kKS; ++ks) {
                    // E for this lane's pixel and 8 staged visibilities: straight-line
                    // phase and sin/cos, then the classes (the same across each half-wave).
                    // Lanes 0-31 take staged visibilities kb0 .. kb0 + 7, lanes 32-63
                    // the next 8; their classes are bits of the wave-uniform masks.
                    const int kb0 = ks * kKS, kb = kb0 + kg * 8;
                    const uint32_t in16 = (uint32_t)(in_m >> kb0) & 0xffffu;
                    const uint32_t st16 = (uint32_t)(st_m >> kb0) & 0xffffu;
                    float cs[8], sn[8];
#pragma unroll
                    for (int t = 0; t < 8; ++t) {
                        const VisS &q = sV[buf][kb + t];
                        const double ph = fma(q.u, s.x, fma(q.v, s.y, fma(q.w, s.z, q.wo)));
                        const float r = (float)__builtin_amdgcn_fract(ph);
                        cs[t] = __builtin_amdgcn_cosf(r);
                        sn[t] = __builtin_amdgcn_sinf(r);
                    }
                    H8 bre_h, bre_l, bim_h, bim_l;
                    if ((in16 | st16) != 0) {  // wave-uniform: some entry is shaded
                        uint64_t lm[8];  // lanes for which entry t is shaded
                        if (st16 != 0) {  // some entry straddles: test pixel by pixel
                            uint32_t shm = (in16 >> (8 * kg)) & 0xffu;
                            uint32_t stm = (st16 >> (8 * kg)) & 0xffu;
                            while (stm) {
                                const int t = __ffs(stm) - 1;
                                stm &= stm - 1;
                                const VisS &q = sV[buf][kb + t];
                                if (pair_shaded(g, s, cap_full(q.a), cap_full(q.b), q.ia, q.ib))
                                    shm |= 1u << t;
                            }
#pragma unroll
                            for (int t = 0; t < 8; ++t) lm[t] = __ballot((shm >> t) & 1u);
                        } else {
                            // entry t is shaded for a whole half-wave or not at all: the
                            // lane mask is formed on the scalar unit
#pragma unroll
                            for (int t = 0; t < 8; ++t)
                                lm[t] = (((in16 >> t) & 1u) ? 0x00000000ffffffffull : 0ull) |
                                        (((in16 >> (8 + t)) & 1u) ? 0xffffffff00000000ull : 0ull);
                        }
#pragma unroll
                        for (int t = 0; t < 8; ++t) {
                            const bool sh = __builtin_amdgcn_inverse_ballot_w64(lm[t]);
                            cs[t] = sh ? 0.0f : cs[t];
                            sn[t] = sh ? 0.0f : sn[t];
                        }
                    }
#pragma unroll
                    for (int t2 = 0; t2 < 4; ++t2) {
                        split2(cs[2 * t2], cs[2 * t2 + 1], bre_h.u[t2], bre_l.u[t2]);
                        split2(sn[2 * t2], sn[2 * t2 + 1], bim_h.u[t2], bim_l.u[t2]);
                    }
#pragma unroll
                    for (int mt = 0; mt < 2; ++mt) {
                        const int row = mt * kMT + col, so = a_slot(row, ks * 2 + kg) * 8;
                        const half8 arh = *reinterpret_cast<const half8 *>(&sA[buf][0][row][so]);
                        const half8 arl = *reinterpret_cast<const half8 *>(&sA[buf][1][row][so]);
                        const half8 aih = *reinterpret_cast<const half8 *>(&sA[buf][2][row][so]);
                        const half8 ail = *reinterpret_cast<const half8 *>(&sA[buf][3][row][so]);
                        acc[mt] = mfma(arh, bre_h.v, acc[mt]);
                        acc[mt] = mfma(arh, bre_l.v, acc[mt]);
                        acc[mt] = mfma(arl, bre_h.v, acc[mt]);
                        acc[mt] = mfma(aih, bim_h.v, acc[mt]);
                        acc[mt] = mfma(aih, bim_l.v, acc[mt]);
                        acc[mt] = mfma(ail, bim_h.v, acc[mt]);
                    }
                }
            }
            // the other buffer was last read before the previous barrier
            if (more && threadIdx.x < kStage) {  // wave 0
                vis_stage_store(g, k, ch0, ch1, raw, sV[buf ^ 1][threadIdx.x]);
                stage_masks(sV[buf ^ 1][threadIdx.x].cls, sMask[buf ^ 1]);
            }
            wait_vm();
            __syncthreads();
        }
        if (live) {
#pragma unroll
            for (int mt = 0; mt < 2; ++mt)
#pragma unroll
                for (int r = 0; r < 16; ++r) {
                    const int64_t m = grp * kMaps + mt * kMT + acc_row(r, lane);
                    if (m < n_maps)
                        part[(m * n_seg + seg) * g.n_pad + pi] = acc[mt][r] * inv_scale[m];
                }
        }
    }
}

// ------------------------------------------------------------------ per-satellite factors
// On an orbit geometry u v w = (r_y - r_x) / lambda for pair (x, y) at one step,
// so the phase s (u l + v m + w (n - 1)) of every pair is phi_y - phi_x with
// phi_x = s r_x . (l, m, n - 1) / lambda (n for the trailing-w convention), and
// E_xy = E_y conj(E_x): with (c, s) = (cos, sin) 2 pi phi,
//   cos(phi_y - phi_x) = c_y c_x + s_y s_x,   sin(phi_y - phi_x) = s_y c_x - c_y s_x.
// The shade mask factorises the same way: (x, y) is clear for a pixel iff
// neither satellite's ray to it is occulted, so zeroing E_x on the pixels that
// satellite x cannot see zeroes exactly the shaded entries of every product.
// Each per-satellite factor (FP64 phase and fract, FP32 sin/cos, its own
// bit-exact occultation test) serves every pair that satellite is in: n_sat
// phases and sin/cos per step instead of n_base.

// Pairs of an n-satellite array (n even) in star order: star c (c = 0 .. n - 2)
// holds the pairs of satellite c with satellite n - 1 and with c + 1 .. c +
// (n - 2) / 2 (mod n - 1), so n - 1 stars of n / 2 pairs cover every pair once
// (a regular tournament on the first n - 1 satellites, each beating the last).
// Entry j of the list is pair (star_x, star_y), x < y; star_center its star's
// centre.  Past the n (n - 1) / 2 pairs the entries are padding (-1).
__host__ __device__ constexpr int star_center(int n, int j) { return j / (n / 2); }
__host__ __device__ constexpr int star_leaf(int n, int j) {
    return j % (n / 2) == 0 ? n - 1 : (star_center(n, j) + j % (n / 2)) % (n - 1);
}
__host__ __device__ constexpr int star_x(int n, int j) {
    return j >= n * (n - 1) / 2 ? -1
                                : (star_center(n, j) < star_leaf(n, j) ? star_center(n, j) : star_leaf(n, j));
}
__host__ __device__ constexpr int star_y(int n, int j) {
    return j >= n * (n - 1) / 2 ? -1
                                : (star_center(n, j) < star_leaf(n, j) ? star_leaf(n, j) : star_center(n, j));
}
// Index of pair (x, y), x < y, in the lexicographic pair list (0,1), (0,2), ...
__host__ __device__ constexpr int lex_pair(int n, int x, int y) { return x * n - x * (x + 1) / 2 + (y - x - 1); }

// The star order of n satellites holds every pair exactly once, each with the
// smaller satellite first, and every pair of star c holds satellite c.
constexpr bool star_order_ok(int n) {
    uint64_t seen[4] = {0, 0, 0, 0};  // bit lex_pair of the pairs met (n <= 16)
    for (int j = 0; j < n * (n - 1) / 2; ++j) {
        const int x = star_x(n, j), y = star_y(n, j), c = star_center(n, j);
        if (x < 0 || x >= y || y >= n || (x != c && y != c)) return false;
        const int b = lex_pair(n, x, y);
        if ((seen[b >> 6] >> (b & 63)) & 1u) return false;
        seen[b >> 6] |= uint64_t(1) << (b & 63);
    }
    return star_x(n, n * (n - 1) / 2) < 0;
}
static_assert(star_order_ok(16) && star_order_ok(8) && star_order_ok(4), "star order of the pairs");

// A buffer index as a type, for stage loops unrolled by two.
template <int B>
struct BufC {
    static constexpr int value = B;
};

// Largest number of (step, satellite) rows one vis-gen block holds per stage.
template <int TB>
struct SatRows {
    static constexpr int value = TB == 2 ? 32 : 44;
};
constexpr int kSatLd = kStage + 2;      // float2 per LDS row: 64 pixels + 16 bytes (rows on shifted banks)
constexpr int kSatSlots = (kWaves - 1) * kNTV;  // visibilities per block: 15 waves x 16

// x = hi + lo in FP16: hi rounded toward zero, lo = x - hi rounded to nearest.
// neg1 is -1 hidden from the optimiser (opaque_neg1): fma(hi, neg1, x) is then
// formed by v_fma_mix_f32, which widens the FP16 operand exactly, instead of a
// widening conversion and a subtraction (4 VALU per pair instead of 6).  Forming
// each half of lo in place (v_fma_mixlo / v_fma_mixhi, 3 VALU per pair) measured
// slower: the second half waits for the first.
__device__ __forceinline__ float opaque_neg1() {
    float m;
    asm("" : "=s"(m) : "0"(-1.0f));
    return m;
}
__device__ __forceinline__ void split2m(float x0, float x1, float neg1, uint32_t &hi, uint32_t &lo) {
    typedef _Float16 half2_t __attribute__((ext_vector_type(2)));
    const half2_t h = __builtin_bit_cast(half2_t, __builtin_amdgcn_cvt_pkrtz(x0, x1));
    half2_t l;
    l[0] = (_Float16)__builtin_fmaf((float)h[0], neg1, x0);
    l[1] = (_Float16)__builtin_fmaf((float)h[1], neg1, x1);
    hi = __builtin_bit_cast(uint32_t, h);
    lo = __builtin_bit_cast(uint32_t, l);
}

// The map tile of stage st (both FP16 planes, 64 maps x 64 pixels, 16 KiB) as
// one 16-byte piece per thread of a 1024-thread block, loaded into registers;
// tile_store puts it at the same place stage_load_lds would (the swizzled
// LDS layout).  Staging through registers keeps the LDS reads of the current
// stage free of waits: a read that may alias an in-flight global->LDS load
// waits for every such load (vmcnt(0)), which exposed the next stage's whole
// load latency at the first K-step.
__device__ __forceinline__ uint4 tile_load(const _Float16 *__restrict__ Ag, int64_t ld, int64_t st) {
    const int i = threadIdx.x;                       // 16-byte slot of the tile
    const int hl = i / (kMaps * (kStage / 8));
    const int rem = i % (kMaps * (kStage / 8));
    const int row = rem / (kStage / 8), slot = rem % (kStage / 8);
    const int seg = a_slot(row, slot);
    return *reinterpret_cast<const uint4 *>(Ag + ((int64_t)hl * kMaps + row) * ld + st * kStage + seg * 8);
}
__device__ __forceinline__ void tile_store(_Float16 *sAbuf, uint4 x) {
    reinterpret_cast<uint4 *>(sAbuf)[threadIdx.x] = x;
}
static_assert(2 * kMaps * kStage * 2 == kThreads * 16, "one 16-byte piece of the map tile per thread");

// A row's satellite position (FP64; w = |r|^2, for the exact predicate), the
// same scaled by s / lambda (the phase of a pixel is one dot product with it)
// and the row's shade cap: staged in LDS once per tile, so that forming a stage
// reads no global memory (LDS reads do not queue behind the vector-memory
// prefetches of the next stage).
struct SatRow {
    double4 pos;
    double4 spos;
    CapRec cap;
};
static_assert(sizeof(SatRow) == 96, "row record layout");

// A stage's pixel inputs, staged in LDS: pixel_piece loads one 16-byte piece of
// them into a register at the top of the stage two before their use, and
// pixel_store puts it in place at the end of that stage.  Pieces: 128 of the FP64
// vectors, 64 of the FP32 vectors, 16 of the RING indices, 2 of the 64-pixel
// group's bounding cap (210 of the block's threads take one each).
struct PixStage {
    double4 p64[kStage];
    float4 p32[kStage];
    int32_t ring[kStage];
    ChunkRec sub;
};
constexpr int kPixPieces = kStage * 2 + kStage + kStage / 4 + 2;
static_assert(sizeof(PixStage) == 16 * kPixPieces, "pixel pieces tile the record");

// Where this thread's piece of every stage lies: stage st's piece at base + st *
// stride (threads past the pieces reload piece 0 and store nothing).  Picked
// once per kernel, so a stage's load is one address computation (a branch per
// source at each stage would serialise the loads into one register behind waits).
struct PixSrc {
    const char *base;
    int32_t stride;
};
__device__ __forceinline__ PixSrc pixel_src(const GeomDev &g) {
    const int i = threadIdx.x;
    if (i < 2 * kStage) return {reinterpret_cast<const char *>(g.pix64) + 16 * i, kStage * 32};
    if (i < 3 * kStage) return {reinterpret_cast<const char *>(g.pix32) + 16 * (i - 2 * kStage), kStage * 16};
    if (i < 3 * kStage + kStage / 4)
        return {reinterpret_cast<const char *>(g.ring) + 16 * (i - 3 * kStage), kStage * 4};
    if (i < kPixPieces)
        return {reinterpret_cast<const char *>(g.sub) + 16 * (i - 3 * kStage - kStage / 4),
                (int32_t)sizeof(ChunkRec)};
    return {reinterpret_cast<const char *>(g.pix64), kStage * 32};
}
static_assert(sizeof(double4) == 32 && sizeof(float4) == 16 && sizeof(ChunkRec) == 32,
              "pixel piece strides");
__device__ __forceinline__ uint4 pixel_piece(const PixSrc &src, int64_t st) {
    return *reinterpret_cast<const uint4 *>(src.base + st * src.stride);
}
__device__ __forceinline__ void pixel_store(PixStage *d, uint4 x) {
    if (threadIdx.x < kPixPieces) reinterpret_cast<uint4 *>(d)[threadIdx.x] = x;
}

// Rows r = wave + kWaves i (i < MR, bit i of `live`) of one stage: for (step
// t0 + r / n_sat, satellite r % n_sat) and the stage's 64 pixels (lane = pixel),
// E_x = (cos, sin) of 2 pi phi_x with the pixels that satellite cannot see
// zeroed into sEm, and with FULL the unzeroed factor into sEf (the Gram sums
// need it; with FULL, padding pixels are zero in both, while vis-gen alone
// needs no padding test: the map operand is zero on them).  Occultation per pixel: the FP32 test
// against the guard band's edges, the exact FP64 predicate inside the band (the
// rule of every kernel here; a per-pixel test costs fewer instructions than
// classifying the row's cap against the 64-pixel group first).  PADS: the
// stage may hold padding pixels.  Returns the AND over the wave's live rows of
// the lane bits of pixels occulted for the row or padding (where the row
// contributes nothing to any product); all ones if no row is live.
template <bool FULL, bool PADS, int MR>
__device__ __forceinline__ uint64_t sat_rows(const GeomDev &g, double dzo, int wave, uint32_t live,
                                             const PixStage *ps, const SatRow *sRow,
                                             float2 (*sEm)[kSatLd], float2 (*sEf)[kSatLd]) {
    const int lane = threadIdx.x & 63;
    const double4 s = ps->p64[lane];
    const float4 s32 = ps->p32[lane];
    const bool pad = PADS && ps->ring[lane] < 0;
    const double szd = s.z - dzo;
    uint64_t none = ~0ull;
#pragma unroll
    for (int i = 0; i < MR; ++i) {
        if (!((live >> i) & 1u)) continue;  // wave-uniform
        const int r = wave + kWaves * i;
        const double4 sp = sRow[r].spos;
        const CapRec cap = sRow[r].cap;
        const double ph = fma(sp.x, s.x, fma(sp.y, s.y, sp.z * szd));
        const float rv = (float)__builtin_amdgcn_fract(ph);
        float2 ef = make_float2(__builtin_amdgcn_cosf(rv), __builtin_amdgcn_sinf(rv));
        if (FULL && PADS && pad) ef = make_float2(0.f, 0.f);
        const float x = fmaf(s32.z, cap.dz, fmaf(s32.y, cap.dy, s32.x * cap.dx));
        bool occ = x >= cap.hi;
        if (x >= cap.lo && !occ) occ = occ64m(sRow[r].pos, s.x, s.y, s.z, g.limb);
        sEm[r][lane] = occ ? make_float2(0.f, 0.f) : ef;
        if (FULL) sEf[r][lane] = ef;
        none &= __ballot(PADS ? occ || pad : occ);
    }
    return none;
}

// 16 x 16 Gram blocks of one step, staged for the consumers: [P1, P2] (full-sky
// sums; see mfma_visgen_sat_kernel).
constexpr int kGramLd16 = 17;

// Vis-gen on per-satellite factors.  Block: TB consecutive steps of every
// baseline (TB * n_base <= 240 visibilities); waves 0-14 hold 16 slots each on
// the 16x16x32 instruction (the products of mfma_visgen_kernel, with E'_xy =
// E'_y conj(E'_x) formed from the two staged rows of each lane's pair); every
// wave forms its share of the next stage's rows after its K-steps.  Slots: j =
// step (j / n_base), baseline (j % n_base); with STAR (every pair of 16
// satellites, two steps per block), wave w holds star w of the star order at
// both steps, so all its pairs hold satellite w, and it skips the K-steps on
// whose 32 pixels satellite w is occulted at both steps: all its products there
// are zero.  Wave w forms rows w and w + 16 itself (satellite w at both steps),
// so the pixels to skip come out of its own row formation.
//
// Shaded-pixel term (TOCC): V = sum_clear I E + t_occ sum_shaded E =
// sum_clear (I - t_occ) E + t_occ U, with U_xy = sum_p E_y conj(E_x) over every
// pixel; the map operand holds I - t_occ.  U is a Gram matrix of the unmasked
// rows of a step, which wave 15 contracts on the matrix cores: with C, S the
// cos and sin rows of the 16 satellites, Re U_xy = P1[y][x] = (C C^T + S
// S^T)[y][x], and sum_p sin(phi_y - phi_x) = P2[x][y] - P2[y][x], P2 = C S^T.
// Split FP16: P1 = K + K^T with K = Hc Hc^T / 2 + Hc Lc^T + Hs Hs^T / 2 + Hs
// Ls^T, and P2 = Hc Hs^T + Hc Ls^T + Lc Hs^T.  Each wave role runs its own
// stage loop (with the same barriers), so the Gram accumulators and the
// consumers' accumulators never share registers.
template <int TB, bool TOCC, bool STAR>
__global__ void __launch_bounds__(kThreads) mfma_visgen_sat_kernel(
    GeomDev g, PhaseKnobs k, double sc, int32_t n_maps, const _Float16 *__restrict__ A16,
    const float *__restrict__ inv_scale, float2 *__restrict__ V, int32_t pads) {
    constexpr int R = SatRows<TB>::value;
    __shared__ _Float16 sA[2][2][kMaps][kStage];  // [buffer][hi/lo][map][pixel]
    __shared__ float2 sE[2][TOCC ? 2 : 1][R][kSatLd];  // [buffer][masked, full][row][pixel]
    __shared__ float sG[TOCC ? TB : 1][2][16][kGramLd16];
    // the Gram wave's running totals: sum and compensation (each lane owns its
    // entries, [q][P1, P2][row][column] as in sG)
    __shared__ float sGt[TOCC && TB == 2 ? TB : 1][2][16][kGramLd16];
    __shared__ float sGc[TOCC && TB == 2 ? TB : 1][2][16][kGramLd16];
    __shared__ SatRow sRow[R];
    __shared__ PixStage sPix[2];  // [stage & 1]
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int col = lane & 15, kg = lane >> 4;
    const int n_sat = g.n_sat, n_base = g.n_base;
    const int n_rows = TB * n_sat;
    const int n_slot = TB * n_base;
    const int n_ct = (n_slot + kNTV - 1) / kNTV;  // consumer waves
    const int64_t n_tb = (g.n_t + TB - 1) / TB;
    const int64_t n_groups = (n_maps + kMaps - 1) / kMaps;
    const int64_t n_tiles = n_tb * n_groups;
    const int64_t n_stages = g.n_pad / kStage;
    const double dzo = k.wfac ? 0.0 : 1.0;
    const float t_occ = k.t_occ;
    const bool consumer = wave < n_ct;             // wave-uniform
    const bool gram = TOCC && wave == kWaves - 1;  // wave-uniform
    constexpr int MR = (R + kWaves - 1) / kWaves;  // most rows one wave forms
    const int nr = wave < n_rows ? (n_rows - 1 - wave) / kWaves + 1 : 0;  // rows this wave forms
    int rtl[MR];  // step (within the block) of each of this wave's rows
#pragma unroll
    for (int i = 0; i < MR; ++i) rtl[i] = (wave + kWaves * i) / n_sat;
    // this lane's slot (fixed over the tiles)
    const int j = wave * kNTV + col;
    int tlv, bi, rx = 0, ry = 0, px = 0, py = 0;
    if (STAR) {
        const int h = n_sat / 2;  // pairs per star
        tlv = col / h;
        const int js = wave * h + col % h;
        px = consumer ? star_x(n_sat, js) : 0;
        py = consumer ? star_y(n_sat, js) : 0;
        bi = lex_pair(n_sat, px, py);
    } else {
        tlv = j / n_base;
        bi = j - tlv * n_base;
        if (consumer && j < n_slot) {
            const int2 pr = g.pairs[bi];
            px = pr.x;
            py = pr.y;
        }
    }
    if (consumer && j < n_slot) {
        rx = tlv * n_sat + px;
        ry = tlv * n_sat + py;
    }
    const float neg1 = opaque_neg1();
    const PixSrc psrc = pixel_src(g);

    // Each wave role runs its own tile loop with the same barriers (two per tile
    // set-up, one per stage, one before the results with TOCC), so that the
    // registers of one role are not held live through another's.
    auto tile_of = [&](int64_t tile, int64_t &grp, int &t0, uint32_t &live) {
        grp = tile / n_tb;
        t0 = (int)(tile % n_tb) * TB;
        live = 0;  // this wave's rows whose step exists
#pragma unroll
        for (int i = 0; i < MR; ++i)
            if (i < nr && t0 + rtl[i] < g.n_t) live |= 1u << i;
    };
    // Rows of one stage (pads: the geometry has padding pixels)
    auto form_rows = [&](const PixStage *ps, uint32_t live, float2(*em)[kSatLd], float2(*ef)[kSatLd]) {
        return pads ? sat_rows<TOCC, true, MR>(g, dzo, wave, live, ps, sRow, em, ef)
                    : sat_rows<TOCC, false, MR>(g, dzo, wave, live, ps, sRow, em, ef);
    };
    auto tile_begin = [&](int t0, uint32_t live, const _Float16 *Ag) {
        // The buffers, sRow and sPix are free: every wave passed the previous
        // tile's last barrier, and the last stage forms no rows.
        if (threadIdx.x < n_rows) {
            const int r = threadIdx.x;
            const int tl = r / n_sat, sat = r - tl * n_sat;
            SatRow e;
            if (t0 + tl < g.n_t) {
                const int64_t row = (int64_t)sat * g.n_t + t0 + tl;
                e.pos = g.pos[row];
                e.cap = g.cap[row];
            } else {
                e.pos = make_double4(0.0, 0.0, 0.0, 0.0);
                e.cap = {0.f, 0.f, 0.f, 2.0f, 0.f, kGuardMin, 2.0f - kGuardMin, 2.0f + kGuardMin};
            }
            e.spos = make_double4(e.pos.x * sc, e.pos.y * sc, e.pos.z * sc, 0.0);
            sRow[r] = e;
        }
        tile_store(&sA[0][0][0][0], tile_load(Ag, g.n_pad, 0));
        pixel_store(&sPix[0], pixel_piece(psrc, 0));
        if (n_stages > 1) pixel_store(&sPix[1], pixel_piece(psrc, 1));
        __syncthreads();
        const uint64_t none = form_rows(&sPix[0], live, sE[0][0], sE[0][TOCC ? 1 : 0]);
        __syncthreads();
        return none;
    };
    // Prefetch into registers, stored at the end of each stage: the map tile
    // of stage st + 1 and the pixel inputs of stage st + 2 (loaded without a
    // branch: the last stages reload their own).  Nothing in a stage waits
    // for them.
    auto stage_end = [&](int64_t st, uint4 tnext, uint4 pnext, auto bufc) {
        constexpr int buf = decltype(bufc)::value;
        // sA[buf ^ 1] was last read before the previous barrier; sPix[buf] holds
        // stage st's inputs, last read while forming stage st
        asm volatile("" ::: "memory");
        if (st + 1 < n_stages) tile_store(&sA[buf ^ 1][0][0][0], tnext);
        if (st + 2 < n_stages) pixel_store(&sPix[buf], pnext);
        __syncthreads();
    };

    if (consumer) {
        for (int64_t tile = blockIdx.x; tile < n_tiles; tile += gridDim.x) {
            int64_t grp;
            int t0;
            uint32_t live;
            tile_of(tile, grp, t0, live);
            const _Float16 *Ag = A16 + grp * 2 * kMaps * g.n_pad;
            uint64_t none = tile_begin(t0, live, Ag);  // of the next stage to consume
            floatx4 acc[kMaps / kMTV][2];
#pragma unroll
            for (int a = 0; a < kMaps / kMTV; ++a)
#pragma unroll
                for (int c = 0; c < 2; ++c) acc[a][c] = floatx4{};
            // the stage loop unrolled by two: buffer indices known at compile time
            auto cstage = [&](int64_t st, auto bufc) {
                constexpr int buf = decltype(bufc)::value;
                const uint4 tnext = tile_load(Ag, g.n_pad, st + 1 < n_stages ? st + 1 : st);
                const uint4 pnext = pixel_piece(psrc, st + 2 < n_stages ? st + 2 : st);
                // K-steps whose products are all zero (STAR): bit ks
                uint32_t skip = 0;
                if (STAR) skip = ((uint32_t)none == ~0u ? 1u : 0u) | ((uint32_t)(none >> 32) == ~0u ? 2u : 0u);
                // E of K-step ks for this lane's 8 pixels (ks * 32 + 8 kg + i, identity
                // K order), read one K-step ahead so that the reads of K-step ks + 1
                // are in flight while the products of K-step ks issue
                float4 ea[4], eb[4];
                auto read_e = [&](int ks) {
                    const float2 *ex = &sE[buf][0][rx][ks * kKSV + kg * 8];
                    const float2 *ey = &sE[buf][0][ry][ks * kKSV + kg * 8];
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        ea[q] = *reinterpret_cast<const float4 *>(ex + 2 * q);
                        eb[q] = *reinterpret_cast<const float4 *>(ey + 2 * q);
                    }
                };
                read_e(0);
#pragma unroll
                for (int ks = 0; ks < kStage / kKSV; ++ks) {
                    if ((skip >> ks) & 1u) {  // wave-uniform
                        if (ks + 1 < kStage / kKSV) read_e(ks + 1);
                        continue;
                    }
                    float cs[8], sn[8];
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        const float4 a = ea[q], b = eb[q];
                        cs[2 * q] = fmaf(b.x, a.x, b.y * a.y);
                        sn[2 * q] = fmaf(b.y, a.x, -(b.x * a.y));
                        cs[2 * q + 1] = fmaf(b.z, a.z, b.w * a.w);
                        sn[2 * q + 1] = fmaf(b.w, a.z, -(b.z * a.w));
                    }
                    if (ks + 1 < kStage / kKSV) read_e(ks + 1);
                    H8 bre_h, bre_l, bim_h, bim_l;
#pragma unroll
                    for (int t2 = 0; t2 < 4; ++t2) {
                        split2m(cs[2 * t2], cs[2 * t2 + 1], neg1, bre_h.u[t2], bre_l.u[t2]);
                        split2m(sn[2 * t2], sn[2 * t2 + 1], neg1, bim_h.u[t2], bim_l.u[t2]);
                    }
#pragma unroll
                    for (int mt = 0; mt < kMaps / kMTV; ++mt) {
                        const int row = mt * kMTV + col, so = a_slot(row, ks * 4 + kg) * 8;
                        const half8 ah = *reinterpret_cast<const half8 *>(&sA[buf][0][row][so]);
                        const half8 al = *reinterpret_cast<const half8 *>(&sA[buf][1][row][so]);
                        acc[mt][0] = mfma16(ah, bre_h.v, acc[mt][0]);
                        acc[mt][0] = mfma16(ah, bre_l.v, acc[mt][0]);
                        acc[mt][0] = mfma16(al, bre_h.v, acc[mt][0]);
                        acc[mt][1] = mfma16(ah, bim_h.v, acc[mt][1]);
                        acc[mt][1] = mfma16(ah, bim_l.v, acc[mt][1]);
                        acc[mt][1] = mfma16(al, bim_h.v, acc[mt][1]);
                    }
                }
                if (st + 1 < n_stages)
                    none = form_rows(&sPix[buf ^ 1], live, sE[buf ^ 1][0], sE[buf ^ 1][TOCC ? 1 : 0]);
                stage_end(st, tnext, pnext, bufc);
            };
            int64_t st = 0;
            for (; st + 1 < n_stages; st += 2) {
                cstage(st, BufC<0>{});
                cstage(st + 1, BufC<1>{});
            }
            if (st < n_stages) cstage(st, BufC<0>{});
            if (TOCC) __syncthreads();  // sG whole
            if (consumer && j < n_slot && t0 + tlv < g.n_t) {
                const int64_t v = (int64_t)bi * g.n_t + t0 + tlv;
                float shc = 0.0f, shs = 0.0f;
                if (TOCC) {
                    // t_occ times the pair's full-sky sum
                    const float(*G)[16][kGramLd16] = sG[tlv];
                    shc = t_occ * (G[0][py][px] + G[0][px][py]);
                    shs = t_occ * (G[1][px][py] - G[1][py][px]);
                }
#pragma unroll
                for (int mt = 0; mt < kMaps / kMTV; ++mt)
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const int64_t m = grp * kMaps + mt * kMTV + 4 * kg + r;
                        // V = sum_clear (I - t_occ) E + t_occ U, E = cos - i sin
                        if (m < n_maps) {
                            const float is = inv_scale[m];
                            V[m * g.n_vis + v] = make_float2(acc[mt][0][r] * is + shc,
                                                             -(acc[mt][1][r] * is) - shs);
                        }
                    }
            }
        }
    } else if (gram) {
        // this lane: satellite gs = lane % 16 (its row and column of the Gram
        // blocks), pixels 8 kg + i of each 32-pixel K-step.  The products of
        // kGramFold stages accumulate in registers, then fold into compensated
        // totals in LDS (two-sum): a full-sky sum of ~1e5 terms would otherwise
        // collect one FP32 rounding per product, at the magnitude of the sum.
        const int gs = lane & 15;
        constexpr int kGramFold = 16;
        constexpr bool kFold = TB == 2;  // four steps' blocks would spill
        auto fold = [&](floatx4 (&p1)[TB], floatx4 (&p2)[TB]) {
#pragma unroll
            for (int q = 0; q < TB; ++q)
#pragma unroll
                for (int k = 0; k < 2; ++k) {
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const float p = k ? p2[q][r] : p1[q][r];
                        float &t = sGt[q][k][4 * kg + r][gs];
                        float &c = sGc[q][k][4 * kg + r][gs];
                        const float t0_ = t, sum = t0_ + p, bp = sum - t0_;
                        c += (t0_ - (sum - bp)) + (p - bp);
                        t = sum;
                    }
                    __builtin_amdgcn_sched_barrier(0);  // one (q, P1 / P2) block at a time
                }
#pragma unroll
            for (int q = 0; q < TB; ++q) p1[q] = p2[q] = floatx4{};
        };
        for (int64_t tile = blockIdx.x; tile < n_tiles; tile += gridDim.x) {
            int64_t grp;
            int t0;
            uint32_t live;
            tile_of(tile, grp, t0, live);
            const _Float16 *Ag = A16 + grp * 2 * kMaps * g.n_pad;
            (void)tile_begin(t0, live, Ag);
            floatx4 p1[TB], p2[TB];
#pragma unroll
            for (int q = 0; q < TB; ++q) {
                p1[q] = p2[q] = floatx4{};
                if (kFold) {
#pragma unroll
                    for (int k = 0; k < 2; ++k)
#pragma unroll
                        for (int r = 0; r < 4; ++r) sGt[q][k][4 * kg + r][gs] = sGc[q][k][4 * kg + r][gs] = 0.0f;
                }
            }
            auto gstage = [&](int64_t st, auto bufc) {
                constexpr int buf = decltype(bufc)::value;
                const uint4 tnext = tile_load(Ag, g.n_pad, st + 1 < n_stages ? st + 1 : st);
                const uint4 pnext = pixel_piece(psrc, st + 2 < n_stages ? st + 2 : st);
#pragma unroll
                for (int q = 0; q < TB; ++q) {
                    if (t0 + q >= g.n_t) continue;  // wave-uniform
#pragma unroll 1
                    for (int kk = 0; kk < kStage / kKSV; ++kk) {
                        // one Gram block pair from the step's unmasked rows
                        float cc[8], ss[8];
                        if (gs < n_sat) {
                            const float2 *e = &sE[buf][TOCC ? 1 : 0][q * n_sat + gs][kk * kKSV + kg * 8];
#pragma unroll
                            for (int i = 0; i < 4; ++i) {
                                const float4 a = *reinterpret_cast<const float4 *>(e + 2 * i);
                                cc[2 * i] = a.x; ss[2 * i] = a.y; cc[2 * i + 1] = a.z; ss[2 * i + 1] = a.w;
                            }
                        } else {
#pragma unroll
                            for (int i = 0; i < 8; ++i) cc[i] = ss[i] = 0.0f;
                        }
                        H8 hc, lc, hs, ls, hc2, hs2;
#pragma unroll
                        for (int i = 0; i < 4; ++i) {
                            split2m(cc[2 * i], cc[2 * i + 1], neg1, hc.u[i], lc.u[i]);
                            split2m(ss[2 * i], ss[2 * i + 1], neg1, hs.u[i], ls.u[i]);
                        }
                        hc2.v = hc.v * (_Float16)0.5f;
                        hs2.v = hs.v * (_Float16)0.5f;
                        p1[q] = mfma16(hc2.v, hc.v, p1[q]);
                        p1[q] = mfma16(hc.v, lc.v, p1[q]);
                        p1[q] = mfma16(hs2.v, hs.v, p1[q]);
                        p1[q] = mfma16(hs.v, ls.v, p1[q]);
                        p2[q] = mfma16(hc.v, hs.v, p2[q]);
                        p2[q] = mfma16(hc.v, ls.v, p2[q]);
                        p2[q] = mfma16(lc.v, hs.v, p2[q]);
                    }
                }
                if (kFold && (st % kGramFold) == kGramFold - 1) fold(p1, p2);  // wave-uniform
                if (st + 1 < n_stages)
                    (void)form_rows(&sPix[buf ^ 1], live, sE[buf ^ 1][0], sE[buf ^ 1][TOCC ? 1 : 0]);
                stage_end(st, tnext, pnext, bufc);
            };
            int64_t st = 0;
            for (; st + 1 < n_stages; st += 2) {
                gstage(st, BufC<0>{});
                gstage(st + 1, BufC<1>{});
            }
            if (st < n_stages) gstage(st, BufC<0>{});
            if (kFold) fold(p1, p2);
            // D of 16x16x32: lane holds rows 4 (lane / 16) + r of column lane % 16
#pragma unroll
            for (int q = 0; q < TB; ++q)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int a = 4 * kg + r;
                    sG[q][0][a][gs] = kFold ? sGt[q][0][a][gs] + sGc[q][0][a][gs] : p1[q][r];
                    sG[q][1][a][gs] = kFold ? sGt[q][1][a][gs] + sGc[q][1][a][gs] : p2[q][r];
                }
            __syncthreads();  // sG whole
        }
    } else {
        for (int64_t tile = blockIdx.x; tile < n_tiles; tile += gridDim.x) {
            int64_t grp;
            int t0;
            uint32_t live;
            tile_of(tile, grp, t0, live);
            const _Float16 *Ag = A16 + grp * 2 * kMaps * g.n_pad;
            (void)tile_begin(t0, live, Ag);
            auto istage = [&](int64_t st, auto bufc) {
                constexpr int buf = decltype(bufc)::value;
                const uint4 tnext = tile_load(Ag, g.n_pad, st + 1 < n_stages ? st + 1 : st);
                const uint4 pnext = pixel_piece(psrc, st + 2 < n_stages ? st + 2 : st);
                if (st + 1 < n_stages)
                    (void)form_rows(&sPix[buf ^ 1], live, sE[buf ^ 1][0], sE[buf ^ 1][TOCC ? 1 : 0]);
                stage_end(st, tnext, pnext, bufc);
            };
            int64_t st = 0;
            for (; st + 1 < n_stages; st += 2) {
                istage(st, BufC<0>{});
                istage(st + 1, BufC<1>{});
            }
            if (st < n_stages) istage(st, BufC<0>{});
            if (TOCC) __syncthreads();  // sG whole
        }
    }
}
// ------------------------------------------------------------------ recon on per-satellite factors

// Layout of the recon operand over (step pair, pair): per step pair tp (steps
// 2 tp and 2 tp + 1) kKTP K-steps of 16 slots; slot 16 ks + 8 h + e holds pair
// 8 ks + e (star order) at step 2 tp + h.  kKTP is a multiple of 8, so each
// step pair fills an even number of 64-slot stages.
template <int NS>
struct SatRecon {
    static constexpr int kPairs = NS * (NS - 1) / 2;
    static constexpr int kKT = (kPairs + 7) / 8;
    static constexpr int kKTP = (kKT + 7) / 8 * 8;
    static constexpr int kSlots = 16 * kKTP;  // per step pair
    static constexpr int kStages = kKTP / 4;  // per step pair: even, so a stage's
                                              // buffer parity is known at compile time
};

// V [n_maps][n_vis] complex -> hi/lo FP16 in the step-pair slot order above,
// weighted by w * kappa and scaled per map, stored as the recon kernel's LDS
// tiles: [group][tile of 64 slots][re hi, re lo, im hi, im lo][map][64 slots],
// each 64-slot row in the a_slot swizzle, so a tile is one contiguous 32 KiB run
// that the kernel copies into LDS unchanged.
// Block item: one map row and kTPB step pairs.  The 2 kTPB steps of every pair
// (contiguous in V) are staged in LDS, weighted, then written slot by slot, so
// both the reads of V and the writes of each plane are contiguous runs.
template <int NS>
__global__ void __launch_bounds__(256) split_vis_sat_kernel(int64_t n_vis, int32_t n_t, int64_t n_slots,
                                                            int32_t n_maps, const float2 *__restrict__ V,
                                                            const float *__restrict__ w, float kappa,
                                                            const unsigned int *__restrict__ amax,
                                                            _Float16 *__restrict__ out,
                                                            float *__restrict__ inv_scale) {
    using L = SatRecon<NS>;
    constexpr int kTPB = 8, kT = 2 * kTPB;  // step pairs and steps per item
    __shared__ float2 sv[L::kPairs][kT];    // [lexicographic pair][step]: (Re, -Im) weighted
    const int64_t n_tp = n_slots / L::kSlots;
    const int64_t n_tb = (n_tp + kTPB - 1) / kTPB;
    const int64_t n_groups = (n_maps + kMaps - 1) / kMaps;
    const int64_t n_items = n_groups * kMaps * n_tb;
    for (int64_t item = blockIdx.x; item < n_items; item += gridDim.x) {
        const int64_t tb = item % n_tb, mrow = item / n_tb;
        const int64_t grp = mrow / kMaps, row = mrow % kMaps;
        const int64_t m = grp * kMaps + row;
        const int64_t t0 = tb * kT;
        const float sc = m < n_maps ? row_scale(amax[m], fabsf(kappa)) : 1.0f;
        __syncthreads();  // the previous item's reads of sv are done
        for (int i = threadIdx.x; i < L::kPairs * kT; i += blockDim.x) {
            const int p = i / kT, ss = i % kT;
            const int64_t t = t0 + ss;
            float2 q = make_float2(0.0f, 0.0f);
            if (m < n_maps && t < n_t) {
                const int64_t v = (int64_t)p * n_t + t;
                const float2 x = V[m * n_vis + v];
                const float ww = (w ? w[v] : 1.0f) * kappa * sc;
                q = make_float2(x.x * ww, -(x.y * ww));
            }
            sv[p][ss] = q;
        }
        __syncthreads();
        const int64_t s0 = tb * kTPB * L::kSlots;
        const int nsl = (int)(n_slots - s0 < kTPB * L::kSlots ? n_slots - s0 : kTPB * L::kSlots);
        for (int i = threadIdx.x; i < nsl; i += blockDim.x) {
            const int tpl = i / L::kSlots, rem = i % L::kSlots;
            const int j = 8 * (rem / 16) + rem % 8;  // pair (star order)
            const int h = (rem / 8) % 2;             // step 2 tp + h
            float xr = 0.0f, xi = 0.0f;
            if (j < L::kPairs) {
                const float2 q = sv[lex_pair(NS, star_x(NS, j), star_y(NS, j))][2 * tpl + h];
                xr = q.x;
                xi = q.y;
            }
            const _Float16 hr = (_Float16)xr, hi_ = (_Float16)xi;
            const _Float16 lr = (_Float16)(xr - (float)hr), li = (_Float16)(xi - (float)hi_);
            const int64_t o = s0 + i;
            const int64_t tile = o / kStage;
            const int c = (int)(o % kStage);
            _Float16 *t = out + (grp * (n_slots / kStage) + tile) * (4 * kMaps * kStage) + row * kStage +
                          a_slot((int)row, c >> 3) * 8 + (c & 7);
            t[0 * kMaps * kStage] = hr;
            t[1 * kMaps * kStage] = lr;
            t[2 * kMaps * kStage] = hi_;
            t[3 * kMaps * kStage] = li;
        }
        if (tb == 0 && threadIdx.x == 0 && m < n_maps) inv_scale[m] = 1.0f / sc;
    }
}

// Satellite positions scaled by s / lambda (the phase of a pixel is one dot
// product with them); zero-filled samples stay zero.
__global__ void __launch_bounds__(256) scale_pos_kernel(int64_t n, const double4 *__restrict__ pos,
                                                        double sc, double4 *__restrict__ spos) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        const double4 p = pos[i];
        spos[i] = make_double4(p.x * sc, p.y * sc, p.z * sc, 0.0);
    }
}

// Recon block on per-satellite factors: 8 waves (two per SIMD: the factors of
// 16 satellites take 32 registers, and at three waves per SIMD the kernel
// spilled, its reloads waiting behind the stage's loads in flight) x 32 pixels.
constexpr int kWavesRS = 8;
constexpr int kThreadsRS = kWavesRS * 64;

// The weighted-V tile of stage st (4 FP16 planes x 64 maps x 64 slots, 32 KiB,
// stored contiguous and already swizzled) straight from global memory into LDS
// (global_load_lds_dwordx4): 32 blocks of 1 KiB, wave w moving blocks w,
// w + kWavesRS, ...  The two stage buffers are distinct LDS arrays, chosen at
// compile time, so that the compiler can tell the K-steps' reads of one from
// the loads in flight into the other (reads it cannot tell apart would wait
// for every load in flight).
__device__ __forceinline__ void vtile_dma(const _Float16 *__restrict__ Ag, int64_t st, _Float16 *sAbuf) {
    constexpr int kBlocks = 4 * kMaps * (kStage / 8) / 64;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const _Float16 *tile = Ag + st * (4 * kMaps * kStage);  // stored in the LDS layout
#pragma unroll
    for (int r = 0; r < (kBlocks + kWavesRS - 1) / kWavesRS; ++r) {
        const int blk = wave + r * kWavesRS;
        if (blk < kBlocks)
            __builtin_amdgcn_global_load_lds(tile + blk * 512 + lane * 8, (lds_ptr_t)(sAbuf + blk * 512), 16, 0, 0);
    }
}

// A satellite's scaled position, shade cap and position (for the exact
// predicate) at one step (96 bytes): forming the factors reads no global memory,
// whose loads would wait behind the V tile in flight.
struct RecRow {
    double4 spos;
    CapRec cap;
    double4 pos;
};
static_assert(sizeof(RecRow) == 96, "recon row layout");

// Recon on per-satellite factors (NS satellites, the lexicographic pair list).
// Block: 8 waves x 32 pixels; K runs over step pairs, 4 K-steps of 16 slots per
// stage: lanes 0-31 hold step 2 tp, lanes 32-63 step 2 tp + 1, and each lane's 8
// entries of a K-step are the 8 pairs of one star (16 satellites; two stars for
// 8).  At the start of each step pair every lane forms its pixel's factors of
// all NS satellites at its step in registers (FP64 phase and fract, FP32 sin and
// cos, zeroed where that satellite's ray to the pixel is occulted: the exact
// per-pixel rule), so each entry is E'_xy = E'_y conj(E'_x), one complex
// product, and a shaded entry is zero by construction.
template <int NS>
__global__ void __launch_bounds__(kThreadsRS) mfma_recon_sat_kernel(
    GeomDev g, PhaseKnobs k, int32_t n_maps, int64_t n_slots, const _Float16 *__restrict__ A16,
    const float *__restrict__ inv_scale, const double4 *__restrict__ spos, int32_t n_seg,
    float *__restrict__ part) {
    using L = SatRecon<NS>;
    __shared__ _Float16 sA0[4][kMaps][kStage];  // even stages: [re hi, re lo, im hi, im lo][map][slot]
    __shared__ _Float16 sA1[4][kMaps][kStage];  // odd stages
    __shared__ RecRow sRow[2][2 * NS];          // [step pair & 1][h * NS + satellite]
    __shared__ double4 sPix[kWavesRS * kNT];     // the block's pixels (FP64; w: the FP32 copy's
                                                 // z is not needed, x y z FP32 from pix32)
    __shared__ float4 sPix32[kWavesRS * kNT];
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int col = lane & 31, kg = lane >> 5;
    constexpr int kPix = kWavesRS * kNT;  // pixels per block
    const int64_t n_ptiles = (g.n_pad + kPix - 1) / kPix;
    const int64_t n_groups = (n_maps + kMaps - 1) / kMaps;
    const int64_t n_tp = (g.n_t + 1) / 2;
    const int64_t per_seg = (n_tp + n_seg - 1) / n_seg;
    const int64_t n_items = n_ptiles * n_groups * n_seg;
    const double dzo = k.wfac ? 0.0 : 1.0;
    const float neg1 = opaque_neg1();

    // the rows of step pair tp: 2 NS rows of 6 pieces, threads 0 .. 12 NS - 1;
    // this thread's piece: row (satellite rs, step 2 tp + rh), 16-byte part rq
    constexpr int kRowPieces = sizeof(RecRow) / 16;
    const int rr_ = threadIdx.x / kRowPieces, rq = threadIdx.x % kRowPieces;
    const int rh = rr_ / NS, rs = rr_ % NS;
    const int64_t re = (int64_t)rs * g.n_t + rh;
    const uint4 *rsrc = (rq < 2   ? reinterpret_cast<const uint4 *>(spos + re)
                         : rq < 4 ? reinterpret_cast<const uint4 *>(g.cap + re)
                                  : reinterpret_cast<const uint4 *>(g.pos + re)) +
                        (rq & 1);
    auto rows_load = [&](int64_t tp) {
        uint4 x = make_uint4(0u, 0u, 0u, 0u);
        if (threadIdx.x < 2 * NS * kRowPieces && 2 * tp + rh < g.n_t) x = rsrc[2 * (2 * tp)];  // 2 pieces per element
        return x;
    };
    auto rows_store = [&](RecRow *dst, uint4 x) {
        if (threadIdx.x < 2 * NS * kRowPieces) reinterpret_cast<uint4 *>(dst)[threadIdx.x] = x;
    };

    for (int64_t item = blockIdx.x; item < n_items; item += gridDim.x) {
        const int64_t seg = item % n_seg;
        const int64_t rest = item / n_seg;
        const int64_t pt = rest % n_ptiles;
        const int64_t grp = rest / n_ptiles;
        const int64_t pi = pt * kPix + wave * kNT + col;
        const bool live = pt * kPix + wave * kNT < g.n_pad;  // wave-uniform
        // the block's pixels into LDS (read at each step pair's factors: kept in
        // registers they would hold 11 of them across every K-step)
        if (kg == 0) {
            sPix[wave * kNT + col] = live ? g.pix64[pi] : make_double4(0.0, 0.0, 0.0, 0.0);
            sPix32[wave * kNT + col] = live ? g.pix32[pi] : make_float4(0.f, 0.f, 0.f, 0.f);
        }
        const int64_t tpb = seg * per_seg;
        const int64_t tpe = tpb + per_seg < n_tp ? tpb + per_seg : n_tp;
        const _Float16 *Ag = A16 + grp * 4 * kMaps * n_slots;
        floatx16 acc[2];
        acc[0] = floatx16{};
        acc[1] = floatx16{};
        if (tpb < tpe) {
            // Both buffers are free: every wave passed the previous item's last barrier.
            vtile_dma(Ag, tpb * L::kStages, &sA0[0][0][0]);
            rows_store(&sRow[tpb & 1][0], rows_load(tpb));
            wait_vm();
            __syncthreads();
        }
        for (int64_t tp = tpb; tp < tpe; ++tp) {
            // this lane's factors at step 2 tp + kg
            float ec[NS], es[NS];
            uint32_t occm = 0;
            {
                const double4 s64 = sPix[wave * kNT + col];
                const float4 s32 = sPix32[wave * kNT + col];
                const double szd = s64.z - dzo;
                const RecRow *rr = &sRow[tp & 1][kg * NS];
#pragma unroll
                for (int s = 0; s < NS; ++s) {
                    const double4 sp = rr[s].spos;
                    const CapRec cap = rr[s].cap;
                    const double ph = fma(sp.x, s64.x, fma(sp.y, s64.y, sp.z * szd));
                    const float rv = (float)__builtin_amdgcn_fract(ph);
                    const float x = fmaf(s32.z, cap.dz, fmaf(s32.y, cap.dy, s32.x * cap.dx));
                    bool occ = x >= cap.hi;
                    if (x >= cap.lo && !occ) {  // inside the guard band: the exact predicate
                        const int64_t t = 2 * tp + kg;
                        occ = t < g.n_t && occ64m(rr[s].pos, s64.x, s64.y, s64.z, g.limb);
                    }
                    ec[s] = occ ? 0.0f : __builtin_amdgcn_cosf(rv);
                    es[s] = occ ? 0.0f : __builtin_amdgcn_sinf(rv);
                    occm |= (uint32_t)occ << s;
                    // one satellite at a time: hoisting the row reads of later
                    // satellites would hold their 64 bytes per lane in registers
                    if (s % 4 == 3) __builtin_amdgcn_sched_barrier(0);
                }
            }
#pragma unroll
            for (int sq = 0; sq < L::kStages; ++sq) {
                const int64_t st = (tp - tpb) * L::kStages + sq;
                // stage parity = sq & 1 (kStages is even)
                const _Float16(*sAc)[kMaps][kStage] = (sq & 1) ? sA1 : sA0;
                const bool more = tp + 1 < tpe || sq + 1 < L::kStages;
                // the next step pair's rows at its last stage (stored at the end of
                // the stage), then the next stage's V tile into the other buffer
                const bool rows_next = sq + 1 == L::kStages && tp + 1 < tpe;
                const uint4 rn = rows_load(rows_next ? tp + 1 : tp);
                if (more) vtile_dma(Ag, tpb * L::kStages + st + 1, (sq & 1) ? &sA0[0][0][0] : &sA1[0][0][0]);
                {
#pragma unroll
                    for (int kq = 0; kq < 4; ++kq) {
                        const int ks = sq * 4 + kq;
                        if (ks >= L::kKT) continue;
                        float cs[8], sn[8];
#pragma unroll
                        for (int e = 0; e < 8; ++e) {
                            const int j = 8 * ks + e;
                            const int x = star_x(NS, j), y = star_y(NS, j);
                            if (x < 0) {
                                cs[e] = 0.0f;
                                sn[e] = 0.0f;
                            } else {
                                cs[e] = fmaf(ec[y], ec[x], es[y] * es[x]);
                                sn[e] = fmaf(es[y], ec[x], -(ec[y] * es[x]));
                            }
                        }
                        H8 bre_h, bre_l, bim_h, bim_l;
#pragma unroll
                        for (int t2 = 0; t2 < 4; ++t2) {
                            split2m(cs[2 * t2], cs[2 * t2 + 1], neg1, bre_h.u[t2], bre_l.u[t2]);
                            split2m(sn[2 * t2], sn[2 * t2 + 1], neg1, bim_h.u[t2], bim_l.u[t2]);
                        }
#pragma unroll
                        for (int mt = 0; mt < 2; ++mt) {
                            const int row = mt * kMT + col, so = a_slot(row, kq * 2 + kg) * 8;
                            const half8 arh = *reinterpret_cast<const half8 *>(&sAc[0][row][so]);
                            const half8 arl = *reinterpret_cast<const half8 *>(&sAc[1][row][so]);
                            const half8 aih = *reinterpret_cast<const half8 *>(&sAc[2][row][so]);
                            const half8 ail = *reinterpret_cast<const half8 *>(&sAc[3][row][so]);
                            acc[mt] = mfma(arh, bre_h.v, acc[mt]);
                            acc[mt] = mfma(arh, bre_l.v, acc[mt]);
                            acc[mt] = mfma(arl, bre_h.v, acc[mt]);
                            acc[mt] = mfma(aih, bim_h.v, acc[mt]);
                            acc[mt] = mfma(aih, bim_l.v, acc[mt]);
                            acc[mt] = mfma(ail, bim_h.v, acc[mt]);
                        }
                        // K-steps stay apart: interleaving the unrolled K-steps of a stage
                        // (their operand reads hoisted) costs more registers than it hides
                        __builtin_amdgcn_sched_barrier(0);
                    }
                }
                // the other buffer was last read before the previous barrier; sRow[(tp + 1) & 1]
                // was last read when step pair tp - 1 formed its factors
                if (rows_next) rows_store(&sRow[(tp + 1) & 1][0], rn);
                wait_vm();
                __syncthreads();
            }
        }
        if (live) {
#pragma unroll
            for (int mt = 0; mt < 2; ++mt)
#pragma unroll
                for (int r = 0; r < 16; ++r) {
                    const int64_t m = grp * kMaps + mt * kMT + acc_row(r, lane);
                    if (m < n_maps) part[(m * n_seg + seg) * g.n_pad + pi] = acc[mt][r] * inv_scale[m];
                }
        }
    }
}

// ------------------------------------------------------------------ one-map recon as a quadratic form
//
// With one map the recon sum over the pairs of a step is a Hermitian form in
// the step's per-satellite factors e_s = (cos, sin) of 2 pi phi_s (masked:
// zero where satellite s cannot see the pixel):
//   F(p) = sum_t Re sum_b V'_b e_yb conj(e_xb) = sum_t sum_x c_x Zr_x + s_x Zi_x,
//   Z_x = sum_y W_xy e_y,  W_xy = V'_b for pair b = (x, y), V' = V w kappa,
// so each (step, pixel) costs the 16 factors and one 16 x 32 real matrix
// product (Z from [c; s]) on the matrix cores, instead of a product per pair.
// The real form of W: Zr = [Wr, -Wi] [c; s], Zi = [Wi, Wr] [c; s].  K order of
// the product (the B operand a lane forms): positions 8 q .. 8 q + 7 hold c then
// s of satellites 4 q .. 4 q + 3, so lane group q (lanes 16 q .. 16 q + 15) forms
// the factors of satellites 4 q .. 4 q + 3 for pixel lane % 16, and the result
// rows it receives (4 q + r, r < 4) are Z of those same satellites.

// Satellite row of one step for the quadratic-form recon (48 bytes): position
// scaled by s / lambda, and the cap fields the FP32 test reads.
struct QRow {
    double sx, sy, sz;
    float dx, dy, dz, lo, hi, unused;
};
static_assert(sizeof(QRow) == 48, "quadratic-form row layout");
constexpr int kQGroups = 4;  // 16-pixel groups per wave (the W operand is read once per 64 pixels)

// Per step t (one 64-thread block): W from V' (duplicate pairs add), split into
// hi + lo FP16 and stored as the A operands of the 16x16x32 instruction,
// [t][Zr hi, Zr lo, Zi hi, Zi lo][lane] 16 bytes each (lane: row lane % 16, K
// positions 8 (lane / 16) ..), scaled by the power of two that keeps max |V' |
// under 2^14; and the step's 16 satellite rows (absent satellites: a cap that
// covers every pixel, so their factors are zero).
__global__ void __launch_bounds__(64) quad_prep_kernel(GeomDev g, double sc, const float2 *__restrict__ V,
                                                       const float *__restrict__ w, float kappa,
                                                       const unsigned int *__restrict__ amax,
                                                       uint4 *__restrict__ wop, QRow *__restrict__ rows,
                                                       float *__restrict__ inv_scale) {
    __shared__ float2 sW[16][17];
    const int t = blockIdx.x, l = threadIdx.x;
    const float scl = row_scale(amax[0], fabsf(kappa));
    for (int i = l; i < 16 * 17; i += 64) (&sW[0][0])[i] = make_float2(0.0f, 0.0f);
    if (l < 16) {
        QRow q = {0.0, 0.0, 0.0, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
        if (l < g.n_sat) {
            const int64_t r = (int64_t)l * g.n_t + t;
            const double4 p = g.pos[r];
            const CapRec c = g.cap[r];
            q = {p.x * sc, p.y * sc, p.z * sc, c.dx, c.dy, c.dz, c.lo, c.hi, 0.f};
        }
        rows[(int64_t)t * 16 + l] = q;
    }
    __syncthreads();
    for (int b = l; b < g.n_base; b += 64) {
        const int2 pr = g.pairs[b];
        const int64_t v = (int64_t)b * g.n_t + t;
        const float2 x = V[v];
        const float ww = (w ? w[v] : 1.0f) * kappa * scl;
        atomicAdd(&sW[pr.x][pr.y].x, x.x * ww);
        atomicAdd(&sW[pr.x][pr.y].y, x.y * ww);
    }
    __syncthreads();
    const int row = l & 15, q = l >> 4;
    H8 o[4];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        const float2 x = sW[row][4 * q + (j & 3)];
        const float ar = j < 4 ? x.x : -x.y, ai = j < 4 ? x.y : x.x;
        const _Float16 rh = (_Float16)ar, ih = (_Float16)ai;
        o[0].v[j] = rh;
        o[1].v[j] = (_Float16)(ar - (float)rh);
        o[2].v[j] = ih;
        o[3].v[j] = (_Float16)(ai - (float)ih);
    }
#pragma unroll
    for (int pl = 0; pl < 4; ++pl)
        wop[((int64_t)t * 4 + pl) * 64 + l] = make_uint4(o[pl].u[0], o[pl].u[1], o[pl].u[2], o[pl].u[3]);
    if (t == 0 && l == 0) inv_scale[0] = 1.0f / scl;
}

// Quadratic-form recon: each wave holds 64 pixels (4 groups of 16) over one
// segment of steps and needs no LDS and no barrier: per step it reads the W
// operand (4 KiB, shared through the caches by every wave at that step) and
// the rows of its lane groups' satellites, forms each lane's 16 factors
// (4 satellites x 4 pixels; FP64 phase and fract, FP32 sin/cos, the FP32 cap
// test with the exact FP64 predicate inside the guard band), and issues per
// group three products for Zr and three for Zi (hi*hi + hi*lo + lo*hi).  The
// next step's operand and rows load while the current step computes.  part:
// [n_seg][n_pad] partial images (unscaled by inv_scale).
__global__ void __launch_bounds__(256, 2) quad_recon_kernel(GeomDev g, PhaseKnobs k,
                                                            const uint4 *__restrict__ wop,
                                                            const QRow *__restrict__ rows,
                                                            const float *__restrict__ inv_scale,
                                                            int32_t n_seg, float *__restrict__ part) {
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int col = lane & 15, q = lane >> 4;
    const int64_t pt = blockIdx.x / n_seg, seg = blockIdx.x % n_seg;
    const int64_t base = (pt * 4 + wave) * (16 * kQGroups);
    if (base >= g.n_pad) return;  // wave-uniform; the kernel has no barrier
    const int64_t per = ((int64_t)g.n_t + n_seg - 1) / n_seg;
    const int64_t t0 = seg * per;
    const int64_t t1 = t0 + per < g.n_t ? t0 + per : g.n_t;
    const double dzo = k.wfac ? 0.0 : 1.0;
    const float neg1 = opaque_neg1();
    double px[kQGroups], py[kQGroups], pz[kQGroups], pzd[kQGroups];
    float fx[kQGroups], fy[kQGroups], fz[kQGroups], F[kQGroups];
#pragma unroll
    for (int gi = 0; gi < kQGroups; ++gi) {
        const int64_t p = base + 16 * gi + col;
        const double4 s = g.pix64[p];
        const float4 s32 = g.pix32[p];
        px[gi] = s.x;
        py[gi] = s.y;
        pz[gi] = s.z;
        pzd[gi] = s.z - dzo;
        fx[gi] = s32.x;
        fy[gi] = s32.y;
        fz[gi] = s32.z;
        F[gi] = 0.0f;
    }
    struct StepIn {
        uint4 a[4];
        QRow r[4];
    };
    auto load_step = [&](int64_t t) {
        StepIn s;
        const uint4 *wt = wop + t * 4 * 64 + lane;
#pragma unroll
        for (int pl = 0; pl < 4; ++pl) s.a[pl] = wt[pl * 64];
        const QRow *rr = rows + t * 16 + 4 * q;
#pragma unroll
        for (int r = 0; r < 4; ++r) s.r[r] = rr[r];
        return s;
    };
    StepIn cur = load_step(t0 < t1 ? t0 : 0);
    for (int64_t t = t0; t < t1; ++t) {
        const StepIn nxt = load_step(t + 1 < t1 ? t + 1 : t);
        float mc[kQGroups][4], ms[kQGroups][4];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const QRow &R = cur.r[r];
#pragma unroll
            for (int gi = 0; gi < kQGroups; ++gi) {
                const double ph = fma(R.sx, px[gi], fma(R.sy, py[gi], R.sz * pzd[gi]));
                const float rv = (float)__builtin_amdgcn_fract(ph);
                const float x = fmaf(fz[gi], R.dz, fmaf(fy[gi], R.dy, fx[gi] * R.dx));
                bool occ = x >= R.hi;
                if (x >= R.lo && !occ)  // inside the guard band: the exact predicate
                    occ = occ64m(g.pos[(int64_t)(4 * q + r) * g.n_t + t], px[gi], py[gi], pz[gi], g.limb);
                mc[gi][r] = occ ? 0.0f : __builtin_amdgcn_cosf(rv);
                ms[gi][r] = occ ? 0.0f : __builtin_amdgcn_sinf(rv);
            }
        }
        H8 arh, arl, aih, ail;
        arh.v = __builtin_bit_cast(half8, cur.a[0]);
        arl.v = __builtin_bit_cast(half8, cur.a[1]);
        aih.v = __builtin_bit_cast(half8, cur.a[2]);
        ail.v = __builtin_bit_cast(half8, cur.a[3]);
#pragma unroll
        for (int gi = 0; gi < kQGroups; ++gi) {
            H8 bh, bl;
            split2m(mc[gi][0], mc[gi][1], neg1, bh.u[0], bl.u[0]);
            split2m(mc[gi][2], mc[gi][3], neg1, bh.u[1], bl.u[1]);
            split2m(ms[gi][0], ms[gi][1], neg1, bh.u[2], bl.u[2]);
            split2m(ms[gi][2], ms[gi][3], neg1, bh.u[3], bl.u[3]);
            floatx4 zr = mfma16(arh.v, bh.v, floatx4{});
            zr = mfma16(arh.v, bl.v, zr);
            zr = mfma16(arl.v, bh.v, zr);
            floatx4 zi = mfma16(aih.v, bh.v, floatx4{});
            zi = mfma16(aih.v, bl.v, zi);
            zi = mfma16(ail.v, bh.v, zi);
            float f = F[gi];
#pragma unroll
            for (int r = 0; r < 4; ++r) f = fmaf(mc[gi][r], zr[r], fmaf(ms[gi][r], zi[r], f));
            F[gi] = f;
        }
        cur = nxt;
    }
    // each pixel's sum over the four lane groups; lane group gi writes group gi
    float out = 0.0f;
#pragma unroll
    for (int gi = 0; gi < kQGroups; ++gi) {
        float v = F[gi];
        v += __shfl_xor(v, 16, 64);
        v += __shfl_xor(v, 32, 64);
        if (gi == q) out = v;
    }
    part[seg * g.n_pad + base + 16 * q + col] = out * inv_scale[0];
}

// ------------------------------------------------------------------ one-map vis-gen as Gram matrices
//
// With one map, the visibilities of a step are entries of 16 x 16 Gram
// matrices of the step's satellite factors over the pixels (K = pixels):
//   V_xy = sum_clear (I - t_occ) e_y conj(e_x) + t_occ U_xy, conjugated as
// everywhere here (E = cos - i sin).  With A = w m (c, s) (w = (I - t_occ) x
// scale, m the satellite's visibility of the pixel) and B = m (c, s):
//   R = A_c B_c^T + A_s B_s^T,  Q = A_c B_s^T:  Re = R[x][y], Im = Q[x][y] - Q[y][x];
// and the full-sky sums U as in mfma_visgen_sat_kernel's Gram wave (K + K^T and
// P[x][y] - P[y][x]).  Each product is split hi + lo FP16 (three products).
// The factor lane layout of the 16x16x32 instruction is the same for A and B
// (row or column = satellite lane % 16, K = pixels 8 (lane / 16) ..), so every
// lane forms one satellite's factors for 8 pixels of a K-step and holds both
// operands.
constexpr int kGTB = 2;  // steps per block

__global__ void __launch_bounds__(256) qrows_kernel(GeomDev g, double sc, QRow *__restrict__ rows) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)g.n_t * 16) return;
    const int sat = (int)(i % 16);
    const int64_t t = i / 16;
    QRow q = {0.0, 0.0, 0.0, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    if (sat < g.n_sat) {
        const int64_t r = (int64_t)sat * g.n_t + t;
        const double4 p = g.pos[r];
        const CapRec c = g.cap[r];
        q = {p.x * sc, p.y * sc, p.z * sc, c.dx, c.dy, c.dz, c.lo, c.hi, 0.f};
    }
    rows[i] = q;
}

// Block: kGTB steps x one segment of the K-steps (32 pixels each), which the 4
// waves take in turn; the waves' Gram blocks are summed in LDS and stored as
// part[t][seg][R, Q, K, P][16][16].
__global__ void __launch_bounds__(256, 2) gram_visgen_kernel(GeomDev g, PhaseKnobs k,
                                                             const float *__restrict__ I,
                                                             const unsigned int *__restrict__ amax,
                                                             const QRow *__restrict__ rows,
                                                             int32_t n_seg, float *__restrict__ part) {
    __shared__ float sP[4][kGTB * 4 * 256];
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int gs = lane & 15, kg = lane >> 4;
    const int64_t tg = blockIdx.x / n_seg, seg = blockIdx.x % n_seg;
    const int64_t t0 = tg * kGTB;
    const int64_t n_ks = g.n_pad / kKSV;
    const int64_t ks0 = seg * n_ks / n_seg, ks1 = (seg + 1) * n_ks / n_seg;
    const double dzo = k.wfac ? 0.0 : 1.0;
    const float t_occ = k.t_occ;
    const float scl = row_scale(amax[0], 1.0f);
    const float neg1 = opaque_neg1();
    QRow R[kGTB];
    int64_t tq[kGTB];
#pragma unroll
    for (int q = 0; q < kGTB; ++q) {
        tq[q] = t0 + q < g.n_t ? t0 + q : g.n_t - 1;  // a step past the end: its sums are not stored
        R[q] = rows[tq[q] * 16 + gs];
    }
    floatx4 acc[kGTB][4];
#pragma unroll
    for (int q = 0; q < kGTB; ++q)
#pragma unroll
        for (int m = 0; m < 4; ++m) acc[q][m] = floatx4{};
    for (int64_t ks = ks0 + wave; ks < ks1; ks += 4) {
        double px[8], py[8], pz[8];
        float fx[8], fy[8], fz[8], w[8];
        uint32_t padm = 0;
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const int64_t j = ks * kKSV + 8 * kg + i;
            const double4 s = g.pix64[j];
            const float4 s32 = g.pix32[j];
            const int32_t rg = g.ring[j];
            px[i] = s.x;
            py[i] = s.y;
            pz[i] = s.z;
            fx[i] = s32.x;
            fy[i] = s32.y;
            fz[i] = s32.z;
            w[i] = rg >= 0 ? (I[rg] - t_occ) * scl : 0.0f;
            padm |= (uint32_t)(rg < 0) << i;
        }
#pragma unroll
        for (int q = 0; q < kGTB; ++q) {
            float c[8], sn[8], mc[8], ms[8];
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                const double ph = fma(R[q].sx, px[i], fma(R[q].sy, py[i], R[q].sz * (pz[i] - dzo)));
                const float rv = (float)__builtin_amdgcn_fract(ph);
                const float x = fmaf(fz[i], R[q].dz, fmaf(fy[i], R[q].dy, fx[i] * R[q].dx));
                bool occ = x >= R[q].hi;
                if (x >= R[q].lo && !occ)  // inside the guard band: the exact predicate
                    occ = occ64m(g.pos[(int64_t)gs * g.n_t + tq[q]], px[i], py[i], pz[i], g.limb);
                const bool pad = (padm >> i) & 1u;
                c[i] = pad ? 0.0f : __builtin_amdgcn_cosf(rv);
                sn[i] = pad ? 0.0f : __builtin_amdgcn_sinf(rv);
                mc[i] = occ ? 0.0f : c[i];
                ms[i] = occ ? 0.0f : sn[i];
            }
            H8 ahc, alc, ahs, als, bhc, blc, bhs, bls, hc, lc, hs, ls, hc2, hs2;
#pragma unroll
            for (int p = 0; p < 4; ++p) {
                split2m(w[2 * p] * mc[2 * p], w[2 * p + 1] * mc[2 * p + 1], neg1, ahc.u[p], alc.u[p]);
                split2m(w[2 * p] * ms[2 * p], w[2 * p + 1] * ms[2 * p + 1], neg1, ahs.u[p], als.u[p]);
                split2m(mc[2 * p], mc[2 * p + 1], neg1, bhc.u[p], blc.u[p]);
                split2m(ms[2 * p], ms[2 * p + 1], neg1, bhs.u[p], bls.u[p]);
                split2m(c[2 * p], c[2 * p + 1], neg1, hc.u[p], lc.u[p]);
                split2m(sn[2 * p], sn[2 * p + 1], neg1, hs.u[p], ls.u[p]);
            }
            hc2.v = hc.v * (_Float16)0.5f;
            hs2.v = hs.v * (_Float16)0.5f;
            floatx4 a = acc[q][0];
            a = mfma16(ahc.v, bhc.v, a);
            a = mfma16(ahc.v, blc.v, a);
            a = mfma16(alc.v, bhc.v, a);
            a = mfma16(ahs.v, bhs.v, a);
            a = mfma16(ahs.v, bls.v, a);
            a = mfma16(als.v, bhs.v, a);
            acc[q][0] = a;
            a = acc[q][1];
            a = mfma16(ahc.v, bhs.v, a);
            a = mfma16(ahc.v, bls.v, a);
            a = mfma16(alc.v, bhs.v, a);
            acc[q][1] = a;
            a = acc[q][2];
            a = mfma16(hc2.v, hc.v, a);
            a = mfma16(hc.v, lc.v, a);
            a = mfma16(hs2.v, hs.v, a);
            a = mfma16(hs.v, ls.v, a);
            acc[q][2] = a;
            a = acc[q][3];
            a = mfma16(hc.v, hs.v, a);
            a = mfma16(hc.v, ls.v, a);
            a = mfma16(lc.v, hs.v, a);
            acc[q][3] = a;
        }
    }
    // D of 16x16x32: lane holds rows 4 (lane / 16) + r of column lane % 16
#pragma unroll
    for (int q = 0; q < kGTB; ++q)
#pragma unroll
        for (int m = 0; m < 4; ++m)
#pragma unroll
            for (int r = 0; r < 4; ++r) sP[wave][(q * 4 + m) * 256 + (4 * kg + r) * 16 + gs] = acc[q][m][r];
    __syncthreads();
    for (int i = threadIdx.x; i < kGTB * 4 * 256; i += 256) {
        const int q = i / 1024;
        if (t0 + q < g.n_t)
            part[((t0 + q) * n_seg + seg) * 1024 + (i % 1024)] = (sP[0][i] + sP[1][i]) + (sP[2][i] + sP[3][i]);
    }
}

// V of each (pair, step) from the segments' Gram blocks (summed in FP64).
__global__ void __launch_bounds__(256) gram_visgen_finish_kernel(GeomDev g, PhaseKnobs k, int32_t n_seg,
                                                                 const unsigned int *__restrict__ amax,
                                                                 const float *__restrict__ part,
                                                                 float2 *__restrict__ V) {
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= g.n_vis) return;
    const int64_t b = v / g.n_t, t = v % g.n_t;
    const int2 pr = g.pairs[b];
    const int x = pr.x, y = pr.y;
    double r = 0.0, qa = 0.0, qb = 0.0, ka = 0.0, kb = 0.0, pa = 0.0, pb = 0.0;
    for (int s = 0; s < n_seg; ++s) {
        const float *G = part + (t * n_seg + s) * 1024;
        r += G[x * 16 + y];
        qa += G[256 + x * 16 + y];
        qb += G[256 + y * 16 + x];
        ka += G[512 + x * 16 + y];
        kb += G[512 + y * 16 + x];
        pa += G[768 + x * 16 + y];
        pb += G[768 + y * 16 + x];
    }
    const double inv = 1.0 / (double)row_scale(amax[0], 1.0f);
    const double to = k.t_occ;
    V[v] = make_float2((float)(r * inv + to * (ka + kb)), (float)(-((qa - qb) * inv) - to * (pa - pb)));
}

}  // namespace

int64_t gram_visgen_segments(int64_t n_pad, int32_t n_t) {
    const int64_t n_tg = ((int64_t)n_t + kGTB - 1) / kGTB;
    const int64_t most = n_pad / kKSV / 4 > 1 ? n_pad / kKSV / 4 : 1;  // at least one K-step per wave
    const int64_t s = (8192 + n_tg - 1) / n_tg;
    return s < 1 ? 1 : (s > most ? most : s);
}
size_t gram_visgen_part_bytes(int32_t n_t, int64_t n_seg) { return (size_t)n_t * n_seg * 1024 * sizeof(float); }

hipError_t launch_gram_visgen(const GeomDev &g, const PhaseKnobs &k, double inv_lambda_km, int64_t n_pix,
                              const float *I, unsigned int *amax, void *rows, int32_t n_seg, float *part,
                              float2 *V, hipStream_t st) {
    if (g.n_sat > 16 || g.n_pad % kKSV != 0) return hipErrorInvalidValue;
    const double sc = k.s * inv_lambda_km;
    hipLaunchKernelGGL(zero_u32_kernel, dim3(1), dim3(256), 0, st, amax, 1);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(row_absmax_kernel, dim3(64, 1), dim3(256), 0, st, I, n_pix, n_pix, 0, nullptr,
                       amax, k.t_occ);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    const int64_t nr = (int64_t)g.n_t * 16;
    hipLaunchKernelGGL(qrows_kernel, dim3((unsigned)((nr + 255) / 256)), dim3(256), 0, st, g, sc, (QRow *)rows);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    const int64_t n_tg = ((int64_t)g.n_t + kGTB - 1) / kGTB;
    hipLaunchKernelGGL(gram_visgen_kernel, dim3((unsigned)(n_tg * n_seg)), dim3(256), 0, st, g, k, I,
                       (const unsigned int *)amax, (const QRow *)rows, n_seg, part);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    hipLaunchKernelGGL(gram_visgen_finish_kernel, dim3((unsigned)((g.n_vis + 255) / 256)), dim3(256), 0, st,
                       g, k, n_seg, (const unsigned int *)amax, (const float *)part, V);
    return hipGetLastError();
}

int64_t quad_recon_segments(int64_t n_pad, int32_t n_t) {
    const int64_t waves = n_pad / (16 * kQGroups);
    int64_t n_seg = (12288 + waves - 1) / waves;
    return n_seg < 1 ? 1 : (n_seg > n_t ? n_t : n_seg);
}
size_t quad_recon_wop_bytes(int32_t n_t) { return (size_t)n_t * 4 * 64 * sizeof(uint4); }
size_t quad_recon_rows_bytes(int32_t n_t) { return (size_t)n_t * 16 * sizeof(QRow); }

hipError_t launch_quad_recon(const GeomDev &g, const PhaseKnobs &k, double inv_lambda_km,
                             const float2 *V, const float *w, float kappa, unsigned int *amax,
                             float *inv_scale, void *wop, void *rows, int32_t n_seg, float *part,
                             hipStream_t st) {
    if (g.n_sat > 16 || g.n_pad % (16 * kQGroups) != 0) return hipErrorInvalidValue;
    const double sc = k.s * inv_lambda_km;
    hipLaunchKernelGGL(zero_u32_kernel, dim3(1), dim3(256), 0, st, amax, 1);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(row_absmax_kernel, dim3(64, 1), dim3(256), 0, st, (const float *)V, g.n_vis,
                       2 * g.n_vis, 1, w, amax, 0.0f);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    hipLaunchKernelGGL(quad_prep_kernel, dim3(g.n_t), dim3(64), 0, st, g, sc, V, w, kappa, amax,
                       (uint4 *)wop, (QRow *)rows, inv_scale);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    const int64_t ptiles = (g.n_pad / (16 * kQGroups) + 3) / 4;
    hipLaunchKernelGGL(quad_recon_kernel, dim3((unsigned)(ptiles * n_seg)), dim3(256), 0, st, g, k,
                       (const uint4 *)wop, (const QRow *)rows, (const float *)inv_scale, n_seg, part);
    return hipGetLastError();
}

int mfma_vis_pad(int64_t n_vis, int64_t *n_vis_pad) {
    *n_vis_pad = (n_vis + kStage - 1) / kStage * kStage;
    return 0;
}

size_t mfma_a16_bytes_visgen(int32_t n_maps, int64_t n_pad) {
    const int64_t groups = (n_maps + kMaps - 1) / kMaps;
    return (size_t)groups * 2 * kMaps * n_pad * sizeof(_Float16);
}

size_t mfma_a16_bytes_recon(int32_t n_maps, int64_t n_vis_pad) {
    const int64_t groups = (n_maps + kMaps - 1) / kMaps;
    return (size_t)groups * 4 * kMaps * n_vis_pad * sizeof(_Float16);
}

hipError_t launch_mfma_visgen(const GeomDev &g, const PhaseKnobs &k, int32_t n_maps, int64_t n_pix,
                              const float *I, unsigned int *amax, float *inv_scale, void *a16,
                              float2 *V, int grid, hipStream_t st) {
    hipLaunchKernelGGL(zero_u32_kernel, dim3((n_maps + 255) / 256), dim3(256), 0, st, amax, n_maps);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(row_absmax_kernel, dim3(64, n_maps), dim3(256), 0, st, I, n_pix, n_pix, 0,
                       (const float *)nullptr, amax, 0.0f);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    hipLaunchKernelGGL(split_maps_kernel, dim3(2048), dim3(256), 0, st, g, n_maps, n_pix, I, amax,
                       (_Float16 *)a16, inv_scale, 1, 0.0f);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    hipLaunchKernelGGL(mfma_visgen_kernel, dim3(grid), dim3(kThreads), 0, st, g, k, n_maps,
                       (const _Float16 *)a16, inv_scale, V);
    return hipGetLastError();
}

int mfma_sat_tb(int32_t n_sat, int32_t n_base) {
    if (n_sat < 2 || n_sat > 16 || n_base < 1) return 0;
    if (2 * n_base <= kSatSlots && 2 * n_sat <= SatRows<2>::value &&
        (4 * n_base > kSatSlots || 4 * n_sat > SatRows<4>::value))
        return 2;
    if (4 * n_base <= kSatSlots && 4 * n_sat <= SatRows<4>::value) return 4;
    return 0;
}

template <int TB, bool TOCC>
static void launch_sat_visgen_t(const GeomDev &g, const PhaseKnobs &k, double sc, int32_t n_maps,
                                const void *a16, const float *inv_scale, float2 *V, bool star,
                                int32_t pads, int grid, hipStream_t st) {
    if constexpr (TB == 2) {  // the star layout holds two steps
        if (star) {
            hipLaunchKernelGGL((mfma_visgen_sat_kernel<TB, TOCC, true>), dim3(grid), dim3(kThreads), 0, st,
                               g, k, sc, n_maps, (const _Float16 *)a16, inv_scale, V, pads);
            return;
        }
    }
    hipLaunchKernelGGL((mfma_visgen_sat_kernel<TB, TOCC, false>), dim3(grid), dim3(kThreads), 0, st,
                           g, k, sc, n_maps, (const _Float16 *)a16, inv_scale, V, pads);
}

int mfma_visgen_sat_star(int32_t n_sat, int32_t n_base) {
    // two steps of 15 stars of 8 pairs: one star per consumer wave, whose rows
    // (satellite w at both steps: rows w and w + 16) that wave forms itself
    return n_sat == 16 && n_base == 120 && mfma_sat_tb(n_sat, n_base) == 2;
}

hipError_t launch_mfma_visgen_sat(const GeomDev &g, const PhaseKnobs &k, double inv_lambda_km,
                                  int32_t n_maps, int64_t n_pix, const float *I,
                                  unsigned int *amax, float *inv_scale, void *a16, float2 *V,
                                  bool star, int grid, hipStream_t st) {
    const int tb = mfma_sat_tb(g.n_sat, g.n_base);
    if (tb == 0 || (star && !mfma_visgen_sat_star(g.n_sat, g.n_base))) return hipErrorInvalidValue;
    // maps I - t_occ: V = sum_clear (I - t_occ) E + t_occ sum_all E
    const float shift = k.t_occ;
    hipLaunchKernelGGL(zero_u32_kernel, dim3((n_maps + 255) / 256), dim3(256), 0, st, amax, n_maps);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(row_absmax_kernel, dim3(64, n_maps), dim3(256), 0, st, I, n_pix, n_pix, 0,
                       (const float *)nullptr, amax, shift);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    hipLaunchKernelGGL(split_maps_kernel, dim3(2048), dim3(256), 0, st, g, n_maps, n_pix, I, amax,
                       (_Float16 *)a16, inv_scale, 0, shift);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    const double sc = k.s * inv_lambda_km;
    const bool tocc = k.t_occ != 0.0f;
    const int32_t pads = n_pix != g.n_pad;  // padding pixels: the pixel count is not whole chunks
    if (tb == 2) {
        if (tocc) launch_sat_visgen_t<2, true>(g, k, sc, n_maps, a16, inv_scale, V, star, pads, grid, st);
        else launch_sat_visgen_t<2, false>(g, k, sc, n_maps, a16, inv_scale, V, star, pads, grid, st);
    } else {
        if (tocc) launch_sat_visgen_t<4, true>(g, k, sc, n_maps, a16, inv_scale, V, star, pads, grid, st);
        else launch_sat_visgen_t<4, false>(g, k, sc, n_maps, a16, inv_scale, V, star, pads, grid, st);
    }
    return hipGetLastError();
}

int mfma_recon_sat_ns(int32_t n_sat) { return (n_sat == 16 || n_sat == 8) ? n_sat : 0; }

int64_t mfma_recon_sat_slots(int32_t n_sat, int32_t n_t) {
    const int64_t n_tp = ((int64_t)n_t + 1) / 2;
    return n_sat == 16 ? n_tp * SatRecon<16>::kSlots : n_tp * SatRecon<8>::kSlots;
}

template <int NS>
static hipError_t launch_recon_sat_t(const GeomDev &g, const PhaseKnobs &k, double sc, int32_t n_maps,
                                     const float2 *V, const float *w, float kappa, int64_t n_slots,
                                     unsigned int *amax, float *inv_scale, void *a16, double4 *spos,
                                     int32_t n_seg, float *part, int grid, hipStream_t st) {
    const int64_t nst = (int64_t)g.n_sat * g.n_t;
    hipLaunchKernelGGL(scale_pos_kernel, dim3((unsigned)((nst + 255) / 256)), dim3(256), 0, st, nst, g.pos,
                       sc, spos);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(zero_u32_kernel, dim3((n_maps + 255) / 256), dim3(256), 0, st, amax, n_maps);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    hipLaunchKernelGGL(row_absmax_kernel, dim3(64, n_maps), dim3(256), 0, st,
                       (const float *)V, g.n_vis, 2 * g.n_vis, 1, w, amax, 0.0f);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    hipLaunchKernelGGL((split_vis_sat_kernel<NS>), dim3(2048), dim3(256), 0, st, g.n_vis, g.n_t, n_slots,
                       n_maps, V, w, kappa, amax, (_Float16 *)a16, inv_scale);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    hipLaunchKernelGGL((mfma_recon_sat_kernel<NS>), dim3(grid), dim3(kThreadsRS), 0, st, g, k, n_maps,
                       n_slots, (const _Float16 *)a16, inv_scale, (const double4 *)spos, n_seg, part);
    return hipGetLastError();
}

hipError_t launch_mfma_recon_sat(const GeomDev &g, const PhaseKnobs &k, double inv_lambda_km,
                                 int32_t n_maps, const float2 *V, const float *w, float kappa,
                                 int64_t n_slots, unsigned int *amax, float *inv_scale, void *a16,
                                 void *spos, int32_t n_seg, float *part, int grid, hipStream_t st) {
    const double sc = k.s * inv_lambda_km;
    if (g.n_sat == 16)
        return launch_recon_sat_t<16>(g, k, sc, n_maps, V, w, kappa, n_slots, amax, inv_scale, a16,
                                      (double4 *)spos, n_seg, part, grid, st);
    if (g.n_sat == 8)
        return launch_recon_sat_t<8>(g, k, sc, n_maps, V, w, kappa, n_slots, amax, inv_scale, a16,
                                     (double4 *)spos, n_seg, part, grid, st);
    return hipErrorInvalidValue;
}

int mfma_recon_sat_threads() { return kThreadsRS; }
int mfma_recon_sat_pix_per_block() { return kWavesRS * kNT; }

const void *mfma_recon_sat_kernel_ptr(int ns) {
    return ns == 16 ? (const void *)mfma_recon_sat_kernel<16> : (const void *)mfma_recon_sat_kernel<8>;
}

const void *mfma_visgen_sat_kernel_ptr(int tb) {
    return tb == 2 ? (const void *)mfma_visgen_sat_kernel<2, true, true>
                   : (const void *)mfma_visgen_sat_kernel<4, true, false>;
}

hipError_t launch_mfma_recon(const GeomDev &g, const PhaseKnobs &k, int32_t n_maps,
                             const float2 *V, const float *w, float kappa, int64_t n_vis_pad,
                             unsigned int *amax, float *inv_scale, void *a16, int32_t n_seg,
                             float *part, int grid, hipStream_t st) {
    hipLaunchKernelGGL(zero_u32_kernel, dim3((n_maps + 255) / 256), dim3(256), 0, st, amax, n_maps);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(row_absmax_kernel, dim3(64, n_maps), dim3(256), 0, st,
                       (const float *)V, g.n_vis, 2 * g.n_vis, 1, w, amax, 0.0f);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    hipLaunchKernelGGL(split_vis_kernel, dim3(2048), dim3(256), 0, st, g.n_vis, n_vis_pad, n_maps,
                       V, w, kappa, amax, (_Float16 *)a16, inv_scale);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    hipLaunchKernelGGL(mfma_recon_kernel, dim3(grid), dim3(kThreadsR), 0, st, g, k, n_maps,
                       n_vis_pad, (const _Float16 *)a16, inv_scale, n_seg, part);
    return hipGetLastError();
}

const void *mfma_visgen_kernel_ptr() { return (const void *)mfma_visgen_kernel; }
const void *mfma_recon_kernel_ptr() { return (const void *)mfma_recon_kernel; }
int mfma_block_threads() { return kThreads; }
int mfma_recon_block_threads() { return kThreadsR; }
int mfma_vis_per_block() { return kWaves * kNTV; }
int mfma_pix_per_block() { return kWavesR * kNT; }

}  // namespace vrs

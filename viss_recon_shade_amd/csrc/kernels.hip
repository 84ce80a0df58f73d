// gfx950 kernels of the four hot-path stages.
//
// Work layout (all kernels): lanes own pixels of one spatially compact chunk
// (kPixPerLane pixels per lane, kChunk per chunk, coalesced float4 loads);
// the visibility side is wave-uniform and comes from LDS tiles.  For every
// (visibility, chunk) pair the chunk's bounding cap is classified against the
// two satellites' shade caps: wholly outside (no mask work), wholly inside
// (shaded: t_occ in vis-gen, skipped in recon) or straddling (per-pixel FP32
// test; pixels inside the FP32 guard band are decided by the exact FP64
// predicate, the same sequence of fma()s as the FP64 oracle).
//
// Phase: revolutions.  phi = u l + v m + w n + wo with wo = -w (n - 1
// convention) or 0 (trailing exp(-2 pi i s w) folded in); fract() reduces it
// to [0, 1) and v_sin_f32 / v_cos_f32, which take revolutions, finish it.
// The FP64-phase variant forms phi and the fract in FP64 and only the reduced
// value is rounded to FP32.
#include <hip/hip_runtime.h>

#include "kernels.hpp"

namespace vrs {

namespace {

__device__ __forceinline__ float wave_sum(float x) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, kWave);
    return x;
}
__device__ __forceinline__ int wave_sum_i(int x) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, kWave);
    return x;
}
__device__ __forceinline__ uint32_t wave_xor(uint32_t x) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) x ^= (uint32_t)__shfl_xor((int)x, o, kWave);
    return x;
}

// Classify a chunk's bounding cap against one satellite's shade cap:
// kIn = every pixel is inside by more than the guard band, kOut = every pixel
// is outside by more than the guard band, kStraddle otherwise.
__device__ __forceinline__ int classify_cap(const ChunkRec &ch, const CapRec &a) {
    if (ch.flags & kChunkAlwaysStraddle) return kStraddle;
    const float x = fmaf(ch.gz, a.dz, fmaf(ch.gy, a.dy, ch.gx * a.dx));
    const float mg = a.g + kChunkMargin;
    // all pixels within angle alpha - rho of the cap centre: x >= cos(alpha - rho)
    if ((x >= fmaf(a.c, ch.crho, a.sa * ch.srho) + mg) && (a.c < ch.crho)) return kIn;
    // all pixels beyond angle alpha + rho: x < cos(alpha + rho)
    if (x < fmaf(a.c, ch.crho, -a.sa * ch.srho) - mg) return kOut;
    return kStraddle;
}

// The same test as two flags (no class encoding): in = classify_cap == kIn,
// out = classify_cap == kOut; `always` = the chunk's always-straddle flag.
__device__ __forceinline__ void cap_in_out(const ChunkRec &ch, bool always, const CapRec &a,
                                           bool &in, bool &out) {
    const float x = fmaf(ch.gz, a.dz, fmaf(ch.gy, a.dy, ch.gx * a.dx));
    const float mg = a.g + kChunkMargin;
    in = !always && (x >= fmaf(a.c, ch.crho, a.sa * ch.srho) + mg) && (a.c < ch.crho);
    out = !always && !in && (x < fmaf(a.c, ch.crho, -a.sa * ch.srho) - mg);
}

// Classify a chunk against the pair: shaded by either satellite, by neither, or mixed.
__device__ __forceinline__ int classify(const ChunkRec &ch, const CapRec &a, const CapRec &b) {
    const int ca = classify_cap(ch, a), cb = classify_cap(ch, b);
    return (ca == kIn || cb == kIn) ? kIn : ((ca == kOut && cb == kOut) ? kOut : kStraddle);
}

// FP32 cap test against the guard band's edges: shaded when dot >= hi (the
// FP32 dot of two unit vectors is within 4e-7 of the exact one, g >= 4e-6);
// flags pixels with lo <= dot < hi, which the FP64 predicate decides.
__device__ __forceinline__ bool occ32(float sx, float sy, float sz, const CapRec &a, bool &unc) {
    const float x = fmaf(sz, a.dz, fmaf(sy, a.dy, sx * a.dx));
    const bool in_hi = x >= a.hi;
    unc = unc || ((x >= a.lo) && !in_hi);
    return in_hi;
}

// Exact predicate, identical operation sequence to the FP64 oracle:
// d = r.s < 0 and |r|^2 - d^2 <= R^2; near-limb when |r|^2 - d^2 is within
// ((R - e)^2, (R + e)^2), e = 1e-6 R.
__device__ __forceinline__ bool occ64(const double4 r, double sx, double sy, double sz,
                                      const LimbConsts &L, int &near) {
    const double d = fma(r.z, sz, fma(r.y, sy, r.x * sx));
    if (!(d < 0.0)) return false;
    const double q = fma(-d, d, r.w);
    if (q > L.lo2 && q < L.hi2) near = 1;
    return q <= L.R2;
}

__device__ __forceinline__ bool occ_pair64(const double4 *__restrict__ pos,
                                        const double4 *__restrict__ pix64, int64_t pi, int ia,
                                        int ib, LimbConsts L, int *near) {
    const double4 s = pix64[pi];
    int na = 0, nb = 0;
    const bool m = occ64(pos[ia], s.x, s.y, s.z, L, na) | occ64(pos[ib], s.x, s.y, s.z, L, nb);
    *near = na | nb;
    return m;
}

// cos(0.5): bounding caps wider than 0.5 rad are not used for whole-pixel decisions
constexpr float kFineWideCos = 0.87758256f;

// Sub-pixel shading: the pixel's effective intensity for one visibility is the
// mean over its fine cells of (cell shaded ? t_occ : FF) (vissGen10M_shade.cu:121-138).
// The pixel's own bounding cap decides wholly shaded / wholly clear pixels;
// only pixels on a limb test their cells (FP32 with the FP64 guard band).
__device__ __noinline__ float fine_value(const GeomDev &g, int64_t pi, float px, float py, float pz,
                            const CapRec &A, const CapRec &B, int ia, int ib, float t_occ,
                            float cbar) {
    const float2 pc = g.fine.pixcap[pi];
    ChunkRec pr;
    pr.gx = px; pr.gy = py; pr.gz = pz; pr.crho = pc.x; pr.srho = pc.y;
    pr.flags = (pc.x < kFineWideCos) ? kChunkAlwaysStraddle : 0;  // wide pixel: cell by cell
    pr.n_valid = 1; pr.hash_all = 0;
    const int ca = classify_cap(pr, A), cb = classify_cap(pr, B);
    if (ca == kIn || cb == kIn) return t_occ;
    if (ca == kOut && cb == kOut) return cbar;
    const int n = g.fine.count[pi];
    const int s0 = g.fine.start[pi];
    float acc = 0.0f;
    for (int k = 0; k < n; ++k) {
        const int c = g.fine.idx[s0 + k];
        const float4 cc = g.fine.cell32[c];
        bool unc = false;
        bool sh = occ32(cc.x, cc.y, cc.z, A, unc) | occ32(cc.x, cc.y, cc.z, B, unc);
        if (unc) {
            const double4 s64 = g.fine.cell64[c];
            int na = 0, nb = 0;
            sh = occ64(g.pos[ia], s64.x, s64.y, s64.z, g.limb, na) |
                 occ64(g.pos[ib], s64.x, s64.y, s64.z, g.limb, nb);
        }
        acc += sh ? t_occ : cc.w;
    }
    return n > 0 ? acc / (float)n : 0.0f;
}

__device__ __forceinline__ float fract_rev(float x) { return __builtin_amdgcn_fractf(x); }
__device__ __forceinline__ float fract_rev(double x) { return (float)__builtin_amdgcn_fract(x); }

// ------------------------------------------------------------------ stage 1
__global__ void __launch_bounds__(256) orbit_kernel(OrbitArgs a, double4 *__restrict__ pos,
                                                    CapRec *__restrict__ cap) {
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (int64_t)a.n_sat * a.n_t) return;
    const int k = (int)(idx / a.n_t);
    const int j = (int)(idx % a.n_t);
    const int64_t ts = a.t_idx ? a.t_idx[j] : (int64_t)j;
    if (ts < 0) {
        // a sample past the end of the propagated span: the reference zero-fills
        // it (uvwPosition, MIncline_generate_uvw_float.cu:134-139); it shades nothing
        pos[idx] = make_double4(0.0, 0.0, 0.0, 0.0);
        cap[idx] = {0.f, 0.f, 0.f, 2.0f, 0.f, kGuardMin, 2.0f - kGuardMin, 2.0f + kGuardMin};
        return;
    }
    // circular orbit: E = M = nu (calculateOrbit, MIncline_generate_uvw_float.cu:93-104)
    const double w = a.m0[k] + a.n_mot * ((double)ts * a.dt);
    const double cw = cos(w), sw = sin(w), cO = cos(a.raan[k]), sO = sin(a.raan[k]);
    const double x = a.a * (cw * cO - sw * a.cos_i * sO);
    const double y = a.a * (cw * sO + sw * a.cos_i * cO);
    const double z = a.a * (sw * a.sin_i);
    const double r2 = fma(z, z, fma(y, y, x * x));
    pos[idx] = make_double4(x, y, z, r2);
    const double rn = sqrt(r2);
    const double R = a.r_moon;
    const double c = sqrt(fmax(r2 - R * R, 0.0)) / rn;
    CapRec cr;
    cr.dx = (float)(-x / rn);
    cr.dy = (float)(-y / rn);
    cr.dz = (float)(-z / rn);
    cr.c = (float)c;
    cr.sa = (float)(R / rn);
    // near-limb band |delta| < 1e-6 R maps to |dot - c| < 1e-6 R^2 / (|r|^2 c)
    const double band = kNearLimbRel * R * R / (r2 * fmax(c, 1e-12));
    cr.g = (float)fmax((double)kGuardMin, 2.0 * band + 1e-6);
    cr.lo = cr.c - cr.g;
    cr.hi = cr.c + cr.g;
    cap[idx] = cr;
}

__global__ void __launch_bounds__(256) uvw_kernel(GeomDev g, double inv_lambda_km, int ceil_uvw,
                                                  double4 *__restrict__ uvw64,
                                                  int2 *__restrict__ vis_cap) {
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= g.n_vis) return;
    const int b = (int)(v / g.n_t), t = (int)(v % g.n_t);
    const int2 pr = g.pairs[b];
    const double4 ra = g.pos[(int64_t)pr.x * g.n_t + t], rb = g.pos[(int64_t)pr.y * g.n_t + t];
    // computeUVW (MIncline_generate_uvw_float.cu:156-158); its ceil only with ceil_uvw
    double u = (rb.x - ra.x) * inv_lambda_km, vv = (rb.y - ra.y) * inv_lambda_km,
           w = (rb.z - ra.z) * inv_lambda_km;
    if (ceil_uvw) {
        u = ceil(u);
        vv = ceil(vv);
        w = ceil(w);
    }
    uvw64[v] = make_double4(u, vv, w, 0.0);
    vis_cap[v] = make_int2(pr.x * g.n_t + t, pr.y * g.n_t + t);
}

// ------------------------------------------------------------------ stage 2
// One wave per visibility; lanes classify 64 chunks at a time, then sweep the
// straddling (and, for the bitmap, the shaded) chunks pixel by pixel.
__global__ void __launch_bounds__(kBlock) shade_kernel(GeomDev g, int32_t *__restrict__ count,
                                                       uint32_t *__restrict__ hash,
                                                       uint32_t *__restrict__ bits,
                                                       int64_t n_words,
                                                       unsigned long long *__restrict__ tot) {
    const int lane = threadIdx.x & (kWave - 1);
    const int64_t wave0 = (int64_t)blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6);
    const int64_t nwaves = (int64_t)gridDim.x * kWavesPerBlock;
    unsigned long long t_shaded = 0, t_near = 0, t_f64 = 0;
    for (int64_t v = wave0; v < g.n_vis; v += nwaves) {
        const int2 vc = g.vis_cap[v];
        const int ia = vc.x, ib = vc.y;
        const CapRec A = g.cap[ia], B = g.cap[ib];
        int cnt = 0, near = 0, f64 = 0;
        uint32_t h = 0;
        for (int c0 = 0; c0 < g.n_chunk; c0 += kWave) {
            const int c = c0 + lane;
            int cls = kOut;
            bool limb = false;  // the chunk straddles either satellite's limb
            if (c < g.n_chunk) {
                const ChunkRec ch = g.chunk[c];
                const int ca = classify_cap(ch, A), cb = classify_cap(ch, B);
                cls = (ca == kIn || cb == kIn) ? kIn : ((ca == kOut && cb == kOut) ? kOut : kStraddle);
                limb = (ca == kStraddle) || (cb == kStraddle);
                if (cls == kIn) {
                    cnt += ch.n_valid;
                    h ^= ch.hash_all;
                }
            }
            uint64_t todo = __ballot(limb);
            if (bits) todo |= __ballot(cls == kIn);
            while (todo) {
                const int cs = __ffsll((unsigned long long)todo) - 1;
                todo &= todo - 1;
                const int cc = c0 + cs;
                const bool whole = __shfl(cls, cs, kWave) == kIn;
                const bool test = __shfl((int)limb, cs, kWave) != 0;
#pragma unroll
                for (int j = 0; j < kPixPerLane; ++j) {
                    const int64_t pi = (int64_t)cc * kChunk + j * kWave + lane;
                    const int rp = g.ring[pi];
                    if (rp < 0) continue;
                    bool m = whole;  // wholly shaded chunks were counted above
                    if (test) {
                        // near-limb pixels are counted for either satellite's limb,
                        // also inside chunks that the other satellite shades wholly
                        const float4 s = g.pix32[pi];
                        bool unc = false;
                        bool m32 = occ32(s.x, s.y, s.z, A, unc) | occ32(s.x, s.y, s.z, B, unc);
                        if (unc) {
                            int nl = 0;
                            m32 = occ_pair64(g.pos, g.pix64, pi, ia, ib, g.limb, &nl);
                            near += nl;
                            ++f64;
                        }
                        if (!whole) {
                            m = m32;
                            if (m) {
                                ++cnt;
                                h ^= pix_hash_u64((uint64_t)rp);
                            }
                        }
                    }
                    if (m && bits)
                        atomicOr(&bits[v * n_words + (rp >> 5)], 1u << (rp & 31));
                }
            }
        }
        cnt = wave_sum_i(cnt);
        near = wave_sum_i(near);
        f64 = wave_sum_i(f64);
        h = wave_xor(h);
        if (lane == 0) {
            if (count) count[v] = cnt;
            if (hash) hash[v] = h;
            t_shaded += (unsigned long long)cnt;
            t_near += (unsigned long long)near;
            t_f64 += (unsigned long long)f64;
        }
    }
    if (lane == 0 && tot) {
        atomicAdd(&tot[0], t_shaded);
        atomicAdd(&tot[1], t_near);
        atomicAdd(&tot[2], t_f64);
    }
}

__global__ void __launch_bounds__(256) to_chunk_order_kernel(GeomDev g, int32_t n_maps,
                                                             int64_t n_pix,
                                                             const float *__restrict__ I,
                                                             float *__restrict__ Ich) {
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (int64_t)n_maps * g.n_pad) return;
    const int64_t m = idx / g.n_pad, pi = idx % g.n_pad;
    const int r = g.ring[pi];
    Ich[idx] = (r >= 0) ? I[m * n_pix + r] : 0.0f;
}

// ------------------------------------------------------------------ stage 3
// List mode (ragged batches): g.vis_sat != nullptr.  Visibility v is then the
// v-th entry of a concatenation of per-map (per-channel) sample lists, each
// padded to a multiple of kListPad entries; padding entries have vis_sat.x < 0.
// uvw = (r_b - r_a) * inv_lambda[map], formed exactly as stage 1 forms it.
constexpr int kListPad = 32;

// FP32 cap test that ORs the wave's guard-band pixels into a lane mask (kept
// in scalar registers) instead of a per-lane flag.
__device__ __forceinline__ bool occ32w(float sx, float sy, float sz, const CapRec &a,
                                       uint64_t &unc_w) {
    const float x = fmaf(sz, a.dz, fmaf(sy, a.dy, sx * a.dx));
    const uint64_t hi_w = __ballot(x >= a.hi);
    unc_w |= __ballot(x >= a.lo) & ~hi_w;
    return x >= a.hi;
}

// One straddling pair over this lane's pixels: the FP32 test of cap A (and of
// cap B when TWO) pixel by pixel, T_occ where shaded, I elsewhere.  Returns the
// lanes with a pixel inside a tested cap's guard band (called by the whole wave).
template <bool TWO, typename T>
__device__ __forceinline__ uint64_t straddle_terms(const T (&l)[kPixPerLane],
                                                   const T (&m)[kPixPerLane],
                                                   const T (&n)[kPixPerLane],
                                                   const float (&I)[kPixPerLane],
                                                   const CapRec &A, const CapRec &B, T u, T vv,
                                                   T w, T wo, float t_occ, float &are,
                                                   float &aim) {
    uint64_t unc_w = 0;
#pragma unroll
    for (int j = 0; j < kPixPerLane; ++j) {
        const float sx = (float)l[j], sy = (float)m[j], sz = (float)n[j];
        bool sh = occ32w(sx, sy, sz, A, unc_w);
        if (TWO) sh = occ32w(sx, sy, sz, B, unc_w) | sh;
        const float ie = sh ? t_occ : I[j];
        const T ph = fma(u, l[j], fma(vv, m[j], fma(w, n[j], wo)));
        const float r = fract_rev(ph);
        are = fmaf(ie, __builtin_amdgcn_cosf(r), are);
        aim = fmaf(ie, __builtin_amdgcn_sinf(r), aim);
    }
    return unc_w;
}

template <bool F64> struct PhaseT { using T = float; };
template <> struct PhaseT<true> { using T = double; };

template <bool F64> struct VisEnt {
    typename PhaseT<F64>::T u, v, w, wo;
    CapRec a, b;
    int32_t ia, ib, pad0, pad1;
};

// (ia, ib, uvw) of visibility v; returns false for padding / tail entries.
__device__ __forceinline__ bool vis_geometry(const GeomDev &g, int64_t v, int64_t map, int &ia,
                                             int &ib, double &u, double &vv, double &w) {
    if (v >= g.n_vis) return false;
    if (g.vis_sat) {
        const int2 s = g.vis_sat[v];
        if (s.x < 0) return false;
        ia = s.x;
        ib = s.y;
        const double4 ra = g.pos[ia], rb = g.pos[ib];
        const double il = g.inv_lambda[map];
        u = (rb.x - ra.x) * il;
        vv = (rb.y - ra.y) * il;
        w = (rb.z - ra.z) * il;
    } else {
        const int2 vc = g.vis_cap[v];
        ia = vc.x;
        ib = vc.y;
        const double4 q = g.uvw64[v];
        u = q.x; vv = q.y; w = q.z;
    }
    return true;
}

template <bool F64>
__device__ __forceinline__ void fill_vis_ent(const GeomDev &g, const PhaseKnobs &k, int64_t v,
                                             int64_t map, VisEnt<F64> &e) {
    using T = typename PhaseT<F64>::T;
    int ia = 0, ib = 0;
    double u, vv, w;
    if (vis_geometry(g, v, map, ia, ib, u, vv, w)) {
        e.ia = ia;
        e.ib = ib;
        e.a = g.cap[ia];
        e.b = g.cap[ib];
        if (F64) {
            e.u = (T)(u * k.s);
            e.v = (T)(vv * k.s);
            e.w = (T)(w * k.s);
        } else {
            // the FP32 path rounds uvw to FP32 first
            const float s = (float)k.s;
            e.u = (T)((float)u * s);
            e.v = (T)((float)vv * s);
            e.w = (T)((float)w * s);
        }
        e.wo = k.wfac ? (T)0 : -e.w;
    } else {
        // tail / padding: no shading (c = 2 is never reached), zero phase; never written out
        e.ia = e.ib = 0;
        CapRec none = {0.f, 0.f, 0.f, 2.0f, 0.f, 0.f, 2.0f, 2.0f};
        e.a = none;
        e.b = none;
        e.u = e.v = e.w = e.wo = (T)0;
    }
}

// Four waves per SIMD: the FP64-phase tiles then spill a few registers, which
// costs less than a quarter of the waves.  The FP32 32-visibility tile keeps
// its registers (three waves) rather than spill at the same speed.
template <bool F64, int VT, bool FINE = false>
__global__ void __launch_bounds__(kBlock, (!F64 && VT == 32) ? 3 : 4) visgen_kernel(GeomDev g, PhaseKnobs k, int32_t n_maps,
                                                        const float *__restrict__ Ich,
                                                        float2 *__restrict__ V) {
    using T = typename PhaseT<F64>::T;
    static_assert(VT == 8 || VT == 16 || VT == 32, "tile of 8, 16 or 32 visibilities");
    static_assert(kListPad % VT == 0, "list padding must be whole tiles");
    __shared__ VisEnt<F64> ents[VT];
    __shared__ float red[kWavesPerBlock][VT][2];
    __shared__ float corr[kWavesPerBlock][VT][2];  // straddling pairs, per wave
    constexpr uint32_t kTileMask = (VT == 32) ? 0xffffffffu : ((1u << VT) - 1u);
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // wave-uniform
    const bool list = g.vis_sat != nullptr;
    const int64_t n_vtiles = (g.n_vis + VT - 1) / VT;
    const int64_t n_tiles = list ? n_vtiles : n_vtiles * n_maps;
    const float t_occ = k.t_occ;

    for (int64_t tile = blockIdx.x; tile < n_tiles; tile += gridDim.x) {
        const int64_t map = list ? g.tile_map[(tile * VT) / kListPad] : tile / n_vtiles;
        const int64_t v0 = (list ? tile : tile % n_vtiles) * VT;
        const float *__restrict__ Im = Ich + map * g.n_pad;
        if (threadIdx.x < VT) fill_vis_ent<F64>(g, k, v0 + threadIdx.x, map, ents[threadIdx.x]);
        if (threadIdx.x < kWavesPerBlock * VT * 2) (&corr[0][0][0])[threadIdx.x] = 0.0f;
        __syncthreads();

        float are[VT], aim[VT];
#pragma unroll
        for (int q = 0; q < VT; ++q) are[q] = aim[q] = 0.0f;

        for (int c = wave; c < g.n_chunk; c += kWavesPerBlock) {
            // The tile entries are invariant over this loop; keep the compiler
            // from hoisting all of them into registers.
            asm volatile("" ::: "memory");
            const ChunkRec ch = g.chunk[c];
            T l[kPixPerLane], m[kPixPerLane], n[kPixPerLane];
            float I[kPixPerLane];
#pragma unroll
            for (int j = 0; j < kPixPerLane; ++j) {
                const int64_t pi = (int64_t)c * kChunk + j * kWave + lane;
                if (F64) {
                    const double4 p = g.pix64[pi];
                    l[j] = (T)p.x; m[j] = (T)p.y; n[j] = (T)p.z;
                } else {
                    const float4 p = g.pix32[pi];
                    l[j] = (T)p.x; m[j] = (T)p.y; n[j] = (T)p.z;
                }
                I[j] = Im[pi];
            }
            if constexpr (!FINE) {
                // Lane-parallel classification of the tile's pairs against this
                // chunk, per cap: the results become wave-uniform bit masks.
                // sa_m / sb_m: pairs whose cap A / B must be tested pixel by pixel.
                uint32_t in_m, out_m, sa_m, sb_m;
                {
                    // lane q + VT * s tests cap s (0 = A, 1 = B) of pair q
                    const int q = lane & (VT - 1), s = (lane / VT) & 1;
                    bool cin, cout;
                    cap_in_out(ch, (ch.flags & kChunkAlwaysStraddle) != 0, (&ents[q].a)[s], cin,
                               cout);
                    const uint64_t bin = __ballot(cin), bout = __ballot(cout);
                    const uint32_t ain = (uint32_t)bin & kTileMask;
                    const uint32_t bin_ = (uint32_t)(bin >> VT) & kTileMask;
                    const uint32_t aout = (uint32_t)bout & kTileMask;
                    const uint32_t bout_ = (uint32_t)(bout >> VT) & kTileMask;
                    in_m = ain | bin_;
                    out_m = aout & bout_;
                    sa_m = ~(in_m | aout) & kTileMask;
                    sb_m = ~(in_m | bout_) & kTileMask;
                }
                uint32_t unc_m = 0;  // straddling pairs with a pixel inside the FP32 guard band
#pragma unroll
                for (int q = 0; q < VT; ++q) {
                    const T u = ents[q].u, vv = ents[q].v, w = ents[q].w, wo = ents[q].wo;
                    if ((in_m >> q) & 1u) {          // wholly shaded: T_occ everywhere
#pragma unroll
                        for (int j = 0; j < kPixPerLane; ++j) {
                            const T ph = fma(u, l[j], fma(vv, m[j], fma(w, n[j], wo)));
                            const float r = fract_rev(ph);
                            are[q] = fmaf(t_occ, __builtin_amdgcn_cosf(r), are[q]);
                            aim[q] = fmaf(t_occ, __builtin_amdgcn_sinf(r), aim[q]);
                        }
                    } else if ((out_m >> q) & 1u) {  // wholly clear
#pragma unroll
                        for (int j = 0; j < kPixPerLane; ++j) {
                            const T ph = fma(u, l[j], fma(vv, m[j], fma(w, n[j], wo)));
                            const float r = fract_rev(ph);
                            are[q] = fmaf(I[j], __builtin_amdgcn_cosf(r), are[q]);
                            aim[q] = fmaf(I[j], __builtin_amdgcn_sinf(r), aim[q]);
                        }
                    } else {                         // straddling: FP32 cap test per pixel
                        const bool ta = (sa_m >> q) & 1u, tb = (sb_m >> q) & 1u;
                        const uint64_t unc_w =
                            (ta && tb)
                                ? straddle_terms<true>(l, m, n, I, ents[q].a, ents[q].b, u, vv, w,
                                                       wo, t_occ, are[q], aim[q])
                                // one cap straddles; the other is wholly clear of the chunk
                                : straddle_terms<false>(l, m, n, I, ta ? ents[q].a : ents[q].b,
                                                        ents[q].b, u, vv, w, wo, t_occ, are[q],
                                                        aim[q]);
                        if (unc_w) unc_m |= 1u << q;
                    }
                }
                // Pixels inside the guard band: replace the FP32 decision by the
                // exact FP64 predicate; the difference goes to the wave's LDS
                // accumulator (only this wave touches corr[wave]: deterministic).
                while (unc_m) {
                    const int q = __ffs(unc_m) - 1;
                    unc_m &= unc_m - 1;
                    const bool ta = (sa_m >> q) & 1u, tb = (sb_m >> q) & 1u;
                    const CapRec A = ents[q].a, B = ents[q].b;
                    const T u = ents[q].u, vv = ents[q].v, w = ents[q].w, wo = ents[q].wo;
                    float dre = 0.0f, dim = 0.0f;
#pragma unroll
                    for (int j = 0; j < kPixPerLane; ++j) {
                        const float sx = (float)l[j], sy = (float)m[j], sz = (float)n[j];
                        bool unc = false, sh = false;
                        if (ta) sh = occ32(sx, sy, sz, A, unc);
                        if (tb) sh = occ32(sx, sy, sz, B, unc) || sh;
                        if (unc) {
                            int nl;
                            const int64_t pi = (int64_t)c * kChunk + j * kWave + lane;
                            const bool ex = occ_pair64(g.pos, g.pix64, pi, ents[q].ia, ents[q].ib,
                                                       g.limb, &nl);
                            if (ex != sh) {
                                const float d = ex ? (t_occ - I[j]) : (I[j] - t_occ);
                                const T ph = fma(u, l[j], fma(vv, m[j], fma(w, n[j], wo)));
                                const float r = fract_rev(ph);
                                dre = fmaf(d, __builtin_amdgcn_cosf(r), dre);
                                dim = fmaf(d, __builtin_amdgcn_sinf(r), dim);
                            }
                        }
                    }
                    dre = wave_sum(dre);
                    dim = wave_sum(dim);
                    if (lane == 0) {
                        corr[wave][q][0] += dre;
                        corr[wave][q][1] += dim;
                    }
                }
            } else {
                // Fine-cell sky: chunk bounds cover every cell; straddling pairs
                // go pixel by pixel through the pixel's own bounding cap and cells.
                uint32_t in_m, out_m;
                {
                    const int q = lane & (VT - 1);
                    const int cls = classify(ch, ents[q].a, ents[q].b);
                    in_m = (uint32_t)__ballot(cls == kIn) & kTileMask;
                    out_m = (uint32_t)__ballot(cls == kOut) & kTileMask;
                }
                const uint32_t str_m = ~(in_m | out_m) & kTileMask;
#pragma unroll
                for (int q = 0; q < VT; ++q) {
                    if ((str_m >> q) & 1u) continue;
                    const bool sh = (in_m >> q) & 1u;
                    const T u = ents[q].u, vv = ents[q].v, w = ents[q].w, wo = ents[q].wo;
#pragma unroll
                    for (int j = 0; j < kPixPerLane; ++j) {
                        const float ie = sh ? t_occ : I[j];
                        const T ph = fma(u, l[j], fma(vv, m[j], fma(w, n[j], wo)));
                        const float r = fract_rev(ph);
                        are[q] = fmaf(ie, __builtin_amdgcn_cosf(r), are[q]);
                        aim[q] = fmaf(ie, __builtin_amdgcn_sinf(r), aim[q]);
                    }
                }
                uint32_t todo = str_m;
                while (todo) {
                    const int q = __ffs(todo) - 1;
                    todo &= todo - 1;
                    const CapRec A = ents[q].a, B = ents[q].b;
                    const T u = ents[q].u, vv = ents[q].v, w = ents[q].w, wo = ents[q].wo;
                    float dre = 0.0f, dim = 0.0f;
#pragma unroll
                    for (int j = 0; j < kPixPerLane; ++j) {
                        const float sx = (float)l[j], sy = (float)m[j], sz = (float)n[j];
                        const int64_t pi = (int64_t)c * kChunk + j * kWave + lane;
                        const float ie = fine_value(g, pi, sx, sy, sz, A, B, ents[q].ia,
                                                    ents[q].ib, t_occ, I[j]);
                        const T ph = fma(u, l[j], fma(vv, m[j], fma(w, n[j], wo)));
                        const float r = fract_rev(ph);
                        dre = fmaf(ie, __builtin_amdgcn_cosf(r), dre);
                        dim = fmaf(ie, __builtin_amdgcn_sinf(r), dim);
                    }
                    dre = wave_sum(dre);
                    dim = wave_sum(dim);
                    if (lane == 0) {
                        corr[wave][q][0] += dre;
                        corr[wave][q][1] += dim;
                    }
                }
            }
        }
#pragma unroll
        for (int q = 0; q < VT; ++q) {
            const float sr = wave_sum(are[q]);
            const float si = wave_sum(aim[q]);
            if (lane == 0) {
                red[wave][q][0] = sr;
                red[wave][q][1] = si;
            }
        }
        __syncthreads();
        if (threadIdx.x < VT) {
            const int q = threadIdx.x;
            const int64_t v = v0 + q;
            const bool real = list ? (v < g.n_vis && g.vis_sat[v].x >= 0) : (v < g.n_vis);
            if (real) {
                float sr = 0.0f, si = 0.0f;
#pragma unroll
                for (int w = 0; w < kWavesPerBlock; ++w) {
                    sr += red[w][q][0] + corr[w][q][0];
                    si += red[w][q][1] + corr[w][q][1];
                }
                // exp(-i theta): V = sum I (cos - i sin)
                V[(list ? 0 : map * g.n_vis) + v] = make_float2(sr, -si);
            }
        }
        __syncthreads();
    }
}

// ------------------------------------------------------- antipodal pairs
// On the antipodal layout (g.n_half > 0) pixel p of chunk k and its antipode
// p' = -p, held in the same slot of chunk k + n_half, share one phase:
// phi(p') = u.p' + wo = 2 wo - phi(p), so E(p') = exp(-2 pi i phi(p')) =
// conj(E(p)) exp(-4 pi i wo).  One fract and one v_sin / v_cos serve both
// terms.  With E(p) = cos - i sin and C the pixel's intensity (t_occ where
// shaded),
//   V = (A1 - i B1) + exp(-4 pi i wo) (A2 + i B2),
//   A1 = sum C_p cos, B1 = sum C_p sin, A2 = sum C_p' cos, B2 = sum C_p' sin.
// Every shade decision is still made per pixel: the FP32 test of p' uses -p
// (its dot with a cap axis is exactly minus p's), the FP64 predicate of a
// guard-band pixel reads the antipode's own FP64 direction.

// Classification of the tile's VT pairs against one chunk (lane q + VT s tests
// cap s of pair q): wholly shaded, wholly clear, cap A / cap B straddling.
template <bool F64, int VT>
__device__ __forceinline__ void tile_masks(const ChunkRec &ch, const VisEnt<F64> *ents, int lane,
                                           uint32_t &in_m, uint32_t &out_m, uint32_t &sa_m,
                                           uint32_t &sb_m) {
    constexpr uint32_t kTileMask = (VT == 32) ? 0xffffffffu : ((1u << VT) - 1u);
    const int q = lane & (VT - 1), s = (lane / VT) & 1;
    bool cin, cout;
    cap_in_out(ch, (ch.flags & kChunkAlwaysStraddle) != 0, (&ents[q].a)[s], cin, cout);
    const uint64_t bin = __ballot(cin), bout = __ballot(cout);
    const uint32_t ain = (uint32_t)bin & kTileMask;
    const uint32_t bin_ = (uint32_t)(bin >> VT) & kTileMask;
    const uint32_t aout = (uint32_t)bout & kTileMask;
    const uint32_t bout_ = (uint32_t)(bout >> VT) & kTileMask;
    in_m = ain | bin_;
    out_m = aout & bout_;
    sa_m = ~(in_m | aout) & kTileMask;
    sb_m = ~(in_m | bout_) & kTileMask;
}

// acc += a * b in the accumulator's own register (v_fmac_f32): the paired
// vis-gen updates its accumulators from three code paths per step, and the
// tied operand keeps the register allocator from giving each path its own
// copy (which costs register moves at every join).
__device__ __forceinline__ void acc_fmac(float &acc, float a, float b) {
    asm("v_fmac_f32 %0, %1, %2" : "+v"(acc) : "v"(a), "v"(b));
}

// One pair's terms over this lane's pixel pairs: one phase, fract, v_sin and
// v_cos per pixel, accumulated with the pixel's and the antipode's intensity.
template <bool F64>
__device__ __forceinline__ void pair_terms(typename PhaseT<F64>::T u, typename PhaseT<F64>::T vv,
                                           typename PhaseT<F64>::T w, typename PhaseT<F64>::T wo,
                                           const typename PhaseT<F64>::T (&l)[kPixPerLane],
                                           const typename PhaseT<F64>::T (&m)[kPixPerLane],
                                           const typename PhaseT<F64>::T (&n)[kPixPerLane],
                                           const float (&e1)[kPixPerLane],
                                           const float (&e2)[kPixPerLane], float &a1, float &b1,
                                           float &a2, float &b2) {
    using T = typename PhaseT<F64>::T;
#pragma unroll
    for (int j = 0; j < kPixPerLane; ++j) {
        const T ph = fma(u, l[j], fma(vv, m[j], fma(w, n[j], wo)));
        const float r = fract_rev(ph);
        const float cs = __builtin_amdgcn_cosf(r), sn = __builtin_amdgcn_sinf(r);
        acc_fmac(a1, e1[j], cs);
        acc_fmac(b1, e1[j], sn);
        acc_fmac(a2, e2[j], cs);
        acc_fmac(b2, e2[j], sn);
    }
}

// FP32 tests of one cap on this lane's pixel pairs: x = dot(p, dir) for each
// pixel p and -x for its antipode (negation is exact, and -p rounded to FP32
// is the antipode's FP32 direction up to one FP32 ulp, far inside the guard
// band).  t1 / t2: test the pixels / the antipodes (wave-uniform).  Sets the
// shade flags and ORs in the wave's lanes with a pixel inside the guard band.
template <typename T>
__device__ __forceinline__ void pair_cap_tests(const T (&l)[kPixPerLane], const T (&m)[kPixPerLane],
                                               const T (&n)[kPixPerLane], const CapRec &a, bool t1,
                                               bool t2, bool (&sh1)[kPixPerLane],
                                               bool (&sh2)[kPixPerLane], uint64_t &unc1,
                                               uint64_t &unc2) {
    float x[kPixPerLane];
#pragma unroll
    for (int j = 0; j < kPixPerLane; ++j)
        x[j] = fmaf((float)n[j], a.dz, fmaf((float)m[j], a.dy, (float)l[j] * a.dx));
    if (t1) {
#pragma unroll
        for (int j = 0; j < kPixPerLane; ++j) {
            const bool hi = x[j] >= a.hi;
            unc1 |= __ballot(x[j] >= a.lo) & ~__ballot(hi);
            sh1[j] = sh1[j] || hi;
        }
    }
    if (t2) {
#pragma unroll
        for (int j = 0; j < kPixPerLane; ++j) {
            const float y = -x[j];
            const bool hi = y >= a.hi;
            unc2 |= __ballot(y >= a.lo) & ~__ballot(hi);
            sh2[j] = sh2[j] || hi;
        }
    }
}

// Guard-band pixels of one straddling pair in chunk `base` (neg: the antipodal
// chunk, whose FP32 test used the negated FP32 direction of the pixel at
// base - off2): the exact FP64 predicate on the pixel's own direction replaces
// the FP32 decision; returns the intensity difference times the pixel's own E.
template <bool F64>
__device__ __forceinline__ float2 band_fix(const GeomDev &g, int64_t base, int64_t off2, bool neg,
                                           int lane, bool ta, bool tb, const VisEnt<F64> &e,
                                           const float (&I)[kPixPerLane], float t_occ) {
    using T = typename PhaseT<F64>::T;
    float dre = 0.0f, dim = 0.0f;
#pragma unroll
    for (int j = 0; j < kPixPerLane; ++j) {
        const int64_t pi = base + j * kWave + lane;
        float4 p32 = g.pix32[neg ? pi - off2 : pi];
        if (neg) {
            p32.x = -p32.x; p32.y = -p32.y; p32.z = -p32.z;
        }
        bool unc = false, sh = false;
        if (ta) sh = occ32(p32.x, p32.y, p32.z, e.a, unc);
        if (tb) sh = occ32(p32.x, p32.y, p32.z, e.b, unc) || sh;
        if (unc) {
            int nl;
            const bool ex = occ_pair64(g.pos, g.pix64, pi, e.ia, e.ib, g.limb, &nl);
            if (ex != sh) {
                const float d = ex ? (t_occ - I[j]) : (I[j] - t_occ);
                T l, m, n;
                if (F64) {
                    const double4 p = g.pix64[pi];
                    l = (T)p.x; m = (T)p.y; n = (T)p.z;
                } else {
                    const float4 p = g.pix32[pi];
                    l = (T)p.x; m = (T)p.y; n = (T)p.z;
                }
                const T ph = fma(e.u, l, fma(e.v, m, fma(e.w, n, e.wo)));
                const float r = fract_rev(ph);
                dre = fmaf(d, __builtin_amdgcn_cosf(r), dre);
                dim = fmaf(d, __builtin_amdgcn_sinf(r), dim);
            }
        }
    }
    return make_float2(dre, dim);
}

template <bool F64, int VT>
__global__ void __launch_bounds__(kBlock, VT == 8 ? 4 : 3) visgen_pair_kernel(GeomDev g, PhaseKnobs k, int32_t n_maps,
                                                             const float *__restrict__ Ich,
                                                             float2 *__restrict__ V) {
    using T = typename PhaseT<F64>::T;
    static_assert(VT == 8 || VT == 16 || VT == 32, "tile of 8, 16 or 32 visibilities");
    static_assert(kListPad % VT == 0, "list padding must be whole tiles");
    __shared__ VisEnt<F64> ents[VT];
    __shared__ float red[kWavesPerBlock][VT][4];
    __shared__ float corr[kWavesPerBlock][VT][2];  // guard-band fixes, per wave
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // wave-uniform
    const bool list = g.vis_sat != nullptr;
    const int64_t n_vtiles = (g.n_vis + VT - 1) / VT;
    const int64_t n_tiles = list ? n_vtiles : n_vtiles * n_maps;
    const int nh = g.n_half;
    const int64_t off2 = (int64_t)nh * kChunk;  // position of a pixel's antipode, relative
    const float t_occ = k.t_occ;
    const float tv[kPixPerLane] = {t_occ, t_occ, t_occ, t_occ};

    for (int64_t tile = blockIdx.x; tile < n_tiles; tile += gridDim.x) {
        const int64_t map = list ? g.tile_map[(tile * VT) / kListPad] : tile / n_vtiles;
        const int64_t v0 = (list ? tile : tile % n_vtiles) * VT;
        const float *__restrict__ Im = Ich + map * g.n_pad;
        if (threadIdx.x < VT) fill_vis_ent<F64>(g, k, v0 + threadIdx.x, map, ents[threadIdx.x]);
        if (threadIdx.x < kWavesPerBlock * VT * 2) (&corr[0][0][0])[threadIdx.x] = 0.0f;
        __syncthreads();

        float a1[VT], b1[VT], a2[VT], b2[VT];
#pragma unroll
        for (int q = 0; q < VT; ++q) a1[q] = b1[q] = a2[q] = b2[q] = 0.0f;

        for (int c = wave; c < nh; c += kWavesPerBlock) {
            asm volatile("" ::: "memory");
            const ChunkRec ch1 = g.chunk[c], ch2 = g.chunk[c + nh];
            T l[kPixPerLane], m[kPixPerLane], n[kPixPerLane];
            float I1[kPixPerLane], I2[kPixPerLane];
#pragma unroll
            for (int j = 0; j < kPixPerLane; ++j) {
                const int64_t pi = (int64_t)c * kChunk + j * kWave + lane;
                if (F64) {
                    const double4 p = g.pix64[pi];
                    l[j] = (T)p.x; m[j] = (T)p.y; n[j] = (T)p.z;
                } else {
                    const float4 p = g.pix32[pi];
                    l[j] = (T)p.x; m[j] = (T)p.y; n[j] = (T)p.z;
                }
                I1[j] = Im[pi];
                I2[j] = Im[pi + off2];
            }
            uint32_t in1, out1, sa1, sb1, in2, out2, sa2, sb2;
            tile_masks<F64, VT>(ch1, ents, lane, in1, out1, sa1, sb1);
            tile_masks<F64, VT>(ch2, ents, lane, in2, out2, sa2, sb2);
            // Materialise the masks here, in scalar registers: left to itself the
            // compiler re-derives them from the ballots inside every pair's step.
            uint32_t both_out = __builtin_amdgcn_readfirstlane(out1 & out2);
            uint32_t straddle = __builtin_amdgcn_readfirstlane(sa1 | sb1 | sa2 | sb2);
            in1 = __builtin_amdgcn_readfirstlane(in1);
            in2 = __builtin_amdgcn_readfirstlane(in2);
            sa1 = __builtin_amdgcn_readfirstlane(sa1);
            sb1 = __builtin_amdgcn_readfirstlane(sb1);
            sa2 = __builtin_amdgcn_readfirstlane(sa2);
            sb2 = __builtin_amdgcn_readfirstlane(sb2);
            asm volatile("" : "+s"(both_out), "+s"(straddle), "+s"(in1), "+s"(in2));
            asm volatile("" : "+s"(sa1), "+s"(sb1), "+s"(sa2), "+s"(sb2));
            uint32_t unc1 = 0, unc2 = 0;  // pairs with a guard-band pixel in chunk c / c + nh
#pragma unroll
            for (int q = 0; q < VT; ++q) {
                const T u = ents[q].u, vv = ents[q].v, w = ents[q].w, wo = ents[q].wo;
                // Three paths per pair; each updates the accumulators in place
                // (acc_fmac), so that they keep their registers across the joins.
                if ((both_out >> q) & 1u) {
                    pair_terms<F64>(u, vv, w, wo, l, m, n, I1, I2, a1[q], b1[q], a2[q], b2[q]);
                } else if (!((straddle >> q) & 1u)) {
                    // each chunk wholly shaded or wholly clear for this pair, one
                    // of them shaded: t_occ is a wave-uniform operand
                    const bool in1q = (in1 >> q) & 1u, in2q = (in2 >> q) & 1u;
                    if (in1q && in2q)
                        pair_terms<F64>(u, vv, w, wo, l, m, n, tv, tv, a1[q], b1[q], a2[q], b2[q]);
                    else if (in1q)
                        pair_terms<F64>(u, vv, w, wo, l, m, n, tv, I2, a1[q], b1[q], a2[q], b2[q]);
                    else
                        pair_terms<F64>(u, vv, w, wo, l, m, n, I1, tv, a1[q], b1[q], a2[q], b2[q]);
                } else {
                    // a cap's FP32 dot serves the pixel and (negated) its antipode;
                    // the wave-uniform tests branch once per cap, not per pixel
                    const bool ta1 = (sa1 >> q) & 1u, tb1 = (sb1 >> q) & 1u;
                    const bool ta2 = (sa2 >> q) & 1u, tb2 = (sb2 >> q) & 1u;
                    bool sh1[kPixPerLane], sh2[kPixPerLane];
#pragma unroll
                    for (int j = 0; j < kPixPerLane; ++j) {
                        sh1[j] = (in1 >> q) & 1u;
                        sh2[j] = (in2 >> q) & 1u;
                    }
                    uint64_t uw1 = 0, uw2 = 0;
                    if (ta1 || ta2) pair_cap_tests(l, m, n, ents[q].a, ta1, ta2, sh1, sh2, uw1, uw2);
                    if (tb1 || tb2) pair_cap_tests(l, m, n, ents[q].b, tb1, tb2, sh1, sh2, uw1, uw2);
                    if (uw1) unc1 |= 1u << q;
                    if (uw2) unc2 |= 1u << q;
                    float e1[kPixPerLane], e2[kPixPerLane];
#pragma unroll
                    for (int j = 0; j < kPixPerLane; ++j) {
                        e1[j] = sh1[j] ? t_occ : I1[j];
                        e2[j] = sh2[j] ? t_occ : I2[j];
                    }
                    pair_terms<F64>(u, vv, w, wo, l, m, n, e1, e2, a1[q], b1[q], a2[q], b2[q]);
                }
            }
            // Guard-band pixels: the exact predicate; the difference goes, with
            // the pixel's own E, to the wave's LDS accumulator (this wave only:
            // deterministic).
            while (unc1) {
                const int q = __ffs(unc1) - 1;
                unc1 &= unc1 - 1;
                float2 d = band_fix<F64>(g, (int64_t)c * kChunk, off2, false, lane,
                                         (sa1 >> q) & 1u, (sb1 >> q) & 1u, ents[q], I1, t_occ);
                d.x = wave_sum(d.x);
                d.y = wave_sum(d.y);
                if (lane == 0) {
                    corr[wave][q][0] += d.x;
                    corr[wave][q][1] += d.y;
                }
            }
            while (unc2) {
                const int q = __ffs(unc2) - 1;
                unc2 &= unc2 - 1;
                float2 d = band_fix<F64>(g, (int64_t)c * kChunk + off2, off2, true, lane,
                                         (sa2 >> q) & 1u, (sb2 >> q) & 1u, ents[q], I2, t_occ);
                d.x = wave_sum(d.x);
                d.y = wave_sum(d.y);
                if (lane == 0) {
                    corr[wave][q][0] += d.x;
                    corr[wave][q][1] += d.y;
                }
            }
        }
#pragma unroll
        for (int q = 0; q < VT; ++q) {
            const float s0 = wave_sum(a1[q]), s1 = wave_sum(b1[q]);
            const float s2 = wave_sum(a2[q]), s3 = wave_sum(b2[q]);
            if (lane == 0) {
                red[wave][q][0] = s0;
                red[wave][q][1] = s1;
                red[wave][q][2] = s2;
                red[wave][q][3] = s3;
            }
        }
        __syncthreads();
        if (threadIdx.x < VT) {
            const int q = threadIdx.x;
            const int64_t v = v0 + q;
            const bool real = list ? (v < g.n_vis && g.vis_sat[v].x >= 0) : (v < g.n_vis);
            if (real) {
                float A1 = 0.0f, B1 = 0.0f, A2 = 0.0f, B2 = 0.0f, cr = 0.0f, ci = 0.0f;
#pragma unroll
                for (int w = 0; w < kWavesPerBlock; ++w) {
                    A1 += red[w][q][0];
                    B1 += red[w][q][1];
                    A2 += red[w][q][2];
                    B2 += red[w][q][3];
                    cr += corr[w][q][0];
                    ci += corr[w][q][1];
                }
                // exp(-4 pi i wo) = Cr - i Sr
                const float r2 = fract_rev(ents[q].wo + ents[q].wo);
                const float Cr = __builtin_amdgcn_cosf(r2), Sr = __builtin_amdgcn_sinf(r2);
                V[(list ? 0 : map * g.n_vis) + v] =
                    make_float2(A1 + fmaf(Cr, A2, Sr * B2) + cr, fmaf(Cr, B2, -Sr * A2) - B1 - ci);
            }
        }
        __syncthreads();
    }
}

// ------------------------------------------------------------------ stage 4
constexpr int kReconStage = kBlock;  // visibilities staged in LDS per step (one per thread)

// Per-visibility data the pixel loop reads for every staged visibility ...
template <bool F64> struct RecHot {
    typename PhaseT<F64>::T u, v, w, wo;
    float vr, vi;
    int32_t pad0, pad1;
};
// ... and what it reads only when the chunk straddles a limb: the fields of
// the two caps a pixel test reads (not CapRec itself: its 32-byte alignment
// would take LDS that the fifth block per CU needs).
struct CapTest {
    float dx, dy, dz, lo, hi;
};
struct RecCold {
    CapTest a, b;
    int32_t ia, ib;
};
static_assert(sizeof(RecCold) == 48, "recon cold record");

// Work item = (slot, chunk); a slot is a visibility range of one map.  In
// geometry mode slot = map * n_seg + seg over equal segments; in list mode
// the slots come from g.slots.  Each step stages kReconStage visibilities
// (one per thread), classifies them against the block's chunk once, and
// compacts away those that shade the whole chunk: wholly clear ones first,
// then those with one straddling cap (stored as cap a), then those with two,
// so that each kind runs its own branch-free loop.

// One staged straddling visibility over this lane's pixels: the FP32 test of
// cap a (and of cap b when TWO) against the guard band's edges as wave lane
// masks, the exact FP64 pair predicate for the pixels inside a band (one
// wave-uniform branch, rarely taken), shaded pixels dropped.
template <bool F64, bool TWO>
__device__ __forceinline__ void recon_straddle_masks(const GeomDev &g, int c, int lane,
                                               const RecHot<F64> &e, const RecCold &o,
                                               const typename PhaseT<F64>::T (&l)[kPixPerLane],
                                               const typename PhaseT<F64>::T (&m)[kPixPerLane],
                                               const typename PhaseT<F64>::T (&n)[kPixPerLane],
                                               const float (&sx)[kPixPerLane],
                                               const float (&sy)[kPixPerLane],
                                               const float (&sz)[kPixPerLane],
                                               float (&acc)[kPixPerLane]) {
    using T = typename PhaseT<F64>::T;
    uint64_t shm[kPixPerLane], uncm[kPixPerLane];  // per pixel j: shaded / inside a band
    uint64_t unc_any = 0;
#pragma unroll
    for (int j = 0; j < kPixPerLane; ++j) {
        const float xa = fmaf(sz[j], o.a.dz, fmaf(sy[j], o.a.dy, sx[j] * o.a.dx));
        uint64_t sh_w = __ballot(xa >= o.a.hi);
        uint64_t un_w = __ballot(xa >= o.a.lo) & ~sh_w;
        if (TWO) {
            const float xb = fmaf(sz[j], o.b.dz, fmaf(sy[j], o.b.dy, sx[j] * o.b.dx));
            const uint64_t sb = __ballot(xb >= o.b.hi);
            un_w |= __ballot(xb >= o.b.lo) & ~sb;
            sh_w |= sb;
        }
        shm[j] = sh_w;
        uncm[j] = un_w;
        unc_any |= un_w;
    }
    if (unc_any) {  // wave-uniform
#pragma unroll
        for (int j = 0; j < kPixPerLane; ++j) {
            if (uncm[j]) {  // wave-uniform
                bool ex = false;
                if ((uncm[j] >> lane) & 1ull) {
                    int nl;
                    ex = occ_pair64(g.pos, g.pix64, (int64_t)c * kChunk + j * kWave + lane, o.ia,
                                    o.ib, g.limb, &nl);
                }
                shm[j] = (shm[j] & ~uncm[j]) | __ballot(ex);
            }
        }
    }
#pragma unroll
    for (int j = 0; j < kPixPerLane; ++j) {
        const bool sh = __builtin_amdgcn_inverse_ballot_w64(shm[j]);
        const float vr = sh ? 0.0f : e.vr, vi = sh ? 0.0f : e.vi;
        const T ph = fma(e.u, l[j], fma(e.v, m[j], fma(e.w, n[j], e.wo)));
        const float r = fract_rev(ph);
        acc[j] = fmaf(vr, __builtin_amdgcn_cosf(r), acc[j]);
        acc[j] = fmaf(-vi, __builtin_amdgcn_sinf(r), acc[j]);
    }
}

// The same with the FP64 predicate in a per-pixel branch.  Timed against the
// lane-mask form on an MI355X (c2): 0.5 % faster for the FP32 phase, 2 % slower
// for the FP64 phase, so recon_straddle picks one per phase type.
template <bool F64, bool TWO>
__device__ __forceinline__ void recon_straddle_branch(const GeomDev &g, int c, int lane,
                                               const RecHot<F64> &e, const RecCold &o,
                                               const typename PhaseT<F64>::T (&l)[kPixPerLane],
                                               const typename PhaseT<F64>::T (&m)[kPixPerLane],
                                               const typename PhaseT<F64>::T (&n)[kPixPerLane],
                                               const float (&sx)[kPixPerLane],
                                               const float (&sy)[kPixPerLane],
                                               const float (&sz)[kPixPerLane],
                                               float (&acc)[kPixPerLane]) {
    using T = typename PhaseT<F64>::T;
    const CapRec A2 = {o.a.dx, o.a.dy, o.a.dz, 0.f, 0.f, 0.f, o.a.lo, o.a.hi};
    const CapRec B2 = {o.b.dx, o.b.dy, o.b.dz, 0.f, 0.f, 0.f, o.b.lo, o.b.hi};
#pragma unroll
    for (int j = 0; j < kPixPerLane; ++j) {
        bool unc = false;
        bool sh = occ32(sx[j], sy[j], sz[j], A2, unc);
        if (TWO) sh = occ32(sx[j], sy[j], sz[j], B2, unc) | sh;
        if (unc) {
            int nl;
            sh = occ_pair64(g.pos, g.pix64, (int64_t)c * kChunk + j * kWave + lane, o.ia, o.ib,
                            g.limb, &nl);
        }
        const float vr = sh ? 0.0f : e.vr, vi = sh ? 0.0f : e.vi;
        const T ph = fma(e.u, l[j], fma(e.v, m[j], fma(e.w, n[j], e.wo)));
        const float r = fract_rev(ph);
        acc[j] = fmaf(vr, __builtin_amdgcn_cosf(r), acc[j]);
        acc[j] = fmaf(-vi, __builtin_amdgcn_sinf(r), acc[j]);
    }
}

template <bool F64, bool TWO>
__device__ __forceinline__ void recon_straddle(const GeomDev &g, int c, int lane,
                                               const RecHot<F64> &e, const RecCold &o,
                                               const typename PhaseT<F64>::T (&l)[kPixPerLane],
                                               const typename PhaseT<F64>::T (&m)[kPixPerLane],
                                               const typename PhaseT<F64>::T (&n)[kPixPerLane],
                                               const float (&sx)[kPixPerLane],
                                               const float (&sy)[kPixPerLane],
                                               const float (&sz)[kPixPerLane],
                                               float (&acc)[kPixPerLane]) {
    if constexpr (F64)
        recon_straddle_masks<F64, TWO>(g, c, lane, e, o, l, m, n, sx, sy, sz, acc);
    else
        recon_straddle_branch<F64, TWO>(g, c, lane, e, o, l, m, n, sx, sy, sz, acc);
}

// Work queue of the recon kernels: blocks take (slot, chunk) items in index
// order from a device counter (zeroed by the launcher) instead of a fixed
// stride, so that blocks which drew cheap items (chunks most visibilities
// shade wholly) take more of them and the tail of the launch stays short
// whatever the chunk order.  Every block leaves once the counter passes the
// item count.
__device__ __forceinline__ int64_t next_item(unsigned long long *queue, int64_t &slot_s) {
    __syncthreads();  // every wave has read the previous item
    if (threadIdx.x == 0) slot_s = (int64_t)atomicAdd(queue, 1ull);
    __syncthreads();
    return slot_s;
}

__global__ void zero_queue_kernel(unsigned long long *queue) { *queue = 0ull; }

template <bool F64>
__device__ __forceinline__ void recon_body(const GeomDev &g, const PhaseKnobs &k, int32_t n_maps,
                                           const float2 *__restrict__ V,
                                           const float *__restrict__ wgt, float kappa,
                                           int32_t n_seg, float *__restrict__ part,
                                           unsigned long long *__restrict__ queue) {
    using T = typename PhaseT<F64>::T;
    __shared__ RecHot<F64> hot[kReconStage];
    __shared__ RecCold cold[kReconStage];
    __shared__ int4 wave_cnt[kWavesPerBlock];  // (clear, one cap, two caps) staged per wave
    __shared__ float red[kWavesPerBlock][kChunk];
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // wave-uniform
    const bool list = g.vis_sat != nullptr;
    const int64_t seg_len = (g.n_vis + n_seg - 1) / n_seg;
    const int64_t n_slots = list ? g.n_slots : (int64_t)n_maps * n_seg;
    const int64_t n_items = n_slots * g.n_chunk;

    __shared__ int64_t s_item;
    for (int64_t item = next_item(queue, s_item); item < n_items;
         item = next_item(queue, s_item)) {
        // the slot varies fastest: with n_slots a multiple of 8 each XCD keeps to its own slots
        const int64_t slot = item % n_slots;
        const int c = (int)(item / n_slots);
        int64_t map, vb, ve;
        float kap;
        if (list) {
            const int4 sl = g.slots[slot];
            map = sl.x;
            vb = sl.y;
            ve = sl.z;
            kap = g.kappa_map[map];
        } else {
            map = slot / n_seg;
            const int64_t seg = slot % n_seg;
            vb = seg * seg_len;
            ve = (vb + seg_len < g.n_vis) ? vb + seg_len : g.n_vis;
            kap = kappa;
        }
        const int64_t voff = list ? 0 : map * g.n_vis;
        const ChunkRec ch = g.chunk[c];
        T l[kPixPerLane], m[kPixPerLane], n[kPixPerLane];
        float sx[kPixPerLane], sy[kPixPerLane], sz[kPixPerLane], acc[kPixPerLane];
#pragma unroll
        for (int j = 0; j < kPixPerLane; ++j) {
            const int64_t pi = (int64_t)c * kChunk + j * kWave + lane;
            if (F64) {
                const double4 p = g.pix64[pi];
                l[j] = (T)p.x; m[j] = (T)p.y; n[j] = (T)p.z;
                sx[j] = (float)p.x; sy[j] = (float)p.y; sz[j] = (float)p.z;
            } else {
                const float4 p = g.pix32[pi];
                l[j] = (T)p.x; m[j] = (T)p.y; n[j] = (T)p.z;
                sx[j] = p.x; sy[j] = p.y; sz[j] = p.z;
            }
            acc[j] = 0.0f;
        }
        for (int64_t vt = vb; vt < ve; vt += kReconStage) {
            // ---- stage: one visibility per thread, classified, compacted
            const int64_t v = vt + threadIdx.x;
            int ia = 0, ib = 0, ca = kIn, cb = kIn;
            double u = 0, vv = 0, w = 0;
            CapRec A, B;
            if (v < ve && vis_geometry(g, v, map, ia, ib, u, vv, w)) {
                A = g.cap[ia];
                B = g.cap[ib];
                ca = classify_cap(ch, A);
                cb = classify_cap(ch, B);
            }
            const bool live = ca != kIn && cb != kIn;
            const bool clear = ca == kOut && cb == kOut;
            const bool two = ca == kStraddle && cb == kStraddle;
            const bool one = live && !clear && !two;
            const uint64_t bc = __ballot(clear), b1 = __ballot(one), b2 = __ballot(two);
            if (lane == 0)
                wave_cnt[wave] = make_int4(__popcll(bc), __popcll(b1), __popcll(b2), 0);
            __syncthreads();
            int base_c = 0, base_1 = 0, base_2 = 0, n_clear = 0, n_one = 0, n_live = 0;
#pragma unroll
            for (int q = 0; q < kWavesPerBlock; ++q) {
                const int4 wc = wave_cnt[q];
                base_c += (q < wave) ? wc.x : 0;
                base_1 += (q < wave) ? wc.y : 0;
                base_2 += (q < wave) ? wc.z : 0;
                n_clear += wc.x;
                n_one += wc.y;
                n_live += wc.x + wc.y + wc.z;
            }
            // all wave-uniform; saying so keeps the pixel loops below scalar loops
            base_c = __builtin_amdgcn_readfirstlane(base_c);
            base_1 = __builtin_amdgcn_readfirstlane(base_1);
            base_2 = __builtin_amdgcn_readfirstlane(base_2);
            n_clear = __builtin_amdgcn_readfirstlane(n_clear);
            n_one = __builtin_amdgcn_readfirstlane(n_one);
            n_live = __builtin_amdgcn_readfirstlane(n_live);
            if (live) {
                const uint64_t below = (1ull << lane) - 1ull;
                const int pos = clear ? base_c + __popcll(bc & below)
                                : one ? n_clear + base_1 + __popcll(b1 & below)
                                      : n_clear + n_one + base_2 + __popcll(b2 & below);
                RecHot<F64> &e = hot[pos];
                if (F64) {
                    e.u = (T)(u * k.s); e.v = (T)(vv * k.s); e.w = (T)(w * k.s);
                } else {
                    const float s = (float)k.s;
                    e.u = (T)((float)u * s); e.v = (T)((float)vv * s); e.w = (T)((float)w * s);
                }
                e.wo = k.wfac ? (T)0 : -e.w;
                const float2 q = V[voff + v];
                const float sc = kap * (wgt ? wgt[v] : 1.0f);
                e.vr = q.x * sc;
                e.vi = q.y * sc;
                RecCold &o = cold[pos];
                // one straddling cap goes first (cap b is then wholly clear of the chunk)
                const CapRec &t0 = (ca == kOut) ? B : A;
                o.a = {t0.dx, t0.dy, t0.dz, t0.lo, t0.hi};
                o.b = {B.dx, B.dy, B.dz, B.lo, B.hi};
                o.ia = ia;
                o.ib = ib;
            }
            __syncthreads();
            // ---- pixel loops: the wholly clear visibilities, then the straddling ones
            for (int q = wave; q < n_clear; q += kWavesPerBlock) {
                const RecHot<F64> e = hot[q];
#pragma unroll
                for (int j = 0; j < kPixPerLane; ++j) {
                    const T ph = fma(e.u, l[j], fma(e.v, m[j], fma(e.w, n[j], e.wo)));
                    const float r = fract_rev(ph);
                    // Re(V exp(+i theta)) = Vr cos - Vi sin
                    acc[j] = fmaf(e.vr, __builtin_amdgcn_cosf(r), acc[j]);
                    acc[j] = fmaf(-e.vi, __builtin_amdgcn_sinf(r), acc[j]);
                }
            }
            for (int q = n_clear + wave; q < n_clear + n_one; q += kWavesPerBlock)
                recon_straddle<F64, false>(g, c, lane, hot[q], cold[q], l, m, n, sx, sy, sz, acc);
            for (int q = n_clear + n_one + wave; q < n_live; q += kWavesPerBlock)
                recon_straddle<F64, true>(g, c, lane, hot[q], cold[q], l, m, n, sx, sy, sz, acc);
            __syncthreads();
        }
#pragma unroll
        for (int j = 0; j < kPixPerLane; ++j) red[wave][j * kWave + lane] = acc[j];
        __syncthreads();
        {
            const int p = threadIdx.x;  // kBlock == kChunk
            const float s = red[0][p] + red[1][p] + red[2][p] + red[3][p];
            part[slot * g.n_pad + (int64_t)c * kChunk + p] = s;
        }
        __syncthreads();
    }
}
static_assert(kBlock == kChunk, "recon reduction assumes one thread per chunk pixel");

// The FP32 recon runs five blocks per CU at its natural 86 registers; the
// FP64-phase one is held to five waves per SIMD too (a few spilled registers
// cost less than a fifth of the waves).
template <bool F64>
__global__ void __launch_bounds__(kBlock) recon_kernel(GeomDev g, PhaseKnobs k, int32_t n_maps,
                                                       const float2 *__restrict__ V,
                                                       const float *__restrict__ wgt,
                                                       float kappa, int32_t n_seg,
                                                       float *__restrict__ part, unsigned long long *__restrict__ queue) {
    recon_body<F64>(g, k, n_maps, V, wgt, kappa, n_seg, part, queue);
}
template <>
__global__ void __launch_bounds__(kBlock, 5) recon_kernel<true>(GeomDev g, PhaseKnobs k,
                                                                int32_t n_maps,
                                                                const float2 *__restrict__ V,
                                                                const float *__restrict__ wgt,
                                                                float kappa, int32_t n_seg,
                                                                float *__restrict__ part, unsigned long long *__restrict__ queue) {
    recon_body<true>(g, k, n_maps, V, wgt, kappa, n_seg, part, queue);
}

// F[map][ring[p]] = sum over the map's slots of part[slot][p].
__global__ void __launch_bounds__(256) recon_finalize_kernel(GeomDev g, int32_t n_maps,
                                                             int32_t n_seg, int64_t n_pix,
                                                             const float *__restrict__ part,
                                                             float *__restrict__ F) {
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (int64_t)n_maps * g.n_pad) return;
    const int64_t map = idx / g.n_pad, pi = idx % g.n_pad;
    const int r = g.ring[pi];
    if (r < 0) return;
    const int64_t s0 = g.map_slot0 ? g.map_slot0[map] : map * n_seg;
    const int64_t s1 = g.map_slot0 ? g.map_slot0[map + 1] : (map + 1) * n_seg;
    float s = 0.0f;
    for (int64_t q = s0; q < s1; ++q) s += part[q * g.n_pad + pi];
    F[map * n_pix + r] = s;
}

// Per chunk-order pixel: mean FF over its cells (same order as fine_value) and
// the cos / sin of the largest centre-to-cell angle, padded like chunk bounds.
__global__ void __launch_bounds__(256) fine_pixels_kernel(GeomDev g, float *__restrict__ cbar,
                                                          float2 *__restrict__ pixcap) {
    const int64_t pi = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (pi >= g.n_pad) return;
    const int n = g.fine.count[pi], s0 = g.fine.start[pi];
    const double4 p = g.pix64[pi];
    float acc = 0.0f;
    double minc = 1.0;
    for (int k = 0; k < n; ++k) {
        const int c = g.fine.idx[s0 + k];
        acc += g.fine.cell32[c].w;
        const double4 q = g.fine.cell64[c];
        minc = fmin(minc, p.x * q.x + p.y * q.y + p.z * q.z);
    }
    cbar[pi] = n > 0 ? acc / (float)n : 0.0f;
    double rho = acos(fmax(-1.0, fmin(1.0, minc)));
    rho = rho * (1.0 + 1e-6) + 1e-6;
    pixcap[pi] = make_float2((float)cos(rho), (float)sin(rho));
}

// Chunk bounds widened to hold every cell of every pixel in the chunk:
// rho_fine = (rho_chunk + max pixel-cell angle), padded like the chunk bounds.
__global__ void __launch_bounds__(256) fine_chunks_kernel(GeomDev g, const float2 *__restrict__ pixcap,
                                                          ChunkRec *__restrict__ out) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= g.n_chunk) return;
    ChunkRec ch = g.chunk[c];
    double mp = 0.0;
    for (int j = 0; j < kChunk; ++j) {
        const float2 pc = pixcap[(int64_t)c * kChunk + j];
        mp = fmax(mp, atan2((double)pc.y, (double)pc.x));
    }
    double rho = atan2((double)ch.srho, (double)ch.crho) + mp;
    rho = rho * (1.0 + 1e-6) + 1e-6;
    if (rho > 0.5) ch.flags |= kChunkAlwaysStraddle;
    rho = fmin(rho, M_PI);
    ch.crho = (float)cos(rho);
    ch.srho = (float)sin(rho);
    out[c] = ch;
}

__global__ void __launch_bounds__(256) from_chunk_order_kernel(GeomDev g, const float *__restrict__ in,
                                                               float *__restrict__ out) {
    const int64_t pi = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (pi >= g.n_pad) return;
    const int r = g.ring[pi];
    if (r >= 0) out[r] = in[pi];
}

inline unsigned blocks_for(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

}  // namespace

hipError_t launch_from_chunk_order(const GeomDev &g, int64_t n_pix, const float *in, float *out,
                                   hipStream_t st) {
    (void)n_pix;
    hipLaunchKernelGGL(from_chunk_order_kernel, dim3(blocks_for(g.n_pad, 256)), dim3(256), 0, st, g,
                       in, out);
    return hipGetLastError();
}

hipError_t launch_fine_pixels(const GeomDev &g, float *cbar, float2 *pixcap, ChunkRec *fchunk,
                              hipStream_t st) {
    hipLaunchKernelGGL(fine_pixels_kernel, dim3(blocks_for(g.n_pad, 256)), dim3(256), 0, st, g,
                       cbar, pixcap);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(fine_chunks_kernel, dim3(blocks_for(g.n_chunk, 256)), dim3(256), 0, st, g,
                       (const float2 *)pixcap, fchunk);
    return hipGetLastError();
}

hipError_t launch_orbit(const OrbitArgs &a, const GeomDev &g, double4 *pos, CapRec *cap,
                        double4 *uvw64, int2 *vis_cap, hipStream_t st) {
    const int64_t n = (int64_t)a.n_sat * a.n_t;
    hipLaunchKernelGGL(orbit_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, st, a, pos, cap);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    if (g.n_vis > 0)
        hipLaunchKernelGGL(uvw_kernel, dim3(blocks_for(g.n_vis, 256)), dim3(256), 0, st, g,
                           a.inv_lambda_km, (int)a.ceil_uvw, uvw64, vis_cap);
    return hipGetLastError();
}

hipError_t launch_shade(const GeomDev &g, int32_t *count, uint32_t *hash, uint32_t *bits,
                        int64_t n_words, unsigned long long *tot, int grid, hipStream_t st) {
    hipLaunchKernelGGL(shade_kernel, dim3(grid), dim3(kBlock), 0, st, g, count, hash, bits,
                       n_words, tot);
    return hipGetLastError();
}

hipError_t launch_to_chunk_order(const GeomDev &g, int32_t n_maps, int64_t n_pix, const float *I,
                                 float *Ich, hipStream_t st) {
    hipLaunchKernelGGL(to_chunk_order_kernel, dim3(blocks_for((int64_t)n_maps * g.n_pad, 256)),
                       dim3(256), 0, st, g, n_maps, n_pix, I, Ich);
    return hipGetLastError();
}

template <bool F64, int VT>
static hipError_t launch_visgen_t(const GeomDev &g, const PhaseKnobs &k, bool pair, int32_t n_maps,
                                  const float *Ich, float2 *V, int grid, hipStream_t st) {
    if (pair)  // tiles of 8 or 16 (see visgen_kernel_ptr)
        hipLaunchKernelGGL((visgen_pair_kernel<F64, VT == 8 ? 8 : 16>), dim3(grid), dim3(kBlock),
                           0, st, g, k, n_maps, Ich, V);
    else
        hipLaunchKernelGGL((visgen_kernel<F64, VT>), dim3(grid), dim3(kBlock), 0, st, g, k,
                           n_maps, Ich, V);
    return hipGetLastError();
}

hipError_t launch_visgen(const GeomDev &g, const PhaseKnobs &k, bool f64, int vt, bool pair,
                         int32_t n_maps, const float *Ich, float2 *V, int grid, hipStream_t st) {
    pair = pair && g.n_half > 0;
    switch (vt) {
        case 8:
            return f64 ? launch_visgen_t<true, 8>(g, k, pair, n_maps, Ich, V, grid, st)
                       : launch_visgen_t<false, 8>(g, k, pair, n_maps, Ich, V, grid, st);
        case 32:
            return f64 ? launch_visgen_t<true, 32>(g, k, pair, n_maps, Ich, V, grid, st)
                       : launch_visgen_t<false, 32>(g, k, pair, n_maps, Ich, V, grid, st);
        default:
            return f64 ? launch_visgen_t<true, 16>(g, k, pair, n_maps, Ich, V, grid, st)
                       : launch_visgen_t<false, 16>(g, k, pair, n_maps, Ich, V, grid, st);
    }
}

hipError_t launch_recon(const GeomDev &g, const PhaseKnobs &k, bool f64_phase, int32_t n_maps,
                        const float2 *V, const float *w, float kappa, int32_t n_seg, float *part,
                        unsigned long long *queue, int grid, hipStream_t st) {
    // the queue counter is cleared by a kernel, not a memset (graph-safe)
    hipLaunchKernelGGL(zero_queue_kernel, dim3(1), dim3(1), 0, st, queue);
    if (hipError_t e = hipGetLastError()) return e;
    if (f64_phase)
        hipLaunchKernelGGL(recon_kernel<true>, dim3(grid), dim3(kBlock), 0, st, g, k, n_maps, V,
                           w, kappa, n_seg, part, queue);
    else
        hipLaunchKernelGGL(recon_kernel<false>, dim3(grid), dim3(kBlock), 0, st, g, k, n_maps, V,
                           w, kappa, n_seg, part, queue);
    return hipGetLastError();
}

hipError_t launch_recon_finalize(const GeomDev &g, int32_t n_maps, int32_t n_seg, int64_t n_pix,
                                 const float *part, float *F, hipStream_t st) {
    hipLaunchKernelGGL(recon_finalize_kernel, dim3(blocks_for((int64_t)n_maps * g.n_pad, 256)),
                       dim3(256), 0, st, g, n_maps, n_seg, n_pix, part, F);
    return hipGetLastError();
}

hipError_t launch_visgen_fine(const GeomDev &g, const PhaseKnobs &k, bool f64, const float *Ich,
                              float2 *V, int grid, hipStream_t st) {
    if (f64)
        hipLaunchKernelGGL((visgen_kernel<true, 16, true>), dim3(grid), dim3(kBlock), 0, st, g, k,
                           1, Ich, V);
    else
        hipLaunchKernelGGL((visgen_kernel<false, 16, true>), dim3(grid), dim3(kBlock), 0, st, g, k,
                           1, Ich, V);
    return hipGetLastError();
}
const void *visgen_fine_kernel_ptr(bool f64) {
    return f64 ? (const void *)visgen_kernel<true, 16, true> : (const void *)visgen_kernel<false, 16, true>;
}

// ------------------------------------------------------------------ ceiling
// The bare plain-FP32 term at full occupancy, registers only: per (visibility,
// pixel) 3 FMA (phase), fract, v_cos, v_sin, 2 FMA (accumulate) and nothing
// else; the visibilities come from a 16-entry table through the scalar cache.
// Its rate is the VALU/transcendental ceiling of the vis-gen and recon terms.
constexpr int kCeilPix = 8;
__global__ void __launch_bounds__(256) term_ceiling_kernel(const float4 *__restrict__ uvw,
                                                           int32_t iters,
                                                           float *__restrict__ out) {
    const int gid = blockIdx.x * blockDim.x + threadIdx.x;
    float l[kCeilPix], m[kCeilPix], n[kCeilPix], I[kCeilPix], ar[kCeilPix], ai[kCeilPix];
#pragma unroll
    for (int j = 0; j < kCeilPix; ++j) {
        const float a = 0.37f * (float)((gid * kCeilPix + j) & 1023) * (1.0f / 1024.0f);
        l[j] = 0.6f * a;
        m[j] = 0.8f - 0.5f * a;
        n[j] = 0.1f + 0.3f * a;
        I[j] = 1.0f + 0.01f * (float)j;
        ar[j] = ai[j] = 0.0f;
    }
    for (int it = 0; it < iters; ++it) {
        const float4 q = uvw[it & 15];
#pragma unroll
        for (int j = 0; j < kCeilPix; ++j) {
            const float r = __builtin_amdgcn_fractf(fmaf(q.x, l[j], fmaf(q.y, m[j], fmaf(q.z, n[j], q.w))));
            ar[j] = fmaf(I[j], __builtin_amdgcn_cosf(r), ar[j]);
            ai[j] = fmaf(I[j], __builtin_amdgcn_sinf(r), ai[j]);
        }
    }
    float s = 0.0f;
#pragma unroll
    for (int j = 0; j < kCeilPix; ++j) s += ar[j] + ai[j];
    out[gid] = s;
}

hipError_t launch_term_ceiling(const float4 *uvw, int32_t iters, int32_t blocks, float *out,
                               hipStream_t st) {
    hipLaunchKernelGGL(term_ceiling_kernel, dim3(blocks), dim3(256), 0, st, uvw, iters, out);
    return hipGetLastError();
}
int64_t term_ceiling_terms(int32_t iters, int32_t blocks) {
    return (int64_t)blocks * 256 * iters * kCeilPix;
}

int visgen_tile_vis() { return 16; }
int list_pad() { return kListPad; }
int recon_stage_vis() { return kReconStage; }
const void *visgen_kernel_ptr(bool f64, int vt, bool pair) {
    if (pair) {  // tiles of 8 or 16 (32 would spill its four sums per visibility)
        if (vt == 8) return f64 ? (const void *)visgen_pair_kernel<true, 8> : (const void *)visgen_pair_kernel<false, 8>;
        return f64 ? (const void *)visgen_pair_kernel<true, 16> : (const void *)visgen_pair_kernel<false, 16>;
    }
    switch (vt) {
        case 8: return f64 ? (const void *)visgen_kernel<true, 8> : (const void *)visgen_kernel<false, 8>;
        case 32: return f64 ? (const void *)visgen_kernel<true, 32> : (const void *)visgen_kernel<false, 32>;
        default: return f64 ? (const void *)visgen_kernel<true, 16> : (const void *)visgen_kernel<false, 16>;
    }
}
const void *recon_kernel_ptr(bool f64) {
    return f64 ? (const void *)recon_kernel<true> : (const void *)recon_kernel<false>;
}
const void *shade_kernel_ptr() { return (const void *)shade_kernel; }

}  // namespace vrs

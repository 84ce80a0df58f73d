// uvw gridding, de-duplication and cell counts (the reference's stage-1 tail:
// uvw = ceil(dx * 1e3 / lambda) per ordered satellite pair,
// MIncline_generate_uvw_float.cu:153-160, then sort + unique per pair and
// globally, :374-397; the per-cell counts are what the updated_uvw files carry
// as their 4th column and what recon uses as 1/count weights,
// vissGen10M_shade.cu:357-366).
//
// Integer work, HBM-bound: every gridded sample becomes a 63-bit key
// ((u + 2^20) << 42 | (v + 2^20) << 21 | (w + 2^20)), whose unsigned order is
// the lexicographic (u, v, w) order the reference sorts by (SortCompare,
// MIncline_generate_uvw_float.cu:60-70).  Keys are sorted with a stable
// least-significant-digit radix sort over 8-bit digits (per pass: tile
// histograms, their prefix in (digit, tile) order, a stable scatter; 24 bytes
// of HBM traffic per key and pass), skipping digits that are the same for
// every key (the biased fields of nearby cells share their high digits), then
// run-length encoded into (unique key, count); each visibility's count is
// found by binary search.  Indices are 64-bit: the reference's own stage 1 grids
// 56 ordered pairs x 40.32 M samples, 2.3e9 keys.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "grid.hpp"

namespace vrs {

namespace {

constexpr int kGB = 256;                 // threads per block
constexpr int64_t kBias = 1 << 20;
constexpr int64_t kField = (int64_t)1 << 21;

__device__ __forceinline__ int64_t grid_key(double du, double dv, double dw, double sc, int *bad) {
    const double u = ceil(du * sc), v = ceil(dv * sc), w = ceil(dw * sc);
    if (!(fabs(u) < kBias && fabs(v) < kBias && fabs(w) < kBias)) {
        *bad = 1;
        return 0;
    }
    return (((int64_t)u + kBias) << 42) | (((int64_t)v + kBias) << 21) | ((int64_t)w + kBias);
}

// keys[p * n_t + t] over the given (ordered or unordered) pairs: one thread per
// step t walks all pairs, so each satellite's position at t is read from HBM
// about once and then from cache, and each pair's keys are written coalesced.
__global__ void __launch_bounds__(kGB) grid_keys_kernel(const double4 *__restrict__ pos, int n_t,
                                                        const int2 *__restrict__ pairs,
                                                        int n_pairs, double sc,
                                                        uint64_t *__restrict__ keys,
                                                        int *__restrict__ bad) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n_t) return;
    int b = 0;
#pragma unroll 4
    for (int p = 0; p < n_pairs; ++p) {
        const int2 pr = pairs[p];
        const double4 ra = pos[(int64_t)pr.x * n_t + t], rb = pos[(int64_t)pr.y * n_t + t];
        keys[(int64_t)p * n_t + t] = (uint64_t)grid_key(rb.x - ra.x, rb.y - ra.y, rb.z - ra.z, sc, &b);
    }
    if (b) atomicOr(bad, 1);
}

// ---- stable LSD radix sort of 64-bit keys, 8-bit digits
// A pass over digit d: per-tile digit histograms, their exclusive prefix in
// (digit, tile) order -- the first output position of each tile's keys of each
// digit -- then a scatter that ranks each tile's keys stably within their digit.
constexpr int kRB = 256;                  // threads per radix block
constexpr int kRI = 16;                   // keys per thread and tile
constexpr int kRT = kRB * kRI;            // keys per tile
constexpr int kSeg = 4096;                // tiles per scan segment

__device__ __forceinline__ int digit_of(uint64_t k, int d) { return (int)((k >> (8 * d)) & 0xffu); }

// Lanes of the wave that hold a valid key (ok) with the same digit as this lane.
__device__ __forceinline__ uint64_t same_digit_lanes(bool ok, int dg) {
    uint64_t peers = __ballot(ok);
#pragma unroll
    for (int bit = 0; bit < 8; ++bit) {
        const uint64_t m = __ballot((dg >> bit) & 1);
        peers &= ((dg >> bit) & 1) ? m : ~m;
    }
    return peers;
}

// AND and OR of all keys (andor[0] must start at ~0, andor[1] at 0): digit d is
// the same for every key exactly where the two agree.
__global__ void __launch_bounds__(kRB) key_and_or_kernel(const uint64_t *__restrict__ keys, int64_t n,
                                                         unsigned long long *__restrict__ andor) {
    unsigned long long a = ~0ull, o = 0ull;
    for (int64_t i = (int64_t)blockIdx.x * kRB + threadIdx.x; i < n; i += (int64_t)gridDim.x * kRB) {
        const unsigned long long k = keys[i];
        a &= k;
        o |= k;
    }
    for (int off = 32; off > 0; off >>= 1) {
        a &= __shfl_xor(a, off, 64);
        o |= __shfl_xor(o, off, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        atomicAnd(&andor[0], a);
        atomicOr(&andor[1], o);
    }
}

// hist[b * n_tiles + tile] = keys of tile `tile` whose digit d is b.
__global__ void __launch_bounds__(kRB) tile_hist_kernel(const uint64_t *__restrict__ keys, int64_t n,
                                                        int d, int64_t n_tiles,
                                                        int *__restrict__ hist) {
    __shared__ int h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    const int64_t base = (int64_t)blockIdx.x * kRT;
#pragma unroll 4
    for (int r = 0; r < kRI; ++r) {
        const int64_t i = base + r * kRB + threadIdx.x;
        const bool ok = i < n;
        const int dg = ok ? digit_of(keys[i], d) : 0;
        // one LDS add per distinct digit of the wave (skewed digits would serialise)
        const uint64_t peers = same_digit_lanes(ok, dg);
        if (ok && __popcll(peers & ((1ull << (threadIdx.x & 63)) - 1ull)) == 0)
            atomicAdd(&h[dg], __popcll(peers));
    }
    __syncthreads();
    hist[(int64_t)threadIdx.x * n_tiles + blockIdx.x] = h[threadIdx.x];
}

// Block-wide inclusive scan of one value per thread (kRB threads, a multiple of
// the wave); *total = sum.  Waves scan with shuffles, then add the totals of
// the waves before them (s: kRB / 64 entries of LDS).
__device__ __forceinline__ int64_t block_incl_scan(int64_t x, int64_t *s, int64_t *total) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int64_t y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) s[wave] = x;
    __syncthreads();
    int64_t off = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < kRB / 64; ++w) {
        const int64_t v = s[w];
        off += (w < wave) ? v : 0;
        tot += v;
    }
    *total = tot;
    __syncthreads();  // s is reused by the next call
    return x + off;
}

// seg[b * n_seg + g] = sum of hist[b][g * kSeg .. (g + 1) * kSeg); block (g, b).
__global__ void __launch_bounds__(kRB) seg_sum_kernel(const int *__restrict__ hist, int64_t n_tiles,
                                                      int64_t n_seg, int64_t *__restrict__ seg) {
    __shared__ int64_t s[kRB];
    const int64_t g = blockIdx.x, b = blockIdx.y;
    const int *row = hist + b * n_tiles;
    int64_t x = 0;
    for (int j = 0; j < kSeg / kRB; ++j) {
        const int64_t t = g * kSeg + j * kRB + threadIdx.x;
        if (t < n_tiles) x += row[t];
    }
    int64_t tot;
    block_incl_scan(x, s, &tot);
    if (threadIdx.x == 0) seg[b * n_seg + g] = tot;
}

// In-place exclusive scan of a[n] (one block of kRB threads, looping).
__global__ void __launch_bounds__(kRB) scan64_kernel(int64_t *__restrict__ a, int64_t n) {
    __shared__ int64_t s[kRB];
    int64_t carry = 0;
    for (int64_t base = 0; base < n; base += kRB) {
        const int64_t i = base + threadIdx.x;
        const int64_t x = (i < n) ? a[i] : 0;
        int64_t tot;
        const int64_t incl = block_incl_scan(x, s, &tot);
        if (i < n) a[i] = carry + incl - x;
        carry += tot;
    }
}

// offs[b * n_tiles + t] = seg_excl[b][g] + exclusive prefix of hist[b][g * kSeg .. t).
__global__ void __launch_bounds__(kRB) seg_offsets_kernel(const int *__restrict__ hist,
                                                          int64_t n_tiles, int64_t n_seg,
                                                          const int64_t *__restrict__ seg,
                                                          int64_t *__restrict__ offs) {
    __shared__ int64_t s[kRB];
    const int64_t g = blockIdx.x, b = blockIdx.y;
    const int *row = hist + b * n_tiles;
    int64_t carry = seg[b * n_seg + g];
    for (int j = 0; j < kSeg / kRB; ++j) {
        const int64_t t = g * kSeg + j * kRB + threadIdx.x;
        const int64_t x = (t < n_tiles) ? row[t] : 0;
        int64_t tot;
        const int64_t incl = block_incl_scan(x, s, &tot);
        if (t < n_tiles) offs[b * n_tiles + t] = carry + incl - x;
        carry += tot;
    }
}

// Stable scatter of one tile by digit d: round r takes keys r * kRB + thread, so
// tile order is (round, wave, lane); a key's rank within its digit counts the
// same-digit keys of earlier rounds, earlier waves of its round and lower lanes
// of its wave.
__global__ void __launch_bounds__(kRB) radix_scatter_kernel(const uint64_t *__restrict__ in,
                                                            int64_t n, int d, int64_t n_tiles,
                                                            const int64_t *__restrict__ offs,
                                                            uint64_t *__restrict__ out) {
    __shared__ int64_t base_of[256];       // output position of this tile's first key of each digit
    __shared__ int running[256];           // keys of each digit in earlier rounds
    __shared__ int cnt[kRB / 64][256];     // keys of each digit per wave in this round
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    base_of[threadIdx.x] = offs[(int64_t)threadIdx.x * n_tiles + blockIdx.x];
    running[threadIdx.x] = 0;
    const int64_t tbase = (int64_t)blockIdx.x * kRT;
    const uint64_t below = (1ull << lane) - 1ull;
    for (int r = 0; r < kRI; ++r) {
#pragma unroll
        for (int w = 0; w < kRB / 64; ++w) cnt[w][threadIdx.x] = 0;
        __syncthreads();
        const int64_t i = tbase + r * kRB + threadIdx.x;
        const bool ok = i < n;
        const uint64_t k = ok ? in[i] : 0;
        const int dg = digit_of(k, d);
        const uint64_t peers = same_digit_lanes(ok, dg);
        const int rank = __popcll(peers & below);
        if (ok && rank == 0) cnt[wave][dg] = __popcll(peers);
        __syncthreads();
        if (ok) {
            int before = running[dg];
            for (int w = 0; w < wave; ++w) before += cnt[w][dg];
            out[base_of[dg] + before + rank] = k;
        }
        __syncthreads();
        int add = 0;
#pragma unroll
        for (int w = 0; w < kRB / 64; ++w) add += cnt[w][threadIdx.x];
        running[threadIdx.x] += add;
    }
}

// ---- unique runs of sorted keys
constexpr int kUT = 1024;  // keys per unique block (256 threads x 4)

__device__ __forceinline__ bool run_start(const uint64_t *keys, int64_t i) {
    return i == 0 || keys[i] != keys[i - 1];
}

// bsum[blk] = run starts among keys [blk * kUT, (blk + 1) * kUT)
__global__ void __launch_bounds__(kRB) run_count_kernel(const uint64_t *__restrict__ keys,
                                                        int64_t n, int64_t *__restrict__ bsum) {
    __shared__ int64_t s[kRB];
    const int64_t base = (int64_t)blockIdx.x * kUT;
    int64_t x = 0;
#pragma unroll
    for (int r = 0; r < kUT / kRB; ++r) {
        const int64_t i = base + r * kRB + threadIdx.x;
        if (i < n && run_start(keys, i)) ++x;
    }
    int64_t tot;
    block_incl_scan(x, s, &tot);
    if (threadIdx.x == 0) bsum[blockIdx.x] = tot;
}

// run u (in key order) starting at i: ukeys[u] = keys[i], start[u] = i
__global__ void __launch_bounds__(kRB) run_starts_kernel(const uint64_t *__restrict__ keys,
                                                         int64_t n, const int64_t *__restrict__ boff,
                                                         uint64_t *__restrict__ ukeys,
                                                         int64_t *__restrict__ start) {
    __shared__ int cw[kUT / kRB][kRB / 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int64_t base = (int64_t)blockIdx.x * kUT;
    bool f[kUT / kRB];
    int pre[kUT / kRB];
#pragma unroll
    for (int r = 0; r < kUT / kRB; ++r) {
        const int64_t i = base + r * kRB + threadIdx.x;
        f[r] = i < n && run_start(keys, i);
        const uint64_t m = __ballot(f[r]);
        pre[r] = __popcll(m & ((1ull << lane) - 1ull));
        if (lane == 0) cw[r][wave] = __popcll(m);
    }
    __syncthreads();
    int64_t before = boff[blockIdx.x];
#pragma unroll
    for (int r = 0; r < kUT / kRB; ++r) {
        int64_t b = before;
        for (int w = 0; w < wave; ++w) b += cw[r][w];
        const int64_t i = base + r * kRB + threadIdx.x;
        if (f[r]) {
            ukeys[b + pre[r]] = keys[i];
            start[b + pre[r]] = i;
        }
        for (int w = 0; w < kRB / 64; ++w) before += cw[r][w];
    }
}

__global__ void __launch_bounds__(kRB) run_counts_kernel(const int64_t *__restrict__ start,
                                                         int64_t n_unique, int64_t n,
                                                         int32_t *__restrict__ counts) {
    const int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= n_unique) return;
    const int64_t e = (u + 1 < n_unique) ? start[u + 1] : n;
    counts[u] = (int32_t)(e - start[u]);
}

__global__ void __launch_bounds__(kGB) decode_kernel(const uint64_t *__restrict__ ukeys,
                                                     int64_t n_unique, float *__restrict__ uvw) {
    const int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= n_unique) return;
    const uint64_t k = ukeys[u];
    uvw[3 * u] = (float)((int64_t)((k >> 42) & (kField - 1)) - kBias);
    uvw[3 * u + 1] = (float)((int64_t)((k >> 21) & (kField - 1)) - kBias);
    uvw[3 * u + 2] = (float)((int64_t)(k & (kField - 1)) - kBias);
}

// weight[v] = 1 / count of the cell of visibility v's gridded uvw
__global__ void __launch_bounds__(kGB) vis_weight_kernel(const uint64_t *__restrict__ vkeys,
                                                         int64_t n_vis,
                                                         const uint64_t *__restrict__ ukeys,
                                                         const int32_t *__restrict__ counts,
                                                         int64_t n_unique, float *__restrict__ wgt,
                                                         int32_t *__restrict__ vcount) {
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n_vis) return;
    const uint64_t k = vkeys[v];
    int64_t lo = 0, hi = n_unique;  // first index with ukeys[idx] >= k
    while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if (ukeys[mid] < k) lo = mid + 1;
        else hi = mid;
    }
    const int c = (lo < n_unique && ukeys[lo] == k) ? counts[lo] : 0;
    if (vcount) vcount[v] = c;
    if (wgt) wgt[v] = c > 0 ? 1.0f / (float)c : 0.0f;
}

unsigned blocks_of(int64_t n, int per) { return (unsigned)((n + per - 1) / per); }

// ---- fine-cell sky: inverse index (cells grouped by parent pixel)
__global__ void __launch_bounds__(kGB) fine_pack_kernel(const double *__restrict__ lmn,
                                                        const float *__restrict__ ff,
                                                        const int32_t *__restrict__ nx,
                                                        int32_t nx_base, int64_t n, int64_t n_pix,
                                                        float4 *__restrict__ c32,
                                                        double4 *__restrict__ c64,
                                                        uint64_t *__restrict__ keys,
                                                        int *__restrict__ bad) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double x = lmn[3 * i], y = lmn[3 * i + 1], z = lmn[3 * i + 2];
    c32[i] = make_float4((float)x, (float)y, (float)z, ff[i]);
    c64[i] = make_double4(x, y, z, 0.0);
    if (!(fabs(fma(z, z, fma(y, y, x * x)) - 1.0) <= 1e-5)) atomicOr(bad, 4);
    const int64_t p = (int64_t)nx[i] - nx_base;
    if (p < 0 || p >= n_pix) atomicOr(bad, 1);
    keys[i] = ((uint64_t)(uint32_t)(p < 0 ? 0 : p) << 32) | (uint64_t)(uint32_t)i;
}

__global__ void __launch_bounds__(kGB) fine_idx_kernel(const uint64_t *__restrict__ keys, int64_t n,
                                                       int32_t *__restrict__ idx) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) idx[i] = (int32_t)(keys[i] & 0xffffffffull);
}

// per chunk-order pixel: [start, start + count) of its cells in the sorted keys
__global__ void __launch_bounds__(kGB) fine_ranges_kernel(const uint64_t *__restrict__ keys,
                                                          int64_t n, const int32_t *__restrict__ ring,
                                                          int64_t n_pad, int32_t *__restrict__ start,
                                                          int32_t *__restrict__ count,
                                                          int *__restrict__ bad) {
    const int64_t pi = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (pi >= n_pad) return;
    const int r = ring[pi];
    if (r < 0) {
        start[pi] = 0;
        count[pi] = 0;
        return;
    }
    const uint64_t lo_key = (uint64_t)(uint32_t)r << 32, hi_key = (uint64_t)(uint32_t)(r + 1) << 32;
    int64_t a = 0, b = n;
    while (a < b) { const int64_t m = (a + b) >> 1; if (keys[m] < lo_key) a = m + 1; else b = m; }
    int64_t c = a, d = n;
    while (c < d) { const int64_t m = (c + d) >> 1; if (keys[m] < hi_key) c = m + 1; else d = m; }
    start[pi] = (int32_t)a;
    count[pi] = (int32_t)(c - a);
    if (c == a) atomicOr(bad, 2);
}

// countLoc in RING order from the chunk-order counts
__global__ void __launch_bounds__(kGB) fine_count_ring_kernel(const int32_t *__restrict__ count,
                                                              const int32_t *__restrict__ ring,
                                                              int64_t n_pad,
                                                              int32_t *__restrict__ out) {
    const int64_t pi = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (pi >= n_pad) return;
    const int r = ring[pi];
    if (r >= 0) out[r] = count[pi];
}

}  // namespace

namespace {
struct RadixScratch {
    int64_t n_tiles, n_seg;
    int *hist;                  // [256][n_tiles]
    int64_t *offs;              // [256][n_tiles]
    int64_t *seg;               // [256][n_seg]
    unsigned long long *andor;  // [2]: AND and OR of all keys
};
RadixScratch radix_layout(int64_t n, void *scratch) {
    RadixScratch r;
    r.n_tiles = (n + kRT - 1) / kRT;
    r.n_seg = (r.n_tiles + kSeg - 1) / kSeg;
    char *p = (char *)scratch;
    r.andor = (unsigned long long *)p;
    p += 256;
    r.seg = (int64_t *)p;
    p += sizeof(int64_t) * 256 * r.n_seg;
    r.offs = (int64_t *)p;
    p += sizeof(int64_t) * 256 * r.n_tiles;
    r.hist = (int *)p;
    return r;
}
}  // namespace

size_t radix_scratch_bytes(int64_t n) {
    const int64_t n_tiles = (n + kRT - 1) / kRT, n_seg = (n_tiles + kSeg - 1) / kSeg;
    return 256 + sizeof(int64_t) * 256 * n_seg +
           sizeof(int64_t) * 256 * n_tiles + sizeof(int) * 256 * n_tiles + 256;
}

hipError_t radix_sort(uint64_t *keys, uint64_t *tmp, int64_t n, int d_lo, int d_hi, void *scratch,
                      hipStream_t st, int *passes_run) {
    if (passes_run) *passes_run = 0;
    if (n <= 1) return hipSuccess;
    const RadixScratch r = radix_layout(n, scratch);
    const unsigned long long init[2] = {~0ull, 0ull};
    hipError_t e = hipMemcpyAsync(r.andor, init, sizeof init, hipMemcpyHostToDevice, st);
    if (e != hipSuccess) return e;
    const int64_t hb = std::min<int64_t>((n + kRB - 1) / kRB, 2048);
    hipLaunchKernelGGL(key_and_or_kernel, dim3((unsigned)hb), dim3(kRB), 0, st, keys, n, r.andor);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    unsigned long long ao[2];
    if ((e = hipMemcpyAsync(ao, r.andor, sizeof ao, hipMemcpyDeviceToHost, st)) != hipSuccess) return e;
    if ((e = hipStreamSynchronize(st)) != hipSuccess) return e;
    uint64_t *a = keys, *b = tmp;
    for (int d = d_lo; d < d_hi; ++d) {
        // every key has the same digit d: the pass would keep the order
        if ((((ao[0] ^ ao[1]) >> (8 * d)) & 0xffull) == 0) continue;
        hipLaunchKernelGGL(tile_hist_kernel, dim3((unsigned)r.n_tiles), dim3(kRB), 0, st, a, n, d,
                           r.n_tiles, r.hist);
        hipLaunchKernelGGL(seg_sum_kernel, dim3((unsigned)r.n_seg, 256), dim3(kRB), 0, st,
                           (const int *)r.hist, r.n_tiles, r.n_seg, r.seg);
        hipLaunchKernelGGL(scan64_kernel, dim3(1), dim3(kRB), 0, st, r.seg, 256 * r.n_seg);
        hipLaunchKernelGGL(seg_offsets_kernel, dim3((unsigned)r.n_seg, 256), dim3(kRB), 0, st,
                           (const int *)r.hist, r.n_tiles, r.n_seg, (const int64_t *)r.seg, r.offs);
        hipLaunchKernelGGL(radix_scatter_kernel, dim3((unsigned)r.n_tiles), dim3(kRB), 0, st,
                           (const uint64_t *)a, n, d, r.n_tiles, (const int64_t *)r.offs, b);
        if ((e = hipGetLastError()) != hipSuccess) return e;
        std::swap(a, b);
        if (passes_run) ++*passes_run;
    }
    if (a == keys) return hipSuccess;
    return hipMemcpyAsync(keys, a, sizeof(uint64_t) * n, hipMemcpyDeviceToDevice, st);
}

size_t unique_scratch_bytes(int64_t n) {
    return sizeof(int64_t) * (size_t)((n + kUT - 1) / kUT + 1) + 256;
}

hipError_t sorted_unique_count(const uint64_t *keys, int64_t n, void *scratch, int64_t *n_unique,
                               hipStream_t st) {
    *n_unique = 0;
    if (n <= 0) return hipSuccess;
    const int64_t nb = (n + kUT - 1) / kUT;
    int64_t *boff = (int64_t *)scratch;  // [nb + 1]
    hipLaunchKernelGGL(run_count_kernel, dim3((unsigned)nb), dim3(kRB), 0, st, keys, n, boff);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    // boff[nb] = 0 before the scan makes the scan's last entry the total
    const int64_t zero = 0;
    if ((e = hipMemcpyAsync(boff + nb, &zero, sizeof zero, hipMemcpyHostToDevice, st)) != hipSuccess)
        return e;
    hipLaunchKernelGGL(scan64_kernel, dim3(1), dim3(kRB), 0, st, boff, nb + 1);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if ((e = hipMemcpyAsync(n_unique, boff + nb, sizeof(int64_t), hipMemcpyDeviceToHost, st)) !=
        hipSuccess)
        return e;
    return hipStreamSynchronize(st);
}

hipError_t sorted_unique_runs(const uint64_t *keys, int64_t n, const void *scratch, int64_t n_unique,
                              uint64_t *ukeys, int64_t *start, int32_t *counts, hipStream_t st) {
    if (n <= 0 || n_unique <= 0) return hipSuccess;
    const int64_t nb = (n + kUT - 1) / kUT;
    hipLaunchKernelGGL(run_starts_kernel, dim3((unsigned)nb), dim3(kRB), 0, st, keys, n,
                       (const int64_t *)scratch, ukeys, start);
    hipLaunchKernelGGL(run_counts_kernel, dim3(blocks_of(n_unique, kGB)), dim3(kGB), 0, st,
                       (const int64_t *)start, n_unique, n, counts);
    return hipGetLastError();
}

hipError_t fine_build_index(const double *lmn, const float *ff, const int32_t *nx,
                            int32_t nx_base, int64_t n, int64_t n_pix, const int32_t *ring,
                            int64_t n_pad, float4 *c32, double4 *c64, uint64_t *keys,
                            uint64_t *tmp, void *scratch, int *bad, int32_t *idx, int32_t *start,
                            int32_t *count, hipStream_t st) {
    hipLaunchKernelGGL(fine_pack_kernel, dim3(blocks_of(n, kGB)), dim3(kGB), 0, st, lmn, ff, nx,
                       nx_base, n, n_pix, c32, c64, keys, bad);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    // the keys start in cell order, so a stable sort on the parent digits alone
    // (bits 32 .. 63; the parent index is below 2^31) groups them
    if ((e = radix_sort(keys, tmp, n, 4, 8, scratch, st, nullptr)) != hipSuccess) return e;
    hipLaunchKernelGGL(fine_idx_kernel, dim3(blocks_of(n, kGB)), dim3(kGB), 0, st, keys, n, idx);
    hipLaunchKernelGGL(fine_ranges_kernel, dim3(blocks_of(n_pad, kGB)), dim3(kGB), 0, st, keys, n,
                       ring, n_pad, start, count, bad);
    return hipGetLastError();
}

hipError_t fine_count_ring(const int32_t *count, const int32_t *ring, int64_t n_pad, int32_t *out,
                           hipStream_t st) {
    hipLaunchKernelGGL(fine_count_ring_kernel, dim3(blocks_of(n_pad, kGB)), dim3(kGB), 0, st, count,
                       ring, n_pad, out);
    return hipGetLastError();
}

hipError_t grid_build_keys(const double4 *pos, int n_t, const int2 *pairs, int n_pairs, double sc,
                           uint64_t *keys, int *bad, hipStream_t st) {
    const int64_t n = (int64_t)n_pairs * n_t;
    (void)n;
    hipLaunchKernelGGL(grid_keys_kernel, dim3(blocks_of(n_t, kGB)), dim3(kGB), 0, st, pos, n_t, pairs,
                       n_pairs, sc, keys, bad);
    return hipGetLastError();
}

hipError_t grid_decode(const uint64_t *ukeys, int64_t n_unique, float *uvw, hipStream_t st) {
    hipLaunchKernelGGL(decode_kernel, dim3(blocks_of(n_unique, kGB)), dim3(kGB), 0, st, ukeys,
                       n_unique, uvw);
    return hipGetLastError();
}

hipError_t grid_vis_weights(const uint64_t *vkeys, int64_t n_vis, const uint64_t *ukeys,
                            const int32_t *counts, int64_t n_unique, float *wgt, int32_t *vcount,
                            hipStream_t st) {
    hipLaunchKernelGGL(vis_weight_kernel, dim3(blocks_of(n_vis, kGB)), dim3(kGB), 0, st, vkeys,
                       n_vis, ukeys, counts, n_unique, wgt, vcount);
    return hipGetLastError();
}


}  // namespace vrs

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
// least-significant-bit-first split sort (one pass per key bit: count zeros
// per block, scan, scatter), run-length encoded into (unique key, count), and
// each visibility's count is found by binary search.
#include <hip/hip_runtime.h>

#include "grid.hpp"

namespace vrs {

namespace {

constexpr int kGB = 256;                 // threads per block
constexpr int kGItems = 4;               // keys per thread
constexpr int kGTile = kGB * kGItems;    // keys per block
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

// keys[i] for i = p * n_t + t over the given (ordered or unordered) pairs.
__global__ void __launch_bounds__(kGB) grid_keys_kernel(const double4 *__restrict__ pos, int n_t,
                                                        const int2 *__restrict__ pairs,
                                                        int64_t n, double sc,
                                                        uint64_t *__restrict__ keys,
                                                        int *__restrict__ bad) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int p = (int)(i / n_t), t = (int)(i % n_t);
    const int2 pr = pairs[p];
    const double4 ra = pos[(int64_t)pr.x * n_t + t], rb = pos[(int64_t)pr.y * n_t + t];
    int b = 0;
    const int64_t k = grid_key(rb.x - ra.x, rb.y - ra.y, rb.z - ra.z, sc, &b);
    keys[i] = (uint64_t)k;
    if (b) atomicOr(bad, 1);
}

// zeros[blk] = number of keys in block blk whose bit `bit` is 0
__global__ void __launch_bounds__(kGB) split_count_kernel(const uint64_t *__restrict__ keys,
                                                          int64_t n, int bit,
                                                          int *__restrict__ zeros) {
    __shared__ int s[kGB / 64];
    int z = 0;
    const int64_t base = (int64_t)blockIdx.x * kGTile;
#pragma unroll
    for (int r = 0; r < kGItems; ++r) {
        const int64_t i = base + r * kGB + threadIdx.x;
        if (i < n) z += ((keys[i] >> bit) & 1ull) ? 0 : 1;
    }
    for (int o = 32; o > 0; o >>= 1) z += __shfl_xor(z, o, 64);
    if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = z;
    __syncthreads();
    if (threadIdx.x == 0) zeros[blockIdx.x] = s[0] + s[1] + s[2] + s[3];
}

// In-place exclusive scan of a small array (one block); returns the total in *total.
__global__ void __launch_bounds__(1024) scan_small_kernel(int *__restrict__ a, int n,
                                                          int *__restrict__ total) {
    __shared__ int s[1024];
    __shared__ int carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (int base = 0; base < n; base += 1024) {
        const int i = base + threadIdx.x;
        const int x = (i < n) ? a[i] : 0;
        s[threadIdx.x] = x;
        __syncthreads();
        for (int o = 1; o < 1024; o <<= 1) {
            const int y = (threadIdx.x >= o) ? s[threadIdx.x - o] : 0;
            __syncthreads();
            s[threadIdx.x] += y;
            __syncthreads();
        }
        if (i < n) a[i] = carry + s[threadIdx.x] - x;
        __syncthreads();
        if (threadIdx.x == 1023) carry += s[1023];
        __syncthreads();
    }
    if (threadIdx.x == 0) *total = carry;
}

// Stable split by bit: zeros keep their order in front, ones after them.
__global__ void __launch_bounds__(kGB) split_scatter_kernel(const uint64_t *__restrict__ in,
                                                            int64_t n, int bit,
                                                            const int *__restrict__ zoff,
                                                            const int *__restrict__ ztotal,
                                                            uint64_t *__restrict__ out) {
    __shared__ int wz[kGItems][kGB / 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int64_t base = (int64_t)blockIdx.x * kGTile;
    uint64_t k[kGItems];
    bool one[kGItems];
    int pre[kGItems];
#pragma unroll
    for (int r = 0; r < kGItems; ++r) {
        const int64_t i = base + r * kGB + threadIdx.x;
        k[r] = (i < n) ? in[i] : 0;
        one[r] = ((k[r] >> bit) & 1ull) != 0;
        const bool z = (i < n) && !one[r];
        const uint64_t m = __ballot(z);
        pre[r] = __popcll(m & ((1ull << lane) - 1ull));
        if (lane == 0) wz[r][wave] = __popcll(m);
    }
    __syncthreads();
    // zeros before (row r, wave w) in block order: rows first, then waves
    int zero_before = 0;
    const int blk_z0 = zoff[blockIdx.x];
    const int tot_z = *ztotal;
#pragma unroll
    for (int r = 0; r < kGItems; ++r) {
        int zb = zero_before;
        for (int w = 0; w < wave; ++w) zb += wz[r][w];
        const int64_t i = base + r * kGB + threadIdx.x;
        if (i < n) {
            const int64_t local = r * kGB + threadIdx.x;   // position within the block
            const int zpos = zb + pre[r];                   // zeros before this key in the block
            int64_t dst;
            if (!one[r]) {
                dst = blk_z0 + zpos;
            } else {
                const int64_t ones_before_blocks = base - blk_z0;   // ones in earlier blocks
                dst = tot_z + ones_before_blocks + (local - zpos);
            }
            out[dst] = k[r];
        }
        for (int w = 0; w < kGB / 64; ++w) zero_before += wz[r][w];
    }
}

// flags[i] = 1 where a new key starts
__global__ void __launch_bounds__(kGB) run_flags_kernel(const uint64_t *__restrict__ keys,
                                                        int64_t n, int *__restrict__ flags) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) flags[i] = (i == 0 || keys[i] != keys[i - 1]) ? 1 : 0;
}

// Block-level inclusive pieces of a large exclusive scan: block sums first.
__global__ void __launch_bounds__(kGB) block_sum_kernel(const int *__restrict__ a, int64_t n,
                                                        int *__restrict__ sums) {
    __shared__ int s[kGB / 64];
    int x = 0;
    const int64_t base = (int64_t)blockIdx.x * kGTile;
#pragma unroll
    for (int r = 0; r < kGItems; ++r) {
        const int64_t i = base + r * kGB + threadIdx.x;
        if (i < n) x += a[i];
    }
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
    if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = x;
    __syncthreads();
    if (threadIdx.x == 0) sums[blockIdx.x] = s[0] + s[1] + s[2] + s[3];
}

// run starts: where flags[i] = 1, write the run index -> start position
__global__ void __launch_bounds__(kGB) run_starts_kernel(const uint64_t *__restrict__ keys,
                                                         const int *__restrict__ flags, int64_t n,
                                                         const int *__restrict__ boff,
                                                         uint64_t *__restrict__ ukeys,
                                                         int *__restrict__ start) {
    __shared__ int wz[kGItems][kGB / 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int64_t base = (int64_t)blockIdx.x * kGTile;
    int f[kGItems], pre[kGItems];
#pragma unroll
    for (int r = 0; r < kGItems; ++r) {
        const int64_t i = base + r * kGB + threadIdx.x;
        f[r] = (i < n) ? flags[i] : 0;
        const uint64_t m = __ballot(f[r] != 0);
        pre[r] = __popcll(m & ((1ull << lane) - 1ull));
        if (lane == 0) wz[r][wave] = __popcll(m);
    }
    __syncthreads();
    int before = boff[blockIdx.x];
#pragma unroll
    for (int r = 0; r < kGItems; ++r) {
        int b = before;
        for (int w = 0; w < wave; ++w) b += wz[r][w];
        const int64_t i = base + r * kGB + threadIdx.x;
        if (i < n && f[r]) {
            const int u = b + pre[r];
            ukeys[u] = keys[i];
            start[u] = (int)i;
        }
        for (int w = 0; w < kGB / 64; ++w) before += wz[r][w];
    }
}

__global__ void __launch_bounds__(kGB) run_counts_kernel(const int *__restrict__ start,
                                                         int n_unique, int64_t n,
                                                         int32_t *__restrict__ counts) {
    const int u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= n_unique) return;
    const int64_t e = (u + 1 < n_unique) ? start[u + 1] : n;
    counts[u] = (int32_t)(e - start[u]);
}

__global__ void __launch_bounds__(kGB) decode_kernel(const uint64_t *__restrict__ ukeys,
                                                     int n_unique, float *__restrict__ uvw) {
    const int u = blockIdx.x * blockDim.x + threadIdx.x;
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
                                                         int n_unique, float *__restrict__ wgt,
                                                         int32_t *__restrict__ vcount) {
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n_vis) return;
    const uint64_t k = vkeys[v];
    int lo = 0, hi = n_unique;  // first index with ukeys[idx] >= k
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
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

hipError_t grid_sort_keys(uint64_t *keys, uint64_t *tmp, int64_t n, int *small, int bit_lo,
                          int bit_hi, hipStream_t st) {
    const int64_t nb = (n + kGTile - 1) / kGTile;
    int *zoff = small, *total = small + nb;
    uint64_t *a = keys, *b = tmp;
    for (int bit = bit_lo; bit < bit_hi; ++bit) {
        hipLaunchKernelGGL(split_count_kernel, dim3((unsigned)nb), dim3(kGB), 0, st, a, n, bit, zoff);
        hipLaunchKernelGGL(scan_small_kernel, dim3(1), dim3(1024), 0, st, zoff, (int)nb, total);
        hipLaunchKernelGGL(split_scatter_kernel, dim3((unsigned)nb), dim3(kGB), 0, st, a, n, bit,
                           (const int *)zoff, (const int *)total, b);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
        uint64_t *t = a;
        a = b;
        b = t;
    }
    if (a == keys) return hipSuccess;
    return hipMemcpyAsync(keys, a, sizeof(uint64_t) * n, hipMemcpyDeviceToDevice, st);
}

hipError_t fine_build_index(const double *lmn, const float *ff, const int32_t *nx,
                            int32_t nx_base, int64_t n, int64_t n_pix, const int32_t *ring,
                            int64_t n_pad, float4 *c32, double4 *c64, uint64_t *keys,
                            uint64_t *tmp, int *small, int *bad, int32_t *idx, int32_t *start,
                            int32_t *count, hipStream_t st) {
    hipLaunchKernelGGL(fine_pack_kernel, dim3(blocks_of(n, kGB)), dim3(kGB), 0, st, lmn, ff, nx,
                       nx_base, n, n_pix, c32, c64, keys, bad);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    // the keys start in cell order, so a stable sort on the parent bits alone groups them
    int bits = 1;
    while (bits < 31 && ((int64_t)1 << bits) < n_pix) ++bits;
    if ((e = grid_sort_keys(keys, tmp, n, small, 32, 32 + bits, st)) != hipSuccess) return e;
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

size_t grid_scratch_bytes(int64_t n_keys) {
    const int64_t nb = (n_keys + kGTile - 1) / kGTile;
    return (size_t)n_keys * 8 * 2 + (size_t)n_keys * 4 * 2 + (size_t)(nb + 2) * 4 * 2 + 1024;
}

hipError_t grid_build_keys(const double4 *pos, int n_t, const int2 *pairs, int n_pairs, double sc,
                           uint64_t *keys, int *bad, hipStream_t st) {
    const int64_t n = (int64_t)n_pairs * n_t;
    hipLaunchKernelGGL(grid_keys_kernel, dim3(blocks_of(n, kGB)), dim3(kGB), 0, st, pos, n_t, pairs,
                       n, sc, keys, bad);
    return hipGetLastError();
}

hipError_t grid_sort_unique(uint64_t *keys, uint64_t *tmp, int64_t n, int *iscratch, int *small,
                            uint64_t *ukeys, int32_t *counts, int *n_unique_dev, hipStream_t st) {
    const int64_t nb = (n + kGTile - 1) / kGTile;
    int *zoff = small;            // [nb]
    hipError_t e = grid_sort_keys(keys, tmp, n, small, 0, 63, st);
    if (e != hipSuccess) return e;
    int *flags = iscratch;              // [n]
    int *start = iscratch + n;          // [n]
    hipLaunchKernelGGL(run_flags_kernel, dim3(blocks_of(n, kGB)), dim3(kGB), 0, st, keys, n, flags);
    hipLaunchKernelGGL(block_sum_kernel, dim3((unsigned)nb), dim3(kGB), 0, st, flags, n, zoff);
    hipLaunchKernelGGL(scan_small_kernel, dim3(1), dim3(1024), 0, st, zoff, (int)nb, n_unique_dev);
    hipLaunchKernelGGL(run_starts_kernel, dim3((unsigned)nb), dim3(kGB), 0, st, keys, flags, n,
                       zoff, ukeys, start);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    return hipSuccess;
}

hipError_t grid_counts(const int *start, int n_unique, int64_t n, int32_t *counts, hipStream_t st) {
    hipLaunchKernelGGL(run_counts_kernel, dim3(blocks_of(n_unique, kGB)), dim3(kGB), 0, st, start,
                       n_unique, n, counts);
    return hipGetLastError();
}

hipError_t grid_decode(const uint64_t *ukeys, int n_unique, float *uvw, hipStream_t st) {
    hipLaunchKernelGGL(decode_kernel, dim3(blocks_of(n_unique, kGB)), dim3(kGB), 0, st, ukeys,
                       n_unique, uvw);
    return hipGetLastError();
}

hipError_t grid_vis_weights(const uint64_t *vkeys, int64_t n_vis, const uint64_t *ukeys,
                            const int32_t *counts, int n_unique, float *wgt, int32_t *vcount,
                            hipStream_t st) {
    hipLaunchKernelGGL(vis_weight_kernel, dim3(blocks_of(n_vis, kGB)), dim3(kGB), 0, st, vkeys,
                       n_vis, ukeys, counts, n_unique, wgt, vcount);
    return hipGetLastError();
}

int grid_tile() { return kGTile; }

}  // namespace vrs

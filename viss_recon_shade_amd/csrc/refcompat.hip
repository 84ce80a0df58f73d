// The reference's own shade predicate, kept as a compatibility stage: the
// antipodal (theta, phi) index box of each uvw on a RES x RES grid
// (ShadeMatrix, shade10M.cu:73-105, after Cart2SphKernel, shade10M.cu:62-70,
// and the mean-|uvw| set-up of shade10M.cu:218-224).  The reference evaluates
// it in FP32 with CUDA's atan2f / asinf; here the angles are FP64 and each grid
// index is rounded once with a tie rule that an independent host evaluation
// shares, so GPU and host agree on every row.  The geometric cap mask of kernels.hip is the physical one;
// this stage exists to produce shadeM files interchangeable with the
// reference pipeline.
#include <hip/hip_runtime.h>

#include "refcompat.hpp"

namespace vrs {

namespace {

constexpr int kBoxBlock = 256;
constexpr int kBoxPartials = 256;  // blocks of the |uvw| sum: a fixed order of FP64 additions

// |uvw| in FP64: the products of FP32 inputs are exact in FP64, so fused or
// separate multiply-adds round the same way; sqrt is correctly rounded.
__device__ __forceinline__ double row_norm(float u, float v, float w) {
    const double uu = (double)u * u, vv = (double)v * v, ww = (double)w * w;
    return sqrt((uu + vv) + ww);
}

// Nearest integer of x >= 0; values within 1e-9 of a half-integer go up (the
// direction roundf takes at an exact tie).  FP64 atan2 / asin / atan of two
// libraries differ by an ulp or two, ~1e-12 here in units of the grid step, so
// both sides of a comparison take the same branch.
__device__ __forceinline__ int round_shared(double x) {
    const double f = floor(x), d = x - f;
    if (fabs(d - 0.5) <= 1e-9) return (int)f + 1;
    return (int)f + (d > 0.5 ? 1 : 0);
}

// Per-block FP64 partial sums of |uvw| in a fixed order: thread t of block b
// adds rows b * 256 + t + k * 65536 in turn, then a halving tree in the block.
__global__ void __launch_bounds__(kBoxBlock) uvw_norm_partial_kernel(const float *__restrict__ uvw,
                                                                     int64_t n,
                                                                     double *__restrict__ partial) {
    __shared__ double s[kBoxBlock];
    double acc = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * kBoxBlock + threadIdx.x; i < n;
         i += (int64_t)kBoxPartials * kBoxBlock)
        acc += row_norm(uvw[3 * i], uvw[3 * i + 1], uvw[3 * i + 2]);
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int o = kBoxBlock / 2; o > 0; o >>= 1) {
        if (threadIdx.x < o) s[threadIdx.x] += s[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = s[0];
}

__global__ void __launch_bounds__(kBoxBlock) shade_box_kernel(const float *__restrict__ uvw,
                                                              int64_t n, int res, float r_ref,
                                                              const double *__restrict__ partial,
                                                              int32_t *__restrict__ out4) {
    __shared__ int half_width;
    if (threadIdx.x == 0) {
        double sum = 0.0;
        for (int i = 0; i < kBoxPartials; ++i) sum += partial[i];
        const double mean_r = sum / (double)n;
        const double R = (double)r_ref;
        const double r0 = atan(R / (mean_r + R));
        half_width = round_shared(r0 / (2.0 * M_PI / res));
    }
    __syncthreads();
    const double dtheta = 2.0 * M_PI / res, dphi = M_PI / res;
    const int thetaR = half_width;
    const int phiR = thetaR / 2;  // integer division, as in the reference
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const float u = uvw[3 * i], v = uvw[3 * i + 1], w = uvw[3 * i + 2];
        const double r = row_norm(u, v, w);
        const double th = atan2((double)v, (double)u);
        const double ph = r > 0.0 ? asin((double)w / r) : 0.0;  // the zero row has no direction
        // antipode of the baseline direction: theta -> 2 pi - (theta + pi), phi -> pi - (phi + pi/2)
        const int ti = round_shared((2.0 * M_PI - (th + M_PI)) / dtheta);
        const int pi_ = round_shared((M_PI - (ph + M_PI / 2)) / dphi);
        int tS = ti - thetaR, tE = ti + thetaR, pS = pi_ - phiR, pE = pi_ + phiR;
        if (tS <= 0) tS += res;
        if (tE > res) tE -= res;
        if (pS <= 0) pS += res;
        if (pE > res) pE -= res;
        int4 o;
        o.x = tS; o.y = tE; o.z = pS; o.w = pE;
        reinterpret_cast<int4 *>(out4)[i] = o;
    }
}

}  // namespace

int shade_box_partials() { return kBoxPartials; }

hipError_t launch_shade_box(const float *uvw, int64_t n, int res, float r_ref, double *partial,
                            int32_t *out4, hipStream_t st) {
    hipLaunchKernelGGL(uvw_norm_partial_kernel, dim3(kBoxPartials), dim3(kBoxBlock), 0, st, uvw, n,
                       partial);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    int64_t blocks = (n + kBoxBlock - 1) / kBoxBlock;
    if (blocks > 65535) blocks = 65535;
    hipLaunchKernelGGL(shade_box_kernel, dim3((unsigned)blocks), dim3(kBoxBlock), 0, st, uvw, n, res,
                       r_ref, (const double *)partial, out4);
    return hipGetLastError();
}

}  // namespace vrs

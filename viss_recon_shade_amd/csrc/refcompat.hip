// The reference's own shade predicate, kept as a compatibility stage: the
// antipodal (theta, phi) index box of each uvw on a RES x RES grid
// (ShadeMatrix, shade10M.cu:73-105, after Cart2SphKernel, shade10M.cu:62-70,
// and the mean-|uvw| set-up of shade10M.cu:218-224), evaluated in FP32 as the
// reference does.  The geometric cap mask of kernels.hip is the physical one;
// this stage exists to produce shadeM files interchangeable with the
// reference pipeline.
#include <hip/hip_runtime.h>

#include "refcompat.hpp"

namespace vrs {

namespace {

constexpr int kBoxBlock = 256;

// Per-block FP32 partial sums of |uvw|.
__global__ void __launch_bounds__(kBoxBlock) uvw_norm_partial_kernel(const float *__restrict__ uvw,
                                                                     int64_t n,
                                                                     float *__restrict__ partial) {
    __shared__ float s[kBoxBlock];
    float acc = 0.0f;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const float u = uvw[3 * i], v = uvw[3 * i + 1], w = uvw[3 * i + 2];
        acc += sqrtf(u * u + v * v + w * w);
    }
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
                                                              const float *__restrict__ partial,
                                                              int n_partial,
                                                              int32_t *__restrict__ out4) {
    __shared__ float mean_r;
    if (threadIdx.x == 0) {
        float sum = 0.0f;
        for (int i = 0; i < n_partial; ++i) sum += partial[i];
        mean_r = sum / (float)n;
    }
    __syncthreads();
    const float r0 = atanf(r_ref / (mean_r + r_ref));
    const float dtheta = (float)(2.0 * M_PI / res);
    const float dphi = (float)(M_PI / res);
    const int thetaR = (int)roundf(r0 / dtheta);
    const int phiR = (int)roundf((float)(thetaR / 2));  // integer division, as in the reference
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const float u = uvw[3 * i], v = uvw[3 * i + 1], w = uvw[3 * i + 2];
        const float r = sqrtf(u * u + v * v + w * w);
        float th = atan2f(v, u);
        float ph = asinf(w / r);
        // antipode of the baseline direction
        th += (float)M_PI;
        ph += (float)(M_PI / 2);
        th = (float)(2 * M_PI) - th;
        ph = (float)M_PI - ph;
        const int ti = (int)roundf(th / dtheta);
        const int pi_ = (int)roundf(ph / dphi);
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

int shade_box_partials() { return 256; }

hipError_t launch_shade_box(const float *uvw, int64_t n, int res, float r_ref, float *partial,
                            int32_t *out4, hipStream_t st) {
    const int np = shade_box_partials();
    hipLaunchKernelGGL(uvw_norm_partial_kernel, dim3(np), dim3(kBoxBlock), 0, st, uvw, n, partial);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    int64_t blocks = (n + kBoxBlock - 1) / kBoxBlock;
    if (blocks > 65535) blocks = 65535;
    hipLaunchKernelGGL(shade_box_kernel, dim3((unsigned)blocks), dim3(kBoxBlock), 0, st, uvw, n, res,
                       r_ref, (const float *)partial, np, out4);
    return hipGetLastError();
}

}  // namespace vrs

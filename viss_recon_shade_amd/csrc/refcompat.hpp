// Reference-compatible stages (refcompat.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace vrs {
int shade_box_partials();
// uvw: [n][3] FP32 device; partial: [shade_box_partials()] FP64 device scratch; out4: [n][4] device.
hipError_t launch_shade_box(const float *uvw, int64_t n, int res, float r_ref, double *partial,
                            int32_t *out4, hipStream_t st);
}  // namespace vrs

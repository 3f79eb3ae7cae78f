// 8-bit scalar quantization: device helpers shared by sq.hip and the search kernel, and
// the launch interface.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "dkm_device.h"

namespace dkm {

// Code rows are padded with zero bytes to a multiple of 16 so each lane reads whole
// 16-byte chunks; the padding adds nothing to the integer sums.
__host__ __device__ inline int64_t sq_code_stride(int64_t dim) { return (dim + 15) & ~int64_t(15); }

// Integer sums of one 16-byte chunk of two code rows: ab += a.b, aa += a.a, bb += b.b.
__device__ __forceinline__ void sq_dot16(const uint4 &a, const uint4 &b, uint32_t &ab,
                                         uint32_t &aa, uint32_t &bb) {
    ab = __builtin_amdgcn_udot4(a.x, b.x, ab, false);
    ab = __builtin_amdgcn_udot4(a.y, b.y, ab, false);
    ab = __builtin_amdgcn_udot4(a.z, b.z, ab, false);
    ab = __builtin_amdgcn_udot4(a.w, b.w, ab, false);
    aa = __builtin_amdgcn_udot4(a.x, a.x, aa, false);
    aa = __builtin_amdgcn_udot4(a.y, a.y, aa, false);
    aa = __builtin_amdgcn_udot4(a.z, a.z, aa, false);
    aa = __builtin_amdgcn_udot4(a.w, a.w, aa, false);
    bb = __builtin_amdgcn_udot4(b.x, b.x, bb, false);
    bb = __builtin_amdgcn_udot4(b.y, b.y, bb, false);
    bb = __builtin_amdgcn_udot4(b.z, b.z, bb, false);
    bb = __builtin_amdgcn_udot4(b.w, b.w, bb, false);
}

// The f32 finish of CompensatedSquaredL2 / CompensatedIP / CompensatedCosineNormalized
// (diskann-quantization/src/scalar/vectors.rs:220-245, 324-345, 424-436) as distances:
// bss = ((1/255) * (1/255)) * scale^2, l2 = sum (a - b)^2, ab = sum a b.
__device__ __forceinline__ float sq_finish(int metric, float bss, float ssn, uint32_t ab,
                                           uint32_t l2, float comp_a, float comp_b) {
    if (metric == METRIC_IP) return -(fma_(bss, (float)ab, ssn) + (comp_b + comp_a));
    const float v = bss * (float)l2;
    if (metric == METRIC_COSNORM) return 1.0f - (1.0f - v / 2.0f);
    return v;
}

// (1/255)^2 * scale^2 in the reference's f32 order.
__host__ __device__ inline float sq_bss(float scale) {
    const float inv = 1.0f / 255.0f;
    const float bs = inv * inv;
    return bs * (scale * scale);
}

hipError_t launch_sq_train(const float *x, int64_t n, int dim, int64_t stride, double stdevs,
                           double *mean, double *var, double *norms, float *scale,
                           float *shift, float *mean_norm, hipStream_t st);
hipError_t launch_sq_shift_norm(const float *shift, int dim, float *out, hipStream_t st);
hipError_t launch_sq_encode(const float *x, int64_t stride, int64_t n, int dim, float scale,
                            const float *shift, uint8_t *codes, int64_t cstride, float *comp,
                            uint32_t *nan_flag, hipStream_t st);
hipError_t launch_sq_pairs(int metric, float bss, float ssn, const uint8_t *left,
                           const float *left_comp, const uint8_t *right,
                           const float *right_comp, int64_t cstride, const uint32_t *ids,
                           int64_t nl, int64_t m, float *out, hipStream_t st);

}  // namespace dkm

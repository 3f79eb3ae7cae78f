// Closest centres of the partitioned build (compute_closest_centers).
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace dkm {
// Largest k (centres per point) the kernel keeps in registers.
constexpr int kMaxClosest = 16;
size_t closest_centers_lds();
// out[i * k + t]: the t-th closest of the K centres c (K x dim) to row i of x (n x dim,
// contiguous); cnorm (K floats) receives the centres' squared norms.
hipError_t launch_closest_centers(const float *x, int64_t n, int dim, const float *c, int K,
                                  int k, float *cnorm, uint32_t *out, hipStream_t st);
}  // namespace dkm

// Medoid and exact FP64 brute-force kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace dkm {
hipError_t medoid_device(const float *x, int64_t n, int64_t stride, int dim, float *mean,
                         unsigned long long *best, hipStream_t stream);
hipError_t bruteforce_device(int metric, const float *x, int64_t n, int64_t stride, int dim,
                             const float *queries, int64_t nq, int k, uint32_t *out_ids,
                             double *out_d, hipStream_t stream);
// Plain device-to-device copy (one 16-byte load and store per lane step); `bytes` must
// be a multiple of 16 and both pointers 16-byte aligned.
hipError_t copy_device(void *dst, const void *src, size_t bytes, hipStream_t stream);
}  // namespace dkm

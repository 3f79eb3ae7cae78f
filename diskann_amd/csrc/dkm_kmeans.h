// Large-dimension k-means (FP32 reference assignment, BF16 MFMA assignment, FP64 update).
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace dkm {
hipError_t kmeans_to_bf16(const float *x, int64_t rows, int dim, int64_t stride, int dp,
                          uint16_t *out, hipStream_t st);
hipError_t kmeans_norms(const float *x, int64_t rows, int dim, int64_t stride, float *out,
                        hipStream_t st);
hipError_t kmeans_assign_f32(const float *x, int64_t n, int dim, int64_t stride, const float *c,
                             int K, const float *cnorm, uint32_t *assign, float *best,
                             hipStream_t st);
// Padded dimension of the BF16 copies (the register-resident kernel's step counts up to 768).
int kmeans_bf16_dp(int dim);
size_t kmeans_mfma_lds(int dp, int K);
// Whether the BF16 assignment for (dp, K) takes the register-resident kernel, whose operands
// are in the fragment layout (kmeans_to_bf16_frag) with rows padded to kmeans_bf16_rows.
bool kmeans_bf16_frag(int dp, int K);
int64_t kmeans_bf16_rows(int64_t n, int dp, int K, bool centres);
hipError_t kmeans_to_bf16_frag(const float *x, int64_t rows, int dim, int64_t stride, int dp,
                               int64_t rows_pad, uint16_t *out, hipStream_t st);
hipError_t kmeans_assign_mfma(const uint16_t *xb, int64_t n, int dp, const uint16_t *cb,
                              const float *cnorm, int K, uint32_t *assign, float *best,
                              hipStream_t st);
size_t kmeans_sort_tmp_bytes(int64_t n, int K);
hipError_t kmeans_update(const float *x, int64_t n, int dim, int64_t stride,
                         const uint32_t *assign, int K, uint64_t *keys, uint64_t *keys_tmp,
                         void *sort_tmp, size_t sort_bytes, int64_t *start, float *c,
                         hipStream_t st);
}  // namespace dkm

// Candidate pre-scoring (BF16 or FP8 MFMA approximate distances, exact FP32 rescore).
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

// Largest sample of base rows the per-query threshold is read from (LDS of the selection).
#define PS_SAMPLE_MAX 15360
// Passing (row, query) pairs each filter wave stages in LDS before appending them
// (a multiple of 64).
#define PS_POOL 1536
// Query q's list counter sits at count[q * PS_COUNT_STRIDE]: one 128-byte line per query, so
// the filter's atomics on different queries never contend for a line.
#define PS_COUNT_STRIDE 32

namespace dkm {
bool ps_supported_dp(int dp);
// Padded dimension of the FP8 operands (a multiple of 128 up to 768), -1 when dim > 768.
int ps_fp8_dp(int dim);
size_t ps_filter_lds(int dp, bool fp8, int nqp);
hipError_t ps_dense(int dp, bool fp8, const uint16_t *xf, int64_t ns, const float *xn,
                    const uint16_t *qf, int nq, int nqp, bool l2, float inv_s, float *out,
                    int64_t ld, hipStream_t st);
hipError_t ps_select(const float *dense, int64_t ld, int ns, int j0, int j1, int j2, float *tau,
                     int nq, hipStream_t st);
hipError_t ps_filter(int dp, bool fp8, const uint16_t *xf, int64_t n, const float *xn,
                     const uint16_t *qf, int nq, int nqp, bool l2, float s, const float *tau,
                     const uint32_t *qmap, uint32_t *count, uint64_t *lists, int64_t cap,
                     hipStream_t st);
hipError_t ps_to_fp8_frag(const float *x, int64_t rows, int dim, int64_t stride, int dp,
                          int64_t rows_pad, float scale, uint8_t *out, hipStream_t st);
hipError_t ps_absmax(const float *x, int64_t rows, int dim, int64_t stride, uint32_t *out,
                     hipStream_t st);
hipError_t ps_norms(const float *x, int64_t n, int dim, const int64_t *rows, float *out,
                    hipStream_t st);
hipError_t ps_gather_rows(const float *x, int dim, const int64_t *rows, int64_t ns, float *out,
                          hipStream_t st);
hipError_t ps_sort_tmp_bytes(int64_t total, int segments, size_t *bytes);
hipError_t ps_sort(void *tmp, size_t tmp_bytes, const uint64_t *in, uint64_t *out, int64_t total,
                   int segments, const uint32_t *begin, const uint32_t *end, hipStream_t st);
// Rows of the C smallest keys of each query's list (count[q * PS_COUNT_STRIDE] entries at
// lists + q * cap, count >= C), in no particular order.
hipError_t ps_take_smallest(const uint64_t *lists, const uint32_t *count, int nq, int64_t cap,
                            int C, uint32_t *ids, hipStream_t st);
hipError_t ps_final_keys(const float *d, const uint32_t *ids, int64_t total, uint64_t *keys,
                         hipStream_t st);
hipError_t ps_output(const uint64_t *sorted, int nq, int C, int k, uint32_t *ids, float *dists,
                     hipStream_t st);
// Final top k (k <= 1024) of each query's C rescored keys, sorted: the same output as a
// segmented sort of the keys followed by ps_output.
bool ps_topk_supported(int k);
hipError_t ps_topk(const uint64_t *keys, int nq, int C, int k, uint32_t *ids, float *dists,
                   hipStream_t st);
}  // namespace dkm

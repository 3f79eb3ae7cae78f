// Launch helpers for the distance / PQ / k-means kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace dkm {

// out[i*m + j] = distance(q_i, base[ids[i*m + j]])
hipError_t launch_distance_batch(int metric, int dim, const float *queries, int64_t nq,
                                 const float *base, int64_t base_stride, const uint32_t *ids,
                                 int64_t m, float *out, hipStream_t stream);

// lut[q][m*256 + c]; ip -> negated inner-product table (TableIP), else SquaredL2.
hipError_t launch_pq_lut(bool ip, const float *pivots, int dim, const int32_t *offsets, int M,
                         const float *queries, int64_t nq, float *lut, hipStream_t stream);

// Direct PQ distances, out[i*m + j] = d(left_i, pivots selected by codes[ids[i*m + j]]):
// left_i is the full-precision row queries[i] (query_codes == nullptr) or the pivots of
// query_codes[i] (queries == nullptr). Resumable f32x8 accumulation over the chunks.
hipError_t launch_pq_direct(int metric, const float *pivots, int dim, const int32_t *offsets,
                            int M, const float *queries, const uint8_t *query_codes, int64_t nq,
                            const uint8_t *codes, const uint32_t *ids, int64_t m, float *out,
                            hipStream_t stream);

// codes[i*code_stride + m]; *bad_row receives the smallest failing row (init 0xFFFFFFFF).
hipError_t launch_pq_encode(const float *pivots, int K, int dim, const int32_t *offsets, int M,
                            const float *x, int64_t x_stride, int64_t n, uint8_t *codes,
                            int64_t code_stride, uint32_t *bad_row, hipStream_t stream);

// Full PQ training on the device (host data in, host pivots out). draws: M x K uniforms in
// [0, 1) for k-means++; chunk_status: 2 x M words (failure reason, centres picked).
hipError_t pq_train_device(const float *data, int64_t n, int dim, const int32_t *offsets, int M,
                           int K, int iters, const double *draws, float *pivots,
                           uint32_t *chunk_status);
// Draw `ctr` of chunk `chunk` of the library's k-means++ stream for `seed`.
double pq_train_draw(uint64_t seed, int chunk, uint64_t ctr);

}  // namespace dkm

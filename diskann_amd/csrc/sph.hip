// 1-bit spherical quantization kernels (diskann-quantization/src/spherical):
//   encoding      CompressIntoWith<DataMut<1>> (spherical/quantizer.rs:591-642, 779-857)
//   distances     data x data, and query x data for the same-as-data, 4-bit transposed and
//                 full-precision query layouts (spherical/vectors.rs:476-897, iface.rs
//                 fused_query_computer)
//   training      the centroid / mean norm of SphericalQuantizer::train (quantizer.rs:225-320)
// One wave prepares one vector (dkm_sph.h); bit rows are ballots of t > 0.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "dkm_sph.h"

namespace dkm {

__global__ __launch_bounds__(64) void sph_shift_sq_kernel(const float *shift, int dim, float *out) {
    const float v = wave_ip(shift, shift, dim);
    if (threadIdx.x == 0) *out = v;
}

// One block (one wave) per row: stored bit rows (cstride bytes, zero padded) + SphMeta
// and / or canonical rows (bits, f16 correction, f16 metric term, u16 bit sum).
__global__ __launch_bounds__(64) void sph_encode_kernel(SphParams q, const float *x,
                                                        int64_t stride, int64_t n,
                                                        uint8_t *bits, SphMeta *meta,
                                                        uint8_t *canon, uint32_t *err) {
    extern __shared__ __align__(16) float lds[];
    const int64_t r = blockIdx.x;
    if (r >= n) return;
    const int lane = threadIdx.x;
    const int nb = (q.out_dim + 7) / 8;
    uint8_t *brow = bits ? bits + r * q.cstride : nullptr;
    uint8_t *crow = canon ? canon + r * sph_canon_bytes(q.out_dim) : nullptr;
    float sn = 0.0f, ms = 0.0f;
    uint32_t e = 0;
    const float *t = sph_prepare(q, x + r * stride, lds, sn, ms, e);
    const uint32_t bitsum = sph_write_bits(q, e ? nullptr : t, brow, crow);
    uint16_t h[3];
    const SphMeta m = sph_data_meta(q, e ? nullptr : t, sn, ms, bitsum, h, e);
    if (brow)
        for (int i = nb + lane; i < q.cstride; i += 64) brow[i] = 0;
    if (lane == 0) {
        if (meta) meta[r] = m;
        if (crow)
            for (int k = 0; k < 3; ++k) {
                crow[nb + 2 * k] = (uint8_t)(h[k] & 0xFFu);
                crow[nb + 2 * k + 1] = (uint8_t)(h[k] >> 8);
            }
        if (e) atomicOr(err, e);
    }
}

__global__ __launch_bounds__(256) void sph_pairs_kernel(SphParams q, const uint8_t *lbits,
                                                        const SphMeta *lmeta,
                                                        const uint8_t *rbits,
                                                        const SphMeta *rmeta,
                                                        const uint32_t *ids, int64_t nl,
                                                        int64_t m, float *out) {
    const int64_t pair = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (pair >= nl * m) return;
    const int64_t i = pair / m;
    const uint32_t id = ids[pair];
    const uint64_t *a = (const uint64_t *)(lbits + i * q.cstride);
    const uint64_t *b = (const uint64_t *)(rbits + (int64_t)id * q.cstride);
    uint32_t ip = 0;
    for (int w = 0; w < q.cstride / 8; ++w) ip += (uint32_t)__builtin_popcountll(a[w] & b[w]);
    out[pair] = sph_data_data(q.metric, q.shift_sq, (float)q.out_dim, ip, lmeta[i], rmeta[id]);
}

// InnerProduct(&[f32], USlice<1>) in the x86-64-v3 order (bits/distances.rs:2295-2436):
// two f32x8 accumulators alternate over the four 8-lane groups of each 32-element block,
// are combined, the remainder's groups go into the combined vector, then sum_tree.
__device__ inline float sph_ip_full(const float *v, const uint8_t *bits, int len) {
    float s[8], s0[8], s1[8];
#pragma unroll
    for (int l = 0; l < 8; ++l) s[l] = s0[l] = s1[l] = 0.0f;
    const int blocks = len / 32;
    for (int i = 0; i < blocks; ++i) {
        const uint32_t w = (uint32_t)bits[4 * i] | ((uint32_t)bits[4 * i + 1] << 8) |
                           ((uint32_t)bits[4 * i + 2] << 16) | ((uint32_t)bits[4 * i + 3] << 24);
#pragma unroll
        for (int g = 0; g < 4; ++g)
#pragma unroll
            for (int l = 0; l < 8; ++l) {
                const float bv = (w >> (8 * g + l)) & 1u ? 1.0f : 0.0f;
                if (g & 1) s1[l] = fma_(v[32 * i + 8 * g + l], bv, s1[l]);
                else s0[l] = fma_(v[32 * i + 8 * g + l], bv, s0[l]);
            }
    }
    if (blocks > 0)
#pragma unroll
        for (int l = 0; l < 8; ++l) s[l] = s0[l] + s1[l];
    const int rem = len % 32;
    if (rem) {
        const int tail = len % 8 == 0 ? 8 : len % 8;
        const int nv = rem <= 8 ? 1 : rem <= 16 ? 2 : rem <= 24 ? 3 : 4;
        for (int g = 0; g < nv; ++g) {
            const int cnt = g == nv - 1 ? tail : 8;
            for (int l = 0; l < 8; ++l) {
                const int e = 32 * blocks + 8 * g + l;
                const float xv = l < cnt ? v[e] : 0.0f;
                const float bv = l < cnt && ((bits[e >> 3] >> (e & 7)) & 1) ? 1.0f : 0.0f;
                s[l] = fma_(xv, bv, s[l]);
            }
        }
    }
    const float u0 = s[0] + s[4], u1 = s[1] + s[5], u2 = s[2] + s[6], u3 = s[3] + s[7];
    return (u0 + u2) + (u1 + u3);
}

// Query x data distances for one query per block: the query is compressed in the
// block's layout (after the inner-product rescale when asked), then the block's lanes
// take the query's m pairs.
__global__ __launch_bounds__(64) void sph_query_pairs_kernel(SphParams q, int layout, int rescale,
                                                             const float *queries, int64_t nq,
                                                             const uint8_t *bits,
                                                             const SphMeta *meta,
                                                             const uint32_t *ids, int64_t m,
                                                             float *out, uint32_t *err) {
    extern __shared__ __align__(16) float lds[];
    const int64_t qi = blockIdx.x;
    if (qi >= nq) return;
    const int lane = threadIdx.x;
    uint8_t *qb = (uint8_t *)(lds + sph_query_lds_floats(q));  // query code
    __shared__ float qm[4];
    uint32_t e = 0;
    sph_query_code(q, layout, rescale != 0, queries + qi * q.dim, lds, qb, qm, e);
    const SphMeta qdm = {qm[0], qm[1], qm[2], qm[3]};
    if (lane == 0 && e) atomicOr(err, e);
    const float dim = (float)q.out_dim;
    for (int64_t j = lane; j < m; j += 64) {
        const uint32_t id = ids[qi * m + j];
        const uint8_t *row = bits + (int64_t)id * q.cstride;
        const SphMeta y = meta[id];
        float d;
        if (layout == SPH_QUERY_SAME_AS_DATA) {
            uint32_t ip = 0;
            for (int w = 0; w < q.cstride / 8; ++w)
                ip += (uint32_t)__builtin_popcountll(((const uint64_t *)qb)[w] &
                                                     ((const uint64_t *)row)[w]);
            d = sph_data_data(q.metric, q.shift_sq, dim, ip, qdm, y);
        } else if (layout == SPH_QUERY_FOUR_BIT) {
            uint32_t ip = 0;
            for (int w = 0; w < q.cstride / 8; ++w) {
                const uint64_t b = ((const uint64_t *)row)[w];
#pragma unroll
                for (int k = 0; k < 4; ++k)
                    ip += (uint32_t)__builtin_popcountll(((const uint64_t *)(qb + k * q.cstride))[w] & b) << k;
            }
            d = sph_query4_data(q.metric, q.shift_sq, dim, ip, qm, y);
        } else {
            const float s = sph_ip_full((const float *)qb, row, q.out_dim);
            if (q.metric == METRIC_L2) {
                d = qm[2] + y.ms - 2.0f * qm[1] * y.ipc * (s - qm[0] * 0.5f);
            } else {
                const float v = qm[1] * y.ipc * (s - qm[0] * 0.5f) + qm[2] + y.ms + q.shift_sq;
                d = q.metric == METRIC_COSINE ? 1.0f - v : -v;
            }
        }
        out[qi * m + j] = d;
    }
}

// Traversal queries compressed ahead of the search (one block per query, with the
// inner-product rescale of the reference's search strategy): record q holds the layout's
// bit row / planes and the 4 metadata floats.
__global__ __launch_bounds__(64) void sph_query_prep_kernel(SphParams q, int layout, int rescale,
                                                            const float *queries, int64_t qstride,
                                                            int64_t nq, uint8_t *records) {
    extern __shared__ __align__(16) float lds[];
    const int64_t qi = blockIdx.x;
    if (qi >= nq) return;
    uint8_t *qb = (uint8_t *)(lds + sph_query_lds_floats(q));
    __shared__ float qm[4];
    uint32_t e = 0;
    sph_query_code(q, layout, rescale != 0, queries + qi * qstride, lds, qb, qm, e);
    uint8_t *rec = records + qi * sph_query_record_bytes(q);
    const int nbytes = layout == SPH_QUERY_SAME_AS_DATA ? q.cstride : 4 * q.cstride;
    for (int i = threadIdx.x; i < nbytes / 16; i += 64)
        ((uint4 *)rec)[i] = ((const uint4 *)qb)[i];
    if (threadIdx.x < 4) ((float *)(rec + 4 * q.cstride))[threadIdx.x] = qm[threadIdx.x];
}

// ---- training statistics ----------------------------------------------------------------

__global__ __launch_bounds__(256) void sph_row_norm_kernel(const float *x, int64_t n, int dim,
                                                           double *norms) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    const float *row = x + r * dim;
    double acc = 0.0;
    for (int j = 0; j < dim; ++j) {
        const double v = (double)row[j];
        acc += v * v;
    }
    norms[r] = __builtin_sqrt(acc);
}

// Column means in row order; cosine: rows weighted by 1 / norm (1 for a zero row).
__global__ __launch_bounds__(64) void sph_column_mean_kernel(const float *x, int64_t n, int dim,
                                                             const double *norms, int cosine,
                                                             double *mean) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= dim) return;
    double acc = 0.0;
    for (int64_t r = 0; r < n; ++r) {
        const double v = (double)x[r * dim + j];
        if (cosine) {
            const double nr = norms[r];
            acc += v * (nr == 0.0 ? 1.0 : 1.0 / nr);
        } else {
            acc += v;
        }
    }
    mean[j] = acc / (double)n;
}

__global__ __launch_bounds__(64) void sph_finish_stats_kernel(const double *mean, const double *norms,
                                                              int64_t n, int dim, int cosine,
                                                              float *centroid, float *mean_norm) {
    for (int j = threadIdx.x; j < dim; j += blockDim.x) centroid[j] = (float)mean[j];
    if (threadIdx.x == 0) {
        double s = 0.0;
        if (!cosine)
            for (int64_t r = 0; r < n; ++r) s += norms[r];
        *mean_norm = cosine ? 1.0f : (float)(s / (double)n);
    }
}

// ---- launchers ------------------------------------------------------------------------

hipError_t launch_sph_shift_sq(const float *shift, int dim, float *out, hipStream_t st) {
    hipLaunchKernelGGL(sph_shift_sq_kernel, dim3(1), dim3(64), 0, st, shift, dim, out);
    return hipGetLastError();
}

hipError_t launch_sph_encode(const SphParams &q, const float *x, int64_t stride, int64_t n,
                             uint8_t *bits, SphMeta *meta, uint8_t *canon, uint32_t *err,
                             hipStream_t st) {
    if (n <= 0) return hipSuccess;
    const size_t lds = sizeof(float) * (size_t)sph_lds_floats(q);
    hipError_t e = hipFuncSetAttribute((const void *)sph_encode_kernel,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(sph_encode_kernel, dim3((unsigned)n), dim3(64), lds, st, q, x, stride, n,
                       bits, meta, canon, err);
    return hipGetLastError();
}

hipError_t launch_sph_pairs(const SphParams &q, const uint8_t *lbits, const SphMeta *lmeta,
                            const uint8_t *rbits, const SphMeta *rmeta, const uint32_t *ids,
                            int64_t nl, int64_t m, float *out, hipStream_t st) {
    const int64_t total = nl * m;
    if (total == 0) return hipSuccess;
    hipLaunchKernelGGL(sph_pairs_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st,
                       q, lbits, lmeta, rbits, rmeta, ids, nl, m, out);
    return hipGetLastError();
}

hipError_t launch_sph_query_pairs(const SphParams &q, int layout, int rescale, const float *queries,
                                  int64_t nq, const uint8_t *bits, const SphMeta *meta,
                                  const uint32_t *ids, int64_t m, float *out, uint32_t *err,
                                  hipStream_t st) {
    if (nq <= 0) return hipSuccess;
    const size_t lds = sizeof(float) * (size_t)sph_query_lds_floats(q) +
                       (size_t)sph_query_code_bytes(q) + 64;
    hipError_t e = hipFuncSetAttribute((const void *)sph_query_pairs_kernel,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(sph_query_pairs_kernel, dim3((unsigned)nq), dim3(64), lds, st, q, layout,
                       rescale, queries, nq, bits, meta, ids, m, out, err);
    return hipGetLastError();
}

hipError_t launch_sph_query_prep(const SphParams &q, int layout, int rescale, const float *queries,
                                 int64_t qstride, int64_t nq, uint8_t *records, hipStream_t st) {
    if (nq <= 0) return hipSuccess;
    const size_t lds = sizeof(float) * (size_t)sph_query_lds_floats(q) + 4 * (size_t)q.cstride + 64;
    hipError_t e = hipFuncSetAttribute((const void *)sph_query_prep_kernel,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(sph_query_prep_kernel, dim3((unsigned)nq), dim3(64), lds, st, q, layout,
                       rescale, queries, qstride, nq, records);
    return hipGetLastError();
}

hipError_t launch_sph_train_stats(int metric, const float *x, int64_t n, int dim, double *mean,
                                  double *norms, float *centroid, float *mean_norm,
                                  hipStream_t st) {
    const int cosine = metric == METRIC_COSINE;
    hipLaunchKernelGGL(sph_row_norm_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, x,
                       n, dim, norms);
    hipLaunchKernelGGL(sph_column_mean_kernel, dim3((dim + 63) / 64), dim3(64), 0, st, x, n, dim,
                       (const double *)norms, cosine, mean);
    hipLaunchKernelGGL(sph_finish_stats_kernel, dim3(1), dim3(64), 0, st, (const double *)mean,
                       (const double *)norms, n, dim, cosine, centroid, mean_norm);
    return hipGetLastError();
}

}  // namespace dkm

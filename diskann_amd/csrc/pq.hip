// Distance-batch, PQ lookup-table, PQ encoding and PQ training kernels (gfx950).
//
// Reference semantics restated here:
//   populate_chunk_distances_impl  diskann-providers/src/model/pq/fixed_chunk_pq_table.rs:149-184
//   BasicTable::compress_into      diskann-quantization/src/product/tables/basic.rs:161-193
//   LightPQTrainingParameters::train  diskann-quantization/src/product/train.rs:109-231
//   kmeans_plusplus_into_inner     diskann-quantization/src/algorithms/kmeans/plusplus.rs:381-498
//   distances_in_place / update_centroids / lloyds_inner
//                                  diskann-quantization/src/algorithms/kmeans/lloyds.rs:28-426
//   square_norm                    diskann-quantization/src/algorithms/kmeans/common.rs:9-65
// k-means++ draws its samples from a counter-based SplitMix64 stream of this library
// (the reference uses rand/ChaCha12), so seeded pivots differ from the reference's while
// every distance, argmin and centroid update follows the reference's arithmetic.
#include <vector>

#include "dkm_device.h"
#include "dkm_kernels.h"

namespace dkm {

// ---------------------------------------------------------------------------------
// distance batch: one half-wave per (query, candidate) pair
// ---------------------------------------------------------------------------------
template <int METRIC>
__global__ __launch_bounds__(256) void distance_batch_kernel(int dim, const float *queries,
                                                             int64_t nq, const float *base,
                                                             int64_t base_stride,
                                                             const uint32_t *ids, int64_t m,
                                                             float *out) {
    const int64_t pair = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 5;
    const int p = threadIdx.x & 31;
    const int64_t total = nq * m;
    const bool ok = pair < total;
    const int64_t pr = ok ? pair : 0;
    const int64_t qi = pr / m;
    const float *q = queries + qi * dim;
    const float *x = base + (int64_t)ids[pr] * base_stride;
    float d = half_distance<METRIC>(q, x, dim, p);
    if (ok && p == 0) out[pr] = d;
}

hipError_t launch_distance_batch(int metric, int dim, const float *queries, int64_t nq,
                                 const float *base, int64_t base_stride, const uint32_t *ids,
                                 int64_t m, float *out, hipStream_t stream) {
    const int64_t pairs = nq * m;
    dim3 grid((unsigned)((pairs * 32 + 255) / 256));
    switch (metric) {
    case METRIC_L2:
        hipLaunchKernelGGL(distance_batch_kernel<METRIC_L2>, grid, dim3(256), 0, stream, dim,
                           queries, nq, base, base_stride, ids, m, out);
        break;
    case METRIC_IP:
        hipLaunchKernelGGL(distance_batch_kernel<METRIC_IP>, grid, dim3(256), 0, stream, dim,
                           queries, nq, base, base_stride, ids, m, out);
        break;
    case METRIC_COSNORM:
        hipLaunchKernelGGL(distance_batch_kernel<METRIC_COSNORM>, grid, dim3(256), 0, stream, dim,
                           queries, nq, base, base_stride, ids, m, out);
        break;
    default:
        hipLaunchKernelGGL(distance_batch_kernel<METRIC_COSINE>, grid, dim3(256), 0, stream, dim,
                           queries, nq, base, base_stride, ids, m, out);
        break;
    }
    return hipGetLastError();
}

// ---------------------------------------------------------------------------------
// Direct PQ distances (FixedChunkPQTable::{l2_distance, inner_product, cosine_distance,
// qq_*}, diskann-providers/src/model/pq/fixed_chunk_pq_table.rs:32-56, 220-365): eight
// lanes per pair, lane l holding lane l of the f32x8 accumulator. For each chunk the lane
// computes its lane of the simd_op vector over the chunk's slices (chunk-local element o
// into partial o % 32 for L2/IP with Strategy4x1, o % 16 for cosine with Strategy2x4, then
// the masked remainder FMA), adds it to the running accumulator, and the accumulator is
// reduced by sum_tree (shuffles by 4, 2, 1) after the last chunk.
// ---------------------------------------------------------------------------------
template <int METRIC, bool QQ>
__global__ __launch_bounds__(256) void pq_direct_kernel(const float *pivots, int dim,
                                                        const int32_t *offsets, int M,
                                                        const float *queries,
                                                        const uint8_t *query_codes, int64_t nq,
                                                        const uint8_t *codes,
                                                        const uint32_t *ids, int64_t m,
                                                        float *out) {
    const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t total = nq * m;
    const int64_t pair = gid >> 3;
    const int l = (int)(gid & 7);
    const int64_t pr = pair < total ? pair : total - 1;  // all eight lanes stay in the shuffles
    const int64_t qi = pr / m;
    const uint8_t *code = codes + (int64_t)ids[pr] * M;
    const float res = pq_direct8<METRIC>(pivots, dim, offsets, M, QQ ? nullptr : queries + qi * dim,
                                         QQ ? query_codes + qi * M : nullptr, code, l);
    if (pair < total && l == 0) out[pair] = res;
}

template <int METRIC>
static void launch_direct_t(dim3 grid, hipStream_t st, const float *pivots, int dim,
                            const int32_t *offsets, int M, const float *queries,
                            const uint8_t *query_codes, int64_t nq, const uint8_t *codes,
                            const uint32_t *ids, int64_t m, float *out) {
    if (query_codes)
        hipLaunchKernelGGL((pq_direct_kernel<METRIC, true>), grid, dim3(256), 0, st, pivots, dim,
                           offsets, M, queries, query_codes, nq, codes, ids, m, out);
    else
        hipLaunchKernelGGL((pq_direct_kernel<METRIC, false>), grid, dim3(256), 0, st, pivots, dim,
                           offsets, M, queries, query_codes, nq, codes, ids, m, out);
}

hipError_t launch_pq_direct(int metric, const float *pivots, int dim, const int32_t *offsets,
                            int M, const float *queries, const uint8_t *query_codes, int64_t nq,
                            const uint8_t *codes, const uint32_t *ids, int64_t m, float *out,
                            hipStream_t stream) {
    const int64_t pairs = nq * m;
    if (pairs == 0) return hipSuccess;
    dim3 grid((unsigned)((pairs * 8 + 255) / 256));
    if (metric == METRIC_L2)
        launch_direct_t<METRIC_L2>(grid, stream, pivots, dim, offsets, M, queries, query_codes,
                                   nq, codes, ids, m, out);
    else if (metric == METRIC_IP)
        launch_direct_t<METRIC_IP>(grid, stream, pivots, dim, offsets, M, queries, query_codes,
                                   nq, codes, ids, m, out);
    else
        launch_direct_t<METRIC_COSINE>(grid, stream, pivots, dim, offsets, M, queries,
                                       query_codes, nq, codes, ids, m, out);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------------
// PQ lookup tables
// ---------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void pq_lut_kernel(bool ip, const float *pivots, int dim,
                                                     const int32_t *offsets, int M,
                                                     const float *queries, float *lut) {
    extern __shared__ float qs[];
    const int64_t qi = blockIdx.x;
    for (int e = threadIdx.x; e < dim; e += blockDim.x) qs[e] = queries[qi * dim + e];
    __syncthreads();
    for (int e = threadIdx.x; e < M * 256; e += blockDim.x) {
        int m = e >> 8, c = e & 255;
        int o0 = offsets[m], o1 = offsets[m + 1];
        const float *pv = pivots + (size_t)c * dim + o0;
        float v = ip ? -scalar_simd_4x1<true>(qs + o0, pv, o1 - o0)
                     : scalar_simd_4x1<false>(qs + o0, pv, o1 - o0);
        lut[qi * M * 256 + e] = v;
    }
}

hipError_t launch_pq_lut(bool ip, const float *pivots, int dim, const int32_t *offsets, int M,
                         const float *queries, int64_t nq, float *lut, hipStream_t stream) {
    hipLaunchKernelGGL(pq_lut_kernel, dim3((unsigned)nq), dim3(256), sizeof(float) * dim, stream,
                       ip, pivots, dim, offsets, M, queries, lut);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------------
// PQ encoding: block = 256 rows x one chunk; pivot slice of the chunk staged in LDS
// ---------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void pq_encode_kernel(const float *pivots, int K, int dim,
                                                        const int32_t *offsets, const float *x,
                                                        int64_t x_stride, int64_t n,
                                                        uint8_t *codes, int64_t code_stride,
                                                        uint32_t *bad_row) {
    extern __shared__ float sm[];
    const int m = blockIdx.y;
    const int o0 = offsets[m], o1 = offsets[m + 1], ds = o1 - o0;
    float *piv = sm;                 // K * ds
    float *xs = sm + K * ds;         // 256 * ds
    for (int e = threadIdx.x; e < K * ds; e += blockDim.x) {
        int c = e / ds, d = e - c * ds;
        piv[e] = pivots[(size_t)c * dim + o0 + d];
    }
    const int64_t row = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const bool ok = row < n;
    for (int d = 0; d < ds; ++d) xs[threadIdx.x * ds + d] = ok ? x[row * x_stride + o0 + d] : 0.0f;
    __syncthreads();
    if (!ok) return;
    float best = __builtin_inff();
    int arg = 0;
    const float *xv = xs + threadIdx.x * ds;
    for (int c = 0; c < K; ++c) {
        float d = scalar_simd_4x1<false>(xv, piv + c * ds, ds);
        if (d < best) {
            best = d;
            arg = c;
        }
    }
    if (best == __builtin_inff()) atomicMin(bad_row, (uint32_t)row);
    codes[row * code_stride + m] = (uint8_t)arg;
}

hipError_t launch_pq_encode(const float *pivots, int K, int dim, const int32_t *offsets, int M,
                            const float *x, int64_t x_stride, int64_t n, uint8_t *codes,
                            int64_t code_stride, uint32_t *bad_row, hipStream_t stream) {
    // largest chunk width bounds the LDS need
    int maxds = (dim + M - 1) / M;
    size_t lds = sizeof(float) * (size_t)(K + 256) * maxds;
    hipError_t e = hipFuncSetAttribute((const void *)pq_encode_kernel,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    dim3 grid((unsigned)((n + 255) / 256), (unsigned)M);
    hipLaunchKernelGGL(pq_encode_kernel, grid, dim3(256), lds, stream, pivots, K, dim, offsets, x,
                       x_stride, n, codes, code_stride, bad_row);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------------
// PQ training
// ---------------------------------------------------------------------------------

static uint64_t splitmix64(uint64_t x) {
    x += 0x9e3779b97f4a7c15ull;
    x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
    x = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
    return x ^ (x >> 31);
}

// Uniform double in [0, 1) (53-bit grid) for draw `ctr` of chunk `chunk`: the library's
// own counter-based SplitMix64 stream in place of the reference's rand/ChaCha12.
double pq_train_draw(uint64_t seed, int chunk, uint64_t ctr) {
    uint64_t s = splitmix64(seed ^ splitmix64(0x5851f42d4c957f2dull * (uint64_t)(chunk + 1)));
    uint64_t v = splitmix64(s + ctr);
    return (double)(v >> 11) * (1.0 / 9007199254740992.0);
}

// Per-chunk copy of the training rows: xc[(m * n + r) * maxds + d] = x[r, off_m + d], and
// norms[m * n + r] = square_norm of that slice (diskann-quantization/src/algorithms/kmeans/
// common.rs:9-65); the reference takes the same strided view per chunk (train.rs:155-166).
__global__ __launch_bounds__(256) void chunk_rows_kernel(const float *x, int64_t n, int dim,
                                                         const int32_t *offsets, int maxds,
                                                         float *xc, float *norms) {
    const int m = blockIdx.y;
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    const int o0 = offsets[m], ds = offsets[m + 1] - o0;
    float *dst = xc + ((int64_t)m * n + r) * maxds;
    for (int d = 0; d < maxds; ++d) dst[d] = d < ds ? x[r * dim + o0 + d] : 0.0f;
    norms[(int64_t)m * n + r] = square_norm_ref(dst, ds);
}

// rand 0.9 UniformFloat::<f64>::new(0.0, s): an empty range (!(0 < s), NaN included) or a
// non-finite scale fail; otherwise the scale s is lowered one ulp at a time while
// scale * (1 - 2^-52) + 0 >= s, and a draw u in [0, 1) maps to u * scale + 0.
// Returns 0 ok, 1 empty range, 2 non-finite.
__device__ inline int uniform_scale(double s, double *scale) {
    if (!(0.0 < s)) return 1;
    double sc = s - 0.0;
    if (!__builtin_isfinite(sc)) return 2;
    const double max_rand = 1.0 - 2.220446049250313e-16;
    while (sc * max_rand + 0.0 >= s) sc = __longlong_as_double(__double_as_longlong(sc) - 1);
    *scale = sc;
    return 0;
}

constexpr int kKppTile = 4096;

// k-means++ seeding (kmeans_plusplus_into_inner + update_distances, diskann-quantization/
// src/algorithms/kmeans/plusplus.rs:238-498, micro-kernel :84-231), one 256-thread block per
// chunk, draws[m * K + c] standing in for the reference's RNG:
//  - first centre: row floor(draws[0] * n);
//  - each step: distance = (norm_x + norm_c) + dot * -2 with dot a sequential fma over the
//    chunk's dims, min_d lowered where distance < min_d; one thread per 16-row block forms
//    sum_{j<8} ((f64) min_d[j] + (f64) min_d[j + 8]) and one lane adds the blocks in order
//    (the micro-kernel's rolling f64 sum);
//  - threshold = draws[c] * scale(0, s); one lane walks the rows in order with an f64
//    running sum (rows staged through LDS) and picks the first row with running >= threshold,
//    min_d > 0 and not yet picked.
// status[2m] = 0 ok / 1 DatasetTooSmall / 2 InsufficientDiversity / 3 SawInfinity,
// status[2m + 1] = centres picked. Unpicked centres stay zero.
__global__ __launch_bounds__(256) void kmeanspp_kernel(const float *xc, int64_t n,
                                                       const int32_t *offsets, const float *norms,
                                                       int K, const double *draws, float *mind_all,
                                                       uint8_t *picked_all, double *bsum_all,
                                                       float *centers, int maxds, uint32_t *status) {
    const int m = blockIdx.x;
    const int ds = offsets[m + 1] - offsets[m];
    const int t = threadIdx.x;
    const int64_t nblk = (n + 15) / 16;
    const float *xm = xc + (int64_t)m * n * maxds;
    float *mind = mind_all + (int64_t)m * n;
    uint8_t *picked = picked_all + (int64_t)m * n;
    double *bsum = bsum_all + (int64_t)m * nblk;
    const float *nrm = norms + (int64_t)m * n;
    float *cent = centers + (int64_t)m * K * maxds;
    const double *u = draws + (int64_t)m * K;
    __shared__ float cur[64];
    __shared__ float cur_norm;
    __shared__ double thr;
    __shared__ int64_t found;
    __shared__ int state;
    __shared__ int selected;
    __shared__ float tmind[kKppTile];
    __shared__ uint8_t tpick[kKppTile];
    for (int64_t r = t; r < n; r += 256) {
        mind[r] = __builtin_inff();
        picked[r] = 0;
    }
    for (int e = t; e < K * maxds; e += 256) cent[e] = 0.0f;
    __syncthreads();
    if (t == 0) {
        int64_t i0 = (int64_t)(u[0] * (double)n);
        if (i0 >= n) i0 = n - 1;
        if (i0 < 0) i0 = 0;
        picked[i0] = 1;
        for (int d = 0; d < ds; ++d) {
            cur[d] = xm[i0 * maxds + d];
            cent[d] = cur[d];
        }
        cur_norm = nrm[i0];
        state = 0;
        selected = 1;
    }
    __syncthreads();
    const int steps = (int)((int64_t)K < n ? (int64_t)K : n);
    for (int c = 1; c < steps; ++c) {
        for (int64_t b = t; b < nblk; b += 256) {
            float dv[16];
#pragma unroll
            for (int j = 0; j < 16; ++j) {
                const int64_t r = 16 * b + j;
                float md = 0.0f;
                if (r < n) {
                    const float *xr = xm + r * maxds;
                    float dot = 0.0f;
                    for (int d = 0; d < ds; ++d) dot = fma_(xr[d], cur[d], dot);
                    const float dist = (nrm[r] + cur_norm) + dot * -2.0f;
                    md = mind[r];
                    if (dist < md) md = dist;
                    mind[r] = md;
                }
                dv[j] = md;
            }
            double bs = 0.0;
#pragma unroll
            for (int j = 0; j < 8; ++j) bs += (double)dv[j] + (double)dv[j + 8];
            bsum[b] = bs;
        }
        __syncthreads();
        if (t == 0) {
            double s = 0.0;
            int64_t b = 0;
            for (; b + 8 <= nblk; b += 8) {
                double v[8];
#pragma unroll
                for (int j = 0; j < 8; ++j) v[j] = bsum[b + j];
#pragma unroll
                for (int j = 0; j < 8; ++j) s = s + v[j];
            }
            for (; b < nblk; ++b) s = s + bsum[b];
            double scale = 0.0;
            const int us = uniform_scale(s, &scale);
            if (us == 2) state = 3;       // SawInfinity
            else if (us == 1) state = 2;  // empty range: no pick, InsufficientDiversity
            else thr = u[c] * scale + 0.0;
            found = -1;
        }
        __syncthreads();
        if (state) break;
        double rolling = 0.0;  // lane 0's running sum across tiles
        for (int64_t base = 0; base < n; base += kKppTile) {
            const int len = (int)((n - base) < kKppTile ? (n - base) : kKppTile);
            for (int i = t; i < len; i += 256) {
                tmind[i] = mind[base + i];
                tpick[i] = picked[base + i];
            }
            __syncthreads();
            if (t == 0) {
                const double th = thr;
                for (int i = 0; i < len; ++i) {
                    const float md = tmind[i];
                    rolling += (double)md;
                    if (rolling >= th && md > 0.0f && !tpick[i]) {
                        found = base + i;
                        break;
                    }
                }
            }
            __syncthreads();
            if (found >= 0) break;
        }
        if (t == 0) {
            if (found < 0) {
                state = 2;
            } else {
                picked[found] = 1;
                for (int d = 0; d < ds; ++d) {
                    cur[d] = xm[found * maxds + d];
                    cent[c * maxds + d] = cur[d];
                }
                cur_norm = nrm[found];
                selected = c + 1;
            }
        }
        __syncthreads();
        if (state) break;
    }
    if (t == 0) {
        const int reason = state ? state : (selected != K ? 1 : 0);
        status[2 * m] = (uint32_t)reason;
        status[2 * m + 1] = (uint32_t)selected;
    }
}

// Lloyd assignment (distances_in_place, lloyds.rs:28-328): one thread per (row, chunk);
// distance = ((|c|^2 - x.c) - x.c) + |x|^2 with x.c a sequential fma over the chunk's dims,
// the first strictly smaller distance in centre order wins. DS > 0: the chunk width is
// DS (row held in registers); 0: any width.
template <int DS>
__global__ __launch_bounds__(256) void lloyd_assign_kernel(const float *xc, int64_t n,
                                                           const int32_t *offsets,
                                                           const float *norms,
                                                           const float *centers, int K, int maxds,
                                                           uint32_t *assign) {
    extern __shared__ float sm[];
    const int m = blockIdx.y;
    const int ds = DS > 0 ? DS : offsets[m + 1] - offsets[m];
    float *cs = sm;               // K * ds
    float *cn = sm + K * ds;      // K
    const float *cent = centers + (int64_t)m * K * maxds;
    for (int e = threadIdx.x; e < K * ds; e += blockDim.x) {
        int c = e / ds, d = e - c * ds;
        cs[e] = cent[c * maxds + d];
    }
    __syncthreads();
    for (int c = threadIdx.x; c < K; c += blockDim.x) cn[c] = square_norm_ref(cs + c * ds, ds);
    __syncthreads();
    const int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (r >= n) return;
    const float *xr = xc + ((int64_t)m * n + r) * maxds;
    const float xn = norms[(int64_t)m * n + r];
    float best = __builtin_inff();
    uint32_t arg = 0xFFFFFFFFu;
    if constexpr (DS > 0) {
        float xv[DS];
#pragma unroll
        for (int d = 0; d < DS; ++d) xv[d] = xr[d];
        for (int c = 0; c < K; ++c) {
            float dot = 0.0f;
#pragma unroll
            for (int d = 0; d < DS; ++d) dot = fma_(cs[c * DS + d], xv[d], dot);
            const float dist = ((cn[c] - dot) - dot) + xn;
            if (dist < best) {
                best = dist;
                arg = (uint32_t)c;
            }
        }
    } else {
        for (int c = 0; c < K; ++c) {
            float dot = 0.0f;
            for (int d = 0; d < ds; ++d) dot = fma_(cs[c * ds + d], xr[d], dot);
            const float dist = ((cn[c] - dot) - dot) + xn;
            if (dist < best) {
                best = dist;
                arg = (uint32_t)c;
            }
        }
    }
    assign[(int64_t)m * n + r] = arg;
}

// Lloyd update (update_centroids, lloyds.rs:345-366): one thread per (chunk, centre, dim)
// walks every row in order through LDS tiles and adds the rows assigned to its centre in
// f64, so each sum has the reference's row order; centre = (f32)(sum / max(count, 1))
// (empty clusters become zero). Grid: (M, ceil(K * maxds / 256)).
__global__ __launch_bounds__(256) void lloyd_update_kernel(const float *xc, int64_t n,
                                                           const int32_t *offsets,
                                                           const uint32_t *assign, int K,
                                                           int maxds, int tile, float *centers) {
    extern __shared__ float sm[];
    const int m = blockIdx.x;
    const int ds = offsets[m + 1] - offsets[m];
    uint32_t *ta = (uint32_t *)sm;       // tile
    float *tx = sm + tile;               // tile * maxds
    const int idx = blockIdx.y * 256 + threadIdx.x;
    const int c = idx / maxds, d = idx - c * maxds;
    const bool mine = c < K && d < ds;
    double sum = 0.0;
    uint32_t cnt = 0;
    const uint32_t *am = assign + (int64_t)m * n;
    const float *xm = xc + (int64_t)m * n * maxds;
    for (int64_t base = 0; base < n; base += tile) {
        const int len = (int)((n - base) < tile ? (n - base) : tile);
        __syncthreads();
        for (int i = threadIdx.x; i < len; i += 256) ta[i] = am[base + i];
        for (int e = threadIdx.x; e < len * maxds; e += 256) tx[e] = xm[base * maxds + e];
        __syncthreads();
        if (mine) {
            for (int i = 0; i < len; ++i) {
                if (ta[i] == (uint32_t)c) {
                    ++cnt;
                    sum += (double)tx[i * maxds + d];
                }
            }
        }
    }
    if (mine) {
        const double den = (double)(cnt > 1u ? cnt : 1u);
        centers[((int64_t)m * K + c) * maxds + d] = (float)(sum / den);
    }
}

#define TRAIN_TRY(expr)                   \
    do {                                  \
        hipError_t e_ = (expr);           \
        if (e_ != hipSuccess) {           \
            err = e_;                     \
            goto done;                    \
        }                                 \
    } while (0)

hipError_t pq_train_device(const float *data, int64_t n, int dim, const int32_t *offsets, int M,
                           int K, int iters, const double *draws, float *pivots,
                           uint32_t *chunk_status) {
    hipError_t err = hipSuccess;
    int maxds = 0;
    for (int m = 0; m < M; ++m) maxds = offsets[m + 1] - offsets[m] > maxds ? offsets[m + 1] - offsets[m] : maxds;
    if (maxds > 64 || maxds < 1 || K > 256 || K < 1 || n < 1) return hipErrorInvalidValue;
    const int64_t nblk = (n + 15) / 16;
    float *dx = nullptr, *dxc = nullptr, *dn = nullptr, *dmind = nullptr, *dcent = nullptr;
    double *dbsum = nullptr, *ddraw = nullptr;
    uint8_t *dpick = nullptr;
    uint32_t *dassign = nullptr, *dstatus = nullptr;
    int32_t *doff = nullptr;
    std::vector<float> cent;
    bool uniform = true;  // every chunk has width maxds
    // Lloyd update row tile: (4 + 4 maxds) bytes a row within 64 KiB of LDS, a multiple of 64
    // rows and never 0 (192 rows at the widest chunk, maxds = 64)
    const int tile_fit = (65536 / (4 + 4 * maxds)) / 64 * 64;
    const int tile = tile_fit > 64 ? tile_fit : 64;
    const size_t upd_lds = (sizeof(uint32_t) + sizeof(float) * maxds) * (size_t)tile;
    const size_t asg_lds = sizeof(float) * (size_t)K * (maxds + 1);
    const unsigned upd_y = (unsigned)((K * maxds + 255) / 256);
    TRAIN_TRY(hipMalloc(&dx, sizeof(float) * n * dim));
    TRAIN_TRY(hipMalloc(&dxc, sizeof(float) * n * M * maxds));
    TRAIN_TRY(hipMalloc(&dn, sizeof(float) * n * M));
    TRAIN_TRY(hipMalloc(&dmind, sizeof(float) * n * M));
    TRAIN_TRY(hipMalloc(&dpick, (size_t)n * M));
    TRAIN_TRY(hipMalloc(&dbsum, sizeof(double) * nblk * M));
    TRAIN_TRY(hipMalloc(&ddraw, sizeof(double) * (size_t)M * K));
    TRAIN_TRY(hipMalloc(&dcent, sizeof(float) * (size_t)M * K * maxds));
    TRAIN_TRY(hipMalloc(&dassign, sizeof(uint32_t) * n * M));
    TRAIN_TRY(hipMalloc(&dstatus, sizeof(uint32_t) * 2 * M));
    TRAIN_TRY(hipMalloc(&doff, sizeof(int32_t) * (M + 1)));
    TRAIN_TRY(hipMemcpy(dx, data, sizeof(float) * n * dim, hipMemcpyHostToDevice));
    TRAIN_TRY(hipMemcpy(doff, offsets, sizeof(int32_t) * (M + 1), hipMemcpyHostToDevice));
    TRAIN_TRY(hipMemcpy(ddraw, draws, sizeof(double) * (size_t)M * K, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(chunk_rows_kernel, dim3((unsigned)((n + 255) / 256), M), dim3(256), 0, 0,
                       dx, n, dim, doff, maxds, dxc, dn);
    TRAIN_TRY(hipGetLastError());
    hipLaunchKernelGGL(kmeanspp_kernel, dim3(M), dim3(256), 0, 0, dxc, n, doff, dn, K, ddraw,
                       dmind, dpick, dbsum, dcent, maxds, dstatus);
    TRAIN_TRY(hipGetLastError());
    TRAIN_TRY(hipMemcpy(chunk_status, dstatus, sizeof(uint32_t) * 2 * M, hipMemcpyDeviceToHost));
    for (int m = 0; m < M; ++m) {
        // SawInfinity, or nothing picked, is not recoverable (train.rs:180-196)
        if (chunk_status[2 * m] == 3 || chunk_status[2 * m + 1] == 0) {
            err = hipErrorInvalidValue;
            goto done;
        }
    }
    TRAIN_TRY(hipFuncSetAttribute((const void *)lloyd_update_kernel,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)upd_lds));
    // the generic assignment stages K x (maxds + 1) floats: past 64 KiB at K 256, maxds 64
    TRAIN_TRY(hipFuncSetAttribute((const void *)lloyd_assign_kernel<0>,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)asg_lds));
    for (int m = 0; m < M; ++m) uniform = uniform && offsets[m + 1] - offsets[m] == maxds;
    for (int it = 0; it < iters; ++it) {
        const dim3 ag((unsigned)((n + 255) / 256), M);
        switch (uniform && maxds <= 8 ? maxds : 0) {
        case 1: hipLaunchKernelGGL(lloyd_assign_kernel<1>, ag, dim3(256), asg_lds, 0, dxc, n, doff, dn, dcent, K, maxds, dassign); break;
        case 2: hipLaunchKernelGGL(lloyd_assign_kernel<2>, ag, dim3(256), asg_lds, 0, dxc, n, doff, dn, dcent, K, maxds, dassign); break;
        case 3: hipLaunchKernelGGL(lloyd_assign_kernel<3>, ag, dim3(256), asg_lds, 0, dxc, n, doff, dn, dcent, K, maxds, dassign); break;
        case 4: hipLaunchKernelGGL(lloyd_assign_kernel<4>, ag, dim3(256), asg_lds, 0, dxc, n, doff, dn, dcent, K, maxds, dassign); break;
        case 8: hipLaunchKernelGGL(lloyd_assign_kernel<8>, ag, dim3(256), asg_lds, 0, dxc, n, doff, dn, dcent, K, maxds, dassign); break;
        default: hipLaunchKernelGGL(lloyd_assign_kernel<0>, ag, dim3(256), asg_lds, 0, dxc, n, doff, dn, dcent, K, maxds, dassign); break;
        }
        TRAIN_TRY(hipGetLastError());
        hipLaunchKernelGGL(lloyd_update_kernel, dim3(M, upd_y), dim3(256), upd_lds, 0, dxc, n,
                           doff, dassign, K, maxds, tile, dcent);
        TRAIN_TRY(hipGetLastError());
    }
    TRAIN_TRY(hipDeviceSynchronize());
    cent.resize((size_t)M * K * maxds);
    TRAIN_TRY(hipMemcpy(cent.data(), dcent, sizeof(float) * cent.size(), hipMemcpyDeviceToHost));
    for (int m = 0; m < M; ++m)
        for (int c = 0; c < K; ++c)
            for (int d = 0; d < offsets[m + 1] - offsets[m]; ++d)
                pivots[(size_t)c * dim + offsets[m] + d] = cent[((size_t)m * K + c) * maxds + d];
done:
    if (dx) (void)hipFree(dx);
    if (dxc) (void)hipFree(dxc);
    if (dn) (void)hipFree(dn);
    if (dmind) (void)hipFree(dmind);
    if (dpick) (void)hipFree(dpick);
    if (dbsum) (void)hipFree(dbsum);
    if (ddraw) (void)hipFree(ddraw);
    if (dcent) (void)hipFree(dcent);
    if (dassign) (void)hipFree(dassign);
    if (dstatus) (void)hipFree(dstatus);
    if (doff) (void)hipFree(doff);
    return err;
}

}  // namespace dkm

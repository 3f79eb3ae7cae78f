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

__device__ __host__ inline uint64_t splitmix64(uint64_t x) {
    x += 0x9e3779b97f4a7c15ull;
    x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
    x = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
    return x ^ (x >> 31);
}
// Uniform double in [0, 1) for draw `ctr` of chunk `chunk`.
__device__ __host__ inline double draw01(uint64_t seed, int chunk, uint64_t ctr) {
    uint64_t s = splitmix64(seed ^ splitmix64(0x5851f42d4c957f2dull * (uint64_t)(chunk + 1)));
    uint64_t v = splitmix64(s + ctr);
    return (double)(v >> 11) * (1.0 / 9007199254740992.0);
}

// norms[m*n + r] = square_norm(x[r, chunk m])
__global__ void chunk_norms_kernel(const float *x, int64_t n, int dim, const int32_t *offsets,
                                   float *norms) {
    const int m = blockIdx.y;
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    const int o0 = offsets[m], o1 = offsets[m + 1];
    norms[(int64_t)m * n + r] = square_norm_ref(x + r * dim + o0, o1 - o0);
}

// k-means++ seeding, one 256-thread block per chunk; rows split in 256 contiguous
// segments (thread t owns rows [t*n/256, (t+1)*n/256)).
__global__ __launch_bounds__(256) void kmeanspp_kernel(const float *x, int64_t n, int dim,
                                                       const int32_t *offsets, const float *norms,
                                                       int K, uint64_t seed, float *mind_all,
                                                       uint8_t *picked_all, float *centers,
                                                       int maxds, uint32_t *status) {
    const int m = blockIdx.x;
    const int o0 = offsets[m], o1 = offsets[m + 1], ds = o1 - o0;
    const int t = threadIdx.x;
    float *mind = mind_all + (int64_t)m * n;
    uint8_t *picked = picked_all + (int64_t)m * n;
    const float *nrm = norms + (int64_t)m * n;
    float *cent = centers + (int64_t)m * K * maxds;
    __shared__ double segsum[256];
    __shared__ double segoff[256];
    __shared__ float cur[64];
    __shared__ float cur_norm;
    __shared__ uint32_t best;
    __shared__ int stop;
    const int64_t lo = n * t / 256, hi = n * (t + 1) / 256;
    for (int64_t r = lo; r < hi; ++r) {
        mind[r] = __builtin_inff();
        picked[r] = 0;
    }
    for (int e = t; e < K * maxds; e += 256) cent[e] = 0.0f;
    __syncthreads();
    if (t == 0) {
        uint64_t i0 = (uint64_t)(draw01(seed, m, 0) * (double)n);
        if (i0 >= (uint64_t)n) i0 = n - 1;
        picked[i0] = 1;
        for (int d = 0; d < ds; ++d) {
            cur[d] = x[i0 * dim + o0 + d];
            cent[d] = cur[d];
        }
        cur_norm = nrm[i0];
        stop = 0;
    }
    __syncthreads();
    const int steps = (int)(K < n ? K : n);
    for (int c = 1; c < steps; ++c) {
        double acc = 0.0;
        for (int64_t r = lo; r < hi; ++r) {
            const float *xr = x + r * dim + o0;
            float dot = 0.0f;
            for (int d = 0; d < ds; ++d) dot = fma_(xr[d], cur[d], dot);
            float dist = (nrm[r] + cur_norm) + dot * -2.0f;
            float md = mind[r];
            if (dist < md) md = dist;
            mind[r] = md;
            acc += (double)md;
        }
        segsum[t] = acc;
        if (t == 0) best = 0xFFFFFFFFu;
        __syncthreads();
        if (t == 0) {
            double s = 0.0;
            for (int i = 0; i < 256; ++i) {
                segoff[i] = s;
                s += segsum[i];
            }
            segsum[0] = s;  // total (segoff already holds the exclusive offsets)
            if (!(s < __builtin_inf())) {
                stop = 2;  // non-finite total: SawInfinity
            } else if (!(s > 0.0)) {
                stop = 1;  // empty range: no further center (recoverable)
            }
        }
        __syncthreads();
        if (stop) break;
        const double threshold = draw01(seed, m, (uint64_t)c) * segsum[0];
        {
            double run = segoff[t];
            for (int64_t r = lo; r < hi; ++r) {
                float md = mind[r];
                run += (double)md;
                if (run >= threshold && md > 0.0f && !picked[r]) {
                    atomicMin(&best, (uint32_t)r);
                    break;
                }
            }
        }
        __syncthreads();
        if (t == 0) {
            if (best == 0xFFFFFFFFu) {
                stop = 1;  // insufficient diversity (recoverable)
            } else {
                picked[best] = 1;
                for (int d = 0; d < ds; ++d) {
                    cur[d] = x[(int64_t)best * dim + o0 + d];
                    cent[c * maxds + d] = cur[d];
                }
                cur_norm = nrm[best];
            }
        }
        __syncthreads();
        if (stop) break;
    }
    if (t == 0 && stop == 2) atomicOr(status, 1u);
}

// Lloyd assignment (distances_in_place): thread per row, block = 256 rows x one chunk.
__global__ __launch_bounds__(256) void lloyd_assign_kernel(const float *x, int64_t n, int dim,
                                                           const int32_t *offsets,
                                                           const float *norms,
                                                           const float *centers, int K, int maxds,
                                                           uint32_t *assign) {
    extern __shared__ float sm[];
    const int m = blockIdx.y;
    const int o0 = offsets[m], o1 = offsets[m + 1], ds = o1 - o0;
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
    const float *xr = x + r * dim + o0;
    float xv[64];
    for (int d = 0; d < ds; ++d) xv[d] = xr[d];
    const float xn = norms[(int64_t)m * n + r];
    float best = __builtin_inff();
    uint32_t arg = 0xFFFFFFFFu;
    for (int c = 0; c < K; ++c) {
        float dot = 0.0f;
        for (int d = 0; d < ds; ++d) dot = fma_(cs[c * ds + d], xv[d], dot);
        float dist = ((cn[c] - dot) - dot) + xn;
        if (dist < best) {
            best = dist;
            arg = (uint32_t)c;
        }
    }
    assign[(int64_t)m * n + r] = arg;
}

// Lloyd update (update_centroids): one block per chunk, thread c owns center c and walks
// every row in order, so each f64 sum is accumulated in row order exactly as the
// reference does; empty clusters become the zero vector.
__global__ __launch_bounds__(256) void lloyd_update_kernel(const float *x, int64_t n, int dim,
                                                           const int32_t *offsets,
                                                           const uint32_t *assign, int K,
                                                           int maxds, int tile,
                                                           float *centers) {
    extern __shared__ double smd[];
    const int m = blockIdx.x;
    const int o0 = offsets[m], o1 = offsets[m + 1], ds = o1 - o0;
    double *sums = smd;                                   // 256 * ds
    uint32_t *ta = (uint32_t *)(smd + 256 * ds);          // tile
    float *tx = (float *)(ta + tile);                     // tile * ds
    const int c = threadIdx.x;
    for (int d = 0; d < ds; ++d) sums[c * ds + d] = 0.0;
    uint32_t cnt = 0;
    const uint32_t *am = assign + (int64_t)m * n;
    for (int64_t base = 0; base < n; base += tile) {
        const int len = (int)((n - base) < tile ? (n - base) : tile);
        __syncthreads();
        for (int i = threadIdx.x; i < len; i += blockDim.x) ta[i] = am[base + i];
        for (int e = threadIdx.x; e < len * ds; e += blockDim.x) {
            int i = e / ds, d = e - i * ds;
            tx[e] = x[(base + i) * dim + o0 + d];
        }
        __syncthreads();
        if (c < K) {
            for (int i = 0; i < len; ++i) {
                if (ta[i] == (uint32_t)c) {
                    cnt++;
                    for (int d = 0; d < ds; ++d) sums[c * ds + d] += (double)tx[i * ds + d];
                }
            }
        }
    }
    if (c < K) {
        double den = (double)(cnt > 1 ? cnt : 1);
        float *cent = centers + (int64_t)m * K * maxds + c * maxds;
        for (int d = 0; d < ds; ++d) cent[d] = (float)(sums[c * ds + d] / den);
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
                           int K, int iters, uint64_t seed, float *pivots) {
    hipError_t err = hipSuccess;
    int maxds = 0;
    for (int m = 0; m < M; ++m) maxds = offsets[m + 1] - offsets[m] > maxds ? offsets[m + 1] - offsets[m] : maxds;
    if (maxds > 64 || K > 256) return hipErrorInvalidValue;
    float *dx = nullptr, *dn = nullptr, *dmind = nullptr, *dcent = nullptr;
    uint8_t *dpick = nullptr;
    uint32_t *dassign = nullptr, *dstatus = nullptr;
    int32_t *doff = nullptr;
    uint32_t status = 0;
    std::vector<float> cent;
    const int tile = 1024;
    size_t upd_lds = sizeof(double) * 256 * maxds + (sizeof(uint32_t) + sizeof(float) * maxds) * tile;
    size_t asg_lds = sizeof(float) * (size_t)K * (maxds + 1);
    TRAIN_TRY(hipMalloc(&dx, sizeof(float) * n * dim));
    TRAIN_TRY(hipMalloc(&dn, sizeof(float) * n * M));
    TRAIN_TRY(hipMalloc(&dmind, sizeof(float) * n * M));
    TRAIN_TRY(hipMalloc(&dpick, (size_t)n * M));
    TRAIN_TRY(hipMalloc(&dcent, sizeof(float) * (size_t)M * K * maxds));
    TRAIN_TRY(hipMalloc(&dassign, sizeof(uint32_t) * n * M));
    TRAIN_TRY(hipMalloc(&dstatus, sizeof(uint32_t)));
    TRAIN_TRY(hipMalloc(&doff, sizeof(int32_t) * (M + 1)));
    TRAIN_TRY(hipMemcpy(dx, data, sizeof(float) * n * dim, hipMemcpyHostToDevice));
    TRAIN_TRY(hipMemcpy(doff, offsets, sizeof(int32_t) * (M + 1), hipMemcpyHostToDevice));
    TRAIN_TRY(hipMemcpy(dstatus, &status, sizeof(uint32_t), hipMemcpyHostToDevice));
    hipLaunchKernelGGL(chunk_norms_kernel, dim3((unsigned)((n + 255) / 256), M), dim3(256), 0, 0,
                       dx, n, dim, doff, dn);
    TRAIN_TRY(hipGetLastError());
    hipLaunchKernelGGL(kmeanspp_kernel, dim3(M), dim3(256), 0, 0, dx, n, dim, doff, dn, K, seed,
                       dmind, dpick, dcent, maxds, dstatus);
    TRAIN_TRY(hipGetLastError());
    TRAIN_TRY(hipFuncSetAttribute((const void *)lloyd_update_kernel,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)upd_lds));
    for (int it = 0; it < iters; ++it) {
        hipLaunchKernelGGL(lloyd_assign_kernel, dim3((unsigned)((n + 255) / 256), M), dim3(256),
                           asg_lds, 0, dx, n, dim, doff, dn, dcent, K, maxds, dassign);
        TRAIN_TRY(hipGetLastError());
        hipLaunchKernelGGL(lloyd_update_kernel, dim3(M), dim3(256), upd_lds, 0, dx, n, dim, doff,
                           dassign, K, maxds, tile, dcent);
        TRAIN_TRY(hipGetLastError());
    }
    TRAIN_TRY(hipDeviceSynchronize());
    TRAIN_TRY(hipMemcpy(&status, dstatus, sizeof(uint32_t), hipMemcpyDeviceToHost));
    if (status) {
        err = hipErrorInvalidValue;
        goto done;
    }
    cent.resize((size_t)M * K * maxds);
    TRAIN_TRY(hipMemcpy(cent.data(), dcent, sizeof(float) * cent.size(), hipMemcpyDeviceToHost));
    for (int m = 0; m < M; ++m)
        for (int c = 0; c < K; ++c)
            for (int d = 0; d < offsets[m + 1] - offsets[m]; ++d)
                pivots[(size_t)c * dim + offsets[m] + d] = cent[((size_t)m * K + c) * maxds + d];
done:
    if (dx) (void)hipFree(dx);
    if (dn) (void)hipFree(dn);
    if (dmind) (void)hipFree(dmind);
    if (dpick) (void)hipFree(dpick);
    if (dcent) (void)hipFree(dcent);
    if (dassign) (void)hipFree(dassign);
    if (dstatus) (void)hipFree(dstatus);
    if (doff) (void)hipFree(doff);
    return err;
}

}  // namespace dkm

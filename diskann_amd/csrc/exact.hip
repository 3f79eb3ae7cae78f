// Medoid start point and exact FP64 brute-force top-k (the recall oracle of the
// benchmark harness) on gfx950.
//   compute_medoid   diskann-utils/src/sampling/medoid.rs:15-47 (f64 column sums in row
//                    order, mean cast to f32, first row with the smallest SquaredL2,
//                    strict <, starting from f32::MAX)
//   brute force      exact top-k in f64 (ties: smaller id first)
#include <algorithm>
#include "dkm_device.h"
#include "dkm_exact.h"

namespace dkm {

// thread j owns column j and adds the rows in order; the block stages row tiles in LDS
__global__ __launch_bounds__(256) void column_mean_kernel(const float *x, int64_t n,
                                                          int64_t stride, int dim,
                                                          float *mean) {
    __shared__ float tile[32 * 256];
    const int col0 = blockIdx.x * 256;
    const int j = col0 + threadIdx.x;
    const int w = dim - col0 < 256 ? dim - col0 : 256;
    double s = 0.0;
    for (int64_t r0 = 0; r0 < n; r0 += 32) {
        const int rows = (int)(n - r0 < 32 ? n - r0 : 32);
        __syncthreads();
        for (int e = threadIdx.x; e < rows * w; e += 256) {
            int r = e / w, c = e - r * w;
            tile[r * 256 + c] = x[(r0 + r) * stride + col0 + c];
        }
        __syncthreads();
        if (threadIdx.x < w)
            for (int r = 0; r < rows; ++r) s += (double)tile[r * 256 + threadIdx.x];
    }
    if (j < dim) mean[j] = (float)(s / (double)n);
}

__device__ inline uint32_t okey(float d) {
    uint32_t u = __float_as_uint(d == 0.0f ? 0.0f : d);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

__global__ __launch_bounds__(256) void nearest_row_kernel(const float *x, int64_t n,
                                                          int64_t stride, int dim,
                                                          const float *mean,
                                                          unsigned long long *best) {
    const int64_t row = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 5;
    const int hp = threadIdx.x & 31;
    const bool ok = row < n;
    const float *xr = x + (ok ? row : 0) * stride;
    float d = half_simd_4x1<false>(mean, xr, dim, hp);
    // reference: min_dist starts at f32::MAX and only d < min_dist replaces it
    if (ok && hp == 0 && d < 3.40282347e+38f)
        atomicMin(best, ((unsigned long long)okey(d) << 32) | (uint32_t)row);
}

hipError_t medoid_device(const float *x, int64_t n, int64_t stride, int dim, float *mean,
                         unsigned long long *best, hipStream_t stream) {
    hipLaunchKernelGGL(column_mean_kernel, dim3((dim + 255) / 256), dim3(256), 0, stream, x, n,
                       stride, dim, mean);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(nearest_row_kernel, dim3((unsigned)((n * 32 + 255) / 256)), dim3(256), 0,
                       stream, x, n, stride, dim, mean, best);
    return hipGetLastError();
}

__device__ inline bool key_less(double a, uint32_t ia, double b, uint32_t ib) {
    return a < b || (a == b && ia < ib);
}

// One block per query: rows are scanned in tiles; tile candidates beating the current
// k-th entry are merged into the block's sorted top-k list.
__global__ __launch_bounds__(256) void bruteforce_kernel(int metric, const float *x, int64_t n,
                                                         int64_t stride, int dim,
                                                         const float *queries, int k,
                                                         uint32_t *out_ids, double *out_d) {
    extern __shared__ __align__(16) char sm[];
    double *top_d = (double *)sm;                       // k
    uint32_t *top_i = (uint32_t *)(top_d + k);          // k
    double *cd = (double *)(top_i + ((k + 1) & ~1));    // 256 candidates
    uint32_t *ci = (uint32_t *)(cd + 256);
    float *q = (float *)(ci + 256);                     // dim
    __shared__ int ncand;
    __shared__ int have;
    const int qi = blockIdx.x;
    const int t = threadIdx.x;
    for (int e = t; e < dim; e += 256) q[e] = queries[(int64_t)qi * dim + e];
    if (t == 0) {
        have = 0;
        ncand = 0;
    }
    __syncthreads();
    double qq = 0.0;
    if (metric != METRIC_L2)
        for (int e = 0; e < dim; ++e) qq += (double)q[e] * (double)q[e];
    for (int64_t r0 = 0; r0 < n; r0 += 256) {
        const int64_t r = r0 + t;
        double d = 0.0;
        bool ok = r < n;
        if (ok) {
            const float *xr = x + r * stride;
            if (metric == METRIC_L2) {
                for (int e = 0; e < dim; ++e) {
                    double c = (double)q[e] - (double)xr[e];
                    d += c * c;
                }
            } else {
                double xy = 0.0, yy = 0.0;
                for (int e = 0; e < dim; ++e) {
                    xy += (double)q[e] * (double)xr[e];
                    yy += (double)xr[e] * (double)xr[e];
                }
                if (metric == METRIC_IP) d = -xy;
                else if (metric == METRIC_COSNORM) d = 1.0 - xy;
                else d = (qq == 0.0 || yy == 0.0) ? 1.0 : 1.0 - xy / (sqrt(qq) * sqrt(yy));
            }
        }
        const int h = have;
        const bool cand = ok && (h < k || key_less(d, (uint32_t)r, top_d[k - 1], top_i[k - 1]));
        if (cand) {
            int slot = atomicAdd(&ncand, 1);
            cd[slot] = d;
            ci[slot] = (uint32_t)r;
        }
        __syncthreads();
        if (t == 0) {
            int hv = have;
            for (int c = 0; c < ncand; ++c) {
                double dv = cd[c];
                uint32_t iv = ci[c];
                if (hv == k && !key_less(dv, iv, top_d[k - 1], top_i[k - 1])) continue;
                int pos = hv < k ? hv : k - 1;
                while (pos > 0 && key_less(dv, iv, top_d[pos - 1], top_i[pos - 1])) {
                    top_d[pos] = top_d[pos - 1];
                    top_i[pos] = top_i[pos - 1];
                    --pos;
                }
                top_d[pos] = dv;
                top_i[pos] = iv;
                if (hv < k) ++hv;
            }
            have = hv;
            ncand = 0;
        }
        __syncthreads();
    }
    for (int j = t; j < k; j += 256) {
        out_ids[(int64_t)qi * k + j] = j < have ? top_i[j] : 0xFFFFFFFFu;
        out_d[(int64_t)qi * k + j] = j < have ? top_d[j] : __builtin_inf();
    }
}

hipError_t bruteforce_device(int metric, const float *x, int64_t n, int64_t stride, int dim,
                             const float *queries, int64_t nq, int k, uint32_t *out_ids,
                             double *out_d, hipStream_t stream) {
    size_t lds = sizeof(double) * k + sizeof(uint32_t) * ((k + 1) & ~1) +
                 (sizeof(double) + sizeof(uint32_t)) * 256 + sizeof(float) * dim + 64;
    hipError_t e = hipFuncSetAttribute((const void *)bruteforce_kernel,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(bruteforce_kernel, dim3((unsigned)nq), dim3(256), lds, stream, metric, x, n,
                       stride, dim, queries, k, out_ids, out_d);
    return hipGetLastError();
}

__global__ __launch_bounds__(256) void copy_kernel(uint4 *__restrict__ dst,
                                                   const uint4 *__restrict__ src, size_t n) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) dst[i] = src[i];
}

hipError_t copy_device(void *dst, const void *src, size_t bytes, hipStream_t stream) {
    const size_t n = bytes / 16;
    if (n == 0) return hipSuccess;
    const unsigned blocks = (unsigned)std::min<size_t>((n + 255) / 256, 256 * 64);
    hipLaunchKernelGGL(copy_kernel, dim3(blocks), dim3(256), 0, stream, (uint4 *)dst,
                       (const uint4 *)src, n);
    return hipGetLastError();
}

}  // namespace dkm

// 8-bit MinMax quantization kernels (diskann-quantization/src/minmax): per-vector range
// codes after a distance-preserving transform (the spherical quantizer's transforms),
// and the MinMax distances between code rows and against full-precision queries. One
// wave per vector; the reference's sequential f32 sums run on one lane.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "dkm_sph.h"

namespace dkm {

// MinMaxQuantizer::compress (minmax/quantizer.rs:45-160), 8 bits.
__global__ __launch_bounds__(64) void minmax_encode_kernel(SphParams t, float grid, const float *x,
                                                           int64_t stride, int64_t n, uint8_t *out,
                                                           float *loss, uint32_t *err) {
    extern __shared__ __align__(16) float lds[];
    const int64_t r = blockIdx.x;
    if (r >= n) return;
    const int lane = threadIdx.x;
    const float *row = x + r * stride;
    for (int i = lane; i < t.dim; i += 64) lds[i] = row[i];
    __syncthreads();
    const float *v = sph_apply_transform(t, lds);
    float mn = __builtin_nanf(""), mx = __builtin_nanf("");
    bool nan = false;
    for (int i = lane; i < t.out_dim; i += 64) {
        mn = __builtin_fminf(mn, v[i]);
        mx = __builtin_fmaxf(mx, v[i]);
        nan = nan || v[i] != v[i];
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        mn = __builtin_fminf(mn, __shfl_xor(mn, o));
        mx = __builtin_fmaxf(mx, __shfl_xor(mx, o));
    }
    const float width = (mx - mn) / 2.0f;
    const float mid = mn + width;
    const float lo = mid - width * grid, hi = mid + width * grid;
    const float a = __builtin_fmaxf(hi - lo, 1e-8f) / 255.0f;
    uint8_t *o = out + r * (20 + (int64_t)t.out_dim);
    for (int i = lane; i < t.out_dim; i += 64) {
        float c = (v[i] - lo) / a;
        c = c < 0.0f ? 0.0f : (c > 255.0f ? 255.0f : c);
        c = __builtin_roundf(c);
        o[20 + i] = c != c ? (uint8_t)0 : (uint8_t)c;
    }
    if (lane == 0) {
        float ns = 0.0f, cs = 0.0f, ls = 0.0f;
        for (int i = 0; i < t.out_dim; ++i) {
            float c = (v[i] - lo) / a;
            c = c < 0.0f ? 0.0f : (c > 255.0f ? 255.0f : c);
            c = __builtin_roundf(c);
            const float rv = c * a + lo;
            ns += rv * rv;
            cs += c;
            const float d = rv - v[i];
            ls += d * d;
        }
        const uint32_t dim = (uint32_t)t.out_dim;
        const float meta[4] = {lo, a * cs, a, ns};
        for (int k = 0; k < 4; ++k) o[k] = (uint8_t)((dim >> (8 * k)) & 0xFFu);
        for (int j = 0; j < 4; ++j) {
            const uint32_t u = __float_as_uint(meta[j]);
            for (int k = 0; k < 4; ++k) o[4 + 4 * j + k] = (uint8_t)((u >> (8 * k)) & 0xFFu);
        }
        if (loss) loss[r] = ls;
    }
    if (__any(nan) && lane == 0) atomicOr(err, 1u);
}

__device__ inline void mm_read_meta(const uint8_t *c, float m[4]) {
    for (int j = 0; j < 4; ++j) {
        uint32_t u = 0;
        for (int k = 0; k < 4; ++k) u |= (uint32_t)c[4 + 4 * j + k] << (8 * k);
        m[j] = __uint_as_float(u);
    }
}

__device__ inline float mm_finish(int metric, float v, float xns, float yns) {
    if (metric == METRIC_L2) return -2.0f * v + xns + yns;
    if (metric == METRIC_IP) return -v;
    if (metric == METRIC_COSINE) return 1.0f - v / (__builtin_sqrtf(xns) * __builtin_sqrtf(yns));
    return 1.0f - v;
}

// MinMax{IP, L2Squared, Cosine, CosineNormalized} on two code rows (minmax/vectors.rs:
// 206-229 kernel, 231-470): one thread per pair.
__global__ __launch_bounds__(256) void minmax_pairs_kernel(int metric, const uint8_t *left,
                                                           const uint8_t *right, int rowbytes,
                                                           const uint32_t *ids, int64_t nl,
                                                           int64_t m, float *out) {
    const int64_t pair = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (pair >= nl * m) return;
    const uint8_t *a = left + (pair / m) * rowbytes;
    const uint8_t *b = right + (int64_t)ids[pair] * rowbytes;
    const int dim = rowbytes - 20;
    uint32_t raw = 0;
    for (int i = 0; i < dim; ++i) raw += (uint32_t)a[20 + i] * (uint32_t)b[20 + i];
    float xm[4], ym[4];
    mm_read_meta(a, xm);
    mm_read_meta(b, ym);
    const float t0 = xm[2] * ym[2] * (float)raw;
    const float t1x = xm[1] * ym[0];
    const float t1y = ym[1] * xm[0];
    const float t2 = xm[0] * ym[0] * (float)dim;
    out[pair] = mm_finish(metric, t0 + t1x + t1y + t2, xm[3], ym[3]);
}

// FullQueryRef x DataRef<8> (minmax/quantizer.rs:178-210, vectors.rs:271-470): one block
// per query; the f32 x u8 product is sequential (the x86-64 build's 8-bit path).
__global__ __launch_bounds__(64) void minmax_query_pairs_kernel(int metric, SphParams t,
                                                                const float *queries, int64_t nq,
                                                                const uint8_t *codes, int rowbytes,
                                                                const uint32_t *ids, int64_t m,
                                                                float *out, uint32_t *err) {
    extern __shared__ __align__(16) float lds[];
    __shared__ float qm[2];
    const int64_t qi = blockIdx.x;
    if (qi >= nq) return;
    const int lane = threadIdx.x;
    const float *x = queries + qi * t.dim;
    bool nan = false;
    for (int i = lane; i < t.dim; i += 64) {
        lds[i] = x[i];
        nan = nan || x[i] != x[i];
    }
    __syncthreads();
    const float *v = sph_apply_transform(t, lds);
    if (lane == 0) {
        float ns = 0.0f, sum = 0.0f;
        for (int i = 0; i < t.out_dim; ++i) ns += v[i] * v[i];
        for (int i = 0; i < t.out_dim; ++i) sum += v[i];
        qm[0] = sum;
        qm[1] = ns;
    }
    __syncthreads();
    if (__any(nan) && lane == 0) atomicOr(err, 1u);
    for (int64_t j = lane; j < m; j += 64) {
        const uint8_t *b = codes + (int64_t)ids[qi * m + j] * rowbytes;
        float ym[4];
        mm_read_meta(b, ym);
        float s = 0.0f;
        for (int i = 0; i < t.out_dim; ++i) s += v[i] * (float)b[20 + i];
        const float ip = s * ym[2] + qm[0] * ym[0];
        float d;
        if (metric == METRIC_L2) d = qm[1] + ym[3] - 2.0f * ip;
        else if (metric == METRIC_IP) d = -ip;
        else if (metric == METRIC_COSINE) d = 1.0f - ip / (__builtin_sqrtf(qm[1]) * __builtin_sqrtf(ym[3]));
        else d = 1.0f - ip;
        out[qi * m + j] = d;
    }
}

hipError_t launch_minmax_encode(const SphParams &t, float grid, const float *x, int64_t stride,
                                int64_t n, uint8_t *out, float *loss, uint32_t *err,
                                hipStream_t st) {
    if (n <= 0) return hipSuccess;
    const size_t lds = sizeof(float) * (size_t)sph_lds_floats(t);
    hipError_t e = hipFuncSetAttribute((const void *)minmax_encode_kernel,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(minmax_encode_kernel, dim3((unsigned)n), dim3(64), lds, st, t, grid, x,
                       stride, n, out, loss, err);
    return hipGetLastError();
}

hipError_t launch_minmax_pairs(int metric, const uint8_t *left, const uint8_t *right, int rowbytes,
                               const uint32_t *ids, int64_t nl, int64_t m, float *out,
                               hipStream_t st) {
    const int64_t total = nl * m;
    if (total == 0) return hipSuccess;
    hipLaunchKernelGGL(minmax_pairs_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0,
                       st, metric, left, right, rowbytes, ids, nl, m, out);
    return hipGetLastError();
}

hipError_t launch_minmax_query_pairs(int metric, const SphParams &t, const float *queries,
                                     int64_t nq, const uint8_t *codes, int rowbytes,
                                     const uint32_t *ids, int64_t m, float *out, uint32_t *err,
                                     hipStream_t st) {
    if (nq <= 0) return hipSuccess;
    const size_t lds = sizeof(float) * (size_t)sph_lds_floats(t);
    hipError_t e = hipFuncSetAttribute((const void *)minmax_query_pairs_kernel,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(minmax_query_pairs_kernel, dim3((unsigned)nq), dim3(64), lds, st, metric, t,
                       queries, nq, codes, rowbytes, ids, m, out, err);
    return hipGetLastError();
}

}  // namespace dkm

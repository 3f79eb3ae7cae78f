// 8-bit scalar quantization on gfx950 (diskann-quantization/src/scalar):
//   training   ScalarQuantizationParameters::train (train.rs:75-88) with
//              compute_means_and_average_norm / compute_variances (src/utils.rs:109-197)
//   encoding   ScalarQuantizer::compress + compensation (quantizer.rs:190-240, 407-432)
//   distances  CompensatedSquaredL2 / CompensatedIP / CompensatedCosineNormalized
//              (vectors.rs:182-470)
// The training statistics are f64 sums in row order (one thread per dimension walks the
// rows), the compensation is an fma chain in dimension order (one lane per row walks an
// LDS tile), and the distances are exact integer sums (v_dot4_u32_u8) finished with the
// reference's f32 formulas, so every result is bit-identical to the CPU path.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "dkm_device.h"
#include "dkm_sq.h"

namespace dkm {

// ---- training -------------------------------------------------------------------------

// Column sums in row order; with var == nullptr the column means, otherwise the column
// variances about mean (d = x - mean, acc += d * d).
__global__ __launch_bounds__(64) void sq_column_kernel(const float *x, int64_t n, int dim,
                                                       int64_t stride, const double *mean,
                                                       double *out) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= dim) return;
    const float *c = x + j;
    const double m = mean ? mean[j] : 0.0;
    double acc = 0.0;
    int64_t r = 0;
    for (; r + 8 <= n; r += 8) {
        float v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) v[u] = c[(r + u) * stride];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            if (mean) {
                const double d = (double)v[u] - m;
                acc += d * d;
            } else {
                acc += (double)v[u];
            }
        }
    }
    for (; r < n; ++r) {
        const double v = (double)c[r * stride];
        if (mean) {
            const double d = v - m;
            acc += d * d;
        } else {
            acc += v;
        }
    }
    out[j] = acc / (double)n;
}

// Row norms: sqrt of the f64 sum of squares in dimension order.
__global__ __launch_bounds__(256) void sq_row_norm_kernel(const float *x, int64_t n, int dim,
                                                          int64_t stride, double *norms) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    const float *row = x + r * stride;
    double acc = 0.0;
    for (int j = 0; j < dim; ++j) {
        const double v = (double)row[j];
        acc += v * v;
    }
    norms[r] = __builtin_sqrt(acc);
}

// The norm sum in row order (one lane), the range, scale and shift.
__global__ __launch_bounds__(64) void sq_finish_train_kernel(const double *norms, int64_t n,
                                                             const double *mean,
                                                             const double *var, int dim,
                                                             double stdevs, float *scale,
                                                             float *shift, float *mean_norm) {
    __shared__ double p_sh;
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int64_t r = 0; r < n; ++r) s += norms[r];
        double mx = 0.0;
        for (int j = 0; j < dim; ++j) mx = __builtin_fmax(mx, var[j]);
        const double p = __builtin_sqrt(mx) * stdevs;
        p_sh = p;
        *scale = (float)(2.0 * p);
        *mean_norm = (float)(s / (double)n);
    }
    __syncthreads();
    const double p = p_sh;
    for (int j = threadIdx.x; j < dim; j += blockDim.x) shift[j] = (float)(mean[j] - p);
}

// shift_square_norm = <shift, shift> with the reference's f32 inner-product order.
__global__ __launch_bounds__(32) void sq_shift_norm_kernel(const float *shift, int dim,
                                                           float *out) {
    const float v = half_distance<METRIC_IP>(shift, shift, dim, threadIdx.x);
    if (threadIdx.x == 0) *out = -v;
}

// ---- encoding -------------------------------------------------------------------------

// One wave per 64 rows: for each 64-dimension slice the wave computes the codes of the
// 64 rows (coalesced), writes them, and keeps them as floats in an LDS tile; then lane r
// extends row r's compensation fma chain over the slice. Row padding up to cstride is
// zeroed.
__global__ __launch_bounds__(64) void sq_encode_kernel(const float *x, int64_t stride,
                                                       int64_t n, int dim, float scale,
                                                       const float *shift, uint8_t *codes,
                                                       int64_t cstride, float *comp,
                                                       uint32_t *nan_flag) {
    __shared__ float tile[64][65];
    __shared__ float sh[64];
    const int lane = threadIdx.x;
    const int64_t r0 = (int64_t)blockIdx.x * 64;
    const float inv = 255.0f / scale;
    float dot = 0.0f;
    bool nan = false;
    for (int d0 = 0; d0 < dim; d0 += 64) {
        const int d = d0 + lane;
        const float s = d < dim ? shift[d] : 0.0f;
        sh[lane] = s;
        for (int r = 0; r < 64; ++r) {
            const int64_t row = r0 + r;
            float c = 0.0f;
            if (row < n && d < dim) {
                const float f = x[row * stride + d];
                nan = nan || f != f;
                c = (f - s) * inv;
                if (c < 0.0f) c = 0.0f;
                else if (c > 255.0f) c = 255.0f;
                c = __builtin_roundf(c);
                codes[row * cstride + d] = c != c ? (uint8_t)0 : (uint8_t)c;
            }
            tile[r][lane] = c;
        }
        __syncthreads();
        const int len = dim - d0 < 64 ? dim - d0 : 64;
        for (int e = 0; e < len; ++e) dot = fma_(tile[lane][e], sh[e], dot);
        __syncthreads();
    }
    for (int r = 0; r < 64; ++r) {
        const int64_t row = r0 + r;
        if (row >= n) break;
        for (int64_t e = dim + lane; e < cstride; e += 64) codes[row * cstride + e] = 0;
    }
    if (r0 + lane < n) comp[r0 + lane] = scale * (1.0f / 255.0f) * dot;
    if (__any(nan) && lane == 0) atomicOr(nan_flag, 1u);
}

// ---- pair distances -------------------------------------------------------------------

// One group of 16 lanes per (left row i, right row ids[i * m + j]) pair.
__global__ __launch_bounds__(256) void sq_pairs_kernel(int metric, float bss, float ssn,
                                                       const uint8_t *left,
                                                       const float *left_comp,
                                                       const uint8_t *right,
                                                       const float *right_comp,
                                                       int64_t cstride, const uint32_t *ids,
                                                       int64_t nl, int64_t m, float *out) {
    const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t total = nl * m;
    const int64_t pair = gid >> 4;
    const int l = (int)(gid & 15);
    const int64_t pr = pair < total ? pair : total - 1;
    const int64_t i = pr / m;
    const uint32_t id = ids[pr];
    const uint8_t *a = left + i * cstride;
    const uint8_t *b = right + (int64_t)id * cstride;
    uint32_t ab = 0, aa = 0, bb = 0;
    for (int64_t c = l; c < cstride / 16; c += 16) {
        const uint4 va = *(const uint4 *)(a + 16 * c);
        const uint4 vb = *(const uint4 *)(b + 16 * c);
        sq_dot16(va, vb, ab, aa, bb);
    }
#pragma unroll
    for (int o = 8; o > 0; o >>= 1) {
        ab += __shfl_xor(ab, o, 16);
        aa += __shfl_xor(aa, o, 16);
        bb += __shfl_xor(bb, o, 16);
    }
    if (pair < total && l == 0)
        out[pair] = sq_finish(metric, bss, ssn, ab, aa + bb - 2u * ab,
                              metric == METRIC_IP ? left_comp[i] : 0.0f,
                              metric == METRIC_IP ? right_comp[id] : 0.0f);
}

// ---- host launchers -------------------------------------------------------------------

hipError_t launch_sq_train(const float *x, int64_t n, int dim, int64_t stride, double stdevs,
                           double *mean, double *var, double *norms, float *scale,
                           float *shift, float *mean_norm, hipStream_t st) {
    const int cb = (dim + 63) / 64;
    hipLaunchKernelGGL(sq_column_kernel, dim3(cb), dim3(64), 0, st, x, n, dim, stride,
                       (const double *)nullptr, mean);
    hipLaunchKernelGGL(sq_column_kernel, dim3(cb), dim3(64), 0, st, x, n, dim, stride,
                       (const double *)mean, var);
    hipLaunchKernelGGL(sq_row_norm_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st,
                       x, n, dim, stride, norms);
    hipLaunchKernelGGL(sq_finish_train_kernel, dim3(1), dim3(64), 0, st, (const double *)norms,
                       n, (const double *)mean, (const double *)var, dim, stdevs, scale, shift,
                       mean_norm);
    return hipGetLastError();
}

hipError_t launch_sq_shift_norm(const float *shift, int dim, float *out, hipStream_t st) {
    hipLaunchKernelGGL(sq_shift_norm_kernel, dim3(1), dim3(32), 0, st, shift, dim, out);
    return hipGetLastError();
}

hipError_t launch_sq_encode(const float *x, int64_t stride, int64_t n, int dim, float scale,
                            const float *shift, uint8_t *codes, int64_t cstride, float *comp,
                            uint32_t *nan_flag, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(sq_encode_kernel, dim3((unsigned)((n + 63) / 64)), dim3(64), 0, st, x,
                       stride, n, dim, scale, shift, codes, cstride, comp, nan_flag);
    return hipGetLastError();
}

hipError_t launch_sq_pairs(int metric, float bss, float ssn, const uint8_t *left,
                           const float *left_comp, const uint8_t *right,
                           const float *right_comp, int64_t cstride, const uint32_t *ids,
                           int64_t nl, int64_t m, float *out, hipStream_t st) {
    const int64_t threads = nl * m * 16;
    if (threads == 0) return hipSuccess;
    hipLaunchKernelGGL(sq_pairs_kernel, dim3((unsigned)((threads + 255) / 256)), dim3(256), 0,
                       st, metric, bss, ssn, left, left_comp, right, right_comp, cstride, ids,
                       nl, m, out);
    return hipGetLastError();
}

}  // namespace dkm

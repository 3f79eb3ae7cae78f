// k-means for large dimensions (the 1M x 768, K = 1024 configuration): Lloyd iterations
// whose assignment step runs either in the reference's FP32 arithmetic (VALU) or as a
// BF16 MFMA GEMM with FP32 accumulation and a fused arg-min, and whose centroid update is
// the reference's FP64 row-order sum (diskann-quantization/src/algorithms/kmeans/
// lloyds.rs:28-426: distance c.c - x.c - x.c + x.x with x.c a sequential FMA over the
// dimensions, strict < in centre order, centroid = f64 sums in row order / max(count, 1)).
#include <hip/hip_runtime.h>
#include <rocprim/device/device_radix_sort.hpp>

#include <algorithm>
#include <cstdint>

#include "dkm_device.h"
#include "dkm_kmeans.h"

namespace dkm {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8_t;
typedef __attribute__((__vector_size__(16 * sizeof(float)))) float floatx16_t;

// f32 -> bf16, round to nearest even (NaN kept quiet).
__device__ inline uint16_t to_bf16(float f) {
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

// rows x dim f32 (row stride `stride`) -> rows x dp bf16, zero padded to dp columns.
__global__ __launch_bounds__(256) void to_bf16_kernel(const float *x, int64_t rows, int dim,
                                                      int64_t stride, int dp, uint16_t *out) {
    const int64_t total = rows * dp;
    for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < total;
         e += (int64_t)gridDim.x * 256) {
        const int64_t r = e / dp;
        const int d = (int)(e - r * dp);
        out[e] = d < dim ? to_bf16(x[r * stride + d]) : (uint16_t)0;
    }
}

// Squared norm in the reference order (f32x8 partials, as square_norm).
__global__ __launch_bounds__(256) void row_norms_kernel(const float *x, int64_t rows, int dim,
                                                        int64_t stride, float *out) {
    const int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (r < rows) out[r] = square_norm_ref(x + r * stride, dim);
}

// FP32 assignment in the reference arithmetic: one row per thread, four centres at a time
// for instruction-level parallelism (each dot product keeps its own sequential order).
__global__ __launch_bounds__(256) void assign_f32_kernel(const float *x, int64_t n, int dim,
                                                         int64_t stride, const float *c, int K,
                                                         const float *cnorm, uint32_t *assign,
                                                         float *best_out) {
    const int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (r >= n) return;
    const float *xr = x + r * stride;
    const float xn = square_norm_ref(xr, dim);
    float best = __builtin_inff();
    uint32_t arg = 0;
    int k = 0;
    for (; k + 4 <= K; k += 4) {
        const float *c0 = c + (int64_t)k * dim;
        float d0 = 0.0f, d1 = 0.0f, d2 = 0.0f, d3 = 0.0f;
        for (int j = 0; j < dim; ++j) {
            const float xv = xr[j];
            d0 = fma_(c0[j], xv, d0);
            d1 = fma_(c0[dim + j], xv, d1);
            d2 = fma_(c0[2 * dim + j], xv, d2);
            d3 = fma_(c0[3 * dim + j], xv, d3);
        }
        const float dots[4] = {d0, d1, d2, d3};
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const float d = ((cnorm[k + q] - dots[q]) - dots[q]) + xn;
            if (d < best) {
                best = d;
                arg = (uint32_t)(k + q);
            }
        }
    }
    for (; k < K; ++k) {
        const float *ck = c + (int64_t)k * dim;
        float dot = 0.0f;
        for (int j = 0; j < dim; ++j) dot = fma_(ck[j], xr[j], dot);
        const float d = ((cnorm[k] - dot) - dot) + xn;
        if (d < best) {
            best = d;
            arg = (uint32_t)k;
        }
    }
    assign[r] = arg;
    if (best_out) best_out[r] = best;
}

// BF16 MFMA assignment. A workgroup (4 waves) keeps RB x 32 rows of X in LDS (row stride
// dp + 8 bf16, so the 16-byte fragment reads of 8 consecutive rows hit disjoint banks) and
// sweeps the centres in tiles of 128: wave w computes the RB x 32 x 32 scores of columns
// c0 + 32 w .. + 31 with v_mfma_f32_32x32x16_bf16 (lane l holds A[l % 32][8 (l / 32) ..
// + 7] and B[8 (l / 32) .. + 7][l % 32]; accumulator r of lane l is row 8 (r / 4) +
// 4 (l / 32) + r % 4, column l % 32). Each lane keeps the best (c.c - 2 x.c, centre) of
// its rows; lanes of a half-wave, then the four waves, are reduced at the end, ties to
// the smaller centre index (the reference's strict < in centre order).
template <int RB>
__global__ __launch_bounds__(256) void assign_mfma_kernel(const uint16_t *xb, int64_t n, int dp,
                                                          const uint16_t *cb, const float *cnorm,
                                                          int K, uint32_t *assign,
                                                          float *best_out) {
    extern __shared__ __align__(16) uint16_t xs[];
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int half = lane >> 5, l32 = lane & 31;
    const int ld = dp + 8;
    const int64_t row0 = (int64_t)blockIdx.x * (RB * 32);
    const int vpr = dp / 8;  // 16-byte vectors per row
    for (int e = tid; e < RB * 32 * vpr; e += 256) {
        const int r = e / vpr, v = e - r * vpr;
        const int64_t gr = row0 + r;
        const uint4 val = gr < n ? ((const uint4 *)(xb + gr * dp))[v] : make_uint4(0, 0, 0, 0);
        *(uint4 *)(xs + r * ld + 8 * v) = val;
    }
    __syncthreads();
    float bd[RB * 16];
    uint32_t bc[RB * 16];
#pragma unroll
    for (int i = 0; i < RB * 16; ++i) {
        bd[i] = __builtin_inff();
        bc[i] = 0xFFFFFFFFu;
    }
    for (int c0 = 0; c0 < K; c0 += 128) {
        const int col = c0 + wave * 32 + l32;
        const bool cvalid = col < K;
        const uint16_t *crow = cb + (int64_t)(cvalid ? col : 0) * dp + 8 * half;
        floatx16_t acc[RB];
#pragma unroll
        for (int rb = 0; rb < RB; ++rb) acc[rb] = floatx16_t{};
        for (int k0 = 0; k0 < dp; k0 += 16) {
            const bf16x8_t b = *(const bf16x8_t *)(crow + k0);
#pragma unroll
            for (int rb = 0; rb < RB; ++rb) {
                const bf16x8_t a = *(const bf16x8_t *)(xs + (rb * 32 + l32) * ld + k0 + 8 * half);
                acc[rb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc[rb], 0, 0, 0);
            }
        }
        if (cvalid) {
            const float cn = cnorm[col];
#pragma unroll
            for (int rb = 0; rb < RB; ++rb) {
#pragma unroll
                for (int r = 0; r < 16; ++r) {
                    const float d = cn - 2.0f * acc[rb][r];
                    if (d < bd[rb * 16 + r]) {
                        bd[rb * 16 + r] = d;
                        bc[rb * 16 + r] = (uint32_t)col;
                    }
                }
            }
        }
    }
    // lanes of one half-wave hold the same rows (different columns)
#pragma unroll
    for (int i = 0; i < RB * 16; ++i) {
#pragma unroll
        for (int off = 16; off > 0; off >>= 1) {
            const float od = __shfl_xor(bd[i], off);
            const uint32_t oc = __shfl_xor(bc[i], off);
            if (od < bd[i] || (od == bd[i] && oc < bc[i])) {
                bd[i] = od;
                bc[i] = oc;
            }
        }
    }
    __syncthreads();  // the X tile is no longer read: reuse its LDS
    float *rd = (float *)xs;
    uint32_t *rc = (uint32_t *)(rd + 4 * RB * 32);
    if (l32 == 0) {
#pragma unroll
        for (int rb = 0; rb < RB; ++rb) {
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int row = rb * 32 + (r / 4) * 8 + half * 4 + (r % 4);
                rd[wave * RB * 32 + row] = bd[rb * 16 + r];
                rc[wave * RB * 32 + row] = bc[rb * 16 + r];
            }
        }
    }
    __syncthreads();
    if (tid < RB * 32 && row0 + tid < n) {
        float d = rd[tid];
        uint32_t cidx = rc[tid];
        for (int w = 1; w < 4; ++w) {
            const float od = rd[w * RB * 32 + tid];
            const uint32_t oc = rc[w * RB * 32 + tid];
            if (od < d || (od == d && oc < cidx)) {
                d = od;
                cidx = oc;
            }
        }
        assign[row0 + tid] = cidx;
        if (best_out) best_out[row0 + tid] = d;
    }
}

// (cluster << 32 | row) keys, sorted, give every cluster's rows in increasing order.
__global__ __launch_bounds__(256) void cluster_keys_kernel(const uint32_t *assign, int64_t n,
                                                           uint64_t *keys) {
    const int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (r < n) keys[r] = ((uint64_t)assign[r] << 32) | (uint64_t)r;
}

// start[c] = first position of cluster c in the sorted keys (n if none), start[K] = n.
__global__ __launch_bounds__(256) void cluster_starts_kernel(const uint64_t *keys, int64_t n, int K,
                                                             int64_t *start) {
    const int c = blockIdx.x * 256 + threadIdx.x;
    if (c > K) return;
    int64_t lo = 0, hi = n;
    while (lo < hi) {
        const int64_t mid = (lo + hi) / 2;
        if ((keys[mid] >> 32) < (uint64_t)c) lo = mid + 1;
        else hi = mid;
    }
    start[c] = lo;
}

// One workgroup per cluster: f64 sums of its rows in row order, divided by max(count, 1).
__global__ __launch_bounds__(256) void update_kernel(const float *x, int dim, int64_t stride,
                                                    const uint64_t *keys, const int64_t *start,
                                                    float *c) {
    const int k = blockIdx.x;
    const int64_t b = start[k], e = start[k + 1];
    const double den = (double)(e - b > 1 ? e - b : 1);
    for (int d0 = 0; d0 < dim; d0 += 256 * 4) {
        double s[4] = {0.0, 0.0, 0.0, 0.0};
        for (int64_t i = b; i < e; ++i) {
            const float *xr = x + (int64_t)(keys[i] & 0xFFFFFFFFull) * stride;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int d = d0 + threadIdx.x + 256 * q;
                if (d < dim) s[q] += (double)xr[d];
            }
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int d = d0 + threadIdx.x + 256 * q;
            if (d < dim) c[(int64_t)k * dim + d] = (float)(s[q] / den);
        }
    }
}

static unsigned grid_for(int64_t n, int64_t per) { return (unsigned)((n + per - 1) / per); }

hipError_t kmeans_to_bf16(const float *x, int64_t rows, int dim, int64_t stride, int dp,
                          uint16_t *out, hipStream_t st) {
    const int64_t total = rows * dp;
    const unsigned g = (unsigned)std::min<int64_t>((total + 255) / 256, 65536);
    hipLaunchKernelGGL(to_bf16_kernel, dim3(std::max(g, 1u)), dim3(256), 0, st, x, rows, dim, stride, dp, out);
    return hipGetLastError();
}

hipError_t kmeans_norms(const float *x, int64_t rows, int dim, int64_t stride, float *out,
                        hipStream_t st) {
    if (rows == 0) return hipSuccess;
    hipLaunchKernelGGL(row_norms_kernel, dim3(grid_for(rows, 256)), dim3(256), 0, st, x, rows, dim,
                       stride, out);
    return hipGetLastError();
}

hipError_t kmeans_assign_f32(const float *x, int64_t n, int dim, int64_t stride, const float *c,
                             int K, const float *cnorm, uint32_t *assign, float *best,
                             hipStream_t st) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(assign_f32_kernel, dim3(grid_for(n, 256)), dim3(256), 0, st, x, n, dim,
                       stride, c, K, cnorm, assign, best);
    return hipGetLastError();
}

size_t kmeans_mfma_lds(int dp) {
    const int rb = (size_t)64 * (dp + 8) * 2 <= 128 * 1024 ? 2 : 1;
    return std::max<size_t>((size_t)rb * 32 * (dp + 8) * 2, (size_t)8 * rb * 32 * 4);
}

hipError_t kmeans_assign_mfma(const uint16_t *xb, int64_t n, int dp, const uint16_t *cb,
                              const float *cnorm, int K, uint32_t *assign, float *best,
                              hipStream_t st) {
    if (n == 0) return hipSuccess;
    const size_t lds = kmeans_mfma_lds(dp);
    if ((size_t)64 * (dp + 8) * 2 <= 128 * 1024) {
        hipError_t e = hipFuncSetAttribute((const void *)assign_mfma_kernel<2>,
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(assign_mfma_kernel<2>, dim3(grid_for(n, 64)), dim3(256), lds, st, xb, n,
                           dp, cb, cnorm, K, assign, best);
    } else {
        hipError_t e = hipFuncSetAttribute((const void *)assign_mfma_kernel<1>,
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(assign_mfma_kernel<1>, dim3(grid_for(n, 32)), dim3(256), lds, st, xb, n,
                           dp, cb, cnorm, K, assign, best);
    }
    return hipGetLastError();
}

size_t kmeans_sort_tmp_bytes(int64_t n) {
    size_t bytes = 0;
    (void)rocprim::radix_sort_keys(nullptr, bytes, (uint64_t *)nullptr, (uint64_t *)nullptr,
                                   (unsigned)n, 0, 64);
    return bytes;
}

hipError_t kmeans_update(const float *x, int64_t n, int dim, int64_t stride,
                         const uint32_t *assign, int K, uint64_t *keys, uint64_t *keys_tmp,
                         void *sort_tmp, size_t sort_bytes, int64_t *start, float *c,
                         hipStream_t st) {
    if (n > 0) {
        hipLaunchKernelGGL(cluster_keys_kernel, dim3(grid_for(n, 256)), dim3(256), 0, st, assign, n,
                           keys_tmp);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
        e = rocprim::radix_sort_keys(sort_tmp, sort_bytes, keys_tmp, keys, (unsigned)n, 0, 64, st);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(cluster_starts_kernel, dim3(grid_for(K + 1, 256)), dim3(256), 0, st, keys, n,
                       K, start);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(update_kernel, dim3(K), dim3(256), 0, st, x, dim, stride, keys, start, c);
    return hipGetLastError();
}

}  // namespace dkm

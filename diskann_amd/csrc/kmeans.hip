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
#include "dkm_mfma.h"

namespace dkm {

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

// rows x dim f32 -> the MFMA fragment layout of the register-resident kernel: 16-byte chunk
// ((t * ks + s) * 64 + l) holds row 32 t + l % 32, dims 16 s + 8 (l / 32) .. + 7 (zero past
// `rows` and `dim`), so every 1 KB block is one k-step's operand fragment of 32 rows.
__global__ __launch_bounds__(256) void to_bf16_frag_kernel(const float *x, int64_t rows, int dim,
                                                           int64_t stride, int ks,
                                                           int64_t chunks, uint16_t *out) {
    for (int64_t c = (int64_t)blockIdx.x * 256 + threadIdx.x; c < chunks;
         c += (int64_t)gridDim.x * 256) {
        const int64_t t = c / (ks * 64);
        const int rem = (int)(c - t * ks * 64);
        const int s = rem >> 6, l = rem & 63;
        const int64_t r = 32 * t + (l & 31);
        const int d0 = 16 * s + 8 * (l >> 5);
        uint16_t v[8];
#pragma unroll
        for (int j = 0; j < 8; ++j)
            v[j] = (r < rows && d0 + j < dim) ? to_bf16(x[r * stride + d0 + j]) : (uint16_t)0;
        uint4 w;
        w.x = v[0] | ((uint32_t)v[1] << 16);
        w.y = v[2] | ((uint32_t)v[3] << 16);
        w.z = v[4] | ((uint32_t)v[5] << 16);
        w.w = v[6] | ((uint32_t)v[7] << 16);
        *(uint4 *)(out + 8 * c) = w;
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

// BF16 MFMA assignment with the rows resident in registers (dims up to 768), on the
// register-resident GEMM of dkm_mfma.h (rows = data rows, columns = centres). The
// accumulators start at -c.c / 2, so they end at x.c - c.c / 2 = -score / 2, and each lane
// tracks its row's largest value over its centres in increasing index order (strict >); the
// two half-waves are merged at the end, ties to the smaller centre.

// Both row tiles' accumulators start at -c.c / 2 of the lane's 16 centres of the tile at c0.
__device__ __forceinline__ void res_init(const float *cn, int c0, floatx16_t &acc0,
                                         floatx16_t &acc1) {
#pragma unroll
    for (int g = 0; g < 4; ++g) {
        const float4 c4 = *(const float4 *)(cn + c0 + 8 * g);
        acc0[4 * g + 0] = -0.5f * c4.x;
        acc0[4 * g + 1] = -0.5f * c4.y;
        acc0[4 * g + 2] = -0.5f * c4.z;
        acc0[4 * g + 3] = -0.5f * c4.w;
    }
    acc1 = acc0;
}

// Running best of one row: the largest x.c - c.c / 2 (strict >, centres in increasing order,
// so the first of equal scores is kept).
__device__ __forceinline__ void res_epilogue(const floatx16_t &acc, int c0, float &bs,
                                             uint32_t &bc) {
#pragma unroll
    for (int g = 0; g < 4; ++g) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const float v = acc[4 * g + q];
            if (v > bs) {
                bs = v;
                bc = (uint32_t)(c0 + 8 * g + q);
            }
        }
    }
}

// xf: rows in fragment layout (padded to a multiple of 256 rows), cf: centres in fragment
// layout (padded to Kp, a multiple of 64). LDS: two centre tiles (2 x KS KB) and the centre
// norms padded to Kp with +inf.
template <int KS>
__global__ __launch_bounds__(256, 1) void assign_mfma_res_kernel(const uint16_t *xf, int64_t n,
                                                                 const uint16_t *cf,
                                                                 const float *cnorm, int K, int Kp,
                                                                 uint32_t *assign,
                                                                 float *best_out) {
    extern __shared__ __align__(16) uint16_t sm[];
    float *cn_s = (float *)(sm + 2 * KS * 512);
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int half = lane >> 5, l32 = lane & 31;
    const int64_t rtile = (int64_t)blockIdx.x * 8 + 2 * wave;
    bf16x8_t a0[KS], a1[KS];
    res_load_rows<KS>(xf + rtile * KS * 512 + lane * 8, a0, a1);
    // the first centre tile's DMA goes out with the row loads, ahead of the norm fill (whose
    // LDS stores wait for every load issued before them)
    res_prefetch<KS>(cf + lane * 8, sm, wave);
    for (int c = tid; c < Kp; c += 256) cn_s[c] = c < K ? cnorm[c] : __builtin_inff();
    float bs0 = -__builtin_inff(), bs1 = -__builtin_inff();
    uint32_t bc0 = 0xFFFFFFFFu, bc1 = 0xFFFFFFFFu;
    res_sweep<KS, true>(
        sm, cf, Kp / 32, wave, lane, a0, a1,
        [&](int t, floatx16_t &pa, floatx16_t &pb) { res_init(cn_s, 32 * t + 4 * half, pa, pb); },
        [&](int t, floatx16_t &pa, floatx16_t &pb) {
            res_epilogue(pa, 32 * t + 4 * half, bs0, bc0);
            res_epilogue(pb, 32 * t + 4 * half, bs1, bc1);
        });
    const float bsv[2] = {bs0, bs1};
    const uint32_t bcv[2] = {bc0, bc1};
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        float bs = bsv[h];
        uint32_t bc = bcv[h];
        const float os = __shfl_xor(bs, 32);
        const uint32_t oc = __shfl_xor(bc, 32);
        if (os > bs || (os == bs && oc < bc)) {
            bs = os;
            bc = oc;
        }
        const int64_t row = 32 * (rtile + h) + l32;
        if (half == 0 && row < n) {
            assign[row] = bc;
            if (best_out) best_out[row] = -2.0f * bs;
        }
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

// Centroid update: workgroup (k, y) sums dims 256 y .. + 255 of cluster k's rows in f64, in
// increasing row order (one dim per thread; eight rows' loads are issued before their adds so
// the order of the sums is unchanged), and divides by max(count, 1).
__global__ __launch_bounds__(256) void update_kernel(const float *x, int dim, int64_t stride,
                                                    const uint64_t *keys, const int64_t *start,
                                                    float *c) {
    const int k = blockIdx.x;
    const int d = blockIdx.y * 256 + threadIdx.x;
    const int64_t b = start[k], e = start[k + 1];
    const double den = (double)(e - b > 1 ? e - b : 1);
    const int dd = d < dim ? d : dim - 1;
    double s = 0.0;
    int64_t i = b;
    for (; i + 8 <= e; i += 8) {
        float v[8];
#pragma unroll
        for (int q = 0; q < 8; ++q)
            v[q] = x[(int64_t)(keys[i + q] & 0xFFFFFFFFull) * stride + dd];
#pragma unroll
        for (int q = 0; q < 8; ++q) s += (double)v[q];
    }
    for (; i < e; ++i) s += (double)x[(int64_t)(keys[i] & 0xFFFFFFFFull) * stride + dd];
    if (d < dim) c[(int64_t)k * dim + d] = (float)(s / den);
}

static unsigned grid_for(int64_t n, int64_t per) { return (unsigned)((n + per - 1) / per); }

hipError_t kmeans_to_bf16(const float *x, int64_t rows, int dim, int64_t stride, int dp,
                          uint16_t *out, hipStream_t st) {
    const int64_t total = rows * dp;
    const unsigned g = (unsigned)std::min<int64_t>((total + 255) / 256, 65536);
    hipLaunchKernelGGL(to_bf16_kernel, dim3(std::max(g, 1u)), dim3(256), 0, st, x, rows, dim, stride, dp, out);
    return hipGetLastError();
}

hipError_t kmeans_to_bf16_frag(const float *x, int64_t rows, int dim, int64_t stride, int dp,
                               int64_t rows_pad, uint16_t *out, hipStream_t st) {
    const int64_t chunks = rows_pad * dp / 8;
    const unsigned g = (unsigned)std::min<int64_t>((chunks + 255) / 256, 65536);
    hipLaunchKernelGGL(to_bf16_frag_kernel, dim3(std::max(g, 1u)), dim3(256), 0, st, x, rows, dim,
                       stride, dp / 16, chunks, out);
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

static constexpr int kResSteps[] = {4, 8, 16, 24, 32, 48};

int kmeans_bf16_dp(int dim) {
    for (int ks : kResSteps)
        if (16 * ks >= dim) return 16 * ks;
    return (dim + 15) / 16 * 16;
}

static int64_t res_kp(int K) { return ((int64_t)K + 63) / 64 * 64; }

static bool use_resident(int dp, int K) {
    return dp <= 768 && dp % 64 == 0 && (size_t)2 * dp * 64 + 4 * (size_t)res_kp(K) <= 160 * 1024;
}

bool kmeans_bf16_frag(int dp, int K) { return use_resident(dp, K); }

int64_t kmeans_bf16_rows(int64_t n, int dp, int K, bool centres) {
    if (!use_resident(dp, K)) return n;
    return centres ? res_kp((int)n) : (n + 255) / 256 * 256;
}

size_t kmeans_mfma_lds(int dp, int K) {
    if (use_resident(dp, K)) return (size_t)2 * dp * 64 + 4 * (size_t)res_kp(K);
    const int rb = (size_t)64 * (dp + 8) * 2 <= 128 * 1024 ? 2 : 1;
    return std::max<size_t>((size_t)rb * 32 * (dp + 8) * 2, (size_t)8 * rb * 32 * 4);
}

template <int KS>
static hipError_t launch_res(const uint16_t *xb, int64_t n, const uint16_t *cb, const float *cnorm,
                             int K, uint32_t *assign, float *best, hipStream_t st) {
    const size_t lds = kmeans_mfma_lds(KS * 16, K);
    hipError_t e = hipFuncSetAttribute((const void *)assign_mfma_res_kernel<KS>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(assign_mfma_res_kernel<KS>, dim3(grid_for(n, 256)), dim3(256), lds, st, xb,
                       n, cb, cnorm, K, (int)res_kp(K), assign, best);
    return hipGetLastError();
}

hipError_t kmeans_assign_mfma(const uint16_t *xb, int64_t n, int dp, const uint16_t *cb,
                              const float *cnorm, int K, uint32_t *assign, float *best,
                              hipStream_t st) {
    if (n == 0) return hipSuccess;
    if (use_resident(dp, K)) {
        switch (dp / 16) {
        case 4: return launch_res<4>(xb, n, cb, cnorm, K, assign, best, st);
        case 8: return launch_res<8>(xb, n, cb, cnorm, K, assign, best, st);
        case 16: return launch_res<16>(xb, n, cb, cnorm, K, assign, best, st);
        case 24: return launch_res<24>(xb, n, cb, cnorm, K, assign, best, st);
        case 32: return launch_res<32>(xb, n, cb, cnorm, K, assign, best, st);
        case 48: return launch_res<48>(xb, n, cb, cnorm, K, assign, best, st);
        default: break;
        }
    }
    const size_t lds = kmeans_mfma_lds(dp, K);
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

// Keys (cluster << 32 | row) only carry 32 + ceil(log2 K) significant bits.
static unsigned key_bits(int K) {
    unsigned b = 1;
    while ((1ull << b) < (uint64_t)K) ++b;
    return 32 + b;
}

size_t kmeans_sort_tmp_bytes(int64_t n, int K) {
    size_t bytes = 0;
    (void)rocprim::radix_sort_keys(nullptr, bytes, (uint64_t *)nullptr, (uint64_t *)nullptr,
                                   (unsigned)n, 0, key_bits(K));
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
        e = rocprim::radix_sort_keys(sort_tmp, sort_bytes, keys_tmp, keys, (unsigned)n, 0,
                                     key_bits(K), st);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(cluster_starts_kernel, dim3(grid_for(K + 1, 256)), dim3(256), 0, st, keys, n,
                       K, start);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(update_kernel, dim3(K, (dim + 255) / 256), dim3(256), 0, st, x, dim, stride,
                       keys, start, c);
    return hipGetLastError();
}

}  // namespace dkm

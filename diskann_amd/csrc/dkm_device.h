// Device-side building blocks shared by the search, PQ and build kernels (gfx950,
// wave64). Every floating-point reduction here reproduces the reference's exact
// operation order so results are bit-identical to the CPU path:
//   simd_op + Strategy4x1 over f32x8  (diskann-vector/src/distance/simd.rs:686-749, 321-363)
//   f32x8::sum_tree                    (diskann-wide/src/arch/x86_64/v3/f32x8_.rs:187-213)
//   CosineStateless, Strategy2x4       (diskann-vector/src/distance/simd.rs:2281-2461, 408-480)
// This file must be compiled with -ffp-contract=off: the only fused multiply-adds are
// the explicit __builtin_fmaf calls below.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace dkm {

enum : int { METRIC_COSINE = 0, METRIC_IP = 1, METRIC_L2 = 2, METRIC_COSNORM = 3 };

constexpr uint32_t kEmpty = 0xFFFFFFFFu;
constexpr uint32_t kVisitedBit = 0x80000000u;
constexpr int kWave = 64;

__device__ __forceinline__ float fma_(float a, float b, float c) { return __builtin_fmaf(a, b, c); }

// Shuffle inside 32-lane halves of the wave.
__device__ __forceinline__ float down32(float v, int d) { return __shfl_down(v, d, 32); }

// Final 8-lane combine + remainder + sum_tree for a half-wave whose lane p (0..31) holds
// partial accumulator p. Result valid in lane p == 0.
template <bool IP>
__device__ __forceinline__ float finish32(float acc, const float *q, const float *x, int dim,
                                          int p) {
    float v1 = acc + down32(acc, 8);  // lane l, l+16: acc[l]+acc[l+8], acc[l+16]+acc[l+24]
    float t = v1 + down32(v1, 16);    // lane l < 8: (s0+s1)+(s2+s3)
    const int r = dim & 7;
    if (r != 0) {
        const int d8 = dim - r;
        float a = p < r ? q[d8 + p] : 0.0f;
        float b = p < r ? x[d8 + p] : 0.0f;
        if (IP) {
            t = fma_(a, b, t);
        } else {
            float c = a - b;
            t = fma_(c, c, t);
        }
    }
    float u = t + down32(t, 4);  // (x0+x4), (x1+x5), (x2+x6), (x3+x7)
    float w = u + down32(u, 2);  // (x0+x4)+(x2+x6), (x1+x5)+(x3+x7)
    return w + down32(w, 1);
}

// L2 / inner-product distance between q (LDS or global) and x (global row) computed by
// one half-wave; lane p accumulates elements e = p (mod 32), e < dim - dim%8, in order.
template <bool IP>
__device__ __forceinline__ float half_simd_4x1(const float *q, const float *x, int dim, int p) {
    const int d8 = dim & ~7;
    float acc = 0.0f;
    for (int e = p; e < d8; e += 32) {
        float a = q[e], b = x[e];
        if (IP) {
            acc = fma_(a, b, acc);
        } else {
            float c = a - b;
            acc = fma_(c, c, acc);
        }
    }
    return finish32<IP>(acc, q, x, dim, p);
}

// Same with the element loop unrolled for a compile-time dimension (loads issued first).
template <bool IP, int DIM>
__device__ __forceinline__ float half_simd_4x1_fixed(const float *q, const float *x, int p) {
    constexpr int d8 = DIM & ~7;
    constexpr int n = (d8 + 31) / 32;
    float xv[n], qv[n];
#pragma unroll
    for (int i = 0; i < n; ++i) {
        int e = p + 32 * i;
        bool ok = (32 * i + 32 <= d8) || e < d8;
        xv[i] = ok ? x[e] : 0.0f;
        qv[i] = ok ? q[e] : 0.0f;
    }
    float acc = 0.0f;
#pragma unroll
    for (int i = 0; i < n; ++i) {
        int e = p + 32 * i;
        bool ok = (32 * i + 32 <= d8) || e < d8;
        if (ok) {
            if (IP) {
                acc = fma_(qv[i], xv[i], acc);
            } else {
                float c = qv[i] - xv[i];
                acc = fma_(c, c, acc);
            }
        }
    }
    return finish32<IP>(acc, q, x, DIM, p);
}

__device__ __forceinline__ float sum_tree8_lane0(float t) {
    float u = t + down32(t, 4);
    float w = u + down32(u, 2);
    return w + down32(w, 1);
}

// CosineStateless (V3: f32x8, Strategy2x4): lane p < 16 accumulates e = p (mod 16).
// Result valid in lane p == 0 of the half.
__device__ __forceinline__ float half_cosine(const float *q, const float *x, int dim, int p) {
    const int d8 = dim & ~7;
    float nx = 0.0f, ny = 0.0f, xy = 0.0f;
    if (p < 16) {
        for (int e = p; e < d8; e += 16) {
            float a = q[e], b = x[e];
            nx = fma_(a, a, nx);
            ny = fma_(b, b, ny);
            xy = fma_(a, b, xy);
        }
    }
    float cx = nx + down32(nx, 8);
    float cy = ny + down32(ny, 8);
    float cxy = xy + down32(xy, 8);
    const int r = dim & 7;
    if (r != 0) {
        float a = p < r ? q[d8 + p] : 0.0f;
        float b = p < r ? x[d8 + p] : 0.0f;
        cx = fma_(a, a, cx);
        cy = fma_(b, b, cy);
        cxy = fma_(a, b, cxy);
    }
    float normx = sum_tree8_lane0(cx);
    float normy = sum_tree8_lane0(cy);
    float prod = sum_tree8_lane0(cxy);
    float denom = __builtin_sqrtf(normx) * __builtin_sqrtf(normy);
    float sim;
    if (normx < 1.17549435e-38f || normy < 1.17549435e-38f) {
        sim = 0.0f;
    } else {
        float v = prod / denom;
        sim = __builtin_fmaxf(-1.0f, __builtin_fminf(1.0f, v));
    }
    return 1.0f - sim;
}

// Full-precision distance by one half-wave, with the reference's PostOps. f32
// CosineNormalized has its own functor (inner-product loop, 1 - x); only the integer
// element types redirect it to Cosine (distance_provider.rs:274-299).
template <int METRIC>
__device__ __forceinline__ float half_distance(const float *q, const float *x, int dim, int p) {
    if (METRIC == METRIC_L2) {
        if (dim == 128) return half_simd_4x1_fixed<false, 128>(q, x, p);
        if (dim == 96) return half_simd_4x1_fixed<false, 96>(q, x, p);
        return half_simd_4x1<false>(q, x, dim, p);
    } else if (METRIC == METRIC_IP) {
        if (dim == 768) return -half_simd_4x1_fixed<true, 768>(q, x, p);
        if (dim == 128) return -half_simd_4x1_fixed<true, 128>(q, x, p);
        return -half_simd_4x1<true>(q, x, dim, p);
    } else if (METRIC == METRIC_COSNORM) {
        // CosineNormalized for f32: the inner-product loop with post-op 1 - x
        // (diskann-vector/src/distance/implementations.rs:330-332, 368-402)
        if (dim == 768) return 1.0f - half_simd_4x1_fixed<true, 768>(q, x, p);
        return 1.0f - half_simd_4x1<true>(q, x, dim, p);
    } else {
        return half_cosine(q, x, dim, p);
    }
}

// Scalar (single-lane) simd_op for short spans: used for PQ LUT entries and encoding.
// Chain p accumulates e = p (mod 32), e < len - len%8.
template <bool IP>
__device__ __forceinline__ float scalar_simd_4x1(const float *a, const float *b, int len) {
    const int d8 = len & ~7;
    float t[8];
#pragma unroll
    for (int l = 0; l < 8; ++l) {
        float c4[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            float acc = 0.0f;
            for (int e = l + 8 * j; e < d8; e += 32) {
                if (IP) {
                    acc = fma_(a[e], b[e], acc);
                } else {
                    float c = a[e] - b[e];
                    acc = fma_(c, c, acc);
                }
            }
            c4[j] = acc;
        }
        t[l] = (c4[0] + c4[1]) + (c4[2] + c4[3]);
    }
    const int r = len & 7;
    if (r != 0) {
#pragma unroll
        for (int l = 0; l < 8; ++l) {
            float x = l < r ? a[d8 + l] : 0.0f;
            float y = l < r ? b[d8 + l] : 0.0f;
            if (IP) {
                t[l] = fma_(x, y, t[l]);
            } else {
                float c = x - y;
                t[l] = fma_(c, c, t[l]);
            }
        }
    }
    float a0 = t[0] + t[4], a1 = t[1] + t[5], a2 = t[2] + t[6], a3 = t[3] + t[7];
    float b0 = a0 + a2, b1 = a1 + a3;
    return b0 + b1;
}

__device__ __forceinline__ uint32_t hash_id(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

__device__ __forceinline__ uint64_t ballot(bool p) { return __ballot(p); }
__device__ __forceinline__ int lane_id() { return __lane_id(); }
__device__ __forceinline__ int prefix_count(uint64_t mask) {
    // number of set bits in mask below this lane
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32),
                                     __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0));
}

// square_norm (common.rs:9-65) of a strided span of `len` floats.
__device__ inline float square_norm_ref(const float *x, int len) {
    float s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    int i = 0;
    if (i + 32 <= len) {
        float a[32];
#pragma unroll
        for (int p = 0; p < 32; ++p) a[p] = 0.0f;
        while (i + 32 <= len) {
#pragma unroll
            for (int p = 0; p < 32; ++p) a[p] = fma_(x[i + p], x[i + p], a[p]);
            i += 32;
        }
#pragma unroll
        for (int l = 0; l < 8; ++l) s[l] = (a[l] + a[8 + l]) + (a[16 + l] + a[24 + l]);
    }
    while (i + 8 <= len) {
#pragma unroll
        for (int l = 0; l < 8; ++l) s[l] = fma_(x[i + l], x[i + l], s[l]);
        i += 8;
    }
    const int r = len - i;
    if (r != 0) {
#pragma unroll
        for (int l = 0; l < 8; ++l) {
            float v = l < r ? x[i + l] : 0.0f;
            s[l] = fma_(v, v, s[l]);
        }
    }
    float a0 = s[0] + s[4], a1 = s[1] + s[5], a2 = s[2] + s[6], a3 = s[3] + s[7];
    return (a0 + a2) + (a1 + a3);
}

// Direct PQ distances (FixedChunkPQTable::{l2_distance, inner_product, cosine_distance,
// qq_*}, diskann-providers/src/model/pq/fixed_chunk_pq_table.rs:32-56, 220-365) by a
// group of eight lanes, lane l holding lane l of the f32x8 accumulator. Left operand: the
// full-precision row lq, or (lq == nullptr) the pivots of lcode; right: the pivots of
// rcode. For each chunk the lane computes its lane of the simd_op vector over the chunk's
// slices (chunk-local element o into partial o % 32 for L2/IP with Strategy4x1, o % 16 for
// cosine with Strategy2x4, then the masked remainder FMA) and adds it to the running
// accumulator, which sum_tree (shuffles by 4, 2, 1 inside the group) reduces after the last
// chunk. Result valid in lane 0 of the group. METRIC_COSNORM uses the cosine functions.
__device__ __forceinline__ float down8(float v, int d) { return __shfl_down(v, d, 8); }

__device__ __forceinline__ float tree8(float v) {
    float u = v + down8(v, 4);
    float w = u + down8(u, 2);
    return w + down8(w, 1);
}

template <int METRIC>
__device__ inline float pq_direct8(const float *pivots, int dim, const int32_t *offsets, int M,
                                   const float *lq, const uint8_t *lcode, const uint8_t *rcode,
                                   int l) {
    constexpr bool COS = METRIC == METRIC_COSINE || METRIC == METRIC_COSNORM;
    float acc = 0.0f, ax = 0.0f, ay = 0.0f, axy = 0.0f;
    for (int c = 0; c < M; ++c) {
        const int b = offsets[c], len = offsets[c + 1] - b;
        const float *L = lq ? lq + b : pivots + (int64_t)lcode[c] * dim + b;
        const float *R = pivots + (int64_t)rcode[c] * dim + b;
        const int d8 = len & ~7, rem = len & 7;
        if (COS) {
            float nx[2] = {0.0f, 0.0f}, ny[2] = {0.0f, 0.0f}, xy[2] = {0.0f, 0.0f};
            for (int v = 0; 8 * v < d8; ++v) {
                const float a = L[8 * v + l], r = R[8 * v + l];
                nx[v & 1] = fma_(a, a, nx[v & 1]);
                ny[v & 1] = fma_(r, r, ny[v & 1]);
                xy[v & 1] = fma_(a, r, xy[v & 1]);
            }
            float cx = nx[0] + nx[1], cy = ny[0] + ny[1], cxy = xy[0] + xy[1];
            if (rem) {
                const float a = l < rem ? L[d8 + l] : 0.0f, r = l < rem ? R[d8 + l] : 0.0f;
                cx = fma_(a, a, cx);
                cy = fma_(r, r, cy);
                cxy = fma_(a, r, cxy);
            }
            ax = ax + cx;
            ay = ay + cy;
            axy = axy + cxy;
        } else {
            float s[4] = {0.0f, 0.0f, 0.0f, 0.0f};
            for (int v = 0; 8 * v < d8; ++v) {
                const float a = L[8 * v + l], r = R[8 * v + l];
                if (METRIC == METRIC_IP) {
                    s[v & 3] = fma_(a, r, s[v & 3]);
                } else {
                    const float d = a - r;
                    s[v & 3] = fma_(d, d, s[v & 3]);
                }
            }
            float t = (s[0] + s[1]) + (s[2] + s[3]);
            if (rem) {
                const float a = l < rem ? L[d8 + l] : 0.0f, r = l < rem ? R[d8 + l] : 0.0f;
                if (METRIC == METRIC_IP) {
                    t = fma_(a, r, t);
                } else {
                    const float d = a - r;
                    t = fma_(d, d, t);
                }
            }
            acc = acc + t;
        }
    }
    if (COS) {
        const float normx = tree8(ax), normy = tree8(ay), prod = tree8(axy);
        const float denom = __builtin_sqrtf(normx) * __builtin_sqrtf(normy);
        float sim;
        if (normx < 1.17549435e-38f || normy < 1.17549435e-38f) {
            sim = 0.0f;
        } else {
            const float v = prod / denom;
            sim = __builtin_fmaxf(-1.0f, __builtin_fminf(1.0f, v));
        }
        return 1.0f - sim;
    }
    const float sum = tree8(acc);
    return METRIC == METRIC_IP ? -sum : sum;
}

__device__ inline float pq_direct8_rt(int metric, const float *pivots, int dim,
                                      const int32_t *offsets, int M, const float *lq,
                                      const uint8_t *lcode, const uint8_t *rcode, int l) {
    if (metric == METRIC_L2) return pq_direct8<METRIC_L2>(pivots, dim, offsets, M, lq, lcode, rcode, l);
    if (metric == METRIC_IP) return pq_direct8<METRIC_IP>(pivots, dim, offsets, M, lq, lcode, rcode, l);
    return pq_direct8<METRIC_COSINE>(pivots, dim, offsets, M, lq, lcode, rcode, l);
}


}  // namespace dkm

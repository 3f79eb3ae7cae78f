// 1-bit spherical quantization (diskann-quantization/src/spherical): device helpers shared
// by sph.hip and the search kernel, and the launch interface. One wave (64 lanes, one
// block) prepares one vector in LDS: preprocessing, the distance-preserving transform and
// the code; every floating-point step follows the reference's order so codes, metadata
// and distances are bit-identical to the CPU path.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "dkm_device.h"

namespace dkm {

enum : int { SPH_NULL = 0, SPH_PADDING_HADAMARD = 1, SPH_DOUBLE_HADAMARD = 2 };
enum : int { SPH_QUERY_SAME_AS_DATA = 0, SPH_QUERY_FOUR_BIT = 1, SPH_QUERY_FULL = 2 };
enum : uint32_t { SPH_ERR_NAN = 1u, SPH_ERR_META = 2u };

// Device copy of a SphericalQuantizer.
struct SphParams {
    int metric;               // METRIC_L2 / METRIC_IP / METRIC_COSINE
    int dim, out_dim;
    int kind;                 // SPH_*
    int n;                    // transform working length (padded / intermediate dimension)
    int split;                // double Hadamard: largest power of two <= n
    const float *shift;       // dim
    const uint32_t *signs0;   // dim f32 sign masks
    const uint32_t *signs1;   // n (double Hadamard)
    const uint32_t *subsample;  // out_dim indices or nullptr
    float mean_norm, pre_scale;
    float shift_sq;           // <shift, shift>
    float sub_rescale;        // sqrt(n / out_dim) when subsampling
    int cstride;              // bytes per stored bit row (multiple of 16)
};

__host__ __device__ inline int sph_bit_stride(int out_dim) { return ((out_dim + 7) / 8 + 15) & ~15; }
__host__ __device__ inline int64_t sph_canon_bytes(int out_dim) { return (out_dim + 7) / 8 + 6; }

// LDS floats one prepared vector needs: the shifted input / gathered output
// (max(dim, out_dim)) and the transform buffer (n).
__host__ __device__ inline int sph_front(const SphParams &q) {
    return q.dim > q.out_dim ? q.dim : q.out_dim;
}
__host__ __device__ inline int sph_lds_floats(const SphParams &q) { return sph_front(q) + q.n + 8; }

__device__ __forceinline__ float sph_flip(float v, uint32_t s) {
    return __uint_as_float(__float_as_uint(v) ^ s);
}

// hadamard_transform (diskann-quantization/src/algorithms/hadamard.rs) as the x86-64-v3
// build runs it, by one wave on LDS: for length >= 64 each 64-element block first gets
// its eight 8-point transforms as sequential +-sums (micro_kernel_64), then butterflies
// (l + r, l - r) at strides 8, 16, ...; shorter lengths use butterflies from stride 1.
// Each butterfly output is one rounding of a + b or a - b, so the result does not depend
// on the order the butterflies of a stage run in. Finally x *= 1 / sqrt(len).
__device__ inline void wave_hadamard(float *x, int len) {
    const int lane = threadIdx.x;
    if (len <= 1) return;
    int h0 = 1;
    if (len >= 64) {
        const int j = lane >> 3, k = lane & 7;
        for (int b = 0; b < len; b += 64) {
            float acc = 0.0f;
#pragma unroll
            for (int m = 0; m < 8; ++m) {
                const float hv = (__builtin_popcount((unsigned)(m & k)) & 1) ? -1.0f : 1.0f;
                acc = fma_(x[b + 8 * j + m], hv, acc);
            }
            __syncthreads();
            x[b + lane] = acc;
            __syncthreads();
        }
        h0 = 8;
    }
    for (int h = h0; h < len; h <<= 1) {
        for (int p = lane; p < len / 2; p += 64) {
            const int i = (p / h) * 2 * h + (p % h);
            const float a = x[i], c = x[i + h];
            x[i] = a + c;
            x[i + h] = a - c;
        }
        __syncthreads();
    }
    const float m = 1.0f / __builtin_sqrtf((float)len);
    for (int i = lane; i < len; i += 64) x[i] *= m;
    __syncthreads();
}

// Inner product (mathematical value) of two LDS / global vectors with the reference's
// f32 order, result broadcast to every lane of the wave.
__device__ __forceinline__ float wave_ip(const float *a, const float *b, int len) {
    const float v = half_distance<METRIC_IP>(a, b, len, threadIdx.x & 31);
    return -__shfl(v, 0);
}

// The distance-preserving transform (NullTransform / PaddingHadamard / DoubleHadamard
// ::transform_into, algorithms/transforms/) of the dim floats at the front of lds by one
// wave; returns the out_dim transformed floats (front area or transform buffer).
__device__ inline float *sph_apply_transform(const SphParams &q, float *lds) {
    const int lane = threadIdx.x;
    float *sh = lds;
    float *buf = lds + sph_front(q);
    if (q.kind == SPH_NULL) return sh;
    for (int i = lane; i < q.n; i += 64) buf[i] = i < q.dim ? sph_flip(sh[i], q.signs0[i]) : 0.0f;
    __syncthreads();
    if (q.kind == SPH_PADDING_HADAMARD) {
        wave_hadamard(buf, q.n);
    } else {
        wave_hadamard(buf, q.split);
        for (int i = lane; i < q.n; i += 64) buf[i] = sph_flip(buf[i], q.signs1[i]);
        __syncthreads();
        wave_hadamard(buf + (q.n - q.split), q.split);
    }
    if (q.subsample) {
        // gather into the front area (max(dim, out_dim) floats)
        for (int i = lane; i < q.out_dim; i += 64) sh[i] = buf[q.subsample[i]] * q.sub_rescale;
        __syncthreads();
        return sh;
    }
    return buf;
}

// SphericalQuantizer::preprocess + normalisation + transform (spherical/quantizer.rs:
// 333-377, 806-845) of x (dim floats, global or LDS) by one wave. lds: sph_lds_floats.
// Returns the transformed vector (out_dim floats in LDS) or nullptr when the shifted norm
// is zero; sn = shifted norm, ms = metric term; err gets SPH_ERR_NAN for a non-finite norm.
__device__ inline float *sph_prepare(const SphParams &q, const float *x, float *lds, float &sn,
                                     float &ms, uint32_t &err) {
    const int lane = threadIdx.x;
    float *sh = lds;
    float mul = q.pre_scale;
    if (q.metric == METRIC_COSINE) {
        const float norm = __builtin_sqrtf(wave_ip(x, x, q.dim));
        mul = norm == 0.0f ? 1.0f : 1.0f / norm;
    }
    for (int i = lane; i < q.dim; i += 64) sh[i] = mul * x[i] - q.shift[i];
    __syncthreads();
    const float norm = __builtin_sqrtf(wave_ip(sh, sh, q.dim));
    sn = norm;
    ms = q.metric == METRIC_L2 ? norm * norm : wave_ip(sh, q.shift, q.dim);
    if (!__builtin_isfinite(norm)) {
        err |= SPH_ERR_NAN;
        return nullptr;
    }
    if (norm == 0.0f) return nullptr;
    for (int i = lane; i < q.dim; i += 64) sh[i] = sh[i] / norm;
    __syncthreads();
    return sph_apply_transform(q, lds);
}

__device__ __forceinline__ uint16_t f16_bits(float v) {
    const _Float16 h = (_Float16)v;
    return __builtin_bit_cast(uint16_t, h);
}
__device__ __forceinline__ float f16_value(uint16_t b) {
    return (float)__builtin_bit_cast(_Float16, b);
}

// Data metadata as f32 (the f16 values widened exactly) and the bit count.
struct SphMeta {
    float ipc, ms, bs, pad;
};

// The f32 finish of the data x data kernel and the compensated metrics
// (spherical/vectors.rs:493-535, 570-595, 730-759, 867-897).
__device__ __forceinline__ float sph_data_data(int metric, float shift_sq, float dim, uint32_t ip,
                                               const SphMeta &x, const SphMeta &y) {
    const float k = x.ipc * y.ipc * ((float)ip - 0.5f * (x.bs + y.bs) + 0.25f * dim);
    if (metric == METRIC_L2) return x.ms + y.ms - 2.0f * k;
    const float v = x.ms + y.ms + k + shift_sq;
    return metric == METRIC_COSINE ? 1.0f - v : -v;
}

// Query<4> x Data<1> (spherical/vectors.rs:597-641, 761-806): qm = {ip correction, bit sum,
// offset, metric term}.
__device__ __forceinline__ float sph_query4_data(int metric, float shift_sq, float dim, uint32_t ip,
                                                 const float *qm, const SphMeta &y) {
    const float cip = y.ipc * qm[0] * ((float)ip - 0.5f * qm[1] + qm[2] * y.bs - 0.5f * qm[2] * dim);
    if (metric == METRIC_L2) return y.ms + qm[3] - 2.0f * cip;
    const float v = cip + y.ms + qm[3] + shift_sq;
    return metric == METRIC_COSINE ? 1.0f - v : -v;
}

__host__ __device__ inline int sph_query_lds_floats(const SphParams &q) {
    return ((q.dim + 3) & ~3) + ((sph_lds_floats(q) + 3) & ~3);
}

// Bytes of the compressed query area: 4 bit planes, or the full-precision vector.
__host__ __device__ inline int sph_query_code_bytes(const SphParams &q) {
    const int planes = 4 * q.cstride, full = 4 * q.out_dim;
    return planes > full ? planes : full;
}

// Bits of t > 0 for out_dim values into one or two destinations (bytes [0, nb)).
__device__ inline uint32_t sph_write_bits(const SphParams &q, const float *t, uint8_t *a,
                                          uint8_t *b) {
    const int lane = threadIdx.x;
    const int nb = (q.out_dim + 7) / 8;
    uint32_t bitsum = 0;
    for (int g = 0; g < q.out_dim; g += 64) {
        const int i = g + lane;
        const uint64_t mask = __ballot(t && i < q.out_dim && t[i] > 0.0f);
        bitsum += (uint32_t)__builtin_popcountll(mask);
        const int byte = g / 8 + lane;
        if (lane < 8 && byte < nb) {
            const uint8_t v = (uint8_t)((mask >> (8 * lane)) & 0xFFu);
            if (a) a[byte] = v;
            if (b) b[byte] = v;
        }
    }
    return bitsum;
}

// The 1-bit finish: correction 2 |t| |shifted| / sum |t| (sum in index order), f16 meta.
__device__ inline SphMeta sph_data_meta(const SphParams &q, const float *t, float sn, float ms,
                                        uint32_t bitsum, uint16_t *h, uint32_t &err) {
    SphMeta m = {0.0f, 0.0f, 0.0f, 0.0f};
    h[0] = h[1] = h[2] = 0;
    if (!t) return m;
    const float tn = q.subsample ? __builtin_sqrtf(wave_ip(t, t, q.out_dim)) : 1.0f;
    float raw = 0.0f;
    if (threadIdx.x == 0)
        for (int i = 0; i < q.out_dim; ++i) raw += __builtin_fabsf(t[i]);
    raw = __shfl(raw, 0);
    const float ipc = 2.0f * tn * sn / raw;
    h[0] = f16_bits(ipc);
    h[1] = f16_bits(ms);
    h[2] = (uint16_t)bitsum;
    if ((h[0] & 0x7C00u) == 0x7C00u || (h[1] & 0x7C00u) == 0x7C00u) err |= SPH_ERR_META;
    m.ipc = f16_value(h[0]);
    m.ms = f16_value(h[1]);
    m.bs = (float)bitsum;
    return m;
}

// Compress a query for one of the layouts of SphericalQuantizer's fused_query_computer
// (spherical/iface.rs:1898-1950), after the inner-product rescale when `rescale`
// (quantizer.rs:322-326). lds: sph_query_lds_floats; qb: sph_query_code_bytes. qm (LDS,
// 4 floats): same-as-data {correction, metric term, bit sum, 0} with qb the bit row;
// 4-bit {correction, bit sum, offset, metric term} with qb the 4 bit planes (plane k at
// k * cstride); full precision {sum, shifted norm, metric term, 0} with qb the vector.
__device__ inline void sph_query_code(const SphParams &q, int layout, bool rescale, const float *x,
                                      float *lds, uint8_t *qb, float *qm, uint32_t &err) {
    const int lane = threadIdx.x;
    float *xs = lds;
    float *work = lds + ((q.dim + 3) & ~3);
    if (rescale && q.metric == METRIC_IP) {
        const float norm = __builtin_sqrtf(wave_ip(x, x, q.dim));
        const float mul = q.mean_norm / norm;
        for (int i = lane; i < q.dim; i += 64) xs[i] = x[i] * mul;
        __syncthreads();
        x = xs;
    }
    float sn = 0.0f, ms = 0.0f;
    uint32_t e = 0;
    const float *t = sph_prepare(q, x, work, sn, ms, e);
    if (e) t = nullptr;
    err |= e;
    if (layout == SPH_QUERY_SAME_AS_DATA) {
        for (int i = lane; i < q.cstride; i += 64) qb[i] = 0;
        __syncthreads();
        const uint32_t bitsum = sph_write_bits(q, t, qb, nullptr);
        uint16_t h[3];
        const SphMeta dm = sph_data_meta(q, t, sn, ms, bitsum, h, e);
        if (lane == 0) {
            qm[0] = dm.ipc;
            qm[1] = dm.ms;
            qm[2] = dm.bs;
            qm[3] = 0.0f;
        }
    } else if (layout == SPH_QUERY_FOUR_BIT) {
        for (int i = lane; i < 4 * q.cstride; i += 64) qb[i] = 0;
        __syncthreads();
        float mn = 3.40282347e+38f, mx = -3.40282347e+38f;
        if (t)
            for (int i = lane; i < q.out_dim; i += 64) {
                mn = __builtin_fminf(t[i], mn);
                mx = __builtin_fmaxf(t[i], mx);
            }
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            mn = __builtin_fminf(mn, __shfl_xor(mn, o));
            mx = __builtin_fmaxf(mx, __shfl_xor(mx, o));
        }
        const float scale = (mx - mn) / 15.0f;
        uint32_t bs = 0;
        for (int g = 0; g < q.out_dim; g += 64) {
            const int i = g + lane;
            uint32_t c = 0;
            if (t && i < q.out_dim) {
                float v = __builtin_roundf((t[i] - mn) / scale);
                v = v < 0.0f ? 0.0f : (v > 15.0f ? 15.0f : v);
                c = (uint32_t)v;
            }
            bs += c;
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const uint64_t mask = __ballot((c >> k) & 1u);
                if (lane < 8 && g / 8 + lane < q.cstride)
                    qb[k * q.cstride + g / 8 + lane] = (uint8_t)((mask >> (8 * lane)) & 0xFFu);
            }
        }
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) bs += __shfl_xor(bs, o);
        if (lane == 0) {
            qm[0] = t ? sn * scale : 0.0f;
            qm[1] = t ? (float)bs : 0.0f;
            qm[2] = t ? mn / scale : 0.0f;
            qm[3] = t ? ms : 0.0f;
        }
    } else {
        float *fv = (float *)qb;  // out_dim floats (the caller sized the area for them)
        for (int i = lane; i < q.out_dim; i += 64) fv[i] = t ? t[i] : 0.0f;
        __syncthreads();
        if (lane == 0) {
            float sum = 0.0f;
            for (int i = 0; i < q.out_dim; ++i) sum += fv[i];
            qm[0] = t ? sum : 0.0f;
            qm[1] = t ? sn : 0.0f;
            qm[2] = t ? ms : 0.0f;
            qm[3] = 0.0f;
        }
    }
    __syncthreads();
}

// A prepared traversal query in HBM: the bit row (same-as-data) or the 4 bit planes,
// then its 4 metadata floats.
__host__ __device__ inline int sph_query_record_bytes(const SphParams &q) { return 4 * q.cstride + 16; }

hipError_t launch_sph_query_prep(const SphParams &q, int layout, int rescale, const float *queries,
                                 int64_t qstride, int64_t nq, uint8_t *records, hipStream_t st);
// 8-bit MinMax (diskann-quantization/src/minmax) with a SphParams transform: canonical
// rows of 20 + out_dim bytes (u32 dim, f32 b, f32 n, f32 a, f32 |r|^2, codes).
hipError_t launch_minmax_encode(const SphParams &t, float grid, const float *x, int64_t stride,
                                int64_t n, uint8_t *out, float *loss, uint32_t *err,
                                hipStream_t st);
hipError_t launch_minmax_pairs(int metric, const uint8_t *left, const uint8_t *right, int rowbytes,
                               const uint32_t *ids, int64_t nl, int64_t m, float *out,
                               hipStream_t st);
hipError_t launch_minmax_query_pairs(int metric, const SphParams &t, const float *queries,
                                     int64_t nq, const uint8_t *codes, int rowbytes,
                                     const uint32_t *ids, int64_t m, float *out, uint32_t *err,
                                     hipStream_t st);
hipError_t launch_sph_shift_sq(const float *shift, int dim, float *out, hipStream_t st);
hipError_t launch_sph_encode(const SphParams &q, const float *x, int64_t stride, int64_t n,
                             uint8_t *bits, SphMeta *meta, uint8_t *canon, uint32_t *err,
                             hipStream_t st);
hipError_t launch_sph_pairs(const SphParams &q, const uint8_t *lbits, const SphMeta *lmeta,
                            const uint8_t *rbits, const SphMeta *rmeta, const uint32_t *ids,
                            int64_t nl, int64_t m, float *out, hipStream_t st);
hipError_t launch_sph_query_pairs(const SphParams &q, int layout, int rescale, const float *queries,
                                  int64_t nq, const uint8_t *bits, const SphMeta *meta,
                                  const uint32_t *ids, int64_t m, float *out, uint32_t *err,
                                  hipStream_t st);
hipError_t launch_sph_train_stats(int metric, const float *x, int64_t n, int dim, double *mean,
                                  double *norms, float *centroid, float *mean_norm,
                                  hipStream_t st);

}  // namespace dkm

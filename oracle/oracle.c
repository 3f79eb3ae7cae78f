/*
 * CPU oracle for the DiskANN search-and-quantize path -- TEST INFRASTRUCTURE ONLY.
 * Plain C restatement of the reference algorithm; see oracle.h for the parity status of
 * every function. Built with -ffp-contract=off so that the only fused operations are the
 * explicit fmaf() calls that mirror the reference's mul_add_simd.
 */
#define _GNU_SOURCE
#include "oracle.h"

#include <float.h>
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

/* ------------------------------------------------------------------------------------ */
/* Distances                                                                            */
/* ------------------------------------------------------------------------------------ */

/* f32x8::sum_tree, diskann-wide/src/arch/x86_64/v3/f32x8_.rs:187-213:
 * ((x0+x4)+(x2+x6)) + ((x1+x5)+(x3+x7)). */
static float sum_tree8(const float *t) {
    float a0 = t[0] + t[4], a1 = t[1] + t[5], a2 = t[2] + t[6], a3 = t[3] + t[7];
    float b0 = a0 + a2, b1 = a1 + a3;
    return b0 + b1;
}

/* simd_op with Strategy4x1 over f32x8 (diskann-vector/src/distance/simd.rs:686-749,
 * 321-363; L2 schema :817-845, IP schema :1588-1600). The main loop and the full-width
 * epilogues together place element e (e < len - len%8) into partial accumulator e % 32,
 * in increasing e; the accumulators combine as (s0+s1)+(s2+s3) lane-wise; the masked
 * remainder (zero-filled) is accumulated into the combined vector; then sum_tree. */
static float simd_4x1(int ip, const float *x, const float *y, int64_t len) {
    float acc[32];
    for (int p = 0; p < 32; ++p) acc[p] = 0.0f;
    int64_t d8 = len - (len % 8);
    int64_t e = 0;
    for (; e + 32 <= d8; e += 32) {
        for (int p = 0; p < 32; ++p) {
            float a = x[e + p], b = y[e + p];
            if (ip) {
                acc[p] = fmaf(a, b, acc[p]);
            } else {
                float c = a - b;
                acc[p] = fmaf(c, c, acc[p]);
            }
        }
    }
    for (int p = 0; e + p < d8; ++p) {
        float a = x[e + p], b = y[e + p];
        if (ip) {
            acc[p] = fmaf(a, b, acc[p]);
        } else {
            float c = a - b;
            acc[p] = fmaf(c, c, acc[p]);
        }
    }
    float t[8];
    for (int l = 0; l < 8; ++l) t[l] = (acc[l] + acc[8 + l]) + (acc[16 + l] + acc[24 + l]);
    int64_t r = len % 8;
    if (r != 0) {
        for (int l = 0; l < 8; ++l) {
            float a = l < r ? x[d8 + l] : 0.0f;
            float b = l < r ? y[d8 + l] : 0.0f;
            if (ip) {
                t[l] = fmaf(a, b, t[l]);
            } else {
                float c = a - b;
                t[l] = fmaf(c, c, t[l]);
            }
        }
    }
    return sum_tree8(t);
}

/* CosineStateless over f32x8 with Strategy2x4 (simd.rs:2281-2461, 408-480): element e
 * (e < len - len%8) goes to partial e % 16 for each of normx, normy, xy; combine s0+s1;
 * masked remainder; FullCosineAccumulator::sum with subnormal-norm guard and clamp. */
static float cosine_v3(const float *x, const float *y, int64_t len) {
    float nx[16], ny[16], xy[16];
    for (int p = 0; p < 16; ++p) nx[p] = ny[p] = xy[p] = 0.0f;
    int64_t d8 = len - (len % 8);
    for (int64_t e = 0; e < d8; ++e) {
        int p = (int)(e % 16);
        float a = x[e], b = y[e];
        nx[p] = fmaf(a, a, nx[p]);
        ny[p] = fmaf(b, b, ny[p]);
        xy[p] = fmaf(a, b, xy[p]);
    }
    float cx[8], cy[8], cxy[8];
    for (int l = 0; l < 8; ++l) {
        cx[l] = nx[l] + nx[8 + l];
        cy[l] = ny[l] + ny[8 + l];
        cxy[l] = xy[l] + xy[8 + l];
    }
    int64_t r = len % 8;
    if (r != 0) {
        for (int l = 0; l < 8; ++l) {
            float a = l < r ? x[d8 + l] : 0.0f;
            float b = l < r ? y[d8 + l] : 0.0f;
            cx[l] = fmaf(a, a, cx[l]);
            cy[l] = fmaf(b, b, cy[l]);
            cxy[l] = fmaf(a, b, cxy[l]);
        }
    }
    float normx = sum_tree8(cx);
    float normy = sum_tree8(cy);
    float denom = sqrtf(normx) * sqrtf(normy);
    float prod = sum_tree8(cxy);
    float sim;
    if (normx < FLT_MIN || normy < FLT_MIN) {
        sim = 0.0f;
    } else {
        float v = prod / denom;
        sim = fmaxf(-1.0f, fminf(1.0f, v));
    }
    return 1.0f - sim;
}

/* The f32x8 vector a simd_op leaves before its final reduction, for one pair of slices
 * (the per-chunk step of the resumable schemas, simd.rs:609-666): L2/IP with Strategy4x1
 * (slice-local element o < len - len%8 into partial o % 32, (s0+s1)+(s2+s3), masked
 * remainder FMA), cosine with Strategy2x4 (partial o % 16, s0+s1, masked remainder) for
 * the three accumulators normx, normy, xy. */
static void chunk_vec_4x1(int ip, const float *x, const float *y, int64_t len, float t[8]) {
    float acc[32];
    for (int p = 0; p < 32; ++p) acc[p] = 0.0f;
    int64_t d8 = len - (len % 8);
    for (int64_t e = 0; e < d8; ++e) {
        int p = (int)(e % 32);
        float a = x[e], b = y[e];
        if (ip) {
            acc[p] = fmaf(a, b, acc[p]);
        } else {
            float c = a - b;
            acc[p] = fmaf(c, c, acc[p]);
        }
    }
    for (int l = 0; l < 8; ++l) t[l] = (acc[l] + acc[8 + l]) + (acc[16 + l] + acc[24 + l]);
    int64_t r = len % 8;
    if (r != 0) {
        for (int l = 0; l < 8; ++l) {
            float a = l < r ? x[d8 + l] : 0.0f;
            float b = l < r ? y[d8 + l] : 0.0f;
            if (ip) {
                t[l] = fmaf(a, b, t[l]);
            } else {
                float c = a - b;
                t[l] = fmaf(c, c, t[l]);
            }
        }
    }
}

static void chunk_vec_cos(const float *x, const float *y, int64_t len, float cx[8], float cy[8],
                          float cxy[8]) {
    float nx[16], ny[16], xy[16];
    for (int p = 0; p < 16; ++p) nx[p] = ny[p] = xy[p] = 0.0f;
    int64_t d8 = len - (len % 8);
    for (int64_t e = 0; e < d8; ++e) {
        int p = (int)(e % 16);
        float a = x[e], b = y[e];
        nx[p] = fmaf(a, a, nx[p]);
        ny[p] = fmaf(b, b, ny[p]);
        xy[p] = fmaf(a, b, xy[p]);
    }
    for (int l = 0; l < 8; ++l) {
        cx[l] = nx[l] + nx[8 + l];
        cy[l] = ny[l] + ny[8 + l];
        cxy[l] = xy[l] + xy[8 + l];
    }
    int64_t r = len % 8;
    if (r != 0) {
        for (int l = 0; l < 8; ++l) {
            float a = l < r ? x[d8 + l] : 0.0f;
            float b = l < r ? y[d8 + l] : 0.0f;
            cx[l] = fmaf(a, a, cx[l]);
            cy[l] = fmaf(b, b, cy[l]);
            cxy[l] = fmaf(a, b, cxy[l]);
        }
    }
}

/* FullCosineAccumulator::sum: similarity with the subnormal-norm guard and clamp. */
static float cos_sim(const float cx[8], const float cy[8], const float cxy[8]) {
    float normx = sum_tree8(cx);
    float normy = sum_tree8(cy);
    float denom = sqrtf(normx) * sqrtf(normy);
    float prod = sum_tree8(cxy);
    if (normx < FLT_MIN || normy < FLT_MIN) return 0.0f;
    float v = prod / denom;
    return fmaxf(-1.0f, fminf(1.0f, v));
}

/* Direct PQ distances, diskann-providers/src/model/pq/fixed_chunk_pq_table.rs:32-56
 * (direct_distance_impl: full-precision vector against the pivots a code selects) and
 * :290-365 (self_distance: the pivots of two codes), metric dispatch as in
 * .../pq/distance/dynamic.rs:111-135. Per chunk the slices are reduced to an f32x8 vector
 * (chunk_vec_*), which is added lane-wise into the running accumulator (zero at the
 * start, ResumableL2/IP/Cosine::combine_with); the accumulator is reduced once at the end.
 * L2 -> sum; inner product -> -sum; cosine and cosine-normalized -> 1 - similarity. Left
 * operand: the full-precision vector (or the first code's pivots), right: the code's. */
static float pq_direct_core(int metric, const float *pivots, int64_t dim, const int64_t *offsets,
                            int64_t M, const float *q, const uint8_t *qcode,
                            const uint8_t *code) {
    float acc[8], ax[8], ay[8], axy[8];
    for (int l = 0; l < 8; ++l) acc[l] = ax[l] = ay[l] = axy[l] = 0.0f;
    const int cos = metric == DKO_COSINE || metric == DKO_COSINE_NORMALIZED;
    for (int64_t c = 0; c < M; ++c) {
        int64_t b = offsets[c], e = offsets[c + 1];
        const float *left = q ? q + b : pivots + (int64_t)qcode[c] * dim + b;
        const float *right = pivots + (int64_t)code[c] * dim + b;
        if (cos) {
            float cx[8], cy[8], cxy[8];
            chunk_vec_cos(left, right, e - b, cx, cy, cxy);
            for (int l = 0; l < 8; ++l) {
                ax[l] = ax[l] + cx[l];
                ay[l] = ay[l] + cy[l];
                axy[l] = axy[l] + cxy[l];
            }
        } else {
            float t[8];
            chunk_vec_4x1(metric == DKO_INNER_PRODUCT, left, right, e - b, t);
            for (int l = 0; l < 8; ++l) acc[l] = acc[l] + t[l];
        }
    }
    if (cos) return 1.0f - cos_sim(ax, ay, axy);
    float s = sum_tree8(acc);
    return metric == DKO_INNER_PRODUCT ? -s : s;
}

float dko_pq_direct(int metric, const float *pivots, int64_t dim, const int64_t *offsets,
                    int64_t M, const float *query, const uint8_t *code) {
    return pq_direct_core(metric, pivots, dim, offsets, M, query, NULL, code);
}

float dko_pq_qq(int metric, const float *pivots, int64_t dim, const int64_t *offsets, int64_t M,
                const uint8_t *a, const uint8_t *b) {
    return pq_direct_core(metric, pivots, dim, offsets, M, NULL, a, b);
}

/* PostOps: diskann-vector/src/distance/implementations.rs:215-405. For f32 data the
 * DistanceProvider builds the CosineNormalized functor (distance_provider.rs:266-270):
 * the inner-product loop (architecture_hook!(CosineNormalized, simd::IP),
 * implementations.rs:400) with post-op cosine_transformation(x) = 1 - x (:330-332,
 * :368-380). Only the integer element types redirect CosineNormalized to Cosine
 * (distance_provider.rs:274-299). */
float dko_distance(int metric, const float *q, const float *x, int64_t dim) {
    switch (metric) {
    case DKO_L2: return simd_4x1(0, q, x, dim);
    case DKO_INNER_PRODUCT: return -simd_4x1(1, q, x, dim);
    case DKO_COSINE_NORMALIZED: return 1.0f - simd_4x1(1, q, x, dim);
    case DKO_COSINE: return cosine_v3(q, x, dim);
    default: return NAN;
    }
}

/* square_norm, diskann-quantization/src/algorithms/kmeans/common.rs:9-65. */
float dko_square_norm(const float *x, int64_t len) {
    float s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    int64_t i = 0;
    if (i + 32 <= len) {
        float a[32];
        for (int p = 0; p < 32; ++p) a[p] = 0.0f;
        while (i + 32 <= len) {
            for (int p = 0; p < 32; ++p) a[p] = fmaf(x[i + p], x[i + p], a[p]);
            i += 32;
        }
        for (int l = 0; l < 8; ++l) s[l] = (a[l] + a[8 + l]) + (a[16 + l] + a[24 + l]);
    }
    while (i + 8 <= len) {
        for (int l = 0; l < 8; ++l) s[l] = fmaf(x[i + l], x[i + l], s[l]);
        i += 8;
    }
    int64_t r = len - i;
    if (r != 0) {
        for (int l = 0; l < 8; ++l) {
            float v = l < r ? x[i + l] : 0.0f;
            s[l] = fmaf(v, v, s[l]);
        }
    }
    return sum_tree8(s);
}

/* ------------------------------------------------------------------------------------ */
/* PQ                                                                                   */
/* ------------------------------------------------------------------------------------ */

/* ChunkOffsets::partition, diskann-quantization/src/views.rs:233-302: the first
 * dim % M chunks get one extra dimension. */
void dko_chunk_offsets(int64_t dim, int64_t nchunks, int64_t *offsets) {
    int64_t off = 0;
    offsets[0] = 0;
    for (int64_t c = 0; c < nchunks; ++c) {
        off += dim / nchunks;
        if (c < dim % nchunks) off += 1;
        offsets[c + 1] = off;
    }
}

/* populate_chunk_distances_impl, diskann-providers/src/model/pq/fixed_chunk_pq_table.rs:
 * 149-184: LUT[m*K + c] = SquaredL2(q_m, pivot_c,m) (TableL2) or InnerProduct (TableIP,
 * negated). CosineNormalized uses the L2 table (pq/distance/dynamic.rs:73-81). */
void dko_pq_lut(int metric, const float *pivots, int64_t K, int64_t dim,
                const int64_t *offsets, int64_t M, const float *query, float *lut) {
    int ip = metric == DKO_INNER_PRODUCT;
    for (int64_t c = 0; c < K; ++c) {
        for (int64_t m = 0; m < M; ++m) {
            int64_t s = offsets[m], e = offsets[m + 1];
            float v = simd_4x1(ip, query + s, pivots + c * dim + s, e - s);
            lut[m * K + c] = ip ? -v : v;
        }
    }
}

/* TransposedTable::process_into, diskann-quantization/src/product/tables/transposed/
 * table.rs:285-313 with the compute kernels of pivots.rs:785-820 (a fold over the chunk's
 * dimensions from a zero accumulator) and :915-948 (SquaredL2: d = p - q, acc = fma(d, d,
 * acc); InnerProduct: acc = fma(-q, p, acc)). The disk index's quantizer_preprocess
 * (diskann-disk/src/search/pq/quantizer_preprocess.rs:32-50) uses the L2 kernel for L2,
 * cosine and cosine-normalized indexes. */
void dko_pq_lut_transposed(int metric, const float *pivots, int64_t K, int64_t dim,
                           const int64_t *offsets, int64_t M, const float *query, float *lut) {
    const int ip = metric == DKO_INNER_PRODUCT;
    for (int64_t m = 0; m < M; ++m) {
        const int64_t s = offsets[m], e = offsets[m + 1];
        for (int64_t c = 0; c < K; ++c) {
            const float *p = pivots + c * dim;
            float acc = 0.0f;
            for (int64_t d = s; d < e; ++d) {
                if (ip) {
                    acc = fmaf(-query[d], p[d], acc);
                } else {
                    const float t = p[d] - query[d];
                    acc = fmaf(t, t, acc);
                }
            }
            lut[m * K + c] = acc;
        }
    }
}

/* compute_closest_centers, diskann-disk/src/utils/math_util.rs:246-385, with
 * compute_closest_centers_in_block :133-238: norms by FastL2NormSquared (the inner
 * product's simd order, diskann-vector/src/norm.rs:45-56), the three sgemm steps
 * d = (|x|^2 + |c|^2) + (-2 <x, c>), then k = 1: the first index with the strictly smallest
 * d (the in-block loop from f32::MAX); k > 1: ascending d by the PivotContainer heap, whose
 * order among equal d is unspecified (here: lower index first). The dot product inside
 * faer's sgemm (third-party, its blocking order is not pinned) is an fma chain in dimension
 * order here. */
void dko_closest_centers(const float *x, int64_t n, int64_t dim, const float *c, int64_t K,
                         int64_t k, uint32_t *out) {
    float *cn = (float *)malloc(sizeof(float) * (size_t)K);
    float *d = (float *)malloc(sizeof(float) * (size_t)K);
    int64_t *ord = (int64_t *)malloc(sizeof(int64_t) * (size_t)K);
    for (int64_t j = 0; j < K; ++j) cn[j] = simd_4x1(1, c + j * dim, c + j * dim, dim);
    for (int64_t i = 0; i < n; ++i) {
        const float *xi = x + i * dim;
        const float xn = simd_4x1(1, xi, xi, dim);
        for (int64_t j = 0; j < K; ++j) {
            float dot = 0.0f;
            for (int64_t e = 0; e < dim; ++e) dot = fmaf(xi[e], c[j * dim + e], dot);
            d[j] = (xn + cn[j]) + (-2.0f * dot);
        }
        if (k == 1) {
            float best = FLT_MAX;
            int64_t bi = 0;
            for (int64_t j = 0; j < K; ++j)
                if (d[j] < best) {
                    best = d[j];
                    bi = j;
                }
            out[i] = (uint32_t)bi;
            continue;
        }
        /* stable insertion sort of the indices by d */
        for (int64_t j = 0; j < K; ++j) {
            int64_t t = j;
            while (t > 0 && d[ord[t - 1]] > d[j]) {
                ord[t] = ord[t - 1];
                t--;
            }
            ord[t] = j;
        }
        for (int64_t t = 0; t < k; ++t) out[i * k + t] = (uint32_t)ord[t];
    }
    free(cn);
    free(d);
    free(ord);
}

/* pq_dist_lookup_single, fixed_chunk_pq_table.rs:79-95: sequential sum in chunk order. */
float dko_pq_adc(const uint8_t *code, const float *lut, int64_t M, int64_t K) {
    float acc = 0.0f;
    for (int64_t m = 0; m < M; ++m) acc += lut[m * K + code[m]];
    return acc;
}

/* BasicTable::compress_into, diskann-quantization/src/product/tables/basic.rs:161-193:
 * per chunk argmin over pivots of SquaredL2 with strict <, first index wins. An
 * infinite/NaN minimum is an error (returns the failing row + 1). */
int dko_pq_encode(const float *pivots, int64_t K, int64_t dim, const int64_t *offsets,
                  int64_t M, const float *x, int64_t n, uint8_t *codes) {
    for (int64_t i = 0; i < n; ++i) {
        const float *v = x + i * dim;
        for (int64_t m = 0; m < M; ++m) {
            int64_t s = offsets[m], e = offsets[m + 1];
            float best = INFINITY;
            int64_t arg = 0;
            for (int64_t c = 0; c < K; ++c) {
                float d = simd_4x1(0, v + s, pivots + c * dim + s, e - s);
                if (d < best) {
                    best = d;
                    arg = c;
                }
            }
            if (!isfinite(best)) return (int)(i + 1);
            codes[i * M + m] = (uint8_t)arg;
        }
    }
    return 0;
}

/* ------------------------------------------------------------------------------------ */
/* Neighbor priority queue, diskann/src/neighbor/queue.rs:68-475                        */
/* ------------------------------------------------------------------------------------ */

typedef struct {
    uint32_t *id;
    float *d;
    uint8_t *vis;
    int64_t size, cap, cursor, search_l;
} pqueue;

static void q_init(pqueue *q, int64_t cap) {
    q->id = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)(cap + 1));
    q->d = (float *)malloc(sizeof(float) * (size_t)(cap + 1));
    q->vis = (uint8_t *)malloc((size_t)(cap + 1));
    q->size = 0;
    q->cap = cap;
    q->cursor = 0;
    q->search_l = cap;
}
static void q_free(pqueue *q) {
    free(q->id);
    free(q->d);
    free(q->vis);
}

/* insert, queue.rs:130-171; lower bound = first index with d >= new.d (:229-280). */
static void q_insert(pqueue *q, uint32_t id, float d) {
    if (isnan(d)) return;
    if (q->size == q->cap && q->d[q->size - 1] < d) return;
    int64_t idx = 0;
    if (q->size > 0) {
        idx = q->size;
        for (int64_t i = 0; i < q->size; ++i) {
            if (q->d[i] >= d) {
                idx = i;
                break;
            }
        }
    }
    if (q->size == q->cap) q->size -= 1;
    int64_t n = q->size - idx;
    if (n > 0) {
        memmove(q->id + idx + 1, q->id + idx, sizeof(uint32_t) * (size_t)n);
        memmove(q->d + idx + 1, q->d + idx, sizeof(float) * (size_t)n);
        memmove(q->vis + idx + 1, q->vis + idx, (size_t)n);
    }
    q->id[idx] = id;
    q->d[idx] = d;
    q->vis[idx] = 0;
    q->size += 1;
    if (idx < q->cursor) q->cursor = idx;
}

static int q_has_notvisited(const pqueue *q) {
    int64_t lim = q->search_l < q->size ? q->search_l : q->size;
    return q->cursor < lim;
}

/* closest_notvisited, queue.rs:297-313. */
static int64_t q_closest_notvisited(pqueue *q) {
    if (!q_has_notvisited(q)) return -1;
    int64_t cur = q->cursor;
    q->vis[cur] = 1;
    q->cursor += 1;
    while (q->cursor < q->size && q->vis[q->cursor]) q->cursor += 1;
    return cur;
}

/* ------------------------------------------------------------------------------------ */
/* Visited set (exact membership; NotInMut, diskann/src/graph/glue.rs:524-561)          */
/* ------------------------------------------------------------------------------------ */

typedef struct {
    uint32_t *slot; /* stores id+1; 0 = empty */
    int64_t cap, n;
} idset;

static void set_init(idset *s, int64_t cap) {
    int64_t c = 1024;
    while (c < cap) c <<= 1;
    s->slot = (uint32_t *)calloc((size_t)c, sizeof(uint32_t));
    s->cap = c;
    s->n = 0;
}
static void set_free(idset *s) { free(s->slot); }
static inline uint32_t hash32(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7feb352dU;
    x ^= x >> 15;
    x *= 0x846ca68bU;
    x ^= x >> 16;
    return x;
}
static int set_insert(idset *s, uint32_t id);
static void set_grow(idset *s) {
    uint32_t *old = s->slot;
    int64_t oc = s->cap;
    s->cap = oc * 2;
    s->slot = (uint32_t *)calloc((size_t)s->cap, sizeof(uint32_t));
    s->n = 0;
    for (int64_t i = 0; i < oc; ++i)
        if (old[i]) set_insert(s, old[i] - 1);
    free(old);
}
/* returns 1 if newly inserted */
static int set_insert(idset *s, uint32_t id) {
    if (2 * (s->n + 1) > s->cap) set_grow(s);
    uint64_t mask = (uint64_t)s->cap - 1;
    uint64_t h = hash32(id) & mask;
    for (;;) {
        uint32_t v = s->slot[h];
        if (v == 0) {
            s->slot[h] = id + 1;
            s->n++;
            return 1;
        }
        if (v == id + 1) return 0;
        h = (h + 1) & mask;
    }
}
static void set_clear(idset *s) {
    memset(s->slot, 0, sizeof(uint32_t) * (size_t)s->cap);
    s->n = 0;
}

/* ------------------------------------------------------------------------------------ */
/* Search                                                                               */
/* ------------------------------------------------------------------------------------ */

int dko_knn_validate(int64_t k, int64_t L, int64_t bw) {
    if (k == 0) return DKO_ERR_KZERO;
    if (L == 0) return DKO_ERR_LZERO;
    if (k > L) return DKO_ERR_L_LESS_THAN_K;
    if (bw == 0) return DKO_ERR_BEAM_ZERO;
    if (k < 0 || L < 0 || bw < 0) return DKO_ERR_ARG;
    return DKO_OK;
}

typedef struct {
    pqueue q;
    idset vis;
    uint32_t *beam;
    uint32_t *nb_id;
    float *nb_d;
    float *lut;
    int64_t nb_cap;
} scratch;

static void scratch_init(scratch *s, const dko_index *ix, int64_t L, int64_t bw) {
    q_init(&s->q, L + ix->nstart);
    set_init(&s->vis, 4096);
    s->beam = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)bw);
    s->nb_cap = bw * ix->max_degree + 1;
    s->nb_id = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)s->nb_cap);
    s->nb_d = (float *)malloc(sizeof(float) * (size_t)s->nb_cap);
    s->lut = ix->sph  ? (float *)malloc((size_t)ix->sph->out_dim + 64)
             : ix->sq ? (float *)malloc((size_t)ix->sq->dim + 2 * sizeof(float))
             : ix->M > 0 ? (float *)malloc(sizeof(float) * (size_t)(ix->M * ix->K))
                         : NULL;
}
static void scratch_free(scratch *s) {
    q_free(&s->q);
    set_free(&s->vis);
    free(s->beam);
    free(s->nb_id);
    free(s->nb_d);
    free(s->lut);
}

/* ------------------------------------------------------------------------------------
 * Scalar quantization, 8 bits
 * ---------------------------------------------------------------------------------- */

/* ScalarQuantizationParameters::train (diskann-quantization/src/scalar/train.rs:75-88)
 * with compute_means_and_average_norm and compute_variances (src/utils.rs:109-197): f64
 * column sums in row order, row norms as sequential f64 sums of squares, the norm sum in
 * row order; range p = sqrt(max variance) * stdevs, scale = 2p, shift = mean - p. */
void dko_sq_train(const float *x, int64_t n, int64_t dim, int64_t stride, double stdevs,
                  float *scale, float *shift, float *mean_norm) {
    double *mean = (double *)calloc((size_t)dim, sizeof(double));
    double *var = (double *)calloc((size_t)dim, sizeof(double));
    double norm_sum = 0.0;
    for (int64_t r = 0; r < n; ++r) {
        const float *row = x + r * stride;
        double sq = 0.0;
        for (int64_t j = 0; j < dim; ++j) {
            const double v = (double)row[j];
            mean[j] += v;
            sq += v * v;
        }
        norm_sum += sqrt(sq);
    }
    for (int64_t j = 0; j < dim; ++j) mean[j] /= (double)n;
    for (int64_t r = 0; r < n; ++r) {
        const float *row = x + r * stride;
        for (int64_t j = 0; j < dim; ++j) {
            const double d = (double)row[j] - mean[j];
            var[j] += d * d;
        }
    }
    double mx = 0.0;
    for (int64_t j = 0; j < dim; ++j) {
        var[j] /= (double)n;
        mx = fmax(mx, var[j]);
    }
    const double p = sqrt(mx) * stdevs;
    *scale = (float)(2.0 * p);
    for (int64_t j = 0; j < dim; ++j) shift[j] = (float)(mean[j] - p);
    *mean_norm = (float)(norm_sum / (double)n);
    free(mean);
    free(var);
}

/* ScalarQuantizer::new: shift_square_norm = InnerProduct::evaluate(shift, shift), the
 * mathematical value of the f32 inner product (quantizer.rs:116-127). */
float dko_sq_shift_square_norm(const float *shift, int64_t dim) {
    return -dko_distance(DKO_INNER_PRODUCT, shift, shift, dim);
}

/* ScalarQuantizer::compress + the compensation of CompressInto<MutCompensatedVectorRef>
 * (quantizer.rs:190-240, 407-432): code = round(clamp((x - shift) * (255 / scale), 0,
 * 255)) (round half away from zero), compensation = scale * (1/255) * dot with
 * dot = fma(code, shift, dot) in dimension order. */
int dko_sq_compress(float scale, const float *shift, int64_t dim, const float *x,
                    uint8_t *code, float *comp) {
    const float inv = 255.0f / scale;
    float dot = 0.0f;
    int nan = 0;
    for (int64_t i = 0; i < dim; ++i) {
        const float f = x[i];
        nan |= f != f;
        float c = (f - shift[i]) * inv;
        if (c < 0.0f) c = 0.0f;
        else if (c > 255.0f) c = 255.0f;
        c = roundf(c);
        dot = fmaf(c, shift[i], dot);
        code[i] = c != c ? 0 : (uint8_t)c;
    }
    *comp = scale * (1.0f / 255.0f) * dot;
    return nan;
}

/* ScalarQuantizer::rescale (quantizer.rs:300-313): x *= to_norm / |x| unless |x| = 0. */
void dko_sq_rescale(float *x, int64_t dim, float to_norm) {
    const float nsq = -dko_distance(DKO_INNER_PRODUCT, x, x, dim);
    const float norm = sqrtf(nsq);
    if (norm == 0.0f) return;
    const float s = to_norm / norm;
    for (int64_t i = 0; i < dim; ++i) x[i] *= s;
}

/* CompensatedSquaredL2 / CompensatedIP / CompensatedCosineNormalized
 * (diskann-quantization/src/scalar/vectors.rs:182-470) on 8-bit codes: exact integer
 * sums, then bs = (1/255)^2 in f32 and
 *   L2: (bs * scale^2) * sum (a - b)^2
 *   IP: -(fma(bs * scale^2, sum a b, shift_square_norm) + (comp_b + comp_a))
 *   CosineNormalized: 1 - (1 - ((bs * scale^2) * sum (a - b)^2) / 2). */
float dko_sq_distance(int metric, float scale, float shift_square_norm, const uint8_t *a,
                      float comp_a, const uint8_t *b, float comp_b, int64_t dim) {
    const float inv = 1.0f / 255.0f;
    const float bs = inv * inv;
    const float ss = scale * scale;
    if (metric == DKO_INNER_PRODUCT) {
        uint32_t ip = 0;
        for (int64_t i = 0; i < dim; ++i) ip += (uint32_t)a[i] * (uint32_t)b[i];
        const float r = fmaf(bs * ss, (float)ip, shift_square_norm) + (comp_b + comp_a);
        return -r;
    }
    uint32_t l2 = 0;
    for (int64_t i = 0; i < dim; ++i) {
        const int32_t d = (int32_t)a[i] - (int32_t)b[i];
        l2 += (uint32_t)(d * d);
    }
    const float v = bs * ss * (float)l2;
    if (metric == DKO_COSINE_NORMALIZED) return 1.0f - (1.0f - v / 2.0f);
    return v;
}

/* ------------------------------------------------------------------------------------
 * Spherical quantization, 1 bit
 * ---------------------------------------------------------------------------------- */

/* f32 -> f16 round to nearest even (the meta conversion of DataMeta::new,
 * diskann-quantization/src/spherical/vectors.rs:261-294) and back (exact). */
uint16_t dko_f32_to_f16(float f) {
    uint32_t x;
    memcpy(&x, &f, 4);
    const uint32_t sign = (x >> 16) & 0x8000u;
    const uint32_t ex = (x >> 23) & 0xFFu;
    uint32_t mant = x & 0x7FFFFFu;
    if (ex == 0xFFu) return (uint16_t)(sign | 0x7C00u | (mant ? 0x200u : 0u));
    const int32_t e = (int32_t)ex - 127 + 15;
    if (e >= 31) return (uint16_t)(sign | 0x7C00u);
    if (e <= 0) {
        if (e < -10) return (uint16_t)sign;
        mant |= 0x800000u;
        const uint32_t sh = (uint32_t)(14 - e);
        uint32_t h = mant >> sh;
        const uint32_t rem = mant & ((1u << sh) - 1u), hw = 1u << (sh - 1u);
        if (rem > hw || (rem == hw && (h & 1u))) h++;
        return (uint16_t)(sign | h);
    }
    uint32_t h = ((uint32_t)e << 10) | (mant >> 13);
    const uint32_t rem = mant & 0x1FFFu;
    if (rem > 0x1000u || (rem == 0x1000u && (h & 1u))) h++;
    return (uint16_t)(sign | h);
}

float dko_f16_to_f32(uint16_t h) {
    const uint32_t sign = (uint32_t)(h & 0x8000u) << 16;
    const uint32_t ex = (h >> 10) & 0x1Fu, mant = h & 0x3FFu;
    uint32_t x;
    if (ex == 0x1Fu) {
        x = sign | 0x7F800000u | (mant << 13);
    } else if (ex == 0) {
        if (mant == 0) {
            x = sign;
        } else {
            const float v = (float)mant * (1.0f / 16777216.0f);   /* mant * 2^-24, exact */
            memcpy(&x, &v, 4);
            x |= sign;
        }
    } else {
        x = sign | ((ex + 112u) << 23) | (mant << 13);
    }
    float f;
    memcpy(&f, &x, 4);
    return f;
}

void dko_sph_train_stats(int metric, const float *x, int64_t n, int64_t dim, float *centroid,
                         float *mean_norm) {
    double *m = (double *)calloc((size_t)dim, sizeof(double));
    double norm_sum = 0.0;
    for (int64_t r = 0; r < n; ++r) {
        const float *row = x + r * dim;
        double sq = 0.0;
        for (int64_t j = 0; j < dim; ++j) sq += (double)row[j] * (double)row[j];
        const double norm = sqrt(sq);
        if (metric == DKO_COSINE) {
            /* compute_normalized_means (src/utils.rs:141-178) */
            const double inv = norm == 0.0 ? 1.0 : 1.0 / norm;
            for (int64_t j = 0; j < dim; ++j) m[j] += (double)row[j] * inv;
        } else {
            for (int64_t j = 0; j < dim; ++j) m[j] += (double)row[j];
            norm_sum += norm;
        }
    }
    for (int64_t j = 0; j < dim; ++j) centroid[j] = (float)(m[j] / (double)n);
    *mean_norm = metric == DKO_COSINE ? 1.0f : (float)(norm_sum / (double)n);
    free(m);
}

int64_t dko_sph_data_bytes(const dko_sph *q) { return (q->out_dim + 7) / 8 + 6; }

/* hadamard_transform (diskann-quantization/src/algorithms/hadamard.rs) as the x86-64-v3
 * build runs it: recursive halving with (l + r, l - r) butterflies; a 64-element block is
 * done by micro_kernel_64 (eight 8-point transforms as sequential +-sums of the block's
 * 8-element groups, then butterflies at strides 8, 16, 32); finally x *= 1/sqrt(len). */
static void had_rec(float *x, int64_t len) {
    if (len == 64) {
        float d[8][8];
        for (int j = 0; j < 8; ++j)
            for (int k = 0; k < 8; ++k) {
                float acc = 0.0f;
                for (int m = 0; m < 8; ++m) {
                    /* H8[m][k] = (-1)^popcount(m & k) */
                    const float hv = (__builtin_popcount((unsigned)(m & k)) & 1) ? -1.0f : 1.0f;
                    acc = fmaf(x[8 * j + m], hv, acc);
                }
                d[j][k] = acc;
            }
        for (int k = 0; k < 8; ++k) {
            const float e0 = d[0][k] + d[1][k], e1 = d[0][k] - d[1][k];
            const float e2 = d[2][k] + d[3][k], e3 = d[2][k] - d[3][k];
            const float e4 = d[4][k] + d[5][k], e5 = d[4][k] - d[5][k];
            const float e6 = d[6][k] + d[7][k], e7 = d[6][k] - d[7][k];
            const float f0 = e0 + e2, f1 = e1 + e3, f2 = e0 - e2, f3 = e1 - e3;
            const float f4 = e4 + e6, f5 = e5 + e7, f6 = e4 - e6, f7 = e5 - e7;
            x[k] = f0 + f4;
            x[8 + k] = f1 + f5;
            x[16 + k] = f2 + f6;
            x[24 + k] = f3 + f7;
            x[32 + k] = f0 - f4;
            x[40 + k] = f1 - f5;
            x[48 + k] = f2 - f6;
            x[56 + k] = f3 - f7;
        }
        return;
    }
    if (len == 2) {
        const float l = x[0], r = x[1];
        x[0] = l + r;
        x[1] = l - r;
        return;
    }
    const int64_t h = len / 2;
    had_rec(x, h);
    had_rec(x + h, h);
    for (int64_t i = 0; i < h; ++i) {
        const float a = x[i] + x[h + i], b = x[i] - x[h + i];
        x[i] = a;
        x[h + i] = b;
    }
}

void dko_sph_hadamard(float *x, int64_t len) {
    if (len <= 1) return;
    had_rec(x, len);
    const float m = 1.0f / sqrtf((float)len);
    for (int64_t i = 0; i < len; ++i) x[i] *= m;
}

static float flip(float v, uint32_t sign) {
    uint32_t u;
    memcpy(&u, &v, 4);
    u ^= sign;
    memcpy(&v, &u, 4);
    return v;
}

/* PaddingHadamard / DoubleHadamard / NullTransform ::transform_into
 * (algorithms/transforms/{padding_hadamard.rs:204-272, double_hadamard.rs:177-230,
 * null.rs:38-42}). */
void dko_sph_transform(const dko_sph *q, const float *src, float *dst) {
    if (q->kind == DKO_SPH_NULL) {
        memcpy(dst, src, sizeof(float) * (size_t)q->dim);
        return;
    }
    int64_t n;
    float *tmp;
    if (q->kind == DKO_SPH_PADDING_HADAMARD) {
        n = q->padded_dim;
        tmp = (float *)calloc((size_t)n, sizeof(float));
        for (int64_t i = 0; i < q->dim; ++i) tmp[i] = flip(src[i], q->signs0[i]);
        dko_sph_hadamard(tmp, n);
    } else {
        n = q->dim > q->out_dim ? q->dim : q->out_dim;
        if (q->subsample) n = q->dim;  /* subsampling keeps the input dimension */
        tmp = (float *)calloc((size_t)n, sizeof(float));
        for (int64_t i = 0; i < q->dim; ++i) tmp[i] = flip(src[i], q->signs0[i]);
        int64_t split = 1;
        while (split * 2 <= n) split *= 2;
        dko_sph_hadamard(tmp, split);
        for (int64_t i = 0; i < n; ++i) tmp[i] = flip(tmp[i], q->signs1[i]);
        dko_sph_hadamard(tmp + (n - split), split);
    }
    if (q->subsample) {
        const float rescale = sqrtf((float)n / (float)q->nsub);
        for (int64_t i = 0; i < q->nsub; ++i) dst[i] = tmp[q->subsample[i]] * rescale;
    } else {
        memcpy(dst, tmp, sizeof(float) * (size_t)n);
    }
    free(tmp);
}

static int sph_preserves_norms(const dko_sph *q) { return q->subsample == NULL; }

/* SphericalQuantizer::rescale (spherical/quantizer.rs:322-326): v *= mean_norm / |v|. */
void dko_sph_rescale(const dko_sph *q, float *x) {
    const float norm = sqrtf(simd_4x1(1, x, x, q->dim));
    const float m = q->mean_norm / norm;
    for (int64_t i = 0; i < q->dim; ++i) x[i] *= m;
}

/* SphericalQuantizer::preprocess (spherical/quantizer.rs:333-377) then normalisation and
 * the transform. Returns the transformed vector (out_dim, malloc'd) or NULL when the
 * shifted norm is zero; *sn = shifted norm, *ms = metric term. */
static float *sph_prepare(const dko_sph *q, const float *x, float *sn, float *ms, int *err) {
    *err = 0;
    float mul = q->pre_scale;
    if (q->metric == DKO_COSINE) {
        const float norm = sqrtf(simd_4x1(1, x, x, q->dim));
        mul = norm == 0.0f ? 1.0f : 1.0f / norm;
    }
    float *sh = (float *)malloc(sizeof(float) * (size_t)q->dim);
    for (int64_t i = 0; i < q->dim; ++i) sh[i] = mul * x[i] - q->shift[i];
    const float norm = sqrtf(simd_4x1(1, sh, sh, q->dim));
    if (!isfinite(norm)) {
        *err = 1;
        free(sh);
        return NULL;
    }
    *sn = norm;
    *ms = q->metric == DKO_L2 ? norm * norm : simd_4x1(1, sh, q->shift, q->dim);
    if (norm == 0.0f) {
        free(sh);
        return NULL;
    }
    for (int64_t i = 0; i < q->dim; ++i) sh[i] /= norm;
    float *t = (float *)malloc(sizeof(float) * (size_t)(q->out_dim + 1));
    dko_sph_transform(q, sh, t);
    free(sh);
    return t;
}

static void put_u16(uint8_t *p, uint16_t v) {
    p[0] = (uint8_t)(v & 0xFF);
    p[1] = (uint8_t)(v >> 8);
}
static uint16_t get_u16(const uint8_t *p) { return (uint16_t)(p[0] | (p[1] << 8)); }

/* CompressIntoWith<DataMut<1>> + FinishCompressing for 1 bit (spherical/quantizer.rs:
 * 591-642, 779-857): bit = t > 0, quant_raw_ip = sequential sum of |t|, correction =
 * 2 |t| |shifted| / quant_raw_ip; meta stored as f16, f16, u16. */
int dko_sph_compress_data(const dko_sph *q, const float *x, uint8_t *out) {
    const int64_t nb = (q->out_dim + 7) / 8;
    memset(out, 0, (size_t)(nb + 6));
    float sn = 0.0f, ms = 0.0f;
    int err;
    float *t = sph_prepare(q, x, &sn, &ms, &err);
    if (err) return err;
    if (!t) return 0;  /* the centroid itself: default meta, no bits */
    const float tn = sph_preserves_norms(q) ? 1.0f : sqrtf(simd_4x1(1, t, t, q->out_dim));
    float raw = 0.0f;
    uint32_t bits = 0;
    for (int64_t i = 0; i < q->out_dim; ++i) {
        const int b = t[i] > 0.0f;
        raw += fabsf(t[i]);
        bits += (uint32_t)b;
        if (b) out[i >> 3] |= (uint8_t)(1u << (i & 7));
    }
    free(t);
    const float ipc = 2.0f * tn * sn / raw;
    const uint16_t h0 = dko_f32_to_f16(ipc), h1 = dko_f32_to_f16(ms);
    if ((h0 & 0x7C00u) == 0x7C00u || (h1 & 0x7C00u) == 0x7C00u || bits > 0xFFFFu) return 2;
    put_u16(out + nb, h0);
    put_u16(out + nb + 2, h1);
    put_u16(out + nb + 4, (uint16_t)bits);
    return 0;
}

/* CompressIntoWith<QueryMut<4, _>> (spherical/quantizer.rs:1101-1189): codes
 * round((t - min) / scale) clamped to [0, 15] with scale = (max - min) / 15. */
int dko_sph_compress_query4(const dko_sph *q, const float *x, uint8_t *codes, float *meta) {
    memset(codes, 0, (size_t)q->out_dim);
    meta[0] = meta[1] = meta[2] = meta[3] = 0.0f;
    float sn = 0.0f, ms = 0.0f;
    int err;
    float *t = sph_prepare(q, x, &sn, &ms, &err);
    if (err) return err;
    if (!t) return 0;
    float mn = FLT_MAX, mx = -FLT_MAX;
    for (int64_t i = 0; i < q->out_dim; ++i) {
        mn = fminf(t[i], mn);
        mx = fmaxf(t[i], mx);
    }
    const float scale = (mx - mn) / 15.0f;
    float bs = 0.0f;
    for (int64_t i = 0; i < q->out_dim; ++i) {
        float c = roundf((t[i] - mn) / scale);
        if (c < 0.0f) c = 0.0f;
        else if (c > 15.0f) c = 15.0f;
        bs += c;
        codes[i] = (uint8_t)c;
    }
    free(t);
    meta[0] = sn * scale;
    meta[1] = bs;
    meta[2] = mn / scale;
    meta[3] = ms;
    return 0;
}

/* CompressIntoWith<FullQueryMut> (spherical/quantizer.rs:702-777). */
int dko_sph_compress_full(const dko_sph *q, const float *x, float *v, float *meta) {
    memset(v, 0, sizeof(float) * (size_t)q->out_dim);
    meta[0] = meta[1] = meta[2] = 0.0f;
    float sn = 0.0f, ms = 0.0f;
    int err;
    float *t = sph_prepare(q, x, &sn, &ms, &err);
    if (err) return err;
    if (!t) return 0;
    float sum = 0.0f;
    for (int64_t i = 0; i < q->out_dim; ++i) {
        v[i] = t[i];
        sum += t[i];
    }
    free(t);
    meta[0] = sum;
    meta[1] = sn;
    meta[2] = ms;
    return 0;
}

static void data_meta(const dko_sph *q, const uint8_t *c, float *ipc, float *ms, float *bs) {
    const int64_t nb = (q->out_dim + 7) / 8;
    *ipc = dko_f16_to_f32(get_u16(c + nb));
    *ms = dko_f16_to_f32(get_u16(c + nb + 2));
    *bs = (float)get_u16(c + nb + 4);
}

static float sph_shift_sq(const dko_sph *q) { return simd_4x1(1, q->shift, q->shift, q->dim); }

/* the metric's distance from the mathematical inner-product / L2 value */
static float sph_finish(const dko_sph *q, float v) {
    if (q->metric == DKO_L2) return v;
    if (q->metric == DKO_COSINE) return 1.0f - v;
    return -v;
}

/* CompensatedSquaredL2 / CompensatedIP / CompensatedCosine on two data codes
 * (spherical/vectors.rs:493-535, 570-595, 730-759). */
float dko_sph_dist_data(const dko_sph *q, const uint8_t *a, const uint8_t *b) {
    const int64_t nb = (q->out_dim + 7) / 8;
    uint32_t ip = 0;
    for (int64_t i = 0; i < nb; ++i) ip += (uint32_t)__builtin_popcount((unsigned)(a[i] & b[i]));
    float xi, xm, xb, yi, ym, yb;
    data_meta(q, a, &xi, &xm, &xb);
    data_meta(q, b, &yi, &ym, &yb);
    const float dim = (float)q->out_dim;
    const float k = xi * yi * ((float)ip - 0.5f * (xb + yb) + 0.25f * dim);
    if (q->metric == DKO_L2) return xm + ym - 2.0f * k;
    return sph_finish(q, xm + ym + k + sph_shift_sq(q));
}

/* QueryRef<4> x DataRef<1> (spherical/vectors.rs:597-641, 761-806). */
float dko_sph_dist_query4(const dko_sph *q, const uint8_t *codes, const float *meta,
                          const uint8_t *b) {
    uint32_t ip = 0;
    for (int64_t i = 0; i < q->out_dim; ++i)
        if (b[i >> 3] & (1u << (i & 7))) ip += codes[i];
    float yi, ym, yb;
    data_meta(q, b, &yi, &ym, &yb);
    const float dim = (float)q->out_dim;
    const float cip = yi * meta[0] * ((float)ip - 0.5f * meta[1] + meta[2] * yb -
                                      0.5f * meta[2] * dim);
    if (q->metric == DKO_L2) return ym + meta[3] - 2.0f * cip;
    return sph_finish(q, cip + ym + meta[3] + sph_shift_sq(q));
}

/* InnerProduct(&[f32], USlice<1>) as the x86-64-v3 build computes it
 * (diskann-quantization/src/bits/distances.rs:2295-2436): two f32x8 accumulators
 * alternating over the four 8-lane groups of each 32-element block, combined, then the
 * 8-lane groups of the remainder, then sum_tree. */
static float ip_f32_bits(const float *x, const uint8_t *bits, int64_t len) {
    float s[8] = {0}, s0[8] = {0}, s1[8] = {0};
    const int64_t blocks = len / 32;
    if (blocks > 0) {
        for (int64_t i = 0; i < blocks; ++i)
            for (int v = 0; v < 4; ++v)
                for (int l = 0; l < 8; ++l) {
                    const int64_t e = 32 * i + 8 * v + l;
                    const float bv = (bits[e >> 3] >> (e & 7)) & 1 ? 1.0f : 0.0f;
                    float *acc = (v & 1) ? s1 : s0;
                    acc[l] = fmaf(x[e], bv, acc[l]);
                }
        for (int l = 0; l < 8; ++l) s[l] = s0[l] + s1[l];
    }
    const int64_t rem = len % 32;
    if (rem) {
        const int64_t tail = len % 8 == 0 ? 8 : len % 8;
        const int nv = rem <= 8 ? 1 : rem <= 16 ? 2 : rem <= 24 ? 3 : 4;
        for (int v = 0; v < nv; ++v) {
            const int64_t cnt = v == nv - 1 ? tail : 8;
            for (int l = 0; l < 8; ++l) {
                const int64_t e = 32 * blocks + 8 * v + l;
                const float xv = l < cnt ? x[e] : 0.0f;
                const float bv = l < cnt && ((bits[e >> 3] >> (e & 7)) & 1) ? 1.0f : 0.0f;
                s[l] = fmaf(xv, bv, s[l]);
            }
        }
    }
    return sum_tree8(s);
}

/* FullQueryRef x DataRef<1> (spherical/vectors.rs:643-691, 808-850). */
float dko_sph_dist_full(const dko_sph *q, const float *v, const float *meta, const uint8_t *b) {
    const float s = ip_f32_bits(v, b, q->out_dim);
    float yi, ym, yb;
    data_meta(q, b, &yi, &ym, &yb);
    (void)yb;
    if (q->metric == DKO_L2) {
        const float ip = s - meta[0] * 0.5f;
        return meta[2] + ym - 2.0f * meta[1] * yi * ip;
    }
    const float ip = meta[1] * yi * (s - meta[0] * 0.5f);
    return sph_finish(q, ip + meta[2] + ym + sph_shift_sq(q));
}

/* ------------------------------------------------------------------------------------
 * MinMax quantization, 8 bits
 * ---------------------------------------------------------------------------------- */

int64_t dko_minmax_bytes(const dko_sph *t) { return 20 + t->out_dim; }

/* MinMaxQuantizer::compress (diskann-quantization/src/minmax/quantizer.rs:45-160) for 8
 * bits: range = the transformed vector's min / max widened by grid_scale about its
 * middle; a = max(max - min, 1e-8) / 255; code = round(clamp((v - min) / a, 0, 255));
 * |reconstruction|^2, code sum and loss are sequential f32 sums. */
int dko_minmax_compress(const dko_sph *t, float grid_scale, const float *x, uint8_t *out,
                        float *loss) {
    const int64_t n = t->out_dim;
    float *v = (float *)malloc(sizeof(float) * (size_t)(n + 1));
    dko_sph_transform(t, x, v);
    float mn = NAN, mx = NAN;
    for (int64_t i = 0; i < n; ++i) {
        mn = fminf(mn, v[i]);
        mx = fmaxf(mx, v[i]);
    }
    const float width = (mx - mn) / 2.0f;
    const float mid = mn + width;
    const float lo = mid - width * grid_scale, hi = mid + width * grid_scale;
    const float a = fmaxf(hi - lo, 1e-8f) / 255.0f;
    float ns = 0.0f, cs = 0.0f, ls = 0.0f;
    int nan = 0;
    for (int64_t i = 0; i < n; ++i) {
        nan |= v[i] != v[i];
        float c = (v[i] - lo) / a;
        if (c < 0.0f) c = 0.0f;
        else if (c > 255.0f) c = 255.0f;
        c = roundf(c);
        const float r = c * a + lo;
        ns += r * r;
        cs += c;
        const float d = r - v[i];
        ls += d * d;
        out[20 + i] = c != c ? 0 : (uint8_t)c;
    }
    const uint32_t dim = (uint32_t)n;
    const float meta[4] = {lo, a * cs, a, ns};
    memcpy(out, &dim, 4);
    memcpy(out + 4, meta, 16);
    if (loss) *loss = ls;
    free(v);
    return nan;
}

/* CompressInto<FullQueryMut> (minmax/quantizer.rs:178-210): the transformed query with
 * its sequential sum and square norm. */
int dko_minmax_compress_full(const dko_sph *t, const float *x, float *v, float *meta) {
    for (int64_t i = 0; i < t->dim; ++i)
        if (x[i] != x[i]) return 1;
    dko_sph_transform(t, x, v);
    float ns = 0.0f, sum = 0.0f;
    for (int64_t i = 0; i < t->out_dim; ++i) ns += v[i] * v[i];
    for (int64_t i = 0; i < t->out_dim; ++i) sum += v[i];
    meta[0] = sum;
    meta[1] = ns;
    return 0;
}

static void mm_meta(const uint8_t *c, uint32_t *dim, float m[4]) {
    memcpy(dim, c, 4);
    memcpy(m, c + 4, 16);  /* b, n, a, |r|^2 */
}

static float mm_finish(int metric, float v, float xns, float yns) {
    if (metric == DKO_L2) return -2.0f * v + xns + yns;
    if (metric == DKO_INNER_PRODUCT) return -v;
    if (metric == DKO_COSINE) return 1.0f - v / (sqrtf(xns) * sqrtf(yns));
    return 1.0f - v;
}

/* MinMaxIP / L2Squared / Cosine / CosineNormalized on two data rows
 * (minmax/vectors.rs:206-229 kernel and the functors at 231-470). */
float dko_minmax_dist_data(int metric, const uint8_t *a, const uint8_t *b) {
    uint32_t dim, dim2;
    float xm[4], ym[4];
    mm_meta(a, &dim, xm);
    mm_meta(b, &dim2, ym);
    uint32_t raw = 0;
    for (uint32_t i = 0; i < dim; ++i) raw += (uint32_t)a[20 + i] * (uint32_t)b[20 + i];
    const float t0 = xm[2] * ym[2] * (float)raw;
    const float t1x = xm[1] * ym[0];
    const float t1y = ym[1] * xm[0];
    const float t2 = xm[0] * ym[0] * (float)dim;
    const float v = t0 + t1x + t1y + t2;
    return mm_finish(metric, v, xm[3], ym[3]);
}

/* FullQueryRef x DataRef<8> (minmax/vectors.rs:271-470); the f32 x u8 product is the
 * sequential sum x_i * code_i the x86-64 build runs for 8-bit codes
 * (bits/distances.rs:2675-2700). */
float dko_minmax_dist_full(int metric, const float *v, const float *meta, const uint8_t *b) {
    uint32_t dim;
    float ym[4];
    mm_meta(b, &dim, ym);
    float s = 0.0f;
    for (uint32_t i = 0; i < dim; ++i) s += v[i] * (float)b[20 + i];
    const float ip = s * ym[2] + meta[0] * ym[0];
    if (metric == DKO_L2) return meta[1] + ym[3] - 2.0f * ip;
    if (metric == DKO_INNER_PRODUCT) return -ip;
    if (metric == DKO_COSINE) return 1.0f - ip / (sqrtf(meta[1]) * sqrtf(ym[3]));
    return 1.0f - ip;
}

static inline float node_dist(const dko_index *ix, int pq, const float *query,
                              const float *lut, uint32_t id) {
    if (pq && ix->sph) {
        /* query code staged behind the LUT pointer by search_core */
        const dko_sph *q = ix->sph;
        const int64_t cb = dko_sph_data_bytes(q);
        const uint8_t *row = ix->sph_codes + (int64_t)id * cb;
        if (ix->sph_layout == 0) return dko_sph_dist_data(q, (const uint8_t *)lut, row);
        return dko_sph_dist_query4(q, (const uint8_t *)(lut + 4), lut, row);
    }
    if (pq && ix->sq) {
        /* query code / compensation staged behind the LUT pointer by search_core */
        const dko_sq *q = ix->sq;
        const uint8_t *qc = (const uint8_t *)lut;
        float qcomp;
        memcpy(&qcomp, qc + q->dim, sizeof(float));
        return dko_sq_distance(ix->metric, q->scale, q->shift_square_norm, qc, qcomp,
                               q->codes + (int64_t)id * q->dim, q->comp[id], q->dim);
    }
    if (pq) return dko_pq_adc(ix->codes + (int64_t)id * ix->M, lut, ix->M, ix->K);
    return dko_distance(ix->metric, query, ix->vectors + (int64_t)id * ix->vstride, ix->dim);
}

/* DiskANNIndex::search_internal, diskann/src/graph/index.rs:1938-2005, with the
 * FullAccessor / QuantAccessor expand_beam of
 * diskann-providers/.../inmem/full_precision.rs:274-322 and .../inmem/product.rs:304-335. */
static void search_core(const dko_index *ix, int pq, const float *query, int64_t bw,
                        scratch *s, uint32_t *rec_ids, float *rec_d, uint32_t *rec_len,
                        int64_t rec_cap, dko_stats *st) {
    pqueue *q = &s->q;
    q->size = 0;
    q->cursor = 0;
    set_clear(&s->vis);
    uint32_t cmps = 0, hops = 0, nrec = 0;
    if (pq && ix->sph) {
        /* SphericalStore::query_computer with the search strategy's allow_rescale
         * (diskann-providers/.../inmem/spherical.rs:211-232, 505-531): inner-product
         * queries are rescaled to the mean norm, then compressed in the layout */
        const dko_sph *q = ix->sph;
        float *tmp = (float *)malloc(sizeof(float) * (size_t)q->dim);
        memcpy(tmp, query, sizeof(float) * (size_t)q->dim);
        if (q->metric == DKO_INNER_PRODUCT && !ix->sph_norescale) dko_sph_rescale(q, tmp);
        if (ix->sph_layout == 0)
            dko_sph_compress_data(q, tmp, (uint8_t *)s->lut);
        else
            dko_sph_compress_query4(q, tmp, (uint8_t *)(s->lut + 4), s->lut);
        free(tmp);
    } else if (pq && ix->sq) {
        /* SQStore::query_computer (diskann-providers/.../inmem/scalar.rs:227-253) with
         * allow_rescale: the inner-product query is rescaled to the mean norm, then
         * compressed; code and compensation go into the LUT buffer */
        const dko_sq *q = ix->sq;
        float *tmp = (float *)malloc(sizeof(float) * (size_t)q->dim);
        memcpy(tmp, query, sizeof(float) * (size_t)q->dim);
        if (ix->metric == DKO_INNER_PRODUCT) dko_sq_rescale(tmp, q->dim, q->mean_norm);
        uint8_t *qc = (uint8_t *)s->lut;
        float qcomp;
        dko_sq_compress(q->scale, q->shift, q->dim, tmp, qc, &qcomp);
        memcpy(qc + q->dim, &qcomp, sizeof(float));
        free(tmp);
    } else if (pq) {
        int lut_metric = ix->metric == DKO_INNER_PRODUCT ? DKO_INNER_PRODUCT : DKO_L2;
        dko_pq_lut(lut_metric, ix->pivots, ix->K, ix->dim, ix->offsets, ix->M, query, s->lut);
    }
    for (int64_t i = 0; i < ix->nstart; ++i) {
        uint32_t sp = ix->starts[i];
        float d = node_dist(ix, pq, query, s->lut, sp);
        set_insert(&s->vis, sp);
        q_insert(q, sp, d);
        cmps += 1;
    }
    while (q_has_notvisited(q)) {
        int64_t nbeam = 0;
        while (nbeam < bw) {
            int64_t c = q_closest_notvisited(q);
            if (c < 0) break;
            if (rec_ids && nrec < rec_cap) {
                rec_ids[nrec] = q->id[c];
                rec_d[nrec] = q->d[c];
            }
            nrec++;
            s->beam[nbeam++] = q->id[c];
        }
        int64_t nn = 0;
        for (int64_t b = 0; b < nbeam; ++b) {
            uint32_t node = s->beam[b];
            const uint32_t *adj = ix->adj + (int64_t)node * ix->max_degree;
            uint32_t dg = ix->deg[node];
            for (uint32_t j = 0; j < dg; ++j) {
                uint32_t nbr = adj[j];
                if (!set_insert(&s->vis, nbr)) continue;
                s->nb_id[nn] = nbr;
                s->nb_d[nn] = node_dist(ix, pq, query, s->lut, nbr);
                nn++;
            }
        }
        for (int64_t j = 0; j < nn; ++j) q_insert(q, s->nb_id[j], s->nb_d[j]);
        cmps += (uint32_t)nn;
        hops += (uint32_t)nbeam;
    }
    if (rec_len) *rec_len = nrec;
    st->cmps = cmps;
    st->hops = hops;
}

static int is_start(const dko_index *ix, uint32_t id) {
    for (int64_t i = 0; i < ix->nstart; ++i)
        if (ix->starts[i] == id) return 1;
    return 0;
}

/* Post-processing: FilterStartPoints -> CopyIds / RemoveDeletedIdsAndCopy
 * (diskann/src/graph/glue.rs:730-821) or FilterStartPoints -> Rerank
 * (full_precision.rs:357-404). Rerank's sort_unstable has unspecified tie order; this
 * restatement uses a stable sort (ties keep queue order). */
static void post_process(const dko_index *ix, int mode, const float *query, scratch *s,
                         int64_t k, uint32_t *out_ids, float *out_d, dko_stats *st) {
    pqueue *q = &s->q;
    int64_t lim = q->search_l < q->size ? q->search_l : q->size;
    if (mode != DKO_MODE_HYBRID) {
        int64_t c = 0;
        for (int64_t i = 0; i < lim && c < k; ++i) {
            if (is_start(ix, q->id[i])) continue;
            out_ids[c] = q->id[i];
            out_d[c] = q->d[i];
            c++;
        }
        st->count = (uint32_t)c;
        return;
    }
    uint32_t *ids = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)(lim + 1));
    float *ds = (float *)malloc(sizeof(float) * (size_t)(lim + 1));
    int64_t n = 0;
    for (int64_t i = 0; i < lim; ++i) {
        if (is_start(ix, q->id[i])) continue;
        ids[n] = q->id[i];
        ds[n] = dko_distance(ix->metric, query, ix->vectors + (int64_t)q->id[i] * ix->vstride,
                             ix->dim);
        n++;
    }
    /* stable insertion sort by distance (partial_cmp, NaN compares Equal) */
    for (int64_t i = 1; i < n; ++i) {
        uint32_t id = ids[i];
        float d = ds[i];
        int64_t j = i - 1;
        while (j >= 0 && ds[j] > d) {
            ids[j + 1] = ids[j];
            ds[j + 1] = ds[j];
            j--;
        }
        ids[j + 1] = id;
        ds[j + 1] = d;
    }
    int64_t c = n < k ? n : k;
    for (int64_t i = 0; i < c; ++i) {
        out_ids[i] = ids[i];
        out_d[i] = ds[i];
    }
    st->count = (uint32_t)c;
    free(ids);
    free(ds);
}

/* Disk-index search: DiskANNIndex::search_internal (diskann/src/graph/index.rs:1938-2005)
 * with the DiskAccessor (diskann-disk/src/search/provider/disk_provider.rs:666-721) and
 * RerankAndFilter (:323-388):
 *   - the starting point is the graph header's medoid, here ix->starts[0] (a data row),
 *     at its PQ distance;
 *   - PQ distances: dko_pq_lut_transposed, then pq_dist_lookup (fixed_chunk_pq_table.rs:
 *     469-598: 0 + LUT[m][code_m] in chunk order, as dko_pq_adc);
 *   - expand_beam loads the first (io_limit - io_count) beam ids (ensure_loaded counts
 *     them) and expands only those; the loop ends when no unvisited node remains or
 *     terminate_early (io_count > io_limit);
 *   - RerankAndFilter (AcceptAll): every queue entry (the first min(search_l, size), the
 *     start point included) gets its full-precision distance (the cached one of a loaded
 *     node is the same value), sorted by distance; sort_unstable leaves ties unspecified
 *     and this restatement keeps queue order (stable sort); the first k are returned. */
int dko_disk_search(const dko_index *ix, const float *query, int64_t k, int64_t L, int64_t bw,
                    uint64_t io_limit, uint32_t *out_ids, float *out_d, dko_stats *st) {
    int rc = dko_knn_validate(k, L, bw);
    if (rc) return rc;
    if (ix->M <= 0 || ix->nstart != 1) return DKO_ERR_ARG;
    scratch s;
    scratch_init(&s, ix, L, bw);
    pqueue *q = &s.q;
    q->size = 0;
    q->cursor = 0;
    set_clear(&s.vis);
    const int lut_metric = ix->metric == DKO_INNER_PRODUCT ? DKO_INNER_PRODUCT : DKO_L2;
    dko_pq_lut_transposed(lut_metric, ix->pivots, ix->K, ix->dim, ix->offsets, ix->M, query,
                          s.lut);
    uint32_t cmps = 0, hops = 0;
    uint64_t ios = 0;
    const uint32_t medoid = ix->starts[0];
    set_insert(&s.vis, medoid);
    q_insert(q, medoid, dko_pq_adc(ix->codes + (int64_t)medoid * ix->M, s.lut, ix->M, ix->K));
    cmps += 1;
    while (q_has_notvisited(q) && !(ios > io_limit)) {
        int64_t nbeam = 0;
        while (nbeam < bw) {
            int64_t c = q_closest_notvisited(q);
            if (c < 0) break;
            s.beam[nbeam++] = q->id[c];
        }
        const uint64_t left = io_limit - ios;
        const int64_t nload = (uint64_t)nbeam < left ? nbeam : (int64_t)left;
        ios += (uint64_t)nload;
        int64_t nn = 0;
        for (int64_t b = 0; b < nload; ++b) {
            const uint32_t node = s.beam[b];
            const uint32_t *adj = ix->adj + (int64_t)node * ix->max_degree;
            for (uint32_t j = 0; j < ix->deg[node]; ++j) {
                const uint32_t nbr = adj[j];
                if (!set_insert(&s.vis, nbr)) continue;
                s.nb_id[nn] = nbr;
                s.nb_d[nn] = dko_pq_adc(ix->codes + (int64_t)nbr * ix->M, s.lut, ix->M, ix->K);
                nn++;
            }
        }
        for (int64_t j = 0; j < nn; ++j) q_insert(q, s.nb_id[j], s.nb_d[j]);
        cmps += (uint32_t)nn;
        hops += (uint32_t)nbeam;
    }
    st->cmps = cmps;
    st->hops = hops;
    /* RerankAndFilter */
    const int64_t lim = q->search_l < q->size ? q->search_l : q->size;
    uint32_t *ids = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)(lim + 1));
    float *ds = (float *)malloc(sizeof(float) * (size_t)(lim + 1));
    for (int64_t i = 0; i < lim; ++i) {
        ids[i] = q->id[i];
        ds[i] = dko_distance(ix->metric, query, ix->vectors + (int64_t)q->id[i] * ix->vstride,
                             ix->dim);
    }
    for (int64_t i = 1; i < lim; ++i) {
        uint32_t id = ids[i];
        float d = ds[i];
        int64_t j = i - 1;
        while (j >= 0 && ds[j] > d) {
            ids[j + 1] = ids[j];
            ds[j + 1] = ds[j];
            j--;
        }
        ids[j + 1] = id;
        ds[j + 1] = d;
    }
    const int64_t c = lim < k ? lim : k;
    for (int64_t i = 0; i < c; ++i) {
        out_ids[i] = ids[i];
        out_d[i] = ds[i];
    }
    st->count = (uint32_t)c;
    free(ids);
    free(ds);
    scratch_free(&s);
    return 0;
}

int dko_search(const dko_index *ix, int mode, const float *query, int64_t k, int64_t L,
               int64_t bw, uint32_t *out_ids, float *out_d, dko_stats *st) {
    int rc = dko_knn_validate(k, L, bw);
    if (rc) return rc;
    if (mode != DKO_MODE_FULL && ix->M <= 0 && !ix->sq && !ix->sph) return DKO_ERR_ARG;
    scratch s;
    scratch_init(&s, ix, L, bw);
    search_core(ix, mode != DKO_MODE_FULL, query, bw, &s, NULL, NULL, NULL, 0, st);
    post_process(ix, mode, query, &s, k, out_ids, out_d, st);
    scratch_free(&s);
    return 0;
}

int dko_search_full(const dko_index *ix, int mode, const float *query, int64_t L,
                    int64_t bw, uint32_t *q_ids, float *q_d, uint32_t *q_len,
                    uint32_t *rec_ids, float *rec_d, uint32_t *rec_len, int64_t rec_cap,
                    dko_stats *st) {
    if (L <= 0 || bw <= 0) return DKO_ERR_ARG;
    scratch s;
    scratch_init(&s, ix, L, bw);
    search_core(ix, mode != DKO_MODE_FULL, query, bw, &s, rec_ids, rec_d, rec_len, rec_cap,
                st);
    int64_t lim = s.q.search_l < s.q.size ? s.q.search_l : s.q.size;
    for (int64_t i = 0; i < lim; ++i) {
        q_ids[i] = s.q.id[i];
        q_d[i] = s.q.d[i];
    }
    *q_len = (uint32_t)lim;
    st->count = (uint32_t)lim;
    scratch_free(&s);
    return 0;
}

/* ------------------------------------------------------------------------------------ */
/* Threaded batch search (CPU baseline harness)                                         */
/* ------------------------------------------------------------------------------------ */

typedef struct {
    const dko_index *ix;
    int mode;
    const float *queries;
    int64_t qstride, k, L, bw, lo, hi;
    uint32_t *out_ids, *out_cmps, *out_hops;
    float *out_d;
    double *lat_us;
    const uint32_t *ks, *Ls; /* per-query k and L (mixed batches), or NULL */
    int64_t kstride;         /* output row stride of a mixed batch */
} batch_arg;

static double now_s(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

static void *batch_worker(void *p) {
    batch_arg *a = (batch_arg *)p;
    scratch s;
    if (!a->Ls) scratch_init(&s, a->ix, a->L, a->bw);
    for (int64_t i = a->lo; i < a->hi; ++i) {
        double t0 = now_s();
        dko_stats st;
        const float *query = a->queries + i * a->qstride;
        /* a mixed batch sizes each query's scratch by its own L, as the reference's
         * per-query scratch does (diskann/src/graph/search/scratch.rs:186-217) */
        if (a->Ls) scratch_init(&s, a->ix, a->Ls[i], a->bw);
        const int64_t k = a->ks ? (int64_t)a->ks[i] : a->k;
        const int64_t stride = a->ks ? a->kstride : a->k;
        search_core(a->ix, a->mode != DKO_MODE_FULL, query, a->bw, &s, NULL, NULL, NULL, 0,
                    &st);
        post_process(a->ix, a->mode, query, &s, k, a->out_ids + i * stride,
                     a->out_d + i * stride, &st);
        if (a->Ls) scratch_free(&s);
        for (int64_t j = st.count; j < stride; ++j) {
            a->out_ids[i * stride + j] = 0xFFFFFFFFu;
            a->out_d[i * stride + j] = INFINITY;
        }
        if (a->out_cmps) a->out_cmps[i] = st.cmps;
        if (a->out_hops) a->out_hops[i] = st.hops;
        if (a->lat_us) a->lat_us[i] = (now_s() - t0) * 1e6;
    }
    if (!a->Ls) scratch_free(&s);
    return NULL;
}

static double run_batch(const dko_index *ix, int mode, const float *queries, int64_t nq,
                        int64_t qstride, int64_t k, int64_t L, const uint32_t *ks,
                        const uint32_t *Ls, int64_t kstride, int64_t bw, int nthreads,
                        uint32_t *out_ids, float *out_d, uint32_t *out_cmps,
                        uint32_t *out_hops, double *lat_us) {
    if (nthreads < 1) nthreads = 1;
    if (nthreads > nq) nthreads = (int)(nq > 0 ? nq : 1);
    pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * (size_t)nthreads);
    batch_arg *args = (batch_arg *)malloc(sizeof(batch_arg) * (size_t)nthreads);
    double t0 = now_s();
    for (int t = 0; t < nthreads; ++t) {
        /* static contiguous partition: the first nq % T ranges get one extra query */
        int64_t base = nq / nthreads, extra = nq % nthreads;
        int64_t lo = t * base + (t < extra ? t : extra);
        int64_t hi = lo + base + (t < extra ? 1 : 0);
        batch_arg a = {ix, mode, queries, qstride, k, L, bw, lo, hi,
                       out_ids, out_cmps, out_hops, out_d, lat_us, ks, Ls, kstride};
        args[t] = a;
        pthread_create(&th[t], NULL, batch_worker, &args[t]);
    }
    for (int t = 0; t < nthreads; ++t) pthread_join(th[t], NULL);
    double wall = now_s() - t0;
    free(th);
    free(args);
    return wall;
}

double dko_search_batch(const dko_index *ix, int mode, const float *queries, int64_t nq,
                        int64_t qstride, int64_t k, int64_t L, int64_t bw, int nthreads,
                        uint32_t *out_ids, float *out_d, uint32_t *out_cmps,
                        uint32_t *out_hops, double *lat_us) {
    if (dko_knn_validate(k, L, bw)) return -1.0;
    return run_batch(ix, mode, queries, nq, qstride, k, L, NULL, NULL, k, bw, nthreads,
                     out_ids, out_d, out_cmps, out_hops, lat_us);
}

double dko_search_batch_mixed(const dko_index *ix, int mode, const float *queries, int64_t nq,
                              int64_t qstride, const uint32_t *ks, const uint32_t *Ls,
                              int64_t kstride, int64_t bw, int nthreads, uint32_t *out_ids,
                              float *out_d, uint32_t *out_cmps, uint32_t *out_hops,
                              double *lat_us) {
    for (int64_t i = 0; i < nq; ++i)
        if (dko_knn_validate(ks[i], Ls[i], bw) || (int64_t)ks[i] > kstride) return -1.0;
    return run_batch(ix, mode, queries, nq, qstride, 0, 0, ks, Ls, kstride, bw, nthreads,
                     out_ids, out_d, out_cmps, out_hops, lat_us);
}

/* ------------------------------------------------------------------------------------ */
/* RobustPrune                                                                          */
/* ------------------------------------------------------------------------------------ */

/* PruneKind::update_occlude_factor, diskann/src/graph/config/mod.rs:80-103. */
float dko_update_occlude_factor(int metric, float d_ik, float d_jk, float cur, float alpha) {
    if (metric == DKO_INNER_PRODUCT) {
        if (d_jk < alpha * d_ik) return alpha + 0.01f;
        return cur;
    }
    if (d_jk == 0.0f) return FLT_MAX;
    float r = d_ik / d_jk;
    return fmaxf(cur, r);
}

/* occlude_list, diskann/src/graph/index.rs:2570-2782. The pool must already be sorted
 * and truncated to max_occlusion. */
/* Hybrid prune distance (HybridComputer, diskann-providers/src/model/graph/provider/
 * async_/distances.rs:118-165): both full precision -> the metric; one full precision ->
 * the direct PQ distance of that vector against the other's code; both codes -> code vs
 * code (left = a). */
static float hyb_dist(const dko_build_cfg *cfg, const float *vectors, int64_t vstride,
                      int64_t dim, uint32_t a, int fa, uint32_t b, int fb) {
    const float *va = vectors + (int64_t)a * vstride, *vb = vectors + (int64_t)b * vstride;
    if (fa && fb) return dko_distance(cfg->metric, va, vb, dim);
    const uint8_t *ca = cfg->codes + (int64_t)a * cfg->pq_M;
    const uint8_t *cb = cfg->codes + (int64_t)b * cfg->pq_M;
    if (fa) return dko_pq_direct(cfg->metric, cfg->pivots, dim, cfg->pq_offsets, cfg->pq_M, va, cb);
    if (fb) return dko_pq_direct(cfg->metric, cfg->pivots, dim, cfg->pq_offsets, cfg->pq_M, vb, ca);
    return dko_pq_qq(cfg->metric, cfg->pivots, dim, cfg->pq_offsets, cfg->pq_M, ca, cb);
}

/* Distance between rows a and b in a prune: the metric on the vectors, or, for the SQ
 * build, the compensated distance between their stored codes (scalar.rs:370-446 with the
 * SQStore distance computer, scalar/vectors.rs:182-470). */
static float pair_dist(const dko_build_cfg *cfg, const float *vectors, int64_t vstride,
                       int64_t dim, uint32_t a, uint32_t b) {
    if (cfg->sph) {
        const int64_t cb = dko_sph_data_bytes(cfg->sph);
        return dko_sph_dist_data(cfg->sph, cfg->sph_codes + (int64_t)a * cb,
                                 cfg->sph_codes + (int64_t)b * cb);
    }
    if (cfg->sq) {
        const dko_sq *q = cfg->sq;
        return dko_sq_distance(cfg->metric, q->scale, q->shift_square_norm,
                               q->codes + (int64_t)a * q->dim, q->comp[a],
                               q->codes + (int64_t)b * q->dim, q->comp[b], q->dim);
    }
    return dko_distance(cfg->metric, vectors + (int64_t)a * vstride,
                        vectors + (int64_t)b * vstride, dim);
}

static int in_set(const uint32_t *set, int64_t n, uint32_t id) {
    for (int64_t i = 0; i < n; ++i)
        if (set[i] == id) return 1;
    return 0;
}

static int64_t occlude_h(const dko_build_cfg *cfg, const float *vectors, int64_t vstride,
                         int64_t dim, const uint32_t *pool_ids, const float *pool_d,
                         int64_t pool_len, uint32_t self_id, const uint8_t *isfp,
                         uint32_t *out);

int64_t dko_occlude_list(const dko_build_cfg *cfg, const float *vectors, int64_t vstride,
                         int64_t dim, const uint32_t *pool_ids, const float *pool_d,
                         int64_t pool_len, uint32_t self_id, uint32_t *out) {
    return occlude_h(cfg, vectors, vstride, dim, pool_ids, pool_d, pool_len, self_id, NULL, out);
}

/* isfp (NULL: all full precision) marks the pool entries the Hybrid prune keeps in full
 * precision. */
static int64_t occlude_h(const dko_build_cfg *cfg, const float *vectors, int64_t vstride,
                         int64_t dim, const uint32_t *pool_ids, const float *pool_d,
                         int64_t pool_len, uint32_t self_id, const uint8_t *isfp,
                         uint32_t *out) {
    if (pool_len == 0) return 0;
    float *occ = (float *)calloc((size_t)pool_len, sizeof(float));
    uint32_t *last = (uint32_t *)calloc((size_t)pool_len, sizeof(uint32_t));
    uint32_t *nbr = (uint32_t *)calloc((size_t)pool_len, sizeof(uint32_t));
    uint8_t *valid = (uint8_t *)malloc((size_t)pool_len);
    for (int64_t i = 0; i < pool_len; ++i) valid[i] = pool_ids[i] != self_id;
    float alpha = cfg->alpha;
    int64_t degree = cfg->R;
    float cur_alpha = 1.0f;
    float inc = alpha < 1.2f ? alpha : 1.2f;
    int64_t found = 0;
    while (found < degree) {
        for (int64_t i = 0; i < pool_len; ++i) {
            if (found >= degree) break;
            float of = occ[i];
            uint32_t lc = last[i];
            if (of > cur_alpha) continue;
            if (!valid[i]) {
                occ[i] = FLT_MAX;
                continue;
            }
            while ((int64_t)lc != found) {
                uint32_t rp = nbr[lc];
                lc += 1;
                if ((int64_t)rp >= i) {
                    last[i] = lc;
                    continue;
                }
                float d;
                if (!valid[rp])
                    d = FLT_MAX;
                else if (isfp)
                    d = hyb_dist(cfg, vectors, vstride, dim, pool_ids[i], isfp[i], pool_ids[rp],
                                 isfp[rp]);
                else
                    d = pair_dist(cfg, vectors, vstride, dim, pool_ids[i], pool_ids[rp]);
                of = dko_update_occlude_factor(cfg->metric, pool_d[i], d, of, cur_alpha);
                if (of > cur_alpha) break;
            }
            last[i] = lc;
            if (of > cur_alpha) {
                occ[i] = of;
                continue;
            }
            occ[i] = FLT_MAX;
            nbr[found] = (uint32_t)i;
            found++;
        }
        if (cur_alpha == alpha) break;
        cur_alpha = fminf(cur_alpha * inc, alpha);
    }
    for (int64_t n = 0; n < found; ++n) out[n] = pool_ids[nbr[n]];
    int64_t count = found;
    if (cfg->saturate == 2 || (cfg->saturate && alpha > 1.0f)) {
        for (int64_t i = 0; i < pool_len && count < degree; ++i) {
            if (!valid[i]) continue;
            int dup = 0;
            for (int64_t j = 0; j < count; ++j)
                if (out[j] == pool_ids[i]) dup = 1;
            if (!dup) out[count++] = pool_ids[i];
        }
    }
    free(occ);
    free(last);
    free(nbr);
    free(valid);
    return count;
}

/* Stable sort of (id, d) by distance (the reference's SortedNeighbors uses
 * select_nth_unstable + sort_unstable, whose tie order is unspecified,
 * diskann/src/graph/internal/sorted_neighbors.rs:26-44). */
static void sort_pool(uint32_t *ids, float *ds, int64_t n) {
    if (n < 2) return;
    /* bottom-up merge sort */
    uint32_t *ti = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)n);
    float *td = (float *)malloc(sizeof(float) * (size_t)n);
    for (int64_t w = 1; w < n; w *= 2) {
        for (int64_t lo = 0; lo < n; lo += 2 * w) {
            int64_t mid = lo + w < n ? lo + w : n, hi = lo + 2 * w < n ? lo + 2 * w : n;
            int64_t a = lo, b = mid, o = lo;
            while (a < mid && b < hi) {
                if (ds[b] < ds[a]) {
                    ti[o] = ids[b];
                    td[o++] = ds[b++];
                } else {
                    ti[o] = ids[a];
                    td[o++] = ds[a++];
                }
            }
            while (a < mid) {
                ti[o] = ids[a];
                td[o++] = ds[a++];
            }
            while (b < hi) {
                ti[o] = ids[b];
                td[o++] = ds[b++];
            }
        }
        memcpy(ids, ti, sizeof(uint32_t) * (size_t)n);
        memcpy(ds, td, sizeof(float) * (size_t)n);
    }
    free(ti);
    free(td);
}

/* ------------------------------------------------------------------------------------ */
/* Build                                                                                */
/* ------------------------------------------------------------------------------------ */


/* robust_prune_list, index.rs:2388-2450: distances from `location` to every other id of
 * `list` (duplicates and self removed), sorted pool, occlude_list. */
static int64_t prune_list_opt(const dko_build_cfg *cfg, const float *vectors, int64_t vstride,
                              int64_t dim, uint32_t location, const uint32_t *list, int64_t n,
                              uint32_t *out, int force_saturate);
static int64_t prune_list(const dko_build_cfg *cfg, const float *vectors, int64_t vstride,
                          int64_t dim, uint32_t location, const uint32_t *list, int64_t n,
                          uint32_t *out) {
    return prune_list_opt(cfg, vectors, vstride, dim, location, list, n, out, 0);
}
static int64_t push_unique(uint32_t *a, int64_t n, uint32_t v) {
    for (int64_t i = 0; i < n; ++i)
        if (a[i] == v) return n;
    a[n] = v;
    return n + 1;
}
static int64_t count_unique_targets(const uint32_t *nbrs, const int64_t *cnt, int64_t B,
                                    int64_t stride) {
    int64_t total = 0;
    for (int64_t p = 0; p < B; ++p) total += cnt[p];
    uint32_t *t = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)(total + 1));
    int64_t n = 0;
    for (int64_t p = 0; p < B; ++p)
        for (int64_t j = 0; j < cnt[p]; ++j) t[n++] = nbrs[p * stride + j];
    /* insertion into a small sorted set is fine at oracle sizes */
    int64_t u = 0;
    for (int64_t i = 0; i < n; ++i) {
        int dup = 0;
        for (int64_t j = 0; j < u; ++j)
            if (t[j] == t[i]) {
                dup = 1;
                break;
            }
        if (!dup) t[u++] = t[i];
    }
    free(t);
    return u;
}
static int64_t prune_list_opt(const dko_build_cfg *cfg, const float *vectors, int64_t vstride,
                              int64_t dim, uint32_t location, const uint32_t *list, int64_t n,
                              uint32_t *out, int force_saturate) {
    uint32_t *ids = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)(n + 1));
    float *ds = (float *)malloc(sizeof(float) * (size_t)(n + 1));
    uint8_t *isfp = NULL;
    uint32_t *fpset = NULL;
    int64_t nfp = 0;
    if (cfg->max_fp >= 0) {
        /* fill(chain(once(location), list)): the first max_fp ids are full precision */
        fpset = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)(n + 1));
        if (nfp < cfg->max_fp) fpset[nfp++] = location;
        for (int64_t i = 0; i < n && nfp < cfg->max_fp; ++i) fpset[nfp++] = list[i];
        isfp = (uint8_t *)malloc((size_t)(n + 1));
    }
    const int loc_fp = fpset ? in_set(fpset, nfp, location) : 1;
    int64_t m = 0;
    for (int64_t i = 0; i < n; ++i) {
        if (list[i] == location) continue;
        ids[m] = list[i];
        if (fpset)
            ds[m] = hyb_dist(cfg, vectors, vstride, dim, location, loc_fp, list[i],
                             in_set(fpset, nfp, list[i]));
        else
            ds[m] = pair_dist(cfg, vectors, vstride, dim, location, list[i]);
        m++;
    }
    sort_pool(ids, ds, m);
    if (m > cfg->max_occlusion) m = cfg->max_occlusion;
    if (isfp)
        for (int64_t i = 0; i < m; ++i) isfp[i] = (uint8_t)in_set(fpset, nfp, ids[i]);
    dko_build_cfg c2 = *cfg;
    if (force_saturate) c2.saturate = 2; /* forced: saturate regardless of alpha */
    int64_t c = occlude_h(&c2, vectors, vstride, dim, ids, ds, m, location, isfp, out);
    free(ids);
    free(ds);
    free(isfp);
    free(fpset);
    return c;
}

int64_t dko_prune_list(const dko_build_cfg *cfg, const float *vectors, int64_t vstride,
                       int64_t dim, uint32_t location, const uint32_t *list, int64_t n,
                       int force_saturate, uint32_t *out) {
    if (n == 0) return 0; /* robust_prune_list's early exit */
    return prune_list_opt(cfg, vectors, vstride, dim, location, list, n, out, force_saturate);
}

/* add_edge_and_prune, index.rs:2269-2346: append unique targets; prune when the list
 * exceeds the max degree with slack. */
static void add_edge_and_prune(const dko_build_cfg *cfg, const float *vectors, int64_t vstride,
                               int64_t dim, dko_graph *g, uint32_t source,
                               const uint32_t *targets, int64_t nt) {
    uint32_t *buf = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)(g->max_degree + nt + 1));
    int64_t n = g->deg[source];
    memcpy(buf, g->adj + (int64_t)source * g->max_degree, sizeof(uint32_t) * (size_t)n);
    int64_t added = 0;
    for (int64_t t = 0; t < nt; ++t) {
        int dup = 0;
        for (int64_t j = 0; j < n; ++j)
            if (buf[j] == targets[t]) {
                dup = 1;
                break;
            }
        if (!dup) {
            buf[n++] = targets[t];
            added++;
        }
    }
    if (added == 0) {
        free(buf);
        return;
    }
    uint32_t *dst = g->adj + (int64_t)source * g->max_degree;
    if (n <= cfg->R_max) {
        memcpy(dst, buf, sizeof(uint32_t) * (size_t)n);
        g->deg[source] = (uint32_t)n;
    } else {
        uint32_t *out = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)(cfg->R + 1));
        int64_t c = prune_list(cfg, vectors, vstride, dim, source, buf, n, out);
        memcpy(dst, out, sizeof(uint32_t) * (size_t)c);
        g->deg[source] = (uint32_t)c;
        free(out);
    }
    free(buf);
}

/* search + prune for one point (insert index.rs:226-341, search_and_prune :349-434,
 * robust_prune_with :2481-2537). extras: intra-batch candidates (distances appended to
 * the visited record). Returns the number of new neighbors. */
static int64_t search_and_prune(const dko_build_cfg *cfg, const float *vectors, int64_t vstride,
                                int64_t dim, const dko_graph *g, const uint32_t *starts,
                                int64_t nstart, uint32_t id, const uint32_t *extras,
                                int64_t nextras, uint32_t *out) {
    dko_index ix;
    memset(&ix, 0, sizeof(ix));
    ix.metric = cfg->metric;
    ix.dim = dim;
    ix.vstride = vstride;
    ix.vectors = vectors;
    ix.max_degree = g->max_degree;
    ix.adj = g->adj;
    ix.deg = g->deg;
    ix.nstart = nstart;
    ix.starts = starts;
    ix.sq = cfg->sq;  /* SQ build: the insert search is the Quantized search */
    ix.sph = cfg->sph; /* spherical build: SameAsData queries, no rescale */
    ix.sph_codes = cfg->sph_codes;
    ix.sph_layout = 0;
    ix.sph_norescale = 1;
    const float *v = vectors + (int64_t)id * vstride;
    int64_t L = cfg->L_build;
    scratch s;
    scratch_init(&s, &ix, L, 1);
    int64_t cap = 16384;
    uint32_t *rid = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)(cap + nextras));
    float *rd = (float *)malloc(sizeof(float) * (size_t)(cap + nextras));
    uint32_t rlen = 0;
    dko_stats st;
    search_core(&ix, cfg->sq != NULL || cfg->sph != NULL, v, 1, &s, rid, rd, &rlen, cap, &st);
    if (rlen > cap) rlen = (uint32_t)cap;
    /* Hybrid prune: fill(chain(once(self), chain(extras, record)).take(max_occlusion));
     * the first max_fp of those ids are full precision */
    uint32_t *fpset = NULL;
    int64_t nfp = 0;
    if (cfg->max_fp >= 0) {
        int64_t lim = cfg->max_fp < cfg->max_occlusion ? cfg->max_fp : cfg->max_occlusion;
        fpset = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)(lim + 1));
        if (nfp < lim) fpset[nfp++] = id;
        for (int64_t e = 0; e < nextras && nfp < lim; ++e) fpset[nfp++] = extras[e];
        for (uint32_t r = 0; r < rlen && nfp < lim; ++r) fpset[nfp++] = rid[r];
    }
    const int self_fp = fpset ? in_set(fpset, nfp, id) : 1;
    /* the in-memory providers' prune working set is a pass-through view: every id is
     * available, so all extras join the record before sorting */
    for (int64_t e = 0; e < nextras; ++e) {
        rid[rlen] = extras[e];
        if (fpset)
            rd[rlen] = hyb_dist(cfg, vectors, vstride, dim, id, self_fp, extras[e],
                                in_set(fpset, nfp, extras[e]));
        else
            rd[rlen] = pair_dist(cfg, vectors, vstride, dim, id, extras[e]);
        rlen++;
    }
    sort_pool(rid, rd, rlen);
    int64_t plen = rlen;
    if (plen > cfg->max_occlusion) plen = cfg->max_occlusion;
    uint8_t *isfp = NULL;
    if (fpset) {
        isfp = (uint8_t *)malloc((size_t)(plen + 1));
        for (int64_t i = 0; i < plen; ++i) isfp[i] = (uint8_t)in_set(fpset, nfp, rid[i]);
    }
    int64_t c = occlude_h(cfg, vectors, vstride, dim, rid, rd, plen, id, isfp, out);
    free(isfp);
    free(fpset);
    free(rid);
    free(rd);
    scratch_free(&s);
    return c;
}

int dko_build(const dko_build_cfg *cfg, const float *vectors, int64_t vstride, int64_t dim,
              dko_graph *g, const uint32_t *starts, int64_t nstart, const uint32_t *ids,
              int64_t nids, int64_t batch, int64_t intra_batch) {
    if (batch < 1) return DKO_ERR_ARG;
    uint32_t *nbrs = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)(batch * (cfg->R + 1)));
    int64_t *cnt = (int64_t *)malloc(sizeof(int64_t) * (size_t)batch);
    uint32_t *extras = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)batch);
    for (int64_t b0 = 0; b0 < nids; b0 += batch) {
        int64_t B = nids - b0 < batch ? nids - b0 : batch;
        const uint32_t *bid = ids + b0;
        if (batch == 1) {
            /* single insert: search, prune, set, then back edges one at a time */
            uint32_t *out = nbrs;
            int64_t c = search_and_prune(cfg, vectors, vstride, dim, g, starts, nstart, bid[0],
                                         NULL, 0, out);
            memcpy(g->adj + (int64_t)bid[0] * g->max_degree, out, sizeof(uint32_t) * (size_t)c);
            g->deg[bid[0]] = (uint32_t)c;
            int64_t nb = c < cfg->max_backedges ? c : cfg->max_backedges;
            for (int64_t j = 0; j < nb; ++j)
                add_edge_and_prune(cfg, vectors, vstride, dim, g, out[j], &bid[0], 1);
            continue;
        }
        /* candidate generation on the pre-batch graph */
        int64_t nc = intra_batch < 0 ? B : intra_batch;
        for (int64_t p = 0; p < B; ++p) {
            /* async_tools::around(ids, p, nc): a circular window of min(nc, B-1) ids
             * starting ceil(len/2) before p, skipping p
             * (diskann/src/utils/async_tools.rs:20-130) */
            int64_t ne = 0;
            int64_t want = nc < B - 1 ? nc : B - 1;
            if (want > 0) {
                int64_t half = (want + 1) / 2;
                int64_t pos = p >= half ? p - half : B - (half - p);
                for (int64_t r = 0; r < want; ++r) {
                    int64_t q = pos;
                    pos = pos + 1 == B ? 0 : pos + 1;
                    if (q == p) {
                        q = pos;
                        pos = pos + 1 == B ? 0 : pos + 1;
                    }
                    extras[ne++] = bid[q];
                }
            }
            cnt[p] = search_and_prune(cfg, vectors, vstride, dim, g, starts, nstart, bid[p],
                                      extras, ne, nbrs + p * (cfg->R + 1));
        }
        /* bootstrap (index.rs:595-760, trigger :948-960): when the batch saw only part of
         * itself and the back edges reach few distinct nodes, every item re-prunes its
         * edges together with the other batch items, saturated */
        {
            int64_t resolved = nc < B ? nc : B;
            if (resolved < 1) resolved = 1;
            int64_t uniq = count_unique_targets(nbrs, cnt, B, cfg->R + 1);
            if (resolved < B && (uniq + 7) / 8 <= B) {
                uint32_t *cand = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)(cfg->R + B + 1));
                uint32_t *nb2 = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)(B * (cfg->R + 1)));
                int64_t *cnt2 = (int64_t *)malloc(sizeof(int64_t) * (size_t)B);
                dko_build_cfg sat = *cfg;
                sat.saturate = 1;
                for (int64_t p = 0; p < B; ++p) {
                    int64_t nc2 = 0;
                    for (int64_t j = 0; j < cnt[p]; ++j) nc2 = push_unique(cand, nc2, nbrs[p * (cfg->R + 1) + j]);
                    for (int64_t q = 0; q < B; ++q)
                        if (q != p) nc2 = push_unique(cand, nc2, bid[q]);
                    cnt2[p] = nc2 ? prune_list_opt(&sat, vectors, vstride, dim, bid[p], cand, nc2,
                                                   nb2 + p * (cfg->R + 1), 1)
                                  : 0;
                }
                memcpy(nbrs, nb2, sizeof(uint32_t) * (size_t)(B * (cfg->R + 1)));
                memcpy(cnt, cnt2, sizeof(int64_t) * (size_t)B);
                free(cand);
                free(nb2);
                free(cnt2);
            }
        }
        for (int64_t p = 0; p < B; ++p) {
            memcpy(g->adj + (int64_t)bid[p] * g->max_degree, nbrs + p * (cfg->R + 1),
                   sizeof(uint32_t) * (size_t)cnt[p]);
            g->deg[bid[p]] = (uint32_t)cnt[p];
        }
        /* aggregate back edges: target -> sorted list of batch sources */
        int64_t total = 0;
        for (int64_t p = 0; p < B; ++p) total += cnt[p];
        uint64_t *pairs = (uint64_t *)malloc(sizeof(uint64_t) * (size_t)(total + 1));
        int64_t np = 0;
        for (int64_t p = 0; p < B; ++p)
            for (int64_t j = 0; j < cnt[p]; ++j)
                pairs[np++] = ((uint64_t)nbrs[p * (cfg->R + 1) + j] << 32) | bid[p];
        /* sort pairs (target, source) */
        for (int64_t w = 1; w < np; w *= 2) {
            uint64_t *tmp = (uint64_t *)malloc(sizeof(uint64_t) * (size_t)np);
            for (int64_t lo = 0; lo < np; lo += 2 * w) {
                int64_t mid = lo + w < np ? lo + w : np, hi = lo + 2 * w < np ? lo + 2 * w : np;
                int64_t a = lo, b = mid, o = lo;
                while (a < mid && b < hi) tmp[o++] = pairs[b] < pairs[a] ? pairs[b++] : pairs[a++];
                while (a < mid) tmp[o++] = pairs[a++];
                while (b < hi) tmp[o++] = pairs[b++];
            }
            memcpy(pairs, tmp, sizeof(uint64_t) * (size_t)np);
            free(tmp);
        }
        uint32_t *tg = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)(np + 1));
        for (int64_t i = 0; i < np;) {
            uint32_t target = (uint32_t)(pairs[i] >> 32);
            int64_t j = i, nt = 0;
            while (j < np && (uint32_t)(pairs[j] >> 32) == target) tg[nt++] = (uint32_t)pairs[j++];
            add_edge_and_prune(cfg, vectors, vstride, dim, g, target, tg, nt);
            i = j;
        }
        free(tg);
        free(pairs);
    }
    free(nbrs);
    free(cnt);
    free(extras);
    return 0;
}

/* compute_medoid, diskann-utils/src/sampling/medoid.rs:15-47: f64 column means cast to
 * f32, then the row nearest by SquaredL2 (strict <, first wins). */
int64_t dko_medoid(const float *x, int64_t n, int64_t stride, int64_t dim, float *mean_out) {
    double *sum = (double *)calloc((size_t)dim, sizeof(double));
    for (int64_t i = 0; i < n; ++i)
        for (int64_t j = 0; j < dim; ++j) sum[j] += (double)x[i * stride + j];
    float *mean = (float *)malloc(sizeof(float) * (size_t)dim);
    for (int64_t j = 0; j < dim; ++j) mean[j] = (float)(sum[j] / (double)n);
    int64_t best = -1;
    float bd = FLT_MAX;
    for (int64_t i = 0; i < n; ++i) {
        float d = simd_4x1(0, mean, x + i * stride, dim);
        if (d < bd) {
            bd = d;
            best = i;
        }
    }
    if (mean_out) memcpy(mean_out, mean, sizeof(float) * (size_t)dim);
    free(sum);
    free(mean);
    return best;
}

/* ------------------------------------------------------------------------------------ */
/* k-means (Lloyd)                                                                       */
/* ------------------------------------------------------------------------------------ */

/* distances_in_place, diskann-quantization/src/algorithms/kmeans/lloyds.rs:28-328:
 * d = ||c||^2 - <c,x> - <c,x> + ||x||^2 with <c,x> a sequential FMA over dims (c
 * broadcast times data), argmin strict <. Returns the residual (sum of minimum
 * distances, accumulated per lane of 8 then sum_tree; here the lane is row % 8 within
 * blocks of 16 rows split 8+8, so row r goes to lane r % 8). */
float dko_lloyd_assign(const float *x, int64_t n, int64_t dim, const float *c, int64_t K,
                       uint32_t *nearest) {
    float *cn = (float *)malloc(sizeof(float) * (size_t)K);
    for (int64_t k = 0; k < K; ++k) cn[k] = dko_square_norm(c + k * dim, dim);
    float res[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int64_t i = 0; i < n; ++i) {
        const float *xi = x + i * dim;
        float xn = dko_square_norm(xi, dim);
        float best = INFINITY;
        uint32_t arg = 0;
        for (int64_t k = 0; k < K; ++k) {
            float dot = 0.0f;
            for (int64_t j = 0; j < dim; ++j) dot = fmaf(c[k * dim + j], xi[j], dot);
            float d = cn[k] - dot - dot + xn;
            if (d < best) {
                best = d;
                arg = (uint32_t)k;
            }
        }
        nearest[i] = arg;
        res[i % 8] += best;
    }
    free(cn);
    return sum_tree8(res);
}

/* update_centroids, lloyds.rs:345-366: f64 sums in row order divided by max(count,1),
 * cast to f32; empty clusters become the zero vector. */
void dko_lloyd_update(const float *x, int64_t n, int64_t dim, const uint32_t *nearest,
                      int64_t K, float *c) {
    double *s = (double *)calloc((size_t)(K * dim), sizeof(double));
    int64_t *cnt = (int64_t *)calloc((size_t)K, sizeof(int64_t));
    for (int64_t i = 0; i < n; ++i) {
        uint32_t k = nearest[i];
        cnt[k]++;
        for (int64_t j = 0; j < dim; ++j) s[k * dim + j] += (double)x[i * dim + j];
    }
    for (int64_t k = 0; k < K; ++k) {
        double den = (double)(cnt[k] > 1 ? cnt[k] : 1);
        for (int64_t j = 0; j < dim; ++j) c[k * dim + j] = (float)(s[k * dim + j] / den);
    }
    free(s);
    free(cnt);
}

/* rand 0.9 UniformFloat::<f64>::new(0.0, s) (third-party, rand 0.9.x in Cargo.lock): an
 * empty range (!(0 < s), NaN included) or a non-finite scale fail; otherwise scale = s,
 * lowered one ulp at a time while scale * (1 - 2^-52) + 0 >= s; a draw maps to
 * value0_1 * scale + 0. Returns 0 ok, 1 empty range, 2 non-finite. */
static int uniform_f64_scale(double s, double *scale) {
    if (!(0.0 < s)) return 1;
    double sc = s - 0.0;
    if (!isfinite(sc)) return 2;
    const double max_rand = 1.0 - 2.220446049250313e-16;
    while (sc * max_rand + 0.0 >= s) sc = nextafter(sc, 0.0);
    *scale = sc;
    return 0;
}

/* update_distances (plusplus.rs:238-318, MicroKernel for BlockTransposed<f32, 16> :84-231):
 * dot = sequential fma(x_d, c_d, dot) over the chunk, distance = (norm_x + norm_c) +
 * dot * -2, min_d = distance if distance < min_d; the f64 total runs over blocks of 16
 * rows, each block adding sum_{j<8} ((f64) min_d[16b + j] + (f64) min_d[16b + 8 + j]) in
 * order j (rows past n count 0). */
static double kpp_update(const float *x, int64_t n, int64_t stride, int64_t ds,
                         const float *norms, const float *c, float cnorm, float *mind) {
    for (int64_t r = 0; r < n; ++r) {
        const float *xr = x + r * stride;
        float dot = 0.0f;
        for (int64_t d = 0; d < ds; ++d) dot = fmaf(xr[d], c[d], dot);
        const float dist = (norms[r] + cnorm) + dot * -2.0f;
        if (dist < mind[r]) mind[r] = dist;
    }
    double s = 0.0;
    for (int64_t b = 0; b < n; b += 16) {
        double bs = 0.0;
        for (int64_t j = 0; j < 8; ++j) {
            const double d0 = b + j < n ? (double)mind[b + j] : 0.0;
            const double d1 = b + 8 + j < n ? (double)mind[b + 8 + j] : 0.0;
            bs += d0 + d1;
        }
        s = s + bs;
    }
    return s;
}

/* kmeans_plusplus_into_inner, plusplus.rs:381-498. */
int dko_kmeanspp(const float *x, int64_t n, int64_t stride, int64_t ds, int64_t K,
                 const double *u, float *centers, int64_t *selected) {
    memset(centers, 0, sizeof(float) * (size_t)(K * ds));
    *selected = 0;
    if (n <= 0) return K == 0 ? 0 : 1;
    if (K == 0) return 0;
    float *norms = (float *)malloc(sizeof(float) * (size_t)n);
    float *mind = (float *)malloc(sizeof(float) * (size_t)n);
    uint8_t *picked = (uint8_t *)calloc((size_t)n, 1);
    for (int64_t r = 0; r < n; ++r) {
        norms[r] = dko_square_norm(x + r * stride, ds);
        mind[r] = INFINITY;
    }
    int64_t i0 = (int64_t)(u[0] * (double)n);
    if (i0 >= n) i0 = n - 1;
    memcpy(centers, x + i0 * stride, sizeof(float) * (size_t)ds);
    picked[i0] = 1;
    float prev_norm = norms[i0];
    int64_t sel = 1;
    int reason = 0;
    const int64_t steps = K < n ? K : n;
    for (int64_t cur = 1; cur < steps; ++cur) {
        const double s = kpp_update(x, n, stride, ds, norms, centers + (cur - 1) * ds, prev_norm,
                                    mind);
        double scale = 0.0;
        const int us = uniform_f64_scale(s, &scale);
        if (us == 2) {
            reason = 3;  /* SawInfinity */
            break;
        }
        if (us == 0) {
            const double threshold = u[cur] * scale + 0.0;
            double rolling = 0.0;
            for (int64_t i = 0; i < n; ++i) {
                rolling += (double)mind[i];
                if (rolling >= threshold && mind[i] > 0.0f && !picked[i]) {
                    memcpy(centers + cur * ds, x + i * stride, sizeof(float) * (size_t)ds);
                    picked[i] = 1;
                    prev_norm = norms[i];
                    sel = cur + 1;
                    break;
                }
            }
        }
        if (sel != cur + 1) {
            reason = 2;  /* InsufficientDiversity */
            break;
        }
    }
    if (reason == 0 && sel != K) reason = 1;  /* DatasetTooSmall */
    *selected = sel;
    free(norms);
    free(mind);
    free(picked);
    return reason;
}

/* Train driver, train.rs:109-231 (per chunk: strided view, norms, k-means++ with
 * recoverable errors suppressed, lloyds_inner) and lloyds.rs:372-426 (centre norms
 * recomputed between iterations, not after the last). Chunks are independent (the
 * reference maps them in parallel with rayon); here up to 16 threads take chunks in turn. */
typedef struct {
    const float *x;
    int64_t n, dim, M, K, iters;
    const int64_t *offsets;
    const double *u;
    float *pivots;
    int64_t next;
    int failed;  /* 1 + the first failing chunk, 0 if none */
    pthread_mutex_t mu;
} pqt_job;

static int pq_train_chunk(pqt_job *j, int64_t m) {
    const int64_t n = j->n, K = j->K, o0 = j->offsets[m], ds = j->offsets[m + 1] - j->offsets[m];
    float *xc = (float *)malloc(sizeof(float) * (size_t)(n * ds + 1));
    float *c = (float *)malloc(sizeof(float) * (size_t)(K * ds + 1));
    uint32_t *nearest = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)(n + 1));
    for (int64_t r = 0; r < n; ++r)
        memcpy(xc + r * ds, j->x + r * j->dim + o0, sizeof(float) * (size_t)ds);
    int64_t sel = 0;
    const int reason = dko_kmeanspp(xc, n, ds, ds, K, j->u + m * K, c, &sel);
    const int bad = reason == 3 || (reason != 0 && sel == 0);
    if (!bad) {
        for (int64_t it = 0; it < j->iters; ++it) {
            (void)dko_lloyd_assign(xc, n, ds, c, K, nearest);
            dko_lloyd_update(xc, n, ds, nearest, K, c);
        }
        for (int64_t k = 0; k < K; ++k)
            memcpy(j->pivots + k * j->dim + o0, c + k * ds, sizeof(float) * (size_t)ds);
    }
    free(xc);
    free(c);
    free(nearest);
    return bad;
}

static void *pq_train_worker(void *p) {
    pqt_job *j = (pqt_job *)p;
    for (;;) {
        pthread_mutex_lock(&j->mu);
        const int64_t m = j->next++;
        pthread_mutex_unlock(&j->mu);
        if (m >= j->M) break;
        if (pq_train_chunk(j, m)) {
            pthread_mutex_lock(&j->mu);
            if (j->failed == 0 || 1 + m < j->failed) j->failed = (int)(1 + m);
            pthread_mutex_unlock(&j->mu);
        }
    }
    return NULL;
}

int dko_pq_train(const float *x, int64_t n, int64_t dim, const int64_t *offsets, int64_t M,
                 int64_t K, int64_t iters, const double *u, float *pivots) {
    pqt_job j = {x, n, dim, M, K, iters, offsets, u, pivots, 0, 0, PTHREAD_MUTEX_INITIALIZER};
    int nt = M < 16 ? (int)M : 16;
    pthread_t th[16];
    for (int t = 0; t < nt; ++t) pthread_create(&th[t], NULL, pq_train_worker, &j);
    for (int t = 0; t < nt; ++t) pthread_join(th[t], NULL);
    return j.failed;
}

/* ------------------------------------------------------------------------------------ */
/* Brute force FP64 top-k                                                                */
/* ------------------------------------------------------------------------------------ */

typedef struct {
    int metric;
    const float *base;
    int64_t n, stride, dim;
    const float *queries;
    int64_t k, lo, hi;
    uint32_t *out_ids;
    double *out_d;
} bf_arg;

static double dist_f64(int metric, const float *a, const float *b, int64_t dim) {
    if (metric == DKO_L2) {
        double s = 0;
        for (int64_t j = 0; j < dim; ++j) {
            double c = (double)a[j] - (double)b[j];
            s += c * c;
        }
        return s;
    }
    double xy = 0, xx = 0, yy = 0;
    for (int64_t j = 0; j < dim; ++j) {
        xy += (double)a[j] * (double)b[j];
        xx += (double)a[j] * (double)a[j];
        yy += (double)b[j] * (double)b[j];
    }
    if (metric == DKO_INNER_PRODUCT) return -xy;
    if (metric == DKO_COSINE_NORMALIZED) return 1.0 - xy;
    if (xx == 0 || yy == 0) return 1.0;
    return 1.0 - xy / (sqrt(xx) * sqrt(yy));
}

static void *bf_worker(void *p) {
    bf_arg *a = (bf_arg *)p;
    double *bd = (double *)malloc(sizeof(double) * (size_t)a->k);
    uint32_t *bi = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)a->k);
    for (int64_t q = a->lo; q < a->hi; ++q) {
        int64_t cnt = 0;
        const float *qv = a->queries + q * a->dim;
        for (int64_t i = 0; i < a->n; ++i) {
            double d = dist_f64(a->metric, qv, a->base + i * a->stride, a->dim);
            if (cnt == a->k && !(d < bd[cnt - 1])) continue;
            int64_t pos = cnt < a->k ? cnt : a->k - 1;
            while (pos > 0 && bd[pos - 1] > d) {
                bd[pos] = bd[pos - 1];
                bi[pos] = bi[pos - 1];
                pos--;
            }
            bd[pos] = d;
            bi[pos] = (uint32_t)i;
            if (cnt < a->k) cnt++;
        }
        for (int64_t j = 0; j < a->k; ++j) {
            a->out_ids[q * a->k + j] = j < cnt ? bi[j] : 0xFFFFFFFFu;
            a->out_d[q * a->k + j] = j < cnt ? bd[j] : INFINITY;
        }
    }
    free(bd);
    free(bi);
    return NULL;
}

void dko_bruteforce_f64(int metric, const float *base, int64_t n, int64_t stride,
                        int64_t dim, const float *queries, int64_t nq, int64_t k,
                        int nthreads, uint32_t *out_ids, double *out_d) {
    if (nthreads < 1) nthreads = 1;
    pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * (size_t)nthreads);
    bf_arg *args = (bf_arg *)malloc(sizeof(bf_arg) * (size_t)nthreads);
    for (int t = 0; t < nthreads; ++t) {
        int64_t base_n = nq / nthreads, extra = nq % nthreads;
        int64_t lo = t * base_n + (t < extra ? t : extra);
        int64_t hi = lo + base_n + (t < extra ? 1 : 0);
        bf_arg a = {metric, base, n, stride, dim, queries, k, lo, hi, out_ids, out_d};
        args[t] = a;
        pthread_create(&th[t], NULL, bf_worker, &args[t]);
    }
    for (int t = 0; t < nthreads; ++t) pthread_join(th[t], NULL);
    free(th);
    free(args);
}

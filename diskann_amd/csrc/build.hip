// Vamana graph construction on gfx950 with the reference's multi_insert semantics
// (diskann/src/graph/index.rs:820-1035):
//   1. candidate generation on the pre-batch graph: batched beam search (L_build, beam 1)
//      recording every expanded node (VisitedSearchRecord, diskann/src/graph/search/record.rs)
//   2. robust_prune_with (index.rs:2481-2537): pool = record + intra-batch extras
//      (async_tools::around), sorted by distance, truncated to max_occlusion, occlude_list
//      (index.rs:2570-2782)
//   3. optional bootstrap (index.rs:595-760) and set_neighbors_bulk
//   4. back edges aggregated per target, sources sorted ascending, add_edge_and_prune
//      (index.rs:2269-2346) with robust_prune_list (index.rs:2388-2450) on overflow.
// Equal distances in a candidate pool keep their input order (stable sort); the
// reference's select_nth_unstable / sort_unstable leave that order unspecified.
#include <rocprim/device/device_radix_sort.hpp>

#include <algorithm>
#include <vector>

#include "dkm_build.h"
#include "dkm_device.h"
#include "dkm_sq.h"

namespace dkm {

constexpr int kPruneThreads = 256;
constexpr int kHalves = kPruneThreads / 32;

__device__ inline uint32_t order_key(float d) {
    uint32_t u = __float_as_uint(d == 0.0f ? 0.0f : d);
    if (d != d) return 0xFFFFFFFFu;
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

struct PruneLds {
    uint64_t *keys;   // sort capacity (power of two)
    uint32_t *in_id;  // unsorted candidates
    float *in_d;
    uint32_t *pid;    // sorted pool (<= max_occlusion)
    float *pd;
    float *occ;
    uint32_t *last;
    uint32_t *acc;    // accepted pool positions (<= R)
    float *accv;      // accepted vectors cache (R x dim) or nullptr
    float *cand;      // candidate / location vector (dim)
    float *dist;      // per-half distances
    int32_t *misc;
    uint8_t *in_fp;   // Hybrid prune: full-precision flag of in_id entries
    uint8_t *pfp;     // ... and of the sorted pool
};

__host__ __device__ inline size_t a16(size_t x) { return (x + 15) & ~size_t(15); }

__host__ __device__ inline size_t prune_lds_bytes(const BuildParams &p) {
    size_t b = a16(sizeof(uint64_t) * p.sort_cap) + 2 * a16(sizeof(uint32_t) * p.sort_cap) +
               4 * a16(sizeof(uint32_t) * p.max_occlusion) + a16(sizeof(uint32_t) * p.R) +
               a16(sizeof(float) * p.dim) + a16(sizeof(float) * kHalves) + 64 +
               a16(p.sort_cap) + a16(p.max_occlusion);
    if (p.cache_accepted) b += a16(sizeof(float) * (size_t)p.R * p.dim);
    return b;
}

__device__ inline PruneLds carve_prune(const BuildParams &p, char *base) {
    PruneLds s;
    size_t o = 0;
    auto take = [&](size_t bytes) {
        char *r = base + o;
        o += a16(bytes);
        return r;
    };
    s.keys = (uint64_t *)take(sizeof(uint64_t) * p.sort_cap);
    s.in_id = (uint32_t *)take(sizeof(uint32_t) * p.sort_cap);
    s.in_d = (float *)take(sizeof(float) * p.sort_cap);
    s.pid = (uint32_t *)take(sizeof(uint32_t) * p.max_occlusion);
    s.pd = (float *)take(sizeof(float) * p.max_occlusion);
    s.occ = (float *)take(sizeof(float) * p.max_occlusion);
    s.last = (uint32_t *)take(sizeof(uint32_t) * p.max_occlusion);
    s.acc = (uint32_t *)take(sizeof(uint32_t) * p.R);
    s.cand = (float *)take(sizeof(float) * p.dim);
    s.dist = (float *)take(sizeof(float) * kHalves);
    s.misc = (int32_t *)take(64);
    s.in_fp = (uint8_t *)take(p.sort_cap);
    s.pfp = (uint8_t *)take(p.max_occlusion);
    s.accv = p.cache_accepted ? (float *)take(sizeof(float) * (size_t)p.R * p.dim) : nullptr;
    return s;
}

template <int METRIC>
__device__ inline float pair_distance(const float *a, const float *b, int dim, int hp) {
    return half_distance<METRIC>(a, b, dim, hp);
}

__device__ inline float metric_distance(int metric, const float *a, const float *b, int dim,
                                        int hp) {
    if (metric == METRIC_L2) return half_distance<METRIC_L2>(a, b, dim, hp);
    if (metric == METRIC_IP) return half_distance<METRIC_IP>(a, b, dim, hp);
    if (metric == METRIC_COSNORM) return half_distance<METRIC_COSNORM>(a, b, dim, hp);
    return half_distance<METRIC_COSINE>(a, b, dim, hp);
}

// The SQ build's distance between rows a and b (PruneAccessor's DistanceComputer over two
// stored codes, scalar.rs:370-446): the integer sums of the code rows by one half-wave
// (exact in any order) and the f32 finish of the compensated metric; valid in every lane.
__device__ inline float sq_pair(const BuildParams &p, uint32_t a, uint32_t b, int hp) {
    const uint4 *ra = (const uint4 *)(p.sq_codes + (size_t)a * p.sq_cstride);
    const uint4 *rb = (const uint4 *)(p.sq_codes + (size_t)b * p.sq_cstride);
    uint32_t ab = 0, aa = 0, bb = 0;
    for (int c = hp; c < p.sq_cstride / 16; c += 32) sq_dot16(ra[c], rb[c], ab, aa, bb);
#pragma unroll
    for (int off = 16; off > 0; off >>= 1) {
        ab += __shfl_xor(ab, off, 32);
        aa += __shfl_xor(aa, off, 32);
        bb += __shfl_xor(bb, off, 32);
    }
    // sum (a - b)^2 = aa + bb - 2 ab, exact: each sum is at most dim * 255^2
    return sq_finish(p.metric, p.sq_bss, p.sq_ssn, ab, aa + bb - 2u * ab, p.sq_comp[a],
                     p.sq_comp[b]);
}

// The spherical build's distance between rows a and b (the store's data x data distance
// computer, spherical.rs:293-350): the 1-bit inner product of the code rows by one
// half-wave (an exact popcount sum) and the compensated metric's f32 finish.
__device__ inline float sph_pair(const BuildParams &p, uint32_t a, uint32_t b, int hp) {
    const uint4 *ra = (const uint4 *)(p.sph_bits + (size_t)a * p.sph_cstride);
    const uint4 *rb = (const uint4 *)(p.sph_bits + (size_t)b * p.sph_cstride);
    uint32_t ip = 0;
    for (int c = hp; c < p.sph_cstride / 16; c += 32) {
        const uint4 x = ra[c], y = rb[c];
        ip += __builtin_popcount(x.x & y.x) + __builtin_popcount(x.y & y.y) +
              __builtin_popcount(x.z & y.z) + __builtin_popcount(x.w & y.w);
    }
#pragma unroll
    for (int off = 16; off > 0; off >>= 1) ip += __shfl_xor(ip, off, 32);
    return sph_data_data(p.sph_metric, p.sph_shift_sq, p.sph_dim, ip, p.sph_meta[a],
                         p.sph_meta[b]);
}

// A prune distance between rows a (vector va, staged) and b (vector vb) by one half-wave:
// the metric on the vectors, or the SQ / spherical build's code distance. Valid in hp == 0.
__device__ inline float row_distance(const BuildParams &p, uint32_t a, const float *va,
                                     uint32_t b, const float *vb, int hp) {
    if (p.sph) return sph_pair(p, a, b, hp);
    if (p.sq) return sq_pair(p, a, b, hp);
    return metric_distance(p.metric, va, vb, p.dim, hp);
}

// HybridComputer (diskann-providers/src/model/graph/provider/async_/distances.rs:118-165)
// between rows a and b by one half-wave: both full precision -> the metric on the vectors
// va, vb; one full precision -> the direct PQ distance of that vector against the other's
// code; both quantized -> code against code (a on the left). Each group of eight lanes of
// the half computes the PQ forms (the same value); the result is valid in hp == 0.
__device__ inline float hybrid_distance(const BuildParams &p, uint32_t a, bool fa,
                                        const float *va, uint32_t b, bool fb, const float *vb,
                                        int hp) {
    if (fa && fb) return metric_distance(p.metric, va, vb, p.dim, hp);
    const uint8_t *ca = p.codes + (size_t)a * p.pq_M;
    const uint8_t *cb = p.codes + (size_t)b * p.pq_M;
    const int l = hp & 7;
    if (fa) return pq_direct8_rt(p.metric, p.pivots, p.dim, p.pq_off, p.pq_M, va, nullptr, cb, l);
    if (fb) return pq_direct8_rt(p.metric, p.pivots, p.dim, p.pq_off, p.pq_M, vb, nullptr, ca, l);
    return pq_direct8_rt(p.metric, p.pivots, p.dim, p.pq_off, p.pq_M, nullptr, ca, cb, l);
}

// Full-precision membership (HybridPruneAccessor::fill keeps the first max_fp ids of the
// fill iterator in a hash set, product.rs:184-191): `id`, whose own fill position is pos,
// is full precision iff it is among fill(0) .. fill(nfp - 1).
template <class Fill>
__device__ inline uint8_t fp_member(uint32_t id, int pos, int nfp, Fill fill) {
    if (pos < nfp) return 1;
    for (int f = 0; f < nfp; ++f)
        if (fill(f) == id) return 1;
    return 0;
}

// Stable sort of s.in_id / s.in_d [0, n) by distance into s.pid / s.pd (first `keep`).
__device__ inline void sort_pool(const BuildParams &p, PruneLds &s, int n, int keep) {
    const int t = threadIdx.x;
    int S = 1;
    while (S < n) S <<= 1;
    for (int i = t; i < S; i += blockDim.x)
        s.keys[i] = i < n ? ((uint64_t)order_key(s.in_d[i]) << 32) | (uint32_t)i : ~0ull;
    __syncthreads();
    for (int k = 2; k <= S; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int i = t; i < S; i += blockDim.x) {
                int ixj = i ^ j;
                if (ixj > i) {
                    uint64_t a = s.keys[i], b = s.keys[ixj];
                    bool up = (i & k) == 0;
                    if ((a > b) == up) {
                        s.keys[i] = b;
                        s.keys[ixj] = a;
                    }
                }
            }
            __syncthreads();
        }
    }
    for (int r = t; r < keep; r += blockDim.x) {
        uint32_t src = (uint32_t)(s.keys[r] & 0xFFFFFFFFu);
        s.pid[r] = s.in_id[src];
        s.pd[r] = s.in_d[src];
        if (p.max_fp >= 0) s.pfp[r] = s.in_fp[src];
    }
    __syncthreads();
}

// occlude_list on the sorted pool s.pid/s.pd [0, P); writes accepted ids to `out`
// (count returned, identical in every thread).
__device__ inline int occlude(const BuildParams &p, PruneLds &s, int P, uint32_t self_id,
                              bool force_saturate, uint32_t *out) {
    const int t = threadIdx.x;
    const int half = t >> 5, hp = t & 31;
    const float FMAX = 3.40282347e+38f;
    for (int i = t; i < P; i += blockDim.x) {
        s.occ[i] = 0.0f;
        s.last[i] = 0;
    }
    __syncthreads();
    const float alpha = p.alpha;
    const float inc = alpha < 1.2f ? alpha : 1.2f;
    float cur = 1.0f;
    int found = 0;
    while (found < p.R) {
        for (int i = 0; i < P; ++i) {
            if (found >= p.R) break;
            float of = s.occ[i];
            if (of > cur) continue;
            const uint32_t cid = s.pid[i];
            if (cid == self_id) {
                __syncthreads();
                if (t == 0) s.occ[i] = FMAX;
                __syncthreads();
                continue;
            }
            uint32_t lc = s.last[i];
            // stage the candidate vector
            const float *cv = p.vectors + (size_t)cid * p.vstride;
            for (int e = t; e < p.dim; e += blockDim.x) s.cand[e] = cv[e];
            __syncthreads();
            bool rejected = false;
            while ((int)lc < found && !rejected) {
                // distances for accepted slots lc .. lc+kHalves-1 (one per half-wave)
                const int slot = (int)lc + half;
                const uint32_t rp = slot < found ? s.acc[slot] : 0u;
                const bool need = slot < found && (int)rp < i;
                const float *av = !need ? s.cand
                                  : (s.accv ? s.accv + (size_t)slot * p.dim
                                            : p.vectors + (size_t)s.pid[rp] * p.vstride);
                const float d = p.max_fp >= 0 && need
                                    ? hybrid_distance(p, cid, s.pfp[i] != 0, s.cand, s.pid[rp],
                                                      s.pfp[rp] != 0, av, hp)
                                    : row_distance(p, cid, s.cand, need ? s.pid[rp] : cid, av, hp);
                if (slot < found && hp == 0) s.dist[half] = need ? d : 0.0f;
                __syncthreads();
                if (t == 0) {
                    // sequential rule of occlude_list over this chunk of accepted slots
                    const int lc0 = (int)lc;
                    int end = found < lc0 + kHalves ? found : lc0 + kHalves;
                    int stop = 0;
                    while ((int)lc < end) {
                        const int sl = (int)lc;
                        uint32_t rp = s.acc[sl];
                        lc += 1;
                        if ((int)rp >= i) continue;
                        float d = s.dist[sl - lc0];
                        float dik = s.pd[i];
                        if (p.metric == METRIC_IP) {
                            if (d < cur * dik) of = cur + 0.01f;
                        } else {
                            of = (d == 0.0f) ? FMAX : fmaxf(of, dik / d);
                        }
                        if (of > cur) {
                            stop = 1;
                            break;
                        }
                    }
                    s.misc[0] = stop;
                    s.misc[1] = (int)lc;
                    s.misc[2] = __float_as_int(of);
                }
                __syncthreads();
                rejected = s.misc[0] != 0;
                lc = (uint32_t)s.misc[1];
                of = __int_as_float(s.misc[2]);
                __syncthreads();
            }
            if (t == 0) {
                s.last[i] = lc;
                if (of > cur) {
                    s.occ[i] = of;
                } else {
                    s.occ[i] = FMAX;
                    s.acc[found] = (uint32_t)i;
                }
            }
            bool accept = !(of > cur);
            if (accept && s.accv) {
                for (int e = t; e < p.dim; e += blockDim.x) s.accv[(size_t)found * p.dim + e] = s.cand[e];
            }
            __syncthreads();
            if (accept) found += 1;
        }
        if (cur == alpha) break;
        cur = fminf(cur * inc, alpha);
    }
    for (int n = t; n < found; n += blockDim.x) out[n] = s.pid[s.acc[n]];
    __syncthreads();
    int count = found;
    if (force_saturate || (p.saturate && alpha > 1.0f)) {
        if (t == 0) {
            for (int i = 0; i < P && count < p.R; ++i) {
                uint32_t id = s.pid[i];
                if (id == self_id) continue;
                bool dup = false;
                for (int j = 0; j < count; ++j) dup = dup || out[j] == id;
                if (!dup) out[count++] = id;
            }
            s.misc[3] = count;
        }
        __syncthreads();
        count = s.misc[3];
    }
    return count;
}

// Candidate generation: pool = visited record of the point's search + intra-batch
// extras (circular window around its batch position).
__global__ __launch_bounds__(kPruneThreads) void prune_search_kernel(BuildParams p) {
    extern __shared__ __align__(16) char smem[];
    const int b = blockIdx.x;
    if (b >= p.batch) return;
    PruneLds s = carve_prune(p, smem);
    const int t = threadIdx.x;
    const int half = t >> 5, hp = t & 31;
    const uint32_t self_id = p.batch_ids[b];
    int nrec = (int)p.rec_len[b];
    if (nrec > (int)p.rec_cap) {
        if (t == 0) atomicOr(p.status, BUILD_ERR_RECORD);
        nrec = (int)p.rec_cap;
    }
    const int B = p.batch;
    int want = p.intra_batch < 0 ? B - 1 : (p.intra_batch < B - 1 ? p.intra_batch : B - 1);
    if (want < 0) want = 0;
    int n = nrec + want;
    if (n > p.sort_cap) {
        if (t == 0) atomicOr(p.status, BUILD_ERR_POOL);
        n = p.sort_cap;
        if (want > n - nrec) want = n - nrec;
    }
    for (int i = t; i < nrec; i += blockDim.x) {
        s.in_id[i] = p.rec_ids[(size_t)b * p.rec_cap + i];
        s.in_d[i] = p.rec_d[(size_t)b * p.rec_cap + i];
    }
    if (want > 0) {
        // around(ids, b, want): start ceil(want/2) before b (wrapping), skip b
        const int halfw = (want + 1) / 2;
        const int start = b >= halfw ? b - halfw : B - (halfw - b);
        for (int e = t; e < want; e += blockDim.x) {
            int q = start + e + (e >= halfw ? 1 : 0);
            if (q >= B) q -= B;
            s.in_id[nrec + e] = p.batch_ids[q];
        }
        __syncthreads();
    }
    const float *self_v = p.vectors + (size_t)self_id * p.vstride;
    for (int e = t; e < p.dim; e += blockDim.x) s.cand[e] = self_v[e];
    // Hybrid: fill order chain(once(self), chain(extras, record)).take(max_occlusion)
    const bool hyb = p.max_fp >= 0;
    int nfp = 0;
    if (hyb) {
        __syncthreads();  // record and extras ids visible to every thread
        const int lim = p.max_fp < p.max_occlusion ? p.max_fp : p.max_occlusion;
        nfp = 1 + want + nrec < lim ? 1 + want + nrec : lim;
        auto fill = [&](int f) -> uint32_t {
            return f == 0 ? self_id : (f <= want ? s.in_id[nrec + f - 1] : s.in_id[f - 1 - want]);
        };
        for (int i = t; i < n; i += blockDim.x)
            s.in_fp[i] = fp_member(s.in_id[i], i < nrec ? 1 + want + i : 1 + i - nrec, nfp, fill);
    }
    __syncthreads();
    const bool self_fp = !hyb || nfp >= 1;
    for (int e0 = 0; e0 < want; e0 += kHalves) {
        int e = e0 + half;
        bool ok = e < want;
        uint32_t id = ok ? s.in_id[nrec + e] : self_id;
        const float *v = p.vectors + (size_t)id * p.vstride;
        float d = hyb ? hybrid_distance(p, self_id, self_fp, s.cand, id,
                                        ok ? s.in_fp[nrec + e] != 0 : self_fp, v, hp)
                      : row_distance(p, self_id, s.cand, id, v, hp);
        if (ok && hp == 0) s.in_d[nrec + e] = d;
    }
    __syncthreads();
    int keep = n < p.max_occlusion ? n : p.max_occlusion;
    sort_pool(p, s, n, keep);
    uint32_t *out = p.pend + (size_t)b * p.R;
    int c = occlude(p, s, keep, self_id, false, out);
    if (t == 0) p.pend_cnt[b] = (uint32_t)c;
}

// robust_prune_list over explicit candidate lists (bootstrap: forced saturation).
__global__ __launch_bounds__(kPruneThreads) void prune_list_kernel(BuildParams p,
                                                                   const uint32_t *lists,
                                                                   const uint32_t *list_len,
                                                                   int list_stride,
                                                                   int force_saturate) {
    extern __shared__ __align__(16) char smem[];
    const int b = blockIdx.x;
    if (b >= p.batch) return;
    PruneLds s = carve_prune(p, smem);
    const int t = threadIdx.x;
    const int half = t >> 5, hp = t & 31;
    const uint32_t loc = p.batch_ids[b];
    int n = 0;
    const uint32_t *lst = lists + (size_t)b * list_stride;
    int len = (int)list_len[b];
    if (len > 8 * p.sort_cap) len = 8 * p.sort_cap;
    // Hybrid: fill order chain(once(location), list); flags by list position in the sort
    // keys (unused until the sort)
    const bool hyb = p.max_fp >= 0;
    uint8_t *lfp = (uint8_t *)s.keys;
    int nfp = 0;
    if (hyb) {
        nfp = 1 + len < p.max_fp ? 1 + len : p.max_fp;
        auto fill = [&](int f) -> uint32_t { return f == 0 ? loc : lst[f - 1]; };
        for (int i = t; i < len; i += blockDim.x) lfp[i] = fp_member(lst[i], 1 + i, nfp, fill);
        __syncthreads();
    }
    const bool loc_fp = !hyb || nfp >= 1;
    // drop self occurrences, keep order
    if (t == 0) {
        for (int i = 0; i < len && n < p.sort_cap; ++i)
            if (lst[i] != loc) {
                if (hyb) s.in_fp[n] = lfp[i];
                s.in_id[n++] = lst[i];
            }
        s.misc[4] = n;
    }
    const float *lv = p.vectors + (size_t)loc * p.vstride;
    for (int e = t; e < p.dim; e += blockDim.x) s.cand[e] = lv[e];
    __syncthreads();
    n = s.misc[4];
    for (int e0 = 0; e0 < n; e0 += kHalves) {
        int e = e0 + half;
        bool ok = e < n;
        uint32_t id = ok ? s.in_id[e] : loc;
        const float *v = p.vectors + (size_t)id * p.vstride;
        float d = hyb ? hybrid_distance(p, loc, loc_fp, s.cand, id, ok ? s.in_fp[e] != 0 : loc_fp,
                                        v, hp)
                      : row_distance(p, loc, s.cand, id, v, hp);
        if (ok && hp == 0) s.in_d[e] = d;
    }
    __syncthreads();
    int keep = n < p.max_occlusion ? n : p.max_occlusion;
    sort_pool(p, s, n, keep);
    uint32_t *out = p.pend + (size_t)b * p.R;
    int c = keep > 0 ? occlude(p, s, keep, loc, force_saturate != 0, out) : 0;
    if (t == 0) p.pend_cnt[b] = (uint32_t)c;
}

__global__ void set_lists_kernel(BuildParams p) {
    const int b = blockIdx.x;
    if (b >= p.batch) return;
    uint32_t *row = p.graph + (size_t)p.batch_ids[b] * p.gstride;
    const uint32_t c = p.pend_cnt[b];
    for (int j = threadIdx.x; j < (int)c; j += blockDim.x) row[j] = p.pend[(size_t)b * p.R + j];
    if (threadIdx.x == 0) row[p.gstride - 1] = c;
}

__global__ void make_pairs_kernel(BuildParams p, uint64_t *pairs, uint32_t *npairs) {
    const int b = blockIdx.x;
    if (b >= p.batch) return;
    const uint32_t c = p.pend_cnt[b];
    __shared__ uint32_t base;
    if (threadIdx.x == 0) base = atomicAdd(npairs, c);
    __syncthreads();
    for (int j = threadIdx.x; j < (int)c; j += blockDim.x)
        pairs[base + j] = ((uint64_t)p.pend[(size_t)b * p.R + j] << 32) | p.batch_ids[b];
}

__global__ void count_heads_kernel(const uint64_t *pairs, uint32_t n, uint32_t *nheads) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    bool head = i == 0 || (pairs[i] >> 32) != (pairs[i - 1] >> 32);
    if (head) atomicAdd(nheads, 1u);
}

// add_edge_and_prune for one target per block (the block whose pair index is the head
// of the target's run of sorted (target, source) pairs).
__global__ __launch_bounds__(kPruneThreads) void backedge_kernel(BuildParams p,
                                                                 const uint64_t *pairs,
                                                                 uint32_t npairs) {
    extern __shared__ __align__(16) char smem[];
    const uint32_t i0 = blockIdx.x;
    if (i0 >= npairs) return;
    const uint32_t target = (uint32_t)(pairs[i0] >> 32);
    if (i0 > 0 && (uint32_t)(pairs[i0 - 1] >> 32) == target) return;
    uint32_t i1 = i0 + 1;
    while (i1 < npairs && (uint32_t)(pairs[i1] >> 32) == target) ++i1;
    PruneLds s = carve_prune(p, smem);
    const int t = threadIdx.x;
    const int half = t >> 5, hp = t & 31;
    uint32_t *row = p.graph + (size_t)target * p.gstride;
    if (t == 0) {
        int n = (int)row[p.gstride - 1];
        for (int j = 0; j < n; ++j) s.in_id[j] = row[j];
        int added = 0;
        for (uint32_t k = i0; k < i1; ++k) {
            uint32_t src = (uint32_t)pairs[k];
            bool dup = false;
            for (int j = 0; j < n; ++j) dup = dup || s.in_id[j] == src;
            if (!dup && n < p.sort_cap) {
                s.in_id[n++] = src;
                added++;
            }
        }
        s.misc[4] = n;
        s.misc[5] = added;
    }
    __syncthreads();
    const int n = s.misc[4];
    const int added = s.misc[5];
    if (added == 0) return;
    if (n <= p.R_max) {
        for (int j = t; j < n; j += blockDim.x) row[j] = s.in_id[j];
        if (t == 0) row[p.gstride - 1] = (uint32_t)n;
        return;
    }
    // robust_prune_list(target, merged list): Hybrid fill order chain(once(target), list)
    const bool hyb = p.max_fp >= 0;
    int nfp = 0;
    if (hyb) {
        nfp = 1 + n < p.max_fp ? 1 + n : p.max_fp;
        auto fill = [&](int f) -> uint32_t { return f == 0 ? target : s.in_id[f - 1]; };
        for (int j = t; j < n; j += blockDim.x) s.in_fp[j] = fp_member(s.in_id[j], 1 + j, nfp, fill);
        __syncthreads();
    }
    const bool tgt_fp = !hyb || nfp >= 1;
    // distances from the target, drop self
    if (t == 0) {
        int m = 0;
        for (int j = 0; j < n; ++j)
            if (s.in_id[j] != target) {
                if (hyb) s.in_fp[m] = s.in_fp[j];
                s.in_id[m++] = s.in_id[j];
            }
        s.misc[4] = m;
    }
    const float *lv = p.vectors + (size_t)target * p.vstride;
    for (int e = t; e < p.dim; e += blockDim.x) s.cand[e] = lv[e];
    __syncthreads();
    const int m = s.misc[4];
    for (int e0 = 0; e0 < m; e0 += kHalves) {
        int e = e0 + half;
        bool ok = e < m;
        uint32_t id = ok ? s.in_id[e] : target;
        const float *v = p.vectors + (size_t)id * p.vstride;
        float d = hyb ? hybrid_distance(p, target, tgt_fp, s.cand, id,
                                        ok ? s.in_fp[e] != 0 : tgt_fp, v, hp)
                      : row_distance(p, target, s.cand, id, v, hp);
        if (ok && hp == 0) s.in_d[e] = d;
    }
    __syncthreads();
    int keep = m < p.max_occlusion ? m : p.max_occlusion;
    sort_pool(p, s, m, keep);
    uint32_t *out = s.in_id;  // pool copied out of in_id already; reuse as output
    int c = occlude(p, s, keep, target, false, out);
    for (int j = t; j < c; j += blockDim.x) row[j] = out[j];
    if (t == 0) row[p.gstride - 1] = (uint32_t)c;
}

__global__ void gather_rows_kernel(const float *vectors, int64_t vstride, int dim,
                                   const uint32_t *ids, int n, float *out) {
    const int r = blockIdx.x;
    if (r >= n) return;
    const float *src = vectors + (size_t)ids[r] * vstride;
    for (int e = threadIdx.x; e < dim; e += blockDim.x) out[(size_t)r * dim + e] = src[e];
}

size_t build_prune_lds_bytes(const BuildParams &p) { return prune_lds_bytes(p); }

hipError_t launch_gather_rows(const float *vectors, int64_t vstride, int dim, const uint32_t *ids,
                              int n, float *out, hipStream_t stream) {
    hipLaunchKernelGGL(gather_rows_kernel, dim3(n), dim3(128), 0, stream, vectors, vstride, dim,
                       ids, n, out);
    return hipGetLastError();
}

hipError_t launch_prune_search(const BuildParams &p, hipStream_t stream) {
    size_t lds = prune_lds_bytes(p);
    hipError_t e = hipFuncSetAttribute((const void *)prune_search_kernel,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(prune_search_kernel, dim3(p.batch), dim3(kPruneThreads), lds, stream, p);
    return hipGetLastError();
}

hipError_t launch_prune_lists(const BuildParams &p, const uint32_t *lists, const uint32_t *lens,
                              int stride, int force_saturate, hipStream_t stream) {
    size_t lds = prune_lds_bytes(p);
    hipError_t e = hipFuncSetAttribute((const void *)prune_list_kernel,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(prune_list_kernel, dim3(p.batch), dim3(kPruneThreads), lds, stream, p,
                       lists, lens, stride, force_saturate);
    return hipGetLastError();
}

hipError_t launch_set_lists(const BuildParams &p, hipStream_t stream) {
    hipLaunchKernelGGL(set_lists_kernel, dim3(p.batch), dim3(64), 0, stream, p);
    return hipGetLastError();
}

hipError_t launch_backedges(const BuildParams &p, uint64_t *pairs, uint64_t *pairs_tmp,
                            uint32_t *counter, void *sort_tmp, size_t sort_tmp_bytes,
                            uint32_t *out_npairs, uint32_t *out_nheads, hipStream_t stream) {
    hipError_t e = hipMemsetAsync(counter, 0, 2 * sizeof(uint32_t), stream);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(make_pairs_kernel, dim3(p.batch), dim3(64), 0, stream, p, pairs, counter);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    uint32_t np = 0;
    if ((e = hipMemcpyAsync(&np, counter, sizeof(uint32_t), hipMemcpyDeviceToHost, stream)) != hipSuccess) return e;
    if ((e = hipStreamSynchronize(stream)) != hipSuccess) return e;
    *out_npairs = np;
    if (np == 0) {
        *out_nheads = 0;
        return hipSuccess;
    }
    size_t need = sort_tmp_bytes;
    e = rocprim::radix_sort_keys(sort_tmp, need, pairs, pairs_tmp, np, 0, 64, stream);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(count_heads_kernel, dim3((np + 255) / 256), dim3(256), 0, stream, pairs_tmp,
                       np, counter + 1);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    uint32_t nh = 0;
    if ((e = hipMemcpyAsync(&nh, counter + 1, sizeof(uint32_t), hipMemcpyDeviceToHost, stream)) != hipSuccess) return e;
    if ((e = hipStreamSynchronize(stream)) != hipSuccess) return e;
    *out_nheads = nh;
    return hipSuccess;
}

hipError_t launch_apply_backedges(const BuildParams &p, const uint64_t *sorted_pairs,
                                  uint32_t npairs, hipStream_t stream) {
    size_t lds = prune_lds_bytes(p);
    hipError_t e = hipFuncSetAttribute((const void *)backedge_kernel,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(backedge_kernel, dim3(npairs), dim3(kPruneThreads), lds, stream, p,
                       sorted_pairs, npairs);
    return hipGetLastError();
}

size_t radix_sort_tmp_bytes(uint32_t n) {
    size_t bytes = 0;
    (void)rocprim::radix_sort_keys(nullptr, bytes, (uint64_t *)nullptr, (uint64_t *)nullptr, n, 0, 64);
    return bytes;
}

}  // namespace dkm

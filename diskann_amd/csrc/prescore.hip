// Candidate pre-scoring for exhaustive top-k (the precision layer of the 1M x 768
// inner-product configuration): the base rows and the queries are rounded to bf16 and
// scored as an MFMA GEMM with FP32 accumulation; the best `C` candidates of each query by
// that approximate score are then re-scored exactly in FP32 with the reference's distance
// arithmetic (the same functions as Rerank, diskann-providers/src/model/graph/provider/
// async_/inmem/full_precision.rs:357-404) and the top k of those are returned.
//
// Selecting the best C of n approximate scores per query without materialising n scores:
//   1. the queries are scored against a fixed sample of the base rows (dense output) and a
//      per-query threshold tau is read off the sample's order statistics (the distance the
//      j-th best sample row has, j chosen so that about 1.5 C rows of the base score at or
//      below it);
//   2. the full GEMM appends (orderable distance << 32 | row) for every pair at or below its
//      query's tau to that query's candidate list (atomic counter per query);
//   3. queries whose list holds fewer than C entries (or more than its capacity) are redone
//      with a looser (tighter) threshold; every list is sorted by (distance, row) and its
//      first C rows are the candidates.
// Approximate distance per (query q, row x): inner product / cosine-normalized -q.x,
// squared L2 |x|^2 - 2 q.x (|q|^2 is common to all rows of a query).
#include <hip/hip_runtime.h>
#include <rocprim/device/device_segmented_radix_sort.hpp>

#include <algorithm>
#include <cstdint>

#include "dkm_device.h"
#include "dkm_mfma.h"
#include "dkm_prescore.h"

namespace dkm {

// Ascending order of the key = ascending order of the float (NaN not expected).
__device__ __forceinline__ uint32_t orderable(float f) {
    const uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float from_orderable(uint32_t k) {
    return __uint_as_float((k & 0x80000000u) ? (k & 0x7FFFFFFFu) : ~k);
}

// Approximate distance from the accumulated dot product (see the file comment).
__device__ __forceinline__ float approx_dist(bool l2, float dot, float xn) {
    return l2 ? xn - 2.0f * dot : -dot;
}

// Dense scores of the sample rows against a block of queries: out[q * ld + row] for
// row < ns, q < nq. xf: sample rows (fragment layout of the operand policy G, padded to
// 256), qf: queries (same layout, padded to a multiple of 64), xn: row norms (L2 only).
// inv_s undoes the power-of-two operand scales (1 for bf16).
template <class G>
__global__ __launch_bounds__(256, 1) void ps_dense_kernel(const uint16_t *xf, int64_t ns,
                                                          const float *xn, const uint16_t *qf,
                                                          int nq, int nqp, bool l2, float inv_s,
                                                          float *out, int64_t ld) {
    extern __shared__ __align__(16) uint16_t sm[];
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int half = lane >> 5, l32 = lane & 31;
    const int64_t rtile = (int64_t)blockIdx.x * 8 + 2 * wave;
    G g;
    g.load(xf + rtile * G::kBlocks * 512 + lane * 8);
    const int64_t r0 = 32 * rtile + l32, r1 = r0 + 32;
    const float n0 = l2 && r0 < ns ? xn[r0] : 0.0f, n1 = l2 && r1 < ns ? xn[r1] : 0.0f;
    g.sweep(
        sm, qf, nqp / 32, wave, lane,
        [&](int, floatx16_t &pa, floatx16_t &pb) {
            pa = floatx16_t{};
            pb = floatx16_t{};
        },
        [&](int t, floatx16_t &pa, floatx16_t &pb) {
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int q = 32 * t + 8 * (r >> 2) + 4 * half + (r & 3);
                if (q < nq) {
                    if (r0 < ns) out[(int64_t)q * ld + r0] = approx_dist(l2, pa[r] * inv_s, n0);
                    if (r1 < ns) out[(int64_t)q * ld + r1] = approx_dist(l2, pb[r] * inv_s, n1);
                }
            }
        });
}

// Threshold per query: the ranks j[0..2] smallest of the query's ns sample distances, by a
// radix select over orderable keys held in LDS (ns <= kSampleMax).
constexpr int kSampleMax = PS_SAMPLE_MAX;

__global__ __launch_bounds__(256) void ps_select_kernel(const float *dense, int64_t ld, int ns,
                                                        int j0, int j1, int j2, float *tau,
                                                        int nq) {
    __shared__ uint32_t keys[kSampleMax];
    __shared__ uint32_t hist[256];
    __shared__ uint32_t sel[2];
    __shared__ uint32_t wsum[4];
    const int q = blockIdx.x, tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    for (int i = tid; i < ns; i += 256) keys[i] = orderable(dense[(int64_t)q * ld + i]);
    // the three ranks descend byte by byte together; a rank whose prefix equals the previous
    // rank's reuses that level's histogram (the ranks are close, so the heavily contended
    // top-byte histogram is usually built once instead of three times)
    const int js[3] = {j0, j1, j2};
    uint32_t prefix[3] = {0, 0, 0}, want[3];
#pragma unroll
    for (int w = 0; w < 3; ++w) want[w] = (uint32_t)min(max(js[w], 0), ns - 1);
    uint32_t mask = 0;
    for (int shift = 24; shift >= 0; shift -= 8) {
        bool have = false;   // hist holds this level's histogram of the keys matching `built`
        uint32_t built = 0;  // (uniform: every thread derives it from the same broadcasts)
#pragma unroll
        for (int w = 0; w < 3; ++w) {
            if (!have || prefix[w] != built) {
                hist[tid] = 0;
                __syncthreads();
                for (int i = tid; i < ns; i += 256)
                    if ((keys[i] & mask) == prefix[w])
                        atomicAdd(&hist[(keys[i] >> shift) & 255u], 1u);
                __syncthreads();
                built = prefix[w];
                have = true;
            }
            // the first bin d whose running count passes want (bin 255 if none does): thread
            // t owns bin t and takes it when want lies in [count before t, count through t)
            const uint32_t h = hist[tid];
            uint32_t incl = h;
#pragma unroll
            for (int o = 1; o < 64; o <<= 1) {
                const uint32_t v = __shfl_up(incl, o);
                if (lane >= o) incl += v;
            }
            if (lane == 63) wsum[wave] = incl;
            __syncthreads();
            for (int k = 0; k < wave; ++k) incl += wsum[k];
            const uint32_t excl = incl - h;
            if ((excl <= want[w] && want[w] < incl) || (tid == 255 && incl <= want[w])) {
                sel[0] = (uint32_t)tid;
                sel[1] = want[w] - excl;
            }
            __syncthreads();
            prefix[w] |= sel[0] << shift;
            want[w] = sel[1];
            __syncthreads();  // sel and hist are read before they are written again
        }
        mask |= 255u << shift;
    }
    if (tid == 0) {
#pragma unroll
        for (int w = 0; w < 3; ++w) tau[w * (int64_t)nq + q] = from_orderable(prefix[w]);
    }
}

// Full GEMM with the threshold filter: every (row, query) whose approximate distance d is at
// or below tau[query] appends (key << 32 | row) to the query's list (capacity cap; the
// counter keeps counting past it so overflow is visible), the key ordering the query's
// entries like d. Queries with tau = -inf take no part (used when only some queries are
// redone). qmap (may be null) maps the launch's query q to its list / counter index. After a
// tile, each lane builds the 16-bit masks of its passing accumulators without branches; a
// wave-uniform loop then takes one set bit per lane per round (a pass rate under 1 % makes
// that one or two rounds), selects that accumulator with a binary tree of selects (static
// register indices only) and stages the pair in the lane's own LDS slots (no atomics, no
// ballots). The staged pairs are appended to the global lists once the sweep is over, so
// no returning atomic waits inside the tile loop; pairs beyond a lane's slots are appended
// directly.
constexpr int kPool = PS_POOL;
// a wave's LDS pool holds kPool staged pairs, plus one sink slot per lane in the pipelined
// (FP8) sweep for the stores of a branch-free staging round that stage nothing
constexpr int pool_alloc(bool pipelined) { return kPool + (pipelined ? 64 : 0); }

__device__ __forceinline__ void ps_append(uint32_t *count, uint64_t *lists, int64_t cap, int q,
                                          uint64_t key) {
    const uint32_t pos = atomicAdd(&count[(int64_t)q * PS_COUNT_STRIDE], 1u);
    if (pos < cap) lists[(int64_t)q * cap + pos] = key;
}

// s: the product of the two operands' power-of-two scales (1 for bf16); the accumulators
// then hold s times the dot products, and tau and the norm terms are scaled to match.
template <class G>
__global__ __launch_bounds__(256, 1) void ps_filter_kernel(const uint16_t *xf, int64_t n,
                                                           const float *xn, const uint16_t *qf,
                                                           int nq, int nqp, bool l2, float s,
                                                           const float *tau, const uint32_t *qmap,
                                                           uint32_t *count, uint64_t *lists,
                                                           int64_t cap) {
    extern __shared__ __align__(16) uint16_t sm[];
    float *tau_s = (float *)(sm + 2 * G::kBlocks * 512);
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int half = lane >> 5, l32 = lane & 31;
    constexpr int kPoolAlloc = pool_alloc(G::kPipelined);
    uint64_t *skey = (uint64_t *)(tau_s + nqp) + wave * kPoolAlloc;
    uint16_t *sq = (uint16_t *)((uint64_t *)(tau_s + nqp) + 4 * kPoolAlloc) + wave * kPoolAlloc;
    const int64_t rtile = (int64_t)blockIdx.x * 8 + 2 * wave;
    G g;
    g.load(xf + rtile * G::kBlocks * 512 + lane * 8);
    // the first column tile's DMA goes out with the row loads, ahead of the threshold fill
    // (whose LDS stores wait for every load issued before them)
    res_prefetch<G::kBlocks>(qf + lane * 8, sm, wave);
    for (int c = tid; c < nqp; c += 256) tau_s[c] = c < nq ? tau[c] : -__builtin_inff();
    const int64_t r0 = 32 * rtile + l32, r1 = r0 + 32;
    // the accumulators start at tau (L2: tau / 2), so after the tile they hold tau - d
    // (L2: (tau - d + |x|^2) / 2): a pair passes when that, less the row's |x|^2 / 2 (L2),
    // is >= 0, and its negation orders a query's entries like d
    const float h0 = l2 && r0 < n ? 0.5f * s * xn[r0] : 0.0f;
    const float h1 = l2 && r1 < n ? 0.5f * s * xn[r1] : 0.0f;
    const bool v0 = r0 < n, v1 = r1 < n;
    const float tscale = (l2 ? 0.5f : 1.0f) * s;
    constexpr int kPer = kPool / 64;  // LDS slots per lane: entry j of lane l at j * 64 + l
    uint32_t nst = 0;                  // mask epilogue: this lane's staged entries
    uint32_t filled = 0;               // ballot epilogue: the wave's pool entries (uniform)
    // accumulator r of acc, r lane-varying: a binary tree of bitwise selects under opaque
    // masks (a select tree the compiler can see through is folded into a dynamic element
    // index, which it lowers through scratch memory)
    auto pick = [](const floatx16_t &acc, int r) {
        uint32_t msk[4];
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            msk[b] = ((r >> b) & 1) ? ~0u : 0u;
            asm volatile("" : "+v"(msk[b]));
        }
        auto sel = [](uint32_t m, uint32_t x1, uint32_t x0) { return (x1 & m) | (x0 & ~m); };
        uint32_t a8[8], a4[4], a2[2];
#pragma unroll
        for (int i = 0; i < 8; ++i)
            a8[i] = sel(msk[0], __float_as_uint(acc[2 * i + 1]), __float_as_uint(acc[2 * i]));
#pragma unroll
        for (int i = 0; i < 4; ++i) a4[i] = sel(msk[1], a8[2 * i + 1], a8[2 * i]);
#pragma unroll
        for (int i = 0; i < 2; ++i) a2[i] = sel(msk[2], a4[2 * i + 1], a4[2 * i]);
        return __uint_as_float(sel(msk[3], a2[1], a2[0]));
    };
    // the pair (query q, key) of accumulator r of a row tile after tile t
    auto query_of = [&](int t, int r) { return 32 * t + 8 * (r >> 2) + 4 * half + (r & 3); };
    auto key_of = [](float u, int64_t row) {
        return ((uint64_t)orderable(-u) << 32) | (uint32_t)row;
    };
    // every set bit of m (passing accumulators of acc after tile t): a wave-uniform loop
    // takes one bit per lane per round and stages the pair in the lane's own LDS slots
    auto drain = [&](const floatx16_t &acc, uint32_t m, float h, int64_t row, int t) {
        while (__ballot(m != 0)) {
            if (m) {
                const int r = __builtin_ctz(m);
                m &= m - 1;
                const int q = query_of(t, r);
                const uint64_t key = key_of(pick(acc, r) - h, row);
                if (nst < (uint32_t)kPer) {
                    skey[nst * 64 + lane] = key;
                    sq[nst * 64 + lane] = (uint16_t)q;
                    ++nst;
                } else {
                    ps_append(count, lists, cap, qmap ? (int)qmap[q] : q, key);
                }
            }
        }
    };
    // acc - h >= 0 exactly when acc >= h (h is finite, and the difference of two distinct
    // floats is never zero), so the mask needs no subtraction; the key takes the difference
    auto scan = [&](const floatx16_t &acc, float h, bool valid, int64_t row, int t) {
        uint32_t m = 0;
#pragma unroll
        for (int r = 0; r < 16; ++r) m |= (valid && acc[r] >= h) ? (1u << r) : 0u;
        drain(acc, m, h, row, t);
    };
    // one branch-free round of drain: the lowest set bit of m is staged when the lane has
    // a free slot; otherwise the stores go to the lane's sink slot (kPer) and m keeps the bit
    auto stage_one = [&](const floatx16_t &acc, uint32_t &m, float h, int64_t row, int t) {
        const bool take = m != 0 && nst < (uint32_t)kPer;
        const int r = __builtin_ctz(m | 0x10000u) & 15;
        const uint32_t slot = (take ? nst : (uint32_t)kPer) * 64 + lane;
        skey[slot] = key_of(pick(acc, r) - h, row);
        sq[slot] = (uint16_t)query_of(t, r);
        nst += take ? 1u : 0u;
        m = take ? m & (m - 1) : m;
    };
    // ballot epilogue (kernels whose row fragments leave no registers for the masks): one
    // ballot per accumulator, slots of the wave's pool handed out by lane rank
    auto push = [&](bool pass, int q, uint64_t key) {
        const uint64_t m = __ballot(pass);
        if (m) {
            if (pass) {
                const uint32_t slot =
                    filled + __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                                       __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
                if (slot < (uint32_t)kPool) {
                    skey[slot] = key;
                    sq[slot] = (uint16_t)q;
                } else {
                    ps_append(count, lists, cap, qmap ? (int)qmap[q] : q, key);
                }
            }
            filled += (uint32_t)__popcll(m);
        }
    };
    auto init = [&](int t, floatx16_t &pa, floatx16_t &pb) {
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const float4 t4 = *(const float4 *)(tau_s + 32 * t + 8 * j + 4 * half);
            pa[4 * j + 0] = tscale * t4.x;
            pa[4 * j + 1] = tscale * t4.y;
            pa[4 * j + 2] = tscale * t4.z;
            pa[4 * j + 3] = tscale * t4.w;
        }
        pb = pa;
    };
    if constexpr (G::kPipelined) {
        // software-pipelined sweep: tile t's MFMAs run into one accumulator pair while tile
        // t - 1's threshold masks and two staging rounds per row tile, held in the other
        // pair, are threaded between them; bits left over (a lane with more passes than
        // rounds, or a full pool) go through drain after the tile
        constexpr int KB = G::kBlocks, S = KB / 2, SM = (S + 1) / 2;
        const int ntiles = nqp / 32;
        const uint32_t vm0 = v0 ? 0xFFFFu : 0u, vm1 = v1 ? 0xFFFFu : 0u;
        auto src = [&](int t) { return qf + (int64_t)t * KB * 512 + lane * 8; };
        // tile t's column fragments landed (everyone's), its accumulators started and tile
        // t + 1's prefetch issued into the buffer tile t - 1 used
        auto begin = [&](int t, floatx16_t &x0, floatx16_t &x1) {
            __builtin_amdgcn_s_waitcnt(0);
            __syncthreads();
            init(t, x0, x1);
            if (t + 1 < ntiles) res_prefetch<KB>(src(t + 1), sm + ((t + 1) & 1) * KB * 512, wave);
            return (const uint16_t *)(sm + (t & 1) * KB * 512 + lane * 8);
        };
        auto overlap = [&](int t, floatx16_t &x0, floatx16_t &x1, const floatx16_t &p0,
                           const floatx16_t &p1) {
            const uint16_t *buf = begin(t, x0, x1);
            uint32_t m0 = 0, m1 = 0;
            f8_tile_with<KB>(g.a0, g.a1, buf, x0, x1, [&](auto sc) {
                constexpr int s = decltype(sc)::value;
                if constexpr (s < SM) {
                    // accumulators [32 s / SM, 32 (s + 1) / SM) of the two row tiles
                    static_for<32 * s / SM, 32 * (s + 1) / SM>([&](auto ec) {
                        constexpr int e = decltype(ec)::value, r = e & 15;
                        if constexpr (e < 16) m0 |= p0[r] >= h0 ? (1u << r) : 0u;
                        else m1 |= p1[r] >= h1 ? (1u << r) : 0u;
                    });
                    if constexpr (s == SM - 1) {
                        m0 &= vm0;
                        m1 &= vm1;
                    }
                } else {
                    // staging rounds c = 0 .. 3 (row tile c & 1) spread over steps SM .. S - 1
                    static_for<0, 4>([&](auto cc) {
                        constexpr int c = decltype(cc)::value;
                        if constexpr (SM + c * (S - SM) / 4 == s) {
                            if constexpr ((c & 1) == 0) stage_one(p0, m0, h0, r0, t - 1);
                            else stage_one(p1, m1, h1, r1, t - 1);
                        }
                    });
                }
            });
            // the new chains complete before anything below reads them
            asm volatile("" : "+a"(x0), "+a"(x1));
            if (__ballot((m0 | m1) != 0)) {
                drain(p0, m0, h0, r0, t - 1);
                drain(p1, m1, h1, r1, t - 1);
            }
        };
        floatx16_t pa, pb, qa, qb;
        {
            const uint16_t *buf = begin(0, pa, pb);
            f8_tile<KB>(g.a0, g.a1, buf, pa, pb);
            asm volatile("" : "+a"(pa), "+a"(pb));
        }
        int t = 1;
        for (; t + 1 < ntiles; t += 2) {
            overlap(t, qa, qb, pa, pb);
            overlap(t + 1, pa, pb, qa, qb);
        }
        if (t < ntiles) {
            overlap(t, qa, qb, pa, pb);
            scan(qa, h0, v0, r0, t);
            scan(qb, h1, v1, r1, t);
        } else {
            scan(pa, h0, v0, r0, t - 1);
            scan(pb, h1, v1, r1, t - 1);
        }
    } else {
        g.template sweep<true>(sm, qf, nqp / 32, wave, lane, init,
                               [&](int t, floatx16_t &pa, floatx16_t &pb) {
            if constexpr (G::kMaskEpilogue) {
                scan(pa, h0, v0, r0, t);
                scan(pb, h1, v1, r1, t);
            } else {
#pragma unroll
                for (int r = 0; r < 16; ++r) {
                    const int q = query_of(t, r);
                    const float u0 = pa[r] - h0, u1 = pb[r] - h1;
                    push(v0 && u0 >= 0.0f, q, key_of(u0, r0));
                    push(v1 && u1 >= 0.0f, q, key_of(u1, r1));
                }
            }
        });
    }
    // the staged pairs, counted per query in LDS first (the thresholds are no longer needed):
    // one global atomic per (workgroup, query) reserves the query's range, and each pair
    // lands at that base plus its rank among the workgroup's pairs of the query. In both
    // layouts entry j of this lane sits at j * 64 + lane.
    const uint32_t staged = min(filled, (uint32_t)kPool);
    auto held = [&](int j) {
        return G::kMaskEpilogue ? (uint32_t)j < nst : (uint32_t)(j * 64 + lane) < staged;
    };
    uint32_t *qcnt = (uint32_t *)tau_s;
    __syncthreads();
    for (int c = tid; c < nqp; c += 256) qcnt[c] = 0;
    __syncthreads();
    uint32_t rank[kPer];
#pragma unroll
    for (int j = 0; j < kPer; ++j) {
        const uint32_t i = (uint32_t)(j * 64 + lane);
        if (held(j)) rank[j] = atomicAdd(&qcnt[sq[i]], 1u);
    }
    __syncthreads();
    for (int c = tid; c < nqp; c += 256) {
        const uint32_t k = qcnt[c];  // nonzero only for real queries (padding has tau -inf)
        if (k) qcnt[c] = atomicAdd(&count[(int64_t)(qmap ? qmap[c] : c) * PS_COUNT_STRIDE], k);
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kPer; ++j) {
        const uint32_t i = (uint32_t)(j * 64 + lane);
        if (held(j)) {
            const int q = sq[i];
            const uint32_t pos = qcnt[q] + rank[j];
            if (pos < cap) lists[(int64_t)(qmap ? qmap[q] : q) * cap + pos] = skey[i];
        }
    }
}

// Row norms in f32 (the L2 approximate distance's |x|^2; any fixed order serves).
__global__ __launch_bounds__(256) void ps_norms_kernel(const float *x, int64_t n, int dim,
                                                       const int64_t *rows, float *out) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float *r = x + (rows ? rows[i] : i) * (int64_t)dim;
    float s = 0.0f;
    for (int d = 0; d < dim; ++d) s = fma_(r[d], r[d], s);
    out[i] = s;
}

// Gathers sample rows into a contiguous matrix.
__global__ __launch_bounds__(256) void ps_gather_rows_kernel(const float *x, int dim,
                                                             const int64_t *rows, int64_t ns,
                                                             float *out) {
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= ns * dim) return;
    const int64_t i = e / dim;
    out[e] = x[rows[i] * dim + (e - i * dim)];
}

// The key of rank `want` (0-based, ascending) among the nl distinct keys at l, by an 8-bit
// radix select over the 64 bits; called by all 256 threads of the block, which all get it.
// want < nl is required (it keeps want below the count of keys matching each prefix).
__device__ uint64_t block_kth_key(const uint64_t *l, int64_t nl, uint32_t want, uint32_t *hist,
                                  uint32_t *wsum, uint32_t *sel) {
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    uint64_t prefix = 0, mask = 0;
    for (int shift = 56; shift >= 0; shift -= 8) {
        hist[tid] = 0;
        __syncthreads();
        for (int64_t i = tid; i < nl; i += 256) {
            const uint64_t k = l[i];
            if ((k & mask) == prefix) atomicAdd(&hist[(uint32_t)(k >> shift) & 255u], 1u);
        }
        __syncthreads();
        // the bin holding rank want: thread t owns bin t (as in ps_select_kernel)
        const uint32_t h = hist[tid];
        uint32_t incl = h;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const uint32_t v = __shfl_up(incl, o);
            if (lane >= o) incl += v;
        }
        if (lane == 63) wsum[wave] = incl;
        __syncthreads();
        for (int k = 0; k < wave; ++k) incl += wsum[k];
        const uint32_t excl = incl - h;
        if ((excl <= want && want < incl) || (tid == 255 && incl <= want)) {
            sel[0] = (uint32_t)tid;
            sel[1] = want - excl;
        }
        __syncthreads();
        prefix |= (uint64_t)sel[0] << shift;
        mask |= 255ull << shift;
        want = sel[1];
        __syncthreads();
    }
    return prefix;
}

// The rows of the C smallest keys of each query's list, in no particular order (the rescore
// that follows is re-sorted by exact distance): keys are distinct (the row is in the low 32
// bits), so exactly C keys lie at or below the C-th smallest -- the same set as the first C
// of the sorted list. One workgroup per query; the host guarantees count >= C (lists of
// exactly C are taken whole).
__global__ __launch_bounds__(256) void ps_take_smallest_kernel(const uint64_t *lists,
                                                               const uint32_t *count, int64_t cap,
                                                               int C, uint32_t *ids) {
    __shared__ uint32_t hist[256];
    __shared__ uint32_t wsum[4];
    __shared__ uint32_t sel[2];
    __shared__ uint32_t fill;
    const int q = blockIdx.x, tid = threadIdx.x;
    const uint64_t *l = lists + (int64_t)q * cap;
    const int64_t nl = min((int64_t)count[(int64_t)q * PS_COUNT_STRIDE], cap);
    uint32_t *out = ids + (int64_t)q * C;
    const uint64_t kth = nl > C ? block_kth_key(l, nl, (uint32_t)(C - 1), hist, wsum, sel) : ~0ull;
    if (tid == 0) fill = 0;
    __syncthreads();
    for (int64_t i = tid; i < nl; i += 256) {
        const uint64_t k = l[i];
        if (k <= kth) {
            const uint32_t pos = atomicAdd(&fill, 1u);
            if (pos < (uint32_t)C) out[pos] = (uint32_t)(k & 0xFFFFFFFFull);
        }
    }
}

// Final top k of each query's C rescored keys (orderable exact distance << 32 | row, distinct):
// the k smallest by radix select, then sorted in LDS (bitonic over the next power of two,
// padded with ~0) and written as (row, distance) -- the first k of the sorted keys. k <= 1024.
constexpr int kTopkMax = 1024;
__global__ __launch_bounds__(256) void ps_topk_kernel(const uint64_t *keys, int C, int k,
                                                      uint32_t *ids, float *dists) {
    __shared__ uint32_t hist[256];
    __shared__ uint32_t wsum[4];
    __shared__ uint32_t sel[2];
    __shared__ uint32_t fill;
    __shared__ uint64_t srt[kTopkMax];
    const int q = blockIdx.x, tid = threadIdx.x;
    const uint64_t *l = keys + (int64_t)q * C;
    const uint64_t kth = k < C ? block_kth_key(l, C, (uint32_t)(k - 1), hist, wsum, sel) : ~0ull;
    int P = 1;
    while (P < k) P <<= 1;
    for (int i = tid; i < P; i += 256) srt[i] = ~0ull;
    if (tid == 0) fill = 0;
    __syncthreads();
    for (int i = tid; i < C; i += 256) {
        const uint64_t key = l[i];
        if (key <= kth) {
            const uint32_t pos = atomicAdd(&fill, 1u);
            if (pos < (uint32_t)k) srt[pos] = key;
        }
    }
    __syncthreads();
    for (int size = 2; size <= P; size <<= 1) {
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            for (int i = tid; i < P / 2; i += 256) {
                const int lo = 2 * i - (i & (stride - 1)), hi = lo + stride;
                const bool up = (lo & size) == 0;
                const uint64_t a = srt[lo], b = srt[hi];
                if ((a > b) == up) {
                    srt[lo] = b;
                    srt[hi] = a;
                }
            }
            __syncthreads();
        }
    }
    for (int j = tid; j < k; j += 256) {
        const uint64_t key = srt[j];
        ids[(int64_t)q * k + j] = (uint32_t)(key & 0xFFFFFFFFull);
        dists[(int64_t)q * k + j] = from_orderable((uint32_t)(key >> 32));
    }
}

// (orderable exact distance << 32 | row) for the final selection.
__global__ __launch_bounds__(256) void ps_final_keys_kernel(const float *d, const uint32_t *ids,
                                                            int64_t total, uint64_t *keys) {
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e < total) keys[e] = ((uint64_t)orderable(d[e]) << 32) | ids[e];
}

__global__ __launch_bounds__(256) void ps_output_kernel(const uint64_t *sorted, int nq, int C,
                                                        int k, uint32_t *ids, float *dists) {
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= (int64_t)nq * k) return;
    const int64_t q = e / k, j = e - q * k;
    const uint64_t key = sorted[q * C + j];
    ids[e] = (uint32_t)(key & 0xFFFFFFFFull);
    dists[e] = from_orderable((uint32_t)(key >> 32));
}

static unsigned blocks(int64_t n, int64_t per) { return (unsigned)((n + per - 1) / per); }

template <class G>
static hipError_t dense_t(const uint16_t *xf, int64_t ns, const float *xn, const uint16_t *qf,
                          int nq, int nqp, bool l2, float inv_s, float *out, int64_t ld,
                          hipStream_t st) {
    const size_t lds = (size_t)2 * G::kBlocks * 1024;
    hipError_t e = hipFuncSetAttribute((const void *)ps_dense_kernel<G>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(ps_dense_kernel<G>, dim3(blocks(ns, 256)), dim3(256), lds, st, xf, ns, xn,
                       qf, nq, nqp, l2, inv_s, out, ld);
    return hipGetLastError();
}

template <class G>
static hipError_t filter_t(const uint16_t *xf, int64_t n, const float *xn, const uint16_t *qf,
                           int nq, int nqp, bool l2, float s, const float *tau,
                           const uint32_t *qmap, uint32_t *count, uint64_t *lists, int64_t cap,
                           hipStream_t st) {
    const size_t lds = (size_t)2 * G::kBlocks * 1024 + 4 * (size_t)nqp +
                       (size_t)4 * pool_alloc(G::kPipelined) * 10;
    hipError_t e = hipFuncSetAttribute((const void *)ps_filter_kernel<G>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(ps_filter_kernel<G>, dim3(blocks(n, 256)), dim3(256), lds, st, xf, n, xn,
                       qf, nq, nqp, l2, s, tau, qmap, count, lists, cap);
    return hipGetLastError();
}

// bf16: KS = dp / 16 k-steps of 16 dims; fp8: KB = dp / 32 blocks (k-steps of 64 dims).
#define PS_DISPATCH(FP8_, DP_, CALL)                   \
    if (FP8_) {                                        \
        switch ((DP_) / 32) {                          \
        case 4: return CALL(Fp8Rows<4>);               \
        case 8: return CALL(Fp8Rows<8>);               \
        case 12: return CALL(Fp8Rows<12>);             \
        case 16: return CALL(Fp8Rows<16>);             \
        case 24: return CALL(Fp8Rows<24>);             \
        default: return hipErrorInvalidValue;          \
        }                                              \
    }                                                  \
    switch ((DP_) / 16) {                              \
    case 4: return CALL(Bf16Rows<4>);                  \
    case 8: return CALL(Bf16Rows<8>);                  \
    case 16: return CALL(Bf16Rows<16>);                \
    case 24: return CALL(Bf16Rows<24>);                \
    case 32: return CALL(Bf16Rows<32>);                \
    case 48: return CALL(Bf16Rows<48>);                \
    default: return hipErrorInvalidValue;              \
    }

size_t ps_filter_lds(int dp, bool fp8, int nqp) {
    return (size_t)(fp8 ? 1 : 2) * dp * 64 + 4 * (size_t)nqp + (size_t)4 * pool_alloc(fp8) * 10;
}

bool ps_supported_dp(int dp) {
    const int ks = dp / 16;
    return dp % 16 == 0 && (ks == 4 || ks == 8 || ks == 16 || ks == 24 || ks == 32 || ks == 48);
}

int ps_fp8_dp(int dim) {
    for (int dp : {128, 256, 384, 512, 768})
        if (dp >= dim) return dp;
    return -1;
}

hipError_t ps_dense(int dp, bool fp8, const uint16_t *xf, int64_t ns, const float *xn,
                    const uint16_t *qf, int nq, int nqp, bool l2, float inv_s, float *out,
                    int64_t ld, hipStream_t st) {
#define CALL(G) dense_t<G>(xf, ns, xn, qf, nq, nqp, l2, inv_s, out, ld, st)
    PS_DISPATCH(fp8, dp, CALL)
#undef CALL
}

hipError_t ps_filter(int dp, bool fp8, const uint16_t *xf, int64_t n, const float *xn,
                     const uint16_t *qf, int nq, int nqp, bool l2, float s, const float *tau,
                     const uint32_t *qmap, uint32_t *count, uint64_t *lists, int64_t cap,
                     hipStream_t st) {
#define CALL(G) filter_t<G>(xf, n, xn, qf, nq, nqp, l2, s, tau, qmap, count, lists, cap, st)
    PS_DISPATCH(fp8, dp, CALL)
#undef CALL
}

// Rows (x, row stride `stride`) scaled by `scale` and rounded to E4M3 in the FP8 fragment
// layout of dp dims; rows >= `rows` up to rows_pad and dims >= dim are zero. One thread per
// 16-byte chunk.
__global__ __launch_bounds__(256) void ps_fp8_frag_kernel(const float *x, int64_t rows, int dim,
                                                          int64_t stride, int kb, int64_t chunks,
                                                          float scale, uint4 *out) {
    for (int64_t c = (int64_t)blockIdx.x * 256 + threadIdx.x; c < chunks;
         c += (int64_t)gridDim.x * 256) {
        const int64_t t = c / ((int64_t)kb * 64);
        const int bl = (int)((c / 64) % kb), l = (int)(c % 64);
        const int64_t row = 32 * t + (l & 31);
        const int d0 = 32 * bl + 16 * (l >> 5);
        uint32_t w[4] = {0, 0, 0, 0};
        if (row < rows) {
#pragma unroll
            for (int j = 0; j < 16; ++j) {
                const int d = d0 + j;
                const uint32_t b = d < dim ? to_e4m3(x[row * stride + d] * scale) : 0u;
                w[j >> 2] |= b << (8 * (j & 3));
            }
        }
        out[c] = make_uint4(w[0], w[1], w[2], w[3]);
    }
}

hipError_t ps_to_fp8_frag(const float *x, int64_t rows, int dim, int64_t stride, int dp,
                          int64_t rows_pad, float scale, uint8_t *out, hipStream_t st) {
    if (dp % 128 != 0 || rows_pad % 32 != 0 || rows > rows_pad) return hipErrorInvalidValue;
    const int64_t chunks = rows_pad * dp / 16;
    if (chunks == 0) return hipSuccess;
    const unsigned g = (unsigned)std::min<int64_t>((chunks + 255) / 256, 65536);
    hipLaunchKernelGGL(ps_fp8_frag_kernel, dim3(g), dim3(256), 0, st, x, rows, dim, stride, dp / 32,
                       chunks, scale, (uint4 *)out);
    return hipGetLastError();
}

// Largest |x| over rows x dim as f32 bits (non-negative floats order like their bits);
// *out must start at 0.
__global__ __launch_bounds__(256) void ps_absmax_kernel(const float *x, int64_t rows, int dim,
                                                        int64_t stride, uint32_t *out) {
    float m = 0.0f;
    const int64_t total = rows * dim;
    for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < total;
         e += (int64_t)gridDim.x * 256) {
        const int64_t r = e / dim;
        m = fmaxf(m, fabsf(x[r * stride + (e - r * dim)]));
    }
    for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o));
    if ((threadIdx.x & 63) == 0) atomicMax(out, __float_as_uint(m));
}

hipError_t ps_absmax(const float *x, int64_t rows, int dim, int64_t stride, uint32_t *out,
                     hipStream_t st) {
    hipError_t e = hipMemsetAsync(out, 0, sizeof(uint32_t), st);
    if (e != hipSuccess || rows == 0) return e;
    const unsigned g = (unsigned)std::min<int64_t>((rows * dim + 255) / 256, 4096);
    hipLaunchKernelGGL(ps_absmax_kernel, dim3(g), dim3(256), 0, st, x, rows, dim, stride, out);
    return hipGetLastError();
}

hipError_t ps_select(const float *dense, int64_t ld, int ns, int j0, int j1, int j2, float *tau,
                     int nq, hipStream_t st) {
    if (ns > kSampleMax) return hipErrorInvalidValue;
    hipLaunchKernelGGL(ps_select_kernel, dim3(nq), dim3(256), 0, st, dense, ld, ns, j0, j1, j2, tau,
                       nq);
    return hipGetLastError();
}

hipError_t ps_norms(const float *x, int64_t n, int dim, const int64_t *rows, float *out,
                    hipStream_t st) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(ps_norms_kernel, dim3(blocks(n, 256)), dim3(256), 0, st, x, n, dim, rows, out);
    return hipGetLastError();
}

hipError_t ps_gather_rows(const float *x, int dim, const int64_t *rows, int64_t ns, float *out,
                          hipStream_t st) {
    hipLaunchKernelGGL(ps_gather_rows_kernel, dim3(blocks(ns * dim, 256)), dim3(256), 0, st, x,
                       dim, rows, ns, out);
    return hipGetLastError();
}

hipError_t ps_sort_tmp_bytes(int64_t total, int segments, size_t *bytes) {
    *bytes = 0;
    return rocprim::segmented_radix_sort_keys(nullptr, *bytes, (uint64_t *)nullptr,
                                              (uint64_t *)nullptr, (unsigned)total,
                                              (unsigned)segments, (uint32_t *)nullptr,
                                              (uint32_t *)nullptr, 0, 64);
}

hipError_t ps_sort(void *tmp, size_t tmp_bytes, const uint64_t *in, uint64_t *out, int64_t total,
                   int segments, const uint32_t *begin, const uint32_t *end, hipStream_t st) {
    return rocprim::segmented_radix_sort_keys(tmp, tmp_bytes, in, out, (unsigned)total,
                                              (unsigned)segments, begin, end, 0, 64, st);
}

hipError_t ps_take_smallest(const uint64_t *lists, const uint32_t *count, int nq, int64_t cap,
                            int C, uint32_t *ids, hipStream_t st) {
    hipLaunchKernelGGL(ps_take_smallest_kernel, dim3(nq), dim3(256), 0, st, lists, count, cap, C,
                       ids);
    return hipGetLastError();
}

hipError_t ps_final_keys(const float *d, const uint32_t *ids, int64_t total, uint64_t *keys,
                         hipStream_t st) {
    hipLaunchKernelGGL(ps_final_keys_kernel, dim3(blocks(total, 256)), dim3(256), 0, st, d, ids,
                       total, keys);
    return hipGetLastError();
}

bool ps_topk_supported(int k) { return k >= 1 && k <= kTopkMax; }

hipError_t ps_topk(const uint64_t *keys, int nq, int C, int k, uint32_t *ids, float *dists,
                   hipStream_t st) {
    if (!ps_topk_supported(k) || k > C) return hipErrorInvalidValue;
    hipLaunchKernelGGL(ps_topk_kernel, dim3(nq), dim3(256), 0, st, keys, C, k, ids, dists);
    return hipGetLastError();
}

hipError_t ps_output(const uint64_t *sorted, int nq, int C, int k, uint32_t *ids, float *dists,
                     hipStream_t st) {
    hipLaunchKernelGGL(ps_output_kernel, dim3(blocks((int64_t)nq * k, 256)), dim3(256), 0, st,
                       sorted, nq, C, k, ids, dists);
    return hipGetLastError();
}

}  // namespace dkm

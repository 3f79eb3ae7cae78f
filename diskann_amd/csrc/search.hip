// Batched greedy beam search (Vamana) on gfx950: one wavefront per query.
//
// Restates DiskANNIndex::search_internal (diskann/src/graph/index.rs:1938-2005) with
// the in-memory FullAccessor / QuantAccessor expand_beam
// (diskann-providers/src/model/graph/provider/async_/inmem/full_precision.rs:274-322,
//  .../inmem/product.rs:304-335), the NeighborPriorityQueue insert rule
// (diskann/src/neighbor/queue.rs:130-171) and the Knn post-processing pipelines
// (diskann/src/graph/glue.rs:730-821, full_precision.rs:357-404).
//
// Per-query state lives in LDS:
//   query vector, PQ lookup table (M x 256 f32), candidate queue (capacity L + #start
//   points, visited flag in bit 31 of the id, merged in place), expansion list (ranked
//   in place), and an open-addressing visited set that spills to a per-query region of
//   a global (HBM) table when it fills. Keeping this small keeps more queries resident
//   per CU, which is what hides the latency of the gathers.
// Queue merge: the sequential "insert before equal distances, drop the tail" rule is
// the top-C of (queue U new) under the key (distance asc, insertion order desc), so a
// hop's expansion list is ranked in parallel and merged with two binary searches.
#include "dkm_device.h"
#include "dkm_search.h"
#include "dkm_sq.h"

namespace dkm {

struct Lds {
    float *q;        // dim
    float *lut;      // M*256 (PQ)
    float *qa_d;
    uint32_t *qa_i;
    float *nl_d, *sn_d;     // sn aliases nl unless the expansion list can exceed 256
    uint32_t *nl_i, *sn_i;
    uint32_t *hash;  // H
    int32_t *offs;   // M+1
    uint32_t *starts;
    uint32_t *beam;
    uint32_t *rows;  // beam_width x gstride adjacency rows staged per iteration; the
                     // merge reuses it for the new items' target positions
    int32_t *misc;
    uint8_t *sqc;    // scalar-quantized query code (sq_cstride bytes)
    uint32_t *sqm;   // [0] query compensation (f32 bits), [1] sum of squared query codes
    uint8_t *sphb;   // spherical: query bit row / 4 bit planes
    float *sphm;     // spherical: query metadata (4 floats)
};

// Expansion lists longer than this are ranked through a separate sorted buffer.
constexpr int kInPlaceRank = 4 * 64;

__host__ __device__ inline bool separate_sorted(const SearchParams &p) {
    return p.nlcap > kInPlaceRank;
}

__host__ __device__ inline int rows_entries(const SearchParams &p) {
    const int r = (int)(p.beam_width * p.gstride);
    return r > p.nlcap ? r : p.nlcap;
}

__host__ __device__ inline size_t align16(size_t x) { return (x + 15) & ~size_t(15); }

__host__ __device__ inline size_t lds_bytes(const SearchParams &p) {
    size_t b = 0;
    b += align16(sizeof(float) * p.dim);
    if (p.pq && p.adc_mode == 0) b += align16(sizeof(float) * p.M * 256);
    b += 2 * align16(sizeof(float) * p.qcap);                               // qa_d qa_i
    b += (separate_sorted(p) ? 4 : 2) * align16(sizeof(float) * p.nlcap);  // nl (sn)
    b += align16(sizeof(uint32_t) * p.hslots);
    if (p.pq) b += align16(sizeof(int32_t) * (p.M + 1));
    b += align16(sizeof(uint32_t) * p.nstart);
    b += align16(sizeof(uint32_t) * p.beam_width);
    b += align16(sizeof(uint32_t) * rows_entries(p));
    b += 16;
    if (p.sq) b += align16(p.sq_cstride) + 16;
    if (p.sph) b += align16(4 * p.sphq.cstride) + 16;  // the prepared query record
    return b;
}

__device__ inline Lds carve(const SearchParams &p, char *base) {
    Lds s;
    size_t o = 0;
    auto take = [&](size_t bytes) {
        char *r = base + o;
        o += align16(bytes);
        return r;
    };
    s.q = (float *)take(sizeof(float) * p.dim);
    s.lut = (p.pq && p.adc_mode == 0) ? (float *)take(sizeof(float) * p.M * 256) : nullptr;
    s.qa_d = (float *)take(sizeof(float) * p.qcap);
    s.qa_i = (uint32_t *)take(sizeof(uint32_t) * p.qcap);
    s.nl_d = (float *)take(sizeof(float) * p.nlcap);
    s.nl_i = (uint32_t *)take(sizeof(uint32_t) * p.nlcap);
    if (separate_sorted(p)) {
        s.sn_d = (float *)take(sizeof(float) * p.nlcap);
        s.sn_i = (uint32_t *)take(sizeof(uint32_t) * p.nlcap);
    } else {
        s.sn_d = s.nl_d;
        s.sn_i = s.nl_i;
    }
    s.hash = (uint32_t *)take(sizeof(uint32_t) * p.hslots);
    s.offs = p.pq ? (int32_t *)take(sizeof(int32_t) * (p.M + 1)) : nullptr;
    s.starts = (uint32_t *)take(sizeof(uint32_t) * p.nstart);
    s.beam = (uint32_t *)take(sizeof(uint32_t) * p.beam_width);
    s.rows = (uint32_t *)take(sizeof(uint32_t) * rows_entries(p));
    s.misc = (int32_t *)take(16);
    s.sqc = p.sq ? (uint8_t *)take(p.sq_cstride) : nullptr;
    s.sqm = p.sq ? (uint32_t *)take(16) : nullptr;
    s.sphb = p.sph ? (uint8_t *)take(4 * p.sphq.cstride) : nullptr;
    s.sphm = p.sph ? (float *)take(16) : nullptr;
    return s;
}

// Insert-or-find in this query's region of the HBM visited table (u32 ids, kEmpty when
// free, linear probing, cleared by the query before first use): one compare-and-swap per
// probe. Returns true if the id was not present.
__device__ inline bool hbm_visit(const SearchParams &p, uint32_t *gtab, uint32_t id, uint32_t *err) {
    const uint64_t GH = p.spill_slots;
    uint64_t g = hash_id(id ^ 0x9e3779b9u) & (GH - 1);
    for (uint64_t probe = 0; probe < GH; ++probe) {
        const uint32_t prev = atomicCAS(&gtab[g], kEmpty, id);
        if (prev == kEmpty) return true;
        if (prev == id) return false;
        g = (g + 1) & (GH - 1);
    }
    atomicOr(err, ERR_VISITED_OVERFLOW);
    return false;
}

// Visited-set insert for one id per lane (ids within one call are distinct). Returns
// true if the id was not present (and is now recorded).
// Until the LDS table passes its fill threshold it is the whole set: an array of 4-slot
// buckets probed in sequence from a hashed start; one 16-byte read checks a whole bucket,
// an id goes to the first free slot of the first bucket with room, and since slots are
// never freed a lookup may stop at the first bucket that still has a free slot. When it
// passes the threshold (decided per chunk, uniformly) its ids are copied to the query's
// HBM table (flush_visited), which from then on is the set; the LDS array becomes a
// 4-way cache of recently seen ids (home bucket only, random replacement): a hit there
// proves the id visited without an HBM round trip, a miss asks the HBM table.
__device__ inline bool visit_insert(const SearchParams &p, Lds &s, uint32_t id, bool active,
                                    bool spill, uint32_t *gtab, uint32_t *err) {
    if (!active) return false;
    const uint32_t HB = p.hslots >> 2;  // buckets
    uint32_t b = (uint32_t)(((uint64_t)hash_id(id) * HB) >> 32);
    if (spill) {
        const uint4 v = *(const uint4 *)&s.hash[4 * b];
        if (v.x == id || v.y == id || v.z == id || v.w == id) return false;
        const bool fresh = hbm_visit(p, gtab, id, err);
        s.hash[4 * b + (hash_id(id ^ 0x5bd1e995u) >> 30)] = id;
        return fresh;
    }
    for (uint32_t probe = 0; probe < HB;) {
        const uint4 v = *(const uint4 *)&s.hash[4 * b];
        if (v.x == id || v.y == id || v.z == id || v.w == id) return false;
        const int free_slot = v.x == kEmpty ? 0 : v.y == kEmpty ? 1 : v.z == kEmpty ? 2
                              : v.w == kEmpty ? 3 : -1;
        if (free_slot < 0) {
            b = b + 1 == HB ? 0 : b + 1;
            ++probe;
            continue;
        }
        const uint32_t old = atomicCAS(&s.hash[4 * b + free_slot], kEmpty, id);
        if (old == kEmpty) return true;
        if (old == id) return false;
        // another lane took the slot: re-read the same bucket
    }
    atomicOr(err, ERR_VISITED_OVERFLOW);
    return false;
}

// Clears this query's HBM region, then copies every id of the LDS table into it (once,
// when the LDS table fills). The wave waits for its clearing stores before the first
// compare-and-swap on the region.
__device__ inline void flush_visited(const SearchParams &p, Lds &s, uint32_t *gtab, uint32_t *err) {
    const int lane = lane_id();
    const uint4 empty = make_uint4(kEmpty, kEmpty, kEmpty, kEmpty);
    for (uint64_t i = 4 * (uint64_t)lane; i < p.spill_slots; i += 4 * kWave)
        *(uint4 *)&gtab[i] = empty;
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    for (uint32_t h = lane; h < p.hslots; h += kWave) {
        const uint32_t v = s.hash[h];
        if (v != kEmpty) (void)hbm_visit(p, gtab, v, err);
    }
}

__device__ inline float readlane_f(float v, int l) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}

__device__ inline uint32_t readlane_u(uint32_t v, int l) {
    return (uint32_t)__builtin_amdgcn_readlane((int)v, l);
}

// True if the LDS level proves `id` visited: the exact table (probe until a bucket with a
// free slot) before the HBM table takes over, the home bucket of the cache after.
__device__ inline bool lds_has(const SearchParams &p, const Lds &s, uint32_t id, bool spill) {
    const uint32_t HB = p.hslots >> 2;
    uint32_t b = (uint32_t)(((uint64_t)hash_id(id) * HB) >> 32);
    for (uint32_t probe = 0; probe < HB; ++probe) {
        const uint4 v = *(const uint4 *)&s.hash[4 * b];
        if (v.x == id || v.y == id || v.z == id || v.w == id) return true;
        if (spill || v.x == kEmpty || v.y == kEmpty || v.z == kEmpty || v.w == kEmpty) return false;
        b = b + 1 == HB ? 0 : b + 1;
    }
    return false;
}

// Visited-set step for up to E distinct ids per lane, none of which the LDS level knows
// (fresh[e] = the id was not visited; it is recorded now). With the HBM table every
// round issues the pending loads / compare-and-swaps of all E ids before using any
// result, so one hop costs about two memory round trips instead of two per row.
template <int E>
__device__ inline void visit_group(const SearchParams &p, Lds &s, const uint32_t (&id)[E],
                                   const bool (&act)[E], bool spill, uint32_t *gtab,
                                   uint32_t *err, bool (&fresh)[E]) {
    if (!spill) {
#pragma unroll
        for (int e = 0; e < E; ++e) fresh[e] = visit_insert(p, s, id[e], act[e], false, nullptr, err);
        return;
    }
    const uint64_t GH = p.spill_slots;
    uint64_t g[E];
    int st[E];  // 0 pending, 1 fresh, 2 present / inactive
#pragma unroll
    for (int e = 0; e < E; ++e) {
        g[e] = hash_id(id[e] ^ 0x9e3779b9u) & (GH - 1);
        st[e] = act[e] ? 0 : 2;
    }
    // round: a load per pending id; an id found empty is claimed with a compare-and-swap
    // in the same round (ids already visited cost one load)
    for (uint64_t round = 0; round < GH; ++round) {
        bool pending = false;
#pragma unroll
        for (int e = 0; e < E; ++e) pending = pending || st[e] == 0;
        if (!__builtin_amdgcn_ballot_w64(pending)) break;
        uint32_t cur[E];
#pragma unroll
        for (int e = 0; e < E; ++e)
            cur[e] = st[e] == 0 ? __hip_atomic_load(&gtab[g[e]], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                                : 0u;
        uint32_t prev[E];
#pragma unroll
        for (int e = 0; e < E; ++e)
            prev[e] = (st[e] == 0 && cur[e] == kEmpty) ? atomicCAS(&gtab[g[e]], kEmpty, id[e]) : cur[e];
#pragma unroll
        for (int e = 0; e < E; ++e) {
            if (st[e] != 0) continue;
            if (cur[e] == kEmpty && prev[e] == kEmpty) st[e] = 1;
            else if (prev[e] == id[e]) st[e] = 2;
            else if (prev[e] != kEmpty) g[e] = (g[e] + 1) & (GH - 1);
        }
    }
    const uint32_t HB = p.hslots >> 2;
#pragma unroll
    for (int e = 0; e < E; ++e) {
        if (st[e] == 0) atomicOr(err, ERR_VISITED_OVERFLOW);
        fresh[e] = st[e] == 1;
        if (act[e]) {  // remember it in the cache
            const uint32_t b = (uint32_t)(((uint64_t)hash_id(id[e]) * HB) >> 32);
            s.hash[4 * b + (hash_id(id[e] ^ 0x5bd1e995u) >> 30)] = id[e];
        }
    }
}

// Runs the n distinct ids c[0..n) through the visited set, E ids per lane at a time, and
// compacts the fresh ones to the front of c in order; returns their number.
template <int E>
__device__ inline int visit_chunk(const SearchParams &p, Lds &s, uint32_t *c, int n, bool spill,
                                  uint32_t *gtab, uint32_t *err) {
    const int lane = lane_id();
    int nn = 0;
    for (int t0 = 0; t0 < n; t0 += E * kWave) {
        uint32_t id[E];
        bool act[E], fresh[E];
#pragma unroll
        for (int e = 0; e < E; ++e) {
            const int t = t0 + lane + kWave * e;
            act[e] = t < n;
            id[e] = act[e] ? c[t] : 0u;
        }
        visit_group<E>(p, s, id, act, spill, gtab, err, fresh);
#pragma unroll
        for (int e = 0; e < E; ++e) {
            const uint64_t mk = ballot(fresh[e]);
            if (fresh[e]) c[nn + prefix_count(mk)] = id[e];
            nn += __builtin_popcountll(mk);
        }
    }
    return nn;
}

// One hop's visited filtering, row by row in beam order: a row's ids the LDS level
// proves visited are dropped, the rest go through the visited set together (one row's
// ids are distinct), and the fresh ones are appended to s.nl_i in adjacency order. A
// row sees everything the rows before it recorded, as sequential processing would.
// Returns the number of fresh ids.
__device__ inline int expand_hop(const SearchParams &p, Lds &s, int nbeam, bool spill,
                                 uint32_t *gtab, uint32_t *err, uint32_t &degsum) {
    const int lane = lane_id();
    uint32_t *c = s.nl_i;
    int nn = 0;
    for (int b = 0; b < nbeam; ++b) {
        const uint32_t *row = s.rows + b * p.gstride;
        const int cnt = (int)row[p.gstride - 1];
        degsum += (uint32_t)cnt;
        int m = 0;
        for (int j0 = 0; j0 < cnt; j0 += kWave) {
            const int j = j0 + lane;
            const uint32_t id = j < cnt ? row[j] : kEmpty;
            const bool keep = j < cnt && !lds_has(p, s, id, spill);
            const uint64_t mk = ballot(keep);
            if (keep) c[nn + m + prefix_count(mk)] = id;
            m += __builtin_popcountll(mk);
        }
        __syncthreads();
        if (m <= kWave) nn += visit_chunk<1>(p, s, c + nn, m, spill, gtab, err);
        else nn += visit_chunk<2>(p, s, c + nn, m, spill, gtab, err);
        __syncthreads();
    }
    return nn;
}

// Rank of each new item among the new items (key: d asc, insertion index desc), NaNs
// excluded; writes the sorted list sn_d / sn_i and returns the number of valid items.
// Up to kInPlaceRank items every lane holds its items in registers until all ranks are
// known, so the sorted list may overwrite the expansion list (sn aliases nl).
__device__ inline int rank_new(Lds &s, int nn) {
    const int lane = lane_id();
    if (nn <= kWave) {
        // one item per lane; the others' distances are broadcast with v_readlane
        const bool has = lane < nn;
        const float di = has ? s.nl_d[lane] : 0.0f;
        const uint32_t ii = has ? s.nl_i[lane] : 0u;
        const bool ok = has && !__builtin_isnan(di);
        int r = 0;
        for (int j = 0; j < nn; ++j) {
            const float dj = readlane_f(di, j);
            r += (dj < di) || (dj == di && j > lane);
        }
        if (ok) {
            s.sn_d[r] = di;
            s.sn_i[r] = ii;
        }
        return __builtin_popcountll(ballot(ok));
    }
    if (nn <= kInPlaceRank) {
        constexpr int E = kInPlaceRank / kWave;
        float d[E];
        uint32_t id[E];
        int r[E];
        bool ok[E];
        int nvalid = 0;
#pragma unroll
        for (int e = 0; e < E; ++e) {
            const int i = lane + kWave * e;
            const bool has = i < nn;
            d[e] = has ? s.nl_d[i] : 0.0f;
            id[e] = has ? s.nl_i[i] : 0u;
            ok[e] = has && !__builtin_isnan(d[e]);
            r[e] = 0;
            nvalid += __builtin_popcountll(ballot(ok[e]));
        }
        for (int j = 0; j < nn; ++j) {
            const float dj = s.nl_d[j];
#pragma unroll
            for (int e = 0; e < E; ++e) r[e] += (dj < d[e]) || (dj == d[e] && j > lane + kWave * e);
        }
        __syncthreads();
#pragma unroll
        for (int e = 0; e < E; ++e) {
            if (ok[e]) {
                s.sn_d[r[e]] = d[e];
                s.sn_i[r[e]] = id[e];
            }
        }
        return nvalid;
    }
    // longer lists (sn is a separate buffer then)
    int nvalid = 0;
    for (int base = 0; base < nn; base += kWave) {
        int i = base + lane;
        bool ok = i < nn && !__builtin_isnan(s.nl_d[i < nn ? i : 0]);
        nvalid += __builtin_popcountll(ballot(ok));
    }
    for (int base = 0; base < nn; base += kWave) {
        int i = base + lane;
        if (i < nn) {
            float di = s.nl_d[i];
            if (!__builtin_isnan(di)) {
                int r = 0;
#pragma unroll 8
                for (int j = 0; j < nn; ++j) {
                    float dj = s.nl_d[j];
                    r += (dj < di) || (dj == di && j > i);
                }
                s.sn_d[r] = di;
                s.sn_i[r] = s.nl_i[i];
            }
        }
    }
    return nvalid;
}

__device__ inline int wave_min_i(int v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = min(v, __shfl_xor(v, off));
    return v;
}

// Old queue items j = j0 + lane + 64 e (j >= j_first) move to j + #new with
// d_new <= d_old (the new distances broadcast lane by lane). Blocks go from the end of
// the queue down to j_first and every item of a block is read before any is written, so
// moving right never overwrites an item not yet read. Items before j_first do not move.
template <int E>
__device__ inline void merge_old(Lds &s, const int32_t *lo, int size, int nvalid, int cap, float *qd,
                                 uint32_t *qi, int cursor_old, int &best, int j_first) {
    const int lane = lane_id();
    constexpr int B = E * kWave;
    const int nblk = (size - j_first + B - 1) / B;
    for (int k = nblk - 1; k >= 0; --k) {
        const int j0 = j_first + k * B;
        float dj[E];
        uint32_t ij[E];
        int cnt[E];
#pragma unroll
        for (int e = 0; e < E; ++e) {
            const int j = j0 + lane + kWave * e;
            dj[e] = j < size ? qd[j] : 0.0f;
            ij[e] = j < size ? qi[j] : 0u;
            cnt[e] = 0;
        }
        if (nvalid <= 8) {
            // few new items: count them directly (d_new <= d_old)
            const float vn = lane < nvalid ? s.sn_d[lane] : 0.0f;
            for (int t = 0; t < nvalid; ++t) {
                const float v = readlane_f(vn, t);
#pragma unroll
                for (int e = 0; e < E; ++e) cnt[e] += v <= dj[e];
            }
        } else {
            // d_new[r] <= d_old[j]  <=>  lo[r] <= j, where lo[r] (non-decreasing in r) is
            // the number of old items smaller than new item r: binary search in lo
#pragma unroll
            for (int e = 0; e < E; ++e) {
                const int j = j0 + lane + kWave * e;
                int base = 0, len = nvalid;
                while (len > 0) {
                    const int half = len >> 1;
                    if (lo[base + half] <= j) {
                        base += half + 1;
                        len -= half + 1;
                    } else {
                        len = half;
                    }
                }
                cnt[e] = base;
            }
        }
#pragma unroll
        for (int e = 0; e < E; ++e) {
            const int j = j0 + lane + kWave * e;
            if (j < size) {
                const int pos = j + cnt[e];
                if (pos < cap) {
                    qd[pos] = dj[e];
                    qi[pos] = ij[e];
                    if (j == cursor_old) best = min(best, pos);
                }
            }
        }
    }
}

// Merge the sorted new items into the queue in place. New items go before old items of
// equal distance; positions >= cap are dropped. cursor_old is the first unvisited index
// of the old queue (or size); *cursor receives the first unvisited index of the merged
// queue (or its size). Returns the new size.
__device__ inline int merge_queue(Lds &s, int size, int nvalid, int cap, float *qd, uint32_t *qi,
                                  int cursor_old, int *cursor) {
    const int lane = lane_id();
    int best = 0x7FFFFFFF;
    // new items: position = rank + #old with d_old < d (lower bound in the old queue),
    // computed before the old items move
    int32_t *lo = (int32_t *)s.rows;
    for (int r = lane; r < nvalid; r += kWave) {
        float d = s.sn_d[r];
        int a = 0, b = size;
        while (a < b) {
            int mid = (a + b) >> 1;
            if (qd[mid] < d) a = mid + 1; else b = mid;
        }
        lo[r] = a;
        if (r == 0 && a < cap) best = a;
    }
    __syncthreads();
    // old items in front of the smallest new item keep their place
    const int first = nvalid > 0 ? lo[0] : size;
    const int j_first = (first < size ? first : size) & ~(kWave - 1);
    if (cursor_old < j_first) best = min(best, cursor_old);
    const int range = size - j_first;
    if (range > 0) {
        if (range <= kWave) merge_old<1>(s, lo, size, nvalid, cap, qd, qi, cursor_old, best, j_first);
        else if (range <= 2 * kWave) merge_old<2>(s, lo, size, nvalid, cap, qd, qi, cursor_old, best, j_first);
        else if (range <= 4 * kWave) merge_old<4>(s, lo, size, nvalid, cap, qd, qi, cursor_old, best, j_first);
        else merge_old<8>(s, lo, size, nvalid, cap, qd, qi, cursor_old, best, j_first);
    }
    __syncthreads();
    for (int r = lane; r < nvalid; r += kWave) {
        const int pos = lo[r] + r;
        if (pos < cap) {
            qd[pos] = s.sn_d[r];
            qi[pos] = s.sn_i[r];
        }
    }
    best = wave_min_i(best);
    __syncthreads();
    int ns = size + nvalid;
    ns = ns < cap ? ns : cap;
    *cursor = best < ns ? best : ns;
    return ns;
}

// Full-precision distances for n candidates (ids in LDS) into out (LDS): one half-wave per
// candidate; for DIM a multiple of 32 up to 128, eight candidates per half-wave have their
// rows in flight at once.
template <int METRIC, int DIM>
__device__ inline void fp_distances(const SearchParams &p, const float *q, const uint32_t *ids,
                                    float *out, int n, uint32_t fallback) {
    const int lane = lane_id(), half = lane >> 5, hp = lane & 31;
    if constexpr (METRIC != METRIC_COSINE && DIM > 0 && DIM % 32 == 0 && DIM <= 128) {
        constexpr int NI = DIM / 32;
        constexpr int CPB = 8;
        constexpr bool IP = METRIC == METRIC_IP || METRIC == METRIC_COSNORM;
        float qv[NI];
#pragma unroll
        for (int i = 0; i < NI; ++i) qv[i] = q[hp + 32 * i];
        for (int c0 = 0; c0 < n; c0 += 2 * CPB) {
            float xv[CPB][NI];
#pragma unroll
            for (int j = 0; j < CPB; ++j) {
                const int c = c0 + 2 * j + half;
                const uint32_t id = c < n ? ids[c] : fallback;
                const float *x = p.vectors + (size_t)id * p.vstride;
#pragma unroll
                for (int i = 0; i < NI; ++i) xv[j][i] = x[hp + 32 * i];
            }
#pragma unroll
            for (int j = 0; j < CPB; ++j) {
                float acc = 0.0f;
#pragma unroll
                for (int i = 0; i < NI; ++i) {
                    if (IP) {
                        acc = fma_(qv[i], xv[j][i], acc);
                    } else {
                        float c = qv[i] - xv[j][i];
                        acc = fma_(c, c, acc);
                    }
                }
                float d = finish32<IP>(acc, q, q, DIM, hp);  // DIM % 8 == 0: no remainder
                if (METRIC == METRIC_IP) d = -d;
                else if (METRIC == METRIC_COSNORM) d = 1.0f - d;
                const int c = c0 + 2 * j + half;
                if (c < n && hp == 0) out[c] = d;
            }
        }
    } else {
        for (int c0 = 0; c0 < n; c0 += 2) {
            const int c = c0 + half;
            const uint32_t id = c < n ? ids[c] : fallback;
            const float *x = p.vectors + (size_t)id * p.vstride;
            float d = half_distance<METRIC>(q, x, p.dim, hp);
            if (c < n && hp == 0) out[c] = d;
        }
    }
}

// ADC from the LDS lookup table: acc = sum_m LUT[m][code_m], sequential in m.
__device__ inline void adc_lut(const SearchParams &p, Lds &s, const uint32_t *ids, float *out,
                               int n) {
    const int lane = lane_id();
    const int M = p.M;
    for (int c = lane; c < n; c += kWave) {
        const uint8_t *row = p.codes + (size_t)ids[c] * p.code_stride;
        float acc = 0.0f;
        if ((M & 15) == 0) {
            for (int m0 = 0; m0 < M; m0 += 16) {
                uint4 v = *(const uint4 *)(row + m0);
                uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
                for (int b = 0; b < 16; ++b) {
                    uint32_t code = (w[b >> 2] >> (8 * (b & 3))) & 0xFFu;
                    acc += s.lut[(m0 + b) * 256 + code];
                }
            }
        } else {
            for (int m = 0; m < M; ++m) acc += s.lut[m * 256 + row[m]];
        }
        out[c] = acc;
    }
}

// One chunk term of the transposed pivot table (TransposedTable::process_into,
// diskann-quantization/src/product/tables/transposed/table.rs:285-313 with the kernels of
// pivots.rs:785-820,915-948): a fold over the chunk's dimensions from 0, L2
// acc = fma(p - q, p - q, acc), inner product acc = fma(-q, p, acc).
template <bool IP>
__device__ __forceinline__ float transposed_term(const float *q, const float *pv, int len) {
    float acc = 0.0f;
    for (int d = 0; d < len; ++d) {
        if (IP) {
            acc = __builtin_fmaf(-q[d], pv[d], acc);
        } else {
            const float t = pv[d] - q[d];
            acc = __builtin_fmaf(t, t, acc);
        }
    }
    return acc;
}

// Direct ADC without a table: every chunk term is recomputed from the pivots (held in
// L2) with exactly the arithmetic that fills the table, so the result is bit-identical to
// the table path while the query keeps no 32-48 KiB table in LDS. A group of G lanes
// shares one candidate: each lane computes M/G consecutive chunk terms, then the running
// sum is handed from lane to lane so the terms are added in chunk order. The pivots are
// read from the chunk-major copy: the lanes of one load instruction work on the same chunk
// (G = 1) or on G chunks, so they touch at most 16-32 cache lines of that chunk's
// 256 x LEN block rather than one line per lane; these scattered loads are what bounds
// the traversal's distance step.
template <typename T>
__device__ inline T load_as(const uint8_t *p) { return *(const T *)p; }

template <bool IP, int LEN, int M, int G, bool TL>
__device__ inline void adc_direct_group(const SearchParams &p, Lds &s, const uint32_t *ids,
                                        float *out, int n) {
    constexpr int MPER = M / G;
    uint32_t opaque = 0;
    asm volatile("" : "+v"(opaque));
    const int lane = lane_id() + (int)opaque;
    const int g = lane % G;
    for (int c0 = 0; c0 < n; c0 += kWave / G) {
        const int c = c0 + lane / G;
        const bool ok = c < n;
        const uint32_t id = ids[ok ? c : 0];
        const uint8_t *row = p.codes + (size_t)id * p.code_stride + g * MPER;
        // widest loads the row part's alignment allows (code rows are M bytes at stride M,
        // M a multiple of 16): each load instruction touches one line per lane
        uint8_t code[MPER];
        if constexpr (MPER % 16 == 0) {
#pragma unroll
            for (int q = 0; q < MPER; q += 16) {
                const uint4 v = load_as<uint4>(row + q);
                const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
                for (int b = 0; b < 16; ++b) code[q + b] = (uint8_t)(w[b >> 2] >> (8 * (b & 3)));
            }
        } else if constexpr (MPER % 8 == 0) {
#pragma unroll
            for (int q = 0; q < MPER; q += 8) {
                const uint2 v = load_as<uint2>(row + q);
                const uint32_t w[2] = {v.x, v.y};
#pragma unroll
                for (int b = 0; b < 8; ++b) code[q + b] = (uint8_t)(w[b >> 2] >> (8 * (b & 3)));
            }
        } else if constexpr (MPER % 4 == 0) {
#pragma unroll
            for (int q = 0; q < MPER; q += 4) {
                uint32_t w = load_as<uint32_t>(row + q);
#pragma unroll
                for (int b = 0; b < 4; ++b) code[q + b] = (uint8_t)(w >> (8 * b));
            }
        } else if constexpr (MPER % 2 == 0) {
#pragma unroll
            for (int q = 0; q < MPER; q += 2) {
                uint16_t w = load_as<uint16_t>(row + q);
                code[q] = (uint8_t)w;
                code[q + 1] = (uint8_t)(w >> 8);
            }
        } else {
#pragma unroll
            for (int q = 0; q < MPER; ++q) code[q] = row[q];
        }
        float ent[MPER];
#pragma unroll
        for (int q = 0; q < MPER; ++q) {
            const int m = g * MPER + q;
            const float *pv = p.pivots_t + ((size_t)m * 256 + code[q]) * LEN;
            float pc[LEN];
            if constexpr (LEN == 4) {
                float4 t4 = *(const float4 *)pv;
                pc[0] = t4.x; pc[1] = t4.y; pc[2] = t4.z; pc[3] = t4.w;
            } else if constexpr (LEN == 2) {
                float2 t2 = *(const float2 *)pv;
                pc[0] = t2.x; pc[1] = t2.y;
            } else {
#pragma unroll
                for (int l = 0; l < LEN; ++l) pc[l] = pv[l];
            }
            if (TL) {
                ent[q] = transposed_term<IP>(s.q + m * LEN, pc, LEN);
            } else {
                const float v = scalar_simd_4x1<IP>(s.q + m * LEN, pc, LEN);
                ent[q] = IP ? -v : v;
            }
        }
        float acc = 0.0f;
#pragma unroll
        for (int step = 0; step < G; ++step) {
            const float prev = G > 1 ? __shfl_up(acc, 1, G) : acc;
            if (g == step) {
                acc = step == 0 ? 0.0f : prev;
#pragma unroll
                for (int q = 0; q < MPER; ++q) acc += ent[q];
            }
        }
        if (ok && g == G - 1) out[c] = acc;
    }
}

template <bool IP, int LEN, bool TL>
__device__ inline void adc_direct(const SearchParams &p, Lds &s, const uint32_t *ids, float *out,
                                  int n) {
    if constexpr (LEN == 4) {
        if (p.M == 32) {
            if (n <= 8) return adc_direct_group<IP, 4, 32, 8, TL>(p, s, ids, out, n);
            if (n <= 16) return adc_direct_group<IP, 4, 32, 4, TL>(p, s, ids, out, n);
            if (n <= 32) return adc_direct_group<IP, 4, 32, 2, TL>(p, s, ids, out, n);
            return adc_direct_group<IP, 4, 32, 1, TL>(p, s, ids, out, n);
        }
    } else if constexpr (LEN == 2) {
        if (p.M == 48) {
            if (n <= 8) return adc_direct_group<IP, 2, 48, 8, TL>(p, s, ids, out, n);
            if (n <= 16) return adc_direct_group<IP, 2, 48, 4, TL>(p, s, ids, out, n);
            if (n <= 32) return adc_direct_group<IP, 2, 48, 2, TL>(p, s, ids, out, n);
            return adc_direct_group<IP, 2, 48, 1, TL>(p, s, ids, out, n);
        }
    }
    const int lane = lane_id();
    const int M = p.M;
    for (int c = lane; c < n; c += kWave) {
        const uint8_t *row = p.codes + (size_t)ids[c] * p.code_stride;
        float acc = 0.0f;
        for (int m = 0; m < M; ++m) {
            const int o0 = s.offs[m], o1 = s.offs[m + 1];
            const float *pv = p.pivots + (size_t)row[m] * p.dim + o0;
            if (TL) {
                acc += transposed_term<IP>(s.q + o0, pv, o1 - o0);
            } else {
                float v = scalar_simd_4x1<IP>(s.q + o0, pv, o1 - o0);
                acc += IP ? -v : v;
            }
        }
        out[c] = acc;
    }
}

// Scalar-quantized distances (SQStore's QueryComputer, diskann-providers/src/model/graph/
// provider/async_/inmem/scalar.rs:299-320): a group of G lanes per candidate, lane t of
// the group reads the 16-byte chunks t, t + G, ... of the candidate's code row; each lane
// keeps U candidates' rows in flight. The integer sums are exact in any order; the
// group's first lane applies the f32 finish.
template <int METRIC, int G>
__device__ inline void sq_distances(const SearchParams &p, Lds &s, const uint32_t *ids,
                                    float *out, int n) {
    constexpr int NG = kWave / G;
    constexpr int U = 4;
    const int lane = lane_id();
    const int g = lane % G, grp = lane / G;
    const int chunks = p.sq_cstride / 16;
    const float qcomp = __uint_as_float(s.sqm[0]);
    const uint32_t qq = s.sqm[1];
    for (int c0 = 0; c0 < n; c0 += NG * U) {
        uint32_t ab[U], bb[U], aa[U], id[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int c = c0 + grp + NG * u;
            id[u] = ids[c < n ? c : 0];
            ab[u] = bb[u] = aa[u] = 0;
        }
        for (int t = g; t < chunks; t += G) {
            const uint4 a = *(const uint4 *)(s.sqc + 16 * t);
            uint4 b[U];
#pragma unroll
            for (int u = 0; u < U; ++u)
                b[u] = *(const uint4 *)(p.sq_codes + (size_t)id[u] * p.sq_cstride + 16 * t);
#pragma unroll
            for (int u = 0; u < U; ++u) sq_dot16(a, b[u], ab[u], aa[u], bb[u]);
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
#pragma unroll
            for (int o = G / 2; o > 0; o >>= 1) {
                ab[u] += __shfl_xor(ab[u], o, G);
                bb[u] += __shfl_xor(bb[u], o, G);
            }
            const int c = c0 + grp + NG * u;
            if (c < n && g == 0) {
                const float comp = METRIC == METRIC_IP ? p.sq_comp[id[u]] : 0.0f;
                out[c] = sq_finish(METRIC, p.sq_bss, p.sq_ssn, ab[u], qq + bb[u] - 2u * ab[u],
                                   qcomp, comp);
            }
        }
    }
}

// Compress the query (SQStore::query_computer, scalar.rs:227-253, allow_rescale): for
// inner product the query is first scaled to the training mean norm; codes go to s.sqc
// (zero padded), the compensation fma chain and the code square norm to s.sqm.
template <int METRIC>
__device__ inline void sq_query(const SearchParams &p, Lds &s) {
    const int lane = lane_id();
    float rs = 1.0f;
    bool resc = false;
    if (METRIC == METRIC_IP) {
        const float v = half_distance<METRIC_IP>(s.q, s.q, p.dim, lane & 31);
        const float nrm = __builtin_sqrtf(-__shfl(v, 0));
        resc = nrm != 0.0f;
        if (resc) rs = p.sq_mean_norm / nrm;
    }
    const float inv = 255.0f / p.sq_scale;
    for (int e = lane; e < p.sq_cstride; e += kWave) {
        uint8_t code = 0;
        if (e < p.dim) {
            const float f = resc ? s.q[e] * rs : s.q[e];
            float c = (f - p.sq_shift[e]) * inv;
            if (c < 0.0f) c = 0.0f;
            else if (c > 255.0f) c = 255.0f;
            c = __builtin_roundf(c);
            code = c != c ? (uint8_t)0 : (uint8_t)c;
        }
        s.sqc[e] = code;
    }
    __syncthreads();
    uint32_t qq = 0;
    for (int t = lane; t < p.sq_cstride / 4; t += kWave) {
        const uint32_t w = ((const uint32_t *)s.sqc)[t];
        qq = __builtin_amdgcn_udot4(w, w, qq, false);
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) qq += __shfl_xor(qq, o);
    if (lane == 0) {
        float dot = 0.0f;
        if (METRIC == METRIC_IP)
            for (int e = 0; e < p.dim; ++e) dot = fma_((float)s.sqc[e], p.sq_shift[e], dot);
        s.sqm[0] = __float_as_uint(p.sq_scale * (1.0f / 255.0f) * dot);
        s.sqm[1] = qq;
    }
    __syncthreads();
}

// Spherical 1-bit distances (SphericalStore's query computer, diskann-providers/src/model/
// graph/provider/async_/inmem/spherical.rs:360-380): a group of G lanes per candidate,
// lane t of the group takes the 16-byte chunks t, t + G, ... of the candidate's bit row;
// the bit-level inner product is an exact popcount sum (same-as-data: q & b; 4-bit:
// sum_k 2^k popcount(plane_k & b)), finished with the reference's f32 formulas.
template <int LAYOUT, int G>
__device__ inline void sph_distances(const SearchParams &p, Lds &s, const uint32_t *ids,
                                     float *out, int n) {
    constexpr int NG = kWave / G;
    constexpr int U = 2;
    const SphParams &q = p.sphq;
    const int lane = lane_id();
    const int g = lane % G, grp = lane / G;
    const int chunks = q.cstride / 16;
    const float dim = (float)q.out_dim;
    for (int c0 = 0; c0 < n; c0 += NG * U) {
        uint32_t ip[U], id[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int c = c0 + grp + NG * u;
            id[u] = ids[c < n ? c : 0];
            ip[u] = 0;
        }
        for (int t = g; t < chunks; t += G) {
            uint4 b[U];
#pragma unroll
            for (int u = 0; u < U; ++u)
                b[u] = *(const uint4 *)(p.sph_bits + (size_t)id[u] * q.cstride + 16 * t);
            if constexpr (LAYOUT == SPH_QUERY_SAME_AS_DATA) {
                const uint4 a = *(const uint4 *)(s.sphb + 16 * t);
#pragma unroll
                for (int u = 0; u < U; ++u)
                    ip[u] += __builtin_popcount(a.x & b[u].x) + __builtin_popcount(a.y & b[u].y) +
                             __builtin_popcount(a.z & b[u].z) + __builtin_popcount(a.w & b[u].w);
            } else {
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    const uint4 a = *(const uint4 *)(s.sphb + k * q.cstride + 16 * t);
#pragma unroll
                    for (int u = 0; u < U; ++u)
                        ip[u] += (uint32_t)(__builtin_popcount(a.x & b[u].x) +
                                            __builtin_popcount(a.y & b[u].y) +
                                            __builtin_popcount(a.z & b[u].z) +
                                            __builtin_popcount(a.w & b[u].w)) << k;
                }
            }
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
#pragma unroll
            for (int o = G / 2; o > 0; o >>= 1) ip[u] += __shfl_xor(ip[u], o, G);
            const int c = c0 + grp + NG * u;
            if (c < n && g == 0) {
                const SphMeta y = p.sph_meta[id[u]];
                if constexpr (LAYOUT == SPH_QUERY_SAME_AS_DATA) {
                    const SphMeta x = {s.sphm[0], s.sphm[1], s.sphm[2], 0.0f};
                    out[c] = sph_data_data(q.metric, q.shift_sq, dim, ip[u], x, y);
                } else {
                    out[c] = sph_query4_data(q.metric, q.shift_sq, dim, ip[u], s.sphm, y);
                }
            }
        }
    }
}

// MODE 0: full precision; 1: PQ with the LDS table; 2: PQ direct (PQLEN = chunk width);
// 3: 8-bit scalar quantization (PQLEN = lanes per candidate); 4 / 5: 1-bit spherical with
// the same-as-data / 4-bit query (PQLEN = lanes per candidate); 6: PQ direct with the
// transposed table's arithmetic (disk-index search; MODE 1 builds either table).
template <int METRIC, int MODE, int DIM, int PQLEN>
__device__ inline void compute_distances(const SearchParams &p, Lds &s, int n) {
    if constexpr (MODE == 0) {
        fp_distances<METRIC, DIM>(p, s.q, s.nl_i, s.nl_d, n, s.starts[0]);
    } else if constexpr (MODE == 1) {
        adc_lut(p, s, s.nl_i, s.nl_d, n);
    } else if constexpr (MODE == 2) {
        adc_direct<METRIC == METRIC_IP, PQLEN, false>(p, s, s.nl_i, s.nl_d, n);
    } else if constexpr (MODE == 6) {
        adc_direct<METRIC == METRIC_IP, PQLEN, true>(p, s, s.nl_i, s.nl_d, n);
    } else if constexpr (MODE == 3) {
        sq_distances<METRIC, PQLEN>(p, s, s.nl_i, s.nl_d, n);
    } else if constexpr (MODE == 4) {
        sph_distances<SPH_QUERY_SAME_AS_DATA, PQLEN>(p, s, s.nl_i, s.nl_d, n);
    } else {
        sph_distances<SPH_QUERY_FOUR_BIT, PQLEN>(p, s, s.nl_i, s.nl_d, n);
    }
    __syncthreads();
}

// One query's search with the wave's LDS state; the batched kernel runs one per block
// (the serving lanes launch the same kernel on their batches).
struct QueryJob {
    const float *query;
    int L, k;
    uint32_t *err;
    uint32_t *spill;        // this query's HBM visited region
    uint32_t *rec_ids;      // optional visited record (rec_cap entries)
    float *rec_d;
    uint32_t *out_ids;      // k_stride entries
    float *out_d;
    uint32_t *out_count, *out_cmps, *out_hops, *rec_len, *out_degsum;
    const uint8_t *sph_rec;  // spherical: this query's prepared record
};

// RM: the metric of the FP32 rerank (the index metric; a template so that a kernel holds
// only the rerank arithmetic it runs).
template <int METRIC, int MODE, int DIM, int PQLEN, int RM>
__device__ inline void search_query(const SearchParams &p, Lds &s, const QueryJob &job) {
    constexpr bool PQ = MODE == 1 || MODE == 2 || MODE == 6;
    const int lane = threadIdx.x;
    const int L = job.L;
    const int k = job.k;
    const int cap = L + p.nstart;
    uint32_t *err = job.err;
    // per-query k / L come from device memory: a request outside Knn::new's rules or the
    // launch's LDS queue (sized for max_L) and output stride is refused, not searched
    if (L < 1 || k < 1 || k > L || cap > p.qcap || (int64_t)k > p.k_stride) {
        if (lane == 0) {
            *err = ERR_BAD_REQUEST;
            *job.out_count = 0u;
            if (job.out_cmps) *job.out_cmps = 0u;
            if (job.out_hops) *job.out_hops = 0u;
            if (job.rec_len) *job.rec_len = 0u;
            if (job.out_degsum) *job.out_degsum = 0u;
        }
        return;
    }

    // ---- init --------------------------------------------------------------------
    const float *qg = job.query;
    for (int e = lane; e < p.dim; e += kWave) s.q[e] = qg[e];
    for (uint32_t h = lane; h < p.hslots; h += kWave) s.hash[h] = kEmpty;
    for (int i = lane; i < p.nstart; i += kWave) s.starts[i] = p.starts[i];
    if (lane == 0) {
        s.misc[0] = -1;
        s.misc[1] = 0;
    }
    if (PQ) {
        for (int i = lane; i <= p.M; i += kWave) s.offs[i] = p.offsets[i];
    }
    __syncthreads();
    if constexpr (MODE == 3) sq_query<METRIC>(p, s);
    if constexpr (MODE == 4 || MODE == 5) {
        // the query record prepared by sph_query_prep_kernel (launch_search)
        const uint8_t *rec = job.sph_rec;
        const int nb = MODE == 4 ? p.sphq.cstride : 4 * p.sphq.cstride;
        for (int i = lane; i < nb / 16; i += kWave) ((uint4 *)s.sphb)[i] = ((const uint4 *)rec)[i];
        if (lane < 4) s.sphm[lane] = ((const float *)(rec + 4 * p.sphq.cstride))[lane];
        __syncthreads();
    }
    if (MODE == 1) {
        const bool ip = METRIC == METRIC_IP;
        for (int e = lane; e < p.M * 256; e += kWave) {
            int m = e >> 8, c = e & 255;
            int o0 = s.offs[m], o1 = s.offs[m + 1];
            const float *pv = p.pivots + (size_t)c * p.dim + o0;
            float v;
            if (p.tlut) {
                v = ip ? transposed_term<true>(s.q + o0, pv, o1 - o0)
                       : transposed_term<false>(s.q + o0, pv, o1 - o0);
            } else {
                v = ip ? -scalar_simd_4x1<true>(s.q + o0, pv, o1 - o0)
                       : scalar_simd_4x1<false>(s.q + o0, pv, o1 - o0);
            }
            s.lut[m * 256 + c] = v;
        }
        __syncthreads();
    }

    float *qd = s.qa_d;
    uint32_t *qi = s.qa_i;
    int size = 0;
    uint32_t cmps = 0, hops = 0, nrec = 0, degsum = 0, ios = 0;
    int lds_count = 0;
    const int spill_at = (int)(p.hslots - (p.hslots >> 3)) - (p.nlcap - kWave);
    uint32_t *gtab = nullptr;

    // start points: visited + queue insert in order
    for (int i = lane; i < p.nstart; i += kWave) s.nl_i[i] = s.starts[i];
    if (lane == 0) *err = 0;
    __syncthreads();
    for (int i = 0; i < p.nstart; ++i) {
        // sequential inserts (start ids may repeat only in malformed input)
        bool ins = visit_insert(p, s, s.starts[i], lane == 0, false, nullptr, err);
        (void)ins;
        __syncthreads();
    }
    lds_count = p.nstart;
    compute_distances<METRIC, MODE, DIM, PQLEN>(p, s, p.nstart);
    int cursor = 0;
    {
        int nvalid = rank_new(s, p.nstart);
        __syncthreads();
        size = merge_queue(s, size, nvalid, cap, qd, qi, 0, &cursor);
    }
    cmps = p.nstart;

    // Beam width 1: the next hop's beam node is known once this hop's new items are
    // ranked (the best new item or the old queue's next unvisited entry, whichever the
    // merge puts first), so its adjacency row is loaded into registers (lanes j and
    // j + 64 of a row of at most 128 slots) before the merge and used if the guess holds.
    const bool prefetch = p.beam_width == 1 && p.gstride <= 2 * kWave;
    uint32_t pre_id = kEmpty, pre0 = 0u, pre1 = 0u;

    // ---- main loop -----------------------------------------------------------------
    for (uint32_t iter = 0;; ++iter) {
        if (iter > p.max_iters) {
            if (lane == 0) atomicOr(err, ERR_ITER_LIMIT);
            break;
        }
        // beam: the first beam_width unvisited entries from the cursor (the first unvisited
        // index); the cursor then moves to the next unvisited entry after them
        int nbeam = 0;
        int next = size;
        for (int base = cursor; base < size; base += kWave) {
            const int i = base + lane;
            const bool unv = i < size && !(qi[i < size ? i : 0] & kVisitedBit);
            const uint64_t m = ballot(unv);
            const int before = prefix_count(m);
            const int take = (int)p.beam_width - nbeam;
            if (unv && before < take) {
                const uint32_t rid = qi[i];
                s.beam[nbeam + before] = rid;
                qi[i] = rid | kVisitedBit;
                if (job.rec_ids && nrec + before < p.rec_cap) {
                    job.rec_ids[nrec + before] = rid;
                    job.rec_d[nrec + before] = qd[i];
                }
            }
            const int got = __builtin_popcountll(m);
            if (got > take) {
                const uint64_t mn = ballot(unv && before == take);
                next = base + __builtin_ctzll(mn);
                nbeam += take;
                break;
            }
            nbeam += got;
        }
        __syncthreads();
        if (nbeam == 0) break;
        nrec += nbeam;
        // disk-index search: only the first (io_limit - loaded) beam nodes are loaded and
        // expanded; the rest stay taken (visited, counted as hops) without expansion
        int nload = nbeam;
        if (p.tlut) {
            const uint32_t left = p.io_limit - ios;
            nload = (uint32_t)nbeam < left ? nbeam : (int)left;
            ios += (uint32_t)nload;
        }

        // expand: stage the beam nodes' adjacency rows (ids + count slot) with one round
        // of loads, then filter the neighbours through the visited set
        if (prefetch && nload == 1 && s.beam[0] == pre_id) {
            if (lane < p.gstride) s.rows[lane] = pre0;
            if (lane + kWave < p.gstride) s.rows[lane + kWave] = pre1;
        } else {
            for (int b = 0; b < nload; ++b) {
                const uint32_t *row = p.graph + (size_t)s.beam[b] * p.gstride;
                for (int j = lane; j < p.gstride; j += kWave) s.rows[b * p.gstride + j] = row[j];
            }
        }
        __syncthreads();
        // the exact LDS table keeps room for one more hop's ids; past that the HBM table
        // takes over (decided per hop, uniformly)
        const bool spill = lds_count > spill_at;
        if (spill && gtab == nullptr) {
            gtab = job.spill;
            flush_visited(p, s, gtab, err);
        }
        const int nn = expand_hop(p, s, nload, spill, gtab, err, degsum);
        if (!spill) lds_count += nn;
        __syncthreads();
        compute_distances<METRIC, MODE, DIM, PQLEN>(p, s, nn);
        int nvalid = rank_new(s, nn);
        __syncthreads();
        if (prefetch) {
            // new items go before old ones of equal distance
            const bool have_old = next < size;
            const float d_old = have_old ? qd[next] : 0.0f;
            pre_id = nvalid > 0 && (!have_old || s.sn_d[0] <= d_old) ? s.sn_i[0]
                     : have_old ? qi[next] : kEmpty;
            if (pre_id != kEmpty) {
                const uint32_t *row = p.graph + (size_t)pre_id * p.gstride;
                pre0 = lane < p.gstride ? row[lane] : 0u;
                pre1 = lane + kWave < p.gstride ? row[lane + kWave] : 0u;
            }
        }
        size = merge_queue(s, size, nvalid, cap, qd, qi, next, &cursor);
        cmps += (uint32_t)nn;
        hops += (uint32_t)nbeam;
    }

    // ---- post-processing -----------------------------------------------------------
    uint32_t *oids = job.out_ids;
    float *ods = job.out_d;
    int count = 0;
    if (!p.rerank) {
        // FilterStartPoints -> CopyIds: first k non-start entries in queue order
        for (int base = 0; base < size && count < k; base += kWave) {
            int i = base + lane;
            bool ok = false;
            uint32_t id = 0;
            if (i < size) {
                id = qi[i] & ~kVisitedBit;
                ok = true;
                if (!p.keep_starts)
                    for (int t = 0; t < p.nstart; ++t) ok = ok && (s.starts[t] != id);
            }
            uint64_t m = ballot(ok);
            int pos = count + prefix_count(m);
            if (ok && pos < k) {
                oids[pos] = id;
                ods[pos] = qd[i];
            }
            int got = __builtin_popcountll(m);
            count = count + got < k ? count + got : k;
        }
    } else {
        // FilterStartPoints -> Rerank: exact FP32 distances, stable sort, first k.
        // The candidates are compacted to the front of the queue (in place: an entry
        // only moves left, and every entry of a chunk is read before any is written) and
        // their FP32 distances replace the queue distances.
        int n = 0;
        for (int base = 0; base < size; base += kWave) {
            int i = base + lane;
            bool ok = false;
            uint32_t id = 0;
            if (i < size) {
                id = qi[i] & ~kVisitedBit;
                ok = true;
                if (!p.keep_starts)
                    for (int t = 0; t < p.nstart; ++t) ok = ok && (s.starts[t] != id);
            }
            uint64_t m = ballot(ok);
            if (ok) qi[n + prefix_count(m)] = id;
            n += __builtin_popcountll(m);
        }
        __syncthreads();
        fp_distances<RM, RM == METRIC_COSINE ? 0 : DIM>(p, s.q, qi, qd, n, s.starts[0]);
        __syncthreads();
        // stable order of the (distance, position) keys: every candidate's rank is the
        // number of keys below its own (distances as order-preserving u32 with -0 == +0,
        // NaN above +inf), and the ranks below k are the output positions
        const int kk = n < k ? n : k;
        auto okey = [](float d) -> uint32_t {
            const uint32_t u = __float_as_uint(d == 0.0f ? 0.0f : d);
            return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
        };
        for (int b0 = 0; b0 < n; b0 += 4 * kWave) {
            uint32_t ku[4];
            int rk[4];
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int i = b0 + lane + kWave * e;
                ku[e] = i < n ? okey(qd[i]) : 0xFFFFFFFFu;
                rk[e] = 0;
            }
            for (int j = 0; j < n; ++j) {
                const uint32_t kj = okey(qd[j]);
#pragma unroll
                for (int e = 0; e < 4; ++e)
                    rk[e] += (kj < ku[e]) || (kj == ku[e] && j < b0 + lane + kWave * e);
            }
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int i = b0 + lane + kWave * e;
                if (i < n && rk[e] < kk) {
                    oids[rk[e]] = qi[i];
                    ods[rk[e]] = qd[i];
                }
            }
        }
        count = kk;
    }
    for (int i = count + lane; i < k; i += kWave) {
        oids[i] = kEmpty;
        ods[i] = __builtin_inff();
    }
    if (lane == 0) {
        *job.out_count = (uint32_t)count;
        if (job.out_cmps) *job.out_cmps = cmps;
        if (job.out_hops) *job.out_hops = hops;
        if (job.rec_len) *job.rec_len = nrec;
        if (job.out_degsum) *job.out_degsum = degsum;
    }
}

// Waves (queries) per SIMD the register budget is set for: five (at most 96 VGPRs) for
// direct-ADC PQ traversal over 2-dimensional chunks, whose hops wait on L2 pivot loads
// and gain from more resident queries (with 512 LDS visited slots the LDS holds 20 such
// queries per CU; 10M x 96, PQ 48: 6.7 -> 5.8 ms per 10k launch on a 2M proxy), four (128
// VGPRs) for the other modes (PQ 32 x 4 at L 512 was 3 % slower with five).
template <int MODE, int PQLEN>
constexpr int kSearchWavesPerSimd = MODE == 2 && PQLEN == 2 ? 5 : 4;

template <int METRIC, int MODE, int DIM, int PQLEN, int RM>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(kSearchWavesPerSimd<MODE, PQLEN>, 8))) void
beam_search_kernel(SearchParams p) {
    extern __shared__ __align__(16) char smem[];
    if ((int)blockIdx.x >= p.nq) return;
    const int qidx = p.order ? (int)p.order[blockIdx.x] : (int)blockIdx.x;
    Lds s = carve(p, smem);
    // the query's status bits gather in LDS and are stored once at the end (plain stores
    // only, so the outputs may live in host memory)
    __shared__ uint32_t status;
    QueryJob j;
    j.query = p.queries + (size_t)qidx * p.qstride;
    j.L = (int)p.L[qidx];
    j.k = (int)p.k[qidx];
    j.err = &status;
    j.spill = p.spill_table ? p.spill_table + (size_t)qidx * p.spill_slots : nullptr;
    j.rec_ids = p.rec_ids ? p.rec_ids + (size_t)qidx * p.rec_cap : nullptr;
    j.rec_d = p.rec_d ? p.rec_d + (size_t)qidx * p.rec_cap : nullptr;
    j.out_ids = p.out_ids + (size_t)qidx * p.k_stride;
    j.out_d = p.out_d + (size_t)qidx * p.k_stride;
    j.out_count = p.out_count + qidx;
    j.out_cmps = p.out_cmps ? p.out_cmps + qidx : nullptr;
    j.out_hops = p.out_hops ? p.out_hops + qidx : nullptr;
    j.rec_len = p.rec_len ? p.rec_len + qidx : nullptr;
    j.out_degsum = p.out_degsum ? p.out_degsum + qidx : nullptr;
    j.sph_rec = p.sph_qrec ? p.sph_qrec + (size_t)qidx * sph_query_record_bytes(p.sphq) : nullptr;
    search_query<METRIC, MODE, DIM, PQLEN, RM>(p, s, j);
    __syncthreads();
    if (threadIdx.x == 0) p.status[qidx] = status;
    if (p.done) {
        // every lane's output stores reach memory before the flag does
        __threadfence_system();
        __syncthreads();
        if (threadIdx.x == 0) *(volatile uint32_t *)&p.done[qidx] = 1u;
    }
}

// resident == nullptr: launch one block per query; otherwise report in *resident how many
// blocks of this configuration the current device holds at once.
template <int METRIC, int MODE, int DIM, int PQLEN, int RM>
static hipError_t launch_k(const SearchParams &p, hipStream_t stream, int *resident) {
    size_t lds = lds_bytes(p);
    auto kern = beam_search_kernel<METRIC, MODE, DIM, PQLEN, RM>;
    hipError_t e = hipFuncSetAttribute((const void *)kern,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    if (resident) {
        int per_cu = 0, dev = 0, cus = 0;
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void *)kern, 64, lds);
        if (e == hipSuccess) e = hipGetDevice(&dev);
        if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
        *resident = per_cu * cus;
        return e;
    }
    hipLaunchKernelGGL(kern, dim3((unsigned)p.nq), dim3(64), lds, stream, p);
    return hipGetLastError();
}

// Traversal with full precision or SQ codes reranks (if at all) with the traversal metric;
// PQ and spherical traversal rerank with the index metric. The rerank of a dimension-
// specialized variant is specialized for L2 / inner product only.
template <int METRIC, int MODE, int DIM, int PQLEN>
static hipError_t launch_t(const SearchParams &p, hipStream_t stream, int *resident) {
    if constexpr (MODE == 0 || MODE == 3) {
        return launch_k<METRIC, MODE, DIM, PQLEN, METRIC>(p, stream, resident);
    } else if constexpr (METRIC == METRIC_IP && MODE != 4 && MODE != 5) {
        return launch_k<METRIC, MODE, DIM, PQLEN, METRIC_IP>(p, stream, resident);
    } else {
        switch (p.metric) {
        case METRIC_L2: return launch_k<METRIC, MODE, DIM, PQLEN, METRIC_L2>(p, stream, resident);
        case METRIC_IP:
            if constexpr (DIM == 0) return launch_k<METRIC, MODE, 0, PQLEN, METRIC_IP>(p, stream, resident);
            else return launch_k<METRIC, MODE, DIM, PQLEN, METRIC_IP>(p, stream, resident);
        case METRIC_COSNORM: return launch_k<METRIC, MODE, 0, PQLEN, METRIC_COSNORM>(p, stream, resident);
        default: return launch_k<METRIC, MODE, 0, PQLEN, METRIC_COSINE>(p, stream, resident);
        }
    }
}

template <int METRIC>
static hipError_t dispatch_sq(const SearchParams &p, hipStream_t stream, int *resident) {
    const int chunks = p.sq_cstride / 16;
    if (chunks <= 8) return launch_t<METRIC, 3, 0, 8>(p, stream, resident);
    if (chunks <= 32) return launch_t<METRIC, 3, 0, 16>(p, stream, resident);
    return launch_t<METRIC, 3, 0, 64>(p, stream, resident);
}

template <int MODE>
static hipError_t dispatch_sph(const SearchParams &p, hipStream_t stream, int *resident) {
    const int chunks = p.sphq.cstride / 16;
    if (chunks <= 2) return launch_t<METRIC_L2, MODE, 0, 1>(p, stream, resident);
    if (chunks <= 8) return launch_t<METRIC_L2, MODE, 0, 4>(p, stream, resident);
    return launch_t<METRIC_L2, MODE, 0, 16>(p, stream, resident);
}

static hipError_t dispatch(const SearchParams &p, hipStream_t stream, int *resident) {
    if (p.sph) {
        // the metric is the quantizer's (run time); the FP32 rerank uses p.metric
        if (p.sph_layout == SPH_QUERY_SAME_AS_DATA) return dispatch_sph<4>(p, stream, resident);
        return dispatch_sph<5>(p, stream, resident);
    }
    if (p.sq) {
        if (p.metric == METRIC_IP) return dispatch_sq<METRIC_IP>(p, stream, resident);
        if (p.metric == METRIC_COSNORM) return dispatch_sq<METRIC_COSNORM>(p, stream, resident);
        return dispatch_sq<METRIC_L2>(p, stream, resident);
    }
    if (p.pq) {
        const bool ip = p.metric == METRIC_IP;  // L2 and CosineNormalized use the L2 table
        if (p.adc_mode == 0) {
            if (ip) return launch_t<METRIC_IP, 1, 0, 0>(p, stream, resident);
            if (p.dim == 128) return launch_t<METRIC_L2, 1, 128, 0>(p, stream, resident);
            if (p.dim == 96) return launch_t<METRIC_L2, 1, 96, 0>(p, stream, resident);
            return launch_t<METRIC_L2, 1, 0, 0>(p, stream, resident);
        }
        if (p.tlut) {
            // disk-index search with direct ADC (the transposed table's arithmetic)
            if (ip) return launch_t<METRIC_IP, 6, 0, 0>(p, stream, resident);
            if (p.dim == 128 && p.pq_len == 4) return launch_t<METRIC_L2, 6, 128, 4>(p, stream, resident);
            if (p.dim == 96 && p.pq_len == 2) return launch_t<METRIC_L2, 6, 96, 2>(p, stream, resident);
            return launch_t<METRIC_L2, 6, 0, 0>(p, stream, resident);
        }
        if (ip) return launch_t<METRIC_IP, 2, 0, 0>(p, stream, resident);
        if (p.dim == 128 && p.pq_len == 4) return launch_t<METRIC_L2, 2, 128, 4>(p, stream, resident);
        if (p.dim == 96 && p.pq_len == 2) return launch_t<METRIC_L2, 2, 96, 2>(p, stream, resident);
        return launch_t<METRIC_L2, 2, 0, 0>(p, stream, resident);
    }
    switch (p.metric) {
    case METRIC_L2:
        if (p.dim == 128) return launch_t<METRIC_L2, 0, 128, 0>(p, stream, resident);
        if (p.dim == 96) return launch_t<METRIC_L2, 0, 96, 0>(p, stream, resident);
        return launch_t<METRIC_L2, 0, 0, 0>(p, stream, resident);
    case METRIC_IP:
        if (p.dim == 128) return launch_t<METRIC_IP, 0, 128, 0>(p, stream, resident);
        return launch_t<METRIC_IP, 0, 0, 0>(p, stream, resident);
    case METRIC_COSNORM: return launch_t<METRIC_COSNORM, 0, 0, 0>(p, stream, resident);
    default: return launch_t<METRIC_COSINE, 0, 0, 0>(p, stream, resident);
    }
}

size_t search_lds_bytes(const SearchParams &p) { return lds_bytes(p); }

// Longest-L-first order of a launch's queries (a counting sort over L, ties in any order:
// each query's result does not depend on the order): blocks then start the longest
// searches first and the short ones fill in behind them, which shortens the launch's tail.
// order = nq entries followed by kOrderBins + 1 bins.
__global__ __launch_bounds__(256) void order_clear_kernel(uint32_t *bins) {
    for (int i = threadIdx.x + blockIdx.x * 256; i <= kOrderBins; i += gridDim.x * 256) bins[i] = 0;
}

// Wave-aggregated bin increment: the lanes holding the same key make one atomic between
// them (a batch of one L value would otherwise serialise on a single counter); returns the
// lane's slot = the bin's old value + the lane's rank among those lanes.
__device__ inline uint32_t order_slot(uint32_t *bins, uint32_t key, bool active) {
    const int lane = threadIdx.x & 63;
    uint32_t slot = 0;
    uint64_t pending = __ballot(active);
    while (pending) {
        const int leader = __ffsll((unsigned long long)pending) - 1;
        const uint32_t lk = __shfl(key, leader);
        const uint64_t same = __ballot(active && key == lk) & pending;
        uint32_t base = 0;
        if (lane == leader) base = atomicAdd(&bins[lk], (uint32_t)__popcll(same));
        base = __shfl(base, leader);
        if ((same >> lane) & 1ull)
            slot = base + (uint32_t)__popcll(same & ((1ull << lane) - 1ull));
        pending &= ~same;
    }
    return slot;
}

__global__ __launch_bounds__(256) void order_count_kernel(const uint32_t *L, int nq, uint32_t *bins) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    const uint32_t key = i < nq ? min(L[i], (uint32_t)(kOrderBins - 1)) : 0u;
    (void)order_slot(bins, key, i < nq);
}

// bins[l] becomes the first position of L = l in decreasing L order (one block);
// bins[kOrderBins] = 1 when every query has the same L (the launch then keeps its order).
__global__ __launch_bounds__(1024) void order_scan_kernel(uint32_t *bins, int nq) {
    __shared__ uint32_t part[1024];
    constexpr int per = (kOrderBins + 1023) / 1024;
    const int t = threadIdx.x;
    // thread t owns L values [kOrderBins - 1 - per (t + 1) + 1, kOrderBins - 1 - per t]
    uint32_t sum = 0;
    for (int j = 0; j < per; ++j) {
        const int l = kOrderBins - 1 - (per * t + j);
        if (l >= 0) {
            sum += bins[l];
            if (bins[l] == (uint32_t)nq) bins[kOrderBins] = 1u;  // one writer at most
        }
    }
    part[t] = sum;
    __syncthreads();
    for (int o = 1; o < 1024; o <<= 1) {
        const uint32_t v = t >= o ? part[t - o] : 0u;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    uint32_t run = part[t] - sum;  // exclusive prefix of the larger L values
    for (int j = 0; j < per; ++j) {
        const int l = kOrderBins - 1 - (per * t + j);
        if (l >= 0) {
            const uint32_t c = bins[l];
            bins[l] = run;
            run += c;
        }
    }
}

__global__ __launch_bounds__(256) void order_scatter_kernel(const uint32_t *L, int nq,
                                                            uint32_t *bins, uint32_t *order) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (bins[kOrderBins]) {  // a single L value: keep the launch's own order
        if (i < nq) order[i] = (uint32_t)i;
        return;
    }
    const uint32_t key = i < nq ? min(L[i], (uint32_t)(kOrderBins - 1)) : 0u;
    const uint32_t pos = order_slot(bins, key, i < nq);
    if (i < nq) order[pos] = (uint32_t)i;
}

hipError_t launch_search(const SearchParams &p, hipStream_t stream) {
    if (p.order) {
        uint32_t *bins = p.order + p.nq;
        const unsigned g = (unsigned)((p.nq + 255) / 256);
        hipLaunchKernelGGL(order_clear_kernel, dim3(17), dim3(256), 0, stream, bins);
        hipLaunchKernelGGL(order_count_kernel, dim3(g), dim3(256), 0, stream, p.L, p.nq, bins);
        hipLaunchKernelGGL(order_scan_kernel, dim3(1), dim3(1024), 0, stream, bins, p.nq);
        hipLaunchKernelGGL(order_scatter_kernel, dim3(g), dim3(256), 0, stream, p.L, p.nq, bins,
                           p.order);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    if (p.sph) {
        if (!p.sph_qrec) return hipErrorInvalidValue;
        hipError_t e = launch_sph_query_prep(p.sphq, p.sph_layout, p.sph_rescale, p.queries,
                                             p.qstride, p.nq,
                                             p.sph_qrec, stream);
        if (e != hipSuccess) return e;
    }
    return dispatch(p, stream, nullptr);
}

hipError_t search_resident_blocks(const SearchParams &p, int *blocks) {
    return dispatch(p, nullptr, blocks);
}

}  // namespace dkm

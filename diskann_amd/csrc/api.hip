// Host runtime behind the C ABI in include/diskann_mi355x.h.
//
// Owns the device-resident index (vectors, fixed-stride adjacency, PQ schema and
// codes), validates requests the way the reference does (Knn::new,
// diskann/src/graph/search/knn_search.rs:94-130; dimension checks of QueryComputer::new,
// diskann-providers/src/model/pq/distance/dynamic.rs:62-70) and launches the kernels on
// the handle's HIP stream. Calls on one handle are serialised by a mutex, so the handle
// is safe to share across threads (the reference's DiskANNIndex is Sync).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/diskann_mi355x.h"
#include "dkm_build.h"
#include "dkm_sq.h"
#include "dkm_sph.h"
#include "dkm_exact.h"
#include "dkm_kernels.h"
#include "dkm_kmeans.h"
#include "dkm_partition.h"
#include "dkm_prescore.h"
#include "dkm_search.h"

namespace {

thread_local std::string g_err;
int g_device = 0;

dkm_status fail(dkm_status s, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err = buf;
    return s;
}

#define HIP_TRY(expr)                                                                   \
    do {                                                                                \
        hipError_t e_ = (expr);                                                         \
        if (e_ != hipSuccess)                                                           \
            return fail(e_ == hipErrorOutOfMemory ? DKM_ERR_OUT_OF_MEMORY : DKM_ERR_DEVICE, \
                        "%s failed: %s", #expr, hipGetErrorString(e_));                 \
    } while (0)

int64_t round_up(int64_t x, int64_t m) { return (x + m - 1) / m * m; }

template <typename T>
struct DevBuf {
    T *p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    ~DevBuf() { release(); }
    hipError_t ensure(size_t count) {
        if (count <= n && p) return hipSuccess;
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
        hipError_t e = hipMalloc((void **)&p, sizeof(T) * std::max<size_t>(count, 1));
        if (e == hipSuccess) n = count;
        return e;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
};

}  // namespace

namespace dkm {
// Records the thread's error message for dkm_last_error (other source files of the library).
__attribute__((visibility("hidden"))) dkm_status set_error(dkm_status s, const char *msg) {
    g_err = msg;
    return s;
}
}  // namespace dkm

struct dkm_index {
    int device = 0;
    hipStream_t stream = nullptr;
    std::mutex mu;
    int metric = DKM_L2;
    int64_t dim = 0, vstride = 0;
    int64_t max_points = 0, num_start = 0, rows = 0;
    int64_t R = 0, R_max = 0, gstride = 0;
    float *d_vec = nullptr;
    uint32_t *d_graph = nullptr;
    uint32_t *d_starts = nullptr;
    bool starts_set = false;
    // PQ
    int64_t M = 0, K = 0;
    float *d_pivots = nullptr;
    float *d_pivots_t = nullptr;  // chunk-major copy (uniform 2- or 4-dim chunks): (m * 256 + c) * w
    int32_t *d_offsets = nullptr;
    uint8_t *d_codes = nullptr;
    bool codes_set = false;
    std::vector<int64_t> h_offsets;
    // visited-set spill
    uint32_t *d_spill = nullptr;
    int64_t spill_regions = 0;          // allocated regions (one per query of a launch)
    uint64_t spill_slots = 0;           // slots per region (power of two)
    int64_t tune_hslots = 0;
    uint64_t tune_spill = 0;            // HBM visited slots per query (0 = by max L)
    int tune_adc_mode = 1;              // PQ distances: 0 = LDS table, 1 = direct
    int pq_len = 0;                     // uniform PQ chunk width usable by the fast path
    int64_t build_max_fp = -1;          // Hybrid prune: full-precision ids per prune (-1 = all)
    bool build_sq = false;              // SQ Quantized build strategy (dkm_set_build_sq)
    bool build_sph = false;             // spherical build strategy (dkm_set_build_spherical)
    // 8-bit scalar quantization; when sq_active the quantized strategies traverse with it
    bool sq_active = false, sq_codes_set = false;
    float sq_scale = 0.0f, sq_mean_norm = 0.0f, sq_ssn = 0.0f;
    int64_t sq_cstride = 0;
    DevBuf<float> d_sq_shift, d_sq_comp;
    DevBuf<uint8_t> d_sq_codes;
    // 1-bit spherical quantization; when sph_active the quantized strategies traverse with it
    bool sph_active = false, sph_codes_set = false;
    int sph_layout = 1;
    dkm::SphParams sphq{};
    DevBuf<float> d_sph_shift;
    DevBuf<uint32_t> d_sph_signs0, d_sph_signs1, d_sph_sub;
    DevBuf<uint8_t> d_sph_bits;
    DevBuf<dkm::SphMeta> d_sph_meta;
    DevBuf<uint8_t> d_sph_qrec;          // prepared traversal queries of a launch
    DevBuf<uint32_t> d_order;            // longest-L-first order of a launch (+ bins)
    // staging for the host-buffer API
    DevBuf<float> s_q, s_d;
    DevBuf<uint32_t> s_k, s_L, s_ids, s_cnt, s_cmps, s_hops, s_status;
    DevBuf<uint32_t> d_disk_start;  // the disk-index search's start (medoid) row
    int64_t status_nq = 0;  // queries of the latest device-buffer launch (dkm_search_status)
    // profiling
    bool profile = false;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> events;
    int64_t prof_launches = 0;
    double prof_ms = 0.0;
};

extern "C" {

const char *dkm_last_error(void) { return g_err.c_str(); }
const char *dkm_version(void) { return "diskann_mi355x 0.1.0 (gfx950)"; }

dkm_status dkm_init(int device) {
    int n = 0;
    HIP_TRY(hipGetDeviceCount(&n));
    if (device < 0 || device >= n) return fail(DKM_ERR_INVALID_ARG, "device %d of %d", device, n);
    g_device = device;
    HIP_TRY(hipSetDevice(device));
    return DKM_OK;
}

dkm_status dkm_device_alloc(size_t bytes, void **ptr) {
    if (!ptr) return fail(DKM_ERR_INVALID_ARG, "null output pointer");
    HIP_TRY(hipSetDevice(g_device));
    HIP_TRY(hipMalloc(ptr, std::max<size_t>(bytes, 1)));
    return DKM_OK;
}
dkm_status dkm_device_free(void *ptr) {
    if (ptr) HIP_TRY(hipFree(ptr));
    return DKM_OK;
}
dkm_status dkm_memcpy_h2d(void *dst, const void *src, size_t bytes) {
    HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    return DKM_OK;
}
dkm_status dkm_device_copy(void *dst, const void *src, size_t bytes) {
    if ((!dst || !src) && bytes) return fail(DKM_ERR_INVALID_ARG, "null argument");
    if (bytes % 16 || ((uintptr_t)dst | (uintptr_t)src) % 16)
        return fail(DKM_ERR_INVALID_ARG, "copy needs 16-byte multiples and alignment");
    HIP_TRY(hipSetDevice(g_device));
    HIP_TRY(dkm::copy_device(dst, src, bytes, nullptr));
    HIP_TRY(hipDeviceSynchronize());
    return DKM_OK;
}

dkm_status dkm_memcpy_d2h(void *dst, const void *src, size_t bytes) {
    HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    return DKM_OK;
}
dkm_status dkm_synchronize(void) {
    HIP_TRY(hipDeviceSynchronize());
    return DKM_OK;
}

static bool metric_ok(int m) { return m >= 0 && m <= 3; }

dkm_status dkm_distance_batch(int metric, int64_t dim, const float *queries, int64_t nq,
                              const float *base, int64_t nbase, const uint32_t *ids, int64_t m,
                              float *out) {
    if (!metric_ok(metric)) return fail(DKM_ERR_INVALID_ARG, "unknown metric %d", metric);
    if (dim <= 0 || nq < 0 || nbase <= 0 || m < 0) return fail(DKM_ERR_INVALID_ARG, "bad sizes");
    if (nq == 0 || m == 0) return DKM_OK;
    for (int64_t i = 0; i < nq * m; ++i)
        if (ids[i] >= (uint64_t)nbase)
            return fail(DKM_ERR_INDEX, "id %u out of range (%lld rows)", ids[i], (long long)nbase);
    HIP_TRY(hipSetDevice(g_device));
    float *dq, *db, *dout;
    uint32_t *di;
    HIP_TRY(hipMalloc(&dq, sizeof(float) * nq * dim));
    HIP_TRY(hipMalloc(&db, sizeof(float) * nbase * dim));
    HIP_TRY(hipMalloc(&di, sizeof(uint32_t) * nq * m));
    HIP_TRY(hipMalloc(&dout, sizeof(float) * nq * m));
    HIP_TRY(hipMemcpy(dq, queries, sizeof(float) * nq * dim, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(db, base, sizeof(float) * nbase * dim, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(di, ids, sizeof(uint32_t) * nq * m, hipMemcpyHostToDevice));
    HIP_TRY(dkm::launch_distance_batch(metric, (int)dim, dq, nq, db, dim, di, m, dout, nullptr));
    HIP_TRY(hipMemcpy(out, dout, sizeof(float) * nq * m, hipMemcpyDeviceToHost));
    (void)hipFree(dq);
    (void)hipFree(db);
    (void)hipFree(di);
    (void)hipFree(dout);
    return DKM_OK;
}

dkm_status dkm_pq_chunk_offsets(int64_t dim, int64_t num_chunks, int64_t *offsets) {
    if (dim <= 0 || num_chunks <= 0 || num_chunks > dim)
        return fail(DKM_ERR_PQ, "cannot partition dim %lld into %lld chunks", (long long)dim,
                    (long long)num_chunks);
    int64_t off = 0;
    offsets[0] = 0;
    for (int64_t c = 0; c < num_chunks; ++c) {
        off += dim / num_chunks;
        if (c < dim % num_chunks) off += 1;
        offsets[c + 1] = off;
    }
    return DKM_OK;
}

static dkm_status check_offsets(const int64_t *offsets, int64_t M, int64_t dim) {
    if (M <= 0) return fail(DKM_ERR_PQ, "no PQ chunks");
    if (offsets[0] != 0) return fail(DKM_ERR_PQ, "chunk offsets must start at 0");
    for (int64_t i = 0; i < M; ++i)
        if (offsets[i + 1] <= offsets[i]) return fail(DKM_ERR_PQ, "chunk offsets must increase");
    if (offsets[M] != dim) return fail(DKM_ERR_PQ, "last chunk offset %lld != dim %lld",
                                       (long long)offsets[M], (long long)dim);
    return DKM_OK;
}

dkm_status dkm_pq_lut(int metric, const float *pivots, int64_t K, int64_t dim,
                      const int64_t *offsets, int64_t M, const float *queries, int64_t nq,
                      float *lut) {
    if (K != 256) return fail(DKM_ERR_PQ, "PQ tables need 256 centroids (got %lld)", (long long)K);
    if (metric == DKM_COSINE)
        return fail(DKM_ERR_PQ, "cosine PQ uses direct distances, not tables");
    dkm_status st = check_offsets(offsets, M, dim);
    if (st) return st;
    if (nq == 0) return DKM_OK;
    HIP_TRY(hipSetDevice(g_device));
    std::vector<int32_t> o32(offsets, offsets + M + 1);
    float *dp, *dq, *dl;
    int32_t *dof;
    HIP_TRY(hipMalloc(&dp, sizeof(float) * K * dim));
    HIP_TRY(hipMalloc(&dq, sizeof(float) * nq * dim));
    HIP_TRY(hipMalloc(&dl, sizeof(float) * nq * M * K));
    HIP_TRY(hipMalloc(&dof, sizeof(int32_t) * (M + 1)));
    HIP_TRY(hipMemcpy(dp, pivots, sizeof(float) * K * dim, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(dq, queries, sizeof(float) * nq * dim, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(dof, o32.data(), sizeof(int32_t) * (M + 1), hipMemcpyHostToDevice));
    HIP_TRY(dkm::launch_pq_lut(metric == DKM_INNER_PRODUCT, dp, (int)dim, dof, (int)M, dq, nq, dl,
                               nullptr));
    HIP_TRY(hipMemcpy(lut, dl, sizeof(float) * nq * M * K, hipMemcpyDeviceToHost));
    (void)hipFree(dp);
    (void)hipFree(dq);
    (void)hipFree(dl);
    (void)hipFree(dof);
    return DKM_OK;
}

dkm_status dkm_pq_distance_batch(int metric, const float *pivots, int64_t K, int64_t dim,
                                 const int64_t *offsets, int64_t M, const float *queries,
                                 const uint8_t *query_codes, int64_t nq, const uint8_t *codes,
                                 int64_t n, const uint32_t *ids, int64_t m, float *out) {
    if (!metric_ok(metric)) return fail(DKM_ERR_INVALID_ARG, "unknown metric %d", metric);
    if (K <= 0 || K > 256) return fail(DKM_ERR_PQ, "1..256 centroids supported (got %lld)", (long long)K);
    dkm_status st = check_offsets(offsets, M, dim);
    if (st) return st;
    if ((queries == nullptr) == (query_codes == nullptr))
        return fail(DKM_ERR_INVALID_ARG, "give exactly one of queries / query_codes");
    if (nq < 0 || n <= 0 || m < 0 || !pivots || !codes || !out || (nq * m > 0 && !ids))
        return fail(DKM_ERR_INVALID_ARG, "bad sizes");
    if (nq == 0 || m == 0) return DKM_OK;
    for (int64_t i = 0; i < nq * m; ++i)
        if (ids[i] >= (uint64_t)n)
            return fail(DKM_ERR_INDEX, "id %u out of range (%lld codes)", ids[i], (long long)n);
    for (int64_t i = 0; i < n * M; ++i)
        if (codes[i] >= K) return fail(DKM_ERR_PQ, "code %u >= %lld pivots", codes[i], (long long)K);
    if (query_codes)
        for (int64_t i = 0; i < nq * M; ++i)
            if (query_codes[i] >= K)
                return fail(DKM_ERR_PQ, "code %u >= %lld pivots", query_codes[i], (long long)K);
    HIP_TRY(hipSetDevice(g_device));
    std::vector<int32_t> o32(offsets, offsets + M + 1);
    DevBuf<float> dp, dq, dout;
    DevBuf<int32_t> dof;
    DevBuf<uint8_t> dc, dqc;
    DevBuf<uint32_t> di;
    HIP_TRY(dp.ensure(K * dim));
    HIP_TRY(dof.ensure(M + 1));
    HIP_TRY(dc.ensure(n * M));
    HIP_TRY(di.ensure(nq * m));
    HIP_TRY(dout.ensure(nq * m));
    HIP_TRY(hipMemcpy(dp.p, pivots, sizeof(float) * K * dim, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(dof.p, o32.data(), sizeof(int32_t) * (M + 1), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(dc.p, codes, (size_t)n * M, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(di.p, ids, sizeof(uint32_t) * nq * m, hipMemcpyHostToDevice));
    if (queries) {
        HIP_TRY(dq.ensure(nq * dim));
        HIP_TRY(hipMemcpy(dq.p, queries, sizeof(float) * nq * dim, hipMemcpyHostToDevice));
    } else {
        HIP_TRY(dqc.ensure(nq * M));
        HIP_TRY(hipMemcpy(dqc.p, query_codes, (size_t)nq * M, hipMemcpyHostToDevice));
    }
    HIP_TRY(dkm::launch_pq_direct(metric, dp.p, (int)dim, dof.p, (int)M, queries ? dq.p : nullptr,
                                  query_codes ? dqc.p : nullptr, nq, dc.p, di.p, m, dout.p,
                                  nullptr));
    HIP_TRY(hipMemcpy(out, dout.p, sizeof(float) * nq * m, hipMemcpyDeviceToHost));
    return DKM_OK;
}

dkm_status dkm_pq_encode(const float *pivots, int64_t K, int64_t dim, const int64_t *offsets,
                         int64_t M, const float *x, int64_t n, uint8_t *codes) {
    if (K <= 0 || K > 256) return fail(DKM_ERR_PQ, "1..256 centroids supported (got %lld)", (long long)K);
    dkm_status st = check_offsets(offsets, M, dim);
    if (st) return st;
    if (n == 0) return DKM_OK;
    HIP_TRY(hipSetDevice(g_device));
    std::vector<int32_t> o32(offsets, offsets + M + 1);
    float *dp, *dx;
    int32_t *dof;
    uint8_t *dc;
    uint32_t *dbad;
    HIP_TRY(hipMalloc(&dp, sizeof(float) * K * dim));
    HIP_TRY(hipMalloc(&dx, sizeof(float) * n * dim));
    HIP_TRY(hipMalloc(&dc, n * M));
    HIP_TRY(hipMalloc(&dof, sizeof(int32_t) * (M + 1)));
    HIP_TRY(hipMalloc(&dbad, sizeof(uint32_t)));
    HIP_TRY(hipMemcpy(dp, pivots, sizeof(float) * K * dim, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(dx, x, sizeof(float) * n * dim, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(dof, o32.data(), sizeof(int32_t) * (M + 1), hipMemcpyHostToDevice));
    uint32_t bad = 0xFFFFFFFFu;
    HIP_TRY(hipMemcpy(dbad, &bad, sizeof(uint32_t), hipMemcpyHostToDevice));
    HIP_TRY(dkm::launch_pq_encode(dp, (int)K, (int)dim, dof, (int)M, dx, dim, n, dc, M, dbad,
                                  nullptr));
    HIP_TRY(hipMemcpy(codes, dc, n * M, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(&bad, dbad, sizeof(uint32_t), hipMemcpyDeviceToHost));
    (void)hipFree(dp);
    (void)hipFree(dx);
    (void)hipFree(dc);
    (void)hipFree(dof);
    (void)hipFree(dbad);
    if (bad != 0xFFFFFFFFu)
        return fail(DKM_ERR_PQ, "infinite or NaN distance while encoding row %u", bad);
    return DKM_OK;
}

dkm_status dkm_pq_train_draws(const float *data, int64_t n, int64_t dim, int64_t M, int64_t K,
                              int64_t iters, const double *draws, float *pivots,
                              int64_t *offsets) {
    if (!data || !draws || !pivots || !offsets) return fail(DKM_ERR_INVALID_ARG, "null argument");
    if (n <= 0) return fail(DKM_ERR_PQ, "PQ training needs at least one row");
    if (K <= 0 || K > 256) return fail(DKM_ERR_PQ, "K must be in [1, 256] (got %lld)", (long long)K);
    if (iters < 0) return fail(DKM_ERR_INVALID_ARG, "negative Lloyd iteration count");
    dkm_status st = dkm_pq_chunk_offsets(dim, M, offsets);
    if (st) return st;
    int64_t maxds = 0;
    for (int64_t m = 0; m < M; ++m) maxds = std::max(maxds, offsets[m + 1] - offsets[m]);
    if (maxds > 64)
        return fail(DKM_ERR_PQ, "chunk width %lld above 64 is not supported", (long long)maxds);
    for (int64_t i = 0; i < M * K; ++i)
        if (!(draws[i] >= 0.0 && draws[i] < 1.0))
            return fail(DKM_ERR_INVALID_ARG, "draw %lld is not in [0, 1)", (long long)i);
    HIP_TRY(hipSetDevice(g_device));
    std::vector<int32_t> o32(offsets, offsets + M + 1);
    std::vector<uint32_t> cs(2 * (size_t)M, 0u);
    hipError_t e = dkm::pq_train_device(data, n, (int)dim, o32.data(), (int)M, (int)K, (int)iters,
                                        draws, pivots, cs.data());
    for (int64_t m = 0; m < M; ++m)
        if (cs[2 * m] == 3 || (e == hipErrorInvalidValue && cs[2 * m + 1] == 0))
            return fail(DKM_ERR_PQ, "pq training failed on chunk %lld of %lld: k-means++ %s",
                        (long long)m, (long long)M,
                        cs[2 * m] == 3 ? "saw an infinite or NaN value" : "picked no centre");
    if (e != hipSuccess) return fail(DKM_ERR_DEVICE, "PQ training: %s", hipGetErrorString(e));
    return DKM_OK;
}

dkm_status dkm_pq_train(const float *data, int64_t n, int64_t dim, int64_t M, int64_t K,
                        int64_t iters, uint64_t seed, float *pivots, int64_t *offsets) {
    if (M <= 0 || K <= 0 || K > 256) return fail(DKM_ERR_PQ, "bad PQ training sizes");
    std::vector<double> draws((size_t)(M * K));
    for (int64_t m = 0; m < M; ++m)
        for (int64_t c = 0; c < K; ++c) draws[(size_t)(m * K + c)] = dkm::pq_train_draw(seed, (int)m, (uint64_t)c);
    return dkm_pq_train_draws(data, n, dim, M, K, iters, draws.data(), pivots, offsets);
}

// ------------------------------------------------------------------------------------
// Index
// ------------------------------------------------------------------------------------

dkm_status dkm_index_create(int metric, int64_t dim, int64_t max_points, int64_t num_start,
                            int64_t R, int64_t R_max, dkm_index **out) {
    if (!out) return fail(DKM_ERR_INVALID_ARG, "null output");
    *out = nullptr;
    if (!metric_ok(metric)) return fail(DKM_ERR_INVALID_ARG, "unknown metric %d", metric);
    if (dim <= 0 || max_points <= 0 || num_start <= 0 || num_start > 64 || R <= 0 || R_max < R)
        return fail(DKM_ERR_INVALID_ARG, "bad index shape");
    if (max_points + num_start >= (int64_t)0x7FFFFFFF)
        return fail(DKM_ERR_INVALID_ARG, "too many points for 31-bit ids");
    HIP_TRY(hipSetDevice(g_device));
    dkm_index *ix = new dkm_index();
    ix->device = g_device;
    ix->metric = metric;
    ix->dim = dim;
    ix->vstride = round_up(dim, 16);
    ix->max_points = max_points;
    ix->num_start = num_start;
    ix->rows = max_points + num_start;
    ix->R = R;
    ix->R_max = R_max;
    ix->gstride = round_up(R_max + 1, 16);
    hipError_t e = hipStreamCreateWithFlags(&ix->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipMalloc(&ix->d_vec, sizeof(float) * ix->rows * ix->vstride);
    if (e == hipSuccess) e = hipMemset(ix->d_vec, 0, sizeof(float) * ix->rows * ix->vstride);
    if (e == hipSuccess) e = hipMalloc(&ix->d_graph, sizeof(uint32_t) * ix->rows * ix->gstride);
    if (e == hipSuccess) e = hipMemset(ix->d_graph, 0, sizeof(uint32_t) * ix->rows * ix->gstride);
    // the index's own stream does not wait for the null stream: finish the clears here
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e == hipSuccess) e = hipMalloc(&ix->d_starts, sizeof(uint32_t) * num_start);
    if (e == hipSuccess) {
        std::vector<uint32_t> s(num_start);
        for (int64_t i = 0; i < num_start; ++i) s[i] = (uint32_t)(max_points + i);
        e = hipMemcpy(ix->d_starts, s.data(), sizeof(uint32_t) * num_start, hipMemcpyHostToDevice);
    }
    if (e != hipSuccess) {
        dkm_index_destroy(ix);
        return fail(e == hipErrorOutOfMemory ? DKM_ERR_OUT_OF_MEMORY : DKM_ERR_DEVICE,
                    "index allocation: %s", hipGetErrorString(e));
    }
    *out = ix;
    return DKM_OK;
}

void dkm_index_destroy(dkm_index *ix) {
    if (!ix) return;
    (void)hipSetDevice(ix->device);
    if (ix->stream) (void)hipStreamSynchronize(ix->stream);
    for (auto &ev : ix->events) {
        (void)hipEventDestroy(ev.first);
        (void)hipEventDestroy(ev.second);
    }
    void *ptrs[] = {ix->d_vec, ix->d_graph, ix->d_starts, ix->d_pivots, ix->d_pivots_t,
                    ix->d_offsets, ix->d_codes, ix->d_spill};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    ix->s_q.release();
    ix->s_d.release();
    ix->s_k.release();
    ix->s_L.release();
    ix->s_ids.release();
    ix->s_cnt.release();
    ix->s_cmps.release();
    ix->s_hops.release();
    ix->s_status.release();
    if (ix->stream) (void)hipStreamDestroy(ix->stream);
    delete ix;
}

dkm_status dkm_index_shape(dkm_index *ix, int64_t *dim, int64_t *max_points, int64_t *num_start,
                           int64_t *R, int64_t *R_max) {
    if (!ix) return fail(DKM_ERR_INVALID_ARG, "null index");
    if (dim) *dim = ix->dim;
    if (max_points) *max_points = ix->max_points;
    if (num_start) *num_start = ix->num_start;
    if (R) *R = ix->R;
    if (R_max) *R_max = ix->R_max;
    return DKM_OK;
}

dkm_status dkm_set_vectors(dkm_index *ix, const float *x, int64_t row0, int64_t n) {
    if (!ix || (!x && n > 0)) return fail(DKM_ERR_INVALID_ARG, "null argument");
    if (row0 < 0 || n < 0 || row0 + n > ix->max_points)
        return fail(DKM_ERR_INDEX, "rows [%lld, %lld) outside [0, %lld)", (long long)row0,
                    (long long)(row0 + n), (long long)ix->max_points);
    std::lock_guard<std::mutex> g(ix->mu);
    HIP_TRY(hipSetDevice(ix->device));
    if (n == 0) return DKM_OK;
    HIP_TRY(hipMemcpy2D(ix->d_vec + row0 * ix->vstride, sizeof(float) * ix->vstride, x,
                        sizeof(float) * ix->dim, sizeof(float) * ix->dim, n, hipMemcpyHostToDevice));
    ix->codes_set = false;
    return DKM_OK;
}

dkm_status dkm_set_start_points(dkm_index *ix, const float *vectors) {
    if (!ix || !vectors) return fail(DKM_ERR_INVALID_ARG, "null argument");
    std::lock_guard<std::mutex> g(ix->mu);
    HIP_TRY(hipSetDevice(ix->device));
    HIP_TRY(hipMemcpy2D(ix->d_vec + ix->max_points * ix->vstride, sizeof(float) * ix->vstride,
                        vectors, sizeof(float) * ix->dim, sizeof(float) * ix->dim, ix->num_start,
                        hipMemcpyHostToDevice));
    ix->starts_set = true;
    ix->codes_set = false;
    return DKM_OK;
}

dkm_status dkm_set_graph(dkm_index *ix, const uint32_t *lists, const uint32_t *counts,
                         int64_t stride, int64_t row0, int64_t n) {
    if (!ix || (n > 0 && (!lists || !counts))) return fail(DKM_ERR_INVALID_ARG, "null argument");
    if (row0 < 0 || n < 0 || row0 + n > ix->rows)
        return fail(DKM_ERR_INDEX, "graph rows out of range");
    std::vector<uint32_t> buf(std::max<int64_t>(n, 1) * ix->gstride, 0);
    for (int64_t i = 0; i < n; ++i) {
        uint32_t c = counts[i];
        if (c > ix->R_max || c > stride)
            return fail(DKM_ERR_INDEX, "row %lld has %u neighbors (max degree %lld)",
                        (long long)(row0 + i), c, (long long)ix->R_max);
        uint32_t *dst = buf.data() + i * ix->gstride;
        for (uint32_t j = 0; j < c; ++j) {
            uint32_t id = lists[i * stride + j];
            if (id >= ix->rows)
                return fail(DKM_ERR_INDEX, "row %lld: neighbor id %u out of range",
                            (long long)(row0 + i), id);
            for (uint32_t t = 0; t < j; ++t)
                if (dst[t] == id)
                    return fail(DKM_ERR_INDEX, "row %lld: duplicate neighbor %u",
                                (long long)(row0 + i), id);
            dst[j] = id;
        }
        dst[ix->gstride - 1] = c;
    }
    std::lock_guard<std::mutex> g(ix->mu);
    HIP_TRY(hipSetDevice(ix->device));
    if (n > 0)
        HIP_TRY(hipMemcpy(ix->d_graph + row0 * ix->gstride, buf.data(),
                          sizeof(uint32_t) * n * ix->gstride, hipMemcpyHostToDevice));
    return DKM_OK;
}

dkm_status dkm_get_graph(dkm_index *ix, uint32_t *lists, uint32_t *counts, int64_t stride,
                         int64_t row0, int64_t n) {
    if (!ix) return fail(DKM_ERR_INVALID_ARG, "null index");
    if (row0 < 0 || n < 0 || row0 + n > ix->rows) return fail(DKM_ERR_INDEX, "rows out of range");
    if (n == 0) return DKM_OK;
    std::vector<uint32_t> buf(n * ix->gstride);
    {
        std::lock_guard<std::mutex> g(ix->mu);
        HIP_TRY(hipSetDevice(ix->device));
        HIP_TRY(hipStreamSynchronize(ix->stream));
        HIP_TRY(hipMemcpy(buf.data(), ix->d_graph + row0 * ix->gstride,
                          sizeof(uint32_t) * n * ix->gstride, hipMemcpyDeviceToHost));
    }
    for (int64_t i = 0; i < n; ++i) {
        uint32_t c = buf[i * ix->gstride + ix->gstride - 1];
        counts[i] = c;
        for (uint32_t j = 0; j < c && j < stride; ++j) lists[i * stride + j] = buf[i * ix->gstride + j];
    }
    return DKM_OK;
}

dkm_status dkm_set_pq(dkm_index *ix, const float *pivots, int64_t K, const int64_t *offsets,
                      int64_t M) {
    if (!ix || !pivots || !offsets) return fail(DKM_ERR_INVALID_ARG, "null argument");
    if (K != 256) return fail(DKM_ERR_PQ, "the device tables need 256 centroids (got %lld)", (long long)K);
    dkm_status st = check_offsets(offsets, M, ix->dim);
    if (st) return st;
    std::lock_guard<std::mutex> g(ix->mu);
    HIP_TRY(hipSetDevice(ix->device));
    if (ix->d_pivots) (void)hipFree(ix->d_pivots);
    if (ix->d_pivots_t) (void)hipFree(ix->d_pivots_t);
    if (ix->d_offsets) (void)hipFree(ix->d_offsets);
    if (ix->d_codes) (void)hipFree(ix->d_codes);
    ix->d_pivots = nullptr;
    ix->d_pivots_t = nullptr;
    ix->d_offsets = nullptr;
    ix->d_codes = nullptr;
    ix->M = M;
    ix->K = K;
    ix->h_offsets.assign(offsets, offsets + M + 1);
    ix->pq_len = 0;
    if (M % 16 == 0) {
        int64_t w = offsets[1] - offsets[0];
        bool uniform = true;
        for (int64_t i = 0; i < M; ++i) uniform = uniform && (offsets[i + 1] - offsets[i] == w);
        if (uniform && (w == 2 || w == 4)) ix->pq_len = (int)w;
    }
    std::vector<int32_t> o32(offsets, offsets + M + 1);
    HIP_TRY(hipMalloc(&ix->d_pivots, sizeof(float) * K * ix->dim));
    HIP_TRY(hipMalloc(&ix->d_offsets, sizeof(int32_t) * (M + 1)));
    HIP_TRY(hipMalloc(&ix->d_codes, ix->rows * M + 16));
    HIP_TRY(hipMemsetAsync(ix->d_codes, 0, ix->rows * M + 16, ix->stream));
    HIP_TRY(hipStreamSynchronize(ix->stream));  // cleared before any copy on another stream
    HIP_TRY(hipMemcpy(ix->d_pivots, pivots, sizeof(float) * K * ix->dim, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(ix->d_offsets, o32.data(), sizeof(int32_t) * (M + 1), hipMemcpyHostToDevice));
    if (ix->pq_len) {
        // chunk-major copy for the traversal's direct ADC: the lanes of one load read one
        // chunk's 256 x w block instead of 256 rows of the pivot table
        const int64_t w = ix->pq_len;
        std::vector<float> t((size_t)K * ix->dim);
        for (int64_t m = 0; m < M; ++m)
            for (int64_t c = 0; c < K; ++c)
                for (int64_t d = 0; d < w; ++d)
                    t[(size_t)(m * K + c) * w + d] = pivots[c * ix->dim + m * w + d];
        HIP_TRY(hipMalloc(&ix->d_pivots_t, sizeof(float) * K * ix->dim));
        HIP_TRY(hipMemcpy(ix->d_pivots_t, t.data(), sizeof(float) * K * ix->dim, hipMemcpyHostToDevice));
    }
    ix->codes_set = false;
    return DKM_OK;
}

dkm_status dkm_set_codes(dkm_index *ix, const uint8_t *codes, int64_t row0, int64_t n) {
    if (!ix || (!codes && n > 0)) return fail(DKM_ERR_INVALID_ARG, "null argument");
    if (!ix->d_codes) return fail(DKM_ERR_NOT_READY, "set the PQ schema first");
    if (row0 < 0 || n < 0 || row0 + n > ix->rows) return fail(DKM_ERR_INDEX, "rows out of range");
    std::lock_guard<std::mutex> g(ix->mu);
    HIP_TRY(hipSetDevice(ix->device));
    if (n) HIP_TRY(hipMemcpy(ix->d_codes + row0 * ix->M, codes, n * ix->M, hipMemcpyHostToDevice));
    ix->codes_set = true;
    return DKM_OK;
}

dkm_status dkm_encode_index(dkm_index *ix) {
    if (!ix) return fail(DKM_ERR_INVALID_ARG, "null index");
    if (!ix->d_codes) return fail(DKM_ERR_NOT_READY, "set the PQ schema first");
    std::lock_guard<std::mutex> g(ix->mu);
    HIP_TRY(hipSetDevice(ix->device));
    uint32_t *dbad;
    HIP_TRY(hipMalloc(&dbad, sizeof(uint32_t)));
    uint32_t bad = 0xFFFFFFFFu;
    HIP_TRY(hipMemcpy(dbad, &bad, sizeof(uint32_t), hipMemcpyHostToDevice));
    HIP_TRY(dkm::launch_pq_encode(ix->d_pivots, (int)ix->K, (int)ix->dim, ix->d_offsets, (int)ix->M,
                                  ix->d_vec, ix->vstride, ix->rows, ix->d_codes, ix->M, dbad,
                                  ix->stream));
    HIP_TRY(hipStreamSynchronize(ix->stream));
    HIP_TRY(hipMemcpy(&bad, dbad, sizeof(uint32_t), hipMemcpyDeviceToHost));
    (void)hipFree(dbad);
    if (bad != 0xFFFFFFFFu)
        return fail(DKM_ERR_PQ, "infinite or NaN distance while encoding row %u", bad);
    ix->codes_set = true;
    return DKM_OK;
}

dkm_status dkm_get_codes(dkm_index *ix, uint8_t *codes, int64_t row0, int64_t n) {
    if (!ix || !ix->d_codes) return fail(DKM_ERR_NOT_READY, "no PQ codes");
    if (row0 < 0 || n < 0 || row0 + n > ix->rows) return fail(DKM_ERR_INDEX, "rows out of range");
    std::lock_guard<std::mutex> g(ix->mu);
    HIP_TRY(hipSetDevice(ix->device));
    HIP_TRY(hipStreamSynchronize(ix->stream));
    if (n) HIP_TRY(hipMemcpy(codes, ix->d_codes + row0 * ix->M, n * ix->M, hipMemcpyDeviceToHost));
    return DKM_OK;
}


// ------------------------------------------------------------------------------------
// 8-bit scalar quantization (diskann-quantization/src/scalar)
// ------------------------------------------------------------------------------------

}  // extern "C"

// Device copy of a canonical code buffer (rows of [f32 compensation][dim codes]) into
// padded code rows and a compensation array.
static dkm_status sq_upload(const uint8_t *canon, int64_t n, int64_t dim, DevBuf<uint8_t> &codes,
                            DevBuf<float> &comp) {
    const int64_t cs = dkm::sq_code_stride(dim);
    std::vector<uint8_t> rows((size_t)(n * cs), 0);
    std::vector<float> cp((size_t)n);
    for (int64_t r = 0; r < n; ++r) {
        const uint8_t *src = canon + r * (dim + 4);
        std::memcpy(&cp[(size_t)r], src, sizeof(float));
        std::memcpy(&rows[(size_t)(r * cs)], src + 4, (size_t)dim);
    }
    HIP_TRY(codes.ensure(n * cs));
    HIP_TRY(comp.ensure(n));
    HIP_TRY(hipMemcpy(codes.p, rows.data(), rows.size(), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(comp.p, cp.data(), sizeof(float) * n, hipMemcpyHostToDevice));
    return DKM_OK;
}

static dkm_status sq_download(const uint8_t *d_codes, const float *d_comp, int64_t n, int64_t dim,
                              uint8_t *canon) {
    const int64_t cs = dkm::sq_code_stride(dim);
    std::vector<uint8_t> rows((size_t)(n * cs));
    std::vector<float> cp((size_t)n);
    HIP_TRY(hipMemcpy(rows.data(), d_codes, rows.size(), hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(cp.data(), d_comp, sizeof(float) * n, hipMemcpyDeviceToHost));
    for (int64_t r = 0; r < n; ++r) {
        uint8_t *dst = canon + r * (dim + 4);
        std::memcpy(dst, &cp[(size_t)r], sizeof(float));
        std::memcpy(dst + 4, &rows[(size_t)(r * cs)], (size_t)dim);
    }
    return DKM_OK;
}

static dkm_status sq_check(float scale, const float *shift, int64_t dim) {
    if (!shift || dim <= 0) return fail(DKM_ERR_INVALID_ARG, "bad scalar quantizer");
    if (!(scale > 0.0f) || !std::isfinite(scale))
        return fail(DKM_ERR_INVALID_ARG, "scale must be positive and finite");
    return DKM_OK;
}

extern "C" {

dkm_status dkm_sq_train(const float *data, int64_t n, int64_t dim, double standard_deviations,
                        float *out_scale, float *out_shift, float *out_mean_norm) {
    if (!data || !out_scale || !out_shift || !out_mean_norm) return fail(DKM_ERR_INVALID_ARG, "null argument");
    if (n <= 0 || dim <= 0) return fail(DKM_ERR_INVALID_ARG, "training data cannot be empty");
    if (!(standard_deviations > 0.0)) return fail(DKM_ERR_INVALID_ARG, "standard deviations must be positive");
    HIP_TRY(hipSetDevice(g_device));
    DevBuf<float> dx, dsh, dsc;
    DevBuf<double> dm, dv, dn;
    HIP_TRY(dx.ensure(n * dim));
    HIP_TRY(dsh.ensure(dim));
    HIP_TRY(dsc.ensure(2));
    HIP_TRY(dm.ensure(dim));
    HIP_TRY(dv.ensure(dim));
    HIP_TRY(dn.ensure(n));
    HIP_TRY(hipMemcpy(dx.p, data, sizeof(float) * n * dim, hipMemcpyHostToDevice));
    HIP_TRY(dkm::launch_sq_train(dx.p, n, (int)dim, dim, standard_deviations, dm.p, dv.p, dn.p,
                                 dsc.p, dsh.p, dsc.p + 1, nullptr));
    float sc[2];
    HIP_TRY(hipMemcpy(sc, dsc.p, sizeof(sc), hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(out_shift, dsh.p, sizeof(float) * dim, hipMemcpyDeviceToHost));
    *out_scale = sc[0];
    *out_mean_norm = sc[1];
    return DKM_OK;
}

dkm_status dkm_sq_compress(float scale, const float *shift, int64_t dim, const float *x,
                           int64_t n, uint8_t *out) {
    dkm_status st = sq_check(scale, shift, dim);
    if (st) return st;
    if (n < 0 || (n > 0 && (!x || !out))) return fail(DKM_ERR_INVALID_ARG, "bad sizes");
    if (n == 0) return DKM_OK;
    HIP_TRY(hipSetDevice(g_device));
    const int64_t cs = dkm::sq_code_stride(dim);
    DevBuf<float> dx, dsh, dcp;
    DevBuf<uint8_t> dc;
    DevBuf<uint32_t> dnan;
    HIP_TRY(dx.ensure(n * dim));
    HIP_TRY(dsh.ensure(dim));
    HIP_TRY(dcp.ensure(n));
    HIP_TRY(dc.ensure(n * cs));
    HIP_TRY(dnan.ensure(1));
    HIP_TRY(hipMemcpy(dx.p, x, sizeof(float) * n * dim, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(dsh.p, shift, sizeof(float) * dim, hipMemcpyHostToDevice));
    HIP_TRY(hipMemset(dnan.p, 0, sizeof(uint32_t)));
    HIP_TRY(dkm::launch_sq_encode(dx.p, dim, n, (int)dim, scale, dsh.p, dc.p, cs, dcp.p, dnan.p,
                                  nullptr));
    st = sq_download(dc.p, dcp.p, n, dim, out);
    if (st) return st;
    uint32_t nan = 0;
    HIP_TRY(hipMemcpy(&nan, dnan.p, sizeof(uint32_t), hipMemcpyDeviceToHost));
    if (nan) return fail(DKM_ERR_INVALID_ARG, "input contains NaN");
    return DKM_OK;
}

dkm_status dkm_sq_distance_batch(int metric, float scale, const float *shift, int64_t dim,
                                 const uint8_t *left, int64_t nl, const uint8_t *right,
                                 int64_t nr, const uint32_t *ids, int64_t m, float *out) {
    if (metric != DKM_L2 && metric != DKM_INNER_PRODUCT && metric != DKM_COSINE_NORMALIZED)
        return fail(DKM_ERR_INVALID_ARG, "scalar quantization supports L2, inner product and "
                                         "cosine-normalized (got metric %d)", metric);
    dkm_status st = sq_check(scale, shift, dim);
    if (st) return st;
    if (nl < 0 || nr <= 0 || m < 0 || !left || !right || !out || (nl * m > 0 && !ids))
        return fail(DKM_ERR_INVALID_ARG, "bad sizes");
    if (nl == 0 || m == 0) return DKM_OK;
    for (int64_t i = 0; i < nl * m; ++i)
        if (ids[i] >= (uint64_t)nr)
            return fail(DKM_ERR_INDEX, "id %u out of range (%lld rows)", ids[i], (long long)nr);
    HIP_TRY(hipSetDevice(g_device));
    DevBuf<uint8_t> dl, dr;
    DevBuf<float> dlc, drc, dsh, dssn, dout;
    DevBuf<uint32_t> di;
    st = sq_upload(left, nl, dim, dl, dlc);
    if (st) return st;
    st = sq_upload(right, nr, dim, dr, drc);
    if (st) return st;
    HIP_TRY(dsh.ensure(dim));
    HIP_TRY(dssn.ensure(1));
    HIP_TRY(di.ensure(nl * m));
    HIP_TRY(dout.ensure(nl * m));
    HIP_TRY(hipMemcpy(dsh.p, shift, sizeof(float) * dim, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(di.p, ids, sizeof(uint32_t) * nl * m, hipMemcpyHostToDevice));
    HIP_TRY(dkm::launch_sq_shift_norm(dsh.p, (int)dim, dssn.p, nullptr));
    float ssn = 0.0f;
    HIP_TRY(hipMemcpy(&ssn, dssn.p, sizeof(float), hipMemcpyDeviceToHost));
    HIP_TRY(dkm::launch_sq_pairs(metric, dkm::sq_bss(scale), ssn, dl.p, dlc.p, dr.p, drc.p,
                                 dkm::sq_code_stride(dim), di.p, nl, m, dout.p, nullptr));
    HIP_TRY(hipMemcpy(out, dout.p, sizeof(float) * nl * m, hipMemcpyDeviceToHost));
    return DKM_OK;
}

dkm_status dkm_set_sq(dkm_index *ix, float scale, const float *shift, float mean_norm) {
    if (!ix) return fail(DKM_ERR_INVALID_ARG, "null index");
    dkm_status st = sq_check(scale, shift, ix->dim);
    if (st) return st;
    if (ix->metric != DKM_L2 && ix->metric != DKM_INNER_PRODUCT && ix->metric != DKM_COSINE_NORMALIZED)
        return fail(DKM_ERR_INVALID_ARG, "scalar quantization supports L2, inner product and "
                                         "cosine-normalized");
    std::lock_guard<std::mutex> g(ix->mu);
    HIP_TRY(hipSetDevice(ix->device));
    ix->sq_cstride = dkm::sq_code_stride(ix->dim);
    HIP_TRY(ix->d_sq_shift.ensure(ix->dim));
    HIP_TRY(ix->d_sq_comp.ensure(ix->rows));
    HIP_TRY(ix->d_sq_codes.ensure(ix->rows * ix->sq_cstride + 16));
    HIP_TRY(hipMemsetAsync(ix->d_sq_codes.p, 0, ix->rows * ix->sq_cstride + 16, ix->stream));
    HIP_TRY(hipMemsetAsync(ix->d_sq_comp.p, 0, sizeof(float) * ix->rows, ix->stream));
    HIP_TRY(hipStreamSynchronize(ix->stream));
    HIP_TRY(hipMemcpy(ix->d_sq_shift.p, shift, sizeof(float) * ix->dim, hipMemcpyHostToDevice));
    DevBuf<float> dssn;
    HIP_TRY(dssn.ensure(1));
    HIP_TRY(dkm::launch_sq_shift_norm(ix->d_sq_shift.p, (int)ix->dim, dssn.p, ix->stream));
    HIP_TRY(hipStreamSynchronize(ix->stream));
    HIP_TRY(hipMemcpy(&ix->sq_ssn, dssn.p, sizeof(float), hipMemcpyDeviceToHost));
    ix->sq_scale = scale;
    ix->sq_mean_norm = mean_norm;
    ix->sq_active = true;
    ix->sph_active = false;
    ix->sq_codes_set = false;
    return DKM_OK;
}

dkm_status dkm_encode_index_sq(dkm_index *ix) {
    if (!ix) return fail(DKM_ERR_INVALID_ARG, "null index");
    if (!ix->sq_active) return fail(DKM_ERR_NOT_READY, "set the scalar quantizer first");
    std::lock_guard<std::mutex> g(ix->mu);
    HIP_TRY(hipSetDevice(ix->device));
    DevBuf<uint32_t> dnan;
    HIP_TRY(dnan.ensure(1));
    HIP_TRY(hipMemsetAsync(dnan.p, 0, sizeof(uint32_t), ix->stream));
    HIP_TRY(dkm::launch_sq_encode(ix->d_vec, ix->vstride, ix->rows, (int)ix->dim, ix->sq_scale,
                                  ix->d_sq_shift.p, ix->d_sq_codes.p, ix->sq_cstride,
                                  ix->d_sq_comp.p, dnan.p, ix->stream));
    HIP_TRY(hipStreamSynchronize(ix->stream));
    uint32_t nan = 0;
    HIP_TRY(hipMemcpy(&nan, dnan.p, sizeof(uint32_t), hipMemcpyDeviceToHost));
    if (nan) return fail(DKM_ERR_INVALID_ARG, "input contains NaN");
    ix->sq_codes_set = true;
    return DKM_OK;
}

dkm_status dkm_set_sq_codes(dkm_index *ix, const uint8_t *codes, int64_t row0, int64_t n) {
    if (!ix || (!codes && n > 0)) return fail(DKM_ERR_INVALID_ARG, "null argument");
    if (!ix->sq_active) return fail(DKM_ERR_NOT_READY, "set the scalar quantizer first");
    if (row0 < 0 || n < 0 || row0 + n > ix->rows) return fail(DKM_ERR_INDEX, "rows out of range");
    std::lock_guard<std::mutex> g(ix->mu);
    HIP_TRY(hipSetDevice(ix->device));
    if (n) {
        DevBuf<uint8_t> c;
        DevBuf<float> cp;
        dkm_status st = sq_upload(codes, n, ix->dim, c, cp);
        if (st) return st;
        HIP_TRY(hipMemcpy(ix->d_sq_codes.p + row0 * ix->sq_cstride, c.p, n * ix->sq_cstride,
                          hipMemcpyDeviceToDevice));
        HIP_TRY(hipMemcpy(ix->d_sq_comp.p + row0, cp.p, sizeof(float) * n, hipMemcpyDeviceToDevice));
    }
    ix->sq_codes_set = true;
    return DKM_OK;
}

dkm_status dkm_get_sq_codes(dkm_index *ix, uint8_t *codes, int64_t row0, int64_t n) {
    if (!ix || !ix->sq_active) return fail(DKM_ERR_NOT_READY, "no scalar-quantized codes");
    if (row0 < 0 || n < 0 || row0 + n > ix->rows) return fail(DKM_ERR_INDEX, "rows out of range");
    std::lock_guard<std::mutex> g(ix->mu);
    HIP_TRY(hipSetDevice(ix->device));
    HIP_TRY(hipStreamSynchronize(ix->stream));
    if (n) return sq_download(ix->d_sq_codes.p + row0 * ix->sq_cstride, ix->d_sq_comp.p + row0, n,
                              ix->dim, codes);
    return DKM_OK;
}

dkm_status dkm_use_pq(dkm_index *ix) {
    if (!ix) return fail(DKM_ERR_INVALID_ARG, "null index");
    std::lock_guard<std::mutex> g(ix->mu);
    ix->sq_active = false;
    ix->sph_active = false;
    return DKM_OK;
}

// ------------------------------------------------------------------------------------
// 1-bit spherical quantization (diskann-quantization/src/spherical)
// ------------------------------------------------------------------------------------

}  // extern "C"

struct dkm_sph {
    dkm::SphParams q{};                  // device pointers below
    DevBuf<float> shift;
    DevBuf<uint32_t> signs0, signs1, sub;
};

namespace {

uint16_t host_f32_to_f16(float f) {
    uint32_t x;
    std::memcpy(&x, &f, 4);
    const uint32_t sign = (x >> 16) & 0x8000u, ex = (x >> 23) & 0xFFu;
    uint32_t mant = x & 0x7FFFFFu;
    if (ex == 0xFFu) return (uint16_t)(sign | 0x7C00u | (mant ? 0x200u : 0u));
    const int32_t e = (int32_t)ex - 112;
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

float host_f16_to_f32(uint16_t h) {
    const uint32_t sign = (uint32_t)(h & 0x8000u) << 16, ex = (h >> 10) & 0x1Fu, mant = h & 0x3FFu;
    uint32_t x;
    if (ex == 0x1Fu) {
        x = sign | 0x7F800000u | (mant << 13);
    } else if (ex == 0) {
        if (mant == 0) return sign ? -0.0f : 0.0f;
        const float v = (float)mant * (1.0f / 16777216.0f);
        return sign ? -v : v;
    } else {
        x = sign | ((ex + 112u) << 23) | (mant << 13);
    }
    float f;
    std::memcpy(&f, &x, 4);
    return f;
}

// canonical rows (bits, f16, f16, u16) <-> device bit rows + SphMeta
dkm_status sph_upload(const dkm::SphParams &q, const uint8_t *canon, int64_t n,
                      DevBuf<uint8_t> &bits, DevBuf<dkm::SphMeta> &meta, int64_t cap = -1) {
    const int64_t cb = dkm::sph_canon_bytes(q.out_dim), nb = (q.out_dim + 7) / 8;
    std::vector<uint8_t> rows((size_t)(n * q.cstride), 0);
    std::vector<dkm::SphMeta> ms((size_t)n);
    for (int64_t r = 0; r < n; ++r) {
        const uint8_t *c = canon + r * cb;
        std::memcpy(&rows[(size_t)(r * q.cstride)], c, (size_t)nb);
        if (q.out_dim % 8) rows[(size_t)(r * q.cstride + nb - 1)] &= (uint8_t)((1u << (q.out_dim % 8)) - 1u);
        auto u16 = [&](int k) { return (uint16_t)(c[nb + 2 * k] | (c[nb + 2 * k + 1] << 8)); };
        ms[(size_t)r] = {host_f16_to_f32(u16(0)), host_f16_to_f32(u16(1)), (float)u16(2), 0.0f};
    }
    HIP_TRY(bits.ensure(cap >= 0 ? cap * q.cstride : n * q.cstride));
    HIP_TRY(meta.ensure(cap >= 0 ? cap : n));
    HIP_TRY(hipMemcpy(bits.p, rows.data(), rows.size(), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(meta.p, ms.data(), sizeof(dkm::SphMeta) * n, hipMemcpyHostToDevice));
    return DKM_OK;
}

dkm_status sph_download(const dkm::SphParams &q, const uint8_t *bits, const dkm::SphMeta *meta,
                        int64_t n, uint8_t *canon) {
    const int64_t cb = dkm::sph_canon_bytes(q.out_dim), nb = (q.out_dim + 7) / 8;
    std::vector<uint8_t> rows((size_t)(n * q.cstride));
    std::vector<dkm::SphMeta> ms((size_t)n);
    HIP_TRY(hipMemcpy(rows.data(), bits, rows.size(), hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(ms.data(), meta, sizeof(dkm::SphMeta) * n, hipMemcpyDeviceToHost));
    for (int64_t r = 0; r < n; ++r) {
        uint8_t *c = canon + r * cb;
        std::memcpy(c, &rows[(size_t)(r * q.cstride)], (size_t)nb);
        const uint16_t h[3] = {host_f32_to_f16(ms[(size_t)r].ipc), host_f32_to_f16(ms[(size_t)r].ms),
                               (uint16_t)ms[(size_t)r].bs};
        for (int k = 0; k < 3; ++k) {
            c[nb + 2 * k] = (uint8_t)(h[k] & 0xFFu);
            c[nb + 2 * k + 1] = (uint8_t)(h[k] >> 8);
        }
    }
    return DKM_OK;
}

int sph_metric(int m) {
    if (m == DKM_L2) return dkm::METRIC_L2;
    if (m == DKM_INNER_PRODUCT) return dkm::METRIC_IP;
    if (m == DKM_COSINE) return dkm::METRIC_COSINE;
    return -1;
}

dkm_status sph_errors(uint32_t e) {
    if (e & dkm::SPH_ERR_NAN) return fail(DKM_ERR_INVALID_ARG, "input contains NaN or infinite values");
    if (e & dkm::SPH_ERR_META) return fail(DKM_ERR_INVALID_ARG, "a correction term does not fit in f16");
    return DKM_OK;
}

}  // namespace

extern "C" {

dkm_status dkm_sph_create(int metric, int64_t dim, const float *centroid, float mean_norm,
                          float pre_scale, int transform, const uint8_t *signs0,
                          const uint8_t *signs1, int64_t n_signs1, int64_t padded_dim,
                          const uint32_t *subsample, int64_t n_subsample, dkm_sph **out) {
    if (!out || !centroid || dim <= 0) return fail(DKM_ERR_INVALID_ARG, "bad spherical quantizer");
    *out = nullptr;
    const int m = sph_metric(metric);
    if (m < 0) return fail(DKM_ERR_INVALID_ARG, "spherical quantization supports L2, inner product and cosine");
    if (!(mean_norm > 0.0f) || !(pre_scale > 0.0f)) return fail(DKM_ERR_INVALID_ARG, "mean_norm and pre_scale must be positive");
    dkm::SphParams q{};
    q.metric = m;
    q.dim = (int)dim;
    q.kind = transform;
    if (transform == dkm::SPH_NULL) {
        q.n = q.out_dim = (int)dim;
        if (subsample) return fail(DKM_ERR_INVALID_ARG, "the null transform has no subsample");
    } else if (transform == dkm::SPH_PADDING_HADAMARD) {
        if (!signs0 || padded_dim < dim || (padded_dim & (padded_dim - 1)))
            return fail(DKM_ERR_INVALID_ARG, "padding Hadamard needs dim signs and a power-of-two padded dim >= dim");
        q.n = (int)padded_dim;
        q.out_dim = subsample ? (int)n_subsample : (int)padded_dim;
    } else if (transform == dkm::SPH_DOUBLE_HADAMARD) {
        if (!signs0 || !signs1) return fail(DKM_ERR_INVALID_ARG, "double Hadamard needs both sign vectors");
        q.out_dim = subsample ? (int)n_subsample : (int)n_signs1;
        q.n = std::max<int>((int)dim, q.out_dim);
        if (n_signs1 != q.n) return fail(DKM_ERR_INVALID_ARG, "double Hadamard: signs1 must have max(dim, output dim) entries");
    } else {
        return fail(DKM_ERR_INVALID_ARG, "unknown transform %d", transform);
    }
    if (subsample) {
        if (n_subsample <= 0 || n_subsample >= q.n) return fail(DKM_ERR_INVALID_ARG, "bad subsample length");
        for (int64_t i = 0; i < n_subsample; ++i)
            if (subsample[i] >= (uint32_t)q.n || (i > 0 && subsample[i] <= subsample[i - 1]))
                return fail(DKM_ERR_INVALID_ARG, "subsample indices must increase strictly and stay below %d", q.n);
        q.sub_rescale = std::sqrt((float)q.n / (float)n_subsample);
    }
    if (q.n > 16384) return fail(DKM_ERR_CAPACITY, "transform length above 16384 is not supported");
    q.split = 1;
    while (q.split * 2 <= q.n) q.split *= 2;
    q.mean_norm = mean_norm;
    q.pre_scale = pre_scale;
    q.cstride = dkm::sph_bit_stride(q.out_dim);
    HIP_TRY(hipSetDevice(g_device));
    auto h = std::make_unique<dkm_sph>();
    HIP_TRY(h->shift.ensure(dim));
    HIP_TRY(hipMemcpy(h->shift.p, centroid, sizeof(float) * dim, hipMemcpyHostToDevice));
    if (signs0 && transform != dkm::SPH_NULL) {
        std::vector<uint32_t> s0((size_t)dim);
        for (int64_t i = 0; i < dim; ++i) s0[(size_t)i] = signs0[i] ? 0x80000000u : 0u;
        HIP_TRY(h->signs0.ensure(dim));
        HIP_TRY(hipMemcpy(h->signs0.p, s0.data(), sizeof(uint32_t) * dim, hipMemcpyHostToDevice));
    }
    if (transform == dkm::SPH_DOUBLE_HADAMARD) {
        std::vector<uint32_t> s1((size_t)q.n);
        for (int i = 0; i < q.n; ++i) s1[(size_t)i] = signs1[i] ? 0x80000000u : 0u;
        HIP_TRY(h->signs1.ensure(q.n));
        HIP_TRY(hipMemcpy(h->signs1.p, s1.data(), sizeof(uint32_t) * q.n, hipMemcpyHostToDevice));
    }
    if (subsample) {
        HIP_TRY(h->sub.ensure(n_subsample));
        HIP_TRY(hipMemcpy(h->sub.p, subsample, sizeof(uint32_t) * n_subsample, hipMemcpyHostToDevice));
    }
    q.shift = h->shift.p;
    q.signs0 = h->signs0.p;
    q.signs1 = h->signs1.p;
    q.subsample = subsample ? h->sub.p : nullptr;
    DevBuf<float> dss;
    HIP_TRY(dss.ensure(1));
    HIP_TRY(dkm::launch_sph_shift_sq(q.shift, q.dim, dss.p, nullptr));
    HIP_TRY(hipMemcpy(&q.shift_sq, dss.p, sizeof(float), hipMemcpyDeviceToHost));
    h->q = q;
    *out = h.release();
    return DKM_OK;
}

void dkm_sph_destroy(dkm_sph *q) {
    if (!q) return;
    (void)hipSetDevice(g_device);
    delete q;
}

dkm_status dkm_sph_dims(const dkm_sph *q, int64_t *out_dim, int64_t *code_bytes) {
    if (!q || !out_dim || !code_bytes) return fail(DKM_ERR_INVALID_ARG, "null argument");
    *out_dim = q->q.out_dim;
    *code_bytes = dkm::sph_canon_bytes(q->q.out_dim);
    return DKM_OK;
}

dkm_status dkm_sph_train_stats(int metric, const float *data, int64_t n, int64_t dim,
                               float *out_centroid, float *out_mean_norm) {
    const int m = sph_metric(metric);
    if (m < 0) return fail(DKM_ERR_INVALID_ARG, "spherical quantization supports L2, inner product and cosine");
    if (!data || !out_centroid || !out_mean_norm || n <= 0 || dim <= 0)
        return fail(DKM_ERR_INVALID_ARG, "training data cannot be empty");
    HIP_TRY(hipSetDevice(g_device));
    DevBuf<float> dx, dc;
    DevBuf<double> dm, dn;
    HIP_TRY(dx.ensure(n * dim));
    HIP_TRY(dc.ensure(dim + 1));
    HIP_TRY(dm.ensure(dim));
    HIP_TRY(dn.ensure(n));
    HIP_TRY(hipMemcpy(dx.p, data, sizeof(float) * n * dim, hipMemcpyHostToDevice));
    HIP_TRY(dkm::launch_sph_train_stats(m, dx.p, n, (int)dim, dm.p, dn.p, dc.p, dc.p + dim, nullptr));
    HIP_TRY(hipMemcpy(out_centroid, dc.p, sizeof(float) * dim, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(out_mean_norm, dc.p + dim, sizeof(float), hipMemcpyDeviceToHost));
    if (!(*out_mean_norm > 0.0f) || !std::isfinite(*out_mean_norm))
        return fail(DKM_ERR_INVALID_ARG, "the mean norm is not positive and finite");
    return DKM_OK;
}

dkm_status dkm_sph_compress(const dkm_sph *q, const float *x, int64_t n, uint8_t *out) {
    if (!q || n < 0 || (n > 0 && (!x || !out))) return fail(DKM_ERR_INVALID_ARG, "bad arguments");
    if (n == 0) return DKM_OK;
    HIP_TRY(hipSetDevice(g_device));
    const int64_t cb = dkm::sph_canon_bytes(q->q.out_dim);
    DevBuf<float> dx;
    DevBuf<uint8_t> dout;
    DevBuf<uint32_t> derr;
    HIP_TRY(dx.ensure(n * q->q.dim));
    HIP_TRY(dout.ensure(n * cb));
    HIP_TRY(derr.ensure(1));
    HIP_TRY(hipMemcpy(dx.p, x, sizeof(float) * n * q->q.dim, hipMemcpyHostToDevice));
    HIP_TRY(hipMemset(derr.p, 0, sizeof(uint32_t)));
    HIP_TRY(dkm::launch_sph_encode(q->q, dx.p, q->q.dim, n, nullptr, nullptr, dout.p, derr.p, nullptr));
    HIP_TRY(hipMemcpy(out, dout.p, (size_t)(n * cb), hipMemcpyDeviceToHost));
    uint32_t e = 0;
    HIP_TRY(hipMemcpy(&e, derr.p, sizeof(uint32_t), hipMemcpyDeviceToHost));
    return sph_errors(e);
}

dkm_status dkm_sph_distance_batch(const dkm_sph *q, const uint8_t *left, int64_t nl,
                                  const uint8_t *right, int64_t nr, const uint32_t *ids, int64_t m,
                                  float *out) {
    if (!q || nl < 0 || nr <= 0 || m < 0 || !left || !right || !out || (nl * m > 0 && !ids))
        return fail(DKM_ERR_INVALID_ARG, "bad sizes");
    if (nl == 0 || m == 0) return DKM_OK;
    for (int64_t i = 0; i < nl * m; ++i)
        if (ids[i] >= (uint64_t)nr) return fail(DKM_ERR_INDEX, "id %u out of range (%lld rows)", ids[i], (long long)nr);
    HIP_TRY(hipSetDevice(g_device));
    DevBuf<uint8_t> lb, rb;
    DevBuf<dkm::SphMeta> lm, rm;
    DevBuf<uint32_t> di;
    DevBuf<float> dout;
    dkm_status st = sph_upload(q->q, left, nl, lb, lm);
    if (st) return st;
    if ((st = sph_upload(q->q, right, nr, rb, rm))) return st;
    HIP_TRY(di.ensure(nl * m));
    HIP_TRY(dout.ensure(nl * m));
    HIP_TRY(hipMemcpy(di.p, ids, sizeof(uint32_t) * nl * m, hipMemcpyHostToDevice));
    HIP_TRY(dkm::launch_sph_pairs(q->q, lb.p, lm.p, rb.p, rm.p, di.p, nl, m, dout.p, nullptr));
    HIP_TRY(hipMemcpy(out, dout.p, sizeof(float) * nl * m, hipMemcpyDeviceToHost));
    return DKM_OK;
}

dkm_status dkm_sph_query_distance_batch(const dkm_sph *q, int layout, int allow_rescale,
                                        const float *queries, int64_t nq, const uint8_t *codes,
                                        int64_t n, const uint32_t *ids, int64_t m, float *out) {
    if (!q || nq < 0 || n <= 0 || m < 0 || !queries || !codes || !out || (nq * m > 0 && !ids))
        return fail(DKM_ERR_INVALID_ARG, "bad sizes");
    if (layout != dkm::SPH_QUERY_SAME_AS_DATA && layout != dkm::SPH_QUERY_FOUR_BIT &&
        layout != dkm::SPH_QUERY_FULL)
        return fail(DKM_ERR_INVALID_ARG, "unknown query layout %d", layout);
    if (nq == 0 || m == 0) return DKM_OK;
    for (int64_t i = 0; i < nq * m; ++i)
        if (ids[i] >= (uint64_t)n) return fail(DKM_ERR_INDEX, "id %u out of range (%lld rows)", ids[i], (long long)n);
    HIP_TRY(hipSetDevice(g_device));
    DevBuf<uint8_t> b;
    DevBuf<dkm::SphMeta> mt;
    DevBuf<uint32_t> di, derr;
    DevBuf<float> dq, dout;
    dkm_status st = sph_upload(q->q, codes, n, b, mt);
    if (st) return st;
    HIP_TRY(dq.ensure(nq * q->q.dim));
    HIP_TRY(di.ensure(nq * m));
    HIP_TRY(dout.ensure(nq * m));
    HIP_TRY(derr.ensure(1));
    HIP_TRY(hipMemcpy(dq.p, queries, sizeof(float) * nq * q->q.dim, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(di.p, ids, sizeof(uint32_t) * nq * m, hipMemcpyHostToDevice));
    HIP_TRY(hipMemset(derr.p, 0, sizeof(uint32_t)));
    HIP_TRY(dkm::launch_sph_query_pairs(q->q, layout, allow_rescale, dq.p, nq, b.p, mt.p, di.p, m,
                                        dout.p, derr.p, nullptr));
    HIP_TRY(hipMemcpy(out, dout.p, sizeof(float) * nq * m, hipMemcpyDeviceToHost));
    uint32_t e = 0;
    HIP_TRY(hipMemcpy(&e, derr.p, sizeof(uint32_t), hipMemcpyDeviceToHost));
    return sph_errors(e);
}

dkm_status dkm_set_spherical(dkm_index *ix, const dkm_sph *q, int query_layout) {
    if (!ix || !q) return fail(DKM_ERR_INVALID_ARG, "null argument");
    if (q->q.dim != ix->dim) return fail(DKM_ERR_DIM_MISMATCH, "quantizer dim %d != index dim %lld", q->q.dim, (long long)ix->dim);
    if (sph_metric(ix->metric) != q->q.metric)
        return fail(DKM_ERR_INVALID_ARG, "quantizer metric does not match the index metric");
    if (query_layout != dkm::SPH_QUERY_SAME_AS_DATA && query_layout != dkm::SPH_QUERY_FOUR_BIT)
        return fail(DKM_ERR_INVALID_ARG, "traversal supports the same-as-data and 4-bit query layouts");
    std::lock_guard<std::mutex> g(ix->mu);
    HIP_TRY(hipSetDevice(ix->device));
    dkm::SphParams p = q->q;
    HIP_TRY(ix->d_sph_shift.ensure(p.dim));
    HIP_TRY(hipMemcpy(ix->d_sph_shift.p, p.shift, sizeof(float) * p.dim, hipMemcpyDeviceToDevice));
    p.shift = ix->d_sph_shift.p;
    if (p.signs0) {
        HIP_TRY(ix->d_sph_signs0.ensure(p.dim));
        HIP_TRY(hipMemcpy(ix->d_sph_signs0.p, p.signs0, sizeof(uint32_t) * p.dim, hipMemcpyDeviceToDevice));
        p.signs0 = ix->d_sph_signs0.p;
    }
    if (p.signs1) {
        HIP_TRY(ix->d_sph_signs1.ensure(p.n));
        HIP_TRY(hipMemcpy(ix->d_sph_signs1.p, p.signs1, sizeof(uint32_t) * p.n, hipMemcpyDeviceToDevice));
        p.signs1 = ix->d_sph_signs1.p;
    }
    if (p.subsample) {
        HIP_TRY(ix->d_sph_sub.ensure(p.out_dim));
        HIP_TRY(hipMemcpy(ix->d_sph_sub.p, p.subsample, sizeof(uint32_t) * p.out_dim, hipMemcpyDeviceToDevice));
        p.subsample = ix->d_sph_sub.p;
    }
    HIP_TRY(ix->d_sph_bits.ensure(ix->rows * p.cstride + 16));
    HIP_TRY(ix->d_sph_meta.ensure(ix->rows));
    HIP_TRY(hipMemsetAsync(ix->d_sph_bits.p, 0, ix->rows * p.cstride + 16, ix->stream));
    HIP_TRY(hipMemsetAsync(ix->d_sph_meta.p, 0, sizeof(dkm::SphMeta) * ix->rows, ix->stream));
    HIP_TRY(hipStreamSynchronize(ix->stream));
    ix->sphq = p;
    ix->sph_layout = query_layout;
    ix->sph_active = true;
    ix->sph_codes_set = false;
    ix->sq_active = false;
    return DKM_OK;
}

dkm_status dkm_encode_index_sph(dkm_index *ix) {
    if (!ix) return fail(DKM_ERR_INVALID_ARG, "null index");
    if (!ix->sph_active) return fail(DKM_ERR_NOT_READY, "set the spherical quantizer first");
    std::lock_guard<std::mutex> g(ix->mu);
    HIP_TRY(hipSetDevice(ix->device));
    DevBuf<uint32_t> derr;
    HIP_TRY(derr.ensure(1));
    HIP_TRY(hipMemsetAsync(derr.p, 0, sizeof(uint32_t), ix->stream));
    HIP_TRY(dkm::launch_sph_encode(ix->sphq, ix->d_vec, ix->vstride, ix->rows, ix->d_sph_bits.p,
                                   ix->d_sph_meta.p, nullptr, derr.p, ix->stream));
    HIP_TRY(hipStreamSynchronize(ix->stream));
    uint32_t e = 0;
    HIP_TRY(hipMemcpy(&e, derr.p, sizeof(uint32_t), hipMemcpyDeviceToHost));
    dkm_status st = sph_errors(e);
    if (st) return st;
    ix->sph_codes_set = true;
    return DKM_OK;
}

dkm_status dkm_set_sph_codes(dkm_index *ix, const uint8_t *codes, int64_t row0, int64_t n) {
    if (!ix || (!codes && n > 0)) return fail(DKM_ERR_INVALID_ARG, "null argument");
    if (!ix->sph_active) return fail(DKM_ERR_NOT_READY, "set the spherical quantizer first");
    if (row0 < 0 || n < 0 || row0 + n > ix->rows) return fail(DKM_ERR_INDEX, "rows out of range");
    std::lock_guard<std::mutex> g(ix->mu);
    HIP_TRY(hipSetDevice(ix->device));
    if (n) {
        DevBuf<uint8_t> b;
        DevBuf<dkm::SphMeta> mt;
        dkm_status st = sph_upload(ix->sphq, codes, n, b, mt);
        if (st) return st;
        HIP_TRY(hipMemcpy(ix->d_sph_bits.p + row0 * ix->sphq.cstride, b.p, n * ix->sphq.cstride,
                          hipMemcpyDeviceToDevice));
        HIP_TRY(hipMemcpy(ix->d_sph_meta.p + row0, mt.p, sizeof(dkm::SphMeta) * n, hipMemcpyDeviceToDevice));
    }
    ix->sph_codes_set = true;
    return DKM_OK;
}

dkm_status dkm_get_sph_codes(dkm_index *ix, uint8_t *codes, int64_t row0, int64_t n) {
    if (!ix || !ix->sph_active) return fail(DKM_ERR_NOT_READY, "no spherical codes");
    if (row0 < 0 || n < 0 || row0 + n > ix->rows) return fail(DKM_ERR_INDEX, "rows out of range");
    std::lock_guard<std::mutex> g(ix->mu);
    HIP_TRY(hipSetDevice(ix->device));
    HIP_TRY(hipStreamSynchronize(ix->stream));
    if (n) return sph_download(ix->sphq, ix->d_sph_bits.p + row0 * ix->sphq.cstride,
                               ix->d_sph_meta.p + row0, n, codes);
    return DKM_OK;
}

// ------------------------------------------------------------------------------------
// 8-bit MinMax quantization (diskann-quantization/src/minmax)
// ------------------------------------------------------------------------------------

}  // extern "C"

struct dkm_minmax {
    dkm_sph *t = nullptr;   // the transform (centroid / metric unused)
    float grid = 1.0f;
};

extern "C" {

dkm_status dkm_minmax_create(int64_t dim, float grid_scale, int transform, const uint8_t *signs0,
                             const uint8_t *signs1, int64_t n_signs1, int64_t padded_dim,
                             const uint32_t *subsample, int64_t n_subsample, dkm_minmax **out) {
    if (!out) return fail(DKM_ERR_INVALID_ARG, "null argument");
    *out = nullptr;
    if (!(grid_scale > 0.0f) || !std::isfinite(grid_scale)) return fail(DKM_ERR_INVALID_ARG, "grid_scale must be positive");
    if (dim <= 0) return fail(DKM_ERR_INVALID_ARG, "dim must be positive");
    std::vector<float> zero((size_t)dim, 0.0f);
    auto h = std::make_unique<dkm_minmax>();
    dkm_status st = dkm_sph_create(DKM_L2, dim, zero.data(), 1.0f, 1.0f, transform, signs0, signs1,
                                   n_signs1, padded_dim, subsample, n_subsample, &h->t);
    if (st) return st;
    h->grid = grid_scale;
    *out = h.release();
    return DKM_OK;
}

void dkm_minmax_destroy(dkm_minmax *q) {
    if (!q) return;
    dkm_sph_destroy(q->t);
    delete q;
}

dkm_status dkm_minmax_dims(const dkm_minmax *q, int64_t *out_dim, int64_t *code_bytes) {
    if (!q || !out_dim || !code_bytes) return fail(DKM_ERR_INVALID_ARG, "null argument");
    *out_dim = q->t->q.out_dim;
    *code_bytes = 20 + (int64_t)q->t->q.out_dim;
    return DKM_OK;
}

dkm_status dkm_minmax_compress(const dkm_minmax *q, const float *x, int64_t n, uint8_t *out,
                               float *out_loss) {
    if (!q || n < 0 || (n > 0 && (!x || !out))) return fail(DKM_ERR_INVALID_ARG, "bad arguments");
    if (n == 0) return DKM_OK;
    HIP_TRY(hipSetDevice(g_device));
    const dkm::SphParams &t = q->t->q;
    const int64_t rb = 20 + (int64_t)t.out_dim;
    DevBuf<float> dx, dl;
    DevBuf<uint8_t> dout;
    DevBuf<uint32_t> derr;
    HIP_TRY(dx.ensure(n * t.dim));
    HIP_TRY(dl.ensure(n));
    HIP_TRY(dout.ensure(n * rb));
    HIP_TRY(derr.ensure(1));
    HIP_TRY(hipMemcpy(dx.p, x, sizeof(float) * n * t.dim, hipMemcpyHostToDevice));
    HIP_TRY(hipMemset(derr.p, 0, sizeof(uint32_t)));
    HIP_TRY(dkm::launch_minmax_encode(t, q->grid, dx.p, t.dim, n, dout.p, dl.p, derr.p, nullptr));
    HIP_TRY(hipMemcpy(out, dout.p, (size_t)(n * rb), hipMemcpyDeviceToHost));
    if (out_loss) HIP_TRY(hipMemcpy(out_loss, dl.p, sizeof(float) * n, hipMemcpyDeviceToHost));
    uint32_t e = 0;
    HIP_TRY(hipMemcpy(&e, derr.p, sizeof(uint32_t), hipMemcpyDeviceToHost));
    if (e) return fail(DKM_ERR_INVALID_ARG, "input contains NaN");
    return DKM_OK;
}

static int minmax_metric(int m) {
    if (m == DKM_L2) return dkm::METRIC_L2;
    if (m == DKM_INNER_PRODUCT) return dkm::METRIC_IP;
    if (m == DKM_COSINE) return dkm::METRIC_COSINE;
    if (m == DKM_COSINE_NORMALIZED) return dkm::METRIC_COSNORM;
    return -1;
}

dkm_status dkm_minmax_distance_batch(const dkm_minmax *q, int metric, const uint8_t *left,
                                     int64_t nl, const uint8_t *right, int64_t nr,
                                     const uint32_t *ids, int64_t m, float *out) {
    const int mm = minmax_metric(metric);
    if (!q || mm < 0) return fail(DKM_ERR_INVALID_ARG, "bad quantizer or metric");
    if (nl < 0 || nr <= 0 || m < 0 || !left || !right || !out || (nl * m > 0 && !ids))
        return fail(DKM_ERR_INVALID_ARG, "bad sizes");
    if (nl == 0 || m == 0) return DKM_OK;
    for (int64_t i = 0; i < nl * m; ++i)
        if (ids[i] >= (uint64_t)nr) return fail(DKM_ERR_INDEX, "id %u out of range (%lld rows)", ids[i], (long long)nr);
    const int rb = 20 + q->t->q.out_dim;
    HIP_TRY(hipSetDevice(g_device));
    DevBuf<uint8_t> dl, dr;
    DevBuf<uint32_t> di;
    DevBuf<float> dout;
    HIP_TRY(dl.ensure(nl * rb));
    HIP_TRY(dr.ensure(nr * rb));
    HIP_TRY(di.ensure(nl * m));
    HIP_TRY(dout.ensure(nl * m));
    HIP_TRY(hipMemcpy(dl.p, left, (size_t)(nl * rb), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(dr.p, right, (size_t)(nr * rb), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(di.p, ids, sizeof(uint32_t) * nl * m, hipMemcpyHostToDevice));
    HIP_TRY(dkm::launch_minmax_pairs(mm, dl.p, dr.p, rb, di.p, nl, m, dout.p, nullptr));
    HIP_TRY(hipMemcpy(out, dout.p, sizeof(float) * nl * m, hipMemcpyDeviceToHost));
    return DKM_OK;
}

dkm_status dkm_minmax_query_distance_batch(const dkm_minmax *q, int metric, const float *queries,
                                           int64_t nq, const uint8_t *codes, int64_t n,
                                           const uint32_t *ids, int64_t m, float *out) {
    const int mm = minmax_metric(metric);
    if (!q || mm < 0) return fail(DKM_ERR_INVALID_ARG, "bad quantizer or metric");
    if (nq < 0 || n <= 0 || m < 0 || !queries || !codes || !out || (nq * m > 0 && !ids))
        return fail(DKM_ERR_INVALID_ARG, "bad sizes");
    if (nq == 0 || m == 0) return DKM_OK;
    for (int64_t i = 0; i < nq * m; ++i)
        if (ids[i] >= (uint64_t)n) return fail(DKM_ERR_INDEX, "id %u out of range (%lld rows)", ids[i], (long long)n);
    const dkm::SphParams &t = q->t->q;
    const int rb = 20 + t.out_dim;
    HIP_TRY(hipSetDevice(g_device));
    DevBuf<float> dq, dout;
    DevBuf<uint8_t> dc;
    DevBuf<uint32_t> di, derr;
    HIP_TRY(dq.ensure(nq * t.dim));
    HIP_TRY(dc.ensure(n * rb));
    HIP_TRY(di.ensure(nq * m));
    HIP_TRY(dout.ensure(nq * m));
    HIP_TRY(derr.ensure(1));
    HIP_TRY(hipMemcpy(dq.p, queries, sizeof(float) * nq * t.dim, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(dc.p, codes, (size_t)(n * rb), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(di.p, ids, sizeof(uint32_t) * nq * m, hipMemcpyHostToDevice));
    HIP_TRY(hipMemset(derr.p, 0, sizeof(uint32_t)));
    HIP_TRY(dkm::launch_minmax_query_pairs(mm, t, dq.p, nq, dc.p, rb, di.p, m, dout.p, derr.p, nullptr));
    HIP_TRY(hipMemcpy(out, dout.p, sizeof(float) * nq * m, hipMemcpyDeviceToHost));
    uint32_t e = 0;
    HIP_TRY(hipMemcpy(&e, derr.p, sizeof(uint32_t), hipMemcpyDeviceToHost));
    if (e) return fail(DKM_ERR_INVALID_ARG, "query contains NaN");
    return DKM_OK;
}

// Default LDS visited-set size: a small first level (most of the visited set lives in the
// per-query HBM table); small LDS footprints keep more queries resident per CU, which
// hides the gather latency better than a large LDS table does (measured on MI355X).
static uint32_t default_visited_slots(uint64_t L) {
    (void)L;
    return 512;
}

// Slots of one query's HBM visited region for searches up to list size L: a power of two
// of at least 16 L (a query visits about 8 L ids at these graph degrees; the table stays
// near half full) and at least 16384. A query that still fills it reports
// DKM_ERR_CAPACITY.
static uint64_t region_slots(uint64_t L) {
    uint64_t s = 16384;
    while (s < 16 * L) s <<= 1;
    return s;
}

// Per-query visited-set regions (HBM) for `nq` queries of `slots` entries each; a query
// clears its own region when it starts using it.
static dkm_status ensure_spill(dkm_index *ix, int64_t nq, uint64_t slots) {
    if (ix->spill_regions < nq || ix->spill_slots < slots) {
        if (ix->d_spill) HIP_TRY(hipFree(ix->d_spill));
        ix->d_spill = nullptr;
        ix->spill_regions = 0;
        const uint64_t s = std::max(slots, ix->spill_slots);
        HIP_TRY(hipMalloc(&ix->d_spill, sizeof(uint32_t) * s * (size_t)nq));
        ix->spill_regions = nq;
        ix->spill_slots = s;
    }
    return DKM_OK;
}

// Scalar-quantized and spherical traversal fields of the search parameters.
static void set_sq_params(const dkm_index *ix, bool sq, dkm::SearchParams &p, bool sph = false) {
    p.sph = sph;
    p.sph_layout = ix->sph_layout;
    p.sph_rescale = 1;
    p.sphq = ix->sphq;
    p.sph_bits = ix->d_sph_bits.p;
    p.sph_meta = ix->d_sph_meta.p;
    p.sq = sq;
    p.sq_codes = ix->d_sq_codes.p;
    p.sq_comp = ix->d_sq_comp.p;
    p.sq_shift = ix->d_sq_shift.p;
    p.sq_cstride = (int)ix->sq_cstride;
    p.sq_scale = ix->sq_scale;
    p.sq_bss = dkm::sq_bss(ix->sq_scale);
    p.sq_ssn = ix->sq_ssn;
    p.sq_mean_norm = ix->sq_mean_norm;
}

// Per-launch device buffers a caller owns (a captured graph) instead of the index's.
struct OwnBufs {
    uint32_t *spill = nullptr;  // nq regions of region_slots(max_L) slots
    uint64_t spill_slots = 0;
    uint8_t *sph_qrec = nullptr;  // spherical: nq prepared query records
    uint32_t *order = nullptr;    // launches of kOrderMin queries or more: order + bins
};

// Launches of at least this many queries run them longest L first (below it the device
// holds all of them at once and the order does not matter).
constexpr int64_t kOrderMin = 2048;
static size_t order_words(int64_t nq) { return (size_t)nq + dkm::kOrderBins + 1; }

// Validates a batched search and builds its launch parameters (nq >= 1); the visited-set
// regions and spherical query records are the index's (grown as needed) unless `own`
// supplies them. Caller holds ix->mu.
static dkm_status search_params(dkm_index *ix, int strategy, const float *d_queries, int64_t nq,
                                const uint32_t *d_k, const uint32_t *d_L, uint32_t max_k,
                                uint32_t max_L, uint32_t bw, int64_t k_stride, uint32_t *d_ids,
                                float *d_d, uint32_t *d_cnt, uint32_t *d_cmps, uint32_t *d_hops,
                                uint32_t *d_status, uint32_t *d_degsum, const OwnBufs *own,
                                dkm::SearchParams &p, bool disk = false) {
    if (strategy < 0 || strategy > 2) return fail(DKM_ERR_INVALID_ARG, "unknown strategy %d", strategy);
    if (bw == 0) return fail(DKM_ERR_BEAM_WIDTH_ZERO, "beam width must be >= 1");
    if (bw > 64) return fail(DKM_ERR_INVALID_ARG, "beam width above 64 is not supported");
    if (max_L == 0) return fail(DKM_ERR_L_ZERO, "L must be >= 1");
    if (max_L > 4096) return fail(DKM_ERR_INVALID_ARG, "L above 4096 is not supported");
    if ((int64_t)max_k > k_stride) return fail(DKM_ERR_INVALID_ARG, "k exceeds output stride");
    if (!ix->starts_set) return fail(DKM_ERR_NOT_READY, "start points not set");
    const bool quant = strategy != DKM_STRATEGY_FULL;
    const bool sph = quant && ix->sph_active;
    const bool sq = quant && ix->sq_active;
    const bool pq = quant && !sq && !sph;
    if (sph && !ix->sph_codes_set)
        return fail(DKM_ERR_NOT_READY, "spherical strategy needs the codes of every row");
    if (sq && !ix->sq_codes_set)
        return fail(DKM_ERR_NOT_READY, "scalar-quantized strategy needs the codes of every row");
    if (pq && (!ix->d_codes || !ix->codes_set))
        return fail(DKM_ERR_NOT_READY, "PQ strategy needs a PQ schema and codes");
    // (the disk-index search traverses cosine indexes with the L2 table)
    if (pq && ix->metric == DKM_COSINE && !disk)
        return fail(DKM_ERR_PQ, "cosine PQ traversal (direct distances) is not supported");
    if (nq == 0) return DKM_OK;  // validated; nothing to launch
    if (!own) {
        dkm_status st = ensure_spill(ix, nq, std::max<uint64_t>(region_slots(max_L), ix->tune_spill));
        if (st) return st;
    }
    memset(&p, 0, sizeof(p));  // padding too: captured parameters are compared bytewise
    p.metric = ix->metric;
    p.pq = pq;
    p.rerank = strategy == DKM_STRATEGY_HYBRID;
    p.adc_mode = ix->tune_adc_mode;
    p.pq_len = pq ? ix->pq_len : 0;
    p.dim = (int)ix->dim;
    p.vstride = ix->vstride;
    p.vectors = ix->d_vec;
    p.graph = ix->d_graph;
    p.gstride = (int)ix->gstride;
    p.starts = ix->d_starts;
    p.nstart = (int)ix->num_start;
    p.codes = ix->d_codes;
    p.code_stride = (int)ix->M;
    p.M = (int)ix->M;
    p.pivots = ix->d_pivots;
    p.pivots_t = ix->d_pivots_t;
    p.offsets = ix->d_offsets;
    set_sq_params(ix, sq, p, sph);
    if (sph && own) {
        p.sph_qrec = own->sph_qrec;
    } else if (sph) {
        HIP_TRY(ix->d_sph_qrec.ensure(nq * dkm::sph_query_record_bytes(ix->sphq)));
        p.sph_qrec = ix->d_sph_qrec.p;
    }
    p.queries = d_queries;
    p.qstride = ix->dim;
    p.nq = (int)nq;
    p.k = d_k;
    p.L = d_L;
    p.beam_width = bw;
    p.k_stride = k_stride;
    p.out_ids = d_ids;
    p.out_d = d_d;
    p.out_count = d_cnt;
    p.out_cmps = d_cmps;
    p.out_hops = d_hops;
    p.out_degsum = d_degsum;
    p.status = d_status;
    p.qcap = (int)(max_L + ix->num_start);
    p.nlcap = (int)(bw * ix->R_max + 64);
    p.hslots = ix->tune_hslots > 0 ? (uint32_t)ix->tune_hslots : default_visited_slots(max_L);
    p.spill_table = own ? own->spill : ix->d_spill;
    // a capacity override (dkm_set_visited_capacity) lays the regions out at its own stride,
    // inside an allocation sized for at least that many slots per query
    p.spill_slots = own ? own->spill_slots : (ix->tune_spill ? ix->tune_spill : ix->spill_slots);
    if (nq >= kOrderMin) {
        if (own) {
            p.order = own->order;
        } else {
            HIP_TRY(ix->d_order.ensure((int64_t)order_words(nq)));
            p.order = ix->d_order.p;
        }
    }
    p.max_iters = (uint32_t)std::min<int64_t>(ix->rows + 16, 0x7FFFFFFF);
    size_t lds = dkm::search_lds_bytes(p);
    if (lds > 160 * 1024)
        return fail(DKM_ERR_CAPACITY, "per-query LDS need %zu bytes exceeds 160 KiB", lds);
    return DKM_OK;
}

// Launch on the handle's stream; caller holds ix->mu.
static dkm_status launch_locked(dkm_index *ix, int strategy, const float *d_queries, int64_t nq,
                                const uint32_t *d_k, const uint32_t *d_L, uint32_t max_k,
                                uint32_t max_L, uint32_t bw, int64_t k_stride, uint32_t *d_ids,
                                float *d_d, uint32_t *d_cnt, uint32_t *d_cmps, uint32_t *d_hops,
                                uint32_t *d_status, uint32_t *d_degsum = nullptr) {
    dkm::SearchParams p;
    dkm_status st = search_params(ix, strategy, d_queries, nq, d_k, d_L, max_k, max_L, bw,
                                  k_stride, d_ids, d_d, d_cnt, d_cmps, d_hops, d_status, d_degsum,
                                  nullptr, p);
    if (st || nq == 0) return st;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (ix->profile) {
        HIP_TRY(hipEventCreate(&e0));
        HIP_TRY(hipEventCreate(&e1));
        HIP_TRY(hipEventRecord(e0, ix->stream));
    }
    hipError_t e = dkm::launch_search(p, ix->stream);
    if (e != hipSuccess) return fail(DKM_ERR_DEVICE, "search launch: %s", hipGetErrorString(e));
    if (ix->profile) {
        HIP_TRY(hipEventRecord(e1, ix->stream));
        ix->events.push_back({e0, e1});
    }
    return DKM_OK;
}

// Maps the per-query status words of a launch to the first failing query's error
// (the checks dkm_search_batch makes after its copy-back). *first = -1 when none failed.
static dkm_status status_words(const uint32_t *status, int64_t nq, int64_t *first) {
    if (first) *first = -1;
    for (int64_t i = 0; i < nq; ++i) {
        if (!status[i]) continue;
        if (first) *first = i;
        if (status[i] & dkm::ERR_BAD_REQUEST)
            return fail(DKM_ERR_INVALID_ARG,
                        "query %lld: k / L outside Knn::new's rules or the launch's max_k / max_L",
                        (long long)i);
        if (status[i] & dkm::ERR_VISITED_OVERFLOW)
            return fail(DKM_ERR_CAPACITY, "query %lld: visited set overflowed its LDS and spill capacity",
                        (long long)i);
        return fail(DKM_ERR_INDEX, "query %lld: traversal exceeded the iteration bound", (long long)i);
    }
    return DKM_OK;
}

dkm_status dkm_search_batch_device(dkm_index *ix, int strategy, const float *d_queries,
                                   int64_t nq, const uint32_t *d_k, const uint32_t *d_L,
                                   uint32_t max_k, uint32_t max_L, uint32_t beam_width,
                                   int64_t k_stride, uint32_t *d_out_ids, float *d_out_dists,
                                   uint32_t *d_out_counts, uint32_t *d_out_cmps,
                                   uint32_t *d_out_hops) {
    if (!ix) return fail(DKM_ERR_INVALID_ARG, "null index");
    std::lock_guard<std::mutex> g(ix->mu);
    HIP_TRY(hipSetDevice(ix->device));
    HIP_TRY(ix->s_status.ensure(std::max<int64_t>(nq, 1)));
    ix->status_nq = nq;
    return launch_locked(ix, strategy, d_queries, nq, d_k, d_L, max_k, max_L, beam_width, k_stride,
                         d_out_ids, d_out_dists, d_out_counts, d_out_cmps, d_out_hops,
                         ix->s_status.p);
}

dkm_status dkm_search_batch_device_stats(dkm_index *ix, int strategy, const float *d_queries,
                                         int64_t nq, const uint32_t *d_k, const uint32_t *d_L,
                                         uint32_t max_k, uint32_t max_L, uint32_t beam_width,
                                         int64_t k_stride, uint32_t *d_out_ids,
                                         float *d_out_dists, uint32_t *d_out_counts,
                                         uint32_t *d_out_cmps, uint32_t *d_out_hops,
                                         uint32_t *d_out_degsum) {
    if (!ix) return fail(DKM_ERR_INVALID_ARG, "null index");
    std::lock_guard<std::mutex> g(ix->mu);
    HIP_TRY(hipSetDevice(ix->device));
    HIP_TRY(ix->s_status.ensure(std::max<int64_t>(nq, 1)));
    ix->status_nq = nq;
    return launch_locked(ix, strategy, d_queries, nq, d_k, d_L, max_k, max_L, beam_width, k_stride,
                         d_out_ids, d_out_dists, d_out_counts, d_out_cmps, d_out_hops,
                         ix->s_status.p, d_out_degsum);
}

dkm_status dkm_search_status(dkm_index *ix, int64_t *first_failed) {
    if (!ix) return fail(DKM_ERR_INVALID_ARG, "null index");
    if (first_failed) *first_failed = -1;
    std::lock_guard<std::mutex> g(ix->mu);
    HIP_TRY(hipSetDevice(ix->device));
    HIP_TRY(hipStreamSynchronize(ix->stream));
    if (ix->status_nq <= 0) return DKM_OK;
    std::vector<uint32_t> status((size_t)ix->status_nq);
    HIP_TRY(hipMemcpy(status.data(), ix->s_status.p, sizeof(uint32_t) * status.size(),
                      hipMemcpyDeviceToHost));
    return status_words(status.data(), ix->status_nq, first_failed);
}

}  // extern "C"

// ------------------------------------------------------------------------------------
// Allocation-free batched search: the launch's scratch (status words when the caller
// gives none, visited-set regions, the longest-L-first order, spherical query records)
// lives in this file-scope device array. Each GPU holds its own copy, at its own address:
// host code uses the address hipGetSymbolAddress returns for the current device, never the
// array's name (a host-side stand-in).
// ------------------------------------------------------------------------------------
constexpr size_t kScratchBytes = size_t(64) << 20;
static __device__ __align__(256) uint8_t scratch[kScratchBytes];

static std::mutex g_scratch_mu;
static uint8_t *g_scratch_ptr[64] = {};

// The scratch address on `device` (which must be current); looked up once per device.
static dkm_status scratch_ptr(int device, uint8_t **out) {
    if (device < 0 || device >= 64) return fail(DKM_ERR_INVALID_ARG, "device %d out of range", device);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void *p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(DKM_ERR_DEVICE, "scratch symbol address: %s", hipGetErrorString(e));
        g_scratch_ptr[device] = (uint8_t *)p;
    }
    *out = g_scratch_ptr[device];
    return DKM_OK;
}

static size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

struct dkm_search_graph {
    dkm_index *ix = nullptr;
    int strategy = 0;
    const float *queries = nullptr;
    int64_t nq = 0, k_stride = 0;
    const uint32_t *k = nullptr, *L = nullptr;
    uint32_t max_k = 0, max_L = 0, bw = 0;
    uint32_t *ids = nullptr, *cnt = nullptr, *cmps = nullptr, *hops = nullptr, *degsum = nullptr;
    float *dists = nullptr;
    DevBuf<uint32_t> spill, status, order;
    DevBuf<uint8_t> qrec;
    OwnBufs own;
    dkm::SearchParams p{};  // what was captured
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

extern "C" {

void dkm_search_graph_destroy(dkm_search_graph *g) {
    if (!g) return;
    (void)hipSetDevice(g->ix->device);
    (void)hipStreamSynchronize(g->ix->stream);
    if (g->exec) (void)hipGraphExecDestroy(g->exec);
    if (g->graph) (void)hipGraphDestroy(g->graph);
    delete g;
}

dkm_status dkm_search_graph_create(dkm_index *ix, int strategy, const float *d_queries,
                                   int64_t nq, const uint32_t *d_k, const uint32_t *d_L,
                                   uint32_t max_k, uint32_t max_L, uint32_t beam_width,
                                   int64_t k_stride, uint32_t *d_out_ids, float *d_out_dists,
                                   uint32_t *d_out_counts, uint32_t *d_out_cmps,
                                   uint32_t *d_out_hops, uint32_t *d_out_degsum,
                                   dkm_search_graph **out) {
    if (!ix || !out) return fail(DKM_ERR_INVALID_ARG, "null argument");
    *out = nullptr;
    if (nq <= 0) return fail(DKM_ERR_INVALID_ARG, "a search graph needs at least one query");
    std::lock_guard<std::mutex> lk(ix->mu);
    HIP_TRY(hipSetDevice(ix->device));
    auto *g = new dkm_search_graph();
    g->ix = ix;
    g->strategy = strategy;
    g->queries = d_queries;
    g->nq = nq;
    g->k = d_k;
    g->L = d_L;
    g->max_k = max_k;
    g->max_L = max_L == 0 ? 1 : max_L;
    g->bw = beam_width;
    g->k_stride = k_stride;
    g->ids = d_out_ids;
    g->dists = d_out_dists;
    g->cnt = d_out_counts;
    g->cmps = d_out_cmps;
    g->hops = d_out_hops;
    g->degsum = d_out_degsum;
    auto bad = [&](dkm_status st) {
        delete g;
        return st;
    };
    hipError_t e;
    // the graph's own visited-set regions, status words and query records: later launches
    // through the index may reallocate the index's, never these
    g->own.spill_slots = ix->tune_spill ? ix->tune_spill : region_slots(g->max_L);
    if ((e = g->spill.ensure((int64_t)(g->own.spill_slots * (uint64_t)nq))) != hipSuccess ||
        (e = g->status.ensure(nq)) != hipSuccess)
        return bad(fail(e == hipErrorOutOfMemory ? DKM_ERR_OUT_OF_MEMORY : DKM_ERR_DEVICE,
                        "search graph buffers: %s", hipGetErrorString(e)));
    g->own.spill = g->spill.p;
    if (nq >= kOrderMin) {
        if ((e = g->order.ensure((int64_t)order_words(nq))) != hipSuccess)
            return bad(fail(DKM_ERR_DEVICE, "search graph buffers: %s", hipGetErrorString(e)));
        g->own.order = g->order.p;
    }
    if (ix->sph_active) {
        if ((e = g->qrec.ensure(nq * dkm::sph_query_record_bytes(ix->sphq))) != hipSuccess)
            return bad(fail(DKM_ERR_DEVICE, "search graph buffers: %s", hipGetErrorString(e)));
        g->own.sph_qrec = g->qrec.p;
    }
    dkm_status st = search_params(ix, strategy, d_queries, nq, d_k, d_L, max_k, max_L,
                                  beam_width, k_stride, d_out_ids, d_out_dists, d_out_counts,
                                  d_out_cmps, d_out_hops, g->status.p, d_out_degsum, &g->own, g->p);
    if (st) return bad(st);
    // capture the launch (kernels only: no memset or copy nodes) on the index's stream
    if ((e = hipStreamSynchronize(ix->stream)) != hipSuccess ||
        (e = hipStreamBeginCapture(ix->stream, hipStreamCaptureModeThreadLocal)) != hipSuccess)
        return bad(fail(DKM_ERR_DEVICE, "graph capture: %s", hipGetErrorString(e)));
    hipError_t le = dkm::launch_search(g->p, ix->stream);
    e = hipStreamEndCapture(ix->stream, &g->graph);
    if (le != hipSuccess || e != hipSuccess)
        return bad(fail(DKM_ERR_DEVICE, "graph capture: %s",
                        hipGetErrorString(le != hipSuccess ? le : e)));
    if ((e = hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0)) != hipSuccess)
        return bad(fail(DKM_ERR_DEVICE, "graph instantiate: %s", hipGetErrorString(e)));
    *out = g;
    return DKM_OK;
}

dkm_status dkm_search_graph_launch(dkm_search_graph *g) {
    if (!g) return fail(DKM_ERR_INVALID_ARG, "null graph");
    dkm_index *ix = g->ix;
    std::lock_guard<std::mutex> lk(ix->mu);
    HIP_TRY(hipSetDevice(ix->device));
    // the captured launch is still the one the index would make: same buffers, tuning and
    // quantizer (a changed index needs a new graph)
    dkm::SearchParams now;
    dkm_status st = search_params(ix, g->strategy, g->queries, g->nq, g->k, g->L, g->max_k,
                                  g->max_L, g->bw, g->k_stride, g->ids, g->dists, g->cnt, g->cmps,
                                  g->hops, g->status.p, g->degsum, &g->own, now);
    if (st) return st;
    if (memcmp(&now, &g->p, sizeof(now)) != 0)
        return fail(DKM_ERR_NOT_READY,
                    "the index changed since the search graph was captured; create a new one");
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (ix->profile) {
        HIP_TRY(hipEventCreate(&e0));
        HIP_TRY(hipEventCreate(&e1));
        HIP_TRY(hipEventRecord(e0, ix->stream));
    }
    HIP_TRY(hipGraphLaunch(g->exec, ix->stream));
    if (ix->profile) {
        HIP_TRY(hipEventRecord(e1, ix->stream));
        ix->events.push_back({e0, e1});
    }
    return DKM_OK;
}

dkm_status dkm_search_graph_status(dkm_search_graph *g, int64_t *first_failed) {
    if (!g) return fail(DKM_ERR_INVALID_ARG, "null graph");
    if (first_failed) *first_failed = -1;
    dkm_index *ix = g->ix;
    std::lock_guard<std::mutex> lk(ix->mu);
    HIP_TRY(hipSetDevice(ix->device));
    HIP_TRY(hipStreamSynchronize(ix->stream));
    std::vector<uint32_t> status((size_t)g->nq);
    HIP_TRY(hipMemcpy(status.data(), g->status.p, sizeof(uint32_t) * status.size(),
                      hipMemcpyDeviceToHost));
    return status_words(status.data(), g->nq, first_failed);
}

dkm_status dkm_search_batch_device_noalloc(dkm_index *ix, int strategy, const float *d_queries,
                                           int64_t nq, const uint32_t *d_k, const uint32_t *d_L,
                                           uint32_t max_k, uint32_t max_L, uint32_t beam_width,
                                           int64_t k_stride, uint32_t *d_out_ids,
                                           float *d_out_dists, uint32_t *d_out_counts,
                                           uint32_t *d_out_cmps, uint32_t *d_out_hops,
                                           uint32_t *d_out_status) {
    if (!ix) return fail(DKM_ERR_INVALID_ARG, "null index");
    if (nq < 0) return fail(DKM_ERR_INVALID_ARG, "negative query count");
    std::lock_guard<std::mutex> g(ix->mu);
    HIP_TRY(hipSetDevice(ix->device));
    uint8_t *base = nullptr;
    dkm_status st = scratch_ptr(ix->device, &base);
    if (st) return st;
    // scratch layout: [status words][visited regions][order + bins][spherical records]
    const uint64_t slots = ix->tune_spill ? ix->tune_spill : region_slots(max_L == 0 ? 1 : max_L);
    const size_t status_b = d_out_status ? 0 : align256(sizeof(uint32_t) * (size_t)nq);
    const size_t spill_b = align256(sizeof(uint32_t) * slots * (size_t)nq);
    const size_t order_b = nq >= kOrderMin ? align256(sizeof(uint32_t) * order_words(nq)) : 0;
    const size_t sph_b =
        ix->sph_active ? align256((size_t)nq * dkm::sph_query_record_bytes(ix->sphq)) : 0;
    const size_t need = status_b + spill_b + order_b + sph_b;
    if (need > kScratchBytes)
        return fail(DKM_ERR_CAPACITY,
                    "%lld queries at max L %u need %zu bytes of scratch; the allocation-free "
                    "search holds %zu (use fewer queries per call)",
                    (long long)nq, max_L, need, kScratchBytes);
    OwnBufs own;
    uint32_t *status = d_out_status ? d_out_status : (uint32_t *)base;
    own.spill = (uint32_t *)(base + status_b);
    own.spill_slots = slots;
    own.order = order_b ? (uint32_t *)(base + status_b + spill_b) : nullptr;
    own.sph_qrec = sph_b ? base + status_b + spill_b + order_b : nullptr;
    dkm::SearchParams p;
    st = search_params(ix, strategy, d_queries, nq, d_k, d_L, max_k, max_L, beam_width, k_stride,
                       d_out_ids, d_out_dists, d_out_counts, d_out_cmps, d_out_hops, status,
                       nullptr, &own, p);
    if (st || nq == 0) return st;
    hipError_t e = dkm::launch_search(p, ix->stream);
    if (e != hipSuccess) return fail(DKM_ERR_DEVICE, "search launch: %s", hipGetErrorString(e));
    return DKM_OK;
}

dkm_status dkm_medoid(dkm_index *ix, float *out_vector, int64_t *out_row) {
    if (!ix || !out_vector || !out_row) return fail(DKM_ERR_INVALID_ARG, "null argument");
    std::lock_guard<std::mutex> g(ix->mu);
    HIP_TRY(hipSetDevice(ix->device));
    float *dmean;
    unsigned long long *dbest;
    HIP_TRY(hipMalloc(&dmean, sizeof(float) * ix->dim));
    HIP_TRY(hipMalloc(&dbest, sizeof(unsigned long long)));
    unsigned long long init = ~0ull;
    HIP_TRY(hipMemcpy(dbest, &init, sizeof(init), hipMemcpyHostToDevice));
    HIP_TRY(dkm::medoid_device(ix->d_vec, ix->max_points, ix->vstride, (int)ix->dim, dmean, dbest,
                               ix->stream));
    HIP_TRY(hipStreamSynchronize(ix->stream));
    unsigned long long best = 0;
    HIP_TRY(hipMemcpy(&best, dbest, sizeof(best), hipMemcpyDeviceToHost));
    (void)hipFree(dmean);
    (void)hipFree(dbest);
    if (best == ~0ull) {
        *out_row = -1;
        for (int64_t j = 0; j < ix->dim; ++j) out_vector[j] = 0.0f;
        return DKM_OK;
    }
    *out_row = (int64_t)(best & 0xFFFFFFFFull);
    HIP_TRY(hipMemcpy(out_vector, ix->d_vec + *out_row * ix->vstride, sizeof(float) * ix->dim,
                      hipMemcpyDeviceToHost));
    return DKM_OK;
}

dkm_status dkm_closest_centers(const float *data, int64_t n, int64_t dim, const float *centers,
                               int64_t K, int64_t k, uint32_t *out) {
    if (n < 0 || dim < 1 || K < 1) return fail(DKM_ERR_INVALID_ARG, "bad sizes");
    if (k < 1 || k > K || k > dkm::kMaxClosest)
        return fail(DKM_ERR_INDEX, "k (%lld) must be in [1, min(num_centers, %d)]", (long long)k,
                    dkm::kMaxClosest);
    if (n > 0 && (!data || !centers || !out)) return fail(DKM_ERR_INVALID_ARG, "null argument");
    if (n == 0) return DKM_OK;
    DevBuf<float> dx, dc, dn;
    DevBuf<uint32_t> dout;
    const int64_t chunk = std::min<int64_t>(n, std::max<int64_t>(1, (int64_t(1) << 30) / (4 * dim)));
    HIP_TRY(dx.ensure((size_t)chunk * dim));
    HIP_TRY(dc.ensure((size_t)K * dim));
    HIP_TRY(dn.ensure(K));
    HIP_TRY(dout.ensure((size_t)chunk * k));
    HIP_TRY(hipMemcpy(dc.p, centers, sizeof(float) * K * dim, hipMemcpyHostToDevice));
    for (int64_t r0 = 0; r0 < n; r0 += chunk) {
        const int64_t m = std::min(chunk, n - r0);
        HIP_TRY(hipMemcpy(dx.p, data + r0 * dim, sizeof(float) * m * dim, hipMemcpyHostToDevice));
        HIP_TRY(dkm::launch_closest_centers(dx.p, m, (int)dim, dc.p, (int)K, (int)k, dn.p, dout.p,
                                            nullptr));
        HIP_TRY(hipMemcpy(out + r0 * k, dout.p, sizeof(uint32_t) * m * k, hipMemcpyDeviceToHost));
    }
    return DKM_OK;
}

dkm_status dkm_bruteforce(dkm_index *ix, const float *queries, int64_t nq, int64_t k,
                          uint32_t *out_ids, double *out_dists) {
    if (!ix || (nq > 0 && (!queries || !out_ids || !out_dists)))
        return fail(DKM_ERR_INVALID_ARG, "null argument");
    if (k <= 0 || k > 1024) return fail(DKM_ERR_INVALID_ARG, "k must be in [1, 1024]");
    if (nq == 0) return DKM_OK;
    std::lock_guard<std::mutex> g(ix->mu);
    HIP_TRY(hipSetDevice(ix->device));
    float *dq;
    uint32_t *di;
    double *dd;
    HIP_TRY(hipMalloc(&dq, sizeof(float) * nq * ix->dim));
    HIP_TRY(hipMalloc(&di, sizeof(uint32_t) * nq * k));
    HIP_TRY(hipMalloc(&dd, sizeof(double) * nq * k));
    HIP_TRY(hipMemcpy(dq, queries, sizeof(float) * nq * ix->dim, hipMemcpyHostToDevice));
    HIP_TRY(dkm::bruteforce_device(ix->metric, ix->d_vec, ix->max_points, ix->vstride, (int)ix->dim,
                                   dq, nq, (int)k, di, dd, ix->stream));
    HIP_TRY(hipStreamSynchronize(ix->stream));
    HIP_TRY(hipMemcpy(out_ids, di, sizeof(uint32_t) * nq * k, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(out_dists, dd, sizeof(double) * nq * k, hipMemcpyDeviceToHost));
    (void)hipFree(dq);
    (void)hipFree(di);
    (void)hipFree(dd);
    return DKM_OK;
}

dkm_status dkm_search_batch(dkm_index *ix, int strategy, const float *queries, int64_t nq,
                            const uint32_t *k, const uint32_t *L, uint32_t beam_width,
                            int64_t k_stride, uint32_t *out_ids, float *out_dists,
                            uint32_t *out_counts, uint32_t *out_cmps, uint32_t *out_hops) {
    if (!ix) return fail(DKM_ERR_INVALID_ARG, "null index");
    if (nq < 0) return fail(DKM_ERR_INVALID_ARG, "negative query count");
    if (nq == 0) return DKM_OK;
    if (!queries || !k || !L || !out_ids || !out_dists || !out_counts)
        return fail(DKM_ERR_INVALID_ARG, "null argument");
    uint32_t max_k = 0, max_L = 0;
    for (int64_t i = 0; i < nq; ++i) {
        // Knn::new validation order: k, L, k <= L, beam width
        if (k[i] == 0) return fail(DKM_ERR_K_ZERO, "query %lld: k must be >= 1", (long long)i);
        if (L[i] == 0) return fail(DKM_ERR_L_ZERO, "query %lld: L must be >= 1", (long long)i);
        if (k[i] > L[i])
            return fail(DKM_ERR_L_LESS_THAN_K, "query %lld: L (%u) < k (%u)", (long long)i, L[i], k[i]);
        max_k = std::max(max_k, k[i]);
        max_L = std::max(max_L, L[i]);
    }
    if (beam_width == 0) return fail(DKM_ERR_BEAM_WIDTH_ZERO, "beam width must be >= 1");
    std::lock_guard<std::mutex> g(ix->mu);
    HIP_TRY(hipSetDevice(ix->device));
    HIP_TRY(ix->s_q.ensure(nq * ix->dim));
    HIP_TRY(ix->s_k.ensure(nq));
    HIP_TRY(ix->s_L.ensure(nq));
    HIP_TRY(ix->s_ids.ensure(nq * k_stride));
    HIP_TRY(ix->s_d.ensure(nq * k_stride));
    HIP_TRY(ix->s_cnt.ensure(nq));
    HIP_TRY(ix->s_cmps.ensure(nq));
    HIP_TRY(ix->s_hops.ensure(nq));
    HIP_TRY(ix->s_status.ensure(nq));
    HIP_TRY(hipMemcpyAsync(ix->s_q.p, queries, sizeof(float) * nq * ix->dim, hipMemcpyHostToDevice, ix->stream));
    HIP_TRY(hipMemcpyAsync(ix->s_k.p, k, sizeof(uint32_t) * nq, hipMemcpyHostToDevice, ix->stream));
    HIP_TRY(hipMemcpyAsync(ix->s_L.p, L, sizeof(uint32_t) * nq, hipMemcpyHostToDevice, ix->stream));
    dkm_status st = launch_locked(ix, strategy, ix->s_q.p, nq, ix->s_k.p, ix->s_L.p, max_k, max_L,
                                  beam_width, k_stride, ix->s_ids.p, ix->s_d.p, ix->s_cnt.p,
                                  ix->s_cmps.p, ix->s_hops.p, ix->s_status.p);
    if (st) return st;
    std::vector<uint32_t> status(nq);
    HIP_TRY(hipMemcpyAsync(out_ids, ix->s_ids.p, sizeof(uint32_t) * nq * k_stride, hipMemcpyDeviceToHost, ix->stream));
    HIP_TRY(hipMemcpyAsync(out_dists, ix->s_d.p, sizeof(float) * nq * k_stride, hipMemcpyDeviceToHost, ix->stream));
    HIP_TRY(hipMemcpyAsync(out_counts, ix->s_cnt.p, sizeof(uint32_t) * nq, hipMemcpyDeviceToHost, ix->stream));
    if (out_cmps)
        HIP_TRY(hipMemcpyAsync(out_cmps, ix->s_cmps.p, sizeof(uint32_t) * nq, hipMemcpyDeviceToHost, ix->stream));
    if (out_hops)
        HIP_TRY(hipMemcpyAsync(out_hops, ix->s_hops.p, sizeof(uint32_t) * nq, hipMemcpyDeviceToHost, ix->stream));
    HIP_TRY(hipMemcpyAsync(status.data(), ix->s_status.p, sizeof(uint32_t) * nq, hipMemcpyDeviceToHost, ix->stream));
    HIP_TRY(hipStreamSynchronize(ix->stream));
    ix->status_nq = 0;  // the staging status words now belong to this host call
    return status_words(status.data(), nq, nullptr);
}


dkm_status dkm_disk_search_batch(dkm_index *ix, const float *queries, int64_t nq,
                                 const uint32_t *k, const uint32_t *L, uint32_t beam_width,
                                 uint32_t medoid, uint64_t io_limit, int64_t k_stride,
                                 uint32_t *out_ids, float *out_dists, uint32_t *out_counts,
                                 uint32_t *out_cmps, uint32_t *out_hops) {
    if (!ix) return fail(DKM_ERR_INVALID_ARG, "null index");
    if (nq < 0) return fail(DKM_ERR_INVALID_ARG, "negative query count");
    if (nq == 0) return DKM_OK;
    if (!queries || !k || !L || !out_ids || !out_dists || !out_counts)
        return fail(DKM_ERR_INVALID_ARG, "null argument");
    if (ix->sq_active || ix->sph_active)
        return fail(DKM_ERR_NOT_READY, "disk-index search traverses with PQ codes (use_pq)");
    if ((int64_t)medoid >= ix->max_points)
        return fail(DKM_ERR_INDEX, "medoid %u is not a data row", medoid);
    uint32_t max_k = 0, max_L = 0;
    for (int64_t i = 0; i < nq; ++i) {
        if (k[i] == 0) return fail(DKM_ERR_K_ZERO, "query %lld: k must be >= 1", (long long)i);
        if (L[i] == 0) return fail(DKM_ERR_L_ZERO, "query %lld: L must be >= 1", (long long)i);
        if (k[i] > L[i])
            return fail(DKM_ERR_L_LESS_THAN_K, "query %lld: L (%u) < k (%u)", (long long)i, L[i], k[i]);
        max_k = std::max(max_k, k[i]);
        max_L = std::max(max_L, L[i]);
    }
    if (beam_width == 0) return fail(DKM_ERR_BEAM_WIDTH_ZERO, "beam width must be >= 1");
    std::lock_guard<std::mutex> g(ix->mu);
    HIP_TRY(hipSetDevice(ix->device));
    HIP_TRY(ix->s_q.ensure(nq * ix->dim));
    HIP_TRY(ix->s_k.ensure(nq));
    HIP_TRY(ix->s_L.ensure(nq));
    HIP_TRY(ix->s_ids.ensure(nq * k_stride));
    HIP_TRY(ix->s_d.ensure(nq * k_stride));
    HIP_TRY(ix->s_cnt.ensure(nq));
    HIP_TRY(ix->s_cmps.ensure(nq));
    HIP_TRY(ix->s_hops.ensure(nq));
    HIP_TRY(ix->s_status.ensure(nq));
    HIP_TRY(ix->d_disk_start.ensure(1));
    HIP_TRY(hipMemcpyAsync(ix->s_q.p, queries, sizeof(float) * nq * ix->dim, hipMemcpyHostToDevice, ix->stream));
    HIP_TRY(hipMemcpyAsync(ix->s_k.p, k, sizeof(uint32_t) * nq, hipMemcpyHostToDevice, ix->stream));
    HIP_TRY(hipMemcpyAsync(ix->s_L.p, L, sizeof(uint32_t) * nq, hipMemcpyHostToDevice, ix->stream));
    HIP_TRY(hipMemcpyAsync(ix->d_disk_start.p, &medoid, sizeof(uint32_t), hipMemcpyHostToDevice, ix->stream));
    dkm::SearchParams p;
    dkm_status st = search_params(ix, DKM_STRATEGY_HYBRID, ix->s_q.p, nq, ix->s_k.p, ix->s_L.p, max_k,
                                  max_L, beam_width, k_stride, ix->s_ids.p, ix->s_d.p, ix->s_cnt.p,
                                  ix->s_cmps.p, ix->s_hops.p, ix->s_status.p, nullptr, nullptr, p,
                                  true);
    if (st) return st;
    // the DiskAccessor's single starting point (the medoid row, a data row kept in the
    // results), the transposed table's arithmetic and the load budget
    p.starts = ix->d_disk_start.p;
    p.nstart = 1;
    p.keep_starts = 1;
    p.tlut = 1;
    p.io_limit = (uint32_t)std::min<uint64_t>(io_limit, 0xFFFFFFFFu);
    p.qcap = (int)(max_L + 1);
    hipError_t e = dkm::launch_search(p, ix->stream);
    if (e != hipSuccess) return fail(DKM_ERR_DEVICE, "search launch: %s", hipGetErrorString(e));
    std::vector<uint32_t> status(nq);
    HIP_TRY(hipMemcpyAsync(out_ids, ix->s_ids.p, sizeof(uint32_t) * nq * k_stride, hipMemcpyDeviceToHost, ix->stream));
    HIP_TRY(hipMemcpyAsync(out_dists, ix->s_d.p, sizeof(float) * nq * k_stride, hipMemcpyDeviceToHost, ix->stream));
    HIP_TRY(hipMemcpyAsync(out_counts, ix->s_cnt.p, sizeof(uint32_t) * nq, hipMemcpyDeviceToHost, ix->stream));
    if (out_cmps)
        HIP_TRY(hipMemcpyAsync(out_cmps, ix->s_cmps.p, sizeof(uint32_t) * nq, hipMemcpyDeviceToHost, ix->stream));
    if (out_hops)
        HIP_TRY(hipMemcpyAsync(out_hops, ix->s_hops.p, sizeof(uint32_t) * nq, hipMemcpyDeviceToHost, ix->stream));
    HIP_TRY(hipMemcpyAsync(status.data(), ix->s_status.p, sizeof(uint32_t) * nq, hipMemcpyDeviceToHost, ix->stream));
    HIP_TRY(hipStreamSynchronize(ix->stream));
    ix->status_nq = 0;
    return status_words(status.data(), nq, nullptr);
}

// ------------------------------------------------------------------------------------
// Graph construction (multi_insert batches)
// ------------------------------------------------------------------------------------

dkm_status dkm_set_build_hybrid(dkm_index *ix, int64_t max_fp_vecs_per_prune) {
    if (!ix) return fail(DKM_ERR_INVALID_ARG, "null argument");
    std::lock_guard<std::mutex> g(ix->mu);
    ix->build_max_fp = max_fp_vecs_per_prune < 0 ? -1 : max_fp_vecs_per_prune;
    return DKM_OK;
}

dkm_status dkm_set_build_sq(dkm_index *ix, int enable) {
    if (!ix) return fail(DKM_ERR_INVALID_ARG, "null argument");
    std::lock_guard<std::mutex> g(ix->mu);
    ix->build_sq = enable != 0;
    return DKM_OK;
}

dkm_status dkm_set_build_spherical(dkm_index *ix, int enable) {
    if (!ix) return fail(DKM_ERR_INVALID_ARG, "null argument");
    std::lock_guard<std::mutex> g(ix->mu);
    ix->build_sph = enable != 0;
    return DKM_OK;
}

// The build strategy's checks and its SQ distance parameters (prunes and the insert search).
static dkm_status build_strategy(const dkm_index *ix, dkm::BuildParams &bp) {
    if (ix->build_max_fp >= 0 && (!ix->d_codes || !ix->codes_set))
        return fail(DKM_ERR_NOT_READY, "the Hybrid prune needs the PQ codes of every row");
    if ((int)ix->build_sq + (int)ix->build_sph + (int)(ix->build_max_fp >= 0) > 1)
        return fail(DKM_ERR_INVALID_ARG,
                    "the SQ build, the spherical build and the Hybrid prune exclude each other");
    if (ix->build_sq && (!ix->sq_active || !ix->sq_codes_set))
        return fail(DKM_ERR_NOT_READY, "the SQ build needs the scalar quantizer and the codes of every row");
    if (ix->build_sph && (!ix->sph_active || !ix->sph_codes_set))
        return fail(DKM_ERR_NOT_READY,
                    "the spherical build needs the spherical quantizer and the codes of every row");
    bp.sph = ix->build_sph;
    bp.sph_bits = ix->d_sph_bits.p;
    bp.sph_meta = ix->d_sph_meta.p;
    bp.sph_cstride = ix->sphq.cstride;
    bp.sph_metric = ix->sphq.metric;
    bp.sph_shift_sq = ix->sphq.shift_sq;
    bp.sph_dim = (float)ix->sphq.out_dim;
    bp.sq = ix->build_sq;
    bp.sq_codes = ix->d_sq_codes.p;
    bp.sq_comp = ix->d_sq_comp.p;
    bp.sq_cstride = (int)ix->sq_cstride;
    bp.sq_bss = dkm::sq_bss(ix->sq_scale);
    bp.sq_ssn = ix->sq_ssn;
    return DKM_OK;
}

dkm_status dkm_build(dkm_index *ix, const uint32_t *ids, int64_t n, int64_t batch,
                     int64_t intra_batch, uint32_t L_build, float alpha, int64_t max_occlusion,
                     int saturate) {
    if (!ix || (!ids && n > 0)) return fail(DKM_ERR_INVALID_ARG, "null argument");
    if (batch < 1 || batch > 2048) return fail(DKM_ERR_INVALID_ARG, "batch must be in [1, 2048]");
    if (L_build == 0 || L_build > 1024) return fail(DKM_ERR_INVALID_ARG, "L_build must be in [1, 1024]");
    if (!(alpha >= 1.0f)) return fail(DKM_ERR_INVALID_ARG, "alpha must be >= 1");
    if (max_occlusion < 1 || max_occlusion > 4096) return fail(DKM_ERR_INVALID_ARG, "bad max_occlusion");
    if (ix->R > 256) return fail(DKM_ERR_INVALID_ARG, "pruned degree above 256 is not supported");
    if (!ix->starts_set) return fail(DKM_ERR_NOT_READY, "start points not set");
    {
        dkm::BuildParams probe{};
        dkm_status st = build_strategy(ix, probe);
        if (st) return st;
    }
    for (int64_t i = 0; i < n; ++i)
        if (ids[i] >= ix->max_points) return fail(DKM_ERR_INDEX, "id %u is not a data row", ids[i]);
    if (n == 0) return DKM_OK;
    std::lock_guard<std::mutex> g(ix->mu);
    HIP_TRY(hipSetDevice(ix->device));
    const int B = (int)batch;
    const uint32_t rec_cap = std::max<uint32_t>(1024, 4 * L_build);
    // device work buffers
    DevBuf<uint32_t> bid, k1, Lb, oid, ocnt, ostat, rid, rlen, pend, pcnt, cnt, bstat, lists, lens;
    DevBuf<float> bq, od, rd;
    DevBuf<uint64_t> pairs, pairs2;
    DevBuf<char> stmp;
    HIP_TRY(bid.ensure(B));
    HIP_TRY(k1.ensure(B));
    HIP_TRY(Lb.ensure(B));
    HIP_TRY(oid.ensure(B));
    HIP_TRY(od.ensure(B));
    HIP_TRY(ocnt.ensure(B));
    HIP_TRY(ostat.ensure(B));
    HIP_TRY(bq.ensure((size_t)B * ix->dim));
    HIP_TRY(rid.ensure((size_t)B * rec_cap));
    HIP_TRY(rd.ensure((size_t)B * rec_cap));
    HIP_TRY(rlen.ensure(B));
    HIP_TRY(pend.ensure((size_t)B * ix->R));
    HIP_TRY(pcnt.ensure(B));
    HIP_TRY(cnt.ensure(2));
    HIP_TRY(bstat.ensure(1));
    HIP_TRY(pairs.ensure((size_t)B * ix->R));
    HIP_TRY(pairs2.ensure((size_t)B * ix->R));
    size_t stmp_bytes = dkm::radix_sort_tmp_bytes((uint32_t)(B * ix->R));
    HIP_TRY(stmp.ensure(stmp_bytes));
    {
        std::vector<uint32_t> ones(B, 1u), Ls(B, L_build);
        HIP_TRY(hipMemcpy(k1.p, ones.data(), sizeof(uint32_t) * B, hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(Lb.p, Ls.data(), sizeof(uint32_t) * B, hipMemcpyHostToDevice));
        uint32_t z = 0;
        HIP_TRY(hipMemcpy(bstat.p, &z, sizeof(uint32_t), hipMemcpyHostToDevice));
    }
    auto next_pow2_i = [](int64_t x) {
        int64_t q = 1;
        while (q < x) q <<= 1;
        return q;
    };
    dkm::BuildParams bp{};
    bp.metric = ix->metric;
    bp.dim = (int)ix->dim;
    bp.vstride = ix->vstride;
    bp.vectors = ix->d_vec;
    bp.graph = ix->d_graph;
    bp.gstride = (int)ix->gstride;
    bp.R = (int)ix->R;
    bp.R_max = (int)ix->R_max;
    bp.alpha = alpha;
    bp.max_occlusion = (int)max_occlusion;
    bp.saturate = saturate;
    bp.intra_batch = (int)intra_batch;
    bp.rec_ids = rid.p;
    bp.rec_d = rd.p;
    bp.rec_len = rlen.p;
    bp.rec_cap = rec_cap;
    bp.pend = pend.p;
    bp.pend_cnt = pcnt.p;
    bp.status = bstat.p;
    bp.cache_accepted = (size_t)ix->R * ix->dim * sizeof(float) <= 48 * 1024;
    bp.max_fp = ix->build_max_fp < 0 ? -1 : (int)std::min<int64_t>(ix->build_max_fp, 0x7FFFFFFF);
    bp.codes = ix->d_codes;
    bp.pivots = ix->d_pivots;
    bp.pq_off = ix->d_offsets;
    bp.pq_M = (int)ix->M;
    (void)build_strategy(ix, bp);  // checked on entry
    std::vector<uint32_t> h_pend, h_cnt;
    for (int64_t b0 = 0; b0 < n; b0 += B) {
        const int nb = (int)std::min<int64_t>(B, n - b0);
        HIP_TRY(hipMemcpyAsync(bid.p, ids + b0, sizeof(uint32_t) * nb, hipMemcpyHostToDevice, ix->stream));
        HIP_TRY(dkm::launch_gather_rows(ix->d_vec, ix->vstride, (int)ix->dim, bid.p, nb, bq.p, ix->stream));
        // 1. candidate generation: search the pre-batch graph, record expanded nodes
        {
            dkm_status st = ensure_spill(ix, nb, region_slots(L_build));
            if (st) return st;
        }
        dkm::SearchParams sp{};
        sp.metric = ix->metric;
        sp.pq = 0;
        sp.rerank = 0;
        // the SQ build's insert search is the Quantized search (SQ traversal, no rerank); the
        // spherical build's traverses with SameAsData queries and no inner-product rescale
        // (Quantized::build(), spherical.rs:515-533)
        if (ix->build_sq) set_sq_params(ix, true, sp);
        if (ix->build_sph) {
            set_sq_params(ix, false, sp, true);
            sp.sph_layout = dkm::SPH_QUERY_SAME_AS_DATA;
            sp.sph_rescale = 0;
            HIP_TRY(ix->d_sph_qrec.ensure((int64_t)nb * dkm::sph_query_record_bytes(ix->sphq)));
            sp.sph_qrec = ix->d_sph_qrec.p;
        }
        sp.dim = (int)ix->dim;
        sp.vstride = ix->vstride;
        sp.vectors = ix->d_vec;
        sp.graph = ix->d_graph;
        sp.gstride = (int)ix->gstride;
        sp.starts = ix->d_starts;
        sp.nstart = (int)ix->num_start;
        sp.queries = bq.p;
        sp.qstride = ix->dim;
        sp.nq = nb;
        sp.k = k1.p;
        sp.L = Lb.p;
        sp.beam_width = 1;
        sp.k_stride = 1;
        sp.out_ids = oid.p;
        sp.out_d = od.p;
        sp.out_count = ocnt.p;
        sp.status = ostat.p;
        sp.qcap = (int)(L_build + ix->num_start);
        sp.nlcap = (int)(ix->R_max + 64);
        sp.hslots = ix->tune_hslots > 0 ? (uint32_t)ix->tune_hslots : default_visited_slots(L_build);
        sp.spill_table = ix->d_spill;
        sp.spill_slots = ix->spill_slots;
        sp.max_iters = (uint32_t)std::min<int64_t>(ix->rows + 16, 0x7FFFFFFF);
        sp.rec_ids = rid.p;
        sp.rec_d = rd.p;
        sp.rec_cap = rec_cap;
        sp.rec_len = rlen.p;
        HIP_TRY(dkm::launch_search(sp, ix->stream));
        // 2. prune: record + intra-batch extras
        bp.batch = nb;
        bp.batch_ids = bid.p;
        int64_t want = intra_batch < 0 ? nb - 1 : std::min<int64_t>(intra_batch, nb - 1);
        bp.sort_cap = (int)next_pow2_i(std::max<int64_t>(rec_cap + std::max<int64_t>(want, 0), 2));
        if (dkm::build_prune_lds_bytes(bp) > 160 * 1024)
            return fail(DKM_ERR_CAPACITY, "prune pool for batch %d does not fit in LDS", nb);
        HIP_TRY(dkm::launch_prune_search(bp, ix->stream));
        // 3. back edges (needed for the bootstrap decision), bootstrap, set lists
        uint32_t npairs = 0, nheads = 0;
        HIP_TRY(dkm::launch_backedges(bp, pairs.p, pairs2.p, cnt.p, stmp.p, stmp_bytes, &npairs,
                                      &nheads, ix->stream));
        int64_t resolved = std::max<int64_t>(intra_batch < 0 ? nb : std::min<int64_t>(intra_batch, nb), 1);
        if (resolved < nb && ((int64_t)nheads + 7) / 8 <= nb) {
            // bootstrap: candidates = own edges + every other batch item, saturated prune
            h_pend.resize((size_t)nb * ix->R);
            h_cnt.resize(nb);
            HIP_TRY(hipMemcpyAsync(h_pend.data(), pend.p, sizeof(uint32_t) * nb * ix->R, hipMemcpyDeviceToHost, ix->stream));
            HIP_TRY(hipMemcpyAsync(h_cnt.data(), pcnt.p, sizeof(uint32_t) * nb, hipMemcpyDeviceToHost, ix->stream));
            HIP_TRY(hipStreamSynchronize(ix->stream));
            const int stride = (int)(ix->R + nb);
            std::vector<uint32_t> hl((size_t)nb * stride), hn(nb);
            for (int p = 0; p < nb; ++p) {
                uint32_t *dst = hl.data() + (size_t)p * stride;
                int c = 0;
                auto push = [&](uint32_t v) {
                    for (int j = 0; j < c; ++j)
                        if (dst[j] == v) return;
                    dst[c++] = v;
                };
                for (uint32_t j = 0; j < h_cnt[p]; ++j) push(h_pend[(size_t)p * ix->R + j]);
                for (int q2 = 0; q2 < nb; ++q2)
                    if (q2 != p) push(ids[b0 + q2]);
                hn[p] = (uint32_t)c;
            }
            HIP_TRY(lists.ensure(hl.size()));
            HIP_TRY(lens.ensure(nb));
            HIP_TRY(hipMemcpyAsync(lists.p, hl.data(), sizeof(uint32_t) * hl.size(), hipMemcpyHostToDevice, ix->stream));
            HIP_TRY(hipMemcpyAsync(lens.p, hn.data(), sizeof(uint32_t) * nb, hipMemcpyHostToDevice, ix->stream));
            dkm::BuildParams lp = bp;
            lp.sort_cap = (int)next_pow2_i(stride);
            if (dkm::build_prune_lds_bytes(lp) > 160 * 1024)
                return fail(DKM_ERR_CAPACITY, "bootstrap pool does not fit in LDS");
            HIP_TRY(dkm::launch_prune_lists(lp, lists.p, lens.p, stride, 1, ix->stream));
            HIP_TRY(dkm::launch_backedges(bp, pairs.p, pairs2.p, cnt.p, stmp.p, stmp_bytes, &npairs,
                                          &nheads, ix->stream));
        }
        HIP_TRY(dkm::launch_set_lists(bp, ix->stream));
        // 4. add_edge_and_prune per target
        if (npairs > 0) {
            dkm::BuildParams ep = bp;
            ep.sort_cap = (int)next_pow2_i(ix->R_max + nb + 1);
            if (dkm::build_prune_lds_bytes(ep) > 160 * 1024)
                return fail(DKM_ERR_CAPACITY, "back-edge pool does not fit in LDS");
            HIP_TRY(dkm::launch_apply_backedges(ep, pairs2.p, npairs, ix->stream));
        }
    }
    HIP_TRY(hipStreamSynchronize(ix->stream));
    uint32_t st = 0;
    HIP_TRY(hipMemcpy(&st, bstat.p, sizeof(uint32_t), hipMemcpyDeviceToHost));
    if (st & dkm::BUILD_ERR_RECORD) return fail(DKM_ERR_CAPACITY, "search record exceeded %u entries", rec_cap);
    if (st & dkm::BUILD_ERR_POOL) return fail(DKM_ERR_CAPACITY, "candidate pool exceeded its capacity");
    return DKM_OK;
}

}  // extern "C"

namespace {

// Prune parameters shared by dkm_prune_batch and dkm_prune_range (the index's Hybrid
// prune setting applies as in dkm_build).
dkm_status prune_setup(dkm_index *ix, float alpha, int64_t max_occlusion, int saturate,
                       int64_t list_stride, dkm::BuildParams &bp) {
    if (!(alpha >= 1.0f)) return fail(DKM_ERR_INVALID_ARG, "alpha must be >= 1");
    if (max_occlusion < 1 || max_occlusion > 4096) return fail(DKM_ERR_INVALID_ARG, "bad max_occlusion");
    if (ix->R > 256) return fail(DKM_ERR_INVALID_ARG, "pruned degree above 256 is not supported");
    if (list_stride < 1 || list_stride > 4096)
        return fail(DKM_ERR_INVALID_ARG, "candidate lists must hold 1 to 4096 entries");
    {
        dkm::BuildParams probe{};
        dkm_status st = build_strategy(ix, probe);
        if (st) return st;
    }
    int64_t cap = 2;
    while (cap < list_stride) cap <<= 1;
    bp = dkm::BuildParams{};
    bp.metric = ix->metric;
    bp.dim = (int)ix->dim;
    bp.vstride = ix->vstride;
    bp.vectors = ix->d_vec;
    bp.graph = ix->d_graph;
    bp.gstride = (int)ix->gstride;
    bp.R = (int)ix->R;
    bp.R_max = (int)ix->R_max;
    bp.alpha = alpha;
    bp.max_occlusion = (int)max_occlusion;
    bp.saturate = saturate;
    bp.sort_cap = (int)cap;
    bp.cache_accepted = (size_t)ix->R * ix->dim * sizeof(float) <= 48 * 1024;
    bp.max_fp = ix->build_max_fp < 0 ? -1 : (int)std::min<int64_t>(ix->build_max_fp, 0x7FFFFFFF);
    bp.codes = ix->d_codes;
    bp.pivots = ix->d_pivots;
    bp.pq_off = ix->d_offsets;
    bp.pq_M = (int)ix->M;
    (void)build_strategy(ix, bp);  // checked on entry
    if (dkm::build_prune_lds_bytes(bp) > 160 * 1024)
        return fail(DKM_ERR_CAPACITY, "a candidate pool of %lld ids does not fit in LDS",
                    (long long)list_stride);
    return DKM_OK;
}

// Locations per prune launch (one block each).
constexpr int64_t kPruneChunk = 4096;

}  // namespace

extern "C" {

dkm_status dkm_prune_batch(dkm_index *ix, const uint32_t *locations, int64_t n,
                           const uint32_t *lists, const uint32_t *list_len, int64_t list_stride,
                           float alpha, int64_t max_occlusion, int saturate, int force_saturate,
                           uint32_t *out_lists, uint32_t *out_counts, int64_t out_stride) {
    if (!ix || (n > 0 && (!locations || !lists || !list_len || !out_lists || !out_counts)))
        return fail(DKM_ERR_INVALID_ARG, "null argument");
    if (n < 0) return fail(DKM_ERR_INVALID_ARG, "negative count");
    if (out_stride < ix->R) return fail(DKM_ERR_INVALID_ARG, "output rows hold fewer than R ids");
    dkm::BuildParams bp;
    dkm_status st = prune_setup(ix, alpha, max_occlusion, saturate, list_stride, bp);
    if (st) return st;
    for (int64_t i = 0; i < n; ++i) {
        if (locations[i] >= ix->rows)
            return fail(DKM_ERR_INDEX, "location %u out of range (%lld rows)", locations[i],
                        (long long)ix->rows);
        if (list_len[i] > list_stride)
            return fail(DKM_ERR_INVALID_ARG, "list %lld is longer than the stride", (long long)i);
        for (uint32_t j = 0; j < list_len[i]; ++j)
            if (lists[i * list_stride + j] >= ix->rows)
                return fail(DKM_ERR_INDEX, "candidate %u out of range (%lld rows)",
                            lists[i * list_stride + j], (long long)ix->rows);
    }
    if (n == 0) return DKM_OK;
    std::lock_guard<std::mutex> g(ix->mu);
    HIP_TRY(hipSetDevice(ix->device));
    const int64_t B = std::min<int64_t>(n, kPruneChunk);
    DevBuf<uint32_t> loc, dl, dn, pend, pcnt;
    HIP_TRY(loc.ensure(B));
    HIP_TRY(dl.ensure((size_t)B * list_stride));
    HIP_TRY(dn.ensure(B));
    HIP_TRY(pend.ensure((size_t)B * ix->R));
    HIP_TRY(pcnt.ensure(B));
    bp.pend = pend.p;
    bp.pend_cnt = pcnt.p;
    std::vector<uint32_t> hp((size_t)B * ix->R), hc(B);
    for (int64_t b0 = 0; b0 < n; b0 += B) {
        const int nb = (int)std::min<int64_t>(B, n - b0);
        HIP_TRY(hipMemcpyAsync(loc.p, locations + b0, sizeof(uint32_t) * nb, hipMemcpyHostToDevice, ix->stream));
        HIP_TRY(hipMemcpyAsync(dl.p, lists + b0 * list_stride, sizeof(uint32_t) * nb * list_stride,
                               hipMemcpyHostToDevice, ix->stream));
        HIP_TRY(hipMemcpyAsync(dn.p, list_len + b0, sizeof(uint32_t) * nb, hipMemcpyHostToDevice, ix->stream));
        bp.batch = nb;
        bp.batch_ids = loc.p;
        HIP_TRY(dkm::launch_prune_lists(bp, dl.p, dn.p, (int)list_stride, force_saturate ? 1 : 0,
                                        ix->stream));
        HIP_TRY(hipMemcpyAsync(hp.data(), pend.p, sizeof(uint32_t) * nb * ix->R, hipMemcpyDeviceToHost, ix->stream));
        HIP_TRY(hipMemcpyAsync(hc.data(), pcnt.p, sizeof(uint32_t) * nb, hipMemcpyDeviceToHost, ix->stream));
        HIP_TRY(hipStreamSynchronize(ix->stream));
        for (int b = 0; b < nb; ++b) {
            // an empty list leaves nothing to prune (robust_prune_list's early exit)
            const uint32_t c = list_len[b0 + b] == 0 ? 0u : hc[b];
            out_counts[b0 + b] = c;
            std::memcpy(out_lists + (b0 + b) * out_stride, hp.data() + (size_t)b * ix->R,
                        sizeof(uint32_t) * c);
        }
    }
    return DKM_OK;
}

dkm_status dkm_prune_range(dkm_index *ix, const uint32_t *ids, int64_t n, float alpha,
                           int64_t max_occlusion, int saturate) {
    if (!ix || (n > 0 && !ids)) return fail(DKM_ERR_INVALID_ARG, "null argument");
    if (n < 0) return fail(DKM_ERR_INVALID_ARG, "negative count");
    // an adjacency row holds at most gstride - 1 ids
    const int64_t stride = ix->gstride - 1;
    dkm::BuildParams bp;
    dkm_status st = prune_setup(ix, alpha, max_occlusion, saturate, stride, bp);
    if (st) return st;
    for (int64_t i = 0; i < n; ++i)
        if (ids[i] >= ix->rows)
            return fail(DKM_ERR_INDEX, "id %u out of range (%lld rows)", ids[i], (long long)ix->rows);
    if (n == 0) return DKM_OK;
    std::lock_guard<std::mutex> g(ix->mu);
    HIP_TRY(hipSetDevice(ix->device));
    // the current lists of every id (rows of gstride u32, count in the last slot)
    const int64_t B = std::min<int64_t>(n, kPruneChunk);
    DevBuf<uint32_t> loc, rows, dl, dn, pend, pcnt;
    HIP_TRY(loc.ensure(B));
    HIP_TRY(rows.ensure((size_t)B * ix->gstride));
    HIP_TRY(dl.ensure((size_t)B * stride));
    HIP_TRY(dn.ensure(B));
    HIP_TRY(pend.ensure((size_t)B * ix->R));
    HIP_TRY(pcnt.ensure(B));
    bp.pend = pend.p;
    bp.pend_cnt = pcnt.p;
    std::vector<uint32_t> hrow((size_t)B * ix->gstride), hl((size_t)B * stride), hn(B), hid(B);
    for (int64_t b0 = 0; b0 < n; b0 += B) {
        const int nb = (int)std::min<int64_t>(B, n - b0);
        for (int b = 0; b < nb; ++b)
            HIP_TRY(hipMemcpyAsync(hrow.data() + (size_t)b * ix->gstride,
                                   ix->d_graph + (size_t)ids[b0 + b] * ix->gstride,
                                   sizeof(uint32_t) * ix->gstride, hipMemcpyDeviceToHost, ix->stream));
        HIP_TRY(hipStreamSynchronize(ix->stream));
        // prune_range skips lists within the pruned degree
        int m = 0;
        for (int b = 0; b < nb; ++b) {
            const uint32_t *r = hrow.data() + (size_t)b * ix->gstride;
            const uint32_t c = r[ix->gstride - 1];
            if ((int64_t)c <= ix->R) continue;
            if ((int64_t)c > stride) return fail(DKM_ERR_INDEX, "row %u holds a bad count", ids[b0 + b]);
            hid[m] = ids[b0 + b];
            hn[m] = c;
            std::memcpy(hl.data() + (size_t)m * stride, r, sizeof(uint32_t) * c);
            ++m;
        }
        if (m == 0) continue;
        HIP_TRY(hipMemcpyAsync(loc.p, hid.data(), sizeof(uint32_t) * m, hipMemcpyHostToDevice, ix->stream));
        HIP_TRY(hipMemcpyAsync(dl.p, hl.data(), sizeof(uint32_t) * m * stride, hipMemcpyHostToDevice, ix->stream));
        HIP_TRY(hipMemcpyAsync(dn.p, hn.data(), sizeof(uint32_t) * m, hipMemcpyHostToDevice, ix->stream));
        bp.batch = m;
        bp.batch_ids = loc.p;
        HIP_TRY(dkm::launch_prune_lists(bp, dl.p, dn.p, (int)stride, 0, ix->stream));
        // every pruned list is written back after all of the chunk's prunes read the graph
        HIP_TRY(dkm::launch_set_lists(bp, ix->stream));
        HIP_TRY(hipStreamSynchronize(ix->stream));
    }
    return DKM_OK;
}

// ------------------------------------------------------------------------------------
// Serving: continuous batching over the batch search kernel
// ------------------------------------------------------------------------------------

}  // extern "C"

namespace {

// Serving lanes. Launches on different lanes overlap on the device only when each lane's
// stream has a hardware queue of its own: HIP gives a process GPU_MAX_HW_QUEUES queues
// (default 4) and streams beyond that share them, serialising their launches. The lanes
// plus the index stream and the null stream need GPU_MAX_HW_QUEUES >= kServeStreams + 2
// (measured on 1M x 96 at 200k requests/s: p99 5.9 ms with 4 queues, 1.9 ms with 16
// lanes on 20 queues).
constexpr int kServeStreams = 16;

// One in-flight launch: requests [seq0, seq0 + n) of the ring.
struct ServeLane {
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    bool busy = false;
    uint64_t seq0 = 0;
    int64_t n = 0;
    // A launch's inputs (queries n x dim, k n, L n) and outputs (ids n x max_k, distances
    // n x max_k, counts, cmps, hops, status: n each) are packed back to back for its n in
    // pinned host buffers sized for the largest launch; the kernel reads the inputs and
    // writes the outputs there directly (each query reads its inputs once and stores its
    // outputs once), so a launch issues no copies.
    uint32_t *h_in = nullptr, *h_out = nullptr;
    // per request of the launch: set by the kernel once that request's outputs are written
    // (so it is returned before the rest of its launch ends), and taken by the host
    uint32_t *h_done = nullptr;
    std::vector<uint8_t> taken;
    uint32_t *d_spill = nullptr;
    uint8_t *d_sphq = nullptr;  // spherical: prepared queries of a launch
};

double now_us() {
    return std::chrono::duration<double, std::micro>(
               std::chrono::steady_clock::now().time_since_epoch())
        .count();
}

// Pinned host memory (zeroed); coherent = fine-grained, for buffers the device reads and
// writes directly while the host reuses them between launches.
template <typename T>
hipError_t pinned(T **host, size_t count, bool coherent = false) {
    void *h = nullptr;
    hipError_t e = hipHostMalloc(&h, sizeof(T) * std::max<size_t>(count, 1),
                                 coherent ? hipHostMallocCoherent : hipHostMallocDefault);
    if (e != hipSuccess) return e;
    memset(h, 0, sizeof(T) * std::max<size_t>(count, 1));
    *host = (T *)h;
    return hipSuccess;
}

}  // namespace

struct dkm_server {
    dkm_index *ix = nullptr;
    std::mutex mu;
    uint32_t max_k = 0, max_L = 0;
    int64_t capacity = 0, slots = 0, per_launch = 0;
    dkm::SearchParams base{};  // everything but the per-launch fields
    // ring in pinned host memory, indexed by seq % capacity
    float *queries = nullptr;
    uint32_t *req_k = nullptr, *req_L = nullptr;
    uint32_t *res_ids = nullptr, *res_cnt = nullptr, *res_cmps = nullptr, *res_hops = nullptr,
             *res_status = nullptr;
    float *res_d = nullptr;
    std::vector<uint64_t> tags;
    std::vector<double> t_submit;
    std::vector<uint8_t> busy;      // entry holds a request not yet returned by poll
    std::vector<uint64_t> finished; // seqs whose results are in the ring, not yet returned
    uint64_t tail = 0;              // next seq to assign
    uint64_t launched = 0;          // next seq to launch
    int64_t inflight = 0;           // requests launched and not finished
    ServeLane lane[kServeStreams];
};

extern "C" {

static void server_free(dkm_server *sv) {
    if (!sv) return;
    for (ServeLane &l : sv->lane) {
        if (l.stream) (void)hipStreamSynchronize(l.stream);
        void *dp[] = {l.d_spill, l.d_sphq};
        for (void *q : dp)
            if (q) (void)hipFree(q);
        void *hp[] = {l.h_in, l.h_out, l.h_done};
        for (void *q : hp)
            if (q) (void)hipHostFree(q);
        if (l.done) (void)hipEventDestroy(l.done);
        if (l.stream) (void)hipStreamDestroy(l.stream);
    }
    void *hp[] = {sv->queries, sv->req_k,    sv->req_L,    sv->res_ids,   sv->res_cnt,
                  sv->res_cmps, sv->res_hops, sv->res_status, sv->res_d};
    for (void *q : hp)
        if (q) (void)hipHostFree(q);
    delete sv;
}

// Collects finished launches, then launches waiting requests on free streams while the
// in-flight count stays within `slots`. Caller holds sv->mu.
static dkm_status server_pump(dkm_server *sv) {
    for (ServeLane &l : sv->lane) {
        if (!l.busy) continue;
        hipError_t q = hipEventQuery(l.done);
        if (q != hipSuccess && q != hipErrorNotReady)
            return fail(DKM_ERR_DEVICE, "serving launch: %s", hipGetErrorString(q));
        const bool all = q == hipSuccess;
        // requests whose search has finished go to the ring now (their outputs were written
        // before their flag); the lane itself is free once the whole launch has ended
        const int64_t n = l.n, mk = sv->max_k;
        const uint32_t *o = l.h_out;
        const volatile uint32_t *fl = l.h_done;
        for (int64_t i = 0; i < n; ++i) {
            if (l.taken[i] || !(all || fl[i])) continue;
            std::atomic_thread_fence(std::memory_order_acquire);
            const int64_t a = (int64_t)((l.seq0 + (uint64_t)i) % (uint64_t)sv->capacity);
            memcpy(sv->res_ids + a * mk, o + i * mk, sizeof(uint32_t) * mk);
            memcpy(sv->res_d + a * mk, o + (n + i) * mk, sizeof(float) * mk);
            sv->res_cnt[a] = o[2 * n * mk + i];
            sv->res_cmps[a] = o[2 * n * mk + n + i];
            sv->res_hops[a] = o[2 * n * mk + 2 * n + i];
            sv->res_status[a] = o[2 * n * mk + 3 * n + i];
            sv->finished.push_back(l.seq0 + (uint64_t)i);
            l.taken[i] = 1;
        }
        if (!all) continue;
        sv->inflight -= l.n;
        l.busy = false;
    }
    for (ServeLane &l : sv->lane) {
        if (sv->launched >= sv->tail || sv->inflight >= sv->slots) break;
        if (l.busy) continue;
        const int64_t n = std::min<int64_t>(
            {(int64_t)(sv->tail - sv->launched), sv->per_launch, sv->slots - sv->inflight});
        const uint64_t seq0 = sv->launched;
        const int64_t dim = sv->ix->dim, cap = sv->capacity, mk = sv->max_k;
        // pack the requests' queries, k and L into the lane's pinned input stage
        float *hq = (float *)l.h_in;
        uint32_t *hk = l.h_in + n * dim, *hL = hk + n;
        for (int64_t i = 0; i < n; ++i) {
            const int64_t a = (int64_t)((seq0 + (uint64_t)i) % (uint64_t)cap);
            memcpy(hq + i * dim, sv->queries + a * dim, sizeof(float) * dim);
            hk[i] = sv->req_k[a];
            hL[i] = sv->req_L[a];
        }
        for (int64_t i = 0; i < n; ++i) l.h_done[i] = 0;
        std::fill(l.taken.begin(), l.taken.begin() + n, (uint8_t)0);
        dkm::SearchParams p = sv->base;
        p.done = l.h_done;
        p.queries = (const float *)l.h_in;
        p.nq = (int)n;
        p.k = l.h_in + n * dim;
        p.L = l.h_in + n * dim + n;
        p.out_ids = l.h_out;
        p.out_d = (float *)(l.h_out + n * mk);
        p.out_count = l.h_out + 2 * n * mk;
        p.out_cmps = p.out_count + n;
        p.out_hops = p.out_count + 2 * n;
        p.status = p.out_count + 3 * n;
        p.spill_table = l.d_spill;
        p.sph_qrec = l.d_sphq;
        hipError_t e = dkm::launch_search(p, l.stream);
        if (e != hipSuccess) return fail(DKM_ERR_DEVICE, "serving launch: %s", hipGetErrorString(e));
        HIP_TRY(hipEventRecord(l.done, l.stream));
        l.busy = true;
        l.seq0 = seq0;
        l.n = n;
        sv->launched += (uint64_t)n;
        sv->inflight += n;
    }
    return DKM_OK;
}

dkm_status dkm_server_start(dkm_index *ix, int strategy, uint32_t max_k, uint32_t max_L,
                            uint32_t beam_width, int64_t capacity, int64_t slots,
                            dkm_server **out) {
    if (!ix || !out) return fail(DKM_ERR_INVALID_ARG, "null argument");
    *out = nullptr;
    if (strategy < 0 || strategy > 2) return fail(DKM_ERR_INVALID_ARG, "unknown strategy %d", strategy);
    if (max_k == 0 || max_L == 0 || max_k > max_L || max_L > 4096)
        return fail(DKM_ERR_INVALID_ARG, "bad max_k / max_L");
    if (beam_width == 0) return fail(DKM_ERR_BEAM_WIDTH_ZERO, "beam width must be >= 1");
    if (beam_width > 64) return fail(DKM_ERR_INVALID_ARG, "beam width above 64 is not supported");
    if (capacity < 1 || capacity > (int64_t(1) << 24) || slots < 0 || slots > (int64_t(1) << 20))
        return fail(DKM_ERR_INVALID_ARG, "bad capacity / slots");
    if (!ix->starts_set) return fail(DKM_ERR_NOT_READY, "start points not set");
    const bool quant = strategy != DKM_STRATEGY_FULL;
    const bool sph = quant && ix->sph_active;
    const bool sq = quant && ix->sq_active;
    const bool pq = quant && !sq && !sph;
    if (sph && !ix->sph_codes_set)
        return fail(DKM_ERR_NOT_READY, "spherical strategy needs the codes of every row");
    if (sq && !ix->sq_codes_set)
        return fail(DKM_ERR_NOT_READY, "scalar-quantized strategy needs the codes of every row");
    if (pq && (!ix->d_codes || !ix->codes_set))
        return fail(DKM_ERR_NOT_READY, "PQ strategy needs a PQ schema and codes");
    if (pq && ix->metric == DKM_COSINE) return fail(DKM_ERR_PQ, "cosine PQ traversal is not supported");
    std::lock_guard<std::mutex> ig(ix->mu);
    HIP_TRY(hipSetDevice(ix->device));
    HIP_TRY(hipStreamSynchronize(ix->stream));  // index contents complete before serving
    dkm::SearchParams p{};
    p.metric = ix->metric;
    p.pq = pq;
    p.rerank = strategy == DKM_STRATEGY_HYBRID;
    p.adc_mode = ix->tune_adc_mode;
    p.pq_len = pq ? ix->pq_len : 0;
    p.dim = (int)ix->dim;
    p.vstride = ix->vstride;
    p.vectors = ix->d_vec;
    p.graph = ix->d_graph;
    p.gstride = (int)ix->gstride;
    p.starts = ix->d_starts;
    p.nstart = (int)ix->num_start;
    p.codes = ix->d_codes;
    p.code_stride = (int)ix->M;
    p.M = (int)ix->M;
    p.pivots = ix->d_pivots;
    p.pivots_t = ix->d_pivots_t;
    p.offsets = ix->d_offsets;
    set_sq_params(ix, sq, p, sph);
    p.qstride = ix->dim;
    p.beam_width = beam_width;
    p.k_stride = max_k;
    p.qcap = (int)(max_L + ix->num_start);
    p.nlcap = (int)(beam_width * ix->R_max + 64);
    p.hslots = ix->tune_hslots > 0 ? (uint32_t)ix->tune_hslots : default_visited_slots(max_L);
    p.spill_slots = region_slots(max_L);
    p.max_iters = (uint32_t)std::min<int64_t>(ix->rows + 16, 0x7FFFFFFF);
    if (dkm::search_lds_bytes(p) > 160 * 1024)
        return fail(DKM_ERR_CAPACITY, "per-query LDS for max_L %u exceeds 160 KiB", max_L);
    if (slots == 0) {
        // two device-fulls in flight: queued blocks refill slots as earlier queries finish
        int resident = 0;
        HIP_TRY(dkm::search_resident_blocks(p, &resident));
        if (resident < 1) return fail(DKM_ERR_CAPACITY, "search kernel does not fit on the device");
        slots = 2 * (int64_t)resident;
    }
    dkm_server *sv = new dkm_server();
    sv->ix = ix;
    sv->max_k = max_k;
    sv->max_L = max_L;
    sv->capacity = capacity;
    sv->slots = slots;
    sv->per_launch = std::max<int64_t>(1, (slots + kServeStreams - 1) / kServeStreams);
    sv->base = p;
    const int64_t P = sv->per_launch;
    hipError_t e = hipSuccess;
    if (e == hipSuccess) e = pinned(&sv->queries, (size_t)capacity * ix->dim);
    if (e == hipSuccess) e = pinned(&sv->req_k, capacity);
    if (e == hipSuccess) e = pinned(&sv->req_L, capacity);
    if (e == hipSuccess) e = pinned(&sv->res_ids, (size_t)capacity * max_k);
    if (e == hipSuccess) e = pinned(&sv->res_d, (size_t)capacity * max_k);
    if (e == hipSuccess) e = pinned(&sv->res_cnt, capacity);
    if (e == hipSuccess) e = pinned(&sv->res_cmps, capacity);
    if (e == hipSuccess) e = pinned(&sv->res_hops, capacity);
    if (e == hipSuccess) e = pinned(&sv->res_status, capacity);
    for (ServeLane &l : sv->lane) {
        if (e == hipSuccess) e = hipStreamCreateWithFlags(&l.stream, hipStreamNonBlocking);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&l.done, hipEventDisableTiming);
        const size_t in_words = (size_t)P * (ix->dim + 2), out_words = (size_t)P * (2 * max_k + 4);
        if (e == hipSuccess) e = pinned(&l.h_in, in_words, true);
        if (e == hipSuccess) e = pinned(&l.h_out, out_words, true);
        if (e == hipSuccess) e = pinned(&l.h_done, P, true);
        l.taken.assign(P, 0);
        if (e == hipSuccess) e = hipMalloc(&l.d_spill, sizeof(uint32_t) * p.spill_slots * P);
        if (e == hipSuccess && p.sph)
            e = hipMalloc(&l.d_sphq, (size_t)dkm::sph_query_record_bytes(p.sphq) * P);
    }
    for (ServeLane &l : sv->lane)
        if (e == hipSuccess && l.stream) e = hipStreamSynchronize(l.stream);
    if (e != hipSuccess) {
        server_free(sv);
        return fail(e == hipErrorOutOfMemory ? DKM_ERR_OUT_OF_MEMORY : DKM_ERR_DEVICE,
                    "server allocation: %s", hipGetErrorString(e));
    }
    sv->tags.assign(capacity, 0);
    sv->t_submit.assign(capacity, 0.0);
    sv->busy.assign(capacity, 0);
    *out = sv;
    return DKM_OK;
}

dkm_status dkm_server_slots(dkm_server *sv, int64_t *slots) {
    if (!sv || !slots) return fail(DKM_ERR_INVALID_ARG, "null argument");
    *slots = sv->slots;
    return DKM_OK;
}

// Places one validated request in its ring entry; the caller holds sv->mu.
static bool server_put(dkm_server *sv, const float *query, uint32_t k, uint32_t L, uint64_t tag,
                       double t) {
    const uint64_t seq = sv->tail;
    const uint64_t slot = seq % (uint64_t)sv->capacity;
    if (sv->busy[slot]) return false;
    sv->req_k[slot] = k;
    sv->req_L[slot] = L;
    sv->tags[slot] = tag;
    memcpy(sv->queries + slot * sv->ix->dim, query, sizeof(float) * sv->ix->dim);
    sv->busy[slot] = 1;
    sv->t_submit[slot] = t;
    sv->tail = seq + 1;
    return true;
}

static dkm_status server_check(const dkm_server *sv, uint32_t k, uint32_t L) {
    if (k == 0) return fail(DKM_ERR_K_ZERO, "k must be >= 1");
    if (L == 0) return fail(DKM_ERR_L_ZERO, "L must be >= 1");
    if (k > L) return fail(DKM_ERR_L_LESS_THAN_K, "L (%u) < k (%u)", L, k);
    if (k > sv->max_k || L > sv->max_L)
        return fail(DKM_ERR_INVALID_ARG, "k or L above the server's maximum");
    return DKM_OK;
}

dkm_status dkm_server_submit(dkm_server *sv, const float *query, uint32_t k, uint32_t L,
                             uint64_t tag) {
    if (!sv || !query) return fail(DKM_ERR_INVALID_ARG, "null argument");
    dkm_status st = server_check(sv, k, L);
    if (st != DKM_OK) return st;
    std::lock_guard<std::mutex> g(sv->mu);
    HIP_TRY(hipSetDevice(sv->ix->device));
    if (!server_put(sv, query, k, L, tag, now_us())) return fail(DKM_ERR_CAPACITY, "request ring full");
    return server_pump(sv);
}

dkm_status dkm_server_submit_batch(dkm_server *sv, const float *queries, int64_t n,
                                   const uint32_t *k, const uint32_t *L, const uint64_t *tags,
                                   int64_t *accepted) {
    if (!sv || !accepted || (n > 0 && (!queries || !k || !L || !tags)))
        return fail(DKM_ERR_INVALID_ARG, "null argument");
    *accepted = 0;
    for (int64_t i = 0; i < n; ++i) {
        dkm_status st = server_check(sv, k[i], L[i]);
        if (st != DKM_OK) return st;
    }
    std::lock_guard<std::mutex> g(sv->mu);
    HIP_TRY(hipSetDevice(sv->ix->device));
    const double t = now_us();
    int64_t i = 0;
    for (; i < n; ++i)
        if (!server_put(sv, queries + i * sv->ix->dim, k[i], L[i], tags[i], t)) break;
    *accepted = i;
    return server_pump(sv);
}

dkm_status dkm_server_poll(dkm_server *sv, int64_t max_results, uint64_t *tags, uint32_t *ids,
                           float *dists, uint32_t *counts, uint32_t *cmps, uint32_t *hops,
                           double *latency_us, int64_t *n_out) {
    if (!sv || !n_out) return fail(DKM_ERR_INVALID_ARG, "null argument");
    *n_out = 0;
    std::lock_guard<std::mutex> g(sv->mu);
    HIP_TRY(hipSetDevice(sv->ix->device));
    dkm_status st = server_pump(sv);
    if (st != DKM_OK) return st;
    const double t = now_us();
    int64_t n = 0;
    while (n < max_results && !sv->finished.empty()) {
        const uint64_t seq = sv->finished.back();
        sv->finished.pop_back();
        const uint64_t slot = seq % (uint64_t)sv->capacity;
        sv->busy[slot] = 0;
        const uint32_t status = sv->res_status[slot];
        if (status) {
            // the failing request is consumed; the n results before it are valid
            *n_out = n;
            if (status & dkm::ERR_BAD_REQUEST)
                return fail(DKM_ERR_INVALID_ARG, "request %llu: k / L outside the server's limits",
                            (unsigned long long)seq);
            if (status & dkm::ERR_VISITED_OVERFLOW)
                return fail(DKM_ERR_CAPACITY, "request %llu: visited set overflowed",
                            (unsigned long long)seq);
            return fail(DKM_ERR_INDEX, "request %llu: traversal exceeded the iteration bound",
                        (unsigned long long)seq);
        }
        if (tags) tags[n] = sv->tags[slot];
        if (ids) memcpy(ids + n * sv->max_k, sv->res_ids + slot * sv->max_k, sizeof(uint32_t) * sv->max_k);
        if (dists) memcpy(dists + n * sv->max_k, sv->res_d + slot * sv->max_k, sizeof(float) * sv->max_k);
        if (counts) counts[n] = sv->res_cnt[slot];
        if (cmps) cmps[n] = sv->res_cmps[slot];
        if (hops) hops[n] = sv->res_hops[slot];
        if (latency_us) latency_us[n] = t - sv->t_submit[slot];
        ++n;
    }
    *n_out = n;
    return server_pump(sv);  // entries freed above may admit waiting requests
}

dkm_status dkm_server_pending(dkm_server *sv, int64_t *pending) {
    if (!sv || !pending) return fail(DKM_ERR_INVALID_ARG, "null argument");
    std::lock_guard<std::mutex> g(sv->mu);
    *pending = (int64_t)(sv->tail - sv->launched) + sv->inflight + (int64_t)sv->finished.size();
    return DKM_OK;
}

dkm_status dkm_server_stop(dkm_server *sv) {
    if (!sv) return DKM_OK;
    (void)hipSetDevice(sv->ix->device);
    hipError_t e = hipSuccess;
    for (ServeLane &l : sv->lane) {
        hipError_t e1 = hipStreamSynchronize(l.stream);
        if (e == hipSuccess) e = e1;
    }
    server_free(sv);
    if (e != hipSuccess) return fail(DKM_ERR_DEVICE, "serving stream: %s", hipGetErrorString(e));
    return DKM_OK;
}

// ------------------------------------------------------------------------------------
// Large-dimension k-means
// ------------------------------------------------------------------------------------

static dkm_status kmeans_check(const float *d_x, int64_t n, int64_t dim, const float *d_c,
                               int64_t K, int precision) {
    if (!d_x || !d_c) return fail(DKM_ERR_INVALID_ARG, "null argument");
    if (n <= 0 || dim <= 0 || K <= 0 || n > 0x7FFFFFFF || K > 0x7FFFFFFF || dim > 4096)
        return fail(DKM_ERR_INVALID_ARG, "bad k-means shape");
    if (precision != DKM_PRECISION_F32 && precision != DKM_PRECISION_BF16)
        return fail(DKM_ERR_INVALID_ARG, "unknown precision %d", precision);
    if (precision == DKM_PRECISION_BF16 &&
        dkm::kmeans_mfma_lds(dkm::kmeans_bf16_dp((int)dim), (int)K) > 160 * 1024)
        return fail(DKM_ERR_INVALID_ARG, "dimension %lld too large for the BF16 tile", (long long)dim);
    return DKM_OK;
}

// BF16 copy of rows (or centres) in the layout the assignment kernel for (dp, K) reads.
static hipError_t bf16_copy(const float *x, int64_t rows, int64_t dim, int dp, int K, bool centres,
                            uint16_t *out) {
    if (dkm::kmeans_bf16_frag(dp, K))
        return dkm::kmeans_to_bf16_frag(x, rows, (int)dim, dim, dp,
                                        dkm::kmeans_bf16_rows(rows, dp, K, centres), out, nullptr);
    return dkm::kmeans_to_bf16(x, rows, (int)dim, dim, dp, out, nullptr);
}

dkm_status dkm_kmeans_assign_device(const float *d_x, int64_t n, int64_t dim, const float *d_c,
                                    int64_t K, int precision, uint32_t *d_assign,
                                    float *d_dist) {
    dkm_status st = kmeans_check(d_x, n, dim, d_c, K, precision);
    if (st) return st;
    if (!d_assign) return fail(DKM_ERR_INVALID_ARG, "null argument");
    HIP_TRY(hipSetDevice(g_device));
    DevBuf<float> cn;
    HIP_TRY(cn.ensure(K));
    HIP_TRY(dkm::kmeans_norms(d_c, K, (int)dim, dim, cn.p, nullptr));
    if (precision == DKM_PRECISION_F32) {
        HIP_TRY(dkm::kmeans_assign_f32(d_x, n, (int)dim, dim, d_c, (int)K, cn.p, d_assign, d_dist,
                                       nullptr));
    } else {
        const int dp = dkm::kmeans_bf16_dp((int)dim);
        DevBuf<uint16_t> xb, cb;
        HIP_TRY(xb.ensure(dkm::kmeans_bf16_rows(n, dp, (int)K, false) * dp));
        HIP_TRY(cb.ensure(dkm::kmeans_bf16_rows(K, dp, (int)K, true) * dp));
        HIP_TRY(bf16_copy(d_x, n, dim, dp, (int)K, false, xb.p));
        HIP_TRY(bf16_copy(d_c, K, dim, dp, (int)K, true, cb.p));
        HIP_TRY(dkm::kmeans_assign_mfma(xb.p, n, dp, cb.p, cn.p, (int)K, d_assign, d_dist, nullptr));
    }
    HIP_TRY(hipDeviceSynchronize());
    return DKM_OK;
}

dkm_status dkm_kmeans_device(const float *d_x, int64_t n, int64_t dim, int64_t K, int64_t iters,
                             int precision, float *d_centroids, uint32_t *d_assign,
                             double *ms_assign, double *ms_update) {
    dkm_status st = kmeans_check(d_x, n, dim, d_centroids, K, precision);
    if (st) return st;
    if (!d_assign || iters < 0) return fail(DKM_ERR_INVALID_ARG, "bad argument");
    HIP_TRY(hipSetDevice(g_device));
    const int dp = dkm::kmeans_bf16_dp((int)dim);
    const bool bf = precision == DKM_PRECISION_BF16;
    DevBuf<float> cn;
    DevBuf<uint16_t> xb, cb;
    DevBuf<uint64_t> keys, keys_tmp;
    DevBuf<int64_t> start;
    DevBuf<uint8_t> sort_tmp;
    const size_t sort_bytes = dkm::kmeans_sort_tmp_bytes(n, (int)K);
    HIP_TRY(cn.ensure(K));
    HIP_TRY(keys.ensure(n));
    HIP_TRY(keys_tmp.ensure(n));
    HIP_TRY(start.ensure(K + 1));
    HIP_TRY(sort_tmp.ensure(std::max<size_t>(sort_bytes, 1)));
    if (bf) {
        HIP_TRY(xb.ensure(dkm::kmeans_bf16_rows(n, dp, (int)K, false) * dp));
        HIP_TRY(cb.ensure(dkm::kmeans_bf16_rows(K, dp, (int)K, true) * dp));
        HIP_TRY(bf16_copy(d_x, n, dim, dp, (int)K, false, xb.p));
    }
    hipEvent_t ev[3];
    for (auto &e : ev) HIP_TRY(hipEventCreate(&e));
    double ta = 0.0, tu = 0.0;
    for (int64_t it = 0; it < iters; ++it) {
        HIP_TRY(hipEventRecord(ev[0], nullptr));
        HIP_TRY(dkm::kmeans_norms(d_centroids, K, (int)dim, dim, cn.p, nullptr));
        if (bf) {
            HIP_TRY(bf16_copy(d_centroids, K, dim, dp, (int)K, true, cb.p));
            HIP_TRY(dkm::kmeans_assign_mfma(xb.p, n, dp, cb.p, cn.p, (int)K, d_assign, nullptr, nullptr));
        } else {
            HIP_TRY(dkm::kmeans_assign_f32(d_x, n, (int)dim, dim, d_centroids, (int)K, cn.p, d_assign,
                                           nullptr, nullptr));
        }
        HIP_TRY(hipEventRecord(ev[1], nullptr));
        HIP_TRY(dkm::kmeans_update(d_x, n, (int)dim, dim, d_assign, (int)K, keys.p, keys_tmp.p,
                                   sort_tmp.p, sort_bytes, start.p, d_centroids, nullptr));
        HIP_TRY(hipEventRecord(ev[2], nullptr));
        HIP_TRY(hipEventSynchronize(ev[2]));
        float a = 0.0f, u = 0.0f;
        HIP_TRY(hipEventElapsedTime(&a, ev[0], ev[1]));
        HIP_TRY(hipEventElapsedTime(&u, ev[1], ev[2]));
        ta += a;
        tu += u;
    }
    for (auto &e : ev) (void)hipEventDestroy(e);
    if (ms_assign) *ms_assign = ta;
    if (ms_update) *ms_update = tu;
    return DKM_OK;
}

dkm_status dkm_index_synchronize(dkm_index *ix) {
    if (!ix) return fail(DKM_ERR_INVALID_ARG, "null index");
    HIP_TRY(hipSetDevice(ix->device));
    HIP_TRY(hipStreamSynchronize(ix->stream));
    return DKM_OK;
}

dkm_status dkm_profile(dkm_index *ix, int enable) {
    if (!ix) return fail(DKM_ERR_INVALID_ARG, "null index");
    std::lock_guard<std::mutex> g(ix->mu);
    HIP_TRY(hipStreamSynchronize(ix->stream));
    for (auto &ev : ix->events) {
        (void)hipEventDestroy(ev.first);
        (void)hipEventDestroy(ev.second);
    }
    ix->events.clear();
    ix->prof_launches = 0;
    ix->prof_ms = 0.0;
    ix->profile = enable != 0;
    return DKM_OK;
}

dkm_status dkm_profile_read(dkm_index *ix, int64_t *launches, double *total_ms) {
    if (!ix) return fail(DKM_ERR_INVALID_ARG, "null index");
    std::lock_guard<std::mutex> g(ix->mu);
    HIP_TRY(hipStreamSynchronize(ix->stream));
    for (auto &ev : ix->events) {
        float ms = 0.0f;
        HIP_TRY(hipEventElapsedTime(&ms, ev.first, ev.second));
        ix->prof_ms += ms;
        ix->prof_launches += 1;
        (void)hipEventDestroy(ev.first);
        (void)hipEventDestroy(ev.second);
    }
    ix->events.clear();
    if (launches) *launches = ix->prof_launches;
    if (total_ms) *total_ms = ix->prof_ms;
    return DKM_OK;
}

dkm_status dkm_autotune(dkm_index *ix, int strategy, const float *d_queries, int64_t nq,
                        uint32_t k, uint32_t L, uint32_t beam_width,
                        int64_t *out_visited_lds_slots, int *out_adc_mode) {
    if (!ix || !d_queries || nq <= 0) return fail(DKM_ERR_INVALID_ARG, "bad autotune arguments");
    if (k == 0 || L == 0 || k > L) return fail(DKM_ERR_INVALID_ARG, "bad k / L");
    std::lock_guard<std::mutex> g(ix->mu);
    HIP_TRY(hipSetDevice(ix->device));
    DevBuf<uint32_t> dk, dl, ids, cnt, cm, st;
    DevBuf<float> dd;
    HIP_TRY(dk.ensure(nq));
    HIP_TRY(dl.ensure(nq));
    HIP_TRY(ids.ensure(nq * k));
    HIP_TRY(dd.ensure(nq * k));
    HIP_TRY(cnt.ensure(nq));
    HIP_TRY(cm.ensure(nq));
    HIP_TRY(st.ensure(nq));
    {
        std::vector<uint32_t> hk(nq, k), hl(nq, L);
        HIP_TRY(hipMemcpy(dk.p, hk.data(), 4 * nq, hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(dl.p, hl.data(), 4 * nq, hipMemcpyHostToDevice));
    }
    const int64_t save_h = ix->tune_hslots;
    const int save_a = ix->tune_adc_mode;
    auto timed = [&](int64_t hs, int adc, double *ms) -> dkm_status {
        ix->tune_hslots = hs;
        ix->tune_adc_mode = adc;
        hipEvent_t e0, e1;
        HIP_TRY(hipEventCreate(&e0));
        HIP_TRY(hipEventCreate(&e1));
        dkm_status s2 = launch_locked(ix, strategy, d_queries, nq, dk.p, dl.p, k, L, beam_width, k,
                                      ids.p, dd.p, cnt.p, cm.p, nullptr, st.p);  // warm
        if (s2) return s2;
        HIP_TRY(hipEventRecord(e0, ix->stream));
        for (int rep = 0; rep < 2; ++rep) {
            s2 = launch_locked(ix, strategy, d_queries, nq, dk.p, dl.p, k, L, beam_width, k, ids.p,
                               dd.p, cnt.p, cm.p, nullptr, st.p);
            if (s2) return s2;
        }
        HIP_TRY(hipEventRecord(e1, ix->stream));
        HIP_TRY(hipEventSynchronize(e1));
        float f = 0.0f;
        HIP_TRY(hipEventElapsedTime(&f, e0, e1));
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        *ms = f;
        return DKM_OK;
    };
    // 1. measure comparisons with a generous table
    double ms = 0;
    dkm_status s1 = timed(16384, 1, &ms);
    if (s1) {
        ix->tune_hslots = save_h;
        ix->tune_adc_mode = save_a;
        return s1;
    }
    std::vector<uint32_t> hc(nq);
    HIP_TRY(hipMemcpy(hc.data(), cm.p, 4 * nq, hipMemcpyDeviceToHost));
    std::sort(hc.begin(), hc.end());
    const double p90 = hc[(size_t)((nq - 1) * 0.9)];
    // 2. candidates: small first-level tables (the rest in HBM) and tables holding the
    //    p90 visited count at ~75% load
    std::vector<int64_t> sizes = {256, 512, 1024};
    {
        int64_t hs = ((int64_t)(p90 / 0.75) + 64 + ix->num_start + 63) / 64 * 64;
        sizes.push_back(std::min<int64_t>(32768, std::max<int64_t>(256, hs)));
    }
    const bool pq = strategy != DKM_STRATEGY_FULL;
    double best = 1e30;
    int64_t best_h = 0;
    int best_a = 1;
    for (int adc : pq ? std::vector<int>{0, 1} : std::vector<int>{1}) {
        for (int64_t hs : sizes) {
            double t = 0;
            dkm_status s2 = timed(hs, adc, &t);
            if (s2 == DKM_ERR_CAPACITY) continue;  // configuration does not fit in LDS
            if (s2) {
                ix->tune_hslots = save_h;
                ix->tune_adc_mode = save_a;
                return s2;
            }
            if (t < best) {
                best = t;
                best_h = hs;
                best_a = adc;
            }
        }
    }
    ix->tune_hslots = best_h;
    ix->tune_adc_mode = best_a;
    if (out_visited_lds_slots) *out_visited_lds_slots = best_h;
    if (out_adc_mode) *out_adc_mode = best_a;
    return DKM_OK;
}

dkm_status dkm_set_tuning(dkm_index *ix, int64_t visited_lds_slots, int adc_mode) {
    if (!ix) return fail(DKM_ERR_INVALID_ARG, "null index");
    if (visited_lds_slots < 0 || (visited_lds_slots > 0 && visited_lds_slots < 256) ||
        visited_lds_slots > 32768 || visited_lds_slots % 64)
        return fail(DKM_ERR_INVALID_ARG, "visited slots must be 0 or a multiple of 64 in [256, 32768]");
    if (adc_mode != 0 && adc_mode != 1) return fail(DKM_ERR_INVALID_ARG, "adc_mode must be 0 or 1");
    std::lock_guard<std::mutex> g(ix->mu);
    ix->tune_hslots = visited_lds_slots;
    ix->tune_adc_mode = adc_mode;
    return DKM_OK;
}

dkm_status dkm_set_visited_capacity(dkm_index *ix, int64_t slots) {
    if (!ix) return fail(DKM_ERR_INVALID_ARG, "null index");
    if (slots < 0 || (slots > 0 && (slots < 256 || slots > (int64_t(1) << 24) || (slots & (slots - 1)))))
        return fail(DKM_ERR_INVALID_ARG, "visited capacity must be 0 or a power of two in [256, 2^24]");
    std::lock_guard<std::mutex> g(ix->mu);
    ix->tune_spill = (uint64_t)slots;
    return DKM_OK;
}

// -------------------------------------------------------------------------------------
// Candidate pre-scoring (BF16 MFMA approximate scores, exact FP32 rescore of the best C)
// -------------------------------------------------------------------------------------
struct dkm_prescorer {
    int metric = 0;
    int64_t n = 0, dim = 0, ns = 0;
    int dp = 0;
    bool fp8 = false;                // FP8 (E4M3) operands instead of bf16
    float sx = 1.0f;                 // FP8: power-of-two scale of the rows
    const float *x = nullptr;        // caller's f32 rows (rescore)
    DevBuf<uint16_t> xf, sf;         // fragment layout (bf16 or fp8): rows, sample rows
    DevBuf<float> xn, sn;            // f32 squared norms (L2 only)
    // Work buffers of dkm_prescore_search, grown on demand and kept with the handle so a
    // search allocates nothing once they fit (searches hold mu).
    struct Work {
        DevBuf<uint16_t> qf, r_qf;
        DevBuf<float> dense, tau, dexact, r_q, r_tau;
        DevBuf<uint32_t> cnt, bnd_b, bnd_e, cids, r_map, amax;
        DevBuf<uint64_t> lists, fkeys, fsorted;
        DevBuf<uint8_t> tmp;
        DevBuf<int64_t> r_row;
    } w;
    std::mutex mu;
};

namespace {
constexpr int64_t kPsBatch = 1024;  // queries per filter launch (bounds the LDS stage load)

// Power of two that brings the largest magnitude m to at most 448 (the largest finite E4M3)
// while using as much of the range as possible; exact to undo.
float fp8_scale(float m) {
    if (!(m > 0.0f) || !std::isfinite(m)) return 1.0f;
    const int e = std::max(-100, std::min(100, (int)std::floor(std::log2(448.0 / (double)m))));
    return std::ldexp(1.0f, e);
}

hipError_t absmax_host(const float *x, int64_t rows, int dim, DevBuf<uint32_t> &d, float *m) {
    hipError_t e = d.ensure(1);
    if (e == hipSuccess) e = dkm::ps_absmax(x, rows, dim, dim, d.p, nullptr);
    uint32_t bits = 0;
    if (e == hipSuccess) e = hipMemcpy(&bits, d.p, sizeof(bits), hipMemcpyDeviceToHost);
    float f;
    std::memcpy(&f, &bits, sizeof(f));
    *m = f;
    return e;
}

// Rows in the operand layout of the prescorer's precision (scale: FP8 only).
hipError_t ps_frag(bool fp8, const float *x, int64_t rows, int dim, int dp, int64_t rows_pad,
                   float scale, uint16_t *out) {
    if (fp8)
        return dkm::ps_to_fp8_frag(x, rows, dim, dim, dp, rows_pad, scale, (uint8_t *)out, nullptr);
    return dkm::kmeans_to_bf16_frag(x, rows, dim, dim, dp, rows_pad, out, nullptr);
}
}  // namespace

dkm_status dkm_prescore_create(const float *d_x, int64_t n, int64_t dim, int metric,
                               int precision, dkm_prescorer **out) {
    if (!d_x || !out) return fail(DKM_ERR_INVALID_ARG, "null argument");
    *out = nullptr;
    if (n <= 0 || n > 0x7FFFFFFF || dim <= 0) return fail(DKM_ERR_INVALID_ARG, "bad shape");
    if (metric != DKM_L2 && metric != DKM_INNER_PRODUCT && metric != DKM_COSINE_NORMALIZED)
        return fail(DKM_ERR_INVALID_ARG,
                    "pre-scoring supports L2, inner product and cosine-normalized (got %d)", metric);
    if (precision != DKM_PRECISION_BF16 && precision != DKM_PRECISION_FP8)
        return fail(DKM_ERR_INVALID_ARG, "unknown pre-scoring precision %d", precision);
    const bool fp8 = precision == DKM_PRECISION_FP8;
    const int dp = fp8 ? (dim <= 768 ? dkm::ps_fp8_dp((int)dim) : -1)
                       : dkm::kmeans_bf16_dp((int)dim);
    if (fp8 ? dp < 0 : !dkm::ps_supported_dp(dp))
        return fail(DKM_ERR_INVALID_ARG, "pre-scoring supports dims up to 768 (got %lld)",
                    (long long)dim);
    HIP_TRY(hipSetDevice(g_device));
    auto *ps = new dkm_prescorer();
    ps->metric = metric;
    ps->n = n;
    ps->dim = dim;
    ps->dp = dp;
    ps->fp8 = fp8;
    ps->x = d_x;
    ps->ns = std::min<int64_t>(n, PS_SAMPLE_MAX);
    const bool l2 = metric == DKM_L2;
    auto bad = [&](hipError_t e) {
        delete ps;
        return fail(DKM_ERR_DEVICE, "%s", hipGetErrorString(e));
    };
    hipError_t e;
    const int64_t npad = (n + 255) / 256 * 256, spad = (ps->ns + 255) / 256 * 256;
    if (fp8) {
        float m = 0.0f;
        DevBuf<uint32_t> amax;
        if ((e = absmax_host(d_x, n, (int)dim, amax, &m)) != hipSuccess) return bad(e);
        ps->sx = fp8_scale(m);
    }
    const int64_t per_row = fp8 ? dp / 2 : dp;  // 16-bit units per row
    if ((e = ps->xf.ensure(npad * per_row)) != hipSuccess) return bad(e);
    if ((e = ps_frag(fp8, d_x, n, (int)dim, dp, npad, ps->sx, ps->xf.p)) != hipSuccess)
        return bad(e);
    // the sample: ns distinct rows drawn with a fixed seed (SplitMix64), in increasing order
    std::vector<int64_t> rows;
    if (ps->ns == n) {
        rows.resize(n);
        for (int64_t i = 0; i < n; ++i) rows[i] = i;
    } else {
        std::vector<uint8_t> taken(n, 0);
        uint64_t st = 0x5DEECE66Dull;
        while ((int64_t)rows.size() < ps->ns) {
            uint64_t z = (st += 0x9E3779B97F4A7C15ull);
            z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
            z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
            z ^= z >> 31;
            const int64_t r = (int64_t)(z % (uint64_t)n);
            if (!taken[r]) {
                taken[r] = 1;
                rows.push_back(r);
            }
        }
        std::sort(rows.begin(), rows.end());
    }
    DevBuf<int64_t> drows;
    DevBuf<float> srows;
    if ((e = drows.ensure(ps->ns)) != hipSuccess) return bad(e);
    if ((e = srows.ensure(ps->ns * dim)) != hipSuccess) return bad(e);
    if ((e = hipMemcpy(drows.p, rows.data(), sizeof(int64_t) * ps->ns, hipMemcpyHostToDevice)) !=
        hipSuccess)
        return bad(e);
    if ((e = dkm::ps_gather_rows(d_x, (int)dim, drows.p, ps->ns, srows.p, nullptr)) != hipSuccess)
        return bad(e);
    if ((e = ps->sf.ensure(spad * per_row)) != hipSuccess) return bad(e);
    if ((e = ps_frag(fp8, srows.p, ps->ns, (int)dim, dp, spad, ps->sx, ps->sf.p)) != hipSuccess)
        return bad(e);
    if (l2) {
        if ((e = ps->xn.ensure(n)) != hipSuccess) return bad(e);
        if ((e = ps->sn.ensure(ps->ns)) != hipSuccess) return bad(e);
        if ((e = dkm::ps_norms(d_x, n, (int)dim, nullptr, ps->xn.p, nullptr)) != hipSuccess)
            return bad(e);
        if ((e = dkm::ps_norms(d_x, ps->ns, (int)dim, drows.p, ps->sn.p, nullptr)) != hipSuccess)
            return bad(e);
    }
    if ((e = hipDeviceSynchronize()) != hipSuccess) return bad(e);
    *out = ps;
    return DKM_OK;
}

void dkm_prescore_destroy(dkm_prescorer *ps) {
    if (!ps) return;
    (void)hipSetDevice(g_device);
    (void)hipDeviceSynchronize();
    delete ps;
}

dkm_status dkm_prescore_search(dkm_prescorer *ps, const float *d_q, int64_t nq, int64_t k,
                               int64_t candidates, uint32_t *d_ids, float *d_dists,
                               double *ms_stages) {
    if (!ps || !d_q || !d_ids || !d_dists) return fail(DKM_ERR_INVALID_ARG, "null argument");
    if (k <= 0) return fail(DKM_ERR_K_ZERO, "k must be positive");
    if (nq < 0) return fail(DKM_ERR_INVALID_ARG, "bad query count");
    const int64_t C_req = std::min<int64_t>(std::max<int64_t>(candidates, k), ps->n);
    if (k > ps->n) return fail(DKM_ERR_INVALID_ARG, "k = %lld exceeds %lld rows", (long long)k,
                               (long long)ps->n);
    // Tiny bases (1.5 C >= n): the filter passes every row with the same (infinite) key, so
    // the first C of a list would be the first C rows by id; every row is rescored instead.
    const bool all_pass = 1.5 * (double)C_req >= (double)ps->n;
    const int64_t C = all_pass ? ps->n : C_req;
    if (nq == 0) return DKM_OK;
    std::lock_guard<std::mutex> g(ps->mu);
    HIP_TRY(hipSetDevice(g_device));
    const int dp = ps->dp;
    const bool l2 = ps->metric == DKM_L2;
    const int64_t B = std::min<int64_t>(nq, kPsBatch);
    const int64_t Bp = (B + 63) / 64 * 64;
    if (dkm::ps_filter_lds(dp, ps->fp8, (int)Bp) > 160 * 1024)
        return fail(DKM_ERR_INVALID_ARG, "pre-scoring tile does not fit LDS");
    // Threshold ranks in the sample: j0 aims at ~1.5 C passing rows but is at least 15 (an
    // order statistic of rank < 16 scatters too widely); retries use 3x (short lists) and
    // 3/4x (overflow) of it. The list capacity holds 4x what rank j0 + 1 predicts and at
    // least 8 C.
    const double rows_per_sample = (double)ps->n / (double)ps->ns;
    const double want = 1.5 * (double)C / rows_per_sample;
    const int j0 = (int)std::min<int64_t>(ps->ns - 1,
                                          std::max<int64_t>(15, (int64_t)std::ceil(want) - 1));
    const int j_loose = (int)std::min<int64_t>(ps->ns - 1, 3 * ((int64_t)j0 + 1) - 1);
    const int j_tight = (int)std::max<int64_t>(0, ((int64_t)j0 + 1) * 3 / 4 - 1);
    const int64_t cap = std::min<int64_t>(
        ps->n, std::max<int64_t>(8 * C, (int64_t)std::ceil(4.0 * (j0 + 1) * rows_per_sample)));
    auto &w = ps->w;
    auto &qf = w.qf;
    auto &dense = w.dense, &tau = w.tau, &dexact = w.dexact;
    auto &cnt = w.cnt, &bnd_b = w.bnd_b, &bnd_e = w.bnd_e, &cids = w.cids;
    auto &lists = w.lists, &fkeys = w.fkeys, &fsorted = w.fsorted;
    auto &tmp = w.tmp;
    auto &r_q = w.r_q, &r_tau = w.r_tau;  // redo launches
    auto &r_qf = w.r_qf;
    auto &r_map = w.r_map;
    auto &r_row = w.r_row;
    size_t tmp2 = 0;
    HIP_TRY(dkm::ps_sort_tmp_bytes(B * C, (int)B, &tmp2));
    HIP_TRY(qf.ensure(Bp * dp));
    HIP_TRY(dense.ensure(B * ps->ns));
    HIP_TRY(tau.ensure(3 * B));
    HIP_TRY(cnt.ensure(B * PS_COUNT_STRIDE));  // one cache line per query's counter
    HIP_TRY(bnd_b.ensure(B));
    HIP_TRY(bnd_e.ensure(B));
    HIP_TRY(lists.ensure(B * cap));
    HIP_TRY(cids.ensure(B * C));
    HIP_TRY(dexact.ensure(B * C));
    HIP_TRY(fkeys.ensure(B * C));
    HIP_TRY(fsorted.ensure(B * C));
    HIP_TRY(tmp.ensure(std::max<size_t>(tmp2, 1)));
    hipEvent_t ev[5];
    for (auto &e : ev) HIP_TRY(hipEventCreate(&e));
    struct EvGuard {
        hipEvent_t *e;
        ~EvGuard() {
            for (int i = 0; i < 5; ++i) (void)hipEventDestroy(e[i]);
        }
    } evg{ev};
    double acc_ms[4] = {0, 0, 0, 0};
    std::vector<uint32_t> hcnt(B * PS_COUNT_STRIDE);  // query i's count at i * PS_COUNT_STRIDE
    std::vector<float> htau(3 * B);
    for (int64_t q0 = 0; q0 < nq; q0 += B) {
        const int b = (int)std::min<int64_t>(B, nq - q0);
        const int bp = (int)((b + 63) / 64 * 64);
        const float *q = d_q + q0 * ps->dim;
        HIP_TRY(hipEventRecord(ev[0], nullptr));
        // FP8: the batch's own power-of-two scale; the accumulators hold s times the dot
        // products (s = row scale x query scale)
        float sq = 1.0f;
        if (ps->fp8) {
            float m = 0.0f;
            HIP_TRY(absmax_host(q, b, (int)ps->dim, ps->w.amax, &m));
            sq = fp8_scale(m);
        }
        const float S = ps->fp8 ? ps->sx * sq : 1.0f;
        HIP_TRY(ps_frag(ps->fp8, q, b, (int)ps->dim, dp, bp, sq, qf.p));
        if (!all_pass) {
            HIP_TRY(dkm::ps_dense(dp, ps->fp8, ps->sf.p, ps->ns, ps->sn.p, qf.p, b, bp, l2,
                                  1.0f / S, dense.p, ps->ns, nullptr));
            HIP_TRY(dkm::ps_select(dense.p, ps->ns, (int)ps->ns, j0, j_loose, j_tight, tau.p, b,
                                   nullptr));
            HIP_TRY(hipMemcpy(htau.data(), tau.p, sizeof(float) * 3 * b, hipMemcpyDeviceToHost));
        } else {
            for (int i = 0; i < b; ++i) htau[i] = __builtin_inff();
            HIP_TRY(hipMemcpy(tau.p, htau.data(), sizeof(float) * b, hipMemcpyHostToDevice));
        }
        HIP_TRY(hipMemset(cnt.p, 0, sizeof(uint32_t) * b * PS_COUNT_STRIDE));
        HIP_TRY(dkm::ps_filter(dp, ps->fp8, ps->xf.p, ps->n, ps->xn.p, qf.p, b, bp, l2, S, tau.p,
                               nullptr, cnt.p, lists.p, cap, nullptr));
        // Redo the queries whose list is short (looser threshold, then every row) or overflowed
        // (tighter threshold): only those queries, gathered into a small launch of their own.
        for (int round = 0;; ++round) {
            // the padded counters whole (a strided copy moves one 4-byte row at a time)
            HIP_TRY(hipMemcpy(hcnt.data(), cnt.p, sizeof(uint32_t) * b * PS_COUNT_STRIDE,
                              hipMemcpyDeviceToHost));
            std::vector<uint32_t> ridx;
            std::vector<int64_t> rrow;
            std::vector<float> rtau;
            for (int i = 0; i < b; ++i) {
                uint32_t &ci = hcnt[(size_t)i * PS_COUNT_STRIDE];
                if (ci < (uint32_t)C) {
                    rtau.push_back(round == 0 ? htau[b + i] : __builtin_inff());
                } else if (ci > (uint64_t)cap) {
                    if (round > 0)
                        return fail(DKM_ERR_CAPACITY, "query %lld: %u candidates exceed %lld",
                                    (long long)(q0 + i), ci, (long long)cap);
                    rtau.push_back(htau[2 * b + i]);
                } else {
                    continue;
                }
                ridx.push_back((uint32_t)i);
                rrow.push_back(i);
                ci = 0;
            }
            if (ridx.empty()) break;
            if (round == 2) return fail(DKM_ERR_CAPACITY, "pre-scoring threshold did not converge");
            const int f = (int)ridx.size(), fp = (f + 63) / 64 * 64;
            HIP_TRY(r_q.ensure((size_t)f * ps->dim));
            HIP_TRY(r_qf.ensure((size_t)fp * dp));
            HIP_TRY(r_map.ensure(f));
            HIP_TRY(r_row.ensure(f));
            HIP_TRY(r_tau.ensure(f));
            HIP_TRY(hipMemcpy(cnt.p, hcnt.data(), sizeof(uint32_t) * b * PS_COUNT_STRIDE,
                              hipMemcpyHostToDevice));
            HIP_TRY(hipMemcpy(r_map.p, ridx.data(), sizeof(uint32_t) * f, hipMemcpyHostToDevice));
            HIP_TRY(hipMemcpy(r_row.p, rrow.data(), sizeof(int64_t) * f, hipMemcpyHostToDevice));
            HIP_TRY(hipMemcpy(r_tau.p, rtau.data(), sizeof(float) * f, hipMemcpyHostToDevice));
            HIP_TRY(dkm::ps_gather_rows(q, (int)ps->dim, r_row.p, f, r_q.p, nullptr));
            HIP_TRY(ps_frag(ps->fp8, r_q.p, f, (int)ps->dim, dp, fp, sq, r_qf.p));
            HIP_TRY(dkm::ps_filter(dp, ps->fp8, ps->xf.p, ps->n, ps->xn.p, r_qf.p, f, fp, l2, S,
                                   r_tau.p, r_map.p, cnt.p, lists.p, cap, nullptr));
        }
        HIP_TRY(hipEventRecord(ev[1], nullptr));
        HIP_TRY(dkm::ps_take_smallest(lists.p, cnt.p, b, cap, (int)C, cids.p, nullptr));
        HIP_TRY(hipEventRecord(ev[2], nullptr));
        HIP_TRY(dkm::launch_distance_batch(ps->metric, (int)ps->dim, q, b, ps->x, ps->dim, cids.p,
                                           C, dexact.p, nullptr));
        HIP_TRY(hipEventRecord(ev[3], nullptr));
        HIP_TRY(dkm::ps_final_keys(dexact.p, cids.p, (int64_t)b * C, fkeys.p, nullptr));
        if (dkm::ps_topk_supported((int)k)) {
            HIP_TRY(dkm::ps_topk(fkeys.p, b, (int)C, (int)k, d_ids + q0 * k, d_dists + q0 * k,
                                 nullptr));
        } else {
            std::vector<uint32_t> fb(b), fe(b);
            for (int i = 0; i < b; ++i) {
                fb[i] = (uint32_t)(i * C);
                fe[i] = (uint32_t)((i + 1) * C);
            }
            HIP_TRY(hipMemcpy(bnd_b.p, fb.data(), sizeof(uint32_t) * b, hipMemcpyHostToDevice));
            HIP_TRY(hipMemcpy(bnd_e.p, fe.data(), sizeof(uint32_t) * b, hipMemcpyHostToDevice));
            HIP_TRY(dkm::ps_sort(tmp.p, tmp2, fkeys.p, fsorted.p, (int64_t)b * C, b, bnd_b.p,
                                 bnd_e.p, nullptr));
            HIP_TRY(dkm::ps_output(fsorted.p, b, (int)C, (int)k, d_ids + q0 * k,
                                   d_dists + q0 * k, nullptr));
        }
        HIP_TRY(hipEventRecord(ev[4], nullptr));
        HIP_TRY(hipEventSynchronize(ev[4]));
        for (int i = 0; i < 4; ++i) {
            float ms = 0.0f;
            HIP_TRY(hipEventElapsedTime(&ms, ev[i], ev[i + 1]));
            acc_ms[i] += ms;
        }
    }
    if (ms_stages)
        for (int i = 0; i < 4; ++i) ms_stages[i] = acc_ms[i];
    return DKM_OK;
}

}  // extern "C"

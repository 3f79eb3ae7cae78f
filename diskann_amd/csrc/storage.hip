// On-disk formats of the reference, read and written on the host.
//
//   .bin matrices   diskann-utils/src/io.rs:60-100 (8-byte header: npoints, ndims as
//                   little-endian u32, then row-major data)
//   graph file      diskann-providers/src/storage/bin.rs:230-380 (24-byte header: file
//                   size u64, max degree u32, start id u32, number of start points u64;
//                   then per node a u32 count and that many u32 neighbours), with the
//                   in-memory save/load of simple_neighbor_provider.rs:143-185 and the
//                   disk-layout adaptor of simple_neighbor_provider.rs:313-356
//   PQ pivots file  diskann-providers/src/storage/pq_storage.rs:93-149,278-349 (a 4 x 1
//                   u64 .bin offset table at 0; from byte 4096 the pivots K x D f32, the
//                   centroid D x 1 f32 and the chunk offsets (M + 1) x 1 u32, each a .bin;
//                   a non-zero centroid is added to every pivot row on load)
//   PQ codes file   a u8 .bin of N x M (pq_storage.rs:245-276)
#include <algorithm>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/diskann_mi355x.h"

namespace dkm {
dkm_status set_error(dkm_status s, const char *msg);
}

namespace {

dkm_status fail(dkm_status s, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    return dkm::set_error(s, buf);
}

constexpr size_t kPivotMetadata = 4096;  // pq/mod.rs:20 METADATA_SIZE
constexpr size_t kGraphHeader = 24;

struct File {
    FILE *f = nullptr;
    explicit File(const char *path, const char *mode) : f(path ? fopen(path, mode) : nullptr) {}
    ~File() {
        if (f) fclose(f);
    }
    bool ok() const { return f != nullptr; }
};

template <typename T>
bool put(FILE *f, T v) {
    return fwrite(&v, sizeof(T), 1, f) == 1;  // the library targets little-endian hosts
}

template <typename T>
bool get(FILE *f, T *v) {
    return fread(v, sizeof(T), 1, f) == 1;
}

bool seek(FILE *f, uint64_t off) { return fseeko(f, (off_t)off, SEEK_SET) == 0; }

uint64_t file_size(FILE *f) {
    const off_t cur = ftello(f);
    fseeko(f, 0, SEEK_END);
    const off_t end = ftello(f);
    fseeko(f, cur, SEEK_SET);
    return end < 0 ? 0 : (uint64_t)end;
}

// .bin block at `off`: header + rows x cols values of `elem` bytes.
dkm_status read_bin_block(FILE *f, uint64_t off, size_t elem, uint32_t *rows, uint32_t *cols,
                          std::vector<uint8_t> *data, const char *path) {
    if (!seek(f, off) || !get(f, rows) || !get(f, cols))
        return fail(DKM_ERR_IO, "%s: cannot read a .bin header at byte %llu", path,
                    (unsigned long long)off);
    const uint64_t bytes = (uint64_t)*rows * *cols * elem;
    if (off + 8 + bytes > file_size(f))
        return fail(DKM_ERR_IO, "%s: block at byte %llu (%u x %u) runs past the end of the file",
                    path, (unsigned long long)off, *rows, *cols);
    if (data) {
        data->resize(bytes);
        if (bytes && fread(data->data(), 1, bytes, f) != bytes)
            return fail(DKM_ERR_IO, "%s: short read", path);
    }
    return DKM_OK;
}

bool write_bin_block(FILE *f, const void *data, size_t elem, uint32_t rows, uint32_t cols) {
    const size_t bytes = (size_t)rows * cols * elem;
    return put(f, rows) && put(f, cols) && (bytes == 0 || fwrite(data, 1, bytes, f) == bytes);
}

}  // namespace

extern "C" {

dkm_status dkm_bin_write(const char *path, const void *data, int64_t elem_size, int64_t nrows,
                         int64_t ncols) {
    if (!path || (!data && nrows * ncols > 0) || elem_size <= 0)
        return fail(DKM_ERR_INVALID_ARG, "bad argument");
    if (nrows < 0 || ncols < 0 || nrows > UINT32_MAX || ncols > UINT32_MAX)
        return fail(DKM_ERR_INVALID_ARG, "shape does not fit the u32 header");
    File f(path, "wb");
    if (!f.ok()) return fail(DKM_ERR_IO, "%s: cannot create", path);
    if (!write_bin_block(f.f, data, (size_t)elem_size, (uint32_t)nrows, (uint32_t)ncols))
        return fail(DKM_ERR_IO, "%s: write failed", path);
    return DKM_OK;
}

dkm_status dkm_bin_read_shape(const char *path, int64_t *nrows, int64_t *ncols) {
    if (!path || !nrows || !ncols) return fail(DKM_ERR_INVALID_ARG, "null argument");
    File f(path, "rb");
    if (!f.ok()) return fail(DKM_ERR_IO, "%s: cannot open", path);
    uint32_t r = 0, c = 0;
    if (!get(f.f, &r) || !get(f.f, &c)) return fail(DKM_ERR_IO, "%s: no .bin header", path);
    *nrows = r;
    *ncols = c;
    return DKM_OK;
}

dkm_status dkm_bin_read(const char *path, void *out, int64_t elem_size, int64_t nrows,
                        int64_t ncols) {
    if (!path || (!out && nrows * ncols > 0) || elem_size <= 0)
        return fail(DKM_ERR_INVALID_ARG, "bad argument");
    File f(path, "rb");
    if (!f.ok()) return fail(DKM_ERR_IO, "%s: cannot open", path);
    uint32_t r = 0, c = 0;
    dkm_status st = read_bin_block(f.f, 0, (size_t)elem_size, &r, &c, nullptr, path);
    if (st) return st;
    if (r != nrows || c != ncols)
        return fail(DKM_ERR_DIM_MISMATCH, "%s holds %u x %u, expected %lld x %lld", path, r, c,
                    (long long)nrows, (long long)ncols);
    const size_t bytes = (size_t)r * c * (size_t)elem_size;
    if (bytes && fread(out, 1, bytes, f.f) != bytes) return fail(DKM_ERR_IO, "%s: short read", path);
    return DKM_OK;
}

dkm_status dkm_set_vectors_from_bin(dkm_index *ix, const char *path) {
    if (!ix || !path) return fail(DKM_ERR_INVALID_ARG, "null argument");
    int64_t dim = 0, max_points = 0, num_start = 0, R = 0, R_max = 0;
    dkm_status st = dkm_index_shape(ix, &dim, &max_points, &num_start, &R, &R_max);
    if (st) return st;
    File f(path, "rb");
    if (!f.ok()) return fail(DKM_ERR_IO, "%s: cannot open", path);
    uint32_t r = 0, c = 0;
    st = read_bin_block(f.f, 0, sizeof(float), &r, &c, nullptr, path);
    if (st) return st;
    if ((int64_t)c != dim) return fail(DKM_ERR_DIM_MISMATCH, "%s has dim %u, index dim %lld", path, c, (long long)dim);
    if ((int64_t)r > max_points)
        return fail(DKM_ERR_INDEX, "%s has %u rows, index holds %lld", path, r, (long long)max_points);
    const int64_t chunk = 1 << 16;
    std::vector<float> buf((size_t)std::min<int64_t>(chunk, std::max<uint32_t>(r, 1)) * c);
    for (int64_t row0 = 0; row0 < (int64_t)r; row0 += chunk) {
        const int64_t n = std::min<int64_t>(chunk, (int64_t)r - row0);
        if (fread(buf.data(), sizeof(float) * c, (size_t)n, f.f) != (size_t)n)
            return fail(DKM_ERR_IO, "%s: short read", path);
        st = dkm_set_vectors(ix, buf.data(), row0, n);
        if (st) return st;
    }
    return DKM_OK;
}

// save_direct (in-memory layout): every row, start points included, and the configured
// maximum degree.
dkm_status dkm_save_graph(dkm_index *ix, const char *path, uint32_t start_point) {
    if (!ix || !path) return fail(DKM_ERR_INVALID_ARG, "null argument");
    int64_t dim = 0, max_points = 0, num_start = 0, R = 0, R_max = 0;
    dkm_status st = dkm_index_shape(ix, &dim, &max_points, &num_start, &R, &R_max);
    if (st) return st;
    File f(path, "wb");
    if (!f.ok()) return fail(DKM_ERR_IO, "%s: cannot create", path);
    const int64_t rows = max_points + num_start, chunk = 1 << 16;
    uint64_t size = kGraphHeader;
    bool ok = put<uint64_t>(f.f, 0) && put<uint32_t>(f.f, (uint32_t)R_max) &&
              put<uint32_t>(f.f, start_point) && put<uint64_t>(f.f, (uint64_t)num_start);
    std::vector<uint32_t> lists((size_t)chunk * R_max), counts((size_t)chunk);
    for (int64_t row0 = 0; ok && row0 < rows; row0 += chunk) {
        const int64_t n = std::min(chunk, rows - row0);
        st = dkm_get_graph(ix, lists.data(), counts.data(), R_max, row0, n);
        if (st) return st;
        for (int64_t i = 0; ok && i < n; ++i) {
            ok = put<uint32_t>(f.f, counts[i]) &&
                 (counts[i] == 0 || fwrite(&lists[i * R_max], 4, counts[i], f.f) == counts[i]);
            size += 4ull * (1 + counts[i]);
        }
    }
    ok = ok && seek(f.f, 0) && put<uint64_t>(f.f, size);
    if (!ok) return fail(DKM_ERR_IO, "%s: write failed", path);
    return DKM_OK;
}

// The disk-index layout: data rows only, the in-memory start id (the first start row)
// replaced by `actual_start_point` in every list, no start points in the header and the
// observed maximum degree.
dkm_status dkm_save_graph_disk(dkm_index *ix, const char *path, uint32_t actual_start_point) {
    if (!ix || !path) return fail(DKM_ERR_INVALID_ARG, "null argument");
    int64_t dim = 0, max_points = 0, num_start = 0, R = 0, R_max = 0;
    dkm_status st = dkm_index_shape(ix, &dim, &max_points, &num_start, &R, &R_max);
    if (st) return st;
    if (actual_start_point >= (uint64_t)max_points)
        return fail(DKM_ERR_INDEX, "actual start point %u is not a data row", actual_start_point);
    File f(path, "wb");
    if (!f.ok()) return fail(DKM_ERR_IO, "%s: cannot create", path);
    const uint32_t inmem_start = (uint32_t)max_points;
    const int64_t chunk = 1 << 16;
    uint64_t size = kGraphHeader;
    uint32_t observed = 0;
    bool ok = put<uint64_t>(f.f, 0) && put<uint32_t>(f.f, 0) && put<uint32_t>(f.f, actual_start_point) &&
              put<uint64_t>(f.f, 0);
    std::vector<uint32_t> lists((size_t)chunk * R_max), counts((size_t)chunk);
    for (int64_t row0 = 0; ok && row0 < max_points; row0 += chunk) {
        const int64_t n = std::min(chunk, max_points - row0);
        st = dkm_get_graph(ix, lists.data(), counts.data(), R_max, row0, n);
        if (st) return st;
        for (int64_t i = 0; ok && i < n; ++i) {
            uint32_t *l = &lists[i * R_max];
            for (uint32_t j = 0; j < counts[i]; ++j)
                if (l[j] == inmem_start) l[j] = actual_start_point;
            ok = put<uint32_t>(f.f, counts[i]) && (counts[i] == 0 || fwrite(l, 4, counts[i], f.f) == counts[i]);
            size += 4ull * (1 + counts[i]);
            observed = counts[i] > observed ? counts[i] : observed;
        }
    }
    ok = ok && seek(f.f, 0) && put<uint64_t>(f.f, size) && put<uint32_t>(f.f, observed);
    if (!ok) return fail(DKM_ERR_IO, "%s: write failed", path);
    return DKM_OK;
}

// load_direct: the file's lists (start points last) must match the index's rows.
dkm_status dkm_load_graph(dkm_index *ix, const char *path, uint32_t *start_point) {
    if (!ix || !path) return fail(DKM_ERR_INVALID_ARG, "null argument");
    int64_t dim = 0, max_points = 0, num_start = 0, R = 0, R_max = 0;
    dkm_status st = dkm_index_shape(ix, &dim, &max_points, &num_start, &R, &R_max);
    if (st) return st;
    File f(path, "rb");
    if (!f.ok()) return fail(DKM_ERR_IO, "%s: cannot open", path);
    uint64_t fsize = 0, nstart = 0;
    uint32_t max_degree = 0, start = 0;
    if (!get(f.f, &fsize) || !get(f.f, &max_degree) || !get(f.f, &start) || !get(f.f, &nstart))
        return fail(DKM_ERR_IO, "%s: no graph header", path);
    if (fsize > file_size(f.f)) return fail(DKM_ERR_IO, "%s: header size %llu exceeds the file", path,
                                            (unsigned long long)fsize);
    // count the lists first (the header does not store it)
    uint64_t pos = kGraphHeader, total = 0;
    while (pos < fsize) {
        uint32_t c = 0;
        if (!get(f.f, &c)) return fail(DKM_ERR_IO, "%s: truncated list at byte %llu", path, (unsigned long long)pos);
        pos += 4ull * (1 + c);
        if (pos > fsize || !seek(f.f, pos)) return fail(DKM_ERR_IO, "%s: list runs past the end", path);
        ++total;
    }
    if ((int64_t)nstart != num_start || (int64_t)total != max_points + num_start)
        return fail(DKM_ERR_INDEX, "%s holds %llu lists with %llu start points; the index has %lld + %lld",
                    path, (unsigned long long)total, (unsigned long long)nstart, (long long)max_points,
                    (long long)num_start);
    if (!seek(f.f, kGraphHeader)) return fail(DKM_ERR_IO, "%s: seek failed", path);
    const int64_t chunk = 1 << 16;
    std::vector<uint32_t> lists((size_t)chunk * R_max), counts((size_t)chunk);
    for (int64_t row0 = 0; row0 < (int64_t)total; row0 += chunk) {
        const int64_t n = std::min<int64_t>(chunk, (int64_t)total - row0);
        for (int64_t i = 0; i < n; ++i) {
            uint32_t c = 0;
            if (!get(f.f, &c)) return fail(DKM_ERR_IO, "%s: short read", path);
            if (c > (uint32_t)R_max)
                return fail(DKM_ERR_INDEX, "%s: row %lld has %u neighbours (max degree %lld)", path,
                            (long long)(row0 + i), c, (long long)R_max);
            counts[i] = c;
            if (c && fread(&lists[i * R_max], 4, c, f.f) != c) return fail(DKM_ERR_IO, "%s: short read", path);
        }
        st = dkm_set_graph(ix, lists.data(), counts.data(), R_max, row0, n);
        if (st) return st;
    }
    if (start_point) *start_point = start;
    return DKM_OK;
}

dkm_status dkm_pq_pivots_write(const char *path, const float *pivots, int64_t K, int64_t dim,
                               const float *centroid, const int64_t *offsets, int64_t M) {
    if (!path || !pivots || !offsets || K <= 0 || dim <= 0 || M <= 0)
        return fail(DKM_ERR_INVALID_ARG, "bad argument");
    File f(path, "wb");
    if (!f.ok()) return fail(DKM_ERR_IO, "%s: cannot create", path);
    std::vector<float> zero;
    if (!centroid) {
        zero.assign((size_t)dim, 0.0f);
        centroid = zero.data();
    }
    std::vector<uint32_t> off32((size_t)M + 1);
    for (int64_t i = 0; i <= M; ++i) off32[i] = (uint32_t)offsets[i];
    uint64_t cumul[4];
    cumul[0] = kPivotMetadata;
    cumul[1] = cumul[0] + 8 + 4ull * K * dim;
    cumul[2] = cumul[1] + 8 + 4ull * dim;
    cumul[3] = cumul[2] + 8 + 4ull * (M + 1);
    std::vector<uint8_t> pad(kPivotMetadata, 0);
    bool ok = write_bin_block(f.f, cumul, 8, 4, 1);
    ok = ok && fwrite(pad.data(), 1, kPivotMetadata - 40, f.f) == kPivotMetadata - 40;
    ok = ok && write_bin_block(f.f, pivots, 4, (uint32_t)K, (uint32_t)dim);
    ok = ok && write_bin_block(f.f, centroid, 4, (uint32_t)dim, 1);
    ok = ok && write_bin_block(f.f, off32.data(), 4, (uint32_t)(M + 1), 1);
    if (!ok) return fail(DKM_ERR_IO, "%s: write failed", path);
    return DKM_OK;
}

static dkm_status pivots_layout(FILE *f, const char *path, uint64_t off[4], uint32_t *K,
                                uint32_t *dim, uint32_t *m1) {
    uint32_t r = 0, c = 0;
    std::vector<uint8_t> tab;
    dkm_status st = read_bin_block(f, 0, 8, &r, &c, &tab, path);
    if (st) return st;
    if (r != 4 || c != 1)
        return fail(DKM_ERR_PQ, "%s: offset table is %u x %u, expected 4 x 1", path, r, c);
    memcpy(off, tab.data(), 32);
    st = read_bin_block(f, off[0], 4, K, dim, nullptr, path);
    if (st) return st;
    if (*K > 256) return fail(DKM_ERR_PQ, "%s: %u centers, at most 256", path, *K);
    st = read_bin_block(f, off[1], 4, &r, &c, nullptr, path);
    if (st) return st;
    if (r != *dim || c != 1)
        return fail(DKM_ERR_PQ, "%s: centroid is %u x %u, expected %u x 1", path, r, c, *dim);
    st = read_bin_block(f, off[2], 4, m1, &c, nullptr, path);
    if (st) return st;
    if (c != 1 || *m1 < 2) return fail(DKM_ERR_PQ, "%s: chunk offsets are %u x %u", path, *m1, c);
    return DKM_OK;
}

dkm_status dkm_pq_pivots_read_shape(const char *path, int64_t *K, int64_t *dim, int64_t *M) {
    if (!path || !K || !dim || !M) return fail(DKM_ERR_INVALID_ARG, "null argument");
    File f(path, "rb");
    if (!f.ok()) return fail(DKM_ERR_IO, "%s: cannot open", path);
    uint64_t off[4];
    uint32_t k = 0, d = 0, m1 = 0;
    dkm_status st = pivots_layout(f.f, path, off, &k, &d, &m1);
    if (st) return st;
    *K = k;
    *dim = d;
    *M = (int64_t)m1 - 1;
    return DKM_OK;
}

// load_pq_pivots_bin: pivots with a non-zero centroid added back, and the chunk offsets.
dkm_status dkm_pq_pivots_read(const char *path, float *pivots, int64_t K, int64_t dim,
                              int64_t *offsets, int64_t M) {
    if (!path || !pivots || !offsets) return fail(DKM_ERR_INVALID_ARG, "null argument");
    File f(path, "rb");
    if (!f.ok()) return fail(DKM_ERR_IO, "%s: cannot open", path);
    uint64_t off[4];
    uint32_t k = 0, d = 0, m1 = 0;
    dkm_status st = pivots_layout(f.f, path, off, &k, &d, &m1);
    if (st) return st;
    if ((int64_t)k != K || (int64_t)d != dim || (int64_t)m1 != M + 1)
        return fail(DKM_ERR_DIM_MISMATCH, "%s: %u x %u pivots with %u chunks, expected %lld x %lld with %lld",
                    path, k, d, m1 - 1, (long long)K, (long long)dim, (long long)M);
    uint32_t r = 0, c = 0;
    std::vector<uint8_t> pv, cen, offs;
    if ((st = read_bin_block(f.f, off[0], 4, &r, &c, &pv, path))) return st;
    if ((st = read_bin_block(f.f, off[1], 4, &r, &c, &cen, path))) return st;
    if ((st = read_bin_block(f.f, off[2], 4, &r, &c, &offs, path))) return st;
    memcpy(pivots, pv.data(), pv.size());
    const float *cf = (const float *)cen.data();
    bool nonzero = false;
    for (int64_t j = 0; j < dim; ++j) nonzero = nonzero || cf[j] != 0.0f;
    if (nonzero)
        for (int64_t i = 0; i < K; ++i)
            for (int64_t j = 0; j < dim; ++j) pivots[i * dim + j] += cf[j];
    const uint32_t *o = (const uint32_t *)offs.data();
    for (int64_t i = 0; i <= M; ++i) offsets[i] = o[i];
    return DKM_OK;
}

}  // extern "C"

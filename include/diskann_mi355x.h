/*
 * diskann_mi355x: MI355X-native (gfx950) search-and-quantize path for DiskANN.
 *
 * C ABI. Plain pointers and sizes; the library owns all device state behind opaque
 * handles. Every entry point returns a dkm_status (0 = success) and leaves a
 * human-readable message in dkm_last_error() (thread-local) on failure.
 *
 * Reference interfaces replaced (paths relative to the SeliMeli/DiskANN v0.55.0 tree):
 *   - DiskANNIndex::search / search_with + Knn                diskann/src/graph/index.rs:2034-2070,
 *                                                             diskann/src/graph/search/knn_search.rs:76-217
 *   - SearchAccessor::{start_point_distances, expand_beam}    diskann/src/graph/glue.rs:125-254
 *   - FullPrecision / Hybrid / Quantized search strategies    diskann-providers/src/model/graph/provider/
 *                                                             async_/inmem/{full_precision,product}.rs
 *   - DistanceProvider::distance_comparer                     diskann-vector/src/distance/distance_provider.rs:44-46
 *   - FixedChunkPQTable::{populate_chunk_distances,
 *     compress_into}, QueryComputer                           diskann-providers/src/model/pq/fixed_chunk_pq_table.rs:149-455
 *   - TrainQuantizer::train (LightPQTrainingParameters)       diskann-quantization/src/product/train.rs:74-231
 *   - occlude_list / multi_insert                             diskann/src/graph/index.rs:2570-2782, 820-1035
 */
#ifndef DISKANN_MI355X_H
#define DISKANN_MI355X_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes mirror ANNErrorKind / KnnSearchError
 * (diskann/src/error/ann_error.rs:770-830, diskann/src/graph/search/knn_search.rs). */
typedef enum {
    DKM_OK = 0,
    DKM_ERR_INVALID_ARG = 1,
    DKM_ERR_K_ZERO = 2,             /* KnnSearchError::KZero */
    DKM_ERR_L_ZERO = 3,             /* KnnSearchError::LZero */
    DKM_ERR_L_LESS_THAN_K = 4,      /* KnnSearchError::LLessThanK */
    DKM_ERR_BEAM_WIDTH_ZERO = 5,    /* KnnSearchError::BeamWidthZero */
    DKM_ERR_DIM_MISMATCH = 6,       /* ANNErrorKind::DimensionMismatchError */
    DKM_ERR_OUT_OF_MEMORY = 7,
    DKM_ERR_DEVICE = 8,             /* HIP runtime failure */
    DKM_ERR_PQ = 9,                 /* ANNErrorKind::PQError */
    DKM_ERR_KMEANS = 10,            /* ANNErrorKind::KMeansError */
    DKM_ERR_INDEX = 11,             /* ANNErrorKind::IndexError (bad id / graph) */
    DKM_ERR_CAPACITY = 12,          /* a per-query device structure overflowed */
    DKM_ERR_NOT_READY = 13,         /* index lacks data the request needs */
    DKM_ERR_IO = 14                 /* ANNErrorKind::IOError (file missing, short, malformed) */
} dkm_status;

/* Metric numbering of diskann-vector/src/distance/metric.rs:11-20 (repr(C)). */
typedef enum { DKM_COSINE = 0, DKM_INNER_PRODUCT = 1, DKM_L2 = 2, DKM_COSINE_NORMALIZED = 3 } dkm_metric;

/* Search strategies: FullPrecision (CopyIds post-processing), Hybrid (PQ traversal +
 * FP32 Rerank), Quantized (PQ only). */
typedef enum { DKM_STRATEGY_FULL = 0, DKM_STRATEGY_HYBRID = 1, DKM_STRATEGY_QUANTIZED = 2 } dkm_strategy;

typedef struct dkm_index dkm_index;

/* -------------------------------------------------------------------------------- */
/* Library                                                                          */
/* -------------------------------------------------------------------------------- */
const char *dkm_last_error(void);
const char *dkm_version(void);
/* Select the HIP device used by subsequently created handles (default 0). */
dkm_status dkm_init(int device);

/* Device memory helpers for callers that keep inputs resident in HBM. */
dkm_status dkm_device_alloc(size_t bytes, void **ptr);
dkm_status dkm_device_free(void *ptr);
dkm_status dkm_memcpy_h2d(void *dst, const void *src, size_t bytes);
dkm_status dkm_memcpy_d2h(void *dst, const void *src, size_t bytes);
/* Device-to-device copy through a plain streaming kernel (synchronous); bytes a multiple of
 * 16, pointers 16-byte aligned. Also the reference point for memory-traffic counters. */
dkm_status dkm_device_copy(void *dst, const void *src, size_t bytes);
dkm_status dkm_synchronize(void);

/* -------------------------------------------------------------------------------- */
/* Distances (DistanceProvider::distance_comparer + evaluate_similarity)             */
/* -------------------------------------------------------------------------------- */
/* out[i*m + j] = distance(queries[i], base[ids[i*m + j]]) with the reference's exact
 * f32 reduction order. Host buffers. */
dkm_status dkm_distance_batch(int metric, int64_t dim, const float *queries, int64_t nq,
                              const float *base, int64_t nbase, const uint32_t *ids,
                              int64_t m, float *out);

/* -------------------------------------------------------------------------------- */
/* Product quantization                                                             */
/* -------------------------------------------------------------------------------- */
/* ChunkOffsets::partition (diskann-quantization/src/views.rs:233-302). */
dkm_status dkm_pq_chunk_offsets(int64_t dim, int64_t num_chunks, int64_t *offsets);
/* populate_chunk_distances: lut[q][m*K + c] for nq queries. Host buffers. */
dkm_status dkm_pq_lut(int metric, const float *pivots, int64_t K, int64_t dim,
                      const int64_t *offsets, int64_t M, const float *queries, int64_t nq,
                      float *lut);
/* compress_into: codes[i*M + m] = argmin_c SquaredL2(x_i,m , pivot_c,m). Host buffers. */
/* Direct PQ distances, replacing pq::distance::DistanceComputer::evaluate_similarity
 * (diskann-providers/src/model/pq/distance/dynamic.rs:170-205) over
 * FixedChunkPQTable::{l2_distance, inner_product, cosine_distance, qq_l2_distance,
 * qq_ip_distance, qq_cosine_distance} (fixed_chunk_pq_table.rs:32-56, 220-365), the
 * distances of the Hybrid prune (HybridComputer, .../async_/distances.rs:118-165).
 * out[i*m + j] = d(left_i, codes[ids[i*m + j]]) where left_i is the full-precision row
 * queries[i] (query_codes NULL) or the pivots of query_codes[i] (queries NULL); pivots
 * K x dim, codes n x M, chunk offsets M + 1. Cosine-normalized uses the cosine functions,
 * as the reference's table does. Same f32 operation order as the reference. */
dkm_status dkm_pq_distance_batch(int metric, const float *pivots, int64_t K, int64_t dim,
                                 const int64_t *offsets, int64_t M, const float *queries,
                                 const uint8_t *query_codes, int64_t nq, const uint8_t *codes,
                                 int64_t n, const uint32_t *ids, int64_t m, float *out);
dkm_status dkm_pq_encode(const float *pivots, int64_t K, int64_t dim, const int64_t *offsets,
                         int64_t M, const float *x, int64_t n, uint8_t *codes);
/* PQ training, LightPQTrainingParameters::train (diskann-quantization/src/product/
 * train.rs:109-231) as generate_pq_pivots_from_membuf calls it
 * (diskann-providers/src/model/pq/pq_construction.rs:163-253): chunk offsets by
 * ChunkOffsets::partition, then per chunk k-means++ seeding (plusplus.rs:381-498;
 * DatasetTooSmall / InsufficientDiversity are suppressed and the unpicked centres stay
 * zero, SawInfinity fails with DKM_ERR_PQ) and `iters` Lloyd iterations (lloyds.rs:372-426).
 * pivots: K x dim (reference layout), offsets: M+1. Host buffers.
 * dkm_pq_train_draws takes the k-means++ random draws from the caller: draws[m * K + c] in
 * [0, 1) for chunk m; the first centre is row floor(draws[m * K] * n) and step c's threshold
 * is draws[m * K + c] times the scale of rand's UniformFloat over [0, total). dkm_pq_train
 * fills the draws from the library's SplitMix64 stream seeded by `seed` (the reference
 * draws from rand / ChaCha12, which this library does not reproduce). */
dkm_status dkm_pq_train(const float *data, int64_t n, int64_t dim, int64_t M, int64_t K,
                        int64_t iters, uint64_t seed, float *pivots, int64_t *offsets);
dkm_status dkm_pq_train_draws(const float *data, int64_t n, int64_t dim, int64_t M, int64_t K,
                              int64_t iters, const double *draws, float *pivots,
                              int64_t *offsets);

/* -------------------------------------------------------------------------------- */
/* Index (DefaultProvider + DiskANNIndex)                                           */
/* -------------------------------------------------------------------------------- */
/* max_points data rows plus `num_start` frozen start rows at ids max_points.. .
 * R = pruned degree, R_max = max degree with slack (floor(1.3 R) in the reference
 * benchmark, diskann-benchmark/src/inputs/graph_index.rs:690-692). */
dkm_status dkm_index_create(int metric, int64_t dim, int64_t max_points, int64_t num_start,
                            int64_t R, int64_t R_max, dkm_index **out);
void dkm_index_destroy(dkm_index *ix);

/* Rows [row0, row0+n) of the base vectors (n x dim, host). */
dkm_status dkm_set_vectors(dkm_index *ix, const float *x, int64_t row0, int64_t n);
/* Frozen start points: copies `num_start` vectors to rows max_points.. and makes them
 * the search entry points (set_start_points, .../inmem/provider.rs:393-421). */
dkm_status dkm_set_start_points(dkm_index *ix, const float *vectors);
/* Adjacency rows [row0, row0+n): lists[i*stride + j] for j < counts[i]. Lists must be
 * duplicate-free, ids < max_points + num_start, counts <= R_max. */
dkm_status dkm_set_graph(dkm_index *ix, const uint32_t *lists, const uint32_t *counts,
                         int64_t stride, int64_t row0, int64_t n);
dkm_status dkm_get_graph(dkm_index *ix, uint32_t *lists, uint32_t *counts, int64_t stride,
                         int64_t row0, int64_t n);

/* PQ schema for the index (pivots K x dim, offsets M+1), then codes for rows. */
dkm_status dkm_set_pq(dkm_index *ix, const float *pivots, int64_t K, const int64_t *offsets,
                      int64_t M);
dkm_status dkm_set_codes(dkm_index *ix, const uint8_t *codes, int64_t row0, int64_t n);
/* Encode every stored row (data + start points) on the device with the index's PQ. */
dkm_status dkm_encode_index(dkm_index *ix);
dkm_status dkm_get_codes(dkm_index *ix, uint8_t *codes, int64_t row0, int64_t n);

/* Batched k-NN search (Knn::new(k, L, bw) per query; k <= L, k, L, bw >= 1).
 * k and L are per-query arrays (length nq); out_ids/out_dists have k_stride entries per
 * query, of which out_counts[i] are valid. out_cmps/out_hops may be NULL.
 * Host buffers; the call blocks until the results are copied back. */
dkm_status dkm_search_batch(dkm_index *ix, int strategy, const float *queries, int64_t nq,
                            const uint32_t *k, const uint32_t *L, uint32_t beam_width,
                            int64_t k_stride, uint32_t *out_ids, float *out_dists,
                            uint32_t *out_counts, uint32_t *out_cmps, uint32_t *out_hops);

/* Disk-index search (DiskIndexSearcher: DiskAccessor, diskann-disk/src/search/provider/
 * disk_provider.rs:666-721, with RerankAndFilter, :323-388) over the index's rows as the
 * disk layout's node records (vector + adjacency list) with the PQ codes in HBM:
 *   - the single start point is the data row `medoid` (the graph header's medoid), with
 *     its PQ distance;
 *   - PQ distances use the transposed pivot table (quantizer_preprocess.rs:17-64,
 *     TransposedTable::process_into: per pivot a fold of fma over the chunk's dimensions;
 *     cosine and cosine-normalized indexes use the L2 table) summed in chunk order
 *     (pq_dist_lookup, fixed_chunk_pq_table.rs:469-598);
 *   - expand_beam loads at most io_limit beam nodes per query in total (search_io_limit);
 *     beam nodes past the budget are taken (hops) but not expanded;
 *   - every queue entry, the start point included, is reranked with its full-precision
 *     distance and sorted (ties keep queue order; the reference's sort_unstable leaves
 *     them unspecified); the first k are returned.
 * The per-query loads (the reference's io count) are min(hops, io_limit). Needs the PQ
 * schema and codes; host buffers as dkm_search_batch. */
dkm_status dkm_disk_search_batch(dkm_index *ix, const float *queries, int64_t nq,
                                 const uint32_t *k, const uint32_t *L, uint32_t beam_width,
                                 uint32_t medoid, uint64_t io_limit, int64_t k_stride,
                                 uint32_t *out_ids, float *out_dists, uint32_t *out_counts,
                                 uint32_t *out_cmps, uint32_t *out_hops);

/* Same, with every buffer in device memory (dkm_device_alloc); asynchronous on the
 * handle's stream. max_k / max_L bound the launch (max_k <= k_stride, max_L <= 4096); each
 * query's own k / L are read on the device and a query breaking Knn::new's rules (k, L >= 1,
 * k <= L) or the bounds (L <= max_L, k <= k_stride) is not searched: its count is 0 and
 * dkm_search_status reports DKM_ERR_INVALID_ARG for it; no other query's outputs are
 * touched. Launches of 2048 queries or more run them longest L first (a device sort of the
 * launch's L values); no query's result depends on that order. */
dkm_status dkm_search_batch_device(dkm_index *ix, int strategy, const float *d_queries,
                                   int64_t nq, const uint32_t *d_k, const uint32_t *d_L,
                                   uint32_t max_k, uint32_t max_L, uint32_t beam_width,
                                   int64_t k_stride, uint32_t *d_out_ids, float *d_out_dists,
                                   uint32_t *d_out_counts, uint32_t *d_out_cmps,
                                   uint32_t *d_out_hops);
dkm_status dkm_index_synchronize(dkm_index *ix);
/* Waits for the handle's stream and checks the per-query status of the most recent
 * dkm_search_batch_device(_stats) launch: DKM_OK, or the first failing query's error as
 * dkm_search_batch reports it (DKM_ERR_INVALID_ARG for a refused k / L, DKM_ERR_CAPACITY for
 * a visited-set overflow, DKM_ERR_INDEX for the iteration bound), with its position in
 * *first_failed (-1 when none; may be NULL). Outputs of a failed query are not valid. */
dkm_status dkm_search_status(dkm_index *ix, int64_t *first_failed);

/* Allocation-free variant of dkm_search_batch_device (same arguments, results and per-query
 * k / L checks) for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc). No reference counterpart.
 * The launch's scratch (visited-set regions, the longest-L-first order, spherical query
 * records, and the status words when d_out_status is NULL) comes from a fixed 64 MiB device
 * array the library holds on each GPU until the process ends; the call makes no device
 * allocation. A launch needing more (about nq x 64 KiB at max_L <= 1024) is refused with
 * DKM_ERR_CAPACITY. d_out_status (nq words, may be NULL) receives each query's status:
 * 0 = ok, 1 = visited-set overflow, 2 = iteration bound, 4 = k / L refused. Every call on a
 * GPU shares that GPU's array: calls that overlap on one GPU (any handle, any stream) race,
 * so the caller serialises them (one stream, or synchronise between calls). */
dkm_status dkm_search_batch_device_noalloc(dkm_index *ix, int strategy, const float *d_queries,
                                           int64_t nq, const uint32_t *d_k, const uint32_t *d_L,
                                           uint32_t max_k, uint32_t max_L, uint32_t beam_width,
                                           int64_t k_stride, uint32_t *d_out_ids,
                                           float *d_out_dists, uint32_t *d_out_counts,
                                           uint32_t *d_out_cmps, uint32_t *d_out_hops,
                                           uint32_t *d_out_status);

/* HIP graph of one batched search launch: dkm_search_graph_create validates and captures
 * the launch dkm_search_batch_device_stats would make with these arguments (the kernels
 * only: no copies or memsets) into a graph with its own visited-set regions and status
 * words; dkm_search_graph_launch replays it on the handle's stream, reading the same query
 * and k / L buffers and writing the same outputs (refresh their contents between replays;
 * d_out_degsum may be NULL). A replay is refused with DKM_ERR_NOT_READY when the index
 * would now launch differently (tuning, quantizer or buffers changed since the capture).
 * Results equal dkm_search_batch_device's. No reference counterpart: the reference has no
 * device launches to capture; this serves callers that repeat fixed-shape batches. */
typedef struct dkm_search_graph dkm_search_graph;
dkm_status dkm_search_graph_create(dkm_index *ix, int strategy, const float *d_queries,
                                   int64_t nq, const uint32_t *d_k, const uint32_t *d_L,
                                   uint32_t max_k, uint32_t max_L, uint32_t beam_width,
                                   int64_t k_stride, uint32_t *d_out_ids, float *d_out_dists,
                                   uint32_t *d_out_counts, uint32_t *d_out_cmps,
                                   uint32_t *d_out_hops, uint32_t *d_out_degsum,
                                   dkm_search_graph **out);
dkm_status dkm_search_graph_launch(dkm_search_graph *g);
/* dkm_search_status for the graph's most recent replay. */
dkm_status dkm_search_graph_status(dkm_search_graph *g, int64_t *first_failed);
void dkm_search_graph_destroy(dkm_search_graph *g);

/* Kernel timing with HIP events on the handle's stream: enable, then read the count and
 * total milliseconds of the search kernel launches since the last reset. */
dkm_status dkm_profile(dkm_index *ix, int enable);
dkm_status dkm_profile_read(dkm_index *ix, int64_t *launches, double *total_ms);

/* Graph construction with multi_insert semantics (diskann/src/graph/index.rs:820-1035):
 * the rows `ids` (vectors already set) are inserted in consecutive batches of `batch`;
 * each batch searches the pre-batch graph with L_build, prunes (RobustPrune, alpha,
 * max_occlusion), sets its lists and applies back edges. intra_batch: -1 = All,
 * 0 = None, n = Max(n) intra-batch candidates. batch == 1 is insert() semantics. */
dkm_status dkm_build(dkm_index *ix, const uint32_t *ids, int64_t n, int64_t batch,
                     int64_t intra_batch, uint32_t L_build, float alpha, int64_t max_occlusion,
                     int saturate);

/* RobustPrune of caller-given candidate lists: the PruneStrategy / PruneAccessor seam
 * (diskann/src/graph/glue.rs:839-966) with robust_prune_list (diskann/src/graph/index.rs:
 * 2402-2450) and occlude_list (:2570-2782). For each location i (a row id), the candidates
 * lists[i * list_stride, + list_len[i]) (ids, self occurrences dropped, order kept) get
 * their distances to the location's vector, the pool is sorted (distance, then list
 * position) and cut to max_occlusion, and occlude_list with alpha picks at most R ids,
 * saturating when `saturate` and alpha > 1 or when force_saturate
 * (prune::Options::force_saturate). The index's Hybrid prune setting
 * (dkm_set_build_hybrid) applies: fill order location, then the list. The pruned lists go
 * to out_lists (n rows of out_stride >= R ids) with their lengths in out_counts; an empty
 * list prunes to nothing. Host buffers; the graph is not read or changed. */
dkm_status dkm_prune_batch(dkm_index *ix, const uint32_t *locations, int64_t n,
                           const uint32_t *lists, const uint32_t *list_len, int64_t list_stride,
                           float alpha, int64_t max_occlusion, int saturate, int force_saturate,
                           uint32_t *out_lists, uint32_t *out_counts, int64_t out_stride);

/* DiskANNIndex::prune_range (diskann/src/graph/index.rs:2788-2830): every id whose
 * adjacency list is longer than R is pruned with robust_prune_list over its own list
 * (no forced saturation) and its list replaced; shorter lists are left alone. A prune reads
 * only vectors, so the ids are pruned together and the result equals the reference's
 * sequential loop. */
dkm_status dkm_prune_range(dkm_index *ix, const uint32_t *ids, int64_t n, float alpha,
                           int64_t max_occlusion, int saturate);

/* 8-bit scalar quantization (diskann-quantization/src/scalar). A compressed vector uses
 * the reference's canonical layout (meta/vector.rs:384-434): an f32 compensation, then one
 * code byte per dimension, 4 + dim bytes per row.
 * dkm_sq_train: ScalarQuantizationParameters::train (scalar/train.rs:75-88) over n x dim
 *   host rows: scale = 2 * standard_deviations * max column std, shift = column mean -
 *   scale / 2, mean_norm = average row norm (f64 statistics in row order).
 * dkm_sq_compress: ScalarQuantizer::compress_into into MutCompensatedVectorRef
 *   (scalar/quantizer.rs:380-432) for n host rows; DKM_ERR_INVALID_ARG when a row holds a
 *   NaN (the codes are written regardless, as the reference does).
 * dkm_sq_distance_batch: CompensatedSquaredL2 / CompensatedIP / CompensatedCosineNormalized
 *   as distances (scalar/vectors.rs:182-470, the DistanceComputer of
 *   diskann-providers/.../inmem/scalar.rs:260-297): out[i * m + j] = d(left[i],
 *   right[ids[i * m + j]]); metric L2, inner product or cosine-normalized. */
dkm_status dkm_sq_train(const float *data, int64_t n, int64_t dim, double standard_deviations,
                        float *out_scale, float *out_shift, float *out_mean_norm);
dkm_status dkm_sq_compress(float scale, const float *shift, int64_t dim, const float *x,
                           int64_t n, uint8_t *out);
dkm_status dkm_sq_distance_batch(int metric, float scale, const float *shift, int64_t dim,
                                 const uint8_t *left, int64_t nl, const uint8_t *right,
                                 int64_t nr, const uint32_t *ids, int64_t m, float *out);

/* Scalar-quantized store of an index (SQStore, diskann-providers/src/model/graph/provider/
 * async_/inmem/scalar.rs:83-253). After dkm_set_sq the quantized strategies traverse with
 * 8-bit codes: DKM_STRATEGY_HYBRID = the reference's Quantized strategy over a
 * full-precision provider (SQ traversal, FilterStartPoints -> Rerank, scalar.rs:601-637),
 * DKM_STRATEGY_QUANTIZED = Quantized without a full-precision store (SQ distances
 * returned, scalar.rs:640-672). The query is compressed per search, rescaled to mean_norm
 * first for inner product (query_computer with allow_rescale). dkm_encode_index_sq
 * compresses every stored row (SQStore::set_vector); dkm_set_sq_codes / dkm_get_sq_codes
 * move canonical rows; dkm_use_pq switches the quantized strategies back to PQ. */
dkm_status dkm_set_sq(dkm_index *ix, float scale, const float *shift, float mean_norm);
dkm_status dkm_encode_index_sq(dkm_index *ix);
dkm_status dkm_set_sq_codes(dkm_index *ix, const uint8_t *codes, int64_t row0, int64_t n);
dkm_status dkm_get_sq_codes(dkm_index *ix, uint8_t *codes, int64_t row0, int64_t n);
dkm_status dkm_use_pq(dkm_index *ix);

/* 1-bit spherical quantization (diskann-quantization/src/spherical). A quantizer is the
 * reference's SphericalQuantizer (quantizer.rs:36-78; serialized form
 * schemas/spherical_v001.fbs): metric (DKM_L2, DKM_INNER_PRODUCT or DKM_COSINE), centroid
 * (dim floats), mean_norm, pre_scale and a transform: 0 = NullTransform, 1 =
 * PaddingHadamard (signs0: dim flags, padded_dim a power of two >= dim), 2 =
 * DoubleHadamard (signs0: dim flags, signs1: max(dim, output dim) flags); flags are 1 for
 * a flipped sign, as in the serialized bool vectors; subsample (optional, strictly
 * increasing) selects the output coordinates. A compressed vector is the canonical row
 * of the reference's Impl<1>::compress (iface.rs:1955-1966, DataMut::from_canonical_back):
 * ceil(out_dim / 8) bytes of bits (little-endian within each byte), then f16 inner-product
 * correction, f16 metric term and u16 bit count.
 * dkm_sph_train_stats: the deterministic part of SphericalQuantizer::train
 *   (quantizer.rs:225-320): centroid (column means in row order; cosine: of the
 *   unit-normalised rows) and mean norm (1 for cosine); the transform's random signs
 *   are the caller's.
 * dkm_sph_compress: CompressIntoWith<DataMut<1>> (quantizer.rs:591-642, 779-857).
 * dkm_sph_distance_batch: the quantizer's distance_computer on two compressed rows
 *   (CompensatedSquaredL2 / CompensatedIP / CompensatedCosine, vectors.rs:476-897).
 * dkm_sph_query_distance_batch: fused_query_computer (iface.rs:1898-1950) for layout 0 =
 *   SameAsData, 1 = FourBitTransposed, 2 = FullPrecision, with allow_rescale. */
typedef struct dkm_sph dkm_sph;
dkm_status dkm_sph_create(int metric, int64_t dim, const float *centroid, float mean_norm,
                          float pre_scale, int transform, const uint8_t *signs0,
                          const uint8_t *signs1, int64_t n_signs1, int64_t padded_dim,
                          const uint32_t *subsample, int64_t n_subsample, dkm_sph **out);
void dkm_sph_destroy(dkm_sph *q);
dkm_status dkm_sph_dims(const dkm_sph *q, int64_t *out_dim, int64_t *code_bytes);
dkm_status dkm_sph_train_stats(int metric, const float *data, int64_t n, int64_t dim,
                               float *out_centroid, float *out_mean_norm);
dkm_status dkm_sph_compress(const dkm_sph *q, const float *x, int64_t n, uint8_t *out);
dkm_status dkm_sph_distance_batch(const dkm_sph *q, const uint8_t *left, int64_t nl,
                                  const uint8_t *right, int64_t nr, const uint32_t *ids,
                                  int64_t m, float *out);
dkm_status dkm_sph_query_distance_batch(const dkm_sph *q, int layout, int allow_rescale,
                                        const float *queries, int64_t nq, const uint8_t *codes,
                                        int64_t n, const uint32_t *ids, int64_t m, float *out);

/* Spherical store of an index (SphericalStore, diskann-providers/src/model/graph/provider/
 * async_/inmem/spherical.rs:83-240). After dkm_set_spherical the quantized strategies
 * traverse with 1-bit codes and the given query layout (0 = SameAsData, 1 =
 * FourBitTransposed), rescaling inner-product queries as the reference's search
 * strategy does (Quantized::search, spherical.rs:505-531): DKM_STRATEGY_HYBRID = search
 * over a full-precision provider (traversal, FilterStartPoints -> Rerank),
 * DKM_STRATEGY_QUANTIZED = without one (1-bit distances returned). The quantizer's metric
 * must be the index metric. dkm_encode_index_sph compresses every stored row;
 * dkm_set_sph_codes / dkm_get_sph_codes move canonical rows; dkm_use_pq switches the
 * quantized strategies back to PQ. */
dkm_status dkm_set_spherical(dkm_index *ix, const dkm_sph *q, int query_layout);
dkm_status dkm_encode_index_sph(dkm_index *ix);
dkm_status dkm_set_sph_codes(dkm_index *ix, const uint8_t *codes, int64_t row0, int64_t n);
dkm_status dkm_get_sph_codes(dkm_index *ix, uint8_t *codes, int64_t row0, int64_t n);

/* 8-bit MinMax quantization (diskann-quantization/src/minmax): MinMaxQuantizer::new(
 * transform, grid_scale) with the transform given as for dkm_sph_create (0 null, 1
 * padding Hadamard, 2 double Hadamard). A code row is the canonical front layout of
 * minmax::Data<8> (vectors.rs:40-115): u32 dim, f32 b, f32 n, f32 a, f32 |reconstruction|^2,
 * then out_dim code bytes.
 * dkm_minmax_compress: CompressInto<DataMutRef<8>> (minmax/quantizer.rs:45-170), with the
 *   squared reconstruction loss per row (out_loss may be NULL).
 * dkm_minmax_distance_batch: MinMaxL2Squared / MinMaxIP / MinMaxCosine /
 *   MinMaxCosineNormalized on two code rows (vectors.rs:206-470), metric L2, inner
 *   product, cosine or cosine-normalized.
 * dkm_minmax_query_distance_batch: the same against full-precision queries
 *   (CompressInto<FullQueryMut>, quantizer.rs:178-210). */
typedef struct dkm_minmax dkm_minmax;
dkm_status dkm_minmax_create(int64_t dim, float grid_scale, int transform, const uint8_t *signs0,
                             const uint8_t *signs1, int64_t n_signs1, int64_t padded_dim,
                             const uint32_t *subsample, int64_t n_subsample, dkm_minmax **out);
void dkm_minmax_destroy(dkm_minmax *q);
dkm_status dkm_minmax_dims(const dkm_minmax *q, int64_t *out_dim, int64_t *code_bytes);
dkm_status dkm_minmax_compress(const dkm_minmax *q, const float *x, int64_t n, uint8_t *out,
                               float *out_loss);
dkm_status dkm_minmax_distance_batch(const dkm_minmax *q, int metric, const uint8_t *left,
                                     int64_t nl, const uint8_t *right, int64_t nr,
                                     const uint32_t *ids, int64_t m, float *out);
dkm_status dkm_minmax_query_distance_batch(const dkm_minmax *q, int metric, const float *queries,
                                           int64_t nq, const uint8_t *codes, int64_t n,
                                           const uint32_t *ids, int64_t m, float *out);

/* compute_closest_centers (diskann-disk/src/utils/math_util.rs:246-385) of the partitioned
 * disk-index build: for each of the n rows of data (n x dim), the k closest of the K
 * centres (K x dim) by d = (|x|^2 + |c|^2) + (-2 <x, c>) with FastL2NormSquared norms
 * (diskann-vector/src/norm.rs:45-56); out[i*k + t] is the t-th closest (ascending d; equal d:
 * lower centre index first; k = 1 is the first index with the strictly smallest d, as
 * compute_closest_centers_in_block's loop). The reference's dot products come from faer's
 * sgemm, whose summation order is not pinned; here they are fma chains in dimension order.
 * 1 <= k <= min(K, 16). Host buffers. */
dkm_status dkm_closest_centers(const float *data, int64_t n, int64_t dim, const float *centers,
                               int64_t K, int64_t k, uint32_t *out);

/* Hybrid prune for dkm_build (common::Hybrid, diskann-providers/src/model/graph/provider/
 * async_/common.rs:385-405; HybridPruneAccessor, inmem/product.rs:143-230): each prune
 * keeps the first max_fp_vecs_per_prune ids of its fill order (robust_prune_with:
 * self, intra-batch extras, search record, at most max_occlusion ids; robust_prune_list:
 * location, list) in full precision and compares every other id through its PQ code
 * (direct distances, HybridComputer). Negative: no Hybrid, every vector full precision
 * (the default). Hybrid::new(None) is any value at least the longest fill order. Needs the
 * PQ schema and the codes of every row before dkm_build. */
dkm_status dkm_set_build_hybrid(dkm_index *ix, int64_t max_fp_vecs_per_prune);

/* Scalar-quantized build: the Quantized strategy over an SQStore (diskann-providers/src/
 * model/graph/provider/async_/inmem/scalar.rs:676-720). dkm_build's insert search traverses
 * with the SQ query computer, as the Quantized search does (insert_search_accessor =
 * search_accessor, diskann/src/graph/glue.rs:933-940; inner-product queries rescaled to the
 * mean norm), and every prune distance of dkm_build, dkm_prune_batch and dkm_prune_range is
 * the compensated distance between the two rows' stored codes (PruneAccessor,
 * scalar.rs:370-446). The reference runs it with single inserts (batch 1); batches use the
 * same distances. Needs dkm_set_sq and the SQ codes of every row, start rows included
 * (dkm_encode_index_sq); excludes the Hybrid prune. enable = 0: full precision. */
dkm_status dkm_set_build_sq(dkm_index *ix, int enable);

/* Spherical build: Quantized::build() over a SphericalStore (diskann-providers/src/model/
 * graph/provider/async_/inmem/spherical.rs:515-533, 605-650). dkm_build's insert search
 * traverses with the spherical codes, the query compressed in the data layout (SameAsData)
 * without the inner-product rescale (is_search = false), and every prune distance of
 * dkm_build, dkm_prune_batch and dkm_prune_range is the data x data distance of the two
 * rows' codes (PruneAccessor, spherical.rs:293-350). Needs dkm_set_spherical and the codes of
 * every row, start rows included (dkm_encode_index_sph); excludes the SQ build and the Hybrid
 * prune. enable = 0: full precision. */
dkm_status dkm_set_build_spherical(dkm_index *ix, int enable);

/* compute_medoid over the data rows (diskann-utils/src/sampling/medoid.rs:15-47):
 * writes the medoid row's vector (dim floats) and its row id (-1 if none qualifies). */
dkm_status dkm_medoid(dkm_index *ix, float *out_vector, int64_t *out_row);

/* Exact FP64 brute-force top-k over the data rows (recall ground truth): out_ids and
 * out_dists hold k entries per query, ties broken by smaller id. Host buffers. */
dkm_status dkm_bruteforce(dkm_index *ix, const float *queries, int64_t nq, int64_t k,
                          uint32_t *out_ids, double *out_dists);

/* Same as dkm_search_batch_device plus the per-query sum of the stored neighbor counts
 * of every expanded node (the deg_read term of the traversal's byte count). */
dkm_status dkm_search_batch_device_stats(dkm_index *ix, int strategy, const float *d_queries,
                                         int64_t nq, const uint32_t *d_k, const uint32_t *d_L,
                                         uint32_t max_k, uint32_t max_L, uint32_t beam_width,
                                         int64_t k_stride, uint32_t *d_out_ids,
                                         float *d_out_dists, uint32_t *d_out_counts,
                                         uint32_t *d_out_cmps, uint32_t *d_out_hops,
                                         uint32_t *d_out_degsum);

/* -------------------------------------------------------------------------------- */
/* Serving: continuous batching                                                        */
/* -------------------------------------------------------------------------------- */
/* The reference serves queries by calling DiskANNIndex::search concurrently from many
 * async workers over one shared index (diskann-benchmark-core/src/search/api.rs:400-430,
 * diskann/src/graph/index.rs:2034-2070); a server is the device-side form of that loop.
 * Requests go into a ring of `capacity` entries in pinned host memory. Every submit and
 * poll call collects finished launches and launches the requests waiting at that moment
 * as one batch (uneven k and L in one launch) on one of eight streams, as long as at most
 * `slots` requests are in flight: requests that arrive while others run start in the
 * next launch and overlap with them on the device, filling compute units as earlier
 * queries finish. The streams overlap on the device only with a hardware queue each: run
 * the process with GPU_MAX_HW_QUEUES >= 10 (HIP's default, 4, makes the streams share two
 * queues and serialise; the Python package sets 12 unless the variable is set).
 * Admission is by device memory: each stream owns the visited-set
 * regions (HBM) of slots / 8 requests, allocated at start; submits beyond the free ring
 * entries are refused with DKM_ERR_CAPACITY (back-pressure). Requests need
 * k <= max_k and k <= L <= max_L (L < k is rejected as Knn::new does,
 * diskann/src/graph/search/knn_search.rs:101-106). Each result equals
 * dkm_search_batch for the same query, k, L, beam width and strategy, and is returned by
 * poll as soon as its own search ends (not when the rest of its launch does). slots = 0 picks
 * twice the number of queries the device holds at once. Progress needs the caller to
 * keep calling submit or poll. The index must not be modified while a server runs.
 * Thread-safe per server. */
typedef struct dkm_server dkm_server;
dkm_status dkm_server_start(dkm_index *ix, int strategy, uint32_t max_k, uint32_t max_L,
                            uint32_t beam_width, int64_t capacity, int64_t slots,
                            dkm_server **out);
dkm_status dkm_server_slots(dkm_server *sv, int64_t *slots);
dkm_status dkm_server_submit(dkm_server *sv, const float *query, uint32_t k, uint32_t L,
                             uint64_t tag);
/* Submits requests in order until the ring is full: *accepted of the n were taken (the
 * rest may be resubmitted later). All n are validated first; an invalid one rejects the
 * whole call with its status. */
dkm_status dkm_server_submit_batch(dkm_server *sv, const float *queries, int64_t n,
                                   const uint32_t *k, const uint32_t *L, const uint64_t *tags,
                                   int64_t *accepted);
/* Returns up to max_results completed requests (any order): tag, max_k ids / distances
 * per result (counts[i] valid), comparisons, hops and the request's latency in
 * microseconds from submit to the poll that returned it. A request whose search failed
 * (visited-set overflow, iteration bound) is consumed and its error returned; the
 * *n_out results before it are valid. */
dkm_status dkm_server_poll(dkm_server *sv, int64_t max_results, uint64_t *tags,
                           uint32_t *ids, float *dists, uint32_t *counts, uint32_t *cmps,
                           uint32_t *hops, double *latency_us, int64_t *n_out);
/* Requests submitted and not yet returned by dkm_server_poll. */
dkm_status dkm_server_pending(dkm_server *sv, int64_t *pending);
/* Waits for the launches in flight, then releases the server and its memory; results
 * not yet returned by dkm_server_poll are dropped. */
dkm_status dkm_server_stop(dkm_server *sv);

/* Index shape as created (any output may be NULL). */
dkm_status dkm_index_shape(dkm_index *ix, int64_t *dim, int64_t *max_points, int64_t *num_start,
                           int64_t *R, int64_t *R_max);

/* -------------------------------------------------------------------------------- */
/* On-disk formats (host side; little-endian)                                       */
/* -------------------------------------------------------------------------------- */
/* .bin matrix: u32 npoints, u32 ndims, then row-major values of elem_size bytes
 * (diskann-utils/src/io.rs:60-100 write_bin / read_bin). */
dkm_status dkm_bin_write(const char *path, const void *data, int64_t elem_size, int64_t nrows,
                         int64_t ncols);
dkm_status dkm_bin_read_shape(const char *path, int64_t *nrows, int64_t *ncols);
dkm_status dkm_bin_read(const char *path, void *out, int64_t elem_size, int64_t nrows,
                        int64_t ncols);
/* Streams an f32 .bin (<= max_points rows of the index's dim) into the data rows. */
dkm_status dkm_set_vectors_from_bin(dkm_index *ix, const char *path);
/* Graph file (diskann-providers/src/storage/bin.rs:310-380 save_graph, :230-308
 * load_graph). dkm_save_graph / dkm_load_graph are the in-memory index's save_direct /
 * load_direct (simple_neighbor_provider.rs:143-185): every row, start points last, the
 * header's max degree = R_max. dkm_save_graph_disk is the disk-index adaptor
 * (simple_neighbor_provider.rs:313-356): data rows only, the in-memory start id replaced
 * by actual_start_point, no start points in the header, observed max degree. */
dkm_status dkm_save_graph(dkm_index *ix, const char *path, uint32_t start_point);
dkm_status dkm_save_graph_disk(dkm_index *ix, const char *path, uint32_t actual_start_point);
dkm_status dkm_load_graph(dkm_index *ix, const char *path, uint32_t *start_point);
/* PQ pivots file (diskann-providers/src/storage/pq_storage.rs:93-149 write_pivot_data,
 * :278-349 load_pq_pivots_bin): centroid may be NULL (written as zeros); on read a
 * non-zero centroid is added to every pivot row. PQ codes files are u8 .bin matrices. */
dkm_status dkm_pq_pivots_write(const char *path, const float *pivots, int64_t K, int64_t dim,
                               const float *centroid, const int64_t *offsets, int64_t M);
dkm_status dkm_pq_pivots_read_shape(const char *path, int64_t *K, int64_t *dim, int64_t *M);
dkm_status dkm_pq_pivots_read(const char *path, float *pivots, int64_t K, int64_t dim,
                              int64_t *offsets, int64_t M);

/* -------------------------------------------------------------------------------- */
/* Large-dimension k-means (FP32 reference assignment or BF16 MFMA assignment)       */
/* -------------------------------------------------------------------------------- */
/* Lloyd iterations of diskann-quantization/src/algorithms/kmeans/lloyds.rs:28-426 on
 * device-resident f32 data (n x dim, row-major): assignment by c.c - x.c - x.c + x.x,
 * strict < in centre order; update = f64 sums in row order / max(count, 1) (empty
 * clusters become zero). DKM_PRECISION_F32 computes the assignment in the reference's
 * FP32 arithmetic (sequential FMA over the dimensions); DKM_PRECISION_BF16 rounds data
 * and centres to bf16 and computes x.c as an MFMA GEMM with FP32 accumulation (c.c stays
 * FP32), ties to the smaller centre index. d_centroids holds the K initial centres on
 * entry and the final ones on return; ms_* receive the summed device time of each step
 * (may be NULL). Synchronous. */
typedef enum { DKM_PRECISION_F32 = 0, DKM_PRECISION_BF16 = 1, DKM_PRECISION_FP8 = 2 } dkm_precision;
dkm_status dkm_kmeans_device(const float *d_x, int64_t n, int64_t dim, int64_t K, int64_t iters,
                             int precision, float *d_centroids, uint32_t *d_assign,
                             double *ms_assign, double *ms_update);
/* One assignment pass (the nearest centre and its score c.c - 2 x.c + x.x, or without the
 * x.x term for BF16) for the error report against FP32 / FP64. d_dist may be NULL. */
dkm_status dkm_kmeans_assign_device(const float *d_x, int64_t n, int64_t dim, const float *d_c,
                                    int64_t K, int precision, uint32_t *d_assign,
                                    float *d_dist);

/* -------------------------------------------------------------------------------- */
/* Candidate pre-scoring (BF16 / FP8 MFMA) with exact FP32 rescore                   */
/* -------------------------------------------------------------------------------- */
/* Exhaustive top-k over device-resident f32 rows (n x dim, dim <= 768) for inner product,
 * L2 or cosine-normalized: every (query, row) pair is scored approximately as an MFMA GEMM
 * with FP32 accumulation (-q.x, or |x|^2 - 2 q.x for L2) on bf16 operands
 * (DKM_PRECISION_BF16) or on FP8 E4M3 operands (DKM_PRECISION_FP8: the rows and each
 * batch of queries are multiplied by a power of two that brings their largest magnitude to
 * at most 448, rounded to nearest even, and the scales are divided out of the scores);
 * the `candidates` best rows
 * of each query by that score are re-scored exactly with the reference's f32 distance
 * functions (as Rerank, diskann-providers/src/model/graph/provider/async_/inmem/
 * full_precision.rs:357-404, and dkm_distance_batch) and the k best of those are returned
 * in increasing exact distance (ties by row). No reference counterpart: this is the
 * precision layer that stands in for exact exhaustive FP32 scoring.
 * dkm_prescore_create keeps d_x (which must stay valid) and builds the bf16 / fp8 copy;
 * dkm_prescore_search takes device queries / outputs (nq x k) and, if ms_stages is not
 * NULL, reports the device time of four stages (approximate scoring with the candidate
 * filter, candidate selection, exact rescore, final top-k). Synchronous; searches on one
 * handle are serialised. The handle keeps its search work buffers (candidate lists of up
 * to 1,024 queries x max(8 C, ...) keys, sample scores, rescore arrays, sort space: about
 * 0.5 GB at C = 4096)
 * from the first search until dkm_prescore_destroy, so later searches allocate nothing.
 * When 1.5 C >= n every row is rescored. */
typedef struct dkm_prescorer dkm_prescorer;
dkm_status dkm_prescore_create(const float *d_x, int64_t n, int64_t dim, int metric,
                               int precision, dkm_prescorer **out);
dkm_status dkm_prescore_search(dkm_prescorer *ps, const float *d_queries, int64_t nq, int64_t k,
                               int64_t candidates, uint32_t *d_ids, float *d_dists,
                               double *ms_stages);
void dkm_prescore_destroy(dkm_prescorer *ps);

/* Tuning knobs (autotuner surface; none changes results): visited-set slots in LDS per
 * query (a multiple of 64 in 256..32768, 0 = automatic) and the PQ distance mode (0 = lookup table in LDS,
 * 1 = direct recomputation from the pivots, the default). */
dkm_status dkm_set_tuning(dkm_index *ix, int64_t visited_lds_slots, int adc_mode);
/* Capacity of a query's HBM visited-set region for the batched and graph searches created
 * after the call (a power of two in [256, 2^24]; 0 = automatic: at least 16384 and 16 x the
 * launch's max L). A query that visits more ids than fit reports DKM_ERR_CAPACITY. */
dkm_status dkm_set_visited_capacity(dkm_index *ix, int64_t slots);

/* Autotuner: times the search of `nq` device-resident queries (k, L, beam_width fixed)
 * over the PQ distance modes and a few visited-set sizes derived from the measured
 * per-query comparison counts, keeps the fastest setting on the handle and reports it.
 * Results are identical under every setting. */
dkm_status dkm_autotune(dkm_index *ix, int strategy, const float *d_queries, int64_t nq,
                        uint32_t k, uint32_t L, uint32_t beam_width,
                        int64_t *out_visited_lds_slots, int *out_adc_mode);

#ifdef __cplusplus
}
#endif
#endif

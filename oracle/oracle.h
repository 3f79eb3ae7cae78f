/*
 * CPU oracle: a plain-C restatement of the reference (SeliMeli/DiskANN v0.55.0)
 * search-and-quantize path. TEST INFRASTRUCTURE ONLY: tests, the smoke check and the
 * CPU-baseline leg of the benchmark call it as the checker; the product path
 * (diskann_amd/, libdiskann_mi355x.so) never links or calls it.
 *
 * Every function cites the reference file:line it restates.
 *
 * Parity status per function (how each one is pinned):
 *   dko_distance (L2/IP)      pinned: grid_search golden distances, numpy FP64 bound,
 *                             hand-derived 32-partial order on adversarial inputs.
 *   dko_distance (cosine)     pinned: closed forms (parallel / orthogonal / zero-norm)
 *                             and FP64 bound; V3 (f32x8, Strategy2x4) order chosen.
 *   queue + dko_search        pinned: grid_search_{1_100,3_5,4_4} golden ids, distances,
 *                             cmps and hops for beam widths 1/2/4 (reference JSON).
 *   dko_pq_lut / dko_pq_adc   pinned: test_compute_pq_distance KAT and FP64 bound.
 *   dko_pq_encode             pinned: brute-force argmin over pivots (FP64 check).
 *   dko_pq_direct / dko_pq_qq pinned: test_quant_quant_distances KAT (seeded integer
 *                             pivot tables: equal to the distance of the reconstructed
 *                             vectors, L2 / IP / cosine exact), integer-query KAT for the
 *                             full-precision side, reference tolerances on float queries.
 *   dko_chunk_offsets         pinned: partition_happy_path KAT.
 *   dko_occlude_list          pinned: update_occlude_factor KATs + invariants;
 *                             grid_insert golden (post-build searches) where the
 *                             reference's unstable sort ties do not matter.
 *   dko_lloyd_*               pinned: closed form on separated clusters, monotone residual.
 *   dko_distance (cos.-norm.) pinned: bitwise 1 - (pinned inner-product sum), FP64 bound,
 *                             closed forms.
 *   dko_kmeanspp              given the uniform draws: pinned by n = K distinct rows (every
 *                             row picked), a numpy f64 cumulative-sum search for the pick,
 *                             the block-pair order of the f64 total on crafted inputs, and
 *                             the all-zero / infinite / too-small error cases. The draws
 *                             themselves (rand 0.9.4 / ChaCha12 in the reference) are parity
 *                             unpinned: the library uses its own SplitMix64 stream.
 *   dko_pq_train              composition of the pinned pieces (square_norm, k-means++,
 *                             Lloyd); pinned against a numpy restatement of the driver.
 */
#ifndef DKO_ORACLE_H
#define DKO_ORACLE_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Metric numbering follows diskann-vector/src/distance/metric.rs:11-20 (repr(C)). */
enum { DKO_COSINE = 0, DKO_INNER_PRODUCT = 1, DKO_L2 = 2, DKO_COSINE_NORMALIZED = 3 };

/* Search modes: FullPrecision (CopyIds), Hybrid (PQ traversal + FP32 Rerank),
 * Quantized (PQ traversal, PQ distances returned). */
enum { DKO_MODE_FULL = 0, DKO_MODE_HYBRID = 1, DKO_MODE_PQ_ONLY = 2 };

float dko_distance(int metric, const float *q, const float *x, int64_t dim);
float dko_square_norm(const float *x, int64_t len);

/* PQ ---------------------------------------------------------------------------- */
void dko_chunk_offsets(int64_t dim, int64_t nchunks, int64_t *offsets /* nchunks+1 */);
void dko_pq_lut(int metric, const float *pivots /* K x dim */, int64_t K, int64_t dim,
                const int64_t *offsets, int64_t M, const float *query, float *lut /* M*K */);
float dko_pq_adc(const uint8_t *code, const float *lut, int64_t M, int64_t K);
/* Direct PQ distances (fixed_chunk_pq_table.rs:32-56, :290-365): full-precision query vs
 * the pivots a code selects, and code vs code, with the resumable f32x8 accumulation. */
float dko_pq_direct(int metric, const float *pivots, int64_t dim, const int64_t *offsets,
                    int64_t M, const float *query, const uint8_t *code);
float dko_pq_qq(int metric, const float *pivots, int64_t dim, const int64_t *offsets, int64_t M,
                const uint8_t *a, const uint8_t *b);
int dko_pq_encode(const float *pivots, int64_t K, int64_t dim, const int64_t *offsets,
                  int64_t M, const float *x, int64_t n, uint8_t *codes);

/* Scalar quantization, 8 bits (diskann-quantization/src/scalar) ------------------------ */
typedef struct {
    int64_t dim;
    float scale;               /* dynamic range of the codes (train.rs:75-88) */
    const float *shift;        /* [dim] per-dimension shift */
    float shift_square_norm;   /* <shift, shift> (quantizer.rs:116-127) */
    float mean_norm;           /* average row norm (query rescale for inner product) */
    const uint8_t *codes;      /* [rows * dim] codes of the index rows */
    const float *comp;         /* [rows] compensation (quantizer.rs:407-432) */
} dko_sq;

void dko_sq_train(const float *x, int64_t n, int64_t dim, int64_t stride, double stdevs,
                  float *scale, float *shift, float *mean_norm);
float dko_sq_shift_square_norm(const float *shift, int64_t dim);
/* returns 1 when x holds a NaN (the codes are written regardless) */
int dko_sq_compress(float scale, const float *shift, int64_t dim, const float *x,
                    uint8_t *code, float *comp);
void dko_sq_rescale(float *x, int64_t dim, float to_norm);
float dko_sq_distance(int metric, float scale, float shift_square_norm, const uint8_t *a,
                      float comp_a, const uint8_t *b, float comp_b, int64_t dim);

/* Spherical quantization, 1 bit (diskann-quantization/src/spherical) ------------------- */
enum { DKO_SPH_NULL = 0, DKO_SPH_PADDING_HADAMARD = 1, DKO_SPH_DOUBLE_HADAMARD = 2 };
typedef struct {
    int metric;                /* DKO_L2, DKO_INNER_PRODUCT or DKO_COSINE */
    int64_t dim;               /* input dimension */
    int64_t out_dim;           /* transform output dimension (bits per code) */
    const float *shift;        /* [dim] centroid */
    float mean_norm, pre_scale;
    int kind;                  /* DKO_SPH_* */
    const uint32_t *signs0;    /* [dim] f32 sign masks (padding Hadamard: its signs) */
    const uint32_t *signs1;    /* double Hadamard: [max(dim, out_dim)] */
    int64_t padded_dim;        /* padding Hadamard: power of two >= dim */
    const uint32_t *subsample; /* optional strictly increasing indices, nsub = out_dim */
    int64_t nsub;
} dko_sph;

/* bytes of one 1-bit data code: bits (little-endian within bytes), then f16 inner-product
 * correction, f16 metric term, u16 bit sum (DataMut::from_canonical_back_mut) */
/* SphericalQuantizer::train's deterministic part (spherical/quantizer.rs:225-320): the
 * centroid (column means in row order; cosine: means of the unit-normalised rows) and the
 * mean norm (1 for cosine). The transform's random signs are not reproduced. */
void dko_sph_train_stats(int metric, const float *x, int64_t n, int64_t dim, float *centroid,
                         float *mean_norm);
int64_t dko_sph_data_bytes(const dko_sph *q);
void dko_sph_hadamard(float *x, int64_t len);
void dko_sph_transform(const dko_sph *q, const float *src, float *dst);
void dko_sph_rescale(const dko_sph *q, float *x);
/* 0 ok, 1 NaN / infinite input, 2 correction not representable in f16 */
int dko_sph_compress_data(const dko_sph *q, const float *x, uint8_t *out);
/* 4-bit query: out_dim codes (one per byte) + meta {ip correction, bit sum, offset,
 * metric term} */
int dko_sph_compress_query4(const dko_sph *q, const float *x, uint8_t *codes, float *meta);
/* full-precision query: out_dim transformed values + meta {sum, shifted norm, metric term} */
int dko_sph_compress_full(const dko_sph *q, const float *x, float *v, float *meta);
float dko_sph_dist_data(const dko_sph *q, const uint8_t *a, const uint8_t *b);
float dko_sph_dist_query4(const dko_sph *q, const uint8_t *codes, const float *meta,
                          const uint8_t *b);
float dko_sph_dist_full(const dko_sph *q, const float *v, const float *meta, const uint8_t *b);
uint16_t dko_f32_to_f16(float f);
float dko_f16_to_f32(uint16_t h);

/* MinMax quantization, 8 bits (diskann-quantization/src/minmax) ------------------------
 * The transform is a dko_sph's (metric, shift, mean_norm, pre_scale unused). A code row is
 * the canonical front layout: u32 dim, f32 b (minimum), f32 n (a * sum of codes), f32 a
 * (step), f32 |reconstruction|^2, then dim code bytes. */
int64_t dko_minmax_bytes(const dko_sph *t);
/* returns 1 when the transformed vector holds a NaN; *loss = squared reconstruction error */
int dko_minmax_compress(const dko_sph *t, float grid_scale, const float *x, uint8_t *out,
                        float *loss);
/* full-precision query: out_dim transformed values + meta {sum, |v|^2} */
int dko_minmax_compress_full(const dko_sph *t, const float *x, float *v, float *meta);
float dko_minmax_dist_data(int metric, const uint8_t *a, const uint8_t *b);
float dko_minmax_dist_full(int metric, const float *v, const float *meta, const uint8_t *b);

/* Graph + search -------------------------------------------------------------------- */
typedef struct {
    int metric;
    int64_t dim;
    int64_t vstride;           /* floats per vector row */
    const float *vectors;      /* rows indexed by internal id */
    int64_t max_degree;        /* adjacency row stride */
    const uint32_t *adj;       /* [rows * max_degree] */
    const uint32_t *deg;       /* [rows] */
    int64_t nstart;
    const uint32_t *starts;    /* start point internal ids */
    /* PQ (optional) */
    int64_t M, K;
    const uint8_t *codes;      /* [rows * M] */
    const float *pivots;       /* [K * dim] */
    const int64_t *offsets;    /* [M+1] */
    /* scalar quantization (optional): when set, the quantized modes traverse with it */
    const dko_sq *sq;
    /* 1-bit spherical quantization (optional, takes precedence): canonical code rows and
     * the traversal query layout (0 same-as-data, 1 4-bit) */
    const dko_sph *sph;
    const uint8_t *sph_codes;
    int sph_layout;
    /* spherical traversal without the inner-product query rescale (the build strategy,
     * Quantized::build, is_search = false: spherical.rs:515-533) */
    int sph_norescale;
} dko_index;

typedef struct {
    uint32_t cmps, hops, count;
} dko_stats;

/* Knn validation codes (diskann/src/graph/search/knn_search.rs:94-130). */
enum { DKO_OK = 0, DKO_ERR_KZERO = 1, DKO_ERR_LZERO = 2, DKO_ERR_L_LESS_THAN_K = 3,
       DKO_ERR_BEAM_ZERO = 4, DKO_ERR_ARG = 5 };
int dko_knn_validate(int64_t k, int64_t L, int64_t bw);

int dko_search(const dko_index *ix, int mode, const float *query, int64_t k, int64_t L,
               int64_t bw, uint32_t *out_ids, float *out_d, dko_stats *st);

/* Same search but also returns the whole final queue (first min(L+nstart,size)) and
 * the visited record (expanded nodes, in expansion order). */
int dko_search_full(const dko_index *ix, int mode, const float *query, int64_t L,
                    int64_t bw, uint32_t *q_ids, float *q_d, uint32_t *q_len,
                    uint32_t *rec_ids, float *rec_d, uint32_t *rec_len, int64_t rec_cap,
                    dko_stats *st);

/* Threaded batch search: static contiguous partition across nthreads (the reference
 * benchmark harness, diskann/src/utils/async_tools.rs:311-340). Per-query latency in
 * microseconds is written to lat_us when non-NULL. Returns wall seconds. */
double dko_search_batch(const dko_index *ix, int mode, const float *queries, int64_t nq,
                        int64_t qstride, int64_t k, int64_t L, int64_t bw, int nthreads,
                        uint32_t *out_ids, float *out_d, uint32_t *out_cmps,
                        uint32_t *out_hops, double *lat_us);
/* Same with a k and an L per query (uneven requests in one batch); outputs have row
 * stride kstride >= every k. Returns -1 when any (k, L) fails Knn validation. */
double dko_search_batch_mixed(const dko_index *ix, int mode, const float *queries, int64_t nq,
                              int64_t qstride, const uint32_t *ks, const uint32_t *Ls,
                              int64_t kstride, int64_t bw, int nthreads, uint32_t *out_ids,
                              float *out_d, uint32_t *out_cmps, uint32_t *out_hops,
                              double *lat_us);

/* Build (RobustPrune / insert / multi_insert) ---------------------------------------- */
typedef struct {
    int metric;
    int64_t R;                 /* pruned degree */
    int64_t R_max;             /* max degree with slack */
    int64_t L_build;
    float alpha;
    int64_t max_occlusion;     /* 750 */
    int saturate;              /* saturate_after_prune */
    int64_t max_backedges;     /* insert(): number of back edges (default R_max) */
    /* Hybrid prune (HybridPruneAccessor, diskann-providers/src/model/graph/provider/
     * async_/inmem/product.rs:144-221): the first max_fp ids a prune fills are full
     * precision, the others use their PQ codes (distances through the direct PQ
     * functions). max_fp < 0: every id full precision (codes unused). */
    int64_t max_fp;
    const uint8_t *codes;      /* rows x pq_M */
    const float *pivots;       /* K x dim */
    const int64_t *pq_offsets; /* pq_M + 1 */
    int64_t pq_M;
    /* 8-bit scalar-quantized build (the Quantized strategy over an SQStore,
     * diskann-providers/src/model/graph/provider/async_/inmem/scalar.rs:676-720): the
     * insert search traverses with the SQ query computer (the Quantized search accessor,
     * insert_search_accessor = search_accessor, glue.rs:933-940) and every prune distance
     * is the compensated distance between two stored codes (PruneAccessor, scalar.rs:
     * 370-446). NULL: full-precision (or Hybrid) build. Codes of every row, the start
     * points included, must be present. */
    const dko_sq *sq;
    /* 1-bit spherical build (Quantized::build() over a SphericalStore,
     * diskann-providers/src/model/graph/provider/async_/inmem/spherical.rs:293-350, 515-533,
     * 605-650): the insert search traverses with SameAsData queries and no inner-product
     * rescale, every prune distance is the data x data distance of the two rows' canonical
     * codes (sph_codes, rows x dko_sph_data_bytes). NULL: not spherical. */
    const dko_sph *sph;
    const uint8_t *sph_codes;
} dko_build_cfg;

/* compute_closest_centers (partitioned build): out[i*k + t] = the t-th closest centre of
 * row i by (|x|^2 + |c|^2) + (-2 <x, c>); ties by lower index. */
void dko_closest_centers(const float *x, int64_t n, int64_t dim, const float *c, int64_t K,
                         int64_t k, uint32_t *out);
/* TransposedTable::process_into LUT (the disk index's PQ table): lut[m*K + c] is a fold
 * of fma over chunk m's dimensions (L2: (p - q)^2, inner product: -q * p). */
void dko_pq_lut_transposed(int metric, const float *pivots, int64_t K, int64_t dim,
                           const int64_t *offsets, int64_t M, const float *query, float *lut);
/* Disk-index search (DiskAccessor + RerankAndFilter) from the data row ix->starts[0]
 * (nstart must be 1) with at most io_limit expanded nodes. */
int dko_disk_search(const dko_index *ix, const float *query, int64_t k, int64_t L, int64_t bw,
                    uint64_t io_limit, uint32_t *out_ids, float *out_d, dko_stats *st);

/* occlude_list on a pool sorted by distance (diskann/src/graph/index.rs:2570-2782).
 * pool_ids/pool_d: candidates; `self_id` excluded. Distances between candidates are
 * computed with the metric on `vectors`. Returns the number of neighbors written. */
int64_t dko_occlude_list(const dko_build_cfg *cfg, const float *vectors, int64_t vstride,
                         int64_t dim, const uint32_t *pool_ids, const float *pool_d,
                         int64_t pool_len, uint32_t self_id, uint32_t *out);

float dko_update_occlude_factor(int metric, float d_ik, float d_jk, float cur, float alpha);

/* robust_prune_list (diskann/src/graph/index.rs:2402-2450): distances from `location` to
 * every id of `list` other than itself (list order), sorted pool cut to max_occlusion,
 * occlude_list; cfg->max_fp >= 0 selects the Hybrid prune (fill order location, then the
 * list). force_saturate: prune::Options::force_saturate. Returns the number of ids
 * written to out (room for cfg->R). */
int64_t dko_prune_list(const dko_build_cfg *cfg, const float *vectors, int64_t vstride,
                       int64_t dim, uint32_t location, const uint32_t *list, int64_t n,
                       int force_saturate, uint32_t *out);

/* Mutable graph used by the build. */
typedef struct {
    int64_t rows, max_degree;
    uint32_t *adj;
    uint32_t *deg;
} dko_graph;

/* Insert `ids` in batches of `batch` with multi_insert semantics
 * (diskann/src/graph/index.rs:820-1035); batch==1 uses single insert semantics
 * (index.rs:226-341). intra_batch: -1 = All, 0 = None, n = Max(n).
 * Vectors for every id (and the start points) must already be present. */
int dko_build(const dko_build_cfg *cfg, const float *vectors, int64_t vstride, int64_t dim,
              dko_graph *g, const uint32_t *starts, int64_t nstart, const uint32_t *ids,
              int64_t nids, int64_t batch, int64_t intra_batch);

/* Medoid start point (diskann-utils/src/sampling/medoid.rs:15-47). */
int64_t dko_medoid(const float *x, int64_t n, int64_t stride, int64_t dim, float *mean_out);

/* k-means (diskann-quantization/src/algorithms/kmeans/lloyds.rs) -------------------- */
float dko_lloyd_assign(const float *x, int64_t n, int64_t dim, const float *c, int64_t K,
                       uint32_t *nearest);
void dko_lloyd_update(const float *x, int64_t n, int64_t dim, const uint32_t *nearest,
                      int64_t K, float *c);

/* k-means++ seeding (diskann-quantization/src/algorithms/kmeans/plusplus.rs:238-498) over
 * n rows of ds floats (row stride `stride`) into centers (K x ds, zero-filled first), with
 * the caller's uniform draws u[0..K) in [0, 1) standing in for the reference's RNG: the
 * first centre is row floor(u[0] n), step c draws its threshold as u[c] times the scale of
 * rand's UniformFloat over [0, s). Returns 0 (ok) or the failure reason: 1
 * DatasetTooSmall, 2 InsufficientDiversity, 3 SawInfinity; *selected = centres picked. */
int dko_kmeanspp(const float *x, int64_t n, int64_t stride, int64_t ds, int64_t K,
                 const double *u, float *centers, int64_t *selected);
/* LightPQTrainingParameters::train (diskann-quantization/src/product/train.rs:109-231) on
 * n x dim rows with chunk offsets (M + 1): per chunk k-means++ (draws u + m*K), recoverable
 * failures suppressed, then `iters` Lloyd iterations (lloyds_inner, lloyds.rs:372-426).
 * pivots: K x dim. Returns 0, or 1 + the chunk whose initialisation failed
 * unrecoverably. */
int dko_pq_train(const float *x, int64_t n, int64_t dim, const int64_t *offsets, int64_t M,
                 int64_t K, int64_t iters, const double *u, float *pivots);

/* Exact FP64 brute-force top-k (recall oracle). */
void dko_bruteforce_f64(int metric, const float *base, int64_t n, int64_t stride,
                        int64_t dim, const float *queries, int64_t nq, int64_t k,
                        int nthreads, uint32_t *out_ids, double *out_d);

#ifdef __cplusplus
}
#endif
#endif

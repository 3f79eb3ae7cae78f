// Launch interface of the graph-construction kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "dkm_sph.h"

namespace dkm {

enum : uint32_t { BUILD_ERR_RECORD = 1u, BUILD_ERR_POOL = 2u };

struct BuildParams {
    int metric;
    int dim;
    int64_t vstride;
    const float *vectors;
    uint32_t *graph;
    int gstride;
    int R, R_max;
    float alpha;
    int max_occlusion;
    int saturate;
    int sort_cap;          // power of two >= candidate pool before truncation
    int cache_accepted;    // stage accepted vectors in LDS
    int batch;
    const uint32_t *batch_ids;
    int intra_batch;       // -1 = All, 0 = None, n = Max(n)
    const uint32_t *rec_ids;
    const float *rec_d;
    const uint32_t *rec_len;
    uint32_t rec_cap;
    uint32_t *pend;        // batch x R
    uint32_t *pend_cnt;
    uint32_t *status;
    // Hybrid prune (HybridPruneAccessor, diskann-providers/src/model/graph/provider/
    // async_/inmem/product.rs:143-230): the first max_fp ids of a prune's fill order are
    // full precision, every other id is its PQ code. max_fp < 0: all full precision.
    int max_fp;
    const uint8_t *codes;  // rows x pq_M
    const float *pivots;   // K x dim
    const int32_t *pq_off; // pq_M + 1 chunk offsets
    int pq_M;
    // 8-bit scalar-quantized build (the Quantized strategy over an SQStore, diskann-providers/
    // src/model/graph/provider/async_/inmem/scalar.rs:370-446, 676-720): every prune distance
    // is the compensated distance between the two rows' stored codes. 0: full precision.
    int sq;
    const uint8_t *sq_codes;  // rows x sq_cstride, zero padded
    const float *sq_comp;     // rows
    int sq_cstride;           // multiple of 16
    float sq_bss, sq_ssn;     // (1/255)^2 scale^2, <shift, shift>
    // 1-bit spherical build (Quantized::build() over a SphericalStore, diskann-providers/src/
    // model/graph/provider/async_/inmem/spherical.rs:293-350, 605-650): every prune distance
    // is the data x data distance of the two rows' codes. 0: not spherical.
    int sph;
    const uint8_t *sph_bits;  // rows x sph_cstride
    const SphMeta *sph_meta;  // rows
    int sph_cstride;
    int sph_metric;           // the quantizer's metric
    float sph_shift_sq, sph_dim;  // <shift, shift>, output dimension
};

size_t build_prune_lds_bytes(const BuildParams &p);
hipError_t launch_gather_rows(const float *vectors, int64_t vstride, int dim, const uint32_t *ids,
                              int n, float *out, hipStream_t stream);
hipError_t launch_prune_search(const BuildParams &p, hipStream_t stream);
hipError_t launch_prune_lists(const BuildParams &p, const uint32_t *lists, const uint32_t *lens,
                              int stride, int force_saturate, hipStream_t stream);
hipError_t launch_set_lists(const BuildParams &p, hipStream_t stream);
hipError_t launch_backedges(const BuildParams &p, uint64_t *pairs, uint64_t *pairs_tmp,
                            uint32_t *counter, void *sort_tmp, size_t sort_tmp_bytes,
                            uint32_t *out_npairs, uint32_t *out_nheads, hipStream_t stream);
hipError_t launch_apply_backedges(const BuildParams &p, const uint64_t *sorted_pairs,
                                  uint32_t npairs, hipStream_t stream);
size_t radix_sort_tmp_bytes(uint32_t n);

}  // namespace dkm

// Launch interface of the beam-search kernel (host <-> device).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace dkm {

enum : uint32_t { ERR_VISITED_OVERFLOW = 1u, ERR_ITER_LIMIT = 2u };

struct SearchParams {
    int metric;            // dkm_metric
    int pq;                // traverse with PQ ADC distances
    int rerank;            // Hybrid: FP32 rerank of the final queue
    int adc_mode;          // PQ: 0 = lookup table in LDS, 1 = direct (no table)
    int pq_len;            // PQ: uniform chunk width with M % 16 == 0, else 0
    int dim;
    int64_t vstride;       // floats per vector row
    const float *vectors;  // rows x vstride
    const uint32_t *graph; // rows x gstride, count in the last slot
    int gstride;
    const uint32_t *starts;
    int nstart;
    const uint8_t *codes;  // rows x code_stride
    int code_stride;
    int M;
    const float *pivots;   // 256 x dim
    const int32_t *offsets;
    const float *queries;
    int64_t qstride;
    int nq;
    const uint32_t *k;     // per query
    const uint32_t *L;     // per query
    uint32_t beam_width;
    int64_t k_stride;
    uint32_t *out_ids;
    float *out_d;
    uint32_t *out_count;
    uint32_t *out_cmps;    // optional
    uint32_t *out_hops;    // optional
    uint32_t *out_degsum;  // optional: sum of the stored counts of expanded lists
    uint32_t *status;      // per query error bits
    int qcap;              // max L + nstart in the launch
    int nlcap;             // expansion list capacity
    uint32_t hslots;       // LDS visited slots (power of two)
    uint64_t *spill_table; // nq x spill_slots (one region per query of the launch)
    uint64_t spill_slots;  // power of two
    uint32_t epoch;
    uint32_t max_iters;
    uint32_t *rec_ids;     // optional visited record (expanded nodes), rec_cap per query
    float *rec_d;
    uint32_t rec_cap;
    uint32_t *rec_len;
};

size_t search_lds_bytes(const SearchParams &p);
hipError_t launch_search(const SearchParams &p, hipStream_t stream);

}  // namespace dkm

// Launch interface of the beam-search kernel (host <-> device).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "dkm_sph.h"

namespace dkm {

// Per-query status bits. ERR_BAD_REQUEST: the query's k / L break Knn::new's rules
// (k, L >= 1, k <= L; knn_search.rs:101-106) or the launch's capacity (L <= max_L,
// k <= k_stride); the query is not searched and returns no results.
enum : uint32_t { ERR_VISITED_OVERFLOW = 1u, ERR_ITER_LIMIT = 2u, ERR_BAD_REQUEST = 4u };

// L values 0..4096 counted by the longest-L-first ordering of a launch.
constexpr int kOrderBins = 4097;

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
    const float *pivots_t; // uniform chunks of pq_len dims: chunk-major copy, entry (m, c)
                           // at (m * 256 + c) * pq_len
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
    uint32_t *order;       // optional: nq entries + (kOrderBins + 1) bins; the launch then
                           // runs the queries longest L first (block b takes query order[b])
    uint32_t *done;        // optional: per query flag set to 1 once its outputs are written,
                           // after a system-scope fence (outputs in host memory)
    int qcap;              // max L + nstart in the launch
    int nlcap;             // expansion list capacity
    uint32_t hslots;       // LDS visited slots (power of two)
    uint32_t *spill_table; // nq x spill_slots (one region per query of the launch); a
                           // query clears its region when it starts using it
    uint64_t spill_slots;  // power of two
    uint32_t max_iters;
    uint32_t *rec_ids;     // optional visited record (expanded nodes), rec_cap per query
    float *rec_d;
    uint32_t rec_cap;
    uint32_t *rec_len;
    // 8-bit scalar quantization (traverse with it when sq is set; pq is then 0)
    int sq;
    const uint8_t *sq_codes;  // rows x sq_cstride, zero padded
    const float *sq_comp;     // rows
    const float *sq_shift;    // dim
    int sq_cstride;           // multiple of 16
    float sq_scale, sq_bss, sq_ssn, sq_mean_norm;
    // 1-bit spherical quantization (traverse with it when sph is set; pq and sq are 0)
    int sph;
    int sph_layout;           // SPH_QUERY_SAME_AS_DATA or SPH_QUERY_FOUR_BIT
    int sph_rescale;          // inner-product queries rescaled to the mean norm (the search
                              // strategies; the build strategy does not)
    SphParams sphq;
    const uint8_t *sph_bits;  // rows x sphq.cstride
    const SphMeta *sph_meta;  // rows
    uint8_t *sph_qrec;        // nq prepared queries (sph_query_record_bytes each), written by
                              // launch_search before the search kernel
    // Disk-index search (the DiskAccessor of diskann-disk) when tlut is set: PQ distances
    // from the transposed table's arithmetic, the start point is a data row kept in the
    // results (keep_starts), and at most io_limit beam nodes per query are expanded
    // (loaded); io_limit is not read otherwise.
    int tlut;
    int keep_starts;
    uint32_t io_limit;
};

size_t search_lds_bytes(const SearchParams &p);
hipError_t launch_search(const SearchParams &p, hipStream_t stream);
// Search blocks (one query each) of this configuration resident on the current device.
hipError_t search_resident_blocks(const SearchParams &p, int *blocks);

}  // namespace dkm

namespace dkm {

// Serving: request ring and control block in pinned host memory (fine-grained), claim
// counter in device memory. A request is published by storing tail = seq + 1 (release);
// a wave claims seq by CAS on `next`, searches, writes its results into the ring entry
// and stores done_seq = seq + 1 (release).
}  // namespace dkm

namespace dkm {

// Serving: request ring and control block in pinned host memory (fine-grained), claim
// counter in device memory. A request is published by storing tail = seq + 1 (release);
// a wave claims seq by CAS on `next`, searches, writes its results into the ring entry
// and stores done_seq = seq + 1 (release).
struct ServeCtrl {
    uint64_t tail;     // published requests (host writes)
    uint32_t stop;     // host asks the waves to exit
    uint32_t pad;
};

// Device-side mirror of the control block: idle waves poll this instead of host memory;
// at most one wave per poll interval (the one that moves last_poll) reads host memory.
struct ServeDev {
    uint64_t next;       // claim counter
    uint64_t tail;       // largest tail read from host memory
    uint64_t last_poll;  // wall clock (100 MHz) of the latest host read
    uint32_t stop;
    uint32_t pad;
};

struct ServeReq {
    uint32_t k, L;
    uint64_t tag;
};

struct ServeParams {
    const ServeCtrl *ctrl;      // pinned
    ServeDev *dev;              // device
    const ServeReq *req;        // pinned, capacity entries
    const float *queries;       // pinned, capacity x dim
    uint32_t *res_ids;          // pinned, capacity x max_k
    float *res_d;
    uint32_t *res_meta;         // pinned, capacity x 4: count, cmps, hops, status
    uint64_t *claim_seq;        // pinned, capacity: seq + 1 once claimed
    uint64_t *done_seq;         // pinned, capacity: seq + 1 once the result is written
    uint32_t *stage_ids;        // device, 2 x grid x max_k: search output before copy-out
    float *stage_d;
    uint32_t *stage_meta;       // device, 2 x grid x 4
    uint32_t *slot_status;      // device, 2 x grid
    uint64_t capacity;
    uint32_t max_k, max_L;
    uint64_t run_ticks;         // a wave takes no new request after running this long
    uint64_t poll_ticks;        // minimum interval between reads of the host's tail
    uint32_t part;              // 0/1: which half of the per-wave buffers this launch uses
};

hipError_t launch_serve(const SearchParams &p, const ServeParams &sp, int blocks,
                        hipStream_t stream);
// Serving blocks of this configuration that are resident on the current device at once.
hipError_t serve_resident_blocks(const SearchParams &p, int *blocks);

}  // namespace dkm

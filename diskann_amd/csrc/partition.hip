// Closest centres of the partitioned (disk-index) build on gfx950.
//
// Restates compute_closest_centers / compute_closest_centers_in_block
// (diskann-disk/src/utils/math_util.rs:133-238, 246-385): squared norms by
// FastL2NormSquared (diskann-vector/src/norm.rs:45-56, the inner product's 4x1 simd order),
// the three sgemm steps per (point, centre) d = (|x|^2 + |c|^2) + (-2 <x, c>), then the k
// smallest d per point (k = 1: the first centre with the strictly smallest d). The dot
// product inside the reference's sgemm (faer) has its own blocking; here it is an fma chain
// in dimension order, the same order as the CPU restatement.
//
// One wave takes 64 points (a lane each); for each block of 32 centres the points' rows and
// the centres go through LDS in chunks of 128 dimensions (rows at stride 129, so
// lane-strided reads hit distinct banks; centres transposed, dimension-major), and every
// dimension step is one lane-private read plus eight broadcast 16-byte reads feeding 32
// fmas. The running k smallest (distance, index) pairs stay in registers. VALU bound:
// 2 n K dim flops.
#include "dkm_device.h"
#include "dkm_partition.h"

namespace dkm {

namespace {

constexpr int kCB = 32;       // centres per LDS block
constexpr int kRows = 64;     // points per block (one wave)
constexpr int kDC = 128;      // dimensions per LDS chunk

__global__ __launch_bounds__(kWave) void closest_centers_kernel(const float *x, int64_t n, int dim,
                                                                 const float *c, int K, int k,
                                                                 const float *cnorm,
                                                                 uint32_t *out) {
    extern __shared__ __align__(16) float sm[];
    const int lane = threadIdx.x;
    constexpr int xs = kDC + 1;
    float *xt = sm;                                   // kRows x (kDC + 1) chunk of the rows
    float *ct = sm + ((kRows * xs + 3) & ~3);         // kDC x kCB (transposed centre chunk)
    const int64_t r0 = (int64_t)blockIdx.x * kRows;
    const int rows = (int)(n - r0 < kRows ? n - r0 : kRows);
    const bool ok = lane < rows;
    // |x|^2 in the inner product's order, one lane per point from its own row
    const float *xg = x + (size_t)(r0 + (ok ? lane : 0)) * dim;
    const float xn = ok ? scalar_simd_4x1<true>(xg, xg, dim) : 0.0f;
    float topd[kMaxClosest];
    uint32_t topi[kMaxClosest];
#pragma unroll
    for (int j = 0; j < kMaxClosest; ++j) {
        topd[j] = __builtin_inff();
        topi[j] = 0u;
    }
    int have = 0;
    for (int cb = 0; cb < K; cb += kCB) {
        const int nc = K - cb < kCB ? K - cb : kCB;
        float acc[kCB];
#pragma unroll
        for (int j = 0; j < kCB; ++j) acc[j] = 0.0f;
        for (int d0 = 0; d0 < dim; d0 += kDC) {
            const int dc = dim - d0 < kDC ? dim - d0 : kDC;
            __syncthreads();
            for (int e = lane; e < rows * dc; e += kWave) {
                const int r = e / dc, d = e - r * dc;
                xt[r * xs + d] = x[(size_t)(r0 + r) * dim + d0 + d];
            }
            for (int e = lane; e < kCB * dc; e += kWave) {
                const int j = e / dc, d = e - j * dc;
                ct[d * kCB + j] = j < nc ? c[(size_t)(cb + j) * dim + d0 + d] : 0.0f;
            }
            __syncthreads();
            const float *xr = xt + lane * xs;
            for (int d = 0; d < dc; ++d) {
                const float xv = xr[d];
                const float4 *cv = (const float4 *)(ct + d * kCB);
#pragma unroll
                for (int q = 0; q < kCB / 4; ++q) {
                    const float4 c4 = cv[q];
                    acc[4 * q + 0] = __builtin_fmaf(xv, c4.x, acc[4 * q + 0]);
                    acc[4 * q + 1] = __builtin_fmaf(xv, c4.y, acc[4 * q + 1]);
                    acc[4 * q + 2] = __builtin_fmaf(xv, c4.z, acc[4 * q + 2]);
                    acc[4 * q + 3] = __builtin_fmaf(xv, c4.w, acc[4 * q + 3]);
                }
            }
        }
        // insert this block's distances into the running k smallest, centres in index
        // order (a distance equal to a kept one goes after it)
#pragma unroll
        for (int j = 0; j < kCB; ++j) {
            if (j >= nc) break;
            const float dist = (xn + cnorm[cb + j]) + (-2.0f * acc[j]);
            const uint32_t ci = (uint32_t)(cb + j);
            if (have < k || dist < topd[k - 1]) {
                // shift larger entries right (registers: unrolled compare-exchange)
                float dv = dist;
                uint32_t iv = ci;
                bool shifting = false;
#pragma unroll
                for (int t = 0; t < kMaxClosest; ++t) {
                    if (t >= k) break;
                    const bool take = shifting || t >= have || dv < topd[t];
                    const float od = topd[t];
                    const uint32_t oi = topi[t];
                    if (take) {
                        topd[t] = dv;
                        topi[t] = iv;
                        dv = od;
                        iv = oi;
                        shifting = true;
                    }
                }
                have = have < k ? have + 1 : k;
            }
        }
    }
    if (ok) {
        for (int t = 0; t < k; ++t) out[(size_t)(r0 + lane) * k + t] = topi[t];
    }
}

__global__ void center_norms_kernel(const float *c, int K, int dim, float *out) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < K) out[j] = scalar_simd_4x1<true>(c + (size_t)j * dim, c + (size_t)j * dim, dim);
}

}  // namespace

size_t closest_centers_lds() {
    return sizeof(float) * ((size_t)((kRows * (kDC + 1) + 3) & ~3) + (size_t)kDC * kCB);
}

hipError_t launch_closest_centers(const float *x, int64_t n, int dim, const float *c, int K,
                                  int k, float *cnorm, uint32_t *out, hipStream_t st) {
    hipLaunchKernelGGL(center_norms_kernel, dim3((K + 255) / 256), dim3(256), 0, st, c, K, dim,
                       cnorm);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    const size_t lds = closest_centers_lds();
    e = hipFuncSetAttribute((const void *)closest_centers_kernel,
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    const int64_t blocks = (n + kRows - 1) / kRows;
    hipLaunchKernelGGL(closest_centers_kernel, dim3((unsigned)blocks), dim3(kWave), lds, st, x, n,
                       dim, c, K, k, cnorm, out);
    return hipGetLastError();
}

}  // namespace dkm

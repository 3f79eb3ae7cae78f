// Register-resident BF16 MFMA GEMM pieces shared by the k-means assignment and the
// candidate pre-scoring kernels (gfx950, v_mfma_f32_32x32x16_bf16).
//
// Operands are stored in the "fragment layout": 16-byte chunk ((t * KS + s) * 64 + l) holds
// row 32 t + l % 32, dims 16 s + 8 (l / 32) .. + 7, so each 1 KB block is one k-step's
// operand fragment of 32 rows and every operand load is contiguous. A workgroup of four
// waves owns 256 "rows"; wave w keeps the B-operand fragments of two 32-row tiles for the
// whole padded dimension (KS k-steps) in registers, so rows are read from memory once. The
// "columns" stream through LDS in tiles of 32 (KS blocks of 1 KB), loaded by global_load_lds
// and double-buffered. v_mfma_f32_32x32x16_bf16(column fragment, row fragment) leaves
// accumulator r of lane l = <row l % 32, column 8 (r / 4) + 4 (l / 32) + r % 4 of the tile>.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

namespace dkm {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8_t;
typedef __attribute__((__vector_size__(16 * sizeof(float)))) float floatx16_t;
typedef __attribute__((address_space(3))) void lds_void_t;
typedef __attribute__((address_space(1))) void global_void_t;

// f32 -> bf16, round to nearest even (NaN kept quiet).
__device__ inline uint16_t to_bf16(float f) {
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

// Wave w loads blocks w, w + 4, ... of a tile; `src` is the tile's first block + 8 lane. The
// loads are kept together, ahead of the tile's LDS reads (each sets M0, and the compiler
// drains outstanding LDS reads before an M0 write).
template <int KS>
__device__ __forceinline__ void res_prefetch(const uint16_t *src, uint16_t *buf, int wave) {
    // the wave index is uniform: as a scalar the LDS addresses (M0) need no readfirstlane
    const int w = __builtin_amdgcn_readfirstlane(wave);
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int s = 0; s < KS / 4; ++s)
        __builtin_amdgcn_global_load_lds((global_void_t *)(src + (4 * s + w) * 512),
                                         (lds_void_t *)(buf + (4 * s + w) * 512), 16, 0, 0);
    __builtin_amdgcn_sched_barrier(0);
}

// The column fragments are read from LDS in batches of B k-steps: batch j + 1 is issued right
// after the first MFMA pair of batch j, so the one wait before each batch finds its reads
// (issued a batch earlier) complete. The wait is a full LDS drain: the compiler counts the
// in-flight global_load_lds of the next tile as outstanding LDS work and cannot wait for a
// partial count. Scheduling group barriers pin the order.
template <int KS>
__device__ __forceinline__ void res_tile(const bf16x8_t *a0, const bf16x8_t *a1,
                                         const uint16_t *buf, floatx16_t &acc0, floatx16_t &acc1) {
    constexpr int B = KS % 8 == 0 ? 8 : 4;
    static_assert(KS % B == 0, "k-steps per tile");
    bf16x8_t b[2 * B];
#pragma unroll
    for (int i = 0; i < B; ++i) b[i] = *(const bf16x8_t *)(buf + i * 512);
    __builtin_amdgcn_sched_group_barrier(0x100, B, 0);
#pragma unroll
    for (int j = 0; j < KS / B; ++j) {
        bf16x8_t *cur = b + (j & 1) * B, *nxt = b + ((j + 1) & 1) * B;
#pragma unroll
        for (int i = 0; i < B; ++i) {
            const int s = j * B + i;
            acc0 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(cur[i], a0[s], acc0, 0, 0, 0);
            acc1 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(cur[i], a1[s], acc1, 0, 0, 0);
            if (i == 0 && j + 1 < KS / B) {
#pragma unroll
                for (int q = 0; q < B; ++q) nxt[q] = *(const bf16x8_t *)(buf + (s + B + q) * 512);
            }
        }
        __builtin_amdgcn_sched_group_barrier(0x008, 2, 0);
        if (j + 1 < KS / B) __builtin_amdgcn_sched_group_barrier(0x100, B, 0);
        __builtin_amdgcn_sched_group_barrier(0x008, 2 * B - 2, 0);
    }
}

// Loads the wave's two row tiles (ap = first tile's block 0 + 8 lane). Where the fragments
// live: the first tile's and eight of the second's in AGPRs (MFMA reads its B operand from
// there directly), the rest in VGPRs beside the column-fragment batches. The empty asm
// statements only constrain the register class.
template <int KS>
__device__ __forceinline__ void res_load_rows(const uint16_t *ap, bf16x8_t *a0, bf16x8_t *a1) {
#pragma unroll
    for (int s = 0; s < KS; ++s) a0[s] = *(const bf16x8_t *)(ap + s * 512);
#pragma unroll
    for (int s = 0; s < KS; ++s) a1[s] = *(const bf16x8_t *)(ap + (KS + s) * 512);
#pragma unroll
    for (int s = 0; s < KS; ++s) asm volatile("" : "+a"(a0[s]));
#pragma unroll
    for (int s = 0; s < KS; ++s) {
        if (s < 8) asm volatile("" : "+a"(a1[s]));
        else asm volatile("" : "+v"(a1[s]));
    }
}

// Sweeps `ntiles` column tiles (cf = column fragments, tile t at cf + t * KS * 512) through
// the double buffer at sm[0 .. 2 * KS * 512): init(t, acc0, acc1) before a tile's MFMAs,
// epi(t, acc0, acc1) after (acc0: first row tile, acc1: second). The first barrier also
// covers whatever the caller wrote to LDS before the call. ISSUED: the caller already
// issued tile 0's DMA (res_prefetch<KS>(cf + 8 lane, sm, wave)), e.g. with its row loads.
template <int KS, bool ISSUED = false, class Init, class Epi>
__device__ __forceinline__ void res_sweep(uint16_t *sm, const uint16_t *cf, int ntiles, int wave,
                                          int lane, const bf16x8_t *a0, const bf16x8_t *a1,
                                          Init init, Epi epi) {
    auto src = [&](int t) { return cf + (int64_t)t * KS * 512 + lane * 8; };
    floatx16_t pa, pb;
    if constexpr (!ISSUED) res_prefetch<KS>(src(0), sm, wave);
    for (int t = 0; t < ntiles; ++t) {
        __builtin_amdgcn_s_waitcnt(0);  // own loads of tile t landed
        __syncthreads();                // everyone's, and tile t - 1 fully read
        uint16_t *cur = sm + (t & 1) * KS * 512;
        init(t, pa, pb);
        if (t + 1 < ntiles) res_prefetch<KS>(src(t + 1), sm + ((t + 1) & 1) * KS * 512, wave);
        res_tile<KS>(a0, a1, cur + lane * 8, pa, pb);
        epi(t, pa, pb);
    }
}

// ---------------------------------------------------------------------------------------
// FP8 (E4M3, OCP) operands for v_mfma_scale_f32_32x32x64_f8f6f4 with unit scales.
//
// FP8 fragment layout: 16-byte chunk ((t * KB + b) * 64 + l) holds row 32 t + l % 32, dims
// 32 b + 16 (l / 32) .. + 15, one byte each; KB = padded dimension / 32, so a 1 KB block
// again holds 32 rows, and the same prefetch (res_prefetch<KB>) streams a column tile. One
// MFMA k-step covers 64 dims: lane l's 32-byte operand is its chunk of block 2 s followed
// by its chunk of block 2 s + 1. Row and column operands use this same assignment of dims
// to (lane half, byte), so every k-step contracts the same 64 dims of both, whatever order
// the matrix core sums them in. The accumulator layout is the one of the bf16 MFMA above.
// ---------------------------------------------------------------------------------------
typedef __attribute__((__vector_size__(8 * sizeof(int)))) int i32x8_t;
typedef __attribute__((__vector_size__(4 * sizeof(int)))) int i32x4_t;

__device__ __forceinline__ floatx16_t mfma_f8(const i32x8_t &col, const i32x8_t &row,
                                              const floatx16_t &acc) {
    // formats 0 / 0: E4M3 for both operands; E8M0 scales 127 = 2^0
    return __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(col, row, acc, 0, 0, 0, 127, 0, 127);
}

__device__ __forceinline__ i32x8_t f8_join(const i32x4_t &lo, const i32x4_t &hi) {
    return __builtin_shufflevector(lo, hi, 0, 1, 2, 3, 4, 5, 6, 7);
}

// Column fragments of one tile from LDS (buf = tile + 8 lane, in 16-bit units), all KB / 2
// k-steps against the two register-resident row tiles.
template <int KB>
__device__ __forceinline__ void f8_tile(const i32x8_t *a0, const i32x8_t *a1,
                                        const uint16_t *buf, floatx16_t &acc0,
                                        floatx16_t &acc1) {
    static_assert(KB % 2 == 0, "k-steps of 64 dims");
#pragma unroll
    for (int s = 0; s < KB / 2; ++s) {
        const i32x8_t b = f8_join(*(const i32x4_t *)(buf + (2 * s) * 512),
                                  *(const i32x4_t *)(buf + (2 * s + 1) * 512));
        acc0 = mfma_f8(b, a0[s], acc0);
        acc1 = mfma_f8(b, a1[s], acc1);
        // keep the two row tiles' chains interleaved (each column fragment feeds both)
        __builtin_amdgcn_sched_group_barrier(0x008, 2, 0);
    }
}

// f(std::integral_constant<int, i>) for i = I .. N - 1, so f can index registers statically.
template <int I, int N, class F>
__device__ __forceinline__ void static_for(F &&f) {
    if constexpr (I < N) {
        f(std::integral_constant<int, I>{});
        static_for<I + 1, N>(f);
    }
}

// f8_tile with independent work threaded between the k-steps: work(integral_constant<s>)
// follows k-step s's two MFMAs in program order, so its VALU and LDS instructions issue
// while the matrix core runs (each 32x32x64 MFMA keeps it busy for 64 cycles).
template <int KB, class Work>
__device__ __forceinline__ void f8_tile_with(const i32x8_t *a0, const i32x8_t *a1,
                                             const uint16_t *buf, floatx16_t &acc0,
                                             floatx16_t &acc1, Work &&work) {
    static_assert(KB % 2 == 0, "k-steps of 64 dims");
    static_for<0, KB / 2>([&](auto sc) {
        constexpr int s = decltype(sc)::value;
        const i32x8_t b = f8_join(*(const i32x4_t *)(buf + (2 * s) * 512),
                                  *(const i32x4_t *)(buf + (2 * s + 1) * 512));
        acc0 = mfma_f8(b, a0[s], acc0);
        acc1 = mfma_f8(b, a1[s], acc1);
        work(sc);
    });
}

// Operand policies for the pre-scoring GEMMs: kBlocks 1 KB blocks per 32-row tile, the
// wave's two row tiles held in registers, and the sweep over the column tiles.
template <int KS>
struct Bf16Rows {
    static constexpr int kBlocks = KS;
    // the 768-dim row fragments fill the register file: no room for the mask epilogue
    static constexpr bool kMaskEpilogue = KS < 48;
    static constexpr bool kPipelined = false;
    bf16x8_t a0[KS], a1[KS];
    __device__ __forceinline__ void load(const uint16_t *ap) { res_load_rows<KS>(ap, a0, a1); }
    template <bool ISSUED = false, class Init, class Epi>
    __device__ __forceinline__ void sweep(uint16_t *sm, const uint16_t *cf, int ntiles, int wave,
                                          int lane, Init init, Epi epi) {
        res_sweep<KS, ISSUED>(sm, cf, ntiles, wave, lane, a0, a1, init, epi);
    }
};

template <int KB>
struct Fp8Rows {
    static constexpr int kBlocks = KB;
    static constexpr bool kMaskEpilogue = true;
    // the filter threads a tile's threshold work between the next tile's MFMAs
    static constexpr bool kPipelined = true;
    static_assert(KB % 4 == 0, "a column tile is prefetched four blocks per wave");
    i32x8_t a0[KB / 2], a1[KB / 2];
    __device__ __forceinline__ void load(const uint16_t *ap) {
#pragma unroll
        for (int s = 0; s < KB / 2; ++s) {
            a0[s] = f8_join(*(const i32x4_t *)(ap + (2 * s) * 512),
                            *(const i32x4_t *)(ap + (2 * s + 1) * 512));
            a1[s] = f8_join(*(const i32x4_t *)(ap + (KB + 2 * s) * 512),
                            *(const i32x4_t *)(ap + (KB + 2 * s + 1) * 512));
        }
#pragma unroll
        for (int s = 0; s < KB / 2; ++s) asm volatile("" : "+a"(a0[s]));
    }
    template <bool ISSUED = false, class Init, class Epi>
    __device__ __forceinline__ void sweep(uint16_t *sm, const uint16_t *cf, int ntiles, int wave,
                                          int lane, Init init, Epi epi) {
        auto src = [&](int t) { return cf + (int64_t)t * KB * 512 + lane * 8; };
        floatx16_t pa, pb;
        if constexpr (!ISSUED) res_prefetch<KB>(src(0), sm, wave);
        for (int t = 0; t < ntiles; ++t) {
            __builtin_amdgcn_s_waitcnt(0);
            __syncthreads();
            uint16_t *cur = sm + (t & 1) * KB * 512;
            init(t, pa, pb);
            if (t + 1 < ntiles) res_prefetch<KB>(src(t + 1), sm + ((t + 1) & 1) * KB * 512, wave);
            f8_tile<KB>(a0, a1, cur + lane * 8, pa, pb);
            // both row tiles' results are complete here: without this the second chain is
            // sunk past the first tile's epilogue (its loop), serialising the two
            asm volatile("" : "+a"(pa), "+a"(pb));
            epi(t, pa, pb);
        }
    }
};

// f32 -> FP8 E4M3 (OCP: bias 7, no infinities, 448 largest finite), round to nearest even;
// magnitudes that round above 448 saturate to 448, NaN stays NaN.
__device__ inline uint8_t to_e4m3(float f) {
    const uint32_t u = __float_as_uint(f);
    const uint8_t sign = (uint8_t)((u >> 24) & 0x80u);
    const float a = fabsf(f);
    if (a != a) return (uint8_t)(sign | 0x7Fu);
    if (a >= 464.0f) return (uint8_t)(sign | 0x7Eu);
    if (a < 0.015625f) {  // below 2^-6: subnormal steps of 2^-9 (8 rounds up to 2^-6)
        return (uint8_t)(sign | (uint32_t)__builtin_rintf(a * 512.0f));
    }
    const uint32_t au = __float_as_uint(a);
    int e = (int)(au >> 23) - 127;
    uint32_t keep = (au >> 20) & 7u;
    const uint32_t rest = au & 0xFFFFFu;
    if (rest > 0x80000u || (rest == 0x80000u && (keep & 1u))) ++keep;
    if (keep == 8u) {
        keep = 0;
        ++e;
    }
    return (uint8_t)(sign | (uint32_t)((e + 7) << 3) | keep);
}

}  // namespace dkm

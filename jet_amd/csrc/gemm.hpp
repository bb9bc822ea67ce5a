// gemm.hpp -- complex GEMM kernels of the contraction path (gfx950).
//
// Every contraction reduces to the batched "NT" product
//     C[b][m][n] = sum_k A[b][m][k] * B[b][n][k]
// with A and B stored k-fastest (the permutation stage produces them in that
// order).  This is the row-major NoTrans/NoTrans product of the reference
// (TensorHelpers.hpp:47-61) with the right operand kept as [right][common]
// so both MFMA operands are read along k.
#pragma once

#include "errors.hpp"

#include <cstdint>

namespace jet_hip {

struct GemmArgs {
    const void *A;
    const void *B;
    void *C;
    int64_t M, N, K, batch;
    int64_t lda, ldb, ldc; // row strides in elements
    int64_t sA, sB, sC;    // batch strides in elements (0 = broadcast)
    int variant = 0;       // tile / thread-mapping variant (GemmChoice)
};

// Kernel families (also reported by jet_hip_plan_info).
enum GemmKernel {
    kGemmSmall = 0,     // VALU, one thread per output element
    kGemmSmallRed = 1,  // VALU, one block per output element (long K)
    kGemmF32Mfma = 2,   // 4M complex on v_mfma_f32_16x16x4_f32
    kGemmF64Mfma = 3,   // 4M complex on v_mfma_f64_16x16x4_f64
    kGemmBf16x3 = 4,    // 3M complex, split-BF16, v_mfma_f32_16x16x32_bf16
    kGemmBf16x4m = 5,   // 4M complex, split-BF16, v_mfma_f32_16x16x32_bf16
    kGemmSplitK = 6,    // VALU, few outputs and long k: k split over blocks,
                        // partial sums combined in FP64
    kGemmF16x4m = 7,    // 4M complex, split-FP16 with per-row power-of-two
                        // scales, v_mfma_f32_16x16x32_f16
};

// A GEMM kernel family and its tile (or thread-mapping) variant.  Variant 0
// lets the launcher pick by shape; the others force one form:
//   kGemmF32Mfma              1: 128x128  2: 64x128  3: 128x64 block tiles
//   kGemmF64Mfma              1: 128x64 block tiles
//   kGemmBf16x4m, kGemmF16x4m 1: 128x128  2: 128x192 block tiles
//                             3: 128x128 persistent (one work-group per
//                             CU walking the tiles)
//   kGemmSmall                1: one thread per output  2: short-K column
//                             blocks (K <= 16)  3: short-K row threads (K <= 16)
//   kGemmSmallRed, kGemmSplitK, kGemmBf16x3   1: their one form
struct GemmChoice {
    int kernel = kGemmSmall;
    int variant = 0;
};

// Variants 1..gemm_variants(kernel) exist for each kernel family.
int gemm_variants(int kernel);
// Whether the choice computes this shape in precision mode `mode` (0 fp32,
// 1 split-bf16 3M, 2 fp64, 3 split-bf16 4M, 4 split-fp16 4M): the data type
// matches, the kernel's grid fits and its operand limits hold.
bool gemm_choice_valid(int mode, int64_t M, int64_t N, int64_t K, int64_t batch,
                       GemmChoice c);

// Split-K: k range per block, and the FP64 partial-sum buffer it needs.
constexpr int64_t kSplitKChunk = 8192;
inline int64_t splitk_chunk(int64_t K)
{
    const int64_t c = ceil_div(K, 65535);
    return c > kSplitKChunk ? c : kSplitKChunk;
}
inline int64_t splitk_partials(int64_t outputs, int64_t K)
{
    return outputs * ceil_div(K, splitk_chunk(K));
}

// Rows and k of the split-BF16 operand planes are padded to these multiples.
constexpr int kSplitRowPad = 16;
constexpr int kSplitKPad = 32;
// Longest padded k of the 4M split kernels: their per-lane load offsets
// within a tile of up to 192 operand rows stay below 2^32 bytes.
constexpr int64_t kX4mMaxK = int64_t(1) << 21;

// A split-FP16 operand read straight from an unpermuted source: the
// permutation to [batch][rows][K] is a permutation of the linear index bits
// (every extent a power of two), and one pass gathers LDS tiles made of all
// K bits and `xr` row bits -- the row bits that hold the source's lowest
// address bits, so reads are contiguous runs -- takes each tile row's
// maximum and writes the four planes and the scale (no permuted copy).
//   u: tile index in source order (bit j -> source bit su[j], ascending);
//   v: tile index in target order (bits 0..log2k-1 = k, then the tile's row
//      bits; u-bit j is v-bit u2v[j]; tile row bit i is row bit vrow[i]);
//   w: tile number (bit j -> source bit wsrc[j] and row bit wrow[j]).
constexpr int kGatherTileBits = 11; // 2048-element tiles (4096: 40% slower)
struct SplitGather {
    int tb = 0, log2k = 0, xr = 0, nw = 0;
    int8_t su[16] = {}, u2v[16] = {}, vrow[16] = {};
    int8_t wsrc[64] = {}, wrow[64] = {};
};

// A small contraction of power-of-two extents in one VALU pass that reads A
// and B and writes C in their own mode orders (no permuted copies, one
// launch instead of up to four).  Each group of modes (batch, left, right,
// common) is linearised with its last mode fastest; bit i of a group index
// sits at bit pos[i] of the element offset in the operand (row-major, every
// extent a power of two, so the modes' digits occupy disjoint bit ranges).
constexpr int kFusedMaxBits = 24;  // bits per group
constexpr int kFusedMaxLog2K = 12; // k offset tables in LDS (2 x 4 KiB entries)
struct FusedBits {
    int n = 0;
    int8_t pos[kFusedMaxBits] = {};
};
struct FusedSmall {
    int lb = 0, lm = 0, ln = 0, lk = 0; // log2 of batch, M, N, K
    int red = 0; // 1: one work-group per output, k split over its threads
    FusedBits a_b, a_m, a_k, b_b, b_n, b_k, c_b, c_m, c_n;
};

// Split-BF16 operand: planes re hi/lo, im hi/lo[, re+im hi/lo].
//  * planar (3M kernel, six planes): plane p is [batch][rows_p][k_p] bf16 at
//    offset p * plane_elems;
//  * interleaved (4M kernel, four planes): [batch][rows_p][k_p / 32][4][32]
//    bf16, i.e. the four planes of one row and one 32-wide k tile form 256
//    contiguous bytes (two full cache lines per row and k tile).
//  * split-FP16 (4M kernel, four interleaved fp16 planes): each row r of
//    batch b is scaled by 2^(14 - e) before the split, e the exponent of the
//    row's largest |re| or |im|; rowmax[b * rows_p + r] holds that largest
//    magnitude's FP32 bit pattern and scale[b * rows_p + r] = 2^(e - 14), the
//    factor the GEMM applies to its outputs.
struct SplitOperand {
    void *base;
    int64_t rows_p, k_p, batch, plane_elems;
    int planes;
    int interleaved;
    uint32_t *rowmax = nullptr; // split-FP16 only
    float *scale = nullptr;     // split-FP16 only
};

inline int64_t split_plane_elems(int64_t batch, int64_t rows, int64_t k)
{
    const int64_t rp = ceil_div(rows, kSplitRowPad) * kSplitRowPad;
    const int64_t kp = ceil_div(k, kSplitKPad) * kSplitKPad;
    return batch * rp * kp;
}

// The built-in kernel rule for a shape and precision mode (0 = fp32/
// complex64, 1 = split-bf16 3M/complex64, 2 = fp64/complex128, 3 = split-bf16
// 4M/complex64, 4 = split-fp16 4M/complex64), variant 0.  Plans use the
// autotuned table (tuning.hpp) and fall back to this rule only for shapes
// the table does not cover.
int choose_gemm_kernel(int mode, int64_t M, int64_t N, int64_t K,
                       int64_t batch);

#if defined(__HIPCC__)
// `partials` (splitk_partials(batch*M*N, K) pairs of doubles) is only used by
// kGemmSplitK.
void launch_gemm(int kernel, int elem_bytes, const GemmArgs &g,
                 hipStream_t stream, void *partials = nullptr);
// src: complex64 [batch][rows][k] with row stride ld and batch stride s.
void launch_split_bf16(const void *src, int64_t rows, int64_t k, int64_t ld,
                       int64_t s, int64_t batch, const SplitOperand &dst,
                       hipStream_t stream);
void launch_gemm_bf16x4m(const SplitOperand &a, const SplitOperand &b, int64_t K_true,
                         void *C, int64_t M, int64_t N, int64_t ldc,
                         int64_t sC, int64_t batch, bool a_bcast, bool b_bcast,
                         hipStream_t stream, int variant = 0);
// Row maxima then the scaled fp16 planes of a split-FP16 operand.
void launch_split_f16(const void *src, int64_t rows, int64_t k, int64_t ld,
                      int64_t s, int64_t batch, const SplitOperand &dst,
                      hipStream_t stream);
// The same planes and scales gathered from an unpermuted source (SplitGather).
void launch_split_f16_gather(const void *src, const SplitGather &g, const SplitOperand &dst,
                             hipStream_t stream);
void launch_gemm_f16x4m(const SplitOperand &a, const SplitOperand &b, int64_t K_true,
                        void *C, int64_t M, int64_t N, int64_t ldc,
                        int64_t sC, int64_t batch, bool a_bcast, bool b_bcast,
                        hipStream_t stream, int variant = 0);
void launch_gemm_bf16x3(const SplitOperand &a, const SplitOperand &b, int64_t K_true,
                        void *C, int64_t M, int64_t N, int64_t ldc,
                        int64_t sC, int64_t batch, bool a_bcast, bool b_bcast,
                        hipStream_t stream);
// C = contract(A, B) through FusedSmall offsets, one thread per output.
void launch_fused_small(const FusedSmall &f, int elem_bytes, const void *a, const void *b,
                        void *c, hipStream_t stream);
#endif

} // namespace jet_hip

// contract_plan.hpp -- host-side planning of one pairwise contraction.
//
// The mode algebra follows Tensor::ContractTensors (Tensor.hpp:708-752) and
// CudaTensor::GetCudaContractionPlan (CudaTensor.hpp:483-644): modes shared by
// A and B and absent from C are summed, modes of one operand only are free,
// and (an extension used by the batched slice loop) modes in A, B and C are
// batch modes.  The plan maps the contraction onto
//     permute(A -> At[batch][left][common]),
//     permute(B -> Bt[batch][right][common]),
//     C' [batch][left][right] = At . Bt^T   (batched NT complex GEMM),
//     permute(C' -> C)
// choosing the order of the batch/left/right modes from C and of the common
// modes from whichever operand avoids the larger permutation; every
// permutation that is a no-op in memory is skipped.
//
// Pure C++ so that the test suite can check plans on the CPU.
#pragma once

#include "gemm.hpp"
#include "permute_plan.hpp"

#include <cstddef>
#include <cstdint>
#include <vector>

namespace jet_hip {

struct ContractionDesc {
    int elem_bytes = 8; // 8 complex64, 16 complex128
    int mode = 0;       // 0 fp32, 1 bf16x3 3M, 2 fp64, 3 bf16x3 4M, 4 f16x3 4M
    int64_t batch = 1, M = 1, N = 1, K = 1;
    int64_t size_a = 1, size_b = 1, size_c = 1;
    bool need_a = false, need_b = false, need_c = false;
    // split-FP16: the permutation of A (B) folded into its split pass
    bool fuse_a = false, fuse_b = false;
    SplitGather ga, gb;
    PermutePlan pa, pb, pc;
    // a small step on the one-thread-per-output VALU kernel whose operands
    // or result need permuting: one fused pass instead (FusedSmall)
    bool fused_small = false;
    FusedSmall fs;
    int gemm_kernel = kGemmSmall;
    int gemm_variant = 0; // tile variant of gemm_kernel (GemmChoice)
    // K split into `chunks` pieces that run as a batch of the MFMA GEMM
    // (leading chunk modes), then summed: a step with many outputs and a
    // long K whose tuned kernel is a VALU one.  batch and K above are then
    // the GEMM's (chunks * result batch, K / chunks); size_c is the result's.
    int64_t chunks = 1;
    // Workspace layout (bytes, 256-aligned offsets).
    size_t off_at = 0, off_bt = 0, off_ct = 0, off_sa = 0, off_sb = 0, off_partials = 0;
    size_t off_chunks = 0; // the GEMM's per-chunk results (chunks > 1)
    // split-FP16 row maxima and scales of A and B (batch * padded rows each)
    size_t off_rmax_a = 0, off_scale_a = 0, off_rmax_b = 0, off_scale_b = 0;
    int64_t split_a_plane = 0, split_b_plane = 0;
    size_t workspace = 0;
    double flops = 0; // 8 * batch * M * N * K
};

// layout: 0 row-major, 1 column-major.  mode as in ContractionDesc.
ContractionDesc plan_contraction(int elem_bytes, int mode, int layout,
                                 int ra, const int32_t *modes_a,
                                 const int64_t *ext_a, int rb,
                                 const int32_t *modes_b, const int64_t *ext_b,
                                 int rc, const int32_t *modes_c);

// Workspace offsets and size of a description whose shape, permutation
// flags and GEMM kernel are set.
void layout_workspace(ContractionDesc &d);

// Gather plan for the split-FP16 operand [batch][rows][K] (target, K the
// product of the last modes' extents) of a source tensor (row-major modes
// and extents); false when the folded pass does not apply: an extent that
// is not a power of two, K outside 32..1024, fewer than 16 rows, or a
// source whose four lowest address bits (a 128-byte line) do not fit in a
// tile with the K bits.
bool plan_split_gather(const std::vector<int32_t> &src, const std::vector<int64_t> &ext,
                       const std::vector<int32_t> &target, int64_t K, SplitGather &g);

// Fused-pass offsets (FusedSmall) of C[batch, left, right] = sum over common
// of A * B with row-major operands A (modes a, extents ea), B and C (modes
// c); false when an extent is not a power of two, a group has more than
// kFusedMaxBits bits, K exceeds 2^kFusedMaxLog2K or an operand 2^31 elements.
bool plan_fused_small(const std::vector<int32_t> &a, const std::vector<int64_t> &ea,
                      const std::vector<int32_t> &b, const std::vector<int64_t> &eb,
                      const std::vector<int32_t> &c, const std::vector<int32_t> &batch,
                      const std::vector<int32_t> &left, const std::vector<int32_t> &right,
                      const std::vector<int32_t> &common, FusedSmall &f);

} // namespace jet_hip

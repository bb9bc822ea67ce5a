// network_plan.hpp -- host-side plan of a contraction sequence.
//
// A network plan turns a list of leaf tensors and an SSA contraction path
// (leaves 0..n-1, step i produces node n+i, the convention of PathInfo,
// PathInfo.hpp:262-297) into one ContractionDesc per step plus a device
// memory layout:
//   * step results keep the GEMM's natural order [batch][left][right], so no
//     result is permuted except the root, which is produced in the caller's
//     mode order;
//   * "open" modes (the sliced indices of a batched slice loop) are never
//     summed by a step: when both operands hold one it becomes a batch mode,
//     so the root carries one value per slice (the per-slice results that
//     TaskBasedContractor stores in results_, TaskBasedContractor.hpp:
//     192-197, 407-426);
//   * intermediates live in one arena; each is released after its consumer
//     step and the offsets come from a best-fit simulation of that sequence;
//   * one workspace, sized for the largest step, serves every step (the
//     steps run in order on one stream).
//
// Pure C++ so that plans can be checked on the CPU.
#pragma once

#include "contract_plan.hpp"

#include <cstdint>
#include <vector>

namespace jet_hip {

struct NetworkStep {
    int a = 0, b = 0;                 // SSA ids of the operands
    std::vector<int32_t> modes;       // modes of the result (row-major order)
    std::vector<int64_t> extents;
    ContractionDesc desc;
    size_t offset = 0;                // arena offset of the result
    bool is_root = false;
};

struct NetworkPlan {
    int elem_bytes = 8;
    int n_leaves = 0;
    std::vector<std::vector<int32_t>> leaf_modes;
    std::vector<std::vector<int64_t>> leaf_extents;
    std::vector<NetworkStep> steps;
    std::vector<int32_t> out_modes;
    std::vector<int64_t> out_extents;
    size_t arena_bytes = 0;
    size_t workspace_bytes = 0;
    double flops = 0;                 // sum of 8*M*N*K*batch over steps
    PermutePlan root_copy;            // used when there are no steps
    // Contraction DAGs (plan_dag): no root and no output; `kept` marks the
    // step results that stay in the arena (a contiguous region at its end,
    // in the order they were listed) after an execution.
    bool dag = false;
    std::vector<char> kept;           // per step
};

NetworkPlan plan_network(int elem_bytes, int mode, int n_leaves,
                         const int32_t *leaf_ranks, const int32_t *leaf_modes,
                         const int64_t *leaf_extents, int n_steps,
                         const int32_t *path, int n_out,
                         const int32_t *out_modes);

// A contraction DAG: step s contracts nodes steps[2s] and steps[2s+1]
// (leaves 0..n_leaves-1; step s's result is node n_leaves + s) over all the
// modes they share; the result holds the first operand's other modes, then
// the second's.  A node may be an operand of several steps.  Results of the
// steps in `keep` stay in the arena after the execution; the others are
// released after their last use.
NetworkPlan plan_dag(int elem_bytes, int mode, int n_leaves, const int32_t *leaf_ranks,
                     const int32_t *leaf_modes, const int64_t *leaf_extents, int n_steps,
                     const int32_t *steps, int n_keep, const int32_t *keep);

} // namespace jet_hip

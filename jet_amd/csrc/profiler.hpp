// profiler.hpp -- optional per-kernel timing with HIP events.
//
// When enabled (jet_hip_profiler_enable), every kernel launch of the library
// is bracketed by two events on its own stream and tagged with the kernel
// family and its algorithmic work (flops and minimum HBM bytes).  Reading the
// profile waits for the recorded events and accumulates, per family, the
// launch count, total device time and total algorithmic work.
#pragma once

#include "common.hpp"

namespace jet_hip {
namespace prof {

enum Kernel {
    kPermuteBits = 0,
    kPermuteGeneric,
    kCopy,
    kGemmSmall,
    kGemmSmallRed,
    kGemmF32,
    kGemmF64,
    kSplitBf16,
    kGemmBf16x3,
    kAdd,
    kFill,
    kReduce,
    kConvert,
    kGemmSplitK,
    kGemmBf16x4m,
    kSplitF16,
    kGemmF16x4m,
    kNumKernels
};

const char *kernel_name(int k);
void set_enabled(bool on);
bool enabled();
void reset();
// Accumulated totals per kernel family (waits for pending events).
void read(int64_t *launches, double *ms, double *flops, double *bytes);

class Scope {
  public:
    Scope(int kernel, hipStream_t s, double flops, double bytes);
    ~Scope();
    Scope(const Scope &) = delete;
    Scope &operator=(const Scope &) = delete;

  private:
    int kernel_;
    hipStream_t stream_;
    double flops_, bytes_;
    hipEvent_t start_ = nullptr;
};

} // namespace prof
} // namespace jet_hip

// common.hpp -- HIP-side shared types and helpers of libjet_hip.
#pragma once

#include <hip/hip_runtime.h>

#include "errors.hpp"

namespace jet_hip {

#define JH_HIP(call)                                                           \
    do {                                                                       \
        hipError_t e_ = (call);                                                \
        if (e_ != hipSuccess) {                                                \
            std::string m_ = std::string(#call) + ": " + hipGetErrorString(e_); \
            JH_FAIL(e_ == hipErrorOutOfMemory ? JET_HIP_ERROR_OUT_OF_MEMORY    \
                                              : JET_HIP_ERROR_DEVICE,          \
                    m_.c_str());                                               \
        }                                                                      \
    } while (0)

} // namespace jet_hip

// errors.hpp -- error plumbing shared by host-only and HIP sources.
#pragma once

#include "jet_hip.h"

#include <cstdint>
#include <sstream>
#include <stdexcept>
#include <string>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define JH_HD __host__ __device__ __forceinline__
#else
#define JH_HD inline
#endif

namespace jet_hip {

// Error raised inside the library; the C ABI turns it into a status code and
// a thread-local message in the format of Jet::Abort (Abort.hpp:90-97).
struct Error : std::runtime_error {
    int status;
    Error(int s, const std::string &m) : std::runtime_error(m), status(s) {}
};

[[noreturn]] inline void raise(int status, const char *msg, const char *file,
                               int line, const char *func)
{
    std::ostringstream os;
    os << "[" << file << "][Line:" << line << "][Method:" << func
       << "]: Error in Jet: " << msg;
    throw Error(status, os.str());
}

#define JH_FAIL(status, msg) ::jet_hip::raise(status, msg, __FILE__, __LINE__, __func__)
#define JH_REQUIRE(cond, msg)                                                  \
    do {                                                                       \
        if (!(cond)) JH_FAIL(JET_HIP_ERROR_INVALID_VALUE, msg);                \
    } while (0)

inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

} // namespace jet_hip

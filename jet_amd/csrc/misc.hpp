// misc.hpp -- elementwise helpers: add, seeded counter-based fill, FP64
// compensated reduction.
#pragma once

#include "errors.hpp"

#include <cmath>
#include <cstdint>

namespace jet_hip {

// splitmix64 finaliser (Steele, Lea, Flood, "Fast splittable pseudorandom
// number generators", OOPSLA 2014).
JH_HD uint64_t splitmix64(uint64_t z)
{
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// Uniform double in [0, 1) for (seed, counter): 53 high bits of the hash.
JH_HD double counter_uniform(uint64_t seed, uint64_t counter)
{
    const uint64_t x = splitmix64(seed ^ splitmix64(counter));
    return double(x >> 11) * (1.0 / 9007199254740992.0);
}

// dist 0: (u0, u1); 1: (2u0-1, 2u1-1); 2: Box-Muller pair with variance 1/2
// per component.
JH_HD void counter_transform(int dist, double u0, double u1, double *re,
                             double *im)
{
    if (dist == 0) {
        *re = u0;
        *im = u1;
    }
    else if (dist == 1) {
        *re = 2.0 * u0 - 1.0;
        *im = 2.0 * u1 - 1.0;
    }
    else {
        const double r = sqrt(-log(1.0 - u0)); // sqrt(-2 ln(1-u0)) * sqrt(1/2)
        const double t = 6.283185307179586 * u1;
        *re = r * cos(t);
        *im = r * sin(t);
    }
}

#if defined(__HIPCC__)
void launch_add(int elem_bytes, const void *a, const void *b, void *c,
                int64_t n, hipStream_t stream);
void launch_fill(int elem_bytes, void *out, int64_t n, uint64_t seed, int dist,
                 hipStream_t stream);
void launch_reduce(int elem_bytes, const void *x, int64_t n, double *out,
                   hipStream_t stream);
// complex64 <-> complex128 (rounding to nearest when narrowing).
void launch_convert(int src_bytes, int dst_bytes, const void *in, void *out,
                    int64_t n, hipStream_t stream);
// out[i] = sum over c < chunks of part[c * n + i], in FP64 (complex64 or
// complex128 by elem_bytes).
void launch_sum_chunks(int elem_bytes, const void *part, void *out, int64_t n,
                       int64_t chunks, hipStream_t stream);
// The FP16 MFMA rate loop on `blocks` work-groups of four waves: sink holds
// blocks * 256 floats, clocks 8 * blocks (shader-clock and constant-rate
// timer deltas per wave); rounds is a multiple of 4.
void launch_mfma_loop(float *sink, long long *clocks, int blocks, int rounds,
                      hipStream_t stream);
#endif

} // namespace jet_hip

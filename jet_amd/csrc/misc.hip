// misc.hip -- elementwise add, seeded fill, the FP64 compensated
// reduction and the FP16 MFMA rate loop (gfx950).
#include "common.hpp"
#include "misc.hpp"
#include "profiler.hpp"

#include <algorithm>

namespace jet_hip {

namespace {

template <class C2>
__global__ __launch_bounds__(256) void add_kernel(const C2 *__restrict__ a,
                                                  const C2 *__restrict__ b,
                                                  C2 *__restrict__ c, int64_t n)
{
    const int64_t stride = int64_t(gridDim.x) * blockDim.x;
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) {
        const C2 x = a[i], y = b[i];
        c[i] = C2{x.x + y.x, x.y + y.y};
    }
}

template <class C2>
__global__ __launch_bounds__(256) void fill_kernel(C2 *__restrict__ out, int64_t n,
                                                   uint64_t seed, int dist)
{
    using R = decltype(C2{}.x);
    const int64_t stride = int64_t(gridDim.x) * blockDim.x;
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) {
        const double u0 = counter_uniform(seed, uint64_t(2 * i));
        const double u1 = counter_uniform(seed, uint64_t(2 * i + 1));
        double re, im;
        counter_transform(dist, u0, u1, &re, &im);
        out[i] = C2{R(re), R(im)};
    }
}

// Error-free sum (Knuth TwoSum): s + e == a + b exactly.
__device__ __forceinline__ void two_sum(double a, double b, double &s, double &e)
{
    s = a + b;
    const double bb = s - a;
    e = (a - (s - bb)) + (b - bb);
}

template <class C2>
__global__ __launch_bounds__(1024) void reduce_kernel(const C2 *__restrict__ x, int64_t n,
                                                      double *__restrict__ out)
{
    __shared__ double hs[2][1024];
    __shared__ double ls[2][1024];
    // Each thread: double-double (hi, lo) accumulation of its strided subset.
    double h[2] = {0.0, 0.0}, l[2] = {0.0, 0.0};
    for (int64_t i = threadIdx.x; i < n; i += 1024) {
        const C2 v = x[i];
        const double vv[2] = {double(v.x), double(v.y)};
#pragma unroll
        for (int p = 0; p < 2; p++) {
            double s, e;
            two_sum(h[p], vv[p], s, e);
            h[p] = s;
            l[p] += e;
        }
    }
#pragma unroll
    for (int p = 0; p < 2; p++) {
        hs[p][threadIdx.x] = h[p];
        ls[p][threadIdx.x] = l[p];
    }
    __syncthreads();
    for (int s = 512; s > 0; s >>= 1) {
        if (int(threadIdx.x) < s) {
#pragma unroll
            for (int p = 0; p < 2; p++) {
                double sum, e;
                two_sum(hs[p][threadIdx.x], hs[p][threadIdx.x + s], sum, e);
                hs[p][threadIdx.x] = sum;
                ls[p][threadIdx.x] += ls[p][threadIdx.x + s] + e;
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[0] = hs[0][0] + ls[0][0];
        out[1] = hs[1][0] + ls[1][0];
    }
}

template <class S, class D>
__global__ __launch_bounds__(256) void convert_kernel(const S *__restrict__ in,
                                                      D *__restrict__ out, int64_t n)
{
    using R = decltype(D{}.x);
    const int64_t stride = int64_t(gridDim.x) * blockDim.x;
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) {
        const S v = in[i];
        out[i] = D{R(v.x), R(v.y)};
    }
}

int grid_for(int64_t n) { return (int)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 65536)); }


// out[i] = sum over c of part[c * n + i] (c ascending, accumulated in FP64):
// the per-chunk results of a GEMM whose K was split into batched chunks.
template <class C2>
__global__ __launch_bounds__(256) void sum_chunks_kernel(const C2 *__restrict__ part,
                                                         C2 *__restrict__ out, int64_t n,
                                                         int64_t chunks)
{
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += int64_t(gridDim.x) * blockDim.x) {
        double re = 0.0, im = 0.0;
        for (int64_t c = 0; c < chunks; c++) {
            const C2 v = part[c * n + i];
            re += double(v.x);
            im += double(v.y);
        }
        out[i].x = decltype(out[i].x)(re);
        out[i].y = decltype(out[i].y)(im);
    }
}

typedef _Float16 mfma_h8 __attribute__((ext_vector_type(8)));
typedef float mfma_f4 __attribute__((ext_vector_type(4)));

// FP16 MFMA rate loop: every wave runs `rounds` (a multiple of 4) rounds of
// eight independent v_mfma_f32_16x16x32_f16 accumulator chains on operands
// held in registers -- the instruction of the split GEMMs, with no memory
// traffic -- and records its shader-clock and constant-rate-timer deltas.
__global__ __launch_bounds__(256) void mfma_loop_kernel(float *__restrict__ sink,
                                                        long long *__restrict__ clocks,
                                                        int rounds)
{
    const int tid = threadIdx.x;
    // Uniform operands in (-1, 1) with full mantissas, like the split
    // operands' planes: the chip's power draw, hence the clock it holds,
    // depends on the operand bits (low-entropy operands run up to ~20%
    // faster).
    mfma_h8 a[4], b[4];
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int e = 0; e < 8; e++) {
            const uint64_t c = (uint64_t(blockIdx.x) * 256 + tid) * 64 + i * 8 + e;
            a[i][e] = (_Float16)(float(2.0 * counter_uniform(11, c) - 1.0));
            b[i][e] = (_Float16)(float(2.0 * counter_uniform(13, c) - 1.0));
        }
    mfma_f4 acc[8];
#pragma unroll
    for (int i = 0; i < 8; i++) acc[i] = mfma_f4{0.f, 0.f, 0.f, 0.f};
    const long long c0 = clock64(), w0 = wall_clock64();
    for (int it = 0; it < rounds; it += 4) {
#pragma unroll
        for (int u = 0; u < 4; u++)
#pragma unroll
            for (int i = 0; i < 8; i++)
                acc[i] = __builtin_amdgcn_mfma_f32_16x16x32_f16(a[i & 3], b[(i + u) & 3], acc[i],
                                                                0, 0, 0);
    }
    const long long c1 = clock64(), w1 = wall_clock64();
    float s = 0.f;
#pragma unroll
    for (int i = 0; i < 8; i++) s += acc[i][0] + acc[i][3];
    sink[blockIdx.x * 256 + tid] = s;
    if ((tid & 63) == 0) {
        const int w = blockIdx.x * 4 + (tid >> 6);
        clocks[2 * w] = c1 - c0;
        clocks[2 * w + 1] = w1 - w0;
    }
}
} // namespace

void launch_add(int elem_bytes, const void *a, const void *b, void *c, int64_t n,
                hipStream_t stream)
{
    if (n == 0) return;
    prof::Scope ps(prof::kAdd, stream, 2.0 * double(n), 3.0 * elem_bytes * double(n));
    if (elem_bytes == 8)
        hipLaunchKernelGGL(add_kernel<float2>, dim3(grid_for(n)), dim3(256), 0, stream,
                           static_cast<const float2 *>(a), static_cast<const float2 *>(b),
                           static_cast<float2 *>(c), n);
    else
        hipLaunchKernelGGL(add_kernel<double2>, dim3(grid_for(n)), dim3(256), 0, stream,
                           static_cast<const double2 *>(a), static_cast<const double2 *>(b),
                           static_cast<double2 *>(c), n);
    JH_HIP(hipGetLastError());
}

void launch_fill(int elem_bytes, void *out, int64_t n, uint64_t seed, int dist,
                 hipStream_t stream)
{
    if (n == 0) return;
    JH_REQUIRE(dist >= 0 && dist <= 2, "Unknown distribution.");
    prof::Scope ps(prof::kFill, stream, 0.0, double(elem_bytes) * double(n));
    if (elem_bytes == 8)
        hipLaunchKernelGGL(fill_kernel<float2>, dim3(grid_for(n)), dim3(256), 0, stream,
                           static_cast<float2 *>(out), n, seed, dist);
    else
        hipLaunchKernelGGL(fill_kernel<double2>, dim3(grid_for(n)), dim3(256), 0, stream,
                           static_cast<double2 *>(out), n, seed, dist);
    JH_HIP(hipGetLastError());
}

void launch_convert(int src_bytes, int dst_bytes, const void *in, void *out, int64_t n,
                    hipStream_t stream)
{
    if (n == 0) return;
    prof::Scope ps(prof::kConvert, stream, 0.0, double(src_bytes + dst_bytes) * double(n));
    if (src_bytes == 8 && dst_bytes == 16)
        hipLaunchKernelGGL((convert_kernel<float2, double2>), dim3(grid_for(n)), dim3(256), 0,
                           stream, static_cast<const float2 *>(in), static_cast<double2 *>(out), n);
    else if (src_bytes == 16 && dst_bytes == 8)
        hipLaunchKernelGGL((convert_kernel<double2, float2>), dim3(grid_for(n)), dim3(256), 0,
                           stream, static_cast<const double2 *>(in), static_cast<float2 *>(out), n);
    else
        JH_FAIL(JET_HIP_ERROR_INVALID_VALUE, "Conversion needs two different data types.");
    JH_HIP(hipGetLastError());
}

void launch_reduce(int elem_bytes, const void *x, int64_t n, double *out,
                   hipStream_t stream)
{
    prof::Scope ps(prof::kReduce, stream, 8.0 * double(n), double(elem_bytes) * double(n));
    if (elem_bytes == 8)
        hipLaunchKernelGGL(reduce_kernel<float2>, dim3(1), dim3(1024), 0, stream,
                           static_cast<const float2 *>(x), n, out);
    else
        hipLaunchKernelGGL(reduce_kernel<double2>, dim3(1), dim3(1024), 0, stream,
                           static_cast<const double2 *>(x), n, out);
    JH_HIP(hipGetLastError());
}

void launch_sum_chunks(int elem_bytes, const void *part, void *out, int64_t n, int64_t chunks,
                       hipStream_t stream)
{
    JH_REQUIRE(n >= 1 && chunks >= 1, "Chunk sum: bad sizes.");
    prof::Scope ps(prof::kAdd, stream, 2.0 * double(n) * double(chunks),
                   double(elem_bytes) * double(n) * double(chunks + 1));
    if (elem_bytes == 8)
        hipLaunchKernelGGL(sum_chunks_kernel<float2>, dim3(grid_for(n)), dim3(256), 0, stream,
                           static_cast<const float2 *>(part), static_cast<float2 *>(out), n,
                           chunks);
    else
        hipLaunchKernelGGL(sum_chunks_kernel<double2>, dim3(grid_for(n)), dim3(256), 0, stream,
                           static_cast<const double2 *>(part), static_cast<double2 *>(out), n,
                           chunks);
    JH_HIP(hipGetLastError());
}

void launch_mfma_loop(float *sink, long long *clocks, int blocks, int rounds, hipStream_t stream)
{
    JH_REQUIRE(blocks > 0 && rounds > 0 && rounds % 4 == 0, "MFMA loop: bad launch shape.");
    hipLaunchKernelGGL(mfma_loop_kernel, dim3(blocks), dim3(256), 0, stream, sink, clocks, rounds);
    JH_HIP(hipGetLastError());
}

} // namespace jet_hip

// permute.hip -- LDS-tiled bit-permutation and gather kernels (gfx950).
#include "common.hpp"
#include "permute_plan.hpp"
#include "profiler.hpp"

#include <algorithm>
#include <numeric>

namespace jet_hip {

namespace {

template <class T>
__global__ __launch_bounds__(kPermThreads) void permute_bits_kernel(
    const T *__restrict__ in, T *__restrict__ out, BitPermParams p,
    int64_t ntiles)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T *lds = reinterpret_cast<T *>(smem_raw);
    constexpr int EMAX = (1 << kTileMaxBits) / kPermThreads;
    const int tsize = 1 << p.tb;
    const int tid = threadIdx.x;

    // Per-thread in-tile offsets: identical for every tile this block visits.
    int64_t ioff[EMAX], ooff[EMAX];
    int lsrc[EMAX], ldst[EMAX];
#pragma unroll
    for (int e = 0; e < EMAX; e++) {
        const int u = tid + e * kPermThreads;
        int uu = 0;
        ioff[e] = bits_tile_in_offset(p, u);
        ooff[e] = bits_tile_out_offset(p, u, &uu);
        lsrc[e] = bits_lds_index(u, (int)sizeof(T));
        ldst[e] = bits_lds_index(uu, (int)sizeof(T));
    }

    for (int64_t w = blockIdx.x; w < ntiles; w += gridDim.x) {
        int64_t bin, bout;
        bits_tile_base(p, w, &bin, &bout);
        __syncthreads();
#pragma unroll
        for (int e = 0; e < EMAX; e++)
            if (tid + e * kPermThreads < tsize) lds[lsrc[e]] = in[bin + ioff[e]];
        __syncthreads();
#pragma unroll
        for (int e = 0; e < EMAX; e++)
            if (tid + e * kPermThreads < tsize) out[bout + ooff[e]] = lds[ldst[e]];
    }
}

template <class T>
__global__ __launch_bounds__(256) void permute_generic_kernel(
    const T *__restrict__ in, T *__restrict__ out, GenericPermParams p,
    int64_t numel)
{
    const int64_t stride = int64_t(gridDim.x) * blockDim.x;
    for (int64_t o = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; o < numel;
         o += stride) {
        const int64_t off = generic_in_offset(p, o);
        out[o] = in[off];
    }
}

} // namespace

template <class T>
static void launch_typed(const PermutePlan &plan, const void *in, void *out,
                         hipStream_t stream)
{
    const double bytes = 2.0 * sizeof(T) * double(plan.numel);
    if (plan.kind == PermutePlan::kCopy) {
        prof::Scope ps(prof::kCopy, stream, 0.0, bytes);
        JH_HIP(hipMemcpyAsync(out, in, sizeof(T) * plan.numel,
                              hipMemcpyDeviceToDevice, stream));
        return;
    }
    if (plan.kind == PermutePlan::kBits) {
        prof::Scope ps(prof::kPermuteBits, stream, 0.0, bytes);
        const int64_t ntiles = int64_t(1) << plan.bits.nr;
        const int grid = (int)std::min<int64_t>(ntiles, 8192);
        const size_t lds = sizeof(T) << plan.bits.tb;
        hipLaunchKernelGGL(permute_bits_kernel<T>, dim3(grid), dim3(kPermThreads), lds,
                           stream, static_cast<const T *>(in), static_cast<T *>(out),
                           plan.bits, ntiles);
        JH_HIP(hipGetLastError());
        return;
    }
    const int grid = (int)std::min<int64_t>((plan.numel + 255) / 256, 16384);
    prof::Scope ps(prof::kPermuteGeneric, stream, 0.0, bytes);
    hipLaunchKernelGGL(permute_generic_kernel<T>, dim3(grid), dim3(256), 0, stream,
                       static_cast<const T *>(in), static_cast<T *>(out), plan.gen,
                       plan.numel);
    JH_HIP(hipGetLastError());
}

void launch_permute(const PermutePlan &plan, int elem_bytes, const void *in,
                    void *out, hipStream_t stream)
{
    if (plan.numel == 0) return;
    if (elem_bytes == 8)
        launch_typed<float2>(plan, in, out, stream);
    else if (elem_bytes == 16)
        launch_typed<double2>(plan, in, out, stream);
    else
        JH_FAIL(JET_HIP_ERROR_INVALID_VALUE, "Unsupported element size.");
}

} // namespace jet_hip

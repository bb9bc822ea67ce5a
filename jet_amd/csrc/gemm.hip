// gemm.hip -- complex GEMM kernels on CDNA4 MFMA (gfx950).
//
// Three arithmetic modes, all computing C[b][m][n] = sum_k A[b][m][k] B[b][n][k]
// (no conjugation, alpha = 1, beta = 0):
//
//  * FP32 / FP64 "4M" (cgemm_embed_kernel): the complex product is the real
//    product of A viewed as an M x 2K real matrix (re/im interleaved along k)
//    and the 2K x 2N real matrix B' with B'[2k][2n]=br, B'[2k][2n+1]=bi,
//    B'[2k+1][2n]=-bi, B'[2k+1][2n+1]=br.  Its output is C with re/im
//    interleaved, so A and C are used in their native layout and B' is formed
//    in registers.  v_mfma_{f32,f64}_16x16x4 each cover 16 rows x 8 complex
//    columns x 2 complex k.
//
//  * split-BF16 "3M" (cgemm_bf16x3_kernel): T1 = Ar.Br, T2 = Ai.Bi,
//    T3 = (Ar+Ai).(Br+Bi), Cr = T1 - T2, Ci = T3 - T1 - T2, where every real
//    product X.Y is Xh.Yh + Xh.Yl + Xl.Yh with X = Xh + Xl split into two
//    BF16 numbers (the Xl.Yl term, ~2^-18 relative, is dropped) and the sums
//    accumulate in FP32 on v_mfma_f32_16x16x32_bf16.  Operands come
//    pre-split (split_bf16_kernel) as six k-contiguous planes.
//
//  * VALU kernels for shapes too small or skinny for MFMA tiles.
#include "common.hpp"
#include "gemm.hpp"
#include "profiler.hpp"

#include <algorithm>

namespace jet_hip {

namespace {

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef double f64x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));

template <class R> struct Vec;
// Output row of accumulator item r held by lane group q (= lane / 16) in a
// 16x16x4 MFMA block: v_mfma_f32_16x16x4_f32 returns D[4q + r][lane % 16],
// v_mfma_f64_16x16x4_f64 returns D[q + 4r][lane % 16] (both measured on
// gfx950).
template <> struct Vec<float> {
    using c = float2;
    using acc = f32x4;
    static __device__ __forceinline__ int row(int q, int r) { return 4 * q + r; }
};
template <> struct Vec<double> {
    using c = double2;
    using acc = f64x4;
    static __device__ __forceinline__ int row(int q, int r) { return q + 4 * r; }
};

__device__ __forceinline__ f32x4 mfma4(float a, float b, f32x4 c)
{
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ f64x4 mfma4(double a, double b, f64x4 c)
{
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// Map the linear work-group id to an output tile.  Work-groups are dispatched
// round-robin over the 8 XCDs (each with its own L2), so first give every XCD
// a contiguous range of tile indices, then order tiles in groups of
// kGroupM tile-rows so that the work-groups resident on one XCD share A and B
// panels in its L2.
constexpr int kXcds = 8;
constexpr int kGroupM = 8;

__device__ __forceinline__ void tile_coords(int64_t id, int64_t tiles_m, int64_t tiles_n,
                                            int64_t &tm, int64_t &tn)
{
    const int64_t total = tiles_m * tiles_n;
    const int64_t q = total / kXcds, r = total % kXcds;
    const int64_t x = id % kXcds, k = id / kXcds;
    const int64_t j = x * q + (x < r ? x : r) + k;
    const int64_t per_group = int64_t(kGroupM) * tiles_n;
    const int64_t group = j / per_group;
    const int64_t first_m = group * kGroupM;
    const int64_t gsz = (tiles_m - first_m) < kGroupM ? (tiles_m - first_m) : kGroupM;
    const int64_t w = j % per_group;
    tm = first_m + w % gsz;
    tn = w / gsz;
}

// ---------------------------------------------------------------------------
// 4M complex GEMM on FP32 / FP64 MFMA.
// Block tile BM x BN complex outputs, BK complex k per LDS stage, 4 waves in
// a WM x WN grid, double-buffered LDS with register prefetch.
// Lane l of a 16x16x4 MFMA holds A[l%16][l/16] and B[l/16][l%16] and gets
// D[Vec<R>::row(l/16, r)][l%16]; here the four k-slots of four consecutive MFMAs cover
// 8 complex k: lane slot q=l/16 of MFMA g is real k 4q+g, i.e. complex k
// 2q+g/2, real part for even g and imaginary part for odd g.
// ---------------------------------------------------------------------------
template <class R, int BM, int BN, int BK, int WM, int WN>
__global__ __launch_bounds__(256) void cgemm_embed_kernel(GemmArgs g)
{
    using C2 = typename Vec<R>::c;
    using Acc = typename Vec<R>::acc;
    constexpr int WTM = BM / WM, WTN = BN / WN;
    constexpr int FM = WTM / 16, FN = WTN / 8;
    constexpr int PAD = sizeof(R) == 4 ? 2 : 1;
    constexpr int LROW = BK + PAD;
    constexpr int A_PER = BM * BK / 256;
    constexpr int B_PER = BN * BK / 256;
    static_assert(WM * WN == 4, "4 waves");
    static_assert(BK % 8 == 0 && A_PER * 256 == BM * BK && B_PER * 256 == BN * BK, "tile");
    __shared__ C2 As[2][BM * LROW];
    __shared__ C2 Bs[2][BN * LROW];

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave / WN, wn = wave % WN;
    const int64_t bz = blockIdx.z;
    int64_t tm, tn;
    tile_coords(blockIdx.x, (g.M + BM - 1) / BM, (g.N + BN - 1) / BN, tm, tn);
    const int64_t m0 = tm * BM, n0 = tn * BN;
    const C2 *__restrict__ A = static_cast<const C2 *>(g.A) + bz * g.sA;
    const C2 *__restrict__ B = static_cast<const C2 *>(g.B) + bz * g.sB;
    R *__restrict__ C = reinterpret_cast<R *>(static_cast<C2 *>(g.C) + bz * g.sC);
    const int64_t M = g.M, N = g.N, K = g.K;

    Acc acc[FM][FN];
#pragma unroll
    for (int i = 0; i < FM; i++)
#pragma unroll
        for (int j = 0; j < FN; j++) acc[i][j] = Acc{0, 0, 0, 0};

    C2 ra[A_PER], rb[B_PER];
    const C2 zero = {R(0), R(0)};
    auto gload = [&](int64_t k0) {
#pragma unroll
        for (int i = 0; i < A_PER; i++) {
            const int e = tid + i * 256, r = e / BK, kk = e % BK;
            const int64_t m = m0 + r, k = k0 + kk;
            ra[i] = (m < M && k < K) ? A[m * g.lda + k] : zero;
        }
#pragma unroll
        for (int i = 0; i < B_PER; i++) {
            const int e = tid + i * 256, r = e / BK, kk = e % BK;
            const int64_t n = n0 + r, k = k0 + kk;
            rb[i] = (n < N && k < K) ? B[n * g.ldb + k] : zero;
        }
    };
    auto swrite = [&](int buf) {
#pragma unroll
        for (int i = 0; i < A_PER; i++) {
            const int e = tid + i * 256;
            As[buf][(e / BK) * LROW + e % BK] = ra[i];
        }
#pragma unroll
        for (int i = 0; i < B_PER; i++) {
            const int e = tid + i * 256;
            Bs[buf][(e / BK) * LROW + e % BK] = rb[i];
        }
    };

    const int64_t nk = (K + BK - 1) / BK;
    gload(0);
    swrite(0);
    __syncthreads();
    const int q = lane >> 4, j = lane & 15, t = lane & 1;
    for (int64_t kt = 0; kt < nk; kt++) {
        const int buf = int(kt & 1);
        if (kt + 1 < nk) gload((kt + 1) * BK);
#pragma unroll
        for (int kg = 0; kg < BK / 8; kg++) {
            C2 a0[FM], a1[FM], b0[FN], b1[FN];
#pragma unroll
            for (int fm = 0; fm < FM; fm++) {
                const int base = (wm * WTM + fm * 16 + j) * LROW + kg * 8 + 2 * q;
                a0[fm] = As[buf][base];
                a1[fm] = As[buf][base + 1];
            }
#pragma unroll
            for (int fn = 0; fn < FN; fn++) {
                const int base = (wn * WTN + fn * 8 + (j >> 1)) * LROW + kg * 8 + 2 * q;
                b0[fn] = Bs[buf][base];
                b1[fn] = Bs[buf][base + 1];
            }
#pragma unroll
            for (int fn = 0; fn < FN; fn++) {
                const R g0 = t ? b0[fn].y : b0[fn].x;
                const R g1 = t ? b0[fn].x : -b0[fn].y;
                const R g2 = t ? b1[fn].y : b1[fn].x;
                const R g3 = t ? b1[fn].x : -b1[fn].y;
#pragma unroll
                for (int fm = 0; fm < FM; fm++) {
                    acc[fm][fn] = mfma4(a0[fm].x, g0, acc[fm][fn]);
                    acc[fm][fn] = mfma4(a0[fm].y, g1, acc[fm][fn]);
                    acc[fm][fn] = mfma4(a1[fm].x, g2, acc[fm][fn]);
                    acc[fm][fn] = mfma4(a1[fm].y, g3, acc[fm][fn]);
                }
            }
        }
        if (kt + 1 < nk) swrite(buf ^ 1);
        __syncthreads();
    }

#pragma unroll
    for (int fm = 0; fm < FM; fm++) {
#pragma unroll
        for (int fn = 0; fn < FN; fn++) {
            const int64_t col = n0 + wn * WTN + fn * 8 + (j >> 1);
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int64_t row = m0 + wm * WTM + fm * 16 + Vec<R>::row(q, r);
                if (row < M && col < N) C[(row * g.ldc + col) * 2 + t] = acc[fm][fn][r];
            }
        }
    }
}

// ---------------------------------------------------------------------------
// VALU kernels.
// ---------------------------------------------------------------------------
template <class R>
__global__ __launch_bounds__(256) void cgemm_small_kernel(GemmArgs g)
{
    using C2 = typename Vec<R>::c;
    const int64_t total = g.batch * g.M * g.N;
    const int64_t stride = int64_t(gridDim.x) * blockDim.x;
    const C2 *__restrict__ A = static_cast<const C2 *>(g.A);
    const C2 *__restrict__ B = static_cast<const C2 *>(g.B);
    C2 *__restrict__ C = static_cast<C2 *>(g.C);
    for (int64_t idx = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; idx < total;
         idx += stride) {
        const int64_t n = idx % g.N;
        const int64_t rest = idx / g.N;
        const int64_t m = rest % g.M;
        const int64_t b = rest / g.M;
        const C2 *a = A + b * g.sA + m * g.lda;
        const C2 *bb = B + b * g.sB + n * g.ldb;
        R sr = 0, si = 0;
        for (int64_t k = 0; k < g.K; k++) {
            const C2 x = a[k], y = bb[k];
            sr = fma(x.x, y.x, sr);
            sr = fma(-x.y, y.y, sr);
            si = fma(x.x, y.y, si);
            si = fma(x.y, y.x, si);
        }
        C[b * g.sC + m * g.ldc + n] = C2{sr, si};
    }
}

// Fused small contraction (FusedSmall): one thread per output element
// (batch b, row m, column n), the element offsets of A, B and C assembled
// from their group bits; the k offsets of A and B come from LDS tables that
// each work-group fills once.  Same summation order and FMA sequence as
// cgemm_small_kernel.
__device__ __forceinline__ int64_t fused_offset(const FusedBits &f, int64_t x)
{
    int64_t o = 0;
    for (int i = 0; i < f.n; i++) o |= ((x >> i) & 1) << f.pos[i];
    return o;
}

template <class R>
__global__ __launch_bounds__(256) void fused_small_kernel(const void *Av, const void *Bv,
                                                          void *Cv, FusedSmall f)
{
    using C2 = typename Vec<R>::c;
    extern __shared__ int32_t koff[]; // [K] offsets in A, then [K] in B
    const int64_t K = int64_t(1) << f.lk;
    int32_t *ka = koff, *kb = koff + K;
    for (int64_t k = threadIdx.x; k < K; k += blockDim.x) {
        ka[k] = int32_t(fused_offset(f.a_k, k));
        kb[k] = int32_t(fused_offset(f.b_k, k));
    }
    __syncthreads();
    const C2 *__restrict__ A = static_cast<const C2 *>(Av);
    const C2 *__restrict__ B = static_cast<const C2 *>(Bv);
    C2 *__restrict__ C = static_cast<C2 *>(Cv);
    const int64_t total = int64_t(1) << (f.lb + f.lm + f.ln);
    const int64_t stride = int64_t(gridDim.x) * blockDim.x;
    for (int64_t idx = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; idx < total;
         idx += stride) {
        const int64_t n = idx & ((int64_t(1) << f.ln) - 1);
        const int64_t m = (idx >> f.ln) & ((int64_t(1) << f.lm) - 1);
        const int64_t b = idx >> (f.ln + f.lm);
        const C2 *a = A + (fused_offset(f.a_b, b) | fused_offset(f.a_m, m));
        const C2 *bb = B + (fused_offset(f.b_b, b) | fused_offset(f.b_n, n));
        R sr = 0, si = 0;
        for (int64_t k = 0; k < K; k++) {
            const C2 x = a[ka[k]], y = bb[kb[k]];
            sr = fma(x.x, y.x, sr);
            sr = fma(-x.y, y.y, sr);
            si = fma(x.x, y.y, si);
            si = fma(x.y, y.x, si);
        }
        C[fused_offset(f.c_b, b) | fused_offset(f.c_m, m) | fused_offset(f.c_n, n)] =
            C2{sr, si};
    }
}

// The one-work-group-per-output form of the fused pass (few outputs, long
// K): threads take k = t, t + 256, ..., the partial sums are reduced across
// the wave with shuffles and across the four waves in LDS.
template <class R>
__global__ __launch_bounds__(256) void fused_small_red_kernel(const void *Av, const void *Bv,
                                                              void *Cv, FusedSmall f)
{
    using C2 = typename Vec<R>::c;
    extern __shared__ int32_t koff[]; // [K] offsets in A, then [K] in B
    __shared__ R red_s[2][4];
    const int64_t K = int64_t(1) << f.lk;
    int32_t *ka = koff, *kb = koff + K;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    for (int64_t k = tid; k < K; k += blockDim.x) {
        ka[k] = int32_t(fused_offset(f.a_k, k));
        kb[k] = int32_t(fused_offset(f.b_k, k));
    }
    __syncthreads();
    const C2 *__restrict__ A = static_cast<const C2 *>(Av);
    const C2 *__restrict__ B = static_cast<const C2 *>(Bv);
    C2 *__restrict__ C = static_cast<C2 *>(Cv);
    const int64_t total = int64_t(1) << (f.lb + f.lm + f.ln);
    for (int64_t idx = blockIdx.x; idx < total; idx += gridDim.x) {
        const int64_t n = idx & ((int64_t(1) << f.ln) - 1);
        const int64_t m = (idx >> f.ln) & ((int64_t(1) << f.lm) - 1);
        const int64_t b = idx >> (f.ln + f.lm);
        const C2 *a = A + (fused_offset(f.a_b, b) | fused_offset(f.a_m, m));
        const C2 *bb = B + (fused_offset(f.b_b, b) | fused_offset(f.b_n, n));
        R sr = 0, si = 0;
        for (int64_t k = tid; k < K; k += blockDim.x) {
            const C2 x = a[ka[k]], y = bb[kb[k]];
            sr = fma(x.x, y.x, sr);
            sr = fma(-x.y, y.y, sr);
            si = fma(x.x, y.y, si);
            si = fma(x.y, y.x, si);
        }
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            sr += __shfl_xor(sr, off);
            si += __shfl_xor(si, off);
        }
        if (lane == 0) {
            red_s[0][wave] = sr;
            red_s[1][wave] = si;
        }
        __syncthreads();
        if (tid == 0)
            C[fused_offset(f.c_b, b) | fused_offset(f.c_m, m) | fused_offset(f.c_n, n)] =
                C2{(red_s[0][0] + red_s[0][1]) + (red_s[0][2] + red_s[0][3]),
                   (red_s[1][0] + red_s[1][1]) + (red_s[1][2] + red_s[1][3])};
        __syncthreads();
    }
}

// Short-K products (K <= kShortK) with many outputs are bound by writing C.
// Column form: a block owns up to kShortRows rows of one batch and 256
// consecutive columns; each thread keeps its column's K values of B in
// registers, reads the (wave-uniform) A values of each row and writes one
// output per row, consecutive threads to consecutive addresses.  No
// division per output.
constexpr int kShortK = 16;
constexpr int kShortRows = 32;

template <class R>
__global__ __launch_bounds__(256) void cgemm_shortk_cols_kernel(GemmArgs g, int64_t m_tiles,
                                                                int64_t n_tiles)
{
    using C2 = typename Vec<R>::c;
    const int64_t t = blockIdx.x;
    const int64_t nt = t % n_tiles, rest = t / n_tiles;
    const int64_t mt = rest % m_tiles, b = rest / m_tiles;
    const int64_t n = nt * 256 + threadIdx.x;
    if (n >= g.N) return;
    const int K = int(g.K);
    const C2 *bb = static_cast<const C2 *>(g.B) + b * g.sB + n * g.ldb;
    C2 y[kShortK];
#pragma unroll
    for (int k = 0; k < kShortK; k++) y[k] = k < K ? bb[k] : C2{0, 0};
    const C2 *A = static_cast<const C2 *>(g.A) + b * g.sA;
    C2 *C = static_cast<C2 *>(g.C) + b * g.sC + n;
    const int64_t m_end = min(g.M, (mt + 1) * int64_t(kShortRows));
    for (int64_t m = mt * kShortRows; m < m_end; m++) {
        const C2 *a = A + m * g.lda;
        R sr = 0, si = 0;
#pragma unroll
        for (int k = 0; k < kShortK; k++) {
            if (k < K) {
                const C2 x = a[k];
                sr = fma(x.x, y[k].x, sr);
                sr = fma(-x.y, y[k].y, sr);
                si = fma(x.x, y[k].y, si);
                si = fma(x.y, y[k].x, si);
            }
        }
        C[m * g.ldc] = C2{sr, si};
    }
}

// Row form, for few columns: a thread owns one row (its K values of A in
// registers) and writes that row's N outputs; B's values are wave-uniform.
template <class R>
__global__ __launch_bounds__(256) void cgemm_shortk_rows_kernel(GemmArgs g, int64_t m_tiles)
{
    using C2 = typename Vec<R>::c;
    const int64_t t = blockIdx.x;
    const int64_t mt = t % m_tiles, b = t / m_tiles;
    const int64_t m = mt * 256 + threadIdx.x;
    if (m >= g.M) return;
    const int K = int(g.K);
    const C2 *a = static_cast<const C2 *>(g.A) + b * g.sA + m * g.lda;
    C2 x[kShortK];
#pragma unroll
    for (int k = 0; k < kShortK; k++) x[k] = k < K ? a[k] : C2{0, 0};
    const C2 *B = static_cast<const C2 *>(g.B) + b * g.sB;
    C2 *C = static_cast<C2 *>(g.C) + b * g.sC + m * g.ldc;
    for (int64_t n = 0; n < g.N; n++) {
        const C2 *bb = B + n * g.ldb;
        R sr = 0, si = 0;
#pragma unroll
        for (int k = 0; k < kShortK; k++) {
            if (k < K) {
                const C2 y = bb[k];
                sr = fma(x[k].x, y.x, sr);
                sr = fma(-x[k].y, y.y, sr);
                si = fma(x[k].x, y.y, si);
                si = fma(x[k].y, y.x, si);
            }
        }
        C[n] = C2{sr, si};
    }
}

template <class R>
__global__ __launch_bounds__(256) void cgemm_smallred_kernel(GemmArgs g)
{
    using C2 = typename Vec<R>::c;
    __shared__ R red[2][256];
    const int64_t o = blockIdx.x;
    const int64_t n = o % g.N;
    const int64_t rest = o / g.N;
    const int64_t m = rest % g.M;
    const int64_t b = rest / g.M;
    const C2 *a = static_cast<const C2 *>(g.A) + b * g.sA + m * g.lda;
    const C2 *bb = static_cast<const C2 *>(g.B) + b * g.sB + n * g.ldb;
    R sr = 0, si = 0;
    for (int64_t k = threadIdx.x; k < g.K; k += 256) {
        const C2 x = a[k], y = bb[k];
        sr = fma(x.x, y.x, sr);
        sr = fma(-x.y, y.y, sr);
        si = fma(x.x, y.y, si);
        si = fma(x.y, y.x, si);
    }
    red[0][threadIdx.x] = sr;
    red[1][threadIdx.x] = si;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (int(threadIdx.x) < s) {
            red[0][threadIdx.x] += red[0][threadIdx.x + s];
            red[1][threadIdx.x] += red[1][threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0)
        static_cast<C2 *>(g.C)[b * g.sC + m * g.ldc + n] = C2{red[0][0], red[1][0]};
}

// ---------------------------------------------------------------------------
// Split-BF16 operand preparation.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void split2(float v, __bf16 &h, __bf16 &l)
{
    h = (__bf16)v;
    l = (__bf16)(v - (float)h);
}

__global__ __launch_bounds__(256) void split_bf16_kernel(
    const float2 *__restrict__ src, int64_t rows, int64_t K, int64_t ld,
    int64_t s, SplitOperand dst)
{
    const int64_t kq_n = dst.k_p / 8;
    const int64_t total = dst.batch * dst.rows_p * kq_n;
    const int64_t stride = int64_t(gridDim.x) * blockDim.x;
    __bf16 *base = static_cast<__bf16 *>(dst.base);
    for (int64_t gi = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; gi < total;
         gi += stride) {
        const int64_t kq = gi % kq_n;
        const int64_t rest = gi / kq_n;
        const int64_t r = rest % dst.rows_p;
        const int64_t b = rest / dst.rows_p;
        bf16x8 o[6];
#pragma unroll
        for (int e = 0; e < 8; e++) {
            const int64_t k = kq * 8 + e;
            float2 x = {0.f, 0.f};
            if (r < rows && k < K) x = src[b * s + r * ld + k];
            __bf16 h, l;
            split2(x.x, h, l);
            o[0][e] = h;
            o[1][e] = l;
            split2(x.y, h, l);
            o[2][e] = h;
            o[3][e] = l;
            split2(x.x + x.y, h, l);
            o[4][e] = h;
            o[5][e] = l;
        }
        const int64_t row = b * dst.rows_p + r;
        if (dst.interleaved) {
            // [row][k tile][plane][32]
            const int64_t off = (row * (dst.k_p / 32) + kq / 4) * 128 + (kq % 4) * 8;
#pragma unroll
            for (int p = 0; p < 4; p++)
                *reinterpret_cast<bf16x8 *>(base + off + p * 32) = o[p];
        }
        else {
            const int64_t off = row * dst.k_p + kq * 8;
#pragma unroll
            for (int p = 0; p < 6; p++)
                if (p < dst.planes)
                    *reinterpret_cast<bf16x8 *>(base + p * dst.plane_elems + off) = o[p];
        }
    }
}

// ---------------------------------------------------------------------------
// Split-FP16 operand preparation.  fp16 keeps 11 significant bits, so
// hi + lo carries 22 (bf16: 16), but its range is narrow: each row is first
// scaled by a power of two that brings its largest |re| or |im| into
// [2^14, 2^15), the scale being undone on the GEMM outputs.
//   1. row_max_clear_kernel: rowmax[] = 0;
//   2. row_absmax_kernel: rowmax[row] = max over the row of the FP32 bit
//      patterns of |re| and |im| (integer order = magnitude order for
//      non-negative floats), one work-group per (row, 2048-element chunk);
//   3. split_f16_kernel: scaled planes and scale[row] = 2^(e - 14).
// ---------------------------------------------------------------------------
constexpr int kRowMaxChunk = 2048;

__global__ __launch_bounds__(256) void row_max_clear_kernel(uint32_t *rowmax, int64_t n)
{
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += int64_t(gridDim.x) * blockDim.x)
        rowmax[i] = 0u;
}

__global__ __launch_bounds__(256) void row_absmax_kernel(
    const float2 *__restrict__ src, int64_t rows, int64_t K, int64_t ld, int64_t s,
    int64_t batch, int64_t rows_p, uint32_t *__restrict__ rowmax)
{
    __shared__ uint32_t red[4];
    const int64_t nch = (K + kRowMaxChunk - 1) / kRowMaxChunk;
    const int64_t units = batch * rows * nch;
    for (int64_t u = blockIdx.x; u < units; u += gridDim.x) {
        const int64_t ch = u % nch;
        const int64_t rest = u / nch;
        const int64_t r = rest % rows;
        const int64_t b = rest / rows;
        const float2 *row = src + b * s + r * ld;
        const int64_t k1 = (ch + 1) * kRowMaxChunk < K ? (ch + 1) * kRowMaxChunk : K;
        uint32_t m = 0u;
        for (int64_t k = ch * kRowMaxChunk + threadIdx.x; k < k1; k += 256) {
            const float2 x = row[k];
            const uint32_t xr = __float_as_uint(x.x) & 0x7fffffffu;
            const uint32_t xi = __float_as_uint(x.y) & 0x7fffffffu;
            m = max(m, max(xr, xi));
        }
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) m = max(m, (uint32_t)__shfl_xor((int)m, off));
        if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
        __syncthreads();
        if (threadIdx.x == 0)
            atomicMax(rowmax + b * rows_p + r, max(max(red[0], red[1]), max(red[2], red[3])));
        __syncthreads();
    }
}

// Power-of-two exponent that scales a row whose largest magnitude has FP32
// bit pattern m into [2^14, 2^15); clamped so that 2^sh stays a normal float.
__device__ __forceinline__ int row_shift(uint32_t m)
{
    if (m == 0u) return 0;
    const int e = int(m >> 23) - 127; // floor(log2) for normal values
    const int sh = 14 - e;
    return sh > 126 ? 126 : (sh < -126 ? -126 : sh);
}

__device__ __forceinline__ void split2h(float v, _Float16 &h, _Float16 &l)
{
    h = (_Float16)v;
    l = (_Float16)(v - (float)h);
}

__global__ __launch_bounds__(256) void split_f16_kernel(
    const float2 *__restrict__ src, int64_t rows, int64_t K, int64_t ld,
    int64_t s, SplitOperand dst)
{
    const int64_t kq_n = dst.k_p / 8;
    const int64_t total = dst.batch * dst.rows_p * kq_n;
    const int64_t stride = int64_t(gridDim.x) * blockDim.x;
    _Float16 *base = static_cast<_Float16 *>(dst.base);
    for (int64_t gi = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; gi < total;
         gi += stride) {
        const int64_t kq = gi % kq_n;
        const int64_t rest = gi / kq_n;
        const int64_t r = rest % dst.rows_p;
        const int64_t b = rest / dst.rows_p;
        const int64_t row = b * dst.rows_p + r;
        const int sh = r < rows ? row_shift(dst.rowmax[row]) : 0;
        const float sc = __builtin_ldexpf(1.f, sh);
        if (kq == 0) dst.scale[row] = __builtin_ldexpf(1.f, -sh);
        f16x8 o[4];
#pragma unroll
        for (int e = 0; e < 8; e++) {
            const int64_t k = kq * 8 + e;
            float2 x = {0.f, 0.f};
            if (r < rows && k < K) x = src[b * s + r * ld + k];
            _Float16 h, l;
            split2h(x.x * sc, h, l);
            o[0][e] = h;
            o[1][e] = l;
            split2h(x.y * sc, h, l);
            o[2][e] = h;
            o[3][e] = l;
        }
        // [row][k tile][plane][32], as the split-BF16 4M layout
        const int64_t off = (row * (dst.k_p / 32) + kq / 4) * 128 + (kq % 4) * 8;
#pragma unroll
        for (int p = 0; p < 4; p++) *reinterpret_cast<f16x8 *>(base + off + p * 32) = o[p];
    }
}

// Rows of at most kSplitOnePassK padded values: a group of G lanes (G a
// power of two, 4 <= G <= 64, the smallest with G >= k_p / 8) holds one row
// in registers, eight consecutive values per lane and chunk, takes its
// maximum with a shuffle reduction inside the group and writes the planes
// and the scale in the same pass (the operand is read once instead of
// twice).  64 / G rows per wave, so short rows (K = 128: 16 lanes) keep
// every lane busy.  `vec` (host-checked: 16-byte aligned source, even
// leading dimension and batch stride, K a multiple of 8) loads the eight
// values as four 16-byte loads.  Same planes and scales as the two-pass
// path.
constexpr int kSplitOnePassK = 2048;

__global__ __launch_bounds__(256) void split_f16_rows_kernel(
    const float2 *__restrict__ src, int64_t rows, int64_t K, int64_t ld,
    int64_t s, SplitOperand dst, int log2_group, int vec)
{
    constexpr int CH = kSplitOnePassK / 8 / 64; // 8-value chunks per lane at G = 64
    const int lane = threadIdx.x & 63;
    const int G = 1 << log2_group;
    const int gl = lane & (G - 1);
    const int64_t row = (int64_t(blockIdx.x) * 4 + (threadIdx.x >> 6)) * (64 >> log2_group) +
                        (lane >> log2_group);
    const int64_t nrow = dst.batch * dst.rows_p;
    const bool live = row < nrow;
    const int64_t b = live ? row / dst.rows_p : 0, r = live ? row % dst.rows_p : 0;
    const bool in_row = live && r < rows;
    const int kq_n = int(dst.k_p / 8);
    const float2 *rp = src + b * s + r * ld;
    float2 x[CH][8];
    uint32_t m = 0u;
#pragma unroll
    for (int c = 0; c < CH; c++) {
        const int kq = c * G + gl;
        const bool any = in_row && kq < kq_n && int64_t(kq) * 8 < K;
        if (vec && any) {
            const float4 *q = reinterpret_cast<const float4 *>(rp + int64_t(kq) * 8);
#pragma unroll
            for (int e = 0; e < 4; e++) {
                const float4 v = q[e];
                x[c][2 * e] = float2{v.x, v.y};
                x[c][2 * e + 1] = float2{v.z, v.w};
            }
        }
        else {
#pragma unroll
            for (int e = 0; e < 8; e++) {
                const int64_t k = int64_t(kq) * 8 + e;
                x[c][e] = (any && k < K) ? rp[k] : float2{0.f, 0.f};
            }
        }
#pragma unroll
        for (int e = 0; e < 8; e++)
            m = max(m, max(__float_as_uint(x[c][e].x) & 0x7fffffffu,
                           __float_as_uint(x[c][e].y) & 0x7fffffffu));
    }
    for (int off = G >> 1; off > 0; off >>= 1) m = max(m, (uint32_t)__shfl_xor((int)m, off));
    if (!live) return;
    const int sh = r < rows ? row_shift(m) : 0;
    const float sc = __builtin_ldexpf(1.f, sh);
    if (gl == 0) dst.scale[row] = __builtin_ldexpf(1.f, -sh);
    _Float16 *base = static_cast<_Float16 *>(dst.base);
#pragma unroll
    for (int c = 0; c < CH; c++) {
        const int kq = c * G + gl;
        if (kq >= kq_n) continue;
        f16x8 o[4];
#pragma unroll
        for (int e = 0; e < 8; e++) {
            _Float16 h, l;
            split2h(x[c][e].x * sc, h, l);
            o[0][e] = h;
            o[1][e] = l;
            split2h(x[c][e].y * sc, h, l);
            o[2][e] = h;
            o[3][e] = l;
        }
        const int64_t off = (row * (dst.k_p / 32) + kq / 4) * 128 + (kq % 4) * 8;
#pragma unroll
        for (int p = 0; p < 4; p++) *reinterpret_cast<f16x8 *>(base + off + p * 32) = o[p];
    }
}

// Rows of kSplitOnePassK < k_p <= kSplitBlockRowK padded values: one
// work-group per row holds it in registers (CH 8-value chunks per thread,
// chunk c * 256 + thread, so each load instruction of a wave reads 4 KiB of
// the row in order), reduces its maximum across the work-group and writes
// the planes and the scale: one read of the operand instead of the two-pass
// path's two.  `vec` as in split_f16_rows_kernel.
constexpr int kSplitBlockRowK = 16384;

template <int CH>
__global__ __launch_bounds__(256) void split_f16_block_rows_kernel(
    const float2 *__restrict__ src, int64_t rows, int64_t K, int64_t ld, int64_t s,
    SplitOperand dst, int vec)
{
    __shared__ uint32_t red[4];
    const int tid = threadIdx.x;
    const int64_t row = blockIdx.x;
    const int64_t b = row / dst.rows_p, r = row % dst.rows_p;
    const bool in_row = r < rows;
    const int kq_n = int(dst.k_p / 8);
    const float2 *rp = src + b * s + r * ld;
    float2 x[CH][8];
    uint32_t m = 0u;
#pragma unroll
    for (int c = 0; c < CH; c++) {
        const int kq = c * 256 + tid;
        const bool any = in_row && kq < kq_n && int64_t(kq) * 8 < K;
        if (vec && any) {
            const float4 *q = reinterpret_cast<const float4 *>(rp + int64_t(kq) * 8);
#pragma unroll
            for (int e = 0; e < 4; e++) {
                const float4 v = q[e];
                x[c][2 * e] = float2{v.x, v.y};
                x[c][2 * e + 1] = float2{v.z, v.w};
            }
        }
        else {
#pragma unroll
            for (int e = 0; e < 8; e++) {
                const int64_t k = int64_t(kq) * 8 + e;
                x[c][e] = (any && k < K) ? rp[k] : float2{0.f, 0.f};
            }
        }
#pragma unroll
        for (int e = 0; e < 8; e++)
            m = max(m, max(__float_as_uint(x[c][e].x) & 0x7fffffffu,
                           __float_as_uint(x[c][e].y) & 0x7fffffffu));
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) m = max(m, (uint32_t)__shfl_xor((int)m, off));
    if ((tid & 63) == 0) red[tid >> 6] = m;
    __syncthreads();
    m = max(max(red[0], red[1]), max(red[2], red[3]));
    const int sh = in_row ? row_shift(m) : 0;
    const float sc = __builtin_ldexpf(1.f, sh);
    if (tid == 0) dst.scale[row] = __builtin_ldexpf(1.f, -sh);
    _Float16 *base = static_cast<_Float16 *>(dst.base);
#pragma unroll
    for (int c = 0; c < CH; c++) {
        const int kq = c * 256 + tid;
        if (kq >= kq_n) continue;
        f16x8 o[4];
#pragma unroll
        for (int e = 0; e < 8; e++) {
            _Float16 h, l;
            split2h(x[c][e].x * sc, h, l);
            o[0][e] = h;
            o[1][e] = l;
            split2h(x[c][e].y * sc, h, l);
            o[2][e] = h;
            o[3][e] = l;
        }
        const int64_t off = (row * (dst.k_p / 32) + kq / 4) * 128 + (kq % 4) * 8;
#pragma unroll
        for (int p = 0; p < 4; p++) *reinterpret_cast<f16x8 *>(base + off + p * 32) = o[p];
    }
}

// Split-FP16 planes gathered from an unpermuted source (SplitGather): one
// work-group per tile of 2^tb elements (all K values of 2^xr rows).
//   1. the tile is read in source order, two neighbouring elements per
//      16-byte load (source bit 0 is always a tile bit), into LDS in source
//      order, XOR-swizzled (gather_lds_index); the loads of a work-group's
//      next tile are issued into registers before the current tile is
//      split (19% faster on a 2^31-element operand);
//   2. each thread takes its 8-value chunks of target rows from LDS into
//      registers and folds them into the row maxima (shuffles within the
//      wave for K <= 512; LDS atomics across waves for longer rows);
//   3. the planes and scales are written from registers as
//      split_f16_rows_kernel writes them.
// Every per-element offset is computed once per thread, outside the tile
// loop.
// LDS slot of tile element u: its low five bits (the bank of an 8-byte
// element in a 256-byte bank row) XORed with bits 5-9 and 10-11, so that the
// lanes of the row phase, which step through arbitrary u bits, spread over
// the banks.
__device__ __forceinline__ int gather_lds_index(int u)
{
    return u ^ (((u >> 5) ^ (u >> 10)) & 31);
}

constexpr int kGatherPairs = (1 << kGatherTileBits) / 2 / 256;
constexpr int kGatherChunks = (1 << kGatherTileBits) / 8 / 256;

__global__ __launch_bounds__(256) void split_f16_gather_kernel(
    const float2 *__restrict__ src, SplitGather g, SplitOperand dst, int vec)
{
    __shared__ float2 tile[1 << kGatherTileBits];
    __shared__ uint32_t rowmax[128];
    const int tid = threadIdx.x;
    const int tsz = 1 << g.tb;
    const int kq_bits = g.log2k - 3;
    // K <= 512: a target row's 8-value chunks lie in one wave, and the
    // shuffles alone give every lane its row maximum (no LDS atomics and
    // no barrier between the maxima and the split)
    const bool wave_rows = kq_bits <= 6;
    const int nrows = 1 << g.xr;
    // v-bit -> u-bit
    int v2u[16];
#pragma unroll
    for (int j = 0; j < 16; j++) v2u[j] = 0;
    for (int j = 0; j < g.tb; j++) v2u[g.u2v[j]] = j;
    // loads: pairs (u, u + 1), u even
    int64_t soff[kGatherPairs];
    int ls0[kGatherPairs], ls1[kGatherPairs];
#pragma unroll
    for (int e = 0; e < kGatherPairs; e++) {
        const int u = 2 * (tid + e * 256);
        int64_t o = 0;
        for (int j = 1; j < g.tb; j++)
            if ((u >> j) & 1) o += int64_t(1) << g.su[j];
        soff[e] = o;
        ls0[e] = gather_lds_index(u);
        ls1[e] = gather_lds_index(u + 1);
    }
    // chunks: 8 consecutive k of one target row
    int ue[8];
#pragma unroll
    for (int e = 0; e < 8; e++) {
        int u = 0;
        for (int j = 0; j < 3; j++)
            if ((e >> j) & 1) u |= 1 << v2u[j];
        ue[e] = u;
    }
    int cbase[kGatherChunks], crow[kGatherChunks], ckq[kGatherChunks];
    int64_t croff[kGatherChunks];
#pragma unroll
    for (int c = 0; c < kGatherChunks; c++) {
        const int q = tid + c * 256;
        const int v = q * 8;
        int u = 0;
        for (int j = 3; j < g.tb; j++)
            if ((v >> j) & 1) u |= 1 << v2u[j];
        cbase[c] = u;
        crow[c] = q >> kq_bits;
        ckq[c] = q & ((1 << kq_bits) - 1);
        int64_t ro = 0;
        for (int i = 0; i < g.xr; i++)
            if ((crow[c] >> i) & 1) ro += int64_t(1) << g.vrow[i];
        croff[c] = ro;
    }
    const int64_t ntiles = int64_t(1) << g.nw;
    const int kt = 1 << (g.log2k - 5); // 32-wide k tiles per row
    _Float16 *base = static_cast<_Float16 *>(dst.base);
    // the next tile's loads are issued before the current tile is split
    auto tile_src = [&](int64_t w) {
        int64_t sb = 0;
        for (int j = 0; j < g.nw; j++)
            if ((w >> j) & 1) sb += int64_t(1) << g.wsrc[j];
        return sb;
    };
    float4 nx[kGatherPairs];
    auto fetch = [&](int64_t w) {
        const int64_t sb = tile_src(w);
#pragma unroll
        for (int e = 0; e < kGatherPairs; e++)
            if (2 * (tid + e * 256) < tsz) {
                const float2 *q = src + sb + soff[e];
                if (vec) {
                    nx[e] = *reinterpret_cast<const float4 *>(q);
                }
                else {
                    const float2 a = q[0], b = q[1];
                    nx[e] = float4{a.x, a.y, b.x, b.y};
                }
            }
    };
    if (int64_t(blockIdx.x) < ntiles) fetch(blockIdx.x);
    for (int64_t w = blockIdx.x; w < ntiles; w += gridDim.x) {
        int64_t rbase = 0;
        for (int j = 0; j < g.nw; j++)
            if ((w >> j) & 1) rbase += int64_t(1) << g.wrow[j];
        if (!wave_rows && tid < nrows) rowmax[tid] = 0u;
#pragma unroll
        for (int e = 0; e < kGatherPairs; e++)
            if (2 * (tid + e * 256) < tsz) {
                tile[ls0[e]] = float2{nx[e].x, nx[e].y};
                tile[ls1[e]] = float2{nx[e].z, nx[e].w};
            }
        if (w + gridDim.x < ntiles) fetch(w + gridDim.x);
        __syncthreads();
        float2 x[kGatherChunks][8];
        uint32_t rm[kGatherChunks];
#pragma unroll
        for (int c = 0; c < kGatherChunks; c++) {
            uint32_t m = 0u;
            const bool live = 8 * (tid + c * 256) < tsz;
#pragma unroll
            for (int e = 0; e < 8; e++) {
                x[c][e] = live ? tile[gather_lds_index(cbase[c] | ue[e])] : float2{0.f, 0.f};
                m = max(m, max(__float_as_uint(x[c][e].x) & 0x7fffffffu,
                               __float_as_uint(x[c][e].y) & 0x7fffffffu));
            }
            // lanes of one row are consecutive: reduce within each group of
            // min(64, chunks per row) lanes; rows longer than a wave then
            // combine their groups with one LDS atomic per group
            const int grp = kq_bits < 6 ? (1 << kq_bits) : 64;
            for (int off = grp >> 1; off > 0; off >>= 1)
                m = max(m, (uint32_t)__shfl_xor((int)m, off));
            rm[c] = m;
            if (!wave_rows && live && (tid & (grp - 1)) == 0) atomicMax(&rowmax[crow[c]], m);
        }
        if (!wave_rows) __syncthreads();
#pragma unroll
        for (int c = 0; c < kGatherChunks; c++) {
            if (8 * (tid + c * 256) >= tsz) continue;
            const int sh = row_shift(wave_rows ? rm[c] : rowmax[crow[c]]);
            const float sc = __builtin_ldexpf(1.f, sh);
            const int64_t row = rbase + croff[c];
            const int kq = ckq[c];
            if (kq == 0) dst.scale[row] = __builtin_ldexpf(1.f, -sh);
            f16x8 o[4];
#pragma unroll
            for (int e = 0; e < 8; e++) {
                _Float16 h, l;
                split2h(x[c][e].x * sc, h, l);
                o[0][e] = h;
                o[1][e] = l;
                split2h(x[c][e].y * sc, h, l);
                o[2][e] = h;
                o[3][e] = l;
            }
            const int64_t off = (row * kt + kq / 4) * 128 + (kq % 4) * 8;
#pragma unroll
            for (int p = 0; p < 4; p++) *reinterpret_cast<f16x8 *>(base + off + p * 32) = o[p];
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// 3M split-BF16 complex GEMM.
// Stage s covers product s/nk (T1, T2, T3) and k tile s%nk: it loads the hi
// and lo planes of that product's operands (BM x 32 and BN x 32 bf16 each)
// and issues Xh.Yh, Xh.Yl, Xl.Yh on v_mfma_f32_16x16x32_bf16, whose lane l
// holds 8 consecutive k (8(l/16)..8(l/16)+7) of row l%16 of A and of column
// l%16 of B, and gets D[4(l/16)+r][l%16].  Two accumulator sets suffice:
// x = T1 and y = T2, then (x, y) <- (T1 - T2, -T1 - T2) and T3 accumulates
// into y, leaving x = Re C and y = Im C.
// ---------------------------------------------------------------------------
template <int BM, int BN, int WM, int WN>
__global__ __launch_bounds__(64 * WM * WN) void cgemm_bf16x3_kernel(
    SplitOperand a, SplitOperand b, float2 *__restrict__ Cout, int64_t M,
    int64_t N, int64_t ldc, int64_t sC, int a_bcast, int b_bcast)
{
    constexpr int NT = 64 * WM * WN;
    constexpr int KT = 32;
    constexpr int LROW = KT + 8; // bf16 per LDS row (16-byte pad)
    constexpr int WTM = BM / WM, WTN = BN / WN;
    constexpr int FM = WTM / 16, FN = WTN / 16;
    constexpr int A_CH = 2 * BM * (KT / 8) / NT; // 16-byte chunks per thread
    constexpr int B_CH = 2 * BN * (KT / 8) / NT;
    static_assert(A_CH * NT == 2 * BM * 4 && B_CH * NT == 2 * BN * 4, "tile");
    __shared__ bf16x8 As[2][2 * BM * LROW / 8];
    __shared__ bf16x8 Bs[2][2 * BN * LROW / 8];

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave / WN, wn = wave % WN;
    const int64_t bz = blockIdx.z;
    int64_t tm, tn;
    tile_coords(blockIdx.x, (M + BM - 1) / BM, (N + BN - 1) / BN, tm, tn);
    const int64_t m0 = tm * BM, n0 = tn * BN;
    const __bf16 *Ab = static_cast<const __bf16 *>(a.base) +
                       (a_bcast ? 0 : bz * a.rows_p * a.k_p);
    const __bf16 *Bb = static_cast<const __bf16 *>(b.base) +
                       (b_bcast ? 0 : bz * b.rows_p * b.k_p);

    f32x4 acc[2][FM][FN];
#pragma unroll
    for (int p = 0; p < 2; p++)
#pragma unroll
        for (int i = 0; i < FM; i++)
#pragma unroll
            for (int jj = 0; jj < FN; jj++) acc[p][i][jj] = f32x4{0, 0, 0, 0};

    bf16x8 ra[A_CH], rb[B_CH];
    const bf16x8 zero = {};
    const int nk = int(a.k_p / KT);
    auto gload = [&](int s) {
        const int64_t k0 = int64_t(s % nk) * KT;
        const int ph = s / nk;
#pragma unroll
        for (int i = 0; i < A_CH; i++) {
            const int c = tid + i * NT;
            const int pl = c / (BM * 4), rem = c % (BM * 4), r = rem / 4, kc = rem % 4;
            const int64_t m = m0 + r;
            ra[i] = (m < a.rows_p)
                        ? *reinterpret_cast<const bf16x8 *>(
                              Ab + (2 * ph + pl) * a.plane_elems + m * a.k_p + k0 + kc * 8)
                        : zero;
        }
#pragma unroll
        for (int i = 0; i < B_CH; i++) {
            const int c = tid + i * NT;
            const int pl = c / (BN * 4), rem = c % (BN * 4), r = rem / 4, kc = rem % 4;
            const int64_t n = n0 + r;
            rb[i] = (n < b.rows_p)
                        ? *reinterpret_cast<const bf16x8 *>(
                              Bb + (2 * ph + pl) * b.plane_elems + n * b.k_p + k0 + kc * 8)
                        : zero;
        }
    };
    auto swrite = [&](int buf) {
#pragma unroll
        for (int i = 0; i < A_CH; i++) {
            const int c = tid + i * NT;
            const int pl = c / (BM * 4), rem = c % (BM * 4), r = rem / 4, kc = rem % 4;
            As[buf][(pl * BM + r) * (LROW / 8) + kc] = ra[i];
        }
#pragma unroll
        for (int i = 0; i < B_CH; i++) {
            const int c = tid + i * NT;
            const int pl = c / (BN * 4), rem = c % (BN * 4), r = rem / 4, kc = rem % 4;
            Bs[buf][(pl * BN + r) * (LROW / 8) + kc] = rb[i];
        }
    };

    const int nstages = nk * 3;
    gload(0);
    swrite(0);
    __syncthreads();
    const int q = lane >> 4, j = lane & 15;
    for (int s = 0; s < nstages; s++) {
        const int buf = s & 1;
        const int ph = s / nk;
        if (s == 2 * nk) {
#pragma unroll
            for (int fm = 0; fm < FM; fm++)
#pragma unroll
                for (int fn = 0; fn < FN; fn++) {
                    const f32x4 t1 = acc[0][fm][fn], t2 = acc[1][fm][fn];
                    acc[0][fm][fn] = t1 - t2;
                    acc[1][fm][fn] = -(t1 + t2);
                }
        }
        if (s + 1 < nstages) gload(s + 1);
        // B fragments stay in registers for the whole stage; A fragments are
        // loaded per row block to keep the register footprint small.
        bf16x8 bh[FN], bl[FN];
#pragma unroll
        for (int fn = 0; fn < FN; fn++) {
            const int r = wn * WTN + fn * 16 + j;
            bh[fn] = Bs[buf][r * (LROW / 8) + q];
            bl[fn] = Bs[buf][(BN + r) * (LROW / 8) + q];
        }
#define JH_BF16_STAGE(P)                                                                    \
    _Pragma("unroll") for (int fm = 0; fm < FM; fm++)                                       \
    {                                                                                       \
        const int r = wm * WTM + fm * 16 + j;                                               \
        const bf16x8 ah = As[buf][r * (LROW / 8) + q];                                      \
        const bf16x8 al = As[buf][(BM + r) * (LROW / 8) + q];                               \
        _Pragma("unroll") for (int fn = 0; fn < FN; fn++)                                   \
        {                                                                                   \
            acc[P][fm][fn] =                                                                \
                __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bh[fn], acc[P][fm][fn], 0, 0, 0); \
            acc[P][fm][fn] =                                                                \
                __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bl[fn], acc[P][fm][fn], 0, 0, 0); \
            acc[P][fm][fn] =                                                                \
                __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, bh[fn], acc[P][fm][fn], 0, 0, 0); \
        }                                                                                   \
    }
        if (ph == 0) {
            JH_BF16_STAGE(0)
        }
        else {
            JH_BF16_STAGE(1)
        }
#undef JH_BF16_STAGE
        if (s + 1 < nstages) swrite(buf ^ 1);
        __syncthreads();
    }

    float2 *C = Cout + bz * sC;
#pragma unroll
    for (int fm = 0; fm < FM; fm++) {
#pragma unroll
        for (int fn = 0; fn < FN; fn++) {
            const int64_t col = n0 + wn * WTN + fn * 16 + j;
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int64_t row = m0 + wm * WTM + fm * 16 + 4 * q + r;
                if (row < M && col < N)
                    C[row * ldc + col] = float2{acc[0][fm][fn][r], acc[1][fm][fn][r]};
            }
        }
    }
}

// Byte offset, from the first row of a tile, of the 16-byte chunk that load
// lane p (tile row p / 16, LDS slot p % 16) reads in k tile 0; rows past the
// padded operand read its last row.  It stays below 2^32 because the host
// keeps k_p <= kX4mMaxK, so each direct-to-LDS load is a scalar tile base
// plus one 32-bit VGPR (64-bit per-lane addresses cost 8-16 more VGPRs and
// a 64-bit add per load).
__device__ __forceinline__ uint32_t x4m_chunk_offset(int p, int64_t t0, int64_t rows,
                                                     int64_t nkt)
{
    const int r = p >> 4, slot = p & 15;
    int64_t m = t0 + r;
    if (m >= rows) m = rows - 1;
    return uint32_t(((m - t0) * nkt * 16 + (slot ^ (r & 15))) * 16);
}

// ---------------------------------------------------------------------------
// 4M split-BF16 complex GEMM.
// Re C = Ar.Br - Ai.Bi and Im C = Ar.Bi + Ai.Br, each real product split into
// the three BF16 terms Xh.Yh + Xh.Yl + Xl.Yh (12 v_mfma_f32_16x16x32_bf16 per
// 16x16 block and 32 k).  Every k tile stages the four planes (re hi/lo, im
// hi/lo) of both operands once, so each byte loaded into LDS feeds twice as
// many MFMAs as in the 3M form, and no T3 - T1 - T2 cancellation occurs.
// Double-buffered LDS with the next k tile prefetched into registers.
// ---------------------------------------------------------------------------
// Sign flip of eight 16-bit floats (bf16 and fp16 keep the sign in bit 15).
template <class V> __device__ __forceinline__ V neg16x8(V v)
{
    typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
    u32x4 u = __builtin_bit_cast(u32x4, v);
    u ^= 0x80008000u;
    return __builtin_bit_cast(V, u);
}

// The 16x16x32 MFMA of each 16-bit operand type: lane l supplies 8
// consecutive k (8(l/16)..8(l/16)+7) of row l%16 of A and of column l%16
// of B, and holds D[4(l/16)+r][l%16].
template <class E> struct Mfma16;
template <> struct Mfma16<__bf16> {
    using vec = bf16x8;
    static __device__ __forceinline__ f32x4 mma(vec a, vec b, f32x4 c)
    {
        return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
    }
};
template <> struct Mfma16<_Float16> {
    using vec = f16x8;
    static __device__ __forceinline__ f32x4 mma(vec a, vec b, f32x4 c)
    {
        return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
    }
};

// E = __bf16: split-BF16; E = _Float16: split-FP16, whose operand rows carry
// power-of-two scales that the epilogue applies (a.scale / b.scale).
//
// Each k tile (32 wide) of a 128-row operand panel is 128 rows x 256 bytes
// (4 planes x 32 values).  It goes from global memory straight into LDS with
// global_load_lds_dwordx4 (no register staging): one instruction moves 64
// 16-byte chunks to 64 consecutive LDS slots, so LDS row r holds its 16
// chunks at slot (plane * 4 + chunk) ^ (r & 15).  That xor makes the 16 rows
// read by one MFMA operand fetch (same plane and chunk) land in 16 different
// bank groups, and lets each load instruction read 4 whole 256-byte rows.
// Rows past the padded operand are read from its last row: they only feed
// output rows or columns that are not stored.  Two LDS buffers (128 KiB):
// the next k tile streams in while the current one is multiplied.
template <class E, int BM, int BN, int WM, int WN>
__global__ __launch_bounds__(64 * WM * WN) void cgemm_x4m_kernel(
    SplitOperand a, SplitOperand b, float2 *__restrict__ Cout, int64_t M,
    int64_t N, int64_t ldc, int64_t sC, int a_bcast, int b_bcast)
{
    using vec = typename Mfma16<E>::vec;
    typedef __attribute__((address_space(3))) void lds_void;
    constexpr int NW = WM * WN;
    constexpr int KT = 32;
    constexpr int WTM = BM / WM, WTN = BN / WN;
    constexpr int FM = WTM / 16, FN = WTN / 16;
    constexpr int A_IN = BM * 16 / (64 * NW); // load instructions per wave per k tile
    constexpr int B_IN = BN * 16 / (64 * NW);
    static_assert(A_IN * 64 * NW == BM * 16 && B_IN * 64 * NW == BN * 16, "tile");
    __shared__ vec As0[BM * 16], As1[BM * 16];
    __shared__ vec Bs0[BN * 16], Bs1[BN * 16];

    const int tid = threadIdx.x, lane = tid & 63;
    // The wave index is uniform: in a scalar register, the LDS destinations
    // of the direct loads are scalar too.
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int wm = wave / WN, wn = wave % WN;
    const int64_t bz = blockIdx.z;
    int64_t tm, tn;
    tile_coords(blockIdx.x, (M + BM - 1) / BM, (N + BN - 1) / BN, tm, tn);
    const int64_t m0 = tm * BM, n0 = tn * BN;
    // interleaved layout: a batch holds rows_p * k_p * 4 values
    const E *Ab = static_cast<const E *>(a.base) + (a_bcast ? 0 : bz * a.rows_p * a.k_p * 4);
    const E *Bb = static_cast<const E *>(b.base) + (b_bcast ? 0 : bz * b.rows_p * b.k_p * 4);
    const int64_t nkt_a = a.k_p / KT, nkt_b = b.k_p / KT;
    const char *At = reinterpret_cast<const char *>(Ab + m0 * nkt_a * 128);
    const char *Bt = reinterpret_cast<const char *>(Bb + n0 * nkt_b * 128);

    uint32_t offa[A_IN], offb[B_IN];
#pragma unroll
    for (int i = 0; i < A_IN; i++)
        offa[i] = x4m_chunk_offset((wave * A_IN + i) * 64 + lane, m0, a.rows_p, nkt_a);
#pragma unroll
    for (int i = 0; i < B_IN; i++)
        offb[i] = x4m_chunk_offset((wave * B_IN + i) * 64 + lane, n0, b.rows_p, nkt_b);
    auto issue = [&](int kt, vec *as, vec *bs) {
        const uint32_t kb = uint32_t(kt) * 256u;
#pragma unroll
        for (int i = 0; i < A_IN; i++)
            __builtin_amdgcn_global_load_lds((const void *)(At + uint64_t(offa[i] + kb)),
                                             (lds_void *)(as + (wave * A_IN + i) * 64), 16, 0, 0);
#pragma unroll
        for (int i = 0; i < B_IN; i++)
            __builtin_amdgcn_global_load_lds((const void *)(Bt + uint64_t(offb[i] + kb)),
                                             (lds_void *)(bs + (wave * B_IN + i) * 64), 16, 0, 0);
    };

    f32x4 cr[FM][FN], ci[FM][FN];
#pragma unroll
    for (int i = 0; i < FM; i++)
#pragma unroll
        for (int jj = 0; jj < FN; jj++) {
            cr[i][jj] = f32x4{0, 0, 0, 0};
            ci[i][jj] = f32x4{0, 0, 0, 0};
        }

    const int q = lane >> 4, j = lane & 15;
    auto compute = [&](const vec *as, const vec *bs) {
        vec brh[FN], brl[FN], bih[FN], bil[FN];
#pragma unroll
        for (int fn = 0; fn < FN; fn++) {
            const int r = wn * WTN + fn * 16 + j;
            brh[fn] = bs[r * 16 + ((0 * 4 + q) ^ j)];
            brl[fn] = bs[r * 16 + ((1 * 4 + q) ^ j)];
            bih[fn] = bs[r * 16 + ((2 * 4 + q) ^ j)];
            bil[fn] = bs[r * 16 + ((3 * 4 + q) ^ j)];
        }
#pragma unroll
        for (int fm = 0; fm < FM; fm++) {
            const int r = wm * WTM + fm * 16 + j;
            const vec arh = as[r * 16 + ((0 * 4 + q) ^ j)];
            const vec arl = as[r * 16 + ((1 * 4 + q) ^ j)];
            const vec aih = as[r * 16 + ((2 * 4 + q) ^ j)];
            const vec ail = as[r * 16 + ((3 * 4 + q) ^ j)];
            const vec naih = neg16x8(aih), nail = neg16x8(ail);
#pragma unroll
            for (int fn = 0; fn < FN; fn++) {
                f32x4 x = cr[fm][fn], y = ci[fm][fn];
                x = Mfma16<E>::mma(arh, brh[fn], x);
                y = Mfma16<E>::mma(arh, bih[fn], y);
                x = Mfma16<E>::mma(arh, brl[fn], x);
                y = Mfma16<E>::mma(arh, bil[fn], y);
                x = Mfma16<E>::mma(arl, brh[fn], x);
                y = Mfma16<E>::mma(arl, bih[fn], y);
                x = Mfma16<E>::mma(naih, bih[fn], x);
                y = Mfma16<E>::mma(aih, brh[fn], y);
                x = Mfma16<E>::mma(naih, bil[fn], x);
                y = Mfma16<E>::mma(aih, brl[fn], y);
                x = Mfma16<E>::mma(nail, bih[fn], x);
                y = Mfma16<E>::mma(ail, brh[fn], y);
                cr[fm][fn] = x;
                ci[fm][fn] = y;
            }
        }
    };

    // vmcnt(0) (expcnt and lgkmcnt left at their maximum): this wave's
    // direct-to-LDS loads have landed; the barrier then covers all waves.
    constexpr int kWaitLoads = 0x0f70;
    const int nk = int(a.k_p / KT);
    issue(0, As0, Bs0);
    __builtin_amdgcn_s_waitcnt(kWaitLoads);
    __syncthreads();
    for (int kt = 0; kt < nk; kt += 2) {
        if (kt + 1 < nk) issue(kt + 1, As1, Bs1);
        compute(As0, Bs0);
        __builtin_amdgcn_s_waitcnt(kWaitLoads);
        __syncthreads();
        if (kt + 1 >= nk) break;
        if (kt + 2 < nk) issue(kt + 2, As0, Bs0);
        compute(As1, Bs1);
        __builtin_amdgcn_s_waitcnt(kWaitLoads);
        __syncthreads();
    }

    float2 *C = Cout + bz * sC;
    const float *sa = a.scale ? a.scale + (a_bcast ? 0 : bz) * a.rows_p : nullptr;
    const float *sb = b.scale ? b.scale + (b_bcast ? 0 : bz) * b.rows_p : nullptr;
#pragma unroll
    for (int fm = 0; fm < FM; fm++) {
#pragma unroll
        for (int fn = 0; fn < FN; fn++) {
            const int64_t col = n0 + wn * WTN + fn * 16 + j;
            const float fb = (sb && col < N) ? sb[col] : 1.f;
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int64_t row = m0 + wm * WTM + fm * 16 + 4 * q + r;
                if (row < M && col < N) {
                    // Power-of-two factors, applied one at a time so that
                    // their product cannot leave the float range early.
                    const float fa = sa ? sa[row] : 1.f;
                    C[row * ldc + col] = float2{cr[fm][fn][r] * fa * fb, ci[fm][fn][r] * fa * fb};
                }
            }
        }
    }
}

// The same product on 128x192 block tiles (LDS: two buffers of 320 rows x
// 256 bytes = 160 KiB, all of it), for N wide enough that the last column
// tile is nearly full: 17% fewer loaded bytes per output than 128x128.  Two
// differences from cgemm_x4m_kernel, both measured (a 2^14-cube GEMM runs
// 4-5% faster than the 128x128 kernel):
//  * the loads of the next k tile are issued in two halves, before the
//    first and the second A fragment's MFMAs, instead of all at once;
//  * the epilogue goes through LDS: the row and column scales are staged
//    there, and each wave writes a 16-row slice of its accumulators to its
//    own LDS area and reads it back row-major before scaling and storing.
//    Scaling the accumulators in registers makes the register allocator
//    rotate them between two register sets inside the k loop, which at this
//    tile size spills.
template <class E, int BM, int BN, int WM, int WN>
__global__ __launch_bounds__(64 * WM * WN) void cgemm_x4m_wide_kernel(
    SplitOperand a, SplitOperand b, float2 *__restrict__ Cout, int64_t M,
    int64_t N, int64_t ldc, int64_t sC, int a_bcast, int b_bcast, int vec_store)
{
    using vec = typename Mfma16<E>::vec;
    typedef __attribute__((address_space(3))) void lds_void;
    constexpr int NW = WM * WN;
    constexpr int KT = 32;
    constexpr int WTM = BM / WM, WTN = BN / WN;
    constexpr int FM = WTM / 16, FN = WTN / 16;
    constexpr int A_IN = BM * 16 / (64 * NW); // load instructions per wave per k tile
    constexpr int B_IN = BN * 16 / (64 * NW);
    constexpr int L = A_IN + B_IN;
    static_assert(A_IN * 64 * NW == BM * 16 && B_IN * 64 * NW == BN * 16, "tile");
    static_assert(FM >= 2, "loads are issued over the first two A fragments");
    __shared__ vec As[2][BM * 16];
    __shared__ vec Bs[2][BN * 16];

    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6); // uniform: scalar
    const int wm = wave / WN, wn = wave % WN;
    const int64_t bz = blockIdx.z;
    int64_t tm, tn;
    tile_coords(blockIdx.x, (M + BM - 1) / BM, (N + BN - 1) / BN, tm, tn);
    const int64_t m0 = tm * BM, n0 = tn * BN;
    const E *Ab = static_cast<const E *>(a.base) + (a_bcast ? 0 : bz * a.rows_p * a.k_p * 4);
    const E *Bb = static_cast<const E *>(b.base) + (b_bcast ? 0 : bz * b.rows_p * b.k_p * 4);
    const int64_t nkt_a = a.k_p / KT, nkt_b = b.k_p / KT;
    const char *At = reinterpret_cast<const char *>(Ab + m0 * nkt_a * 128);
    const char *Bt = reinterpret_cast<const char *>(Bb + n0 * nkt_b * 128);

    // Load i < A_IN moves A chunks, the others B chunks.
    uint32_t off[L];
#pragma unroll
    for (int i = 0; i < L; i++) {
        const bool isa = i < A_IN;
        const int p = (wave * (isa ? A_IN : B_IN) + (isa ? i : i - A_IN)) * 64 + lane;
        off[i] = isa ? x4m_chunk_offset(p, m0, a.rows_p, nkt_a)
                     : x4m_chunk_offset(p, n0, b.rows_p, nkt_b);
    }
    // Loads [lo, hi) of k tile kt into buffer s, kept in place by the
    // scheduling barriers.
    auto issue = [&](int lo, int hi, int kt, int s) {
        __builtin_amdgcn_sched_barrier(0);
        const uint32_t kb = uint32_t(kt) * 256u;
#pragma unroll
        for (int i = 0; i < L; i++) {
            if (i < lo || i >= hi) continue;
            if (i < A_IN)
                __builtin_amdgcn_global_load_lds(
                    (const void *)(At + uint64_t(off[i] + kb)),
                    (lds_void *)(&As[s][(wave * A_IN + i) * 64]), 16, 0, 0);
            else
                __builtin_amdgcn_global_load_lds(
                    (const void *)(Bt + uint64_t(off[i] + kb)),
                    (lds_void *)(&Bs[s][(wave * B_IN + i - A_IN) * 64]), 16, 0, 0);
        }
        __builtin_amdgcn_sched_barrier(0);
    };

    f32x4 cr[FM][FN], ci[FM][FN];
#pragma unroll
    for (int i = 0; i < FM; i++)
#pragma unroll
        for (int jj = 0; jj < FN; jj++) {
            cr[i][jj] = f32x4{0, 0, 0, 0};
            ci[i][jj] = f32x4{0, 0, 0, 0};
        }

    const int q = lane >> 4, j = lane & 15;
    // Multiplies the k tile in buffer s; issues k tile kt2 (if >= 0) into s2.
    auto compute = [&](int s, int kt2, int s2) {
        const vec *as = As[s], *bs = Bs[s];
        vec brh[FN], brl[FN], bih[FN], bil[FN];
#pragma unroll
        for (int fn = 0; fn < FN; fn++) {
            const int r = wn * WTN + fn * 16 + j;
            brh[fn] = bs[r * 16 + ((0 * 4 + q) ^ j)];
            brl[fn] = bs[r * 16 + ((1 * 4 + q) ^ j)];
            bih[fn] = bs[r * 16 + ((2 * 4 + q) ^ j)];
            bil[fn] = bs[r * 16 + ((3 * 4 + q) ^ j)];
        }
#pragma unroll
        for (int fm = 0; fm < FM; fm++) {
            constexpr int H = (L + 1) / 2;
            if (fm < 2 && kt2 >= 0) issue(fm * H, fm * H + H, kt2, s2);
            const int r = wm * WTM + fm * 16 + j;
            const vec arh = as[r * 16 + ((0 * 4 + q) ^ j)];
            const vec arl = as[r * 16 + ((1 * 4 + q) ^ j)];
            const vec aih = as[r * 16 + ((2 * 4 + q) ^ j)];
            const vec ail = as[r * 16 + ((3 * 4 + q) ^ j)];
            const vec naih = neg16x8(aih), nail = neg16x8(ail);
#pragma unroll
            for (int fn = 0; fn < FN; fn++) {
                f32x4 x = cr[fm][fn], y = ci[fm][fn];
                x = Mfma16<E>::mma(arh, brh[fn], x);
                y = Mfma16<E>::mma(arh, bih[fn], y);
                x = Mfma16<E>::mma(arh, brl[fn], x);
                y = Mfma16<E>::mma(arh, bil[fn], y);
                x = Mfma16<E>::mma(arl, brh[fn], x);
                y = Mfma16<E>::mma(arl, bih[fn], y);
                x = Mfma16<E>::mma(naih, bih[fn], x);
                y = Mfma16<E>::mma(aih, brh[fn], y);
                x = Mfma16<E>::mma(naih, bil[fn], x);
                y = Mfma16<E>::mma(aih, brl[fn], y);
                x = Mfma16<E>::mma(nail, bih[fn], x);
                y = Mfma16<E>::mma(ail, brh[fn], y);
                cr[fm][fn] = x;
                ci[fm][fn] = y;
            }
        }
    };

    // vmcnt(0): this wave's direct-to-LDS loads have landed; the barrier
    // then covers all waves.
    constexpr int kWaitLoads = 0x0f70;
    const int nk = int(a.k_p / KT);
    issue(0, L, 0, 0);
    __builtin_amdgcn_s_waitcnt(kWaitLoads);
    __syncthreads();
    for (int kt = 0; kt < nk; kt += 2) {
        compute(0, kt + 1 < nk ? kt + 1 : -1, 1);
        __builtin_amdgcn_s_waitcnt(kWaitLoads);
        __syncthreads();
        if (kt + 1 >= nk) break;
        compute(1, kt + 2 < nk ? kt + 2 : -1, 0);
        __builtin_amdgcn_s_waitcnt(kWaitLoads);
        __syncthreads();
    }

    // Epilogue through LDS (free after the last barrier).
    float *ssa = reinterpret_cast<float *>(&As[0][0]), *ssb = ssa + BM;
    float2 *stage = reinterpret_cast<float2 *>(reinterpret_cast<char *>(&As[0][0]) + 2048) +
                    wave * (16 * WTN);
    static_assert((BM + BN) * 4 <= 2048 && 2048 + NW * 16 * WTN * 8 <= sizeof(As),
                  "epilogue LDS layout");
    {
        const float *sa = a.scale ? a.scale + (a_bcast ? 0 : bz) * a.rows_p : nullptr;
        const float *sb = b.scale ? b.scale + (b_bcast ? 0 : bz) * b.rows_p : nullptr;
        for (int i = tid; i < BM + BN; i += 64 * NW) {
            if (i < BM)
                ssa[i] = (sa && m0 + i < M) ? sa[m0 + i] : 1.f;
            else
                ssb[i - BM] = (sb && n0 + i - BM < N) ? sb[n0 + i - BM] : 1.f;
        }
    }
    __syncthreads();
    float2 *C = Cout + bz * sC;
    // Each lane stores two neighbouring columns with one 16-byte store (5-9%
    // faster than 8-byte stores on the output-heavy short-K products) when
    // the host found C 16-byte aligned with even row and batch strides.
    constexpr int PAIRS = 16 * WTN / 2 / 64;
    static_assert(PAIRS * 128 == 16 * WTN && WTN % 2 == 0, "slice");
    typedef float f32x4s __attribute__((ext_vector_type(4)));
#pragma unroll
    for (int fm = 0; fm < FM; fm++) {
        // The wave's own area: its LDS operations complete in order, so the
        // reads below see these writes and the next slice's writes come
        // after these reads.
#pragma unroll
        for (int fn = 0; fn < FN; fn++)
#pragma unroll
            for (int r = 0; r < 4; r++)
                stage[(4 * q + r) * WTN + fn * 16 + j] = float2{cr[fm][fn][r], ci[fm][fn][r]};
#pragma unroll
        for (int i = 0; i < PAIRS; i++) {
            const int e = i * 64 + lane;
            const int sr = e / (WTN / 2), sc2 = 2 * (e % (WTN / 2));
            const int rl = wm * WTM + fm * 16 + sr, cl = wn * WTN + sc2;
            const int64_t row = m0 + rl, col = n0 + cl;
            if (row >= M) continue;
            // Power-of-two factors, applied one at a time so that their
            // product cannot leave the float range early.
            const float fa = ssa[rl];
            const float2 v0 = stage[sr * WTN + sc2], v1 = stage[sr * WTN + sc2 + 1];
            if (vec_store && col + 1 < N) {
                const float fb0 = ssb[cl], fb1 = ssb[cl + 1];
                *reinterpret_cast<f32x4s *>(C + row * ldc + col) =
                    f32x4s{v0.x * fa * fb0, v0.y * fa * fb0, v1.x * fa * fb1, v1.y * fa * fb1};
            }
            else {
                if (col < N) C[row * ldc + col] = float2{v0.x * fa * ssb[cl], v0.y * fa * ssb[cl]};
                if (col + 1 < N)
                    C[row * ldc + col + 1] =
                        float2{v1.x * fa * ssb[cl + 1], v1.y * fa * ssb[cl + 1]};
            }
        }
    }
}

// Persistent form of the same product for GEMMs with short k loops and many
// output tiles (the skinny M or N = 128-256 steps of sliced circuits): a
// grid of one work-group per CU walks the output tiles (t, t + gridDim.x,
// ...; gridDim.x a multiple of the XCD count keeps every tile of a
// work-group on its XCD).  After a tile's last k step the LDS buffers are
// free: the work-group issues the first k tile of its next output tile into
// LDS, then writes the finished tile straight from the accumulators (each
// 16-lane group stores a 128-byte row segment), so the stores drain while
// the next tile's loads are in flight instead of one after the other.  k
// loop as in cgemm_x4m_wide_kernel (next k tile issued in two halves over
// the first two A fragments).
template <class E, int BM, int BN, int WM, int WN>
__global__ __launch_bounds__(64 * WM * WN) void cgemm_x4m_persist_kernel(
    SplitOperand a, SplitOperand b, float2 *__restrict__ Cout, int64_t M,
    int64_t N, int64_t ldc, int64_t sC, int a_bcast, int b_bcast, int64_t batch)
{
    using vec = typename Mfma16<E>::vec;
    typedef __attribute__((address_space(3))) void lds_void;
    constexpr int NW = WM * WN;
    constexpr int KT = 32;
    constexpr int WTM = BM / WM, WTN = BN / WN;
    constexpr int FM = WTM / 16, FN = WTN / 16;
    constexpr int A_IN = BM * 16 / (64 * NW);
    constexpr int B_IN = BN * 16 / (64 * NW);
    constexpr int L = A_IN + B_IN;
    static_assert(A_IN * 64 * NW == BM * 16 && B_IN * 64 * NW == BN * 16, "tile");
    static_assert(FM >= 2, "loads are issued over the first two A fragments");
    __shared__ vec As[2][BM * 16];
    __shared__ vec Bs[2][BN * 16];

    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6); // uniform: scalar
    const int wm = wave / WN, wn = wave % WN;
    const int64_t tiles_m = (M + BM - 1) / BM, tiles_n = (N + BN - 1) / BN;
    const int64_t tiles = tiles_m * tiles_n, total = tiles * batch;
    const int64_t nkt_a = a.k_p / KT, nkt_b = b.k_p / KT;
    const int nk = int(a.k_p / KT);
    int64_t t = blockIdx.x;
    if (t >= total) return;

    // The tile being loaded: coordinates, operand bases, per-lane offsets.
    int64_t m0 = 0, n0 = 0, bz = 0;
    const char *At = nullptr, *Bt = nullptr;
    uint32_t off[L];
    auto locate = [&](int64_t tile) {
        bz = tile / tiles;
        int64_t tm, tn;
        tile_coords(tile % tiles, tiles_m, tiles_n, tm, tn);
        m0 = tm * BM;
        n0 = tn * BN;
        const E *Ab = static_cast<const E *>(a.base) + (a_bcast ? 0 : bz * a.rows_p * a.k_p * 4);
        const E *Bb = static_cast<const E *>(b.base) + (b_bcast ? 0 : bz * b.rows_p * b.k_p * 4);
        At = reinterpret_cast<const char *>(Ab + m0 * nkt_a * 128);
        Bt = reinterpret_cast<const char *>(Bb + n0 * nkt_b * 128);
#pragma unroll
        for (int i = 0; i < L; i++) {
            const bool isa = i < A_IN;
            const int p = (wave * (isa ? A_IN : B_IN) + (isa ? i : i - A_IN)) * 64 + lane;
            off[i] = isa ? x4m_chunk_offset(p, m0, a.rows_p, nkt_a)
                         : x4m_chunk_offset(p, n0, b.rows_p, nkt_b);
        }
    };
    auto issue = [&](int lo, int hi, int kt, int s) {
        __builtin_amdgcn_sched_barrier(0);
        const uint32_t kb = uint32_t(kt) * 256u;
#pragma unroll
        for (int i = 0; i < L; i++) {
            if (i < lo || i >= hi) continue;
            if (i < A_IN)
                __builtin_amdgcn_global_load_lds(
                    (const void *)(At + uint64_t(off[i] + kb)),
                    (lds_void *)(&As[s][(wave * A_IN + i) * 64]), 16, 0, 0);
            else
                __builtin_amdgcn_global_load_lds(
                    (const void *)(Bt + uint64_t(off[i] + kb)),
                    (lds_void *)(&Bs[s][(wave * B_IN + i - A_IN) * 64]), 16, 0, 0);
        }
        __builtin_amdgcn_sched_barrier(0);
    };

    f32x4 cr[FM][FN], ci[FM][FN];
    const int q = lane >> 4, j = lane & 15;
    auto compute = [&](int s, int kt2, int s2) {
        const vec *as = As[s], *bs = Bs[s];
        vec brh[FN], brl[FN], bih[FN], bil[FN];
#pragma unroll
        for (int fn = 0; fn < FN; fn++) {
            const int r = wn * WTN + fn * 16 + j;
            brh[fn] = bs[r * 16 + ((0 * 4 + q) ^ j)];
            brl[fn] = bs[r * 16 + ((1 * 4 + q) ^ j)];
            bih[fn] = bs[r * 16 + ((2 * 4 + q) ^ j)];
            bil[fn] = bs[r * 16 + ((3 * 4 + q) ^ j)];
        }
#pragma unroll
        for (int fm = 0; fm < FM; fm++) {
            constexpr int H = (L + 1) / 2;
            if (fm < 2 && kt2 >= 0) issue(fm * H, fm * H + H, kt2, s2);
            const int r = wm * WTM + fm * 16 + j;
            const vec arh = as[r * 16 + ((0 * 4 + q) ^ j)];
            const vec arl = as[r * 16 + ((1 * 4 + q) ^ j)];
            const vec aih = as[r * 16 + ((2 * 4 + q) ^ j)];
            const vec ail = as[r * 16 + ((3 * 4 + q) ^ j)];
            const vec naih = neg16x8(aih), nail = neg16x8(ail);
#pragma unroll
            for (int fn = 0; fn < FN; fn++) {
                f32x4 x = cr[fm][fn], y = ci[fm][fn];
                x = Mfma16<E>::mma(arh, brh[fn], x);
                y = Mfma16<E>::mma(arh, bih[fn], y);
                x = Mfma16<E>::mma(arh, brl[fn], x);
                y = Mfma16<E>::mma(arh, bil[fn], y);
                x = Mfma16<E>::mma(arl, brh[fn], x);
                y = Mfma16<E>::mma(arl, bih[fn], y);
                x = Mfma16<E>::mma(naih, bih[fn], x);
                y = Mfma16<E>::mma(aih, brh[fn], y);
                x = Mfma16<E>::mma(naih, bil[fn], x);
                y = Mfma16<E>::mma(aih, brl[fn], y);
                x = Mfma16<E>::mma(nail, bih[fn], x);
                y = Mfma16<E>::mma(ail, brh[fn], y);
                cr[fm][fn] = x;
                ci[fm][fn] = y;
            }
        }
    };

    constexpr int kWaitLoads = 0x0f70; // vmcnt(0)
    locate(t);
    issue(0, L, 0, 0);
    for (;;) {
#pragma unroll
        for (int i = 0; i < FM; i++)
#pragma unroll
            for (int jj = 0; jj < FN; jj++) {
                cr[i][jj] = f32x4{0, 0, 0, 0};
                ci[i][jj] = f32x4{0, 0, 0, 0};
            }
        __builtin_amdgcn_s_waitcnt(kWaitLoads);
        __syncthreads();
        for (int kt = 0; kt < nk; kt += 2) {
            compute(0, kt + 1 < nk ? kt + 1 : -1, 1);
            __builtin_amdgcn_s_waitcnt(kWaitLoads);
            __syncthreads();
            if (kt + 1 >= nk) break;
            compute(1, kt + 2 < nk ? kt + 2 : -1, 0);
            __builtin_amdgcn_s_waitcnt(kWaitLoads);
            __syncthreads();
        }
        // Every wave is past its last LDS read of this tile.  The row and
        // column scales are read first (their loads then complete before the
        // next tile's are issued, so waiting for them does not wait for
        // those), then the next tile's first k step is issued, then this
        // tile is stored; full tiles store without bounds checks.
        const int64_t cm0 = m0, cn0 = n0, cbz = bz;
        const bool full = cm0 + BM <= M && cn0 + BN <= N;
        const float *sa = a.scale ? a.scale + (a_bcast ? 0 : cbz) * a.rows_p : nullptr;
        const float *sb = b.scale ? b.scale + (b_bcast ? 0 : cbz) * b.rows_p : nullptr;
        float fa[FM][4], fb[FN];
#pragma unroll
        for (int fm = 0; fm < FM; fm++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int64_t row = cm0 + wm * WTM + fm * 16 + 4 * q + r;
                fa[fm][r] = (sa && row < M) ? sa[row] : 1.f;
            }
#pragma unroll
        for (int fn = 0; fn < FN; fn++) {
            const int64_t col = cn0 + wn * WTN + fn * 16 + j;
            fb[fn] = (sb && col < N) ? sb[col] : 1.f;
        }
        __builtin_amdgcn_s_waitcnt(kWaitLoads);
        const int64_t next = t + gridDim.x;
        const bool more = next < total;
        if (more) {
            locate(next);
            issue(0, L, 0, 0);
        }
        float2 *C = Cout + cbz * sC;
        // Power-of-two factors, applied one at a time so that their product
        // cannot leave the float range.
#pragma unroll
        for (int fm = 0; fm < FM; fm++)
#pragma unroll
            for (int fn = 0; fn < FN; fn++) {
                const int64_t col = cn0 + wn * WTN + fn * 16 + j;
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    const int64_t row = cm0 + wm * WTM + fm * 16 + 4 * q + r;
                    if (full || (row < M && col < N))
                        C[row * ldc + col] = float2{cr[fm][fn][r] * fa[fm][r] * fb[fn],
                                                    ci[fm][fn][r] * fa[fm][r] * fb[fn]};
                }
            }
        if (!more) break;
        t = next;
    }
}

double gemm_flops(const GemmArgs &g)
{
    return 8.0 * double(g.batch) * double(g.M) * double(g.N) * double(g.K);
}

// Minimum HBM traffic: each operand read once, C written once (a broadcast
// operand, batch stride 0, counts once).
double gemm_bytes(const GemmArgs &g, int elem_bytes)
{
    const double a = double(g.M) * double(g.K) * (g.sA ? double(g.batch) : 1.0);
    const double b = double(g.N) * double(g.K) * (g.sB ? double(g.batch) : 1.0);
    const double c = double(g.M) * double(g.N) * double(g.batch);
    return elem_bytes * (a + b + c);
}

// Split-K dot products for few outputs and long k: block (o, chunk) sums
// output o over one k chunk in FP64 and stores the partial; a second kernel
// adds the partials of each output in FP64.  Consecutive blocks share the k
// chunk, so operand rows reused by several outputs come from L2.
template <class R>
__global__ __launch_bounds__(256) void gemm_splitk_partial_kernel(GemmArgs g, double2 *partials,
                                                                  int64_t chunk, int64_t nchunks)
{
    using C2 = typename Vec<R>::c;
    __shared__ double red[2][256];
    const int64_t o = blockIdx.x;
    const int64_t c = blockIdx.y;
    const int64_t n = o % g.N;
    const int64_t rest = o / g.N;
    const int64_t m = rest % g.M;
    const int64_t b = rest / g.M;
    const C2 *a = static_cast<const C2 *>(g.A) + b * g.sA + m * g.lda;
    const C2 *bb = static_cast<const C2 *>(g.B) + b * g.sB + n * g.ldb;
    const int64_t k0 = c * chunk;
    const int64_t k1 = (k0 + chunk < g.K) ? k0 + chunk : g.K;
    double sr = 0, si = 0;
    for (int64_t k = k0 + threadIdx.x; k < k1; k += 256) {
        const C2 x = a[k], y = bb[k];
        sr = fma(double(x.x), double(y.x), sr);
        sr = fma(-double(x.y), double(y.y), sr);
        si = fma(double(x.x), double(y.y), si);
        si = fma(double(x.y), double(y.x), si);
    }
    red[0][threadIdx.x] = sr;
    red[1][threadIdx.x] = si;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (int(threadIdx.x) < s) {
            red[0][threadIdx.x] += red[0][threadIdx.x + s];
            red[1][threadIdx.x] += red[1][threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) partials[o * nchunks + c] = double2{red[0][0], red[1][0]};
}

template <class R>
__global__ __launch_bounds__(256) void gemm_splitk_final_kernel(GemmArgs g, const double2 *partials,
                                                                int64_t nchunks)
{
    using C2 = typename Vec<R>::c;
    __shared__ double red[2][256];
    const int64_t o = blockIdx.x;
    double sr = 0, si = 0;
    for (int64_t c = threadIdx.x; c < nchunks; c += 256) {
        const double2 v = partials[o * nchunks + c];
        sr += v.x;
        si += v.y;
    }
    red[0][threadIdx.x] = sr;
    red[1][threadIdx.x] = si;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (int(threadIdx.x) < s) {
            red[0][threadIdx.x] += red[0][threadIdx.x + s];
            red[1][threadIdx.x] += red[1][threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const int64_t n = o % g.N;
        const int64_t rest = o / g.N;
        const int64_t m = rest % g.M;
        const int64_t b = rest / g.M;
        static_cast<C2 *>(g.C)[b * g.sC + m * g.ldc + n] = C2{R(red[0][0]), R(red[1][0])};
    }
}

template <class R>
void launch_splitk(const GemmArgs &g, void *partials, hipStream_t stream)
{
    const int64_t outputs = g.batch * g.M * g.N;
    const int64_t chunk = splitk_chunk(g.K);
    const int64_t nchunks = ceil_div(g.K, chunk);
    JH_REQUIRE(partials != nullptr, "Split-K GEMM needs its partial-sum buffer.");
    JH_REQUIRE(outputs <= 0x7fffffff && nchunks <= 65535, "Split-K grid too large.");
    prof::Scope ps(prof::kGemmSplitK, stream, gemm_flops(g), gemm_bytes(g, int(sizeof(R) * 2)));
    hipLaunchKernelGGL(gemm_splitk_partial_kernel<R>, dim3((unsigned)outputs, (unsigned)nchunks),
                       dim3(256), 0, stream, g, static_cast<double2 *>(partials), chunk, nchunks);
    JH_HIP(hipGetLastError());
    hipLaunchKernelGGL(gemm_splitk_final_kernel<R>, dim3((unsigned)outputs), dim3(256), 0, stream,
                       g, static_cast<const double2 *>(partials), nchunks);
    JH_HIP(hipGetLastError());
}

template <class R>
void launch_valu(int kernel, const GemmArgs &g, hipStream_t stream)
{
    const int64_t total = g.batch * g.M * g.N;
    prof::Scope ps(kernel == kGemmSmallRed ? prof::kGemmSmallRed : prof::kGemmSmall, stream,
                   gemm_flops(g), gemm_bytes(g, int(sizeof(R) * 2)));
    if (kernel == kGemmSmallRed) {
        JH_REQUIRE(total <= (int64_t(1) << 31) - 1, "Too many outputs for the reduction kernel.");
        hipLaunchKernelGGL(cgemm_smallred_kernel<R>, dim3((unsigned)total), dim3(256), 0,
                           stream, g);
    }
    else if (g.variant == 2 ||
             (g.variant == 0 && g.K <= kShortK && g.N >= 64 && total >= (int64_t(1) << 16))) {
        JH_REQUIRE(g.K <= kShortK, "The short-K column kernel needs K <= 16.");
        const int64_t m_tiles = ceil_div(g.M, kShortRows), n_tiles = ceil_div(g.N, 256);
        const int64_t blocks = g.batch * m_tiles * n_tiles;
        JH_REQUIRE(blocks <= 0x7fffffff, "Too many blocks.");
        hipLaunchKernelGGL(cgemm_shortk_cols_kernel<R>, dim3((unsigned)blocks), dim3(256), 0,
                           stream, g, m_tiles, n_tiles);
    }
    else if (g.variant == 3 || (g.variant == 0 && g.K <= kShortK && g.M >= 1024 && g.N < 64)) {
        JH_REQUIRE(g.K <= kShortK, "The short-K row kernel needs K <= 16.");
        const int64_t m_tiles = ceil_div(g.M, 256);
        const int64_t blocks = g.batch * m_tiles;
        JH_REQUIRE(blocks <= 0x7fffffff, "Too many blocks.");
        hipLaunchKernelGGL(cgemm_shortk_rows_kernel<R>, dim3((unsigned)blocks), dim3(256), 0,
                           stream, g, m_tiles);
    }
    else {
        const int grid = (int)std::min<int64_t>((total + 255) / 256, 65536);
        hipLaunchKernelGGL(cgemm_small_kernel<R>, dim3(grid), dim3(256), 0, stream, g);
    }
    JH_HIP(hipGetLastError());
}

} // namespace

void launch_gemm(int kernel, int elem_bytes, const GemmArgs &g,
                 hipStream_t stream, void *partials)
{
    if (g.M == 0 || g.N == 0 || g.batch == 0) return;
    if (kernel == kGemmSplitK) {
        if (elem_bytes == 8)
            launch_splitk<float>(g, partials, stream);
        else
            launch_splitk<double>(g, partials, stream);
        return;
    }
    if (kernel == kGemmSmall || kernel == kGemmSmallRed) {
        if (elem_bytes == 8)
            launch_valu<float>(kernel, g, stream);
        else
            launch_valu<double>(kernel, g, stream);
        return;
    }
    JH_REQUIRE(g.batch <= 65535, "Batch count exceeds 65535.");
    if (kernel == kGemmF32Mfma) {
        JH_REQUIRE(elem_bytes == 8, "FP32 MFMA kernel needs complex64 data.");
        // Variant 0: 128x128 tiles; 128x64 or 64x128 when N or M is at most
        // 64 (half of a 128-wide tile would be padding).
        int BM = (g.M <= 64 && g.N > 64) ? 64 : 128, BN = g.N <= 64 ? 64 : 128;
        if (g.variant == 1) BM = 128, BN = 128;
        if (g.variant == 2) BM = 64, BN = 128;
        if (g.variant == 3) BM = 128, BN = 64;
        JH_REQUIRE(ceil_div(g.M, BM) * ceil_div(g.N, BN) <= 0x7fffffff, "Too many tiles.");
        dim3 grid((unsigned)(ceil_div(g.M, BM) * ceil_div(g.N, BN)), 1, (unsigned)g.batch);
        prof::Scope ps(prof::kGemmF32, stream, gemm_flops(g), gemm_bytes(g, 8));
        if (BN == 64)
            hipLaunchKernelGGL((cgemm_embed_kernel<float, 128, 64, 16, 2, 2>), grid, dim3(256), 0,
                               stream, g);
        else if (BM == 64)
            hipLaunchKernelGGL((cgemm_embed_kernel<float, 64, 128, 16, 2, 2>), grid, dim3(256), 0,
                               stream, g);
        else
            hipLaunchKernelGGL((cgemm_embed_kernel<float, 128, 128, 16, 2, 2>), grid, dim3(256),
                               0, stream, g);
    }
    else if (kernel == kGemmF64Mfma) {
        JH_REQUIRE(elem_bytes == 16, "FP64 MFMA kernel needs complex128 data.");
        constexpr int BM = 128, BN = 64;
        JH_REQUIRE(ceil_div(g.M, BM) * ceil_div(g.N, BN) <= 0x7fffffff, "Too many tiles.");
        dim3 grid((unsigned)(ceil_div(g.M, BM) * ceil_div(g.N, BN)), 1, (unsigned)g.batch);
        prof::Scope ps(prof::kGemmF64, stream, gemm_flops(g), gemm_bytes(g, 16));
        hipLaunchKernelGGL((cgemm_embed_kernel<double, BM, BN, 8, 2, 2>), grid, dim3(256), 0,
                           stream, g);
    }
    else {
        JH_FAIL(JET_HIP_ERROR_INTERNAL, "launch_gemm: unknown kernel.");
    }
    JH_HIP(hipGetLastError());
}

void launch_split_bf16(const void *src, int64_t rows, int64_t k, int64_t ld,
                       int64_t s, int64_t batch, const SplitOperand &dst,
                       hipStream_t stream)
{
    (void)batch;
    const int64_t total = dst.batch * dst.rows_p * (dst.k_p / 8);
    const int grid = (int)std::min<int64_t>((total + 255) / 256, 65536);
    // reads 8 bytes and writes 6 bf16 per (padded) element
    prof::Scope ps(prof::kSplitBf16, stream, 0.0,
                   8.0 * double(batch) * double(rows) * double(k) +
                       2.0 * dst.planes * double(dst.batch) * double(dst.rows_p) * double(dst.k_p));
    hipLaunchKernelGGL(split_bf16_kernel, dim3(grid), dim3(256), 0, stream,
                       static_cast<const float2 *>(src), rows, k, ld, s, dst);
    JH_HIP(hipGetLastError());
}

void launch_split_f16(const void *src, int64_t rows, int64_t k, int64_t ld,
                      int64_t s, int64_t batch, const SplitOperand &dst,
                      hipStream_t stream)
{
    JH_REQUIRE(dst.planes == 4 && dst.interleaved && dst.rowmax && dst.scale,
               "Split-FP16 operands need four interleaved planes and row scales.");
    const int64_t nrow = dst.batch * dst.rows_p;
    const int64_t total = nrow * (dst.k_p / 8);
    const int64_t units = batch * rows * ceil_div(k, int64_t(kRowMaxChunk));
    if (dst.k_p <= kSplitOnePassK) {
        // reads the operand once, writes 4 fp16 per (padded) element
        prof::Scope ps(prof::kSplitF16, stream, 0.0,
                       8.0 * double(batch) * double(rows) * double(k) +
                           8.0 * double(nrow) * double(dst.k_p));
        int lg = 2;
        while ((1 << lg) < 64 && (int64_t(1) << lg) * 8 < dst.k_p) lg++;
        const int64_t rows_per_block = 4 * (64 >> lg);
        const int vec = (reinterpret_cast<uintptr_t>(src) % 16 == 0 && ld % 2 == 0 &&
                         s % 2 == 0 && k % 8 == 0) ? 1 : 0;
        JH_REQUIRE(ceil_div(nrow, rows_per_block) <= 0x7fffffff, "Too many rows.");
        hipLaunchKernelGGL(split_f16_rows_kernel, dim3((unsigned)ceil_div(nrow, rows_per_block)),
                           dim3(256), 0, stream, static_cast<const float2 *>(src), rows, k, ld, s,
                           dst, lg, vec);
        JH_HIP(hipGetLastError());
        return;
    }
    if (dst.k_p <= kSplitBlockRowK) {
        prof::Scope ps(prof::kSplitF16, stream, 0.0,
                       8.0 * double(batch) * double(rows) * double(k) +
                           8.0 * double(nrow) * double(dst.k_p));
        const int vec = (reinterpret_cast<uintptr_t>(src) % 16 == 0 && ld % 2 == 0 &&
                         s % 2 == 0 && k % 8 == 0) ? 1 : 0;
        JH_REQUIRE(nrow <= 0x7fffffff, "Too many rows.");
        if (dst.k_p <= 8192)
            hipLaunchKernelGGL(split_f16_block_rows_kernel<4>, dim3((unsigned)nrow), dim3(256), 0,
                               stream, static_cast<const float2 *>(src), rows, k, ld, s, dst, vec);
        else
            hipLaunchKernelGGL(split_f16_block_rows_kernel<8>, dim3((unsigned)nrow), dim3(256), 0,
                               stream, static_cast<const float2 *>(src), rows, k, ld, s, dst, vec);
        JH_HIP(hipGetLastError());
        return;
    }
    // reads the operand twice (row maxima, then the split) and writes 4 fp16
    // per (padded) element
    prof::Scope ps(prof::kSplitF16, stream, 0.0,
                   16.0 * double(batch) * double(rows) * double(k) +
                       8.0 * double(nrow) * double(dst.k_p));
    hipLaunchKernelGGL(row_max_clear_kernel, dim3((unsigned)std::min<int64_t>(ceil_div(nrow, 256), 4096)),
                       dim3(256), 0, stream, dst.rowmax, nrow);
    if (units > 0)
        hipLaunchKernelGGL(row_absmax_kernel, dim3((unsigned)std::min<int64_t>(units, 65536)),
                           dim3(256), 0, stream, static_cast<const float2 *>(src), rows, k, ld, s,
                           batch, dst.rows_p, dst.rowmax);
    hipLaunchKernelGGL(split_f16_kernel, dim3((unsigned)std::min<int64_t>(ceil_div(total, 256), 65536)),
                       dim3(256), 0, stream, static_cast<const float2 *>(src), rows, k, ld, s, dst);
    JH_HIP(hipGetLastError());
}

void launch_split_f16_gather(const void *src, const SplitGather &g, const SplitOperand &dst,
                             hipStream_t stream)
{
    JH_REQUIRE(dst.planes == 4 && dst.interleaved && dst.scale &&
                   dst.k_p == (int64_t(1) << g.log2k) && g.tb <= kGatherTileBits &&
                   g.tb >= 8 && g.xr >= 1 && g.xr <= 7 && g.log2k >= 5 &&
                   dst.batch * dst.rows_p == (int64_t(1) << (g.xr + g.nw)),
               "Gathered split operand does not match its plan.");
    const int64_t n = int64_t(1) << (g.tb + g.nw);
    // reads the source once, writes 4 fp16 per element
    prof::Scope ps(prof::kSplitF16, stream, 0.0, 16.0 * double(n));
    JH_REQUIRE(g.su[0] == 0, "Gathered split tiles must hold source bit 0.");
    const int64_t ntiles = int64_t(1) << g.nw;
    const int vec = reinterpret_cast<uintptr_t>(src) % 16 == 0 ? 1 : 0;
    hipLaunchKernelGGL(split_f16_gather_kernel, dim3((unsigned)std::min<int64_t>(ntiles, 65536)),
                       dim3(256), 0, stream, static_cast<const float2 *>(src), g, dst, vec);
    JH_HIP(hipGetLastError());
}

// 4M kernels: 128x192 block tiles (cgemm_x4m_wide_kernel) where N fills
// the last column tile to within 5% (N >= 4096), else 128x128; 8 waves (two
// per SIMD) in a 2x4 grid either way.  Variants: 1 128x128, 2 128x192,
// 3 persistent 128x128, 4 192x128 (the wide kernel with its long side along
// M, for products with long M and N of 128-256).
constexpr int kX4mWideN = 4096;

template <class E>
void launch_x4m(const SplitOperand &a, const SplitOperand &b, void *C, int64_t M, int64_t N,
                int64_t ldc, int64_t sC, int64_t batch, bool a_bcast, bool b_bcast,
                hipStream_t stream, int variant)
{
    constexpr int BM = 128, WM = 2, WN = 4;
    JH_REQUIRE(a.k_p <= kX4mMaxK, "k exceeds the 4M split kernels' limit (2^21).");
    if (variant == 3) {
        // Persistent 128x128 tiles: one work-group per CU, the grid a
        // multiple of the XCD count.  (128x192 tiles with the register
        // epilogue spill 106 VGPRs, so the persistent form stays at 128x128.)
        constexpr int BN = 128;
        const int64_t total = ceil_div(M, BM) * ceil_div(N, BN) * batch;
        int device = 0, cus = 0;
        JH_HIP(hipGetDevice(&device));
        JH_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device));
        int64_t grid = std::max(kXcds, cus - cus % kXcds);
        grid = std::min<int64_t>(grid, total);
        hipLaunchKernelGGL((cgemm_x4m_persist_kernel<E, BM, BN, WM, WN>), dim3((unsigned)grid),
                           dim3(64 * WM * WN), 0, stream, a, b, static_cast<float2 *>(C), M, N,
                           ldc, sC, a_bcast ? 1 : 0, b_bcast ? 1 : 0, batch);
        JH_HIP(hipGetLastError());
        return;
    }
    const int vec_store = reinterpret_cast<uintptr_t>(C) % 16 == 0 && ldc % 2 == 0 &&
                          (batch == 1 || sC % 2 == 0);
    if (variant == 4) {
        constexpr int TM = 192, TN = 128;
        JH_REQUIRE(ceil_div(M, TM) * ceil_div(N, TN) <= 0x7fffffff, "Too many tiles.");
        dim3 grid((unsigned)(ceil_div(M, TM) * ceil_div(N, TN)), 1, (unsigned)batch);
        hipLaunchKernelGGL((cgemm_x4m_wide_kernel<E, TM, TN, WM, WN>), grid, dim3(64 * WM * WN),
                           0, stream, a, b, static_cast<float2 *>(C), M, N, ldc, sC,
                           a_bcast ? 1 : 0, b_bcast ? 1 : 0, vec_store);
        JH_HIP(hipGetLastError());
        return;
    }
    const int BN = variant == 1 ? 128 : variant == 2 ? 192 : N >= kX4mWideN ? 192 : 128;
    JH_REQUIRE(ceil_div(M, BM) * ceil_div(N, BN) <= 0x7fffffff, "Too many tiles.");
    dim3 grid((unsigned)(ceil_div(M, BM) * ceil_div(N, BN)), 1, (unsigned)batch);
    if (BN == 192) {
        hipLaunchKernelGGL((cgemm_x4m_wide_kernel<E, BM, 192, WM, WN>), grid, dim3(64 * WM * WN),
                           0, stream, a, b, static_cast<float2 *>(C), M, N, ldc, sC,
                           a_bcast ? 1 : 0, b_bcast ? 1 : 0, vec_store);
    }
    else
        hipLaunchKernelGGL((cgemm_x4m_kernel<E, BM, 128, WM, WN>), grid, dim3(64 * WM * WN), 0,
                           stream, a, b, static_cast<float2 *>(C), M, N, ldc, sC,
                           a_bcast ? 1 : 0, b_bcast ? 1 : 0);
    JH_HIP(hipGetLastError());
}

void launch_gemm_f16x4m(const SplitOperand &a, const SplitOperand &b, int64_t K_true,
                        void *C, int64_t M, int64_t N, int64_t ldc,
                        int64_t sC, int64_t batch, bool a_bcast, bool b_bcast,
                        hipStream_t stream, int variant)
{
    JH_REQUIRE(a.k_p == b.k_p, "Split operands disagree on k.");
    JH_REQUIRE(a.planes == 4 && b.planes == 4 && a.interleaved && b.interleaved && a.scale &&
                   b.scale,
               "The split-FP16 kernel needs four interleaved planes and row scales.");
    JH_REQUIRE(batch <= 65535, "Batch count exceeds 65535.");
    const double K = double(K_true);
    prof::Scope ps(prof::kGemmF16x4m, stream, 8.0 * double(batch) * double(M) * double(N) * K,
                   8.0 * double(batch) * (double(M) * K + double(N) * K + double(M) * double(N)));
    launch_x4m<_Float16>(a, b, C, M, N, ldc, sC, batch, a_bcast, b_bcast, stream, variant);
}

void launch_gemm_bf16x4m(const SplitOperand &a, const SplitOperand &b, int64_t K_true,
                         void *C, int64_t M, int64_t N, int64_t ldc,
                         int64_t sC, int64_t batch, bool a_bcast, bool b_bcast,
                         hipStream_t stream, int variant)
{
    JH_REQUIRE(a.k_p == b.k_p, "Split operands disagree on k.");
    JH_REQUIRE(a.planes == 4 && b.planes == 4 && a.interleaved && b.interleaved,
               "The 4M kernel needs four interleaved planes.");
    JH_REQUIRE(batch <= 65535, "Batch count exceeds 65535.");
    // Algorithmic work counts the unpadded K.
    const double K = double(K_true);
    prof::Scope ps(prof::kGemmBf16x4m, stream, 8.0 * double(batch) * double(M) * double(N) * K,
                   8.0 * double(batch) * (double(M) * K + double(N) * K + double(M) * double(N)));
    launch_x4m<__bf16>(a, b, C, M, N, ldc, sC, batch, a_bcast, b_bcast, stream, variant);
}

void launch_fused_small(const FusedSmall &f, int elem_bytes, const void *a, const void *b,
                        void *c, hipStream_t stream)
{
    JH_REQUIRE(f.lk <= kFusedMaxLog2K && f.lb + f.lm + f.ln <= 31, "Fused step out of range.");
    const int64_t total = int64_t(1) << (f.lb + f.lm + f.ln);
    const double K = double(int64_t(1) << f.lk);
    prof::Scope ps(prof::kGemmSmall, stream, 8.0 * double(total) * K,
                   double(elem_bytes) * double(total));
    const size_t lds = size_t(2) * sizeof(int32_t) << f.lk;
    if (f.red) {
        // Eight work-groups per CU at most: each builds its k offset tables
        // once and walks many outputs (one work-group per output rebuilt the
        // tables for every output, about half of the time of a many-output
        // step).
        const unsigned grid = (unsigned)std::min<int64_t>(total, 2048);
        if (elem_bytes == 8)
            hipLaunchKernelGGL(fused_small_red_kernel<float>, dim3(grid), dim3(256), lds, stream,
                               a, b, c, f);
        else
            hipLaunchKernelGGL(fused_small_red_kernel<double>, dim3(grid), dim3(256), lds,
                               stream, a, b, c, f);
        JH_HIP(hipGetLastError());
        return;
    }
    const unsigned grid = (unsigned)std::min<int64_t>(ceil_div(total, int64_t(256)), 4096);
    if (elem_bytes == 8)
        hipLaunchKernelGGL(fused_small_kernel<float>, dim3(grid), dim3(256), lds, stream, a, b,
                           c, f);
    else
        hipLaunchKernelGGL(fused_small_kernel<double>, dim3(grid), dim3(256), lds, stream, a,
                           b, c, f);
    JH_HIP(hipGetLastError());
}

void launch_gemm_bf16x3(const SplitOperand &a, const SplitOperand &b, int64_t K_true,
                        void *C, int64_t M, int64_t N, int64_t ldc,
                        int64_t sC, int64_t batch, bool a_bcast, bool b_bcast,
                        hipStream_t stream)
{
    constexpr int BM = 128, BN = 128, WM = 2, WN = 2;
    JH_REQUIRE(a.k_p == b.k_p, "Split operands disagree on k.");
    JH_REQUIRE(a.planes == 6 && b.planes == 6 && !a.interleaved && !b.interleaved,
               "The 3M kernel needs six planar planes.");
    JH_REQUIRE(batch <= 65535, "Batch count exceeds 65535.");
    JH_REQUIRE(ceil_div(M, BM) * ceil_div(N, BN) <= 0x7fffffff, "Too many tiles.");
    dim3 grid((unsigned)(ceil_div(M, BM) * ceil_div(N, BN)), 1, (unsigned)batch);
    const double K = double(K_true);
    prof::Scope ps(prof::kGemmBf16x3, stream, 8.0 * double(batch) * double(M) * double(N) * K,
                   8.0 * double(batch) * (double(M) * K + double(N) * K + double(M) * double(N)));
    hipLaunchKernelGGL((cgemm_bf16x3_kernel<BM, BN, WM, WN>), grid, dim3(64 * WM * WN), 0, stream,
                       a, b, static_cast<float2 *>(C), M, N, ldc, sC, a_bcast ? 1 : 0,
                       b_bcast ? 1 : 0);
    JH_HIP(hipGetLastError());
}

} // namespace jet_hip

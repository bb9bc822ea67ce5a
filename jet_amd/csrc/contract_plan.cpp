// contract_plan.cpp -- host-side planning of one pairwise contraction.
#include "contract_plan.hpp"
#include "tuning.hpp"

#include <algorithm>
#include <vector>

namespace jet_hip {

namespace {

int find(const std::vector<int32_t> &v, int32_t x)
{
    for (size_t i = 0; i < v.size(); i++)
        if (v[i] == x) return (int)i;
    return -1;
}

size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

// Permutation plan taking `from` (with extents) to `to`; the returned plan is
// kCopy when the two orders give the same memory layout.
PermutePlan plan_reorder(const std::vector<int32_t> &from,
                         const std::vector<int64_t> &ext,
                         const std::vector<int32_t> &to)
{
    std::vector<int32_t> perm(to.size());
    for (size_t i = 0; i < to.size(); i++) perm[i] = find(from, to[i]);
    return make_permute_plan((int)from.size(), ext.data(), perm.data());
}

} // namespace

bool plan_split_gather(const std::vector<int32_t> &src, const std::vector<int64_t> &ext,
                       const std::vector<int32_t> &target, int64_t K, SplitGather &g)
{
    // log2 extents; every extent a power of two
    const int r = (int)src.size();
    std::vector<int> lg(r);
    for (int i = 0; i < r; i++) {
        if (ext[i] < 1 || (ext[i] & (ext[i] - 1))) return false;
        int b = 0;
        while ((int64_t(1) << b) < ext[i]) b++;
        lg[i] = b;
    }
    // source bit offset of each mode (row-major: the last mode is fastest)
    std::vector<int> soff(r);
    int nbits = 0;
    for (int i = r - 1; i >= 0; i--) {
        soff[i] = nbits;
        nbits += lg[i];
    }
    if (nbits > 62 || K < 32 || K > 1024 || (K & (K - 1))) return false;
    int log2k = 0;
    while ((int64_t(1) << log2k) < K) log2k++;
    // source bit of every target bit
    std::vector<int> srcbit(nbits), tgt_of_src(nbits);
    int toff = 0;
    for (int i = (int)target.size() - 1; i >= 0; i--) {
        const int m = find(src, target[i]);
        if (m < 0) return false;
        for (int q = 0; q < lg[m]; q++) srcbit[toff + q] = soff[m] + q;
        toff += lg[m];
    }
    if (toff != nbits) return false;
    for (int t = 0; t < nbits; t++) tgt_of_src[srcbit[t]] = t;
    const int nrow = nbits - log2k;
    if (nrow < 4) return false;
    // Tile row bits: those holding the lowest source bits, until the tile
    // (all K bits + these) reaches kGatherTileBits; then the lowest other
    // row bits.
    std::vector<char> chosen(nrow, 0);
    int budget = kGatherTileBits - log2k, covered = 0;
    for (int sb = 0; sb < nbits; sb++) {
        const int t = tgt_of_src[sb];
        if (t >= log2k) {
            if (budget == 0) break;
            chosen[t - log2k] = 1;
            budget--;
        }
        covered = sb + 1;
    }
    if (covered < 4) return false;
    for (int rr = 0; rr < nrow && budget > 0; rr++)
        if (!chosen[rr]) {
            chosen[rr] = 1;
            budget--;
        }
    g = SplitGather{};
    g.log2k = log2k;
    std::vector<int> vbit_of_t(nbits, -1);
    for (int t = 0; t < log2k; t++) vbit_of_t[t] = t;
    for (int rr = 0; rr < nrow; rr++)
        if (chosen[rr]) {
            g.vrow[g.xr] = (int8_t)rr;
            vbit_of_t[log2k + rr] = log2k + g.xr;
            g.xr++;
        }
    g.tb = log2k + g.xr;
    // u order: the tile's source bits ascending
    int j = 0;
    for (int sb = 0; sb < nbits; sb++) {
        const int t = tgt_of_src[sb];
        if (vbit_of_t[t] >= 0) {
            g.su[j] = (int8_t)sb;
            g.u2v[j] = (int8_t)vbit_of_t[t];
            j++;
        }
        else {
            g.wsrc[g.nw] = (int8_t)sb;
            g.wrow[g.nw] = (int8_t)(t - log2k);
            g.nw++;
        }
    }
    return j == g.tb;
}

bool plan_fused_small(const std::vector<int32_t> &a, const std::vector<int64_t> &ea,
                      const std::vector<int32_t> &b, const std::vector<int64_t> &eb,
                      const std::vector<int32_t> &c, const std::vector<int32_t> &batch,
                      const std::vector<int32_t> &left, const std::vector<int32_t> &right,
                      const std::vector<int32_t> &common, FusedSmall &f)
{
    auto lg = [](int64_t e) {
        int l = 0;
        while ((int64_t(1) << l) < e) l++;
        return (int64_t(1) << l) == e ? l : -1;
    };
    // log2 extents of every mode, and the bit offset (log2 stride) of each
    // mode in each row-major operand
    std::vector<int32_t> modes;
    std::vector<int> lext;
    auto note = [&](const std::vector<int32_t> &m, const std::vector<int64_t> &e) {
        for (size_t i = 0; i < m.size(); i++) {
            const int l = lg(e[i]);
            if (l < 0) return false;
            if (std::find(modes.begin(), modes.end(), m[i]) == modes.end()) {
                modes.push_back(m[i]);
                lext.push_back(l);
            }
        }
        return true;
    };
    if (!note(a, ea) || !note(b, eb)) return false;
    auto lext_of = [&](int32_t m) {
        return lext[std::find(modes.begin(), modes.end(), m) - modes.begin()];
    };
    auto shift_in = [&](const std::vector<int32_t> &op, int32_t m) {
        int sh = 0;
        for (size_t i = op.size(); i-- > 0;) {
            if (op[i] == m) return sh;
            sh += lext_of(op[i]);
        }
        return -1;
    };
    int total_a = 0, total_b = 0, total_c = 0;
    for (auto m : a) total_a += lext_of(m);
    for (auto m : b) total_b += lext_of(m);
    for (auto m : c) total_c += lext_of(m);
    if (total_a > 31 || total_b > 31 || total_c > 31) return false;
    // group bits (last mode fastest) -> bit positions in operand op
    auto group = [&](const std::vector<int32_t> &g, const std::vector<int32_t> &op,
                     FusedBits &fb) {
        fb.n = 0;
        for (size_t i = g.size(); i-- > 0;) {
            const int w = lext_of(g[i]), sh = shift_in(op, g[i]);
            for (int t = 0; t < w; t++) {
                if (fb.n >= kFusedMaxBits) return false;
                fb.pos[fb.n++] = int8_t(sh + t);
            }
        }
        return true;
    };
    f = FusedSmall{};
    if (!group(batch, a, f.a_b) || !group(left, a, f.a_m) || !group(common, a, f.a_k) ||
        !group(batch, b, f.b_b) || !group(right, b, f.b_n) || !group(common, b, f.b_k) ||
        !group(batch, c, f.c_b) || !group(left, c, f.c_m) || !group(right, c, f.c_n))
        return false;
    f.lb = f.a_b.n;
    f.lm = f.a_m.n;
    f.ln = f.b_n.n;
    f.lk = f.a_k.n;
    return f.lk <= kFusedMaxLog2K;
}

int gemm_variants(int kernel)
{
    switch (kernel) {
    case kGemmF32Mfma:
        return 3;
    case kGemmBf16x4m:
    case kGemmF16x4m:
        return 4;
    case kGemmSmall:
        return 3;
    default:
        return 1;
    }
}

bool gemm_choice_valid(int mode, int64_t M, int64_t N, int64_t K, int64_t batch, GemmChoice c)
{
    if (M < 1 || N < 1 || K < 1 || batch < 1) return false;
    if (c.variant < 0 || c.variant > gemm_variants(c.kernel)) return false;
    const bool c128 = mode == 2;
    const int64_t outputs = batch * M * N;
    switch (c.kernel) {
    case kGemmSmall:
        if (c.variant == 2 || c.variant == 3) return K <= 16;
        return true;
    case kGemmSmallRed:
        return outputs <= 0x7fffffff;
    case kGemmSplitK:
        return outputs <= 0x7fffffff && ceil_div(K, splitk_chunk(K)) <= 65535;
    case kGemmF32Mfma:
        return !c128 && batch <= 65535 && ceil_div(M, 64) * ceil_div(N, 64) <= 0x7fffffff;
    case kGemmF64Mfma:
        return c128 && batch <= 65535 && ceil_div(M, 128) * ceil_div(N, 64) <= 0x7fffffff;
    case kGemmBf16x3:
        return mode == 1 && batch <= 65535 && ceil_div(M, 128) * ceil_div(N, 128) <= 0x7fffffff;
    case kGemmBf16x4m:
        return mode == 3 && batch <= 65535 && ceil_div(M, 128) * ceil_div(N, 128) <= 0x7fffffff &&
               ceil_div(K, kSplitKPad) * kSplitKPad <= kX4mMaxK;
    case kGemmF16x4m:
        return mode == 4 && batch <= 65535 && ceil_div(M, 128) * ceil_div(N, 128) <= 0x7fffffff &&
               ceil_div(K, kSplitKPad) * kSplitKPad <= kX4mMaxK;
    default:
        return false;
    }
}

int choose_gemm_kernel(int mode, int64_t M, int64_t N, int64_t K,
                       int64_t batch)
{
    const bool mfma_shape = M >= 32 && N >= 16 && K >= 16;
    if (mfma_shape) {
        if (mode == 2) return kGemmF64Mfma;
        // The split modes pay where the step is bound by MFMA throughput
        // (every dimension large enough to reuse operands from LDS).  The
        // split-FP16 kernel also beats FP32 MFMA from K = 64 (1.4-2.1x on
        // 256 x 2^23 x 64 .. 16384^2 x 64), and at K = 32 on large outputs.
        if (mode == 1 && M >= 128 && N >= 128 && K >= 128) return kGemmBf16x3;
        const bool x4m_k = K <= kX4mMaxK;
        if (mode == 3 && M >= 128 && N >= 128 && K >= 128 && x4m_k) return kGemmBf16x4m;
        if (mode == 4 && M >= 128 && N >= 128 && x4m_k &&
            (K >= 64 || (K >= 32 && batch * M * N >= (int64_t(1) << 24))))
            return kGemmF16x4m;
        return kGemmF32Mfma;
    }
    if (K >= 4 * kSplitKChunk && batch * M * N <= 65536) return kGemmSplitK;
    if (K > 256 && batch * M * N <= 65536) return kGemmSmallRed;
    return kGemmSmall;
}

ContractionDesc plan_contraction(int elem_bytes, int mode, int layout,
                                 int ra, const int32_t *modes_a,
                                 const int64_t *ext_a, int rb,
                                 const int32_t *modes_b, const int64_t *ext_b,
                                 int rc, const int32_t *modes_c)
{
    JH_REQUIRE(elem_bytes == 8 || elem_bytes == 16, "Unsupported data type.");
    JH_REQUIRE(ra >= 0 && rb >= 0 && rc >= 0 && ra <= kMaxRank && rb <= kMaxRank &&
                   rc <= kMaxRank,
               "Tensor rank must be between 0 and 64.");
    JH_REQUIRE(mode >= 0 && mode <= 4, "Unknown precision mode.");
    JH_REQUIRE(!((mode == 1 || mode == 3) && elem_bytes != 8), "Split-BF16 needs complex64 data.");
    JH_REQUIRE(!(mode == 4 && elem_bytes != 8), "Split-FP16 needs complex64 data.");
    JH_REQUIRE(!(mode == 2 && elem_bytes != 16), "FP64 MFMA needs complex128 data.");
    JH_REQUIRE(!(mode == 0 && elem_bytes != 8), "FP32 MFMA needs complex64 data.");

    // Row-major views (a column-major tensor is the row-major tensor with its
    // modes reversed, CudaTensor.hpp:312-323).
    std::vector<int32_t> A(modes_a, modes_a + ra), B(modes_b, modes_b + rb), C(modes_c, modes_c + rc);
    std::vector<int64_t> EA(ext_a, ext_a + ra), EB(ext_b, ext_b + rb);
    if (layout == 1) {
        std::reverse(A.begin(), A.end());
        std::reverse(B.begin(), B.end());
        std::reverse(C.begin(), C.end());
        std::reverse(EA.begin(), EA.end());
        std::reverse(EB.begin(), EB.end());
    }
    for (int i = 0; i < ra; i++) {
        JH_REQUIRE(EA[i] >= 1, "Tensor extents must be positive.");
        JH_REQUIRE(find(A, A[i]) == i, "Duplicate indices found in the first operand.");
    }
    for (int i = 0; i < rb; i++) {
        JH_REQUIRE(EB[i] >= 1, "Tensor extents must be positive.");
        JH_REQUIRE(find(B, B[i]) == i, "Duplicate indices found in the second operand.");
        const int ia = find(A, B[i]);
        JH_REQUIRE(ia < 0 || EA[ia] == EB[i], "Shared index has different extents.");
    }
    for (int i = 0; i < rc; i++) {
        JH_REQUIRE(find(C, C[i]) == i, "Duplicate indices found in the result.");
        JH_REQUIRE(find(A, C[i]) >= 0 || find(B, C[i]) >= 0,
                   "Result index is not an index of either operand.");
    }
    for (int i = 0; i < ra; i++)
        JH_REQUIRE(find(B, A[i]) >= 0 || find(C, A[i]) >= 0,
                   "Index of the first operand is neither contracted nor kept.");
    for (int i = 0; i < rb; i++)
        JH_REQUIRE(find(A, B[i]) >= 0 || find(C, B[i]) >= 0,
                   "Index of the second operand is neither contracted nor kept.");

    auto ext_of = [&](int32_t m) -> int64_t {
        const int ia = find(A, m);
        return ia >= 0 ? EA[ia] : EB[find(B, m)];
    };

    std::vector<int32_t> batch, left, right, common_a, common_b;
    for (int32_t m : C) {
        const bool in_a = find(A, m) >= 0, in_b = find(B, m) >= 0;
        if (in_a && in_b)
            batch.push_back(m);
        else if (in_a)
            left.push_back(m);
        else
            right.push_back(m);
    }
    for (int32_t m : A)
        if (find(B, m) >= 0 && find(C, m) < 0) common_a.push_back(m);
    for (int32_t m : B)
        if (find(A, m) >= 0 && find(C, m) < 0) common_b.push_back(m);

    ContractionDesc d;
    d.elem_bytes = elem_bytes;
    d.mode = mode;
    for (auto m : batch) d.batch *= ext_of(m);
    for (auto m : left) d.M *= ext_of(m);
    for (auto m : right) d.N *= ext_of(m);
    for (auto m : common_a) d.K *= ext_of(m);
    for (auto e : EA) d.size_a *= e;
    for (auto e : EB) d.size_b *= e;
    d.size_c = d.batch * d.M * d.N;

    auto cat = [](std::vector<int32_t> x, const std::vector<int32_t> &y,
                  const std::vector<int32_t> &z) {
        x.insert(x.end(), y.begin(), y.end());
        x.insert(x.end(), z.begin(), z.end());
        return x;
    };
    // Common-mode order: A's (the reference's choice) or B's, whichever
    // leaves fewer elements to permute.
    PermutePlan best_a, best_b;
    int64_t best_cost = -1;
    int best_opt = 0;
    for (int opt = 0; opt < 2; opt++) {
        const auto &common = opt == 0 ? common_a : common_b;
        PermutePlan pa = plan_reorder(A, EA, cat(batch, left, common));
        PermutePlan pb = plan_reorder(B, EB, cat(batch, right, common));
        const int64_t cost = (pa.kind != PermutePlan::kCopy ? d.size_a : 0) +
                             (pb.kind != PermutePlan::kCopy ? d.size_b : 0);
        if (best_cost < 0 || cost < best_cost) {
            best_cost = cost;
            best_a = pa;
            best_b = pb;
            best_opt = opt;
        }
    }
    d.pa = best_a;
    d.pb = best_b;
    d.need_a = d.pa.kind != PermutePlan::kCopy;
    d.need_b = d.pb.kind != PermutePlan::kCopy;
    {
        const std::vector<int32_t> ct = cat(batch, left, right);
        std::vector<int64_t> ect(ct.size());
        for (size_t i = 0; i < ct.size(); i++) ect[i] = ext_of(ct[i]);
        d.pc = plan_reorder(ct, ect, C);
        d.need_c = d.pc.kind != PermutePlan::kCopy;
    }

    const GemmChoice choice = tuned_gemm_choice(mode, d.M, d.N, d.K, d.batch, nullptr);
    d.gemm_kernel = choice.kernel;
    d.gemm_variant = choice.variant;
    d.flops = 8.0 * double(d.batch) * double(d.M) * double(d.N) * double(d.K);
    std::vector<int32_t> gemm_batch = batch;
    std::vector<int32_t> gemm_common = best_opt == 0 ? common_a : common_b;
    {
        // Many outputs and a long K on a VALU kernel (the split-K and
        // one-block-per-output forms are built for few outputs): split K
        // into chunks that become leading batch modes of an MFMA GEMM, then
        // sum the chunks (M and N at least 64).  The chunk modes are the leading common modes
        // (the slowest-varying K digits), about 512 output tiles in all,
        // and at least 2048 k per chunk.
        const bool valu = d.gemm_kernel == kGemmSplitK || d.gemm_kernel == kGemmSmallRed;
        const bool mfma_mode = mode == 0 || mode == 2 || mode == 4;
        if (valu && mfma_mode && d.M >= 64 && d.N >= 64 && d.K >= (int64_t(1) << 15)) {
            const int64_t tiles = d.batch * ceil_div(d.M, 128) * ceil_div(d.N, 128);
            int64_t want = 1;
            while (tiles * want < 512) want *= 2;
            int64_t chunks = 1;
            size_t n_chunk_modes = 0;
            while (n_chunk_modes < gemm_common.size()) {
                const int64_t e = ext_of(gemm_common[n_chunk_modes]);
                if (chunks * e > want || d.K / (chunks * e) < 2048) break;
                chunks *= e;
                n_chunk_modes++;
            }
            if (chunks > 1) {
                const int64_t k2 = d.K / chunks, b2 = d.batch * chunks;
                const GemmChoice c2 = tuned_gemm_choice(mode, d.M, d.N, k2, b2, nullptr);
                if ((c2.kernel == kGemmF32Mfma || c2.kernel == kGemmF64Mfma ||
                     c2.kernel == kGemmF16x4m) &&
                    gemm_choice_valid(mode, d.M, d.N, k2, b2, c2)) {
                    std::vector<int32_t> chunk_modes(gemm_common.begin(),
                                                     gemm_common.begin() + n_chunk_modes);
                    gemm_common.erase(gemm_common.begin(), gemm_common.begin() + n_chunk_modes);
                    gemm_batch = chunk_modes;
                    gemm_batch.insert(gemm_batch.end(), batch.begin(), batch.end());
                    d.chunks = chunks;
                    d.batch = b2;
                    d.K = k2;
                    d.gemm_kernel = c2.kernel;
                    d.gemm_variant = c2.variant;
                    d.pa = plan_reorder(A, EA, cat(gemm_batch, left, gemm_common));
                    d.pb = plan_reorder(B, EB, cat(gemm_batch, right, gemm_common));
                    d.need_a = d.pa.kind != PermutePlan::kCopy;
                    d.need_b = d.pb.kind != PermutePlan::kCopy;
                }
            }
        }
    }
    if (d.gemm_kernel == kGemmF16x4m) {
        // fold a permutation into the split pass when the rows need no
        // padding (the split rows are then the permuted rows)
        if (d.need_a && d.M % kSplitRowPad == 0)
            d.fuse_a = plan_split_gather(A, EA, cat(gemm_batch, left, gemm_common), d.K, d.ga);
        if (d.need_b && d.N % kSplitRowPad == 0)
            d.fuse_b = plan_split_gather(B, EB, cat(gemm_batch, right, gemm_common), d.K, d.gb);
    }

    // Small steps on the one-thread-per-output or the one-block-per-output
    // VALU kernel that would need a permuted operand or result: one fused
    // pass of the same form, no copies.
    const double work = double(d.batch) * double(d.M) * double(d.N) * double(d.K);
    const bool small = d.gemm_kernel == kGemmSmall && d.gemm_variant <= 1 && work <= double(1 << 22);
    const bool red = d.gemm_kernel == kGemmSmallRed && work <= double(1 << 26);
    if ((small || red) && (d.need_a || d.need_b || d.need_c)) {
        const auto &common = best_opt == 0 ? common_a : common_b;
        d.fused_small = plan_fused_small(A, EA, B, EB, C, batch, left, right, common, d.fs);
        d.fs.red = red ? 1 : 0;
    }

    layout_workspace(d);
    return d;
}

void layout_workspace(ContractionDesc &d)
{
    size_t off = 0;
    const size_t eb = (size_t)d.elem_bytes;
    if (d.fused_small) { // reads and writes the tensors in place
        d.workspace = 0;
        return;
    }
    if (d.need_a && !d.fuse_a) {
        d.off_at = off;
        off = align256(off + eb * (size_t)d.size_a);
    }
    if (d.need_b && !d.fuse_b) {
        d.off_bt = off;
        off = align256(off + eb * (size_t)d.size_b);
    }
    if (d.need_c) {
        d.off_ct = off;
        off = align256(off + eb * (size_t)d.size_c);
    }
    if (d.gemm_kernel == kGemmBf16x3 || d.gemm_kernel == kGemmBf16x4m ||
        d.gemm_kernel == kGemmF16x4m) {
        const size_t planes = d.gemm_kernel == kGemmBf16x3 ? 6 : 4;
        d.split_a_plane = split_plane_elems(d.batch, d.M, d.K);
        d.split_b_plane = split_plane_elems(d.batch, d.N, d.K);
        d.off_sa = off;
        off = align256(off + planes * 2 * (size_t)d.split_a_plane);
        d.off_sb = off;
        off = align256(off + planes * 2 * (size_t)d.split_b_plane);
        if (d.gemm_kernel == kGemmF16x4m) {
            const size_t ra = (size_t)d.batch * (size_t)(ceil_div(d.M, kSplitRowPad) * kSplitRowPad);
            const size_t rb = (size_t)d.batch * (size_t)(ceil_div(d.N, kSplitRowPad) * kSplitRowPad);
            d.off_rmax_a = off;
            off = align256(off + 4 * ra);
            d.off_scale_a = off;
            off = align256(off + 4 * ra);
            d.off_rmax_b = off;
            off = align256(off + 4 * rb);
            d.off_scale_b = off;
            off = align256(off + 4 * rb);
        }
    }
    if (d.gemm_kernel == kGemmSplitK) {
        d.off_partials = off;
        off = align256(off + 16 * (size_t)splitk_partials(d.batch * d.M * d.N, d.K));
    }
    if (d.chunks > 1) {
        d.off_chunks = off;
        off = align256(off + eb * (size_t)d.batch * (size_t)d.M * (size_t)d.N);
    }
    d.workspace = off;
}

} // namespace jet_hip

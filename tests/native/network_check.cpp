// network_check.cpp -- CPU checks of network plans (arena layout).
//
// Builds the plan of jet_amd/csrc/network_plan.cpp for a network and
// verifies, before anything runs on a GPU, that every intermediate result
// lies inside the arena, that no two results alive at the same time
// overlap, and that each step's operands are the results of earlier steps
// or leaves.
#include "../../jet_amd/csrc/network_plan.hpp"
#include "../../jet_amd/csrc/tuning.hpp"

#include <cstring>
#include <vector>

using namespace jet_hip;

extern "C" int check_network(int elem_bytes, int mode, int n_leaves, const int32_t *ranks,
                             const int32_t *modes, const int64_t *exts, int n_steps,
                             const int32_t *path, int n_out, const int32_t *out,
                             double *flops, size_t *arena, size_t *ws, char *err, int errlen)
{
    NetworkPlan np;
    try {
        np = plan_network(elem_bytes, mode, n_leaves, ranks, modes, exts, n_steps, path, n_out,
                          out);
    }
    catch (const Error &e) {
        std::strncpy(err, e.what(), errlen - 1);
        err[errlen - 1] = 0;
        return 3;
    }
    *flops = np.flops;
    *arena = np.arena_bytes;
    *ws = np.workspace_bytes;
    const int n = n_leaves + n_steps;
    std::vector<int> last_use(n, -1);
    for (int s = 0; s < n_steps; s++) {
        last_use[path[2 * s]] = s;
        last_use[path[2 * s + 1]] = s;
    }
    struct Live {
        size_t off, len;
        int id;
    };
    std::vector<Live> live;
    for (int s = 0; s < n_steps; s++) {
        const NetworkStep &st = np.steps[s];
        if (st.a >= n_leaves + s || st.b >= n_leaves + s) return 4;
        if (st.desc.workspace > np.workspace_bytes) return 5;
        if (!st.is_root) {
            const size_t len = (size_t)elem_bytes * (size_t)st.desc.size_c;
            if (st.offset + len > np.arena_bytes) return 1;
            for (const Live &l : live)
                if (st.offset < l.off + l.len && l.off < st.offset + len) return 2;
            live.push_back({st.offset, len, n_leaves + s});
        }
        // operands consumed here are released after the step
        std::vector<Live> keep;
        for (const Live &l : live)
            if (!(last_use[l.id] == s && (l.id == st.a || l.id == st.b))) keep.push_back(l);
        live.swap(keep);
    }
    return 0;
}

// ---------------------------------------------------------------------------
// Emulation of a whole planned network on the host in complex128: each
// step's operand permutations (the plan's PermutePlans, replayed by index
// arithmetic), the batched NT product C'[b][m][n] = sum_k A'[b][m][k]
// B'[b][n][k], then the result permuted (its C plan) or stored as is.
// Returns 0 and the root in out, or 1 for an out-of-bounds index, 3 for a
// planner error; with the number of permuted operand elements.
#include <complex>

namespace {

using cd = std::complex<double>;

int apply_permute(const PermutePlan &plan, const cd *src, cd *dst)
{
    const int64_t n = plan.numel;
    if (plan.kind == PermutePlan::kCopy) {
        for (int64_t i = 0; i < n; i++) dst[i] = src[i];
        return 0;
    }
    if (plan.kind == PermutePlan::kGeneric) {
        for (int64_t o = 0; o < n; o++) {
            const int64_t off = generic_in_offset(plan.gen, o);
            if (off < 0 || off >= n) return 1;
            dst[o] = src[off];
        }
        return 0;
    }
    const BitPermParams &p = plan.bits;
    const int tsize = 1 << p.tb;
    const int64_t ntiles = int64_t(1) << p.nr;
    for (int64_t w = 0; w < ntiles; w++) {
        int64_t bin, bout;
        bits_tile_base(p, w, &bin, &bout);
        for (int v = 0; v < tsize; v++) {
            int uu = 0;
            const int64_t ooff = bits_tile_out_offset(p, v, &uu);
            const int64_t ioff = bits_tile_in_offset(p, uu);
            if (bin + ioff >= n || bout + ooff >= n) return 1;
            dst[bout + ooff] = src[bin + ioff];
        }
    }
    return 0;
}

} // namespace

extern "C" int emu_network(int mode, int n_leaves, const int32_t *ranks, const int32_t *modes,
                           const int64_t *exts, int n_steps, const int32_t *path, int n_out,
                           const int32_t *out_modes, const double *const *leaves, double *out,
                           double *permuted, char *err, int errlen)
{
    NetworkPlan np;
    try {
        np = plan_network(8, mode, n_leaves, ranks, modes, exts, n_steps, path, n_out, out_modes);
    }
    catch (const Error &e) {
        std::strncpy(err, e.what(), errlen - 1);
        err[errlen - 1] = 0;
        return 3;
    }
    *permuted = 0;
    std::vector<std::vector<cd>> node(n_leaves + n_steps);
    size_t pos = 0;
    for (int i = 0; i < n_leaves; i++) {
        int64_t n = 1;
        for (int j = 0; j < ranks[i]; j++) n *= exts[pos + j];
        pos += ranks[i];
        node[i].resize(n);
        for (int64_t e = 0; e < n; e++) node[i][e] = cd(leaves[i][2 * e], leaves[i][2 * e + 1]);
    }
    for (int s = 0; s < n_steps; s++) {
        const NetworkStep &st = np.steps[s];
        const ContractionDesc &d = st.desc;
        const std::vector<cd> &A = node[st.a], &B = node[st.b];
        if ((int64_t)A.size() != d.size_a || (int64_t)B.size() != d.size_b) return 1;
        std::vector<cd> At(A), Bt(B);
        if (d.need_a) {
            if (apply_permute(d.pa, A.data(), At.data())) return 1;
            *permuted += double(d.size_a);
        }
        if (d.need_b) {
            if (apply_permute(d.pb, B.data(), Bt.data())) return 1;
            *permuted += double(d.size_b);
        }
        std::vector<cd> Cn((size_t)d.size_c), C((size_t)d.size_c);
        for (int64_t b = 0; b < d.batch; b++)
            for (int64_t m = 0; m < d.M; m++)
                for (int64_t n = 0; n < d.N; n++) {
                    cd acc = 0;
                    const cd *a = At.data() + (b * d.M + m) * d.K;
                    const cd *bb = Bt.data() + (b * d.N + n) * d.K;
                    for (int64_t k = 0; k < d.K; k++) acc += a[k] * bb[k];
                    Cn[(b * d.M + m) * d.N + n] = acc;
                }
        if (d.need_c) {
            if (apply_permute(d.pc, Cn.data(), C.data())) return 1;
        }
        else {
            C.swap(Cn);
        }
        node[n_leaves + s].swap(C);
    }
    const std::vector<cd> &root = n_steps ? node[n_leaves + n_steps - 1] : node[0];
    for (size_t e = 0; e < root.size(); e++) {
        out[2 * e] = root[e].real();
        out[2 * e + 1] = root[e].imag();
    }
    return 0;
}

// Elements the plan permutes (operands and results).
extern "C" int network_permute_stats(int mode, int n_leaves, const int32_t *ranks,
                                     const int32_t *modes, const int64_t *exts, int n_steps,
                                     const int32_t *path, int n_out, const int32_t *out_modes,
                                     double *permuted)
{
    NetworkPlan np;
    try {
        np = plan_network(8, mode, n_leaves, ranks, modes, exts, n_steps, path, n_out, out_modes);
    }
    catch (const Error &) {
        return 3;
    }
    *permuted = 0;
    for (const NetworkStep &st : np.steps) {
        const ContractionDesc &d = st.desc;
        *permuted += (d.need_a ? double(d.size_a) : 0) + (d.need_b ? double(d.size_b) : 0) +
                     (d.need_c ? double(d.size_c) : 0);
    }
    return 0;
}

// Emulates split_f16_gather_kernel's index mapping on the CPU: for a source
// tensor (row-major modes and extents) and the target order of its split
// operand [rows][K], out[source index] = R * K + k, the target row and k that
// the kernel writes the element to.  Returns 1 when the gather does not
// apply, 0 otherwise (info = tb, xr, nw).
extern "C" int split_gather_targets(int rank, const int32_t *modes, const int64_t *exts,
                                    int trank, const int32_t *target, int64_t K, int64_t *out,
                                    int *info)
{
    std::vector<int32_t> src(modes, modes + rank), tgt(target, target + trank);
    std::vector<int64_t> ext(exts, exts + rank);
    SplitGather g;
    if (!plan_split_gather(src, ext, tgt, K, g)) return 1;
    info[0] = g.tb;
    info[1] = g.xr;
    info[2] = g.nw;
    const int64_t ntiles = int64_t(1) << g.nw, tsz = int64_t(1) << g.tb;
    for (int64_t w = 0; w < ntiles; w++) {
        int64_t sbase = 0, rbase = 0;
        for (int j = 0; j < g.nw; j++)
            if ((w >> j) & 1) {
                sbase += int64_t(1) << g.wsrc[j];
                rbase += int64_t(1) << g.wrow[j];
            }
        for (int64_t u = 0; u < tsz; u++) {
            int64_t so = sbase;
            int v = 0;
            for (int j = 0; j < g.tb; j++)
                if ((u >> j) & 1) {
                    so += int64_t(1) << g.su[j];
                    v |= 1 << g.u2v[j];
                }
            const int rl = v >> g.log2k, k = v & ((1 << g.log2k) - 1);
            int64_t row = rbase;
            for (int i = 0; i < g.xr; i++)
                if ((rl >> i) & 1) row += int64_t(1) << g.vrow[i];
            out[so] = row * K + k;
        }
    }
    return 0;
}

// The fused small-contraction pass (plan_fused_small / fused_small_kernel)
// emulated on the CPU in complex128: the same mode groups as
// plan_contraction (batch = in A, B and C; left / right = one operand and C;
// common = A and B only, in A's order), the same group-bit offsets and the
// same per-output summation over k.  a, b, c interleaved (re, im).  Returns
// 1 when the fused pass does not apply.
extern "C" int fused_small_emulate(int ra, const int32_t *ma, const int64_t *ea, int rb,
                                   const int32_t *mb, const int64_t *eb, int rc,
                                   const int32_t *mc, const double *a, const double *b,
                                   double *c)
{
    std::vector<int32_t> A(ma, ma + ra), B(mb, mb + rb), C(mc, mc + rc);
    std::vector<int64_t> EA(ea, ea + ra), EB(eb, eb + rb);
    auto has = [](const std::vector<int32_t> &v, int32_t m) {
        for (auto x : v)
            if (x == m) return true;
        return false;
    };
    std::vector<int32_t> batch, left, right, common;
    for (auto m : C) {
        if (has(A, m) && has(B, m))
            batch.push_back(m);
        else if (has(A, m))
            left.push_back(m);
        else
            right.push_back(m);
    }
    for (auto m : A)
        if (has(B, m) && !has(C, m)) common.push_back(m);
    FusedSmall f;
    if (!plan_fused_small(A, EA, B, EB, C, batch, left, right, common, f)) return 1;
    auto off = [](const FusedBits &fb, int64_t x) {
        int64_t o = 0;
        for (int i = 0; i < fb.n; i++) o |= ((x >> i) & 1) << fb.pos[i];
        return o;
    };
    const int64_t total = int64_t(1) << (f.lb + f.lm + f.ln), K = int64_t(1) << f.lk;
    for (int64_t idx = 0; idx < total; idx++) {
        const int64_t n = idx & ((int64_t(1) << f.ln) - 1);
        const int64_t m = (idx >> f.ln) & ((int64_t(1) << f.lm) - 1);
        const int64_t bb = idx >> (f.ln + f.lm);
        const int64_t oa = off(f.a_b, bb) | off(f.a_m, m), ob = off(f.b_b, bb) | off(f.b_n, n);
        double sr = 0, si = 0;
        for (int64_t k = 0; k < K; k++) {
            const int64_t ia = oa | off(f.a_k, k), ib = ob | off(f.b_k, k);
            const double xr = a[2 * ia], xi = a[2 * ia + 1], yr = b[2 * ib], yi = b[2 * ib + 1];
            sr += xr * yr - xi * yi;
            si += xr * yi + xi * yr;
        }
        const int64_t oc = off(f.c_b, bb) | off(f.c_m, m) | off(f.c_n, n);
        c[2 * oc] = sr;
        c[2 * oc + 1] = si;
    }
    return 0;
}

// Whether plan_contraction (compiled-in GEMM table) fuses this contraction
// of complex64 operands in FP32 (mode 0): 1 fused, 0 not.
extern "C" int contraction_is_fused(int ra, const int32_t *ma, const int64_t *ea, int rb,
                                    const int32_t *mb, const int64_t *eb, int rc,
                                    const int32_t *mc, size_t *workspace)
{
    const ContractionDesc d = plan_contraction(8, 0, 0, ra, ma, ea, rb, mb, eb, rc, mc);
    *workspace = d.workspace;
    return d.fused_small ? 1 : 0;
}

// plan_contraction's GEMM for complex64 (mode 0 or 4) or complex128 (mode
// 2) operands, with the table replaced by `n_entries` entries (7 ints each;
// n_entries < 0 keeps the compiled-in table): kernel, chunks (K split into
// batched chunks and summed), the GEMM's batch and K, the workspace bytes and
// the chunk buffer's offset.
extern "C" int contraction_gemm_plan(int mode, int n_entries, const int32_t *entries, int ra,
                                     const int32_t *ma, const int64_t *ea, int rb,
                                     const int32_t *mb, const int64_t *eb, int rc,
                                     const int32_t *mc, int64_t *out)
{
    if (n_entries >= 0) {
        std::vector<TuneEntry> t(n_entries);
        for (int i = 0; i < n_entries; i++) {
            const int32_t *e = entries + 7 * i;
            t[i] = TuneEntry{e[0], e[1], e[2], e[3], e[4], e[5], e[6]};
        }
        tuning_set(t);
    }
    const int elem = mode == 2 ? 16 : 8;
    const ContractionDesc d = plan_contraction(elem, mode, 0, ra, ma, ea, rb, mb, eb, rc, mc);
    if (n_entries >= 0) tuning_reset();
    out[0] = d.gemm_kernel;
    out[1] = d.chunks;
    out[2] = d.batch;
    out[3] = d.K;
    out[4] = (int64_t)d.workspace;
    out[5] = (int64_t)d.off_chunks;
    out[6] = d.size_c;
    return 0;
}

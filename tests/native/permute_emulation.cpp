// permute_emulation.cpp -- CPU emulation of the HIP permutation kernels.
//
// Runs the planner of jet_amd/csrc/permute_plan.cpp and replays the exact
// index arithmetic of permute_bits_kernel / permute_generic_kernel (the
// shared JH_HD helpers) on the host, checking every load and store against
// the buffer bounds and checking that every output element is written once.
// Used by the CPU test suite so that plans are validated before any GPU run.
#include "../../jet_amd/csrc/permute_plan.hpp"

#include <cstring>
#include <vector>

using namespace jet_hip;

extern "C" {

// Returns 0 on success, 1 for an out-of-bounds access, 2 for an output
// element written 0 or >1 times, 3 for a planner error (message in err).
// kind_out: 0 copy, 1 bits, 2 generic; tile_bits_out: tile bits of the plan.
int emu_permute(int elem, int rank, const int64_t *extents, const int32_t *perm,
                const void *in, void *out, int *kind_out, int *tile_bits_out,
                char *err, int errlen)
{
    PermutePlan plan;
    try {
        plan = make_permute_plan(rank, extents, perm);
    }
    catch (const Error &e) {
        std::strncpy(err, e.what(), errlen - 1);
        err[errlen - 1] = 0;
        return 3;
    }
    const int64_t n = plan.numel;
    const char *src = static_cast<const char *>(in);
    char *dst = static_cast<char *>(out);
    std::vector<int> written(n, 0);
    *kind_out = (int)plan.kind;
    *tile_bits_out = plan.kind == PermutePlan::kBits ? plan.bits.tb : 0;
    if (plan.kind == PermutePlan::kCopy) {
        std::memcpy(dst, src, (size_t)n * elem);
        return 0;
    }
    if (plan.kind == PermutePlan::kGeneric) {
        for (int64_t o = 0; o < n; o++) {
            const int64_t off = generic_in_offset(plan.gen, o);
            if (off < 0 || off >= n) return 1;
            std::memcpy(dst + o * elem, src + off * elem, elem);
            written[o]++;
        }
    }
    else {
        const BitPermParams &p = plan.bits;
        const int tsize = 1 << p.tb;
        if (p.tb > kTileMaxBits || p.tb < 0 || p.nr + p.tb != p.nbits) return 1;
        const int64_t ntiles = int64_t(1) << p.nr;
        std::vector<char> lds((size_t)tsize * elem);
        for (int64_t w = 0; w < ntiles; w++) {
            int64_t bin, bout;
            bits_tile_base(p, w, &bin, &bout);
            // Thread/element loops of the kernel collapse to one loop over u.
            for (int u = 0; u < tsize; u++) {
                const int64_t ioff = bits_tile_in_offset(p, u);
                if (bin + ioff < 0 || bin + ioff >= n) return 1;
                const int slot = bits_lds_index(u, elem);
                if (slot < 0 || slot >= tsize) return 1;
                std::memcpy(&lds[(size_t)slot * elem], src + (bin + ioff) * elem, elem);
            }
            for (int v = 0; v < tsize; v++) {
                int uu = 0;
                const int64_t ooff = bits_tile_out_offset(p, v, &uu);
                const int slot = bits_lds_index(uu, elem);
                if (uu < 0 || uu >= tsize || slot < 0 || slot >= tsize) return 1;
                if (bout + ooff < 0 || bout + ooff >= n) return 1;
                std::memcpy(dst + (bout + ooff) * elem, &lds[(size_t)slot * elem], elem);
                written[bout + ooff]++;
            }
        }
    }
    for (int64_t o = 0; o < n; o++)
        if (written[o] != 1) return 2;
    return 0;
}

} // extern "C"

extern "C" {

// The bit plan of a permutation without data: nbits, tb, ti, nr and the
// tile/outer bit maps (arrays of 16, 16, 16, 64, 64 entries).  Returns the
// plan kind (0 copy, 1 bits, 2 generic) or -1 for a planner error.
int emu_bits_plan(int rank, const int64_t *extents, const int32_t *perm, int *scalars,
                  int8_t *tile_in, int8_t *tile_out, int8_t *v2u, int8_t *outer_in,
                  int8_t *outer_out)
{
    PermutePlan plan;
    try {
        plan = make_permute_plan(rank, extents, perm);
    }
    catch (const Error &) {
        return -1;
    }
    const BitPermParams &p = plan.bits;
    scalars[0] = p.nbits;
    scalars[1] = p.tb;
    scalars[2] = p.ti;
    scalars[3] = p.nr;
    std::memcpy(tile_in, p.tile_in, 16);
    std::memcpy(tile_out, p.tile_out, 16);
    std::memcpy(v2u, p.v2u, 16);
    std::memcpy(outer_in, p.outer_in, 64);
    std::memcpy(outer_out, p.outer_out, 64);
    return (int)plan.kind;
}

} // extern "C"

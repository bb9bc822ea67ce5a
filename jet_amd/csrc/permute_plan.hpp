// permute_plan.hpp -- planning of tensor permutations for gfx950.
//
// A permutation of a row-major tensor (output axis i = input axis perm[i]) is
// first simplified (extent-1 axes dropped, runs of input axes that stay
// adjacent merged).  When every remaining extent is a power of two the
// permutation is a permutation of the bits of the linear index, and the
// LDS-tiled bit kernel moves tiles made of the lowest input bits and the
// lowest output bits, so that both the reads and the writes are contiguous
// runs.  Other shapes use a gather kernel (coalesced writes).
//
// Pure C++: the index helpers below are shared by the kernels and by the CPU
// emulation that checks plans and bounds before anything runs on a GPU.
#pragma once

#include "errors.hpp"

#include <cstdint>

namespace jet_hip {

constexpr int kMaxRank = 64;
constexpr int kTileMaxBits = 11; // 2048 elements per tile
constexpr int kPermThreads = 256;

struct BitPermParams {
    int nbits;            // bits of the linear index
    int tb;               // tile bits
    int ti;               // lowest input bits inside the tile (contiguous reads)
    int nr;               // outer (tile-index) bits
    int8_t tile_in[16];   // u-bit j -> input bit position (ascending)
    int8_t tile_out[16];  // v-bit j -> output bit position (ascending)
    int8_t v2u[16];       // v-bit j -> u-bit position of the same element
    int8_t outer_in[64];  // w-bit j -> input bit position
    int8_t outer_out[64]; // w-bit j -> output bit position
};

struct GenericPermParams {
    int rank;
    int64_t out_ext[kMaxRank];   // extents in output order
    int64_t in_stride[kMaxRank]; // input stride of output axis i
};

struct PermutePlan {
    enum Kind { kCopy, kBits, kGeneric } kind = kCopy;
    int64_t numel = 0;
    BitPermParams bits{};
    GenericPermParams gen{};
};

// Build a plan; extents are the input's (row-major), perm[i] = input axis of
// output axis i.
PermutePlan make_permute_plan(int rank, const int64_t *extents,
                              const int32_t *perm);

#if defined(__HIPCC__)
// elem_bytes: 8 (complex64) or 16 (complex128).
void launch_permute(const PermutePlan &plan, int elem_bytes, const void *in,
                    void *out, hipStream_t stream);
#endif

// In-tile input offset of tile element u (input order).
JH_HD int64_t bits_tile_in_offset(const BitPermParams &p, int u)
{
    int64_t o = u & ((int64_t(1) << p.ti) - 1);
    for (int j = p.ti; j < p.tb; j++)
        if ((u >> j) & 1) o += int64_t(1) << p.tile_in[j];
    return o;
}

// In-tile output offset of tile element v (output order) and the input-order
// tile index of the same element.
JH_HD int64_t bits_tile_out_offset(const BitPermParams &p, int v, int *u)
{
    int64_t o = 0;
    int uu = 0;
    for (int j = 0; j < p.tb; j++) {
        if ((v >> j) & 1) {
            o += int64_t(1) << p.tile_out[j];
            uu |= 1 << p.v2u[j];
        }
    }
    *u = uu;
    return o;
}

// Input and output base offsets of tile w.
JH_HD void bits_tile_base(const BitPermParams &p, int64_t w, int64_t *bin,
                          int64_t *bout)
{
    int64_t i = 0, o = 0;
    for (int j = 0; j < p.nr; j++) {
        if ((w >> j) & 1) {
            i += int64_t(1) << p.outer_in[j];
            o += int64_t(1) << p.outer_out[j];
        }
    }
    *bin = i;
    *bout = o;
}

// LDS slot of tile element u: XOR the low bits with the next group of bits so
// that the strided reads of the output phase spread over the LDS banks
// (32 slots of 8 bytes, or 16 of 16 bytes, per 256-byte bank row).
JH_HD int bits_lds_index(int u, int elem_bytes)
{
    const int s = elem_bytes == 8 ? 5 : 4;
    return u ^ ((u >> s) & ((1 << s) - 1));
}

// Input offset of output element o for the gather kernel.
JH_HD int64_t generic_in_offset(const GenericPermParams &p, int64_t o)
{
    int64_t off = 0;
    for (int i = p.rank - 1; i >= 0; i--) {
        const int64_t e = p.out_ext[i];
        const int64_t idx = o % e;
        o /= e;
        off += idx * p.in_stride[i];
    }
    return off;
}

} // namespace jet_hip

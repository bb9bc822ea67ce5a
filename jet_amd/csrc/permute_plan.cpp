// permute_plan.cpp -- host-side planning of tensor permutations (pure C++,
// also compiled into the CPU emulation used by the test suite).
#include "permute_plan.hpp"

#include <algorithm>
#include <numeric>
#include <vector>

namespace jet_hip {

namespace {

bool is_pow2(int64_t v) { return v > 0 && (v & (v - 1)) == 0; }

int ilog2(int64_t v)
{
    int r = 0;
    while (v > 1) {
        v >>= 1;
        r++;
    }
    return r;
}

} // namespace

PermutePlan make_permute_plan(int rank, const int64_t *extents,
                              const int32_t *perm)
{
    PermutePlan plan;
    JH_REQUIRE(rank >= 0 && rank <= kMaxRank, "Permutation rank exceeds 64.");
    int64_t numel = 1;
    for (int i = 0; i < rank; i++) {
        JH_REQUIRE(extents[i] >= 1, "Tensor extents must be positive.");
        numel *= extents[i];
    }
    plan.numel = numel;
    {
        std::vector<int> seen(rank, 0);
        for (int i = 0; i < rank; i++) {
            JH_REQUIRE(perm[i] >= 0 && perm[i] < rank && !seen[perm[i]],
                       "New indices are an invalid permutation of the existing indices");
            seen[perm[i]] = 1;
        }
    }
    // Drop extent-1 axes.
    std::vector<int> keep_map(rank, -1);
    std::vector<int64_t> e;
    for (int a = 0; a < rank; a++) {
        if (extents[a] > 1) {
            keep_map[a] = (int)e.size();
            e.push_back(extents[a]);
        }
    }
    std::vector<int> p;
    for (int i = 0; i < rank; i++)
        if (keep_map[perm[i]] >= 0) p.push_back(keep_map[perm[i]]);
    // Merge runs of input axes that stay adjacent in the output.
    std::vector<std::vector<int>> groups;
    for (size_t i = 0; i < p.size(); i++) {
        if (!groups.empty() && groups.back().back() + 1 == p[i])
            groups.back().push_back(p[i]);
        else
            groups.push_back({p[i]});
    }
    const int r = (int)groups.size();
    if (r <= 1 || numel <= 1) {
        plan.kind = PermutePlan::kCopy;
        return plan;
    }
    std::vector<int> by_input(r);
    std::iota(by_input.begin(), by_input.end(), 0);
    std::sort(by_input.begin(), by_input.end(),
              [&](int x, int y) { return groups[x][0] < groups[y][0]; });
    std::vector<int> group_to_in(r);
    std::vector<int64_t> me(r);
    for (int a = 0; a < r; a++) {
        group_to_in[by_input[a]] = a;
        int64_t ext = 1;
        for (int ax : groups[by_input[a]]) ext *= e[ax];
        me[a] = ext;
    }
    std::vector<int> mp(r);
    for (int i = 0; i < r; i++) mp[i] = group_to_in[i];

    bool pow2 = true;
    for (int a = 0; a < r; a++) pow2 = pow2 && is_pow2(me[a]);

    if (!pow2) {
        plan.kind = PermutePlan::kGeneric;
        GenericPermParams &g = plan.gen;
        g.rank = r;
        std::vector<int64_t> in_stride(r);
        in_stride[r - 1] = 1;
        for (int a = r - 1; a > 0; a--) in_stride[a - 1] = in_stride[a] * me[a];
        for (int i = 0; i < r; i++) {
            g.out_ext[i] = me[mp[i]];
            g.in_stride[i] = in_stride[mp[i]];
        }
        return plan;
    }

    // Bit permutation: dst[b] = output bit of input bit b.
    std::vector<int> L(r), lo_in(r), lo_out(r);
    for (int a = 0; a < r; a++) L[a] = ilog2(me[a]);
    int acc = 0;
    for (int a = r - 1; a >= 0; a--) {
        lo_in[a] = acc;
        acc += L[a];
    }
    const int nbits = acc;
    JH_REQUIRE(nbits <= 62, "Tensor too large for the bit permutation kernel.");
    acc = 0;
    for (int i = r - 1; i >= 0; i--) {
        lo_out[i] = acc;
        acc += L[mp[i]];
    }
    std::vector<int> dst(nbits), src(nbits);
    for (int i = 0; i < r; i++)
        for (int j = 0; j < L[mp[i]]; j++) dst[lo_in[mp[i]] + j] = lo_out[i] + j;
    for (int b = 0; b < nbits; b++) src[dst[b]] = b;

    // Tile: alternately add the next lowest input bit and the input bit of
    // the next lowest output bit until the tile holds `target` bits.
    const int target = std::min(nbits, kTileMaxBits);
    std::vector<char> in_tile(nbits, 0);
    int count = 0, ti = 0, to = 0;
    while (count < target) {
        if (ti <= to && ti < nbits) {
            const int b = ti++;
            if (!in_tile[b]) {
                in_tile[b] = 1;
                count++;
            }
        }
        else if (to < nbits) {
            const int b = src[to++];
            if (!in_tile[b]) {
                in_tile[b] = 1;
                count++;
            }
        }
        else
            break;
    }
    BitPermParams &bp = plan.bits;
    bp.nbits = nbits;
    bp.tb = count;
    int contiguous = 0;
    while (contiguous < nbits && in_tile[contiguous]) contiguous++;
    bp.ti = contiguous;
    std::vector<int> tile_in, tile_by_out;
    for (int b = 0; b < nbits; b++)
        if (in_tile[b]) tile_in.push_back(b);
    for (int o = 0; o < nbits; o++)
        if (in_tile[src[o]]) tile_by_out.push_back(src[o]);
    for (int j = 0; j < count; j++) {
        bp.tile_in[j] = (int8_t)tile_in[j];
        bp.tile_out[j] = (int8_t)dst[tile_by_out[j]];
        const int upos = (int)(std::find(tile_in.begin(), tile_in.end(), tile_by_out[j]) -
                               tile_in.begin());
        bp.v2u[j] = (int8_t)upos;
    }
    int nr = 0;
    for (int b = 0; b < nbits; b++) {
        if (!in_tile[b]) {
            bp.outer_in[nr] = (int8_t)b;
            bp.outer_out[nr] = (int8_t)dst[b];
            nr++;
        }
    }
    bp.nr = nr;
    plan.kind = PermutePlan::kBits;
    return plan;
}

} // namespace jet_hip

// network_plan.cpp -- host-side plan of a contraction sequence.
#include "network_plan.hpp"

#include <algorithm>
#include <iterator>
#include <map>
#include <set>
#include <unordered_map>

namespace jet_hip {

namespace {

size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

int find(const std::vector<int32_t> &v, int32_t x)
{
    for (size_t i = 0; i < v.size(); i++)
        if (v[i] == x) return (int)i;
    return -1;
}

// Best-fit allocator over a virtual address range; tracks the high-water mark.
class ArenaSim {
  public:
    size_t alloc(size_t bytes)
    {
        bytes = align256(std::max<size_t>(bytes, 256));
        auto best = free_.end();
        for (auto it = free_.begin(); it != free_.end(); ++it)
            if (it->second >= bytes && (best == free_.end() || it->second < best->second)) best = it;
        size_t off;
        if (best != free_.end()) {
            off = best->first;
            const size_t rest = best->second - bytes;
            free_.erase(best);
            if (rest) free_[off + bytes] = rest;
        }
        else {
            // Grow at the end (merging a free block that ends at the top).
            off = top_;
            if (!free_.empty()) {
                auto last = std::prev(free_.end());
                if (last->first + last->second == top_) {
                    off = last->first;
                    free_.erase(last);
                }
            }
            top_ = off + bytes;
        }
        used_[off] = bytes;
        return off;
    }
    void release(size_t off)
    {
        auto it = used_.find(off);
        if (it == used_.end()) return;
        size_t start = off, len = it->second;
        used_.erase(it);
        auto nx = free_.lower_bound(start);
        if (nx != free_.end() && start + len == nx->first) {
            len += nx->second;
            nx = free_.erase(nx);
        }
        if (nx != free_.begin()) {
            auto pv = std::prev(nx);
            if (pv->first + pv->second == start) {
                start = pv->first;
                len += pv->second;
                free_.erase(pv);
            }
        }
        free_[start] = len;
    }
    size_t high_water() const { return top_; }

  private:
    std::map<size_t, size_t> free_; // offset -> length
    std::map<size_t, size_t> used_;
    size_t top_ = 0;
};

} // namespace

NetworkPlan plan_network(int elem_bytes, int mode, int n_leaves,
                         const int32_t *leaf_ranks, const int32_t *leaf_modes,
                         const int64_t *leaf_extents, int n_steps,
                         const int32_t *path, int n_out,
                         const int32_t *out_modes)
{
    JH_REQUIRE(n_leaves >= 1, "An empty tensor network cannot be contracted.");
    JH_REQUIRE(n_steps >= 0 && n_out >= 0, "Invalid network description.");
    NetworkPlan np;
    np.elem_bytes = elem_bytes;
    np.n_leaves = n_leaves;
    np.out_modes.assign(out_modes, out_modes + n_out);
    const int n_nodes = n_leaves + n_steps;

    std::vector<std::vector<int32_t>> modes(n_nodes);
    std::vector<std::vector<int64_t>> exts(n_nodes);
    std::unordered_map<int32_t, int64_t> extent_of;
    size_t pos = 0;
    for (int i = 0; i < n_leaves; i++) {
        JH_REQUIRE(leaf_ranks[i] >= 0 && leaf_ranks[i] <= kMaxRank, "Leaf rank must be 0..64.");
        modes[i].assign(leaf_modes + pos, leaf_modes + pos + leaf_ranks[i]);
        exts[i].assign(leaf_extents + pos, leaf_extents + pos + leaf_ranks[i]);
        pos += leaf_ranks[i];
        for (size_t j = 0; j < modes[i].size(); j++) {
            auto it = extent_of.find(modes[i][j]);
            JH_REQUIRE(it == extent_of.end() || it->second == exts[i][j],
                       "An index has different extents in different tensors.");
            extent_of[modes[i][j]] = exts[i][j];
        }
    }
    np.leaf_modes.assign(modes.begin(), modes.begin() + n_leaves);
    np.leaf_extents.assign(exts.begin(), exts.begin() + n_leaves);
    for (int32_t m : np.out_modes) {
        JH_REQUIRE(extent_of.count(m), "Output index is not an index of the network.");
        np.out_extents.push_back(extent_of[m]);
    }
    const std::set<int32_t> open(np.out_modes.begin(), np.out_modes.end());

    // How many not-yet-consumed nodes hold each mode (hyperedges count > 2).
    std::unordered_map<int32_t, int> holders;
    for (int i = 0; i < n_leaves; i++)
        for (int32_t m : modes[i]) holders[m]++;

    std::vector<char> consumed(n_nodes, 0);
    std::vector<int> last_use(n_nodes, -1);
    for (int s = 0; s < n_steps; s++) {
        const int a = path[2 * s], b = path[2 * s + 1];
        JH_REQUIRE(a >= 0 && b >= 0 && a < n_leaves + s && b < n_leaves + s && a != b,
                   "Node ID in contraction path pair is invalid.");
        JH_REQUIRE(!consumed[a] && !consumed[b], "A node is contracted twice in the path.");
        consumed[a] = consumed[b] = 1;
        last_use[a] = last_use[b] = s;
    }
    if (n_steps > 0) {
        int unconsumed = 0;
        for (int i = 0; i < n_nodes; i++) unconsumed += !consumed[i];
        JH_REQUIRE(unconsumed == 1, "The path does not contract the network to one tensor.");
    }
    else {
        JH_REQUIRE(n_leaves == 1, "The path does not contract the network to one tensor.");
    }

    ArenaSim arena;
    std::vector<size_t> offset(n_nodes, 0);
    np.steps.resize(n_steps);
    for (int s = 0; s < n_steps; s++) {
        NetworkStep &st = np.steps[s];
        st.a = path[2 * s];
        st.b = path[2 * s + 1];
        const auto &A = modes[st.a], &B = modes[st.b];
        for (int32_t m : A) holders[m]--;
        for (int32_t m : B) holders[m]--;
        // Result: batch (shared and kept) in A's order, then A's free modes,
        // then B's free modes -- the GEMM's natural [batch][left][right].
        std::vector<int32_t> batch, left, right;
        for (int32_t m : A) {
            const bool in_b = find(B, m) >= 0;
            const bool keep = open.count(m) || holders[m] > 0;
            if (in_b && keep)
                batch.push_back(m);
            else if (!in_b)
                left.push_back(m);
        }
        for (int32_t m : B)
            if (find(A, m) < 0) right.push_back(m);
        std::vector<int32_t> out = batch;
        out.insert(out.end(), left.begin(), left.end());
        out.insert(out.end(), right.begin(), right.end());
        st.is_root = (s == n_steps - 1);
        if (st.is_root) {
            std::vector<int32_t> sorted_out = out, sorted_req = np.out_modes;
            std::sort(sorted_out.begin(), sorted_out.end());
            std::sort(sorted_req.begin(), sorted_req.end());
            JH_REQUIRE(sorted_out == sorted_req,
                       "Output indices do not match the indices left after the last step.");
            out = np.out_modes;
        }
        for (int32_t m : out) {
            st.modes.push_back(m);
            st.extents.push_back(extent_of[m]);
            holders[m]++;
        }
        modes[n_leaves + s] = st.modes;
        exts[n_leaves + s] = st.extents;
        st.desc = plan_contraction(elem_bytes, mode, 0, (int)A.size(), A.data(),
                                   exts[st.a].data(), (int)B.size(), B.data(),
                                   exts[st.b].data(), (int)st.modes.size(), st.modes.data());
        np.flops += st.desc.flops;
        np.workspace_bytes = std::max(np.workspace_bytes, st.desc.workspace);
        if (!st.is_root) {
            offset[n_leaves + s] = arena.alloc((size_t)elem_bytes * (size_t)st.desc.size_c);
            st.offset = offset[n_leaves + s];
        }
        for (int id : {st.a, st.b})
            if (id >= n_leaves && last_use[id] == s) arena.release(offset[id]);
    }
    np.arena_bytes = arena.high_water();
    if (n_steps == 0) {
        std::vector<int32_t> perm(np.out_modes.size());
        JH_REQUIRE(np.out_modes.size() == modes[0].size(),
                   "Output indices do not match the indices of the single tensor.");
        for (size_t i = 0; i < perm.size(); i++) {
            perm[i] = find(modes[0], np.out_modes[i]);
            JH_REQUIRE(perm[i] >= 0, "Output indices do not match the indices of the single tensor.");
        }
        np.root_copy = make_permute_plan((int)perm.size(), exts[0].data(), perm.data());
    }
    return np;
}


NetworkPlan plan_dag(int elem_bytes, int mode, int n_leaves, const int32_t *leaf_ranks,
                     const int32_t *leaf_modes, const int64_t *leaf_extents, int n_steps,
                     const int32_t *steps, int n_keep, const int32_t *keep)
{
    JH_REQUIRE(n_leaves >= 1 && n_steps >= 1 && n_keep >= 0, "Invalid contraction DAG.");
    NetworkPlan np;
    np.dag = true;
    np.elem_bytes = elem_bytes;
    np.n_leaves = n_leaves;
    const int n_nodes = n_leaves + n_steps;
    std::vector<std::vector<int32_t>> modes(n_nodes);
    std::vector<std::vector<int64_t>> exts(n_nodes);
    std::unordered_map<int32_t, int64_t> extent_of;
    size_t pos = 0;
    for (int i = 0; i < n_leaves; i++) {
        JH_REQUIRE(leaf_ranks[i] >= 0 && leaf_ranks[i] <= kMaxRank, "Leaf rank must be 0..64.");
        modes[i].assign(leaf_modes + pos, leaf_modes + pos + leaf_ranks[i]);
        exts[i].assign(leaf_extents + pos, leaf_extents + pos + leaf_ranks[i]);
        pos += leaf_ranks[i];
        for (size_t j = 0; j < modes[i].size(); j++) {
            auto it = extent_of.find(modes[i][j]);
            JH_REQUIRE(it == extent_of.end() || it->second == exts[i][j],
                       "An index has different extents in different tensors.");
            extent_of[modes[i][j]] = exts[i][j];
        }
    }
    np.leaf_modes.assign(modes.begin(), modes.begin() + n_leaves);
    np.leaf_extents.assign(exts.begin(), exts.begin() + n_leaves);

    std::vector<int> last_use(n_nodes, -1);
    for (int s = 0; s < n_steps; s++) {
        const int a = steps[2 * s], b = steps[2 * s + 1];
        JH_REQUIRE(a >= 0 && b >= 0 && a < n_leaves + s && b < n_leaves + s && a != b,
                   "Node ID in contraction path pair is invalid.");
        last_use[a] = last_use[b] = s;
    }
    np.kept.assign(n_steps, 0);
    std::vector<int> keep_order;
    for (int i = 0; i < n_keep; i++) {
        const int id = keep[i];
        JH_REQUIRE(id >= n_leaves && id < n_nodes, "Only step results can be kept.");
        JH_REQUIRE(!np.kept[id - n_leaves], "A step result is kept twice.");
        np.kept[id - n_leaves] = 1;
        keep_order.push_back(id - n_leaves);
    }

    ArenaSim arena;
    std::vector<size_t> offset(n_nodes, 0);
    np.steps.resize(n_steps);
    for (int s = 0; s < n_steps; s++) {
        NetworkStep &st = np.steps[s];
        st.a = steps[2 * s];
        st.b = steps[2 * s + 1];
        const auto &A = modes[st.a], &B = modes[st.b];
        for (int32_t m : A)
            if (find(B, m) < 0) st.modes.push_back(m);
        for (int32_t m : B)
            if (find(A, m) < 0) st.modes.push_back(m);
        for (int32_t m : st.modes) st.extents.push_back(extent_of[m]);
        modes[n_leaves + s] = st.modes;
        exts[n_leaves + s] = st.extents;
        st.desc = plan_contraction(elem_bytes, mode, 0, (int)A.size(), A.data(),
                                   exts[st.a].data(), (int)B.size(), B.data(),
                                   exts[st.b].data(), (int)st.modes.size(), st.modes.data());
        np.flops += st.desc.flops;
        np.workspace_bytes = std::max(np.workspace_bytes, st.desc.workspace);
        if (!np.kept[s]) {
            offset[n_leaves + s] = arena.alloc((size_t)elem_bytes * (size_t)st.desc.size_c);
            st.offset = offset[n_leaves + s];
        }
        for (int id : {st.a, st.b})
            if (id >= n_leaves && last_use[id] == s && !np.kept[id - n_leaves])
                arena.release(offset[id]);
        // A result nothing uses is released at once (unless kept).
        if (last_use[n_leaves + s] < 0 && !np.kept[s]) arena.release(st.offset);
    }
    // Kept results: one region after the working arena, in `keep` order.
    size_t top = align256(arena.high_water());
    for (int s : keep_order) {
        np.steps[s].offset = top;
        top += align256(std::max<size_t>(256, (size_t)elem_bytes * (size_t)np.steps[s].desc.size_c));
    }
    np.arena_bytes = top;
    return np;
}

} // namespace jet_hip

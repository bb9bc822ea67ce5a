// network_check.cpp -- CPU checks of network plans (arena layout).
//
// Builds the plan of jet_amd/csrc/network_plan.cpp for a network and
// verifies, before anything runs on a GPU, that every intermediate result
// lies inside the arena, that no two results alive at the same time
// overlap, and that each step's operands are the results of earlier steps
// or leaves.
#include "../../jet_amd/csrc/network_plan.hpp"

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

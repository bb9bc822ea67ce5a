// tuning.cpp -- lookup in the autotuned GEMM kernel table (tuning.hpp).
#include "tuning.hpp"

#include <cmath>
#include <mutex>

namespace jet_hip {

namespace {

// The table measured on gfx950, generated at build time from
// jet_amd/data/gemm_tuning_gfx950.json: rows of {mode, lm, ln, lk, lb,
// kernel, variant}.
const int32_t kBuiltinTable[][kTuneEntryInts] = {
#include "tuning_table.inc"
};

std::vector<TuneEntry> builtin()
{
    std::vector<TuneEntry> v;
    for (const auto &r : kBuiltinTable) {
        if (r[0] < 0) continue; // the generated file's end marker
        v.push_back(TuneEntry{r[0], r[1], r[2], r[3], r[4], r[5], r[6]});
    }
    return v;
}

std::mutex g_mutex;
bool g_loaded = false;
std::vector<TuneEntry> g_table;

void ensure_loaded()
{
    if (!g_loaded) {
        g_table = builtin();
        g_loaded = true;
    }
}

double log2d(int64_t x) { return std::log2(double(x < 1 ? 1 : x)); }

} // namespace

void tuning_set(const std::vector<TuneEntry> &entries)
{
    std::lock_guard<std::mutex> lock(g_mutex);
    g_table = entries;
    g_loaded = true;
}

void tuning_reset()
{
    std::lock_guard<std::mutex> lock(g_mutex);
    g_table = builtin();
    g_loaded = true;
}

std::vector<TuneEntry> tuning_get()
{
    std::lock_guard<std::mutex> lock(g_mutex);
    ensure_loaded();
    return g_table;
}

GemmChoice tuned_gemm_choice(int mode, int64_t M, int64_t N, int64_t K, int64_t batch,
                             int *source)
{
    const double qm = log2d(M), qn = log2d(N), qk = log2d(K), qb = log2d(batch);
    const int im = (int)std::lround(qm), in = (int)std::lround(qn), ik = (int)std::lround(qk),
              ib = (int)std::lround(qb);
    GemmChoice best{};
    double best_d = -1.0;
    bool exact = false;
    {
        std::lock_guard<std::mutex> lock(g_mutex);
        ensure_loaded();
        for (const TuneEntry &e : g_table) {
            if (e.mode != mode) continue;
            const GemmChoice c{e.kernel, e.variant};
            if (!gemm_choice_valid(mode, M, N, K, batch, c)) continue;
            if (e.lm == im && e.ln == in && e.lk == ik && e.lb == ib) {
                best = c;
                best_d = 0.0;
                exact = true;
                break;
            }
            const double d = std::fabs(e.lm - qm) + std::fabs(e.ln - qn) +
                             std::fabs(e.lk - qk) + std::fabs(e.lb - qb);
            if (d <= kTuneMaxDistance && (best_d < 0.0 || d < best_d)) {
                best = c;
                best_d = d;
            }
        }
    }
    if (best_d >= 0.0) {
        if (source) *source = exact ? 1 : 2;
        return best;
    }
    if (source) *source = 0;
    return GemmChoice{choose_gemm_kernel(mode, M, N, K, batch), 0};
}

} // namespace jet_hip

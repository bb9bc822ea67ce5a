// profiler.hip -- per-kernel timing with HIP events (see profiler.hpp).
#include "profiler.hpp"

#include <atomic>
#include <mutex>
#include <vector>

namespace jet_hip {
namespace prof {

namespace {

struct Record {
    int kernel;
    hipEvent_t start, stop;
    double flops, bytes;
};

std::atomic<bool> g_enabled{false};
std::mutex g_mu;
std::vector<Record> g_pending;
std::vector<hipEvent_t> g_free_events;
int64_t g_launches[kNumKernels];
double g_ms[kNumKernels], g_flops[kNumKernels], g_bytes[kNumKernels];

const char *const kNames[kNumKernels] = {
    "permute_bits", "permute_generic", "copy",       "gemm_valu",    "gemm_valu_reduce",
    "gemm_fp32_mfma", "gemm_fp64_mfma", "split_bf16", "gemm_bf16x3_mfma", "add",
    "fill_random",  "reduce_compensated", "convert", "gemm_splitk", "gemm_bf16x3_4m_mfma",
    "split_f16",    "gemm_f16x3_4m_mfma"};

hipEvent_t take_event()
{
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_free_events.empty()) {
        hipEvent_t e = g_free_events.back();
        g_free_events.pop_back();
        return e;
    }
    hipEvent_t e;
    JH_HIP(hipEventCreate(&e));
    return e;
}

// Caller holds g_mu.
void drain_locked()
{
    for (auto &r : g_pending) {
        JH_HIP(hipEventSynchronize(r.stop));
        float ms = 0.f;
        JH_HIP(hipEventElapsedTime(&ms, r.start, r.stop));
        g_launches[r.kernel] += 1;
        g_ms[r.kernel] += ms;
        g_flops[r.kernel] += r.flops;
        g_bytes[r.kernel] += r.bytes;
        g_free_events.push_back(r.start);
        g_free_events.push_back(r.stop);
    }
    g_pending.clear();
}

} // namespace

const char *kernel_name(int k) { return (k >= 0 && k < kNumKernels) ? kNames[k] : ""; }

void set_enabled(bool on) { g_enabled.store(on); }

bool enabled() { return g_enabled.load(std::memory_order_relaxed); }

void reset()
{
    std::lock_guard<std::mutex> lk(g_mu);
    drain_locked();
    for (int k = 0; k < kNumKernels; k++) {
        g_launches[k] = 0;
        g_ms[k] = g_flops[k] = g_bytes[k] = 0.0;
    }
}

void read(int64_t *launches, double *ms, double *flops, double *bytes)
{
    std::lock_guard<std::mutex> lk(g_mu);
    drain_locked();
    for (int k = 0; k < kNumKernels; k++) {
        if (launches) launches[k] = g_launches[k];
        if (ms) ms[k] = g_ms[k];
        if (flops) flops[k] = g_flops[k];
        if (bytes) bytes[k] = g_bytes[k];
    }
}

Scope::Scope(int kernel, hipStream_t s, double flops, double bytes)
    : kernel_(kernel), stream_(s), flops_(flops), bytes_(bytes)
{
    if (!enabled()) return;
    start_ = take_event();
    JH_HIP(hipEventRecord(start_, stream_));
}

Scope::~Scope()
{
    if (!start_) return;
    try {
        hipEvent_t stop = take_event();
        if (hipEventRecord(stop, stream_) != hipSuccess) return;
        std::lock_guard<std::mutex> lk(g_mu);
        g_pending.push_back(Record{kernel_, start_, stop, flops_, bytes_});
    }
    catch (...) {
    }
}

} // namespace prof
} // namespace jet_hip

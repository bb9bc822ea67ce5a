// jet_hip_api.hip -- the extern "C" boundary of libjet_hip.so (include/jet_hip.h).
//
// Every entry point catches all C++ exceptions and turns them into a status
// code plus a thread-local message (Jet::Exception cannot cross a C ABI).
#include "common.hpp"
#include "contract_plan.hpp"
#include "gemm.hpp"
#include "misc.hpp"
#include "network_plan.hpp"
#include "permute_plan.hpp"
#include "profiler.hpp"
#include "tuning.hpp"

#include <algorithm>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

using namespace jet_hip;

struct jet_hip_ctx_s {
    int device = 0;
    hipStream_t stream = nullptr;
};

struct jet_hip_plan_s {
    jet_hip_ctx_s *ctx = nullptr;
    ContractionDesc d;
};

struct jet_hip_network_s {
    jet_hip_ctx_s *ctx = nullptr;
    NetworkPlan plan;
    char *arena = nullptr;
    char *workspace = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
};

struct jet_hip_batch_s {
    jet_hip_ctx_s *ctx = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
};

namespace {

thread_local std::string g_last_error;

int fail(int status, const std::string &msg)
{
    g_last_error = msg;
    return status;
}

template <class F> int guarded(F &&f)
{
    try {
        f();
        return JET_HIP_SUCCESS;
    }
    catch (const Error &e) {
        return fail(e.status, e.what());
    }
    catch (const std::bad_alloc &) {
        return fail(JET_HIP_ERROR_OUT_OF_MEMORY, "Error in Jet: host allocation failed.");
    }
    catch (const std::exception &e) {
        return fail(JET_HIP_ERROR_INTERNAL, std::string("Error in Jet: ") + e.what());
    }
    catch (...) {
        return fail(JET_HIP_ERROR_INTERNAL, "Error in Jet: unknown exception.");
    }
}

int elem_bytes_of(jet_hip_dtype_t dt)
{
    if (dt == JET_HIP_C64) return 8;
    if (dt == JET_HIP_C128) return 16;
    JH_FAIL(JET_HIP_ERROR_INVALID_VALUE, "Unsupported data type.");
}

hipStream_t pick_stream(jet_hip_ctx_s *ctx, void *stream)
{
    return stream ? static_cast<hipStream_t>(stream) : ctx->stream;
}

void set_device(jet_hip_ctx_s *ctx)
{
    JH_REQUIRE(ctx != nullptr, "Null context.");
    JH_HIP(hipSetDevice(ctx->device));
}

int mode_of(jet_hip_dtype_t dt, jet_hip_precision_t p)
{
    switch (p) {
    case JET_HIP_PREC_DEFAULT:
        return dt == JET_HIP_C64 ? 0 : 2;
    case JET_HIP_PREC_FP32:
        return 0;
    case JET_HIP_PREC_BF16X3:
        return 1;
    case JET_HIP_PREC_FP64:
        return 2;
    case JET_HIP_PREC_BF16X3_4M:
        return 3;
    case JET_HIP_PREC_F16X3_4M:
        return 4;
    }
    JH_FAIL(JET_HIP_ERROR_INVALID_VALUE, "Unknown precision.");
}

// Row-major permutation plan from mode lists (layout converts col-major).
PermutePlan permute_plan_from_modes(int layout, int rank, const int32_t *in_modes,
                                    const int64_t *extents, const int32_t *out_modes)
{
    std::vector<int32_t> im(in_modes, in_modes + rank), om(out_modes, out_modes + rank);
    std::vector<int64_t> ex(extents, extents + rank);
    if (layout == JET_HIP_COL_MAJOR) {
        std::reverse(im.begin(), im.end());
        std::reverse(om.begin(), om.end());
        std::reverse(ex.begin(), ex.end());
    }
    std::vector<int32_t> perm(rank);
    for (int i = 0; i < rank; i++) {
        int pos = -1;
        for (int j = 0; j < rank; j++)
            if (im[j] == om[i]) {
                JH_REQUIRE(pos < 0, "Duplicate existing indices found. Please ensure indices are unique.");
                pos = j;
            }
        JH_REQUIRE(pos >= 0, "New indices are an invalid permutation of the existing indices");
        perm[i] = pos;
    }
    return make_permute_plan(rank, ex.data(), perm.data());
}

void run_contraction(const ContractionDesc &d, const void *a, const void *b, void *c,
                     char *ws, hipStream_t s)
{
    const int eb = d.elem_bytes;
    if (d.fused_small) {
        launch_fused_small(d.fs, eb, a, b, c, s);
        return;
    }
    const void *At = a;
    const void *Bt = b;
    if (d.need_a && !d.fuse_a) {
        launch_permute(d.pa, eb, a, ws + d.off_at, s);
        At = ws + d.off_at;
    }
    if (d.need_b && !d.fuse_b) {
        launch_permute(d.pb, eb, b, ws + d.off_bt, s);
        Bt = ws + d.off_bt;
    }
    void *Ct = d.need_c ? static_cast<void *>(ws + d.off_ct) : c;
    // K split into batched chunks: the GEMM writes each chunk's result,
    // then they are summed into Ct.
    void *Cg = d.chunks > 1 ? static_cast<void *>(ws + d.off_chunks) : Ct;
    if (d.gemm_kernel == kGemmBf16x3 || d.gemm_kernel == kGemmBf16x4m ||
        d.gemm_kernel == kGemmF16x4m) {
        const int planes = d.gemm_kernel == kGemmBf16x3 ? 6 : 4;
        SplitOperand sa{ws + d.off_sa, ceil_div(d.M, kSplitRowPad) * kSplitRowPad,
                        ceil_div(d.K, kSplitKPad) * kSplitKPad, d.batch, d.split_a_plane, planes,
                        planes == 4 ? 1 : 0};
        SplitOperand sb{ws + d.off_sb, ceil_div(d.N, kSplitRowPad) * kSplitRowPad,
                        ceil_div(d.K, kSplitKPad) * kSplitKPad, d.batch, d.split_b_plane, planes,
                        planes == 4 ? 1 : 0};
        if (d.gemm_kernel == kGemmF16x4m) {
            sa.rowmax = reinterpret_cast<uint32_t *>(ws + d.off_rmax_a);
            sa.scale = reinterpret_cast<float *>(ws + d.off_scale_a);
            sb.rowmax = reinterpret_cast<uint32_t *>(ws + d.off_rmax_b);
            sb.scale = reinterpret_cast<float *>(ws + d.off_scale_b);
            if (d.fuse_a)
                launch_split_f16_gather(a, d.ga, sa, s);
            else
                launch_split_f16(At, d.M, d.K, d.K, d.M * d.K, d.batch, sa, s);
            if (d.fuse_b)
                launch_split_f16_gather(b, d.gb, sb, s);
            else
                launch_split_f16(Bt, d.N, d.K, d.K, d.N * d.K, d.batch, sb, s);
            launch_gemm_f16x4m(sa, sb, d.K, Cg, d.M, d.N, d.N, d.M * d.N, d.batch, false, false, s,
                               d.gemm_variant);
        }
        else {
            launch_split_bf16(At, d.M, d.K, d.K, d.M * d.K, d.batch, sa, s);
            launch_split_bf16(Bt, d.N, d.K, d.K, d.N * d.K, d.batch, sb, s);
            if (planes == 6)
                launch_gemm_bf16x3(sa, sb, d.K, Cg, d.M, d.N, d.N, d.M * d.N, d.batch, false,
                                   false, s);
            else
                launch_gemm_bf16x4m(sa, sb, d.K, Cg, d.M, d.N, d.N, d.M * d.N, d.batch, false,
                                    false, s, d.gemm_variant);
        }
    }
    else {
        GemmArgs g;
        g.A = At;
        g.B = Bt;
        g.C = Cg;
        g.M = d.M;
        g.N = d.N;
        g.K = d.K;
        g.batch = d.batch;
        g.lda = d.K;
        g.ldb = d.K;
        g.ldc = d.N;
        g.sA = d.M * d.K;
        g.sB = d.N * d.K;
        g.sC = d.M * d.N;
        g.variant = d.gemm_variant;
        launch_gemm(d.gemm_kernel, eb, g, s, ws + d.off_partials);
    }
    if (d.chunks > 1) launch_sum_chunks(eb, Cg, Ct, d.size_c, d.chunks, s);
    if (d.need_c) launch_permute(d.pc, eb, Ct, c, s);
}

void run_network(jet_hip_network_s *net, const void *const *leaves, void *out, hipStream_t s)
{
    const NetworkPlan &np = net->plan;
    if (np.steps.empty()) {
        launch_permute(np.root_copy, np.elem_bytes, leaves[0], out, s);
        return;
    }
    std::vector<const void *> node(np.n_leaves + np.steps.size(), nullptr);
    for (int i = 0; i < np.n_leaves; i++) {
        JH_REQUIRE(leaves[i] != nullptr, "Null leaf tensor.");
        node[i] = leaves[i];
    }
    for (size_t k = 0; k < np.steps.size(); k++) {
        const NetworkStep &st = np.steps[k];
        void *c = st.is_root ? out : static_cast<void *>(net->arena + st.offset);
        run_contraction(st.desc, node[st.a], node[st.b], c, net->workspace, s);
        node[np.n_leaves + k] = c;
    }
}

// hipMalloc that first returns the stream-ordered pool's cached blocks to
// the device when the request does not fit.
void *device_alloc(size_t bytes, int device)
{
    void *p = nullptr;
    hipError_t e = hipMalloc(&p, bytes);
    if (e == hipErrorOutOfMemory) {
        (void)hipGetLastError();
        JH_HIP(hipDeviceSynchronize());
        hipMemPool_t pool;
        JH_HIP(hipDeviceGetDefaultMemPool(&pool, device));
        JH_HIP(hipMemPoolTrimTo(pool, 0));
        e = hipMalloc(&p, bytes);
    }
    JH_HIP(e);
    return p;
}

void free_graph(jet_hip_network_s *net)
{
    if (net->graph_exec) (void)hipGraphExecDestroy(net->graph_exec);
    if (net->graph) (void)hipGraphDestroy(net->graph);
    net->graph_exec = nullptr;
    net->graph = nullptr;
}

} // namespace

// Workspace of jet_hip_contract_noalloc: one fixed device array per GPU that
// the library runs on, held until the process ends.  Host code never uses
// the array's name as an address (on the host it is only a stand-in for the
// symbol): scratch_setup looks up each device's own copy once, with that
// device current, and every later use goes through that pointer.
static_assert(JET_HIP_NOALLOC_SCRATCH_BYTES % 256 == 0, "scratch size");
static __device__ __align__(256) char scratch[JET_HIP_NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mutex;
char *g_scratch_ptr[kMaxScratchDevices] = {};

// Device address of this device's scratch array (the device must be
// current); looked up on the first call per device.
char *scratch_setup(int device)
{
    JH_REQUIRE(device >= 0 && device < kMaxScratchDevices, "Invalid device ordinal.");
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[device]) {
        void *p = nullptr;
        JH_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        JH_REQUIRE(p != nullptr, "No device address for the no-allocation scratch array.");
        g_scratch_ptr[device] = static_cast<char *>(p);
    }
    return g_scratch_ptr[device];
}

} // namespace

extern "C" {

int jet_hip_get_version(int *major, int *minor, int *patch)
{
    if (major) *major = JET_HIP_VERSION_MAJOR;
    if (minor) *minor = JET_HIP_VERSION_MINOR;
    if (patch) *patch = JET_HIP_VERSION_PATCH;
    return JET_HIP_SUCCESS;
}

const char *jet_hip_last_error(void) { return g_last_error.c_str(); }

int jet_hip_ctx_create(int device, jet_hip_ctx_t *ctx)
{
    return guarded([&] {
        JH_REQUIRE(ctx != nullptr, "Null output pointer.");
        int count = 0;
        JH_HIP(hipGetDeviceCount(&count));
        JH_REQUIRE(device >= 0 && device < count, "Invalid device ordinal.");
        JH_HIP(hipSetDevice(device));
        auto *c = new jet_hip_ctx_s;
        c->device = device;
        hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
        if (e != hipSuccess) {
            delete c;
            JH_HIP(e);
        }
        // Keep freed stream-ordered allocations cached for reuse.
        hipMemPool_t pool;
        if (hipDeviceGetDefaultMemPool(&pool, device) == hipSuccess) {
            uint64_t threshold = UINT64_MAX;
            (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &threshold);
        }
        try {
            (void)scratch_setup(device);
        }
        catch (...) {
            (void)hipStreamDestroy(c->stream);
            delete c;
            throw;
        }
        *ctx = c;
    });
}

int jet_hip_ctx_destroy(jet_hip_ctx_t ctx)
{
    return guarded([&] {
        if (!ctx) return;
        set_device(ctx);
        JH_HIP(hipStreamSynchronize(ctx->stream));
        JH_HIP(hipStreamDestroy(ctx->stream));
        delete ctx;
    });
}

int jet_hip_ctx_device(jet_hip_ctx_t ctx, int *device)
{
    return guarded([&] {
        JH_REQUIRE(ctx && device, "Null argument.");
        *device = ctx->device;
    });
}

int jet_hip_ctx_stream(jet_hip_ctx_t ctx, void **stream)
{
    return guarded([&] {
        JH_REQUIRE(ctx && stream, "Null argument.");
        *stream = ctx->stream;
    });
}

int jet_hip_malloc(jet_hip_ctx_t ctx, size_t bytes, void **ptr)
{
    return guarded([&] {
        set_device(ctx);
        JH_REQUIRE(ptr != nullptr, "Null output pointer.");
        *ptr = nullptr;
        if (bytes == 0) bytes = 1;
        // Stream-ordered pool allocation on the context stream, made visible
        // to every stream before returning.  The pool keeps freed blocks
        // cached; when a request does not fit, the cache is released to the
        // device and the request made once more.
        hipError_t e = hipMallocAsync(ptr, bytes, ctx->stream);
        if (e == hipErrorOutOfMemory) {
            (void)hipGetLastError();
            JH_HIP(hipDeviceSynchronize());
            hipMemPool_t pool;
            JH_HIP(hipDeviceGetDefaultMemPool(&pool, ctx->device));
            JH_HIP(hipMemPoolTrimTo(pool, 0));
            e = hipMallocAsync(ptr, bytes, ctx->stream);
        }
        JH_HIP(e);
        JH_HIP(hipStreamSynchronize(ctx->stream));
    });
}

int jet_hip_free(jet_hip_ctx_t ctx, void *ptr)
{
    return guarded([&] {
        set_device(ctx);
        // Ordered after all work queued on the context stream; callers that
        // used other streams synchronize them first.
        if (ptr) JH_HIP(hipFreeAsync(ptr, ctx->stream));
    });
}

int jet_hip_memcpy(jet_hip_ctx_t ctx, void *dst, const void *src, size_t bytes, int kind,
                   void *stream, int async)
{
    return guarded([&] {
        set_device(ctx);
        if (bytes == 0) return;
        hipMemcpyKind k = kind == 0   ? hipMemcpyHostToDevice
                          : kind == 1 ? hipMemcpyDeviceToHost
                                      : hipMemcpyDeviceToDevice;
        hipStream_t s = pick_stream(ctx, stream);
        JH_HIP(hipMemcpyAsync(dst, src, bytes, k, s));
        if (!async) JH_HIP(hipStreamSynchronize(s));
    });
}

int jet_hip_memset(jet_hip_ctx_t ctx, void *dst, int value, size_t bytes, void *stream)
{
    return guarded([&] {
        set_device(ctx);
        if (bytes) JH_HIP(hipMemsetAsync(dst, value, bytes, pick_stream(ctx, stream)));
    });
}

int jet_hip_stream_synchronize(jet_hip_ctx_t ctx, void *stream)
{
    return guarded([&] {
        set_device(ctx);
        JH_HIP(hipStreamSynchronize(pick_stream(ctx, stream)));
    });
}

int jet_hip_device_synchronize(jet_hip_ctx_t ctx)
{
    return guarded([&] {
        set_device(ctx);
        JH_HIP(hipDeviceSynchronize());
    });
}

int jet_hip_permute(jet_hip_ctx_t ctx, jet_hip_dtype_t dtype, jet_hip_layout_t layout, int rank,
                    const int32_t *in_modes, const int64_t *extents, const int32_t *out_modes,
                    const void *in, void *out, void *stream)
{
    return guarded([&] {
        set_device(ctx);
        const int eb = elem_bytes_of(dtype);
        JH_REQUIRE(rank >= 1, "Number of indices cannot be zero.");
        JH_REQUIRE(rank <= kMaxRank, "Permutation rank exceeds 64.");
        PermutePlan p = permute_plan_from_modes(layout, rank, in_modes, extents, out_modes);
        launch_permute(p, eb, in, out, pick_stream(ctx, stream));
    });
}

int jet_hip_contraction_plan_create(jet_hip_ctx_t ctx, jet_hip_dtype_t dtype,
                                    jet_hip_precision_t precision, jet_hip_layout_t layout,
                                    int rank_a, const int32_t *modes_a, const int64_t *extents_a,
                                    int rank_b, const int32_t *modes_b, const int64_t *extents_b,
                                    int rank_c, const int32_t *modes_c, jet_hip_plan_t *plan)
{
    return guarded([&] {
        JH_REQUIRE(ctx && plan, "Null argument.");
        const int eb = elem_bytes_of(dtype);
        auto *p = new jet_hip_plan_s;
        try {
            p->ctx = ctx;
            p->d = plan_contraction(eb, mode_of(dtype, precision), layout, rank_a, modes_a,
                                    extents_a, rank_b, modes_b, extents_b, rank_c, modes_c);
        }
        catch (...) {
            delete p;
            throw;
        }
        *plan = p;
    });
}

int jet_hip_plan_workspace_size(jet_hip_plan_t plan, size_t *bytes)
{
    return guarded([&] {
        JH_REQUIRE(plan && bytes, "Null argument.");
        *bytes = plan->d.workspace;
    });
}

int jet_hip_plan_info(jet_hip_plan_t plan, double *flops, int64_t *m, int64_t *n, int64_t *k,
                      int64_t *batch, int *gemm_kernel)
{
    return guarded([&] {
        JH_REQUIRE(plan, "Null plan.");
        if (flops) *flops = plan->d.flops;
        if (m) *m = plan->d.M;
        if (n) *n = plan->d.N;
        if (k) *k = plan->d.K * plan->d.chunks;
        if (batch) *batch = plan->d.batch / plan->d.chunks;
        if (gemm_kernel) *gemm_kernel = plan->d.gemm_kernel;
    });
}

int jet_hip_plan_destroy(jet_hip_plan_t plan)
{
    delete plan;
    return JET_HIP_SUCCESS;
}

int jet_hip_contract(jet_hip_plan_t plan, const void *a, const void *b, void *c, void *workspace,
                     size_t workspace_bytes, void *stream)
{
    return guarded([&] {
        JH_REQUIRE(plan != nullptr, "Null plan.");
        jet_hip_ctx_s *ctx = plan->ctx;
        set_device(ctx);
        const ContractionDesc &d = plan->d;
        JH_REQUIRE(a && b && c, "Null tensor data.");
        hipStream_t s = pick_stream(ctx, stream);
        char *ws = static_cast<char *>(workspace);
        bool pooled = false;
        if (d.workspace > 0 && ws == nullptr) {
            void *p = nullptr;
            JH_HIP(hipMallocAsync(&p, d.workspace, s));
            ws = static_cast<char *>(p);
            pooled = true;
        }
        else if (d.workspace > 0) {
            JH_REQUIRE(workspace_bytes >= d.workspace, "Workspace is smaller than the plan needs.");
        }
        try {
            run_contraction(d, a, b, c, ws, s);
        }
        catch (...) {
            if (pooled) (void)hipFreeAsync(ws, s);
            throw;
        }
        if (pooled) JH_HIP(hipFreeAsync(ws, s));
    });
}

int jet_hip_contract_noalloc(jet_hip_plan_t plan, const void *a, const void *b, void *c,
                             void *stream)
{
    return guarded([&] {
        JH_REQUIRE(plan != nullptr, "Null plan.");
        jet_hip_ctx_s *ctx = plan->ctx;
        set_device(ctx);
        const ContractionDesc &d = plan->d;
        JH_REQUIRE(a && b && c, "Null tensor data.");
        if (d.workspace > JET_HIP_NOALLOC_SCRATCH_BYTES) {
            const std::string msg = "the contraction needs " + std::to_string(d.workspace) +
                                    " bytes of workspace; jet_hip_contract_noalloc holds " +
                                    std::to_string(size_t(JET_HIP_NOALLOC_SCRATCH_BYTES)) + ".";
            JH_FAIL(JET_HIP_ERROR_INVALID_VALUE, msg.c_str());
        }
        // Same kernels as jet_hip_contract, with the device's scratch array
        // as the workspace.  Every part of the workspace is written before it
        // is read within one contraction, so nothing needs clearing.
        char *ws = d.workspace > 0 ? scratch_setup(ctx->device) : nullptr;
        run_contraction(d, a, b, c, ws, pick_stream(ctx, stream));
    });
}

int jet_hip_gemm_choice(jet_hip_dtype_t dtype, jet_hip_precision_t precision, int64_t m,
                        int64_t n, int64_t k, int64_t batch, int *kernel, int *variant,
                        int *source)
{
    return guarded([&] {
        (void)elem_bytes_of(dtype);
        JH_REQUIRE(m >= 1 && n >= 1 && k >= 1 && batch >= 1, "GEMM extents must be positive.");
        const GemmChoice c = tuned_gemm_choice(mode_of(dtype, precision), m, n, k, batch, source);
        if (kernel) *kernel = c.kernel;
        if (variant) *variant = c.variant;
    });
}

int jet_hip_tuning_set(int n, const int32_t *entries)
{
    return guarded([&] {
        if (n < 0) {
            tuning_reset();
            return;
        }
        JH_REQUIRE(n == 0 || entries != nullptr, "Null tuning entries.");
        std::vector<TuneEntry> v;
        for (int i = 0; i < n; i++) {
            const int32_t *r = entries + kTuneEntryInts * i;
            JH_REQUIRE(r[0] >= 0 && r[0] <= 4, "Unknown precision mode in a tuning entry.");
            JH_REQUIRE(r[5] >= 0 && r[5] <= kGemmF16x4m && r[6] >= 0 &&
                           r[6] <= gemm_variants(r[5]),
                       "Unknown GEMM kernel or variant in a tuning entry.");
            v.push_back(TuneEntry{r[0], r[1], r[2], r[3], r[4], r[5], r[6]});
        }
        tuning_set(v);
    });
}

int jet_hip_tuning_get(int capacity, int32_t *entries, int *n)
{
    return guarded([&] {
        JH_REQUIRE(n != nullptr, "Null output pointer.");
        const std::vector<TuneEntry> v = tuning_get();
        *n = (int)v.size();
        for (int i = 0; i < capacity && i < (int)v.size() && entries; i++) {
            int32_t *r = entries + kTuneEntryInts * i;
            const TuneEntry &e = v[i];
            r[0] = e.mode, r[1] = e.lm, r[2] = e.ln, r[3] = e.lk, r[4] = e.lb;
            r[5] = e.kernel, r[6] = e.variant;
        }
    });
}

int jet_hip_autotune_gemm(jet_hip_ctx_t ctx, jet_hip_dtype_t dtype,
                          jet_hip_precision_t precision, int64_t m, int64_t n, int64_t k,
                          int64_t batch, int reps, int capacity, int32_t *kernels,
                          int32_t *variants, float *ms, int *count)
{
    return guarded([&] {
        set_device(ctx);
        const int eb = elem_bytes_of(dtype);
        const int mode = mode_of(dtype, precision);
        JH_REQUIRE(m >= 1 && n >= 1 && k >= 1 && batch >= 1 && reps >= 1,
                   "GEMM extents and repetitions must be positive.");
        JH_REQUIRE(count != nullptr, "Null output pointer.");
        hipStream_t s = ctx->stream;
        // Every valid candidate, each a plan with no permutations.
        std::vector<ContractionDesc> cands;
        size_t ws_max = 0;
        // MFMA kernels first: they set the bar that slower candidates are
        // measured against.
        const int kernels_all[] = {kGemmF16x4m, kGemmBf16x4m, kGemmBf16x3, kGemmF32Mfma,
                                   kGemmF64Mfma, kGemmSmall,  kGemmSplitK, kGemmSmallRed};
        // Forms not timed (far slower than another candidate there, and slow
        // to time): one-output-per-block reductions (SmallRed, split-K)
        // beyond 2^20 outputs; the one-thread-per-output VALU form beyond
        // K = 2^16; short-K row threads with fewer than 256 rows or more
        // than 4096 columns (each thread walks a whole row); block-tiled MFMA
        // kernels with fewer tiles than CUs and K beyond 2^16 (each tile
        // walks K alone); and every VALU form beyond 2^36 flops where an
        // MFMA kernel remains.
        const double flops = 8.0 * double(batch) * double(m) * double(n) * double(k);
        const int64_t outputs = batch * m * n;
        const bool few_tiles_long_k =
            k > (1 << 16) && batch * ceil_div(m, 64) * ceil_div(n, 64) < 256;
        const bool mfma_ok =
            !few_tiles_long_k &&
            (gemm_choice_valid(mode, m, n, k, batch, GemmChoice{kGemmF32Mfma, 1}) ||
             gemm_choice_valid(mode, m, n, k, batch, GemmChoice{kGemmF64Mfma, 1}));
        for (int kern : kernels_all)
            for (int v = 1; v <= gemm_variants(kern); v++) {
                const GemmChoice c{kern, v};
                if (!gemm_choice_valid(mode, m, n, k, batch, c)) continue;
                const bool valu = kern == kGemmSmall || kern == kGemmSmallRed ||
                                  kern == kGemmSplitK;
                if ((kern == kGemmSmallRed || kern == kGemmSplitK) && outputs > (1 << 20))
                    continue;
                if (kern == kGemmSmall && v == 1 && k > (1 << 16)) continue;
                if (kern == kGemmSmall && v == 3 && (m < 256 || n > 4096)) continue;
                if (!valu && few_tiles_long_k) continue;
                if (valu && mfma_ok && flops > 68719476736.0) continue;
                ContractionDesc d;
                d.elem_bytes = eb;
                d.mode = mode;
                d.batch = batch, d.M = m, d.N = n, d.K = k;
                d.size_a = batch * m * k, d.size_b = batch * n * k, d.size_c = batch * m * n;
                d.gemm_kernel = kern;
                d.gemm_variant = v;
                d.flops = 8.0 * double(batch) * double(m) * double(n) * double(k);
                layout_workspace(d);
                ws_max = std::max(ws_max, d.workspace);
                cands.push_back(d);
            }
        JH_REQUIRE(!cands.empty(), "No GEMM kernel candidate for this shape.");
        const size_t ba = (size_t)eb * batch * m * k, bb = (size_t)eb * batch * n * k,
                     bc = (size_t)eb * batch * m * n;
        std::vector<void *> bufs(4, nullptr);
        const size_t sizes[4] = {ba, bb, bc, ws_max};
        // The operands go back to the device, not to the pool's cache: a
        // sweep over many shapes would otherwise keep each size cached.
        auto release = [&] {
            for (void *p : bufs)
                if (p) (void)hipFreeAsync(p, s);
            (void)hipStreamSynchronize(s);
            hipMemPool_t pool;
            if (hipDeviceGetDefaultMemPool(&pool, ctx->device) == hipSuccess)
                (void)hipMemPoolTrimTo(pool, 0);
        };
        std::vector<hipEvent_t> ev(2, nullptr);
        try {
            for (int i = 0; i < 4; i++)
                if (sizes[i]) JH_HIP(hipMallocAsync(&bufs[i], sizes[i], s));
            launch_fill(eb, bufs[0], batch * m * k, 0x5eedull, 2, s);
            launch_fill(eb, bufs[1], batch * n * k, 0x5eedull + 1, 2, s);
            JH_HIP(hipEventCreate(&ev[0]));
            JH_HIP(hipEventCreate(&ev[1]));
            int out = 0;
            float best = -1.f;
            for (const ContractionDesc &d : cands) {
                char *ws = static_cast<char *>(bufs[3]);
                // The warm-up run is timed too: a candidate more than four
                // times slower than the best so far keeps that time and
                // skips its repetitions.
                JH_HIP(hipEventRecord(ev[0], s));
                run_contraction(d, bufs[0], bufs[1], bufs[2], ws, s);
                JH_HIP(hipEventRecord(ev[1], s));
                JH_HIP(hipEventSynchronize(ev[1]));
                float t = 0.f;
                JH_HIP(hipEventElapsedTime(&t, ev[0], ev[1]));
                if (best < 0.f || t <= 4.f * best) {
                    JH_HIP(hipEventRecord(ev[0], s));
                    for (int r = 0; r < reps; r++)
                        run_contraction(d, bufs[0], bufs[1], bufs[2], ws, s);
                    JH_HIP(hipEventRecord(ev[1], s));
                    JH_HIP(hipEventSynchronize(ev[1]));
                    JH_HIP(hipEventElapsedTime(&t, ev[0], ev[1]));
                    t /= float(reps);
                    if (best < 0.f || t < best) best = t;
                }
                if (out < capacity) {
                    if (kernels) kernels[out] = d.gemm_kernel;
                    if (variants) variants[out] = d.gemm_variant;
                    if (ms) ms[out] = t;
                }
                out++;
            }
            *count = out;
        }
        catch (...) {
            for (hipEvent_t e : ev)
                if (e) (void)hipEventDestroy(e);
            release();
            throw;
        }
        for (hipEvent_t e : ev)
            if (e) JH_HIP(hipEventDestroy(e));
        release();
    });
}

int jet_hip_add(jet_hip_ctx_t ctx, jet_hip_dtype_t dtype, jet_hip_layout_t layout, int rank,
                const int32_t *modes_a, const int64_t *extents_a, const int32_t *modes_b,
                const void *a, const void *b, void *c, void *stream)
{
    return guarded([&] {
        set_device(ctx);
        const int eb = elem_bytes_of(dtype);
        hipStream_t s = pick_stream(ctx, stream);
        int64_t n = 1;
        for (int i = 0; i < rank; i++) n *= extents_a[i];
        if (rank == 0) {
            launch_add(eb, a, b, c, 1, s);
            return;
        }
        // b has modes_b; its extents follow from a's.
        std::vector<int64_t> eb_ext(rank);
        for (int i = 0; i < rank; i++) {
            int pos = -1;
            for (int j = 0; j < rank; j++)
                if (modes_a[j] == modes_b[i]) pos = j;
            JH_REQUIRE(pos >= 0, "Tensor addition with disjoint indices is not supported.");
            eb_ext[i] = extents_a[pos];
        }
        PermutePlan p = permute_plan_from_modes(layout, rank, modes_b, eb_ext.data(), modes_a);
        if (p.kind == PermutePlan::kCopy) {
            launch_add(eb, a, b, c, n, s);
            return;
        }
        void *tmp = nullptr;
        JH_HIP(hipMallocAsync(&tmp, (size_t)eb * n, s));
        try {
            launch_permute(p, eb, b, tmp, s);
            launch_add(eb, a, tmp, c, n, s);
        }
        catch (...) {
            (void)hipFreeAsync(tmp, s);
            throw;
        }
        JH_HIP(hipFreeAsync(tmp, s));
    });
}

int jet_hip_slice(jet_hip_ctx_t ctx, jet_hip_dtype_t dtype, jet_hip_layout_t layout, int rank,
                  const int64_t *extents, int axis, int64_t value, const void *in, void *out,
                  void *stream)
{
    return guarded([&] {
        set_device(ctx);
        const int eb = elem_bytes_of(dtype);
        JH_REQUIRE(rank >= 1 && axis >= 0 && axis < rank, "Sliced index does not exist.");
        std::vector<int64_t> ex(extents, extents + rank);
        int ax = axis;
        if (layout == JET_HIP_COL_MAJOR) {
            std::reverse(ex.begin(), ex.end());
            ax = rank - 1 - axis;
        }
        JH_REQUIRE(value >= 0 && value < ex[ax], "Slice value is out of range.");
        int64_t pre = 1, post = 1;
        for (int i = 0; i < ax; i++) pre *= ex[i];
        for (int i = ax + 1; i < rank; i++) post *= ex[i];
        const size_t width = (size_t)post * eb;
        const char *src = static_cast<const char *>(in) + (size_t)value * width;
        JH_HIP(hipMemcpy2DAsync(out, width, src, width * (size_t)ex[ax], width, (size_t)pre,
                                hipMemcpyDeviceToDevice, pick_stream(ctx, stream)));
    });
}

int jet_hip_fill_random(jet_hip_ctx_t ctx, jet_hip_dtype_t dtype, size_t n, uint64_t seed,
                        int distribution, void *out, void *stream)
{
    return guarded([&] {
        set_device(ctx);
        launch_fill(elem_bytes_of(dtype), out, (int64_t)n, seed, distribution,
                    pick_stream(ctx, stream));
    });
}

int jet_hip_reduce_compensated(jet_hip_ctx_t ctx, jet_hip_dtype_t dtype, size_t n, const void *x,
                               double *out, void *stream)
{
    return guarded([&] {
        set_device(ctx);
        JH_REQUIRE(out != nullptr, "Null output.");
        launch_reduce(elem_bytes_of(dtype), x, (int64_t)n, out, pick_stream(ctx, stream));
    });
}

int jet_hip_convert(jet_hip_ctx_t ctx, jet_hip_dtype_t src_dtype, jet_hip_dtype_t dst_dtype,
                    size_t n, const void *in, void *out, void *stream)
{
    return guarded([&] {
        set_device(ctx);
        launch_convert(elem_bytes_of(src_dtype), elem_bytes_of(dst_dtype), in, out, (int64_t)n,
                       pick_stream(ctx, stream));
    });
}

int jet_hip_mfma_ceiling(jet_hip_ctx_t ctx, int rounds, int reps, double *tflops,
                         double *clock_ghz)
{
    return guarded([&] {
        set_device(ctx);
        JH_REQUIRE(tflops != nullptr && clock_ghz != nullptr, "Null output pointer.");
        JH_REQUIRE(rounds > 0 && rounds % 4 == 0 && reps > 0,
                   "rounds must be a positive multiple of 4 and reps positive.");
        int device = 0, cus = 0, wall_khz = 0;
        JH_HIP(hipGetDevice(&device));
        JH_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device));
        JH_HIP(hipDeviceGetAttribute(&wall_khz, hipDeviceAttributeWallClockRate, device));
        JH_REQUIRE(cus > 0 && wall_khz > 0, "Device attributes unavailable.");
        const int blocks = 4 * cus; // four work-groups of four waves per CU
        hipStream_t s = pick_stream(ctx, nullptr);
        float *sink = nullptr;
        long long *clocks = nullptr;
        hipEvent_t e0 = nullptr, e1 = nullptr;
        std::vector<long long> host(size_t(8) * blocks);
        auto release = [&] {
            if (e0) (void)hipEventDestroy(e0);
            if (e1) (void)hipEventDestroy(e1);
            if (sink) (void)hipFree(sink);
            if (clocks) (void)hipFree(clocks);
        };
        try {
            JH_HIP(hipMalloc(&sink, sizeof(float) * 256 * size_t(blocks)));
            JH_HIP(hipMalloc(&clocks, sizeof(long long) * host.size()));
            JH_HIP(hipEventCreate(&e0));
            JH_HIP(hipEventCreate(&e1));
            launch_mfma_loop(sink, clocks, blocks, rounds, s); // warm-up
            JH_HIP(hipEventRecord(e0, s));
            for (int r = 0; r < reps; r++) launch_mfma_loop(sink, clocks, blocks, rounds, s);
            JH_HIP(hipEventRecord(e1, s));
            JH_HIP(hipEventSynchronize(e1));
            float ms = 0.f;
            JH_HIP(hipEventElapsedTime(&ms, e0, e1));
            JH_HIP(hipMemcpy(host.data(), clocks, sizeof(long long) * host.size(),
                             hipMemcpyDeviceToHost));
            double shader = 0.0, wall = 0.0;
            for (int w = 0; w < 4 * blocks; w++) {
                shader += double(host[2 * w]);
                wall += double(host[2 * w + 1]);
            }
            // 16 * 16 * 32 multiply-adds per MFMA, eight per round and wave
            const double flop = 2.0 * 16 * 16 * 32 * 8.0 * double(rounds) * 4.0 * blocks * reps;
            *tflops = flop / (double(ms) * 1e-3) / 1e12;
            *clock_ghz = wall > 0 ? shader / (wall / (double(wall_khz) * 1e3)) / 1e9 : 0.0;
        }
        catch (...) {
            release();
            throw;
        }
        release();
    });
}

int jet_hip_event_create(jet_hip_ctx_t ctx, void **event)
{
    return guarded([&] {
        set_device(ctx);
        JH_REQUIRE(event != nullptr, "Null output pointer.");
        hipEvent_t e;
        JH_HIP(hipEventCreate(&e));
        *event = e;
    });
}

int jet_hip_event_record(jet_hip_ctx_t ctx, void *event, void *stream)
{
    return guarded([&] {
        set_device(ctx);
        JH_HIP(hipEventRecord(static_cast<hipEvent_t>(event), pick_stream(ctx, stream)));
    });
}

int jet_hip_event_elapsed_ms(void *start, void *stop, float *ms)
{
    return guarded([&] {
        JH_REQUIRE(start && stop && ms, "Null argument.");
        JH_HIP(hipEventSynchronize(static_cast<hipEvent_t>(stop)));
        JH_HIP(hipEventElapsedTime(ms, static_cast<hipEvent_t>(start),
                                   static_cast<hipEvent_t>(stop)));
    });
}

int jet_hip_event_destroy(void *event)
{
    return guarded([&] {
        if (event) JH_HIP(hipEventDestroy(static_cast<hipEvent_t>(event)));
    });
}

int jet_hip_profiler_enable(int enable)
{
    prof::set_enabled(enable != 0);
    return JET_HIP_SUCCESS;
}

int jet_hip_profiler_reset(void)
{
    return guarded([&] { prof::reset(); });
}

int jet_hip_profiler_read(int capacity, int64_t *launches, double *ms, double *flops,
                          double *bytes, int *count)
{
    return guarded([&] {
        JH_REQUIRE(capacity >= prof::kNumKernels, "Profiler buffers are too small.");
        prof::read(launches, ms, flops, bytes);
        if (count) *count = prof::kNumKernels;
    });
}

const char *jet_hip_profiler_kernel_name(int kernel) { return prof::kernel_name(kernel); }

int jet_hip_network_create(jet_hip_ctx_t ctx, jet_hip_dtype_t dtype, jet_hip_precision_t precision,
                           int n_leaves, const int32_t *leaf_ranks, const int32_t *leaf_modes,
                           const int64_t *leaf_extents, int n_steps, const int32_t *path,
                           int n_out, const int32_t *out_modes, jet_hip_network_t *net)
{
    return guarded([&] {
        JH_REQUIRE(ctx && net, "Null argument.");
        set_device(ctx);
        auto *n = new jet_hip_network_s;
        try {
            n->ctx = ctx;
            n->plan = plan_network(elem_bytes_of(dtype), mode_of(dtype, precision), n_leaves,
                                   leaf_ranks, leaf_modes, leaf_extents, n_steps, path, n_out,
                                   out_modes);
            if (n->plan.arena_bytes)
                n->arena = static_cast<char *>(device_alloc(n->plan.arena_bytes, ctx->device));
            if (n->plan.workspace_bytes)
                n->workspace =
                    static_cast<char *>(device_alloc(n->plan.workspace_bytes, ctx->device));
        }
        catch (...) {
            if (n->arena) (void)hipFree(n->arena);
            if (n->workspace) (void)hipFree(n->workspace);
            delete n;
            throw;
        }
        *net = n;
    });
}

int jet_hip_network_estimate(jet_hip_dtype_t dtype, jet_hip_precision_t precision, int n_leaves,
                             const int32_t *leaf_ranks, const int32_t *leaf_modes,
                             const int64_t *leaf_extents, int n_steps, const int32_t *path,
                             int n_out, const int32_t *out_modes, double *flops,
                             size_t *arena_bytes, size_t *workspace_bytes)
{
    return guarded([&] {
        NetworkPlan np = plan_network(elem_bytes_of(dtype), mode_of(dtype, precision), n_leaves,
                                      leaf_ranks, leaf_modes, leaf_extents, n_steps, path, n_out,
                                      out_modes);
        if (flops) *flops = np.flops;
        if (arena_bytes) *arena_bytes = np.arena_bytes;
        if (workspace_bytes) *workspace_bytes = np.workspace_bytes;
    });
}

int jet_hip_dag_create(jet_hip_ctx_t ctx, jet_hip_dtype_t dtype, jet_hip_precision_t precision,
                       int n_leaves, const int32_t *leaf_ranks, const int32_t *leaf_modes,
                       const int64_t *leaf_extents, int n_steps, const int32_t *steps,
                       int n_keep, const int32_t *keep, jet_hip_network_t *net)
{
    return guarded([&] {
        JH_REQUIRE(ctx && net && leaf_ranks && steps && (keep || n_keep == 0), "Null argument.");
        set_device(ctx);
        auto *n = new jet_hip_network_s;
        try {
            n->ctx = ctx;
            n->plan = plan_dag(elem_bytes_of(dtype), mode_of(dtype, precision), n_leaves,
                               leaf_ranks, leaf_modes, leaf_extents, n_steps, steps, n_keep,
                               keep);
            if (n->plan.arena_bytes)
                n->arena = static_cast<char *>(device_alloc(n->plan.arena_bytes, ctx->device));
            if (n->plan.workspace_bytes)
                n->workspace =
                    static_cast<char *>(device_alloc(n->plan.workspace_bytes, ctx->device));
        }
        catch (...) {
            if (n->arena) (void)hipFree(n->arena);
            if (n->workspace) (void)hipFree(n->workspace);
            delete n;
            throw;
        }
        *net = n;
    });
}

int jet_hip_dag_estimate(jet_hip_dtype_t dtype, jet_hip_precision_t precision, int n_leaves,
                         const int32_t *leaf_ranks, const int32_t *leaf_modes,
                         const int64_t *leaf_extents, int n_steps, const int32_t *steps,
                         int n_keep, const int32_t *keep, double *flops, size_t *arena_bytes,
                         size_t *workspace_bytes)
{
    return guarded([&] {
        JH_REQUIRE(leaf_ranks && steps && (keep || n_keep == 0), "Null argument.");
        NetworkPlan np = plan_dag(elem_bytes_of(dtype), mode_of(dtype, precision), n_leaves,
                                  leaf_ranks, leaf_modes, leaf_extents, n_steps, steps, n_keep,
                                  keep);
        if (flops) *flops = np.flops;
        if (arena_bytes) *arena_bytes = np.arena_bytes;
        if (workspace_bytes) *workspace_bytes = np.workspace_bytes;
    });
}

int jet_hip_dag_value(jet_hip_network_t net, int node, void **ptr, int *rank, int32_t *modes,
                      int64_t *extents)
{
    return guarded([&] {
        JH_REQUIRE(net, "Null network.");
        const NetworkPlan &np = net->plan;
        JH_REQUIRE(np.dag, "Not a contraction DAG.");
        const int s = node - np.n_leaves;
        JH_REQUIRE(s >= 0 && s < (int)np.steps.size() && np.kept[s],
                   "The node is not a kept step result.");
        const NetworkStep &st = np.steps[s];
        if (ptr) *ptr = net->arena + st.offset;
        if (rank) *rank = (int)st.modes.size();
        for (size_t i = 0; i < st.modes.size(); i++) {
            if (modes) modes[i] = st.modes[i];
            if (extents) extents[i] = st.extents[i];
        }
    });
}

int jet_hip_network_info(jet_hip_network_t net, double *flops, size_t *arena_bytes,
                         size_t *workspace_bytes, int64_t *out_size, int *n_steps)
{
    return guarded([&] {
        JH_REQUIRE(net, "Null network.");
        if (flops) *flops = net->plan.flops;
        if (arena_bytes) *arena_bytes = net->plan.arena_bytes;
        if (workspace_bytes) *workspace_bytes = net->plan.workspace_bytes;
        if (out_size) {
            int64_t n = 1;
            for (auto e : net->plan.out_extents) n *= e;
            *out_size = n;
        }
        if (n_steps) *n_steps = (int)net->plan.steps.size();
    });
}

int jet_hip_network_step_info(jet_hip_network_t net, int step, int64_t *m, int64_t *n, int64_t *k,
                              int64_t *batch, int *gemm_kernel, int *permutes)
{
    return guarded([&] {
        JH_REQUIRE(net && step >= 0 && step < (int)net->plan.steps.size(), "Invalid step.");
        const ContractionDesc &d = net->plan.steps[step].desc;
        if (m) *m = d.M;
        if (n) *n = d.N;
        if (k) *k = d.K * d.chunks;
        if (batch) *batch = d.batch / d.chunks;
        if (gemm_kernel) *gemm_kernel = d.gemm_kernel;
        if (permutes) *permutes = (d.need_a ? 1 : 0) | (d.need_b ? 2 : 0) | (d.need_c ? 4 : 0);
    });
}

int jet_hip_network_execute(jet_hip_network_t net, const void *const *leaves, void *out,
                            void *stream)
{
    return guarded([&] {
        JH_REQUIRE(net && leaves && (out || net->plan.dag), "Null argument.");
        set_device(net->ctx);
        run_network(net, leaves, out, pick_stream(net->ctx, stream));
    });
}

int jet_hip_network_capture(jet_hip_network_t net, const void *const *leaves, void *out,
                            void *stream)
{
    return guarded([&] {
        JH_REQUIRE(net && leaves && (out || net->plan.dag), "Null argument.");
        set_device(net->ctx);
        hipStream_t s = pick_stream(net->ctx, stream);
        // A previous graph may still run (launched on any stream): wait for
        // it before its exec is destroyed, as jet_hip_network_destroy does.
        if (net->graph_exec) JH_HIP(hipDeviceSynchronize());
        free_graph(net);
        const bool prof_on = prof::enabled();
        prof::set_enabled(false);
        JH_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        try {
            run_network(net, leaves, out, s);
        }
        catch (...) {
            hipGraph_t g = nullptr;
            (void)hipStreamEndCapture(s, &g);
            if (g) (void)hipGraphDestroy(g);
            prof::set_enabled(prof_on);
            throw;
        }
        prof::set_enabled(prof_on);
        JH_HIP(hipStreamEndCapture(s, &net->graph));
        JH_HIP(hipGraphInstantiate(&net->graph_exec, net->graph, nullptr, nullptr, 0));
    });
}

int jet_hip_network_launch(jet_hip_network_t net, void *stream)
{
    return guarded([&] {
        JH_REQUIRE(net && net->graph_exec, "The network has no captured graph.");
        set_device(net->ctx);
        JH_HIP(hipGraphLaunch(net->graph_exec, pick_stream(net->ctx, stream)));
    });
}

int jet_hip_batch_capture(jet_hip_ctx_t ctx, int n, const jet_hip_network_t *networks,
                          const void *const *const *leaves, void *const *outs, void *stream,
                          jet_hip_batch_t *batch)
{
    return guarded([&] {
        JH_REQUIRE(ctx && batch && n >= 1 && networks && leaves && outs, "Null argument.");
        for (int i = 0; i < n; i++) {
            JH_REQUIRE(networks[i] && leaves[i] && (outs[i] || networks[i]->plan.dag),
                       "Null network, leaves or output.");
            JH_REQUIRE(networks[i]->ctx == ctx, "Every network of a batch needs the batch's context.");
        }
        set_device(ctx);
        hipStream_t s = pick_stream(ctx, stream);
        auto *b = new jet_hip_batch_s;
        b->ctx = ctx;
        const bool prof_on = prof::enabled();
        prof::set_enabled(false);
        hipError_t st = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
        if (st != hipSuccess) {
            prof::set_enabled(prof_on);
            delete b;
            JH_HIP(st);
        }
        try {
            for (int i = 0; i < n; i++) run_network(networks[i], leaves[i], outs[i], s);
        }
        catch (...) {
            hipGraph_t g = nullptr;
            (void)hipStreamEndCapture(s, &g);
            if (g) (void)hipGraphDestroy(g);
            prof::set_enabled(prof_on);
            delete b;
            throw;
        }
        prof::set_enabled(prof_on);
        st = hipStreamEndCapture(s, &b->graph);
        if (st == hipSuccess)
            st = hipGraphInstantiate(&b->graph_exec, b->graph, nullptr, nullptr, 0);
        if (st != hipSuccess) {
            if (b->graph) (void)hipGraphDestroy(b->graph);
            delete b;
            JH_HIP(st);
        }
        *batch = b;
    });
}

int jet_hip_batch_launch(jet_hip_batch_t batch, void *stream)
{
    return guarded([&] {
        JH_REQUIRE(batch && batch->graph_exec, "Null batch.");
        set_device(batch->ctx);
        JH_HIP(hipGraphLaunch(batch->graph_exec, pick_stream(batch->ctx, stream)));
    });
}

int jet_hip_batch_destroy(jet_hip_batch_t batch)
{
    return guarded([&] {
        if (!batch) return;
        set_device(batch->ctx);
        JH_HIP(hipDeviceSynchronize());
        if (batch->graph_exec) (void)hipGraphExecDestroy(batch->graph_exec);
        if (batch->graph) (void)hipGraphDestroy(batch->graph);
        delete batch;
    });
}

int jet_hip_network_destroy(jet_hip_network_t net)
{
    return guarded([&] {
        if (!net) return;
        set_device(net->ctx);
        JH_HIP(hipDeviceSynchronize());
        free_graph(net);
        if (net->arena) JH_HIP(hipFree(net->arena));
        if (net->workspace) JH_HIP(hipFree(net->workspace));
        delete net;
    });
}

} // extern "C"

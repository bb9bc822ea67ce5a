"""Benchmark of the MI355X contraction path (one JSON line on stdout).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload c2|c1|c3|c4]

Workloads:
  c2  (default) one amplitude <b|C|0> of a seeded 7x7-qubit grid random
      circuit of depth 14 ({sqrt X, sqrt Y, sqrt W} on every qubit, then
      fSim(pi/2, pi/6) on one of four alternating nearest-neighbour edge
      patterns per cycle), complex64, sliced on 12 bond-2 indices (4096
      slices), contracted along the min-fill path refined for batched
      execution (--c2-plan refined, the headline) and along the min-fill
      path itself (the second leg; --c2-plan min-fill swaps them).  The
      slice loop runs on the GPU in batches: 2^b slices per execution (b
      sliced indices kept as batch modes), executions over the other sliced
      indices.  A step is one execution (2^b slices); the timed region runs
      --steps executions with
      the split 16-bit MFMA GEMM of --split (default split-FP16: each
      complex64 operand row scaled by a power of two, split into hi + lo
      FP16, three products per real product, FP32 accumulation), then the
      --compare modes (default split-BF16) and plain FP32 MFMA; one
      execution in complex128 on FP64 MFMA is the FP64 reference of the
      relative errors.  Each execution's per-slice amplitudes are summed by
      the FP64 compensated reduction.
  c1  single large pairwise contraction: complex64 A, B of rank 28 (all
      extents 2, 2^28 elements each), 14 shared indices at seeded random
      positions in both index orders, re/im ~ N(0, 1/2); one
      2^14 x 2^14 x 2^14 complex GEMM behind the index permutations.
      A step is one ContractTensors call with the --split GEMM (the headline
      value); the plain FP32-MFMA path is timed beside it, and complex128 on
      FP64 MFMA gives the FP64 reference the error is measured against.

value (c2) = slices/s, with the operands resident in HBM when the timed region
starts; algorithmic TFLOP/s (8*M*N*K real flops per executed GEMM) beside it
(the value of c1 and c3).  The timed region is bracketed by HIP events on the
library's stream and a full synchronization.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "TFLOP/s & slices/s, 7x7 grid amp, split-BF16 MFMA vs plain FP32, err<=1e-5"
PEAK = {"bf16": 2.5e15, "fp16": 2.5e15, "fp32": 157.3e12, "fp64": 78.6e12, "hbm": 8.0e12}


def mfma_peak(family: str) -> float:
    """Dense MFMA peak of the arithmetic a kernel family runs on."""
    if "bf16" in family:
        return PEAK["bf16"]
    if "f16" in family:
        return PEAK["fp16"]
    if "fp64" in family:
        return PEAK["fp64"]
    return PEAK["fp32"]

# Split 16-bit MFMA modes: precision constant name, network precision string,
# event-profiler kernel family, 16-bit MFMA flops per algorithmic flop.
SPLIT_MODES = {"3m": ("PREC_BF16X3", "bf16x3", "gemm_bf16x3_mfma", 18 / 8),
               "4m": ("PREC_BF16X3_4M", "bf16x3_4m", "gemm_bf16x3_4m_mfma", 24 / 8),
               "f16": ("PREC_F16X3_4M", "f16x3", "gemm_f16x3_4m_mfma", 24 / 8)}
SPLIT_LABEL = {"3m": "split-BF16 3M", "4m": "split-BF16 4M", "f16": "split-FP16 4M"}

# Kernel names in rocprofv3 summaries for the event profiler's kernel families.
ROCPROF_NAME = {"gemm_bf16x3_4m_mfma": ("cgemm_x4m_kernel bf16", "cgemm_x4m_wide_kernel bf16 128x192",
                                        "cgemm_x4m_wide_kernel bf16 192x128",
                                        "cgemm_x4m_persist_kernel bf16"),
                "gemm_f16x3_4m_mfma": ("cgemm_x4m_kernel f16", "cgemm_x4m_wide_kernel f16 128x192",
                                       "cgemm_x4m_wide_kernel f16 192x128",
                                       "cgemm_x4m_persist_kernel f16"),
                "gemm_bf16x3_mfma": ("cgemm_bf16x3_kernel",),
                "gemm_fp32_mfma": ("cgemm_embed_kernel f32 128x128", "cgemm_embed_kernel f32 128x64",
                                   "cgemm_embed_kernel f32 64x128"),
                "gemm_fp64_mfma": ("cgemm_embed_kernel f64 128x64",),
                "permute_bits": ("permute_bits_kernel c64",)}


def profiled_traffic(summary: str, family: str):
    """(HBM bytes per dispatch of a kernel family, dispatch-weighted over its
    kernels; the commit the counters were read at), from the traffic table
    of a committed rocprofv3 summary under profiles/ (FETCH_SIZE and
    WRITE_SIZE passes, scaled by the copy-kernel calibration, and a
    "Measured at commit" line); (None, None) if absent."""
    path = os.path.join(ROOT, "profiles", summary)
    names = ROCPROF_NAME.get(family)
    if names is None or not os.path.exists(path):
        return None, None
    in_traffic = False
    total, dispatches, commit = 0.0, 0, None
    for line in open(path):
        if line.startswith("Measured at commit"):
            commit = line.split()[-1].strip("`.")
        if line.startswith("## HBM traffic"):
            in_traffic = True
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if in_traffic and len(cells) >= 5 and cells[0] in names:
            try:
                n = int(cells[1])
                total += n * float(cells[4]) * 1e9
                dispatches += n
            except ValueError:
                return None, commit
    return (total / dispatches if dispatches else None), commit


def log(msg):
    print(msg, file=sys.stderr, flush=True)


# ---------------------------------------------------------------------------
# C1: one rank-28 x rank-28 contraction
# ---------------------------------------------------------------------------

def c1_problem(seed=2026, rank=28, shared=14):
    rng = np.random.default_rng(seed)
    la = [f"a{i}" for i in range(rank)]
    common = [f"k{i}" for i in range(shared)]
    pos_a = sorted(rng.choice(rank, size=shared, replace=False).tolist())
    for j, p in enumerate(pos_a):
        la[p] = common[j]
    lb = [f"b{i}" for i in range(rank)]
    perm = rng.permutation(shared).tolist()
    pos_b = sorted(rng.choice(rank, size=shared, replace=False).tolist())
    for j, p in enumerate(pos_b):
        lb[p] = common[perm[j]]
    return la, lb


def _bit_offsets(labels, wanted):
    """Linear-offset contribution of each label in `wanted` (all extents 2)."""
    r = len(labels)
    return np.array([1 << (r - 1 - labels.index(w)) for w in wanted], dtype=np.int64)


def c1_sampled_reference(a_host, b_host, la, lb, rows, cols):
    """Exact FP64 entries C[m, n] of the C1 contraction, from the unpermuted
    operands (index arithmetic only; independent of the GPU plan)."""
    left = [x for x in la if x not in lb]
    right = [x for x in lb if x not in la]
    common = [x for x in la if x in lb]
    nk = len(common)
    kbits = (np.arange(1 << nk)[:, None] >> np.arange(nk - 1, -1, -1)[None, :]) & 1
    offA_k = kbits @ _bit_offsets(la, common)
    offB_k = kbits @ _bit_offsets(lb, common)
    offA_m = _bit_offsets(la, left)
    offB_n = _bit_offsets(lb, right)
    out = np.empty(len(rows), dtype=np.complex128)
    for i, (m, n) in enumerate(zip(rows, cols)):
        mb = (m >> np.arange(len(left) - 1, -1, -1)) & 1
        nb = (n >> np.arange(len(right) - 1, -1, -1)) & 1
        av = a_host[int(mb @ offA_m) + offA_k].astype(np.complex128)
        bv = b_host[int(nb @ offB_n) + offB_k].astype(np.complex128)
        out[i] = np.dot(av, bv)
    return out


def run_c1(args):
    from jet_amd import _native as nat
    from jet_amd.tensor import _relabel
    import ctypes

    ctx = nat.default_context()
    L = nat.lib()
    la, lb = c1_problem()
    left = [x for x in la if x not in lb]
    right = [x for x in lb if x not in la]
    lc = left + right
    ma, mb, mc = _relabel(la, lb, lc)
    n = 1 << 28
    ext = [2] * 28

    A = nat.DeviceBuffer(n * 8, ctx)
    B = nat.DeviceBuffer(n * 8, ctx)
    C = nat.DeviceBuffer(n * 8, ctx)
    nat.check(L.jet_hip_fill_random(ctx.handle, nat.C64, n, 1, 2, A.ptr, None))
    nat.check(L.jet_hip_fill_random(ctx.handle, nat.C64, n, 2, 2, B.ptr, None))

    def plan(dtype, prec):
        p = ctypes.c_void_p()
        nat.check(L.jet_hip_contraction_plan_create(ctx.handle, dtype, prec, nat.ROW_MAJOR, 28,
                                                    nat.i32(ma), nat.i64(ext), 28, nat.i32(mb),
                                                    nat.i64(ext), 28, nat.i32(mc),
                                                    ctypes.byref(p)))
        ws = ctypes.c_size_t()
        nat.check(L.jet_hip_plan_workspace_size(p, ctypes.byref(ws)))
        buf = nat.DeviceBuffer(max(ws.value, 256), ctx)
        return p, buf, ws.value

    flops = 8.0 * (1 << 42)

    def timed(p, a, b, c, ws, wsz, steps, warmup, profile=False):
        for _ in range(warmup):
            nat.check(L.jet_hip_contract(p, a.ptr, b.ptr, c.ptr, ws.ptr, wsz, None))
        ctx.synchronize()
        if profile:
            nat.Profiler.reset()
            nat.Profiler.enable(True)
        e0, e1 = nat.Event(ctx), nat.Event(ctx)
        t0 = time.perf_counter()
        e0.record()
        for _ in range(steps):
            nat.check(L.jet_hip_contract(p, a.ptr, b.ptr, c.ptr, ws.ptr, wsz, None))
        e1.record()
        ctx.synchronize()
        wall = time.perf_counter() - t0
        ms = e0.elapsed_ms(e1) / steps
        prof = None
        if profile:
            prof = nat.Profiler.read()
            nat.Profiler.enable(False)
        return ms, wall, prof

    split_prec = getattr(nat, SPLIT_MODES[args.split][0])
    kname = SPLIT_MODES[args.split][2]
    log(f"c1: {SPLIT_LABEL[args.split]} plan")
    p16, ws16, wsz16 = plan(nat.C64, split_prec)
    ms16, wall16, prof16 = timed(p16, A, B, C, ws16, wsz16, args.steps, args.warmup, True)
    log(f"c1: {SPLIT_LABEL[args.split]} {ms16:.2f} ms/step {flops / ms16 / 1e9:.1f} TFLOP/s")
    c16 = np.empty(n, np.complex64)
    C.download(c16)
    del ws16

    # Copy kernel of known size (2 GiB read + 2 GiB written), the calibration
    # point for the FETCH_SIZE / WRITE_SIZE counters.
    nat.check(L.jet_hip_memcpy(ctx.handle, C.ptr, A.ptr, n * 8, 2, None, 0))
    if args.quick:
        for p in (p16,):
            L.jet_hip_plan_destroy(p)
        return {"metric": METRIC, "value": round(flops / (ms16 * 1e-3) / 1e12, 3),
                "unit": "TFLOP/s", "ms_per_step": round(ms16, 3),
                "kernels_split": prof16}, None

    log("c1: FP32 plan")
    p32, ws32, wsz32 = plan(nat.C64, nat.PREC_FP32)
    ms32, _, prof32 = timed(p32, A, B, C, ws32, wsz32, max(1, args.steps // 2), 1, True)
    log(f"c1: fp32 {ms32:.2f} ms/step {flops / ms32 / 1e9:.1f} TFLOP/s")
    c32 = np.empty(n, np.complex64)
    C.download(c32)
    del ws32, C

    log("c1: FP64 reference")
    A64 = nat.DeviceBuffer(n * 16, ctx)
    B64 = nat.DeviceBuffer(n * 16, ctx)
    C64 = nat.DeviceBuffer(n * 16, ctx)
    nat.check(L.jet_hip_convert(ctx.handle, nat.C64, nat.C128, n, A.ptr, A64.ptr, None))
    nat.check(L.jet_hip_convert(ctx.handle, nat.C64, nat.C128, n, B.ptr, B64.ptr, None))
    p64, ws64, wsz64 = plan(nat.C128, nat.PREC_FP64)
    ms64, _, _ = timed(p64, A64, B64, C64, ws64, wsz64, 1, 0)
    log(f"c1: fp64 {ms64:.2f} ms/step {flops / ms64 / 1e9:.1f} TFLOP/s")
    c64ref = np.empty(n, np.complex128)
    C64.download(c64ref)
    del ws64, A64, B64, C64

    # Pin the FP64 GPU reference to exact CPU dot products on sampled entries.
    a_host = np.empty(n, np.complex64)
    b_host = np.empty(n, np.complex64)
    A.download(a_host)
    B.download(b_host)
    rng = np.random.default_rng(7)
    rows = rng.integers(0, 1 << 14, 48)
    cols = rng.integers(0, 1 << 14, 48)
    exact = c1_sampled_reference(a_host, b_host, la, lb, rows, cols)
    got64 = c64ref.reshape(1 << 14, 1 << 14)[rows, cols]
    fp64_vs_cpu = float(np.max(np.abs(got64 - exact)) / np.sqrt(np.mean(np.abs(exact) ** 2)))
    nref = np.linalg.norm(c64ref)
    err16 = float(np.linalg.norm(c16.astype(np.complex128) - c64ref) / nref)
    err32 = float(np.linalg.norm(c32.astype(np.complex128) - c64ref) / nref)
    log(f"c1: err {SPLIT_LABEL[args.split]} {err16:.3e} fp32 {err32:.3e} "
        f"fp64-vs-cpu-sampled {fp64_vs_cpu:.2e}")
    for p in (p16, p32, p64):
        L.jet_hip_plan_destroy(p)

    gemm = prof16.get(kname, {})
    gemm_ms = gemm.get("ms", 0.0) / max(1, gemm.get("launches", 1))
    gemm_flops = gemm.get("flops", 0.0) / max(1, gemm.get("launches", 1))
    achieved = gemm_flops / (gemm_ms * 1e-3) if gemm_ms else 0.0
    loop = mfma_loop_ceiling(ctx)
    log(f"c1: FP16 MFMA loop {loop[0] / 1e12:.1f} TFLOP/s at {loop[1]:.3f} GHz")
    result = {
        "metric": METRIC,
        "value": round(flops / (ms16 * 1e-3) / 1e12, 3),
        "unit": "TFLOP/s",
        "ms_per_step": round(ms16, 3),
        "dtype": f"c64 ({SPLIT_LABEL[args.split]} MFMA, three terms, fp32 accumulate)",
        "config": {"workload": "c1: rank-28 x rank-28 complex64 contraction, 14 shared "
                               "indices at random positions, 2^14 x 2^14 x 2^14 complex GEMM",
                   "m": 1 << 14, "n": 1 << 14, "k": 1 << 14},
        "fp32_mfma": {"ms_per_step": round(ms32, 3),
                      "tflops": round(flops / (ms32 * 1e-3) / 1e12, 3)},
        "fp64_mfma": {"ms_per_step": round(ms64, 3),
                      "tflops": round(flops / (ms64 * 1e-3) / 1e12, 3)},
        "split_mode": SPLIT_LABEL[args.split],
        "speedup_split_vs_fp32": round(ms32 / ms16, 3),
        "rel_err_vs_fp64": {"split": err16, "fp32": err32,
                            "fp64_vs_cpu_exact_sampled": fp64_vs_cpu},
        "kernels_split": {k: {"launches": v["launches"],
                               "ms_avg": round(v["ms"] / v["launches"], 4)}
                           for k, v in prof16.items()},
        "kernels_fp32": {k: {"launches": v["launches"],
                             "ms_avg": round(v["ms"] / v["launches"], 4)}
                         for k, v in prof32.items()},
        "roofline": {"kernel": kname, "bound": "mfma",
                     "achieved": round(achieved / 1e12, 3), "peak": PEAK["bf16"] / 1e12,
                     "unit": "TFLOP/s", "frac": round(achieved / PEAK["bf16"], 4),
                     "hw_16bit_mfma_frac": round(achieved * SPLIT_MODES[args.split][3] /
                                                 PEAK["bf16"], 4),
                     "mfma_loop": mfma_loop_fields(achieved * SPLIT_MODES[args.split][3],
                                                   *loop),
                     "traffic": profiled_traffic("c1_split_contraction.md", kname)[0],
                     "traffic_source": ("profiles/c1_split_contraction.md: rocprofv3 --pmc "
                                        "FETCH_SIZE / WRITE_SIZE, read at commit "
                                        f"{profiled_traffic('c1_split_contraction.md', kname)[1]}")},
    }
    return result, (a_host, b_host, la, lb)


def c1_cpu_baseline(host, threads):
    """Reference CPU path (QFlex transposes + OpenBLAS cgemm) on a bounded
    sample of C1: one of A's free indices fixed to 0, i.e. M = 2^13."""
    import oracle
    a_host, b_host, la, lb = host
    left = [x for x in la if x not in lb]
    fixed = left[:1]
    shape = [2] * 28
    arr = a_host.reshape(shape)
    idx = tuple(0 if la[i] in fixed else slice(None) for i in range(28))
    a_s = np.ascontiguousarray(arr[idx]).reshape(-1)
    la_s = [x for x in la if x not in fixed]
    oracle.load_blas(threads)
    t0 = time.perf_counter()
    _, c = oracle.baseline_contract(la_s, [2] * 27, a_s, lb, [2] * 28, b_host, threads)
    dt = time.perf_counter() - t0
    flops = 8.0 * (1 << 13) * (1 << 14) * (1 << 14)
    return {"value": round(flops / dt / 1e12, 4), "unit": "TFLOP/s", "cores": threads,
            "kind": "port",
            "sample": "c1 with one free index of A fixed (M=2^13, N=K=2^14): QFlex "
                      "transposes + OpenBLAS cblas_cgemm, one ContractTensors call, "
                      f"{dt:.1f} s"}


# ---------------------------------------------------------------------------
# C2: sliced 7x7 depth-14 amplitude, batched slice loop
# ---------------------------------------------------------------------------

C2_GRID = (7, 7, 14)
C2_SEED = 2026
C2_SLICES = 12


# The two C2 plans: the min-fill path sliced on 12 indices (written by
# `python -m jet_amd.paths`), and the same path refined for batched execution
# (`python -m jet_amd.paths --refine ... --batch-log2 7 --extra-log2 1`),
# with the batch size each runs at by default.
C2_PLANS = {"refined": os.path.join(ROOT, "jet_amd", "data", "grid_7x7x14_seed2026_s12_refined.json"),
            "min-fill": os.path.join(ROOT, "jet_amd", "data", "grid_7x7x14_seed2026_s12.json")}
C2_BATCH_LOG2 = {"refined": 8, "min-fill": 7}
# rocprofv3 summaries (with HBM traffic) of the c2 run on each plan.
C2_PROFILE = {"refined": "c2_sliced_amplitude.md", "min-fill": "c2_min_fill_path.md"}


def dname_of(prof):
    """The kernel family with the most device time in an event profile."""
    return max(prof.items(), key=lambda kv: kv[1]["ms"])[0]


def mfma_loop_ceiling(ctx, rounds=200000, reps=3):
    """The library's FP16 MFMA rate loop (v_mfma_f32_16x16x32_f16 on register
    operands, no memory traffic, every CU): the FP16 MFMA TFLOP/s the chip
    sustains under a full MFMA load and the shader clock it holds meanwhile
    -- the practical ceiling of the split GEMMs' FP16 MFMA work (about 0.3 s
    on MI355X)."""
    import ctypes
    from jet_amd import _native as nat
    tf, ghz = ctypes.c_double(), ctypes.c_double()
    nat.check(nat.lib().jet_hip_mfma_ceiling(ctx.handle, rounds, reps, ctypes.byref(tf),
                                             ctypes.byref(ghz)))
    return tf.value * 1e12, ghz.value


def mfma_loop_fields(achieved_hw, loop_flops, loop_ghz):
    """Roofline fields relating a split GEMM's FP16 MFMA work rate to the
    measured loop ceiling."""
    return {"fp16_tflops": round(loop_flops / 1e12, 1), "clock_ghz": round(loop_ghz, 3),
            "hw_16bit_frac_of_loop": round(achieved_hw / loop_flops, 4),
            "note": "FP16 MFMA rate of the library's register-only v_mfma_f32_16x16x32_f16 "
                    "loop over every CU, measured in this run: the rate the chip sustains "
                    "under a full MFMA load (it clocks down from the 2.4 GHz of the nominal "
                    "peak)"}


def hbm_copy_ceiling(ctx, gib=2, reps=10):
    """The library's own streaming copy (the identity permutation) on a
    `gib` GiB complex64 tensor: read + write bytes per second, the practical
    HBM ceiling that permutations and splits are judged against."""
    from jet_amd import _native as nat
    n = (gib << 30) // 8
    src, dst = nat.DeviceBuffer(n * 8, ctx), nat.DeviceBuffer(n * 8, ctx)
    nat.check(nat.lib().jet_hip_fill_random(ctx.handle, nat.C64, n, 7, 1, src.ptr, None))
    modes, ext = nat.i32([0]), nat.i64([n])

    def copy():
        nat.check(nat.lib().jet_hip_permute(ctx.handle, nat.C64, nat.ROW_MAJOR, 1, modes, ext,
                                            modes, src.ptr, dst.ptr, None))
    copy()
    ctx.synchronize()
    e0, e1 = nat.Event(ctx), nat.Event(ctx)
    e0.record()
    for _ in range(reps):
        copy()
    e1.record()
    ctx.synchronize()
    ms = e0.elapsed_ms(e1) / reps
    src.free()
    dst.free()
    return 2.0 * n * 8 / (ms * 1e-3)


def c2_problem(plan_name="min-fill"):
    from jet_amd import circuits, paths
    if plan_name == "min-fill":
        plan = paths.load_or_make_grid_plan(*C2_GRID, C2_SEED, C2_SLICES)
    else:
        with open(C2_PLANS[plan_name]) as f:
            plan = json.load(f)
        plan["path"] = [tuple(p) for p in plan["path"]]
    circ = circuits.random_grid_circuit(*C2_GRID, seed=C2_SEED)
    net = circuits.amplitude_network(circ, dtype=np.complex64)
    return plan, circ, net


C2_AMPLITUDE = os.path.join(ROOT, "tests", "golden", "c2_amplitude_fp64.json")


def c2_reference_amplitude(plan_name):
    """The complex128 FP64-MFMA amplitude of all 4096 slices (any plan: the
    sliced network is the same), as tests/golden/make_c2_amplitude.py wrote
    it; None if that file is absent."""
    if not os.path.exists(C2_AMPLITUDE):
        return None
    with open(C2_AMPLITUDE) as f:
        doc = json.load(f)
    return complex(*doc["amplitude"])


def run_c2(args):
    from jet_amd import _native as nat
    from jet_amd.network import SlicedContraction

    ctx = nat.default_context()
    t0 = time.perf_counter()
    plan, circ, net = c2_problem(args.c2_plan)
    path, sliced = plan["path"], plan["sliced"]
    batch_log2 = C2_BATCH_LOG2[args.c2_plan] if args.batch_log2 is None else args.batch_log2
    batch = sliced[len(sliced) - batch_log2:] if batch_log2 else []
    log(f"c2: {args.c2_plan} plan ready ({time.perf_counter() - t0:.1f}s): {len(net)} tensors, "
        f"{len(path)} steps, per-slice PathInfo flops {plan['per_slice_pathinfo_flops']:.3e}, "
        f"batch {len(batch)} of {len(sliced)} sliced indices")

    def build(prec, dtype, batch=batch):
        data = [d.astype(dtype) for d in net.data]
        return SlicedContraction(net.indices, net.shapes, data, path, sliced, batch=batch,
                                 dtype=dtype, precision=prec, ctx=ctx, use_graph=False)

    def timed(sc, iters, warmup, profile):
        """The timed executions run as one captured hipGraph launch (the
        warm-up executions as another, launched first), followed on the same
        stream by the FP64 compensated sum of the per-slice results, inside
        the timed region.  Execution i covers slice range
        execution_slot(i, n_loop): past the last one the slice loop starts
        again.  With `profile`, one more execution runs eagerly under the
        event profiler afterwards: the per-kernel times of the roofline."""
        done = min(sc.n_loop, warmup + iters) * sc.n_batch
        if args.eager:
            sc.run(range(warmup))
            ctx.synchronize()
            e0, e1 = nat.Event(ctx), nat.Event(ctx)
            e0.record()
            sc.run(range(warmup, warmup + iters))
            sc.reduce_enqueue(done)
            e1.record()
            ctx.synchronize()
            ms = e0.elapsed_ms(e1)
        else:
            if warmup:
                wb = sc.capture(range(warmup))
                wb.launch()
                ctx.synchronize()
                wb.close()
            tb = sc.capture(range(warmup, warmup + iters))
            ctx.synchronize()
            e0, e1 = nat.Event(ctx), nat.Event(ctx)
            e0.record()
            tb.launch()
            sc.reduce_enqueue(done)
            e1.record()
            ctx.synchronize()
            ms = e0.elapsed_ms(e1)
            tb.close()
        total = sc.reduced()
        prof = None
        if profile:
            nat.Profiler.reset()
            nat.Profiler.enable(True)
            sc.run([warmup])
            ctx.synchronize()
            prof = nat.Profiler.read()
            nat.Profiler.enable(False)
        return ms, prof, total, done

    steps, warm = args.steps, args.warmup
    label = SPLIT_LABEL[args.split]
    sc16 = build(SPLIT_MODES[args.split][1], np.complex64)
    n_batch, n_slices = sc16.n_batch, sc16.n_slices
    exec_flops = sc16.flops_per_execution
    log(f"c2: {label} network: {exec_flops:.3e} flops/execution, arena "
        f"{sc16.net.arena_bytes / 2**30:.1f} GiB, workspace {sc16.net.workspace_bytes / 2**30:.1f} GiB, "
        f"{sc16.n_loop} executions of {n_batch} slices per amplitude")
    ms16, prof16, sum16, done16 = timed(sc16, steps, warm, True)
    if args.quick:
        # Profiler runs: the split leg's executions and the 2 GiB streaming
        # copy of known size that calibrates the HBM traffic counters.
        copy_bps = hbm_copy_ceiling(ctx)
        log(f"c2 quick: {ms16 / steps:.1f} ms/execution, copy {copy_bps / 1e12:.2f} TB/s")
        return {"metric": METRIC, "value": round(steps * n_batch / (ms16 * 1e-3), 3),
                "unit": "slices/s", "ms_per_step": round(ms16 / steps, 3),
                "kernels_split": prof16, "copy_ceiling_gbps": round(copy_bps / 1e9, 1)}, None
    if done16 < n_slices:
        # The whole amplitude needs every execution: run the rest untimed.
        sc16.run(range(warm + steps, sc16.n_loop))
        sum16 = sc16.reduce()
    res16 = sc16.slice_results()
    log(f"c2: {label} {ms16 / steps:.1f} ms/execution, "
        f"{exec_flops * steps / ms16 / 1e9:.1f} TFLOP/s")
    sc16.net.close()
    del sc16

    # The other split modes, timed the same way for comparison.
    others = {}
    for mode in args.compare:
        if mode == args.split:
            continue
        sc = build(SPLIT_MODES[mode][1], np.complex64)
        ms, _, _, _ = timed(sc, steps, warm, False)
        others[mode] = (ms, sc.slice_results(n_batch))
        log(f"c2: {SPLIT_LABEL[mode]} {ms / steps:.1f} ms/execution, "
            f"{exec_flops * steps / ms / 1e9:.1f} TFLOP/s")
        sc.net.close()
        del sc

    steps32 = max(1, args.steps32)
    sc32 = build("fp32", np.complex64)
    ms32, prof32, _, _ = timed(sc32, steps32, 1, True)
    res32 = sc32.slice_results(n_batch)
    log(f"c2: fp32 {ms32 / steps32:.1f} ms/execution, "
        f"{exec_flops * steps32 / ms32 / 1e9:.1f} TFLOP/s")
    sc32.net.close()
    del sc32

    # FP64 reference (complex128, FP64 MFMA) of the first execution's first
    # quarter of slices: two fewer sliced indices batched keep its arena (16-byte
    # elements) at half the complex64 one; the slice order is the same, so
    # its first execution holds the first slices of the complex64 runs.
    batch64 = batch[2:] if len(batch) >= 2 else batch
    sc64 = build("fp64", np.complex128, batch64)
    n_ref = sc64.n_batch
    ms64, _, _, _ = timed(sc64, 1, 0, False)
    res64 = sc64.slice_results(n_ref)
    sc64.net.close()
    del sc64
    err16 = float(np.linalg.norm(res16[:n_ref] - res64) / np.linalg.norm(res64))
    err32 = float(np.linalg.norm(res32[:n_ref] - res64) / np.linalg.norm(res64))
    worst16 = float(np.max(np.abs(res16[:n_ref] - res64)) / np.max(np.abs(res64)))
    # The whole 4096-slice amplitude against the complex128 FP64-MFMA
    # amplitude of the same plan (recomputed live by --full-amplitude).
    amp64 = c2_reference_amplitude(args.c2_plan)
    amp_err16 = abs(sum16 - amp64) / abs(amp64) if amp64 is not None else None
    log(f"c2: fp64 {ms64:.1f} ms/execution; rel err per-slice {label} {err16:.2e} "
        f"fp32 {err32:.2e}; whole-amplitude {label} "
        f"{amp_err16 if amp_err16 is None else f'{amp_err16:.2e}'}")
    compare = {}
    for mode, (ms, res) in others.items():
        e = float(np.linalg.norm(res[:n_ref] - res64) / np.linalg.norm(res64))
        compare[SPLIT_LABEL[mode]] = {
            "ms_per_step": round(ms / steps, 3),
            "tflops": round(exec_flops * steps / (ms * 1e-3) / 1e12, 3),
            "slices_per_s": round(steps * n_batch / (ms * 1e-3), 4),
            "speedup_vs_fp32": round((ms32 / steps32) / (ms / steps), 3),
            "rel_err_per_slice_vs_fp64": e}
        log(f"c2: {SPLIT_LABEL[mode]} rel err per-slice {e:.2e}")

    slices16 = steps * n_batch
    tflops16 = exec_flops * steps / (ms16 * 1e-3)
    tflops32 = exec_flops * steps32 / (ms32 * 1e-3)
    sps16 = slices16 / (ms16 * 1e-3)
    sps32 = steps32 * n_batch / (ms32 * 1e-3)
    traffic, traffic_commit = profiled_traffic(C2_PROFILE[args.c2_plan], dname_of(prof16))
    copy_bps = hbm_copy_ceiling(ctx)
    loop = mfma_loop_ceiling(ctx)
    log(f"c2: FP16 MFMA loop {loop[0] / 1e12:.1f} TFLOP/s at {loop[1]:.3f} GHz")
    perm = prof16.get("permute_bits")
    perm_bps = perm["bytes"] / (perm["ms"] * 1e-3) if perm and perm["ms"] > 0 else None
    split = prof16.get("split_f16")
    split_bps = split["bytes"] / (split["ms"] * 1e-3) if split and split["ms"] > 0 else None
    log(f"c2: copy ceiling {copy_bps / 1e12:.2f} TB/s; permutations "
        f"{(perm_bps or 0) / 1e12:.2f} TB/s; splits {(split_bps or 0) / 1e12:.2f} TB/s")
    dom = max(prof16.items(), key=lambda kv: kv[1]["ms"])
    dname, dk = dom
    bound = "mfma" if dk["flops"] > 0 else "hbm"
    if bound == "mfma":
        achieved = dk["flops"] / (dk["ms"] * 1e-3)
        peak = mfma_peak(dname)
        unit, scale = "TFLOP/s", 1e12
    else:
        achieved = dk["bytes"] / (dk["ms"] * 1e-3)
        peak, unit, scale = PEAK["hbm"], "GB/s", 1e9
    other = "min-fill" if args.c2_plan == "refined" else "refined"
    path_desc = {"refined": "min-fill path refined for batched execution",
                 "min-fill": "min-fill path"}[args.c2_plan]
    return {
        "metric": METRIC,
        "value": round(sps16, 3),
        "unit": "slices/s",
        "ms_per_step": round(ms16 / steps, 3),
        "dtype": f"c64 ({label} MFMA, three terms, fp32 accumulate, on MFMA-bound steps; "
                 "fp32 MFMA elsewhere)",
        "config": {"workload": "c2: one amplitude of a seeded 7x7-qubit grid random circuit, "
                               f"depth 14, complex64, {path_desc} sliced on 12 bond-2 indices "
                               "(4096 slices); a step is one batched execution of "
                               f"{n_batch} slices",
                   "plan": os.path.relpath(C2_PLANS[args.c2_plan], ROOT),
                   "qubits": 49, "depth": 14, "sliced_indices": len(sliced),
                   "slices_total": 4096, "slices_per_step": n_batch,
                   "tensors": len(net), "path_steps": len(path),
                   "per_slice_pathinfo_flops": plan["per_slice_pathinfo_flops"],
                   "per_slice_max_log2_size": plan["per_slice_max_log2_size"]},
        "tflops": round(tflops16 / 1e12, 3),
        "slices_per_s": round(sps16, 4),
        "fp32_mfma": {"ms_per_step": round(ms32 / steps32, 3), "tflops": round(tflops32 / 1e12, 3),
                      "slices_per_s": round(sps32, 4)},
        "split_mode": label,
        "speedup_split_vs_fp32": round(tflops16 / tflops32, 3),
        "projected_amplitude_s": {"split": round(4096 / sps16, 1), "fp32": round(4096 / sps32, 1)},
        "rel_err_vs_fp64": {"split_per_slice": err16, "fp32_per_slice": err32,
                            "split_worst_slice_vs_max": worst16,
                            "slices_compared": n_ref,
                            "split_whole_amplitude": amp_err16,
                            "amplitude_split": [sum16.real, sum16.imag],
                            "amplitude_fp64": (None if amp64 is None
                                               else [amp64.real, amp64.imag])},
        "other_split_modes": compare,
        "other_plan": (c2_plan_leg(args, ctx, net, label, other) if args.other_plan
                       else None),
        "full_amplitude": (c2_full_amplitude(ctx, net, plan, batch, SPLIT_MODES[args.split][1])
                           if args.full_amplitude else None),
        "fp64_mfma": {"ms_per_step": round(ms64, 3), "slices_per_step": n_ref},
        "hbm": {"copy_ceiling_gbps": round(copy_bps / 1e9, 1),
                "copy_frac_of_8tbps": round(copy_bps / PEAK["hbm"], 4),
                "permute_gbps": None if perm_bps is None else round(perm_bps / 1e9, 1),
                "permute_frac_of_copy": None if perm_bps is None else round(perm_bps / copy_bps, 4),
                "split_gbps": None if split_bps is None else round(split_bps / 1e9, 1),
                "split_frac_of_copy": None if split_bps is None else round(split_bps / copy_bps, 4),
                "note": "read + write bytes per second of the library's own streaming copy "
                        "(the identity permutation, 2 GiB) and of the execution's "
                        "permutation / split passes (algorithmic bytes)"},
        "kernels_split": {k: {"launches": v["launches"], "ms": round(v["ms"], 3)}
                          for k, v in prof16.items()},
        "kernels_fp32": {k: {"launches": v["launches"], "ms": round(v["ms"], 3)}
                         for k, v in prof32.items()},
        "roofline": {"kernel": dname, "bound": bound, "achieved": round(achieved / scale, 3),
                     "peak": peak / scale, "unit": unit, "frac": round(achieved / peak, 4),
                     "hw_16bit_mfma_frac": (round(achieved * SPLIT_MODES[args.split][3] / peak, 4)
                                            if dname == SPLIT_MODES[args.split][2] else None),
                     "mfma_loop": (mfma_loop_fields(achieved * SPLIT_MODES[args.split][3], *loop)
                                   if dname == SPLIT_MODES[args.split][2] else None),
                     "algorithmic_bytes": round(dk["bytes"] / dk["launches"]),
                     "traffic": traffic,
                     "traffic_source": (f"profiles/{C2_PROFILE[args.c2_plan]}: rocprofv3 --pmc "
                                        f"FETCH_SIZE / WRITE_SIZE of this kernel family, read "
                                        f"at commit {traffic_commit}")},
    }, (plan, net)


def c2_full_amplitude(ctx, net, plan, batch, split_prec):
    """The whole 4096-slice amplitude in the split mode, plain FP32 and
    complex128 on FP64 MFMA (two fewer batched indices): its relative error
    and the per-slice normwise error over every slice."""
    from jet_amd.network import SlicedContraction
    runs = {}
    for name, prec, dtype, b in (("split", split_prec, np.complex64, batch),
                                 ("fp32", "fp32", np.complex64, batch),
                                 ("fp64", "fp64", np.complex128, batch[2:])):
        data = [d.astype(dtype) for d in net.data]
        sc = SlicedContraction(net.indices, net.shapes, data, plan["path"], plan["sliced"],
                               batch=b, dtype=dtype, precision=prec, ctx=ctx, use_graph=False)
        t0 = time.perf_counter()
        chunk = max(1, sc.n_loop // 4)
        for i in range(0, sc.n_loop, chunk):
            sc.run(range(i, min(sc.n_loop, i + chunk)))
            ctx.synchronize()
            log(f"c2 full amplitude: {name} {min(sc.n_loop, i + chunk)}/{sc.n_loop} executions")
        runs[name] = (sc.reduce(), sc.slice_results(), time.perf_counter() - t0)
        sc.net.close()
        del sc
    a64, r64, _ = runs["fp64"]
    out = {"slices": int(r64.size), "amplitude_fp64": [a64.real, a64.imag]}
    for name in ("split", "fp32"):
        a, r, sec = runs[name]
        out[name] = {"amplitude_rel_err": abs(a - a64) / abs(a64),
                     "per_slice_rel_err": float(np.linalg.norm(r - r64) / np.linalg.norm(r64)),
                     "seconds": round(sec, 2)}
    out["fp64"] = {"seconds": round(runs["fp64"][2], 2)}
    log(f"c2 full amplitude: rel err split {out['split']['amplitude_rel_err']:.2e}, "
        f"fp32 {out['fp32']['amplitude_rel_err']:.2e}")
    return out


def c2_plan_leg(args, ctx, net, label, plan_name):
    """The same amplitude along the other C2 plan (min-fill, or refined for
    batched execution), at that plan's default batch: split and FP32 runs
    timed as one captured graph each (eager launches with --eager); errors
    against a complex128 FP64-MFMA run of the first execution's first
    quarter of slices; the split GEMM's own TFLOP/s from the event
    profiler."""
    from jet_amd import _native as nat
    from jet_amd.network import SlicedContraction
    plan, _, _ = c2_problem(plan_name)
    path, sliced = plan["path"], plan["sliced"]
    nb = C2_BATCH_LOG2[plan_name]
    batch = sliced[len(sliced) - nb:]

    def build(prec, dtype, b):
        data = [d.astype(dtype) for d in net.data]
        return SlicedContraction(net.indices, net.shapes, data, path, sliced, batch=b,
                                 dtype=dtype, precision=prec, ctx=ctx, use_graph=False)

    def timed(sc, iters):
        """One warm-up execution, then `iters` timed ones."""
        if args.eager:
            sc.run(range(1))
        else:
            wb = sc.capture(range(1))
            wb.launch()
            ctx.synchronize()
            wb.close()
        ctx.synchronize()
        e0, e1 = nat.Event(ctx), nat.Event(ctx)
        if args.eager:
            e0.record()
            sc.run(range(1, 1 + iters))
            e1.record()
        else:
            tb = sc.capture(range(1, 1 + iters))
            ctx.synchronize()
            e0.record()
            tb.launch()
            e1.record()
        ctx.synchronize()
        ms = e0.elapsed_ms(e1)
        if not args.eager:
            tb.close()
        return ms

    steps = args.steps
    sc = build(SPLIT_MODES[args.split][1], np.complex64, batch)
    n_batch, exec_flops = sc.n_batch, sc.flops_per_execution
    ms16 = timed(sc, steps)
    res16 = sc.slice_results(n_batch)
    # one more execution under the event profiler: per-family kernel times
    nat.Profiler.reset()
    nat.Profiler.enable(True)
    sc.run([0])
    ctx.synchronize()
    prof = nat.Profiler.read()
    nat.Profiler.enable(False)
    sc.net.close()
    del sc
    sc = build("fp32", np.complex64, batch)
    ms32 = timed(sc, 1)
    res32 = sc.slice_results(n_batch)
    sc.net.close()
    del sc
    sc = build("fp64", np.complex128, batch[2:] if len(batch) >= 2 else batch)
    n_ref = sc.n_batch
    sc.run([0])
    res64 = sc.slice_results(n_ref)
    sc.net.close()
    del sc
    err = lambda r: float(np.linalg.norm(r[:n_ref] - res64) / np.linalg.norm(res64))
    sps16 = steps * n_batch / (ms16 * 1e-3)
    sps32 = n_batch / (ms32 * 1e-3)
    gemm = prof.get(SPLIT_MODES[args.split][2])
    gemm_tf = gemm["flops"] / (gemm["ms"] * 1e-3) if gemm and gemm["ms"] > 0 else None
    log(f"c2 {plan_name} path: {label} {ms16 / steps:.1f} ms/execution ({sps16:.1f} slices/s, "
        f"{exec_flops * steps / ms16 / 1e9:.1f} TFLOP/s), fp32 {ms32:.1f} ms/execution; "
        f"rel err per-slice {err(res16):.2e}, fp32 {err(res32):.2e}")
    return {"plan": os.path.relpath(C2_PLANS[plan_name], ROOT),
            "per_slice_pathinfo_flops": plan["per_slice_pathinfo_flops"],
            "per_slice_max_log2_size": plan["per_slice_max_log2_size"],
            "slices_per_step": n_batch, "ms_per_step": round(ms16 / steps, 3),
            "slices_per_s": round(sps16, 4),
            "tflops": round(exec_flops * steps / (ms16 * 1e-3) / 1e12, 3),
            "split_gemm_tflops": round(gemm_tf / 1e12, 3) if gemm_tf else None,
            "split_gemm_frac": round(gemm_tf / mfma_peak(SPLIT_MODES[args.split][2]), 4)
                               if gemm_tf else None,
            "projected_amplitude_s": round(4096 / sps16, 1),
            "fp32_mfma": {"ms_per_step": round(ms32, 3), "slices_per_s": round(sps32, 4)},
            "speedup_split_vs_fp32": round(sps16 / sps32, 3),
            "rel_err_vs_fp64": {"split_per_slice": err(res16), "fp32_per_slice": err(res32),
                                "slices_compared": n_ref},
            "kernels_split": {k: {"launches": v["launches"], "ms": round(v["ms"], 3)}
                              for k, v in prof.items()}}


def c2_cpu_baseline(host, threads, budget_s=20.0):
    """Reference CPU path (QFlex transposes + OpenBLAS cgemm, one
    ContractTensors per step, Tensor.hpp:708-752) on the most expensive
    contraction steps of one slice, with their real index orders and shapes
    and constant operand data, until ~budget_s of CPU work."""
    import oracle
    from jet_amd import paths
    plan, net = host
    sl = set(plan["sliced"])
    labels = [[l for l in ix if l not in sl] for ix in net.indices]
    steps = []
    for (a, b) in plan["path"]:
        la, lb = labels[a], labels[b]
        left = [x for x in la if x not in lb]
        right = [x for x in lb if x not in la]
        common = [x for x in la if x in lb]
        labels.append(left + right)
        steps.append((la, lb, 8.0 * 2 ** (len(left) + len(right) + len(common))))
    order = sorted(range(len(steps)), key=lambda i: -steps[i][2])
    oracle.load_blas(threads)
    done_flops, t_used, ran = 0.0, 0.0, 0
    est_rate = 5e11
    for i in order:
        la, lb, fl = steps[i]
        if t_used + fl / est_rate > budget_s and ran > 0:
            continue
        a = np.full(2 ** len(la), 0.5 + 0.25j, np.complex64)
        b = np.full(2 ** len(lb), 0.5 - 0.25j, np.complex64)
        t0 = time.perf_counter()
        oracle.baseline_contract(la, [2] * len(la), a, lb, [2] * len(lb), b, threads)
        dt = time.perf_counter() - t0
        t_used += dt
        done_flops += fl
        ran += 1
        est_rate = done_flops / t_used
        if t_used > budget_s:
            break
    rate = done_flops / t_used
    per_slice = 4.0 * plan["per_slice_pathinfo_flops"]
    return {"value": round(rate / per_slice, 6), "unit": "slices/s", "cores": threads,
            "kind": "port", "tflops": round(rate / 1e12, 4),
            "sample": f"the {ran} most expensive contraction steps of one c2 slice (same plan) "
                      f"that fit {budget_s:.0f} s (QFlex transposes + OpenBLAS cblas_cgemm per "
                      f"step), {done_flops:.3e} flops in {t_used:.1f} s; slices/s = that rate / "
                      f"the per-slice flops (8 M N K per step)"}


C3_CIRCUITS = [((5, 5, 14), 24, 51), ((6, 6, 12), 16, 52), ((6, 7, 10), 24, 53)]


def run_c3(args):
    """Uneven batch: 64 amplitudes of three seeded grid circuits (5x5 depth
    14, 6x6 depth 12, 6x7 depth 10; bitstrings seeded) plus a range of
    --c3-slices slices of the c2 amplitude, all executions captured in one
    hipGraph (jet_amd.batch.UnevenBatch; each amplitude job's form, one
    execution per bitstring or all bitstrings batched -- the stack padded
    to a power of two or not -- chosen by timing; flops count the requested
    amplitudes only).
    One launch is timed with the --split GEMM; the same batch in complex128
    on FP64 MFMA (the c2 part: its first execution only) gives the
    per-amplitude errors."""
    from jet_amd import _native as nat, circuits, paths
    from jet_amd.batch import AmplitudeJob, SliceJob, UnevenBatch

    ctx = nat.default_context()
    rng = np.random.default_rng(2026)
    jobs = []
    for (r, c, d), n_amp, seed in C3_CIRCUITS:
        circ = circuits.random_grid_circuit(r, c, d, seed=seed)
        # min-fill path after subtree reconfiguration (jet_amd/data/amplitude_paths.json)
        path = paths.load_or_make_amplitude_path(r, c, d, seed)
        bits = [tuple(int(x) for x in rng.integers(0, 2, r * c)) for _ in range(n_amp)]
        jobs.append(AmplitudeJob(circ, bits, path))
    plan, circ2, net2 = c2_problem(args.c2_plan)
    sliced = plan["sliced"]
    nb = C2_BATCH_LOG2[args.c2_plan] if args.batch_log2 is None else args.batch_log2
    batch = sliced[len(sliced) - nb:] if nb else []
    per_exec = 2 ** len(batch)
    n_exec = max(1, args.c3_slices // per_exec)
    data32 = [d.astype(np.complex64) for d in net2.data]
    jobs.append(SliceJob(net2.indices, net2.shapes, data32, plan["path"], sliced, batch, 0,
                         n_exec))
    n_amp = sum(len(j.bitstrings) for j in jobs[:-1])
    split = SPLIT_MODES[args.split][1]
    log(f"c3: {n_amp} amplitudes over {len(jobs) - 1} circuits + {n_exec * per_exec} c2 slices")

    t0 = time.perf_counter()
    ub = UnevenBatch(jobs, dtype=np.complex64, precision=split, ctx=ctx)
    log(f"c3: batch captured ({time.perf_counter() - t0:.1f}s), forms {ub.forms()}")
    flops = 0.0
    for job, part in zip(ub.jobs, ub.parts):
        if isinstance(job, SliceJob):
            flops += part.net.flops * job.executions
        else:
            # requested amplitudes only: a stack padded to a power of two
            # does n_stack bitstrings' work for n results
            flops += part.net.flops * len(part.execs) * part.n / part.n_stack
    e0, e1 = nat.Event(ctx), nat.Event(ctx)
    ctx.synchronize()
    e0.record()
    ub.launch()
    e1.record()
    ctx.synchronize()
    ms = e0.elapsed_ms(e1)
    amps = [ub.amplitudes(j) for j in range(len(jobs) - 1)]
    slices = ub.slice_results(len(jobs) - 1)
    log(f"c3: one launch {ms:.1f} ms, {flops / (ms * 1e-3) / 1e12:.1f} TFLOP/s")
    forms, tuning = ub.forms(), ub.tuning
    ub.close()
    del ub
    # The same batch on FP32 MFMA: the complex64 error floor the split run is
    # judged beside.
    ub32 = UnevenBatch(jobs, dtype=np.complex64, precision="fp32", ctx=ctx)
    ub32.launch()
    ctx.synchronize()
    amps32 = [ub32.amplitudes(j) for j in range(len(jobs) - 1)]
    ub32.close()
    del ub32

    # FP64 reference: every amplitude, and the first 16 c2 slices (slice order
    # does not depend on how many sliced indices are batched).
    ref_batch = sliced[len(sliced) - 4:]
    ref_jobs = jobs[:-1] + [SliceJob(net2.indices, net2.shapes, net2.data, plan["path"],
                                     sliced, ref_batch, 0, 1)]
    ub64 = UnevenBatch(ref_jobs, dtype=np.complex128, precision="fp64", ctx=ctx)
    ub64.launch()
    ctx.synchronize()
    errs, errs32, norm_errs = [], [], []
    for j, a in enumerate(amps):
        ref = ub64.amplitudes(j)
        errs.extend((np.abs(a - ref) / np.abs(ref)).tolist())
        errs32.extend((np.abs(amps32[j] - ref) / np.abs(ref)).tolist())
        norm_errs.append(float(np.linalg.norm(a - ref) / np.linalg.norm(ref)))
    ref_s = ub64.slice_results(len(jobs) - 1)
    slice_err = float(np.linalg.norm(slices[:len(ref_s)] - ref_s) / np.linalg.norm(ref_s))
    ub64.close()
    errs, errs32 = np.array(errs), np.array(errs32)
    log(f"c3: per-amplitude rel err max {errs.max():.2e} median {np.median(errs):.2e} "
        f"(fp32 MFMA: max {errs32.max():.2e} median {np.median(errs32):.2e}); "
        f"per circuit normwise {', '.join(f'{x:.2e}' for x in norm_errs)}; "
        f"c2 slices {slice_err:.2e}")
    return {
        "metric": METRIC,
        "value": round(flops / (ms * 1e-3) / 1e12, 3),
        "unit": "TFLOP/s",
        "ms_per_step": round(ms, 3),
        "dtype": f"c64 ({SPLIT_LABEL[args.split]} MFMA on MFMA-bound steps, fp32 MFMA elsewhere)",
        "config": {"workload": "c3: uneven batch in one hipGraph launch: 24 amplitudes of a "
                               "5x5 depth-14, 16 of a 6x6 depth-12 and 24 of a 6x7 depth-10 "
                               f"seeded grid circuit, plus {n_exec * per_exec} slices of the "
                               "c2 7x7 depth-14 amplitude; a step is one launch",
                   "amplitudes": n_amp, "c2_slices": n_exec * per_exec},
        "forms": forms,
        "tuning_ms": tuning,
        "rel_err_vs_fp64": {"per_circuit_normwise": norm_errs,
                            "fp32_amplitude_max": float(errs32.max()),
                            "fp32_amplitude_median": float(np.median(errs32)),
                            "amplitude_max": float(errs.max()),
                            "amplitude_median": float(np.median(errs)),
                            "per_amplitude": [float(x) for x in errs],
                            "c2_first_16_slices": slice_err},
    }, None


def run_c4(args):
    """Gradient of L = |<b|C(theta)|0>|^2 for a seeded 6x6 depth-10 grid
    circuit whose ~300 fSim angles are ~U(0, 2 pi): one forward contraction
    plus the adjoint (environment) pass on the GPU, planned as one
    contraction DAG (every --c4-checkpoint-th forward intermediate kept, the
    others recomputed inside it) and replayed as a hipGraph, through the
    torch.autograd.Function; a step is one loss + gradient evaluation.
    Checked against central differences (h = 1e-3, complex128 on FP64 MFMA)
    on --c4-fd angles (all of them by default)."""
    import torch
    from jet_amd import circuits, paths
    from jet_amd.gradient import AmplitudeLoss, ParameterizedAmplitude

    circ = circuits.random_grid_circuit(6, 6, 10, seed=2026, random_angles=True)
    # min-fill path after subtree reconfiguration (jet_amd/data/amplitude_paths.json)
    path = paths.load_or_make_amplitude_path(6, 6, 10, 2026, random_angles=True)

    def timed(prob):
        theta = torch.tensor(prob.theta, dtype=torch.float64, requires_grad=True)
        for _ in range(args.warmup):
            theta.grad = None
            AmplitudeLoss.apply(theta, prob).backward()
        prob.ctx.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            theta.grad = None
            loss = AmplitudeLoss.apply(theta, prob)
            loss.backward()
        prob.ctx.synchronize()
        return (time.perf_counter() - t0) * 1e3 / args.steps, theta.grad.numpy().copy()

    # Checkpointed recomputation inside the DAG against keeping every
    # forward intermediate: arena bytes and time per loss + gradient.
    checkpointing = {}
    for k in sorted({0, 2, 4, args.c4_checkpoint}):
        if k == args.c4_checkpoint:
            continue
        p = ParameterizedAmplitude(circ, path, checkpoint_every=k)
        ms_k, _ = timed(p)
        checkpointing[str(k)] = {"arena_bytes": p.program.dag.arena_bytes,
                                 "dag_steps": len(p.program.steps),
                                 "recomputed_steps": p.program.recomputed_steps,
                                 "ms_per_step": round(ms_k, 3)}
        log(f"c4: checkpoint_every {k}: arena {p.program.dag.arena_bytes / 2**20:.1f} MiB, "
            f"{ms_k:.1f} ms")
        p.program.dag.close()
        del p
    prob = ParameterizedAmplitude(circ, path, checkpoint_every=args.c4_checkpoint)
    ms, g = timed(prob)
    checkpointing[str(args.c4_checkpoint)] = {
        "arena_bytes": prob.program.dag.arena_bytes, "dag_steps": len(prob.program.steps),
        "recomputed_steps": prob.program.recomputed_steps, "ms_per_step": round(ms, 3)}
    n_fd = prob.n_params if args.c4_fd <= 0 else min(args.c4_fd, prob.n_params)
    base = prob.theta.copy()
    h = 1e-3
    fd = np.zeros(n_fd)
    for k in range(n_fd):
        tp, tm = base.copy(), base.copy()
        tp[k] += h
        tm[k] -= h
        fd[k] = (prob.loss_fp64(tp) - prob.loss_fp64(tm)) / (2 * h)
    err = float(np.linalg.norm(g[:n_fd] - fd) / np.linalg.norm(fd))
    log(f"c4: {prob.n_params} angles, loss+gradient {ms:.1f} ms, |g - fd|/|fd| = {err:.2e}")
    return {
        "metric": METRIC,
        "value": round(1e3 / ms, 3),
        "unit": "gradients/s",
        "ms_per_step": round(ms, 3),
        "dtype": "c64 (FP32 MFMA / VALU per step), gradient accumulated in FP64",
        "config": {"workload": "c4: |<b|C(theta)|0>|^2 and its gradient for all fSim angles "
                               "of a seeded 6x6-qubit depth-10 grid circuit (angles ~ U(0, 2 pi)); "
                               "a step is one forward + adjoint evaluation",
                   "angles": prob.n_params, "path_steps": len(path)},
        "rel_err_vs_central_differences": err,
        "angles_checked": n_fd,
        "method": prob.method,
        "checkpoint_every": args.c4_checkpoint,
        "dag_steps": len(prob.program.steps),
        "checkpointing": checkpointing,
    }, None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="c2", choices=["c1", "c2", "c3", "c4"])
    ap.add_argument("--c2-plan", default="refined", choices=sorted(C2_PLANS),
                    help="c2: contraction plan of the headline (the other one runs as a "
                         "second leg unless --single-plan)")
    ap.add_argument("--c4-fd", type=int, default=0,
                    help="c4: angles checked by central differences (0: all)")
    ap.add_argument("--c4-checkpoint", type=int, default=2,
                    help="c4: keep every k-th forward intermediate, recompute the others "
                         "inside the DAG (0: keep all)")
    ap.add_argument("--c3-slices", type=int, default=256,
                    help="c3: slices of the c2 amplitude in the batch")
    ap.add_argument("--batch-log2", type=int, default=None,
                    help="c2: sliced indices batched per execution (2^b slices per step; "
                         "default 8 on the refined plan, 7 on the min-fill one)")
    ap.add_argument("--split", default="f16", choices=sorted(SPLIT_MODES),
                    help="split 16-bit MFMA GEMM: f16 (split-FP16 4M), 4m (split-BF16 "
                         "4M), 3m (split-BF16 3M)")
    ap.add_argument("--compare", nargs="*", default=["4m"], choices=sorted(SPLIT_MODES),
                    help="c2: other split modes timed for comparison")
    ap.add_argument("--steps32", type=int, default=1,
                    help="c2: executions timed in plain FP32")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--single-plan", dest="other_plan", action="store_false",
                    help="c2: skip the other plan's leg")
    ap.add_argument("--full-amplitude", action="store_true",
                    help="c2: also contract all 4096 slices in the split mode, FP32 and "
                         "FP64 (minutes) and report the amplitude's errors")
    ap.add_argument("--eager", action="store_true",
                    help="c2: launch the timed executions eagerly instead of as one "
                         "captured hipGraph (for kernel tracers)")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--quick", action="store_true",
                    help="only the timed split region and a 2 GiB streaming copy "
                         "(for profiler runs)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); the slice "
              "loop shards across GPUs by contiguous slice ranges, which this build does "
              "not run.", file=sys.stderr)
        sys.exit(2)

    if args.workload == "c1":
        result, host = run_c1(args)
        if args.quick:
            print(json.dumps(result), flush=True)
            return
    elif args.workload == "c3":
        result, host = run_c3(args)
        args.no_cpu_baseline = True
    elif args.workload == "c4":
        result, host = run_c4(args)
        args.no_cpu_baseline = True
    else:
        result, host = run_c2(args)
        if args.quick:
            print(json.dumps(result), flush=True)
            return
    # Every host core this process may run on.
    threads = args.cpu_threads or len(os.sched_getaffinity(0))
    if not args.no_cpu_baseline:
        log("cpu baseline")
        result["cpu_baseline"] = (c1_cpu_baseline(host, threads) if args.workload == "c1"
                                  else c2_cpu_baseline(host, threads))
    else:
        result["cpu_baseline"] = None
    result.update({"n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
                   "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
                   "data": "synthetic (seeded counter-based N(0,1/2))"})
    order = ["metric", "value", "unit", "n_gpus", "steps", "warmup", "ms_per_step",
             "higher_is_better", "scaling", "vs_baseline", "dtype", "data", "config"]
    out = {k: result[k] for k in order if k in result}
    out.update({k: v for k, v in result.items() if k not in out})
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

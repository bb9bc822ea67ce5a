"""GEMM kernel autotuner: measures which kernel and tile variant runs each
GEMM shape class fastest on this device and writes the table the library
compiles in (jet_amd/data/gemm_tuning_gfx950.json).

    python -m jet_amd.autotune [--out PATH] [--modes fp32 f16x3 fp64 bf16x3_4m]

For every precision mode and every point of the shape grid (log2 M, log2 N,
log2 K, log2 batch), ``jet_hip_autotune_gemm`` times every kernel / variant
that can run the batched NT product (split modes include their operand
passes) and the fastest becomes the entry.  The grid spans the shapes the
contraction paths produce: every extent a power of two, operands and result
up to 2^31 elements, at most 2^40 multiply-adds, plus a batched sub-grid and
the large MFMA-bound shapes of the sliced 7x7 circuit.  This replaces
cuTENSOR's internal kernel selection (include/jet/CudaTensor.hpp:616-640).
Needs the GPU; prints a progress line every 50 points (or 30 s) and
rewrites the output after each precision mode.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

from . import _native as nat

HERE = os.path.dirname(os.path.abspath(__file__))
TABLE = os.path.join(HERE, "data", "gemm_tuning_gfx950.json")

# precision name -> (dtype code, precision code, library mode id)
MODES = {"fp32": (nat.C64, nat.PREC_FP32, 0), "f16x3": (nat.C64, nat.PREC_F16X3_4M, 4),
         "fp64": (nat.C128, nat.PREC_FP64, 2), "bf16x3_4m": (nat.C64, nat.PREC_BF16X3_4M, 3)}

GRID_MN = [0, 2, 4, 5, 6, 7, 8, 10, 12, 14, 17, 20, 23]
GRID_K = [0, 1, 2, 3, 4, 5, 6, 7, 8, 10, 12, 14, 16, 20, 24]
BATCH_MN = [0, 4, 8, 12]
BATCH_K = [2, 4, 6, 8]
BATCHES = [4, 8]
# MFMA-bound shapes of the sliced 7x7 depth-14 amplitude beyond the grid's
# work limit (log2 M, N, K).
LARGE = [(14, 17, 14), (14, 14, 14), (13, 15, 7), (12, 19, 8)]


def grid_points(elem_bytes: int):
    """(lm, ln, lk, lb) of the sweep: operands and result <= 2^31 elements
    (2^30 for complex128), <= 2^40 multiply-adds (2^38 for complex128)."""
    cap = 31 if elem_bytes == 8 else 30
    work = 40 if elem_bytes == 8 else 38
    pts = []
    for lm in GRID_MN:
        for ln in GRID_MN:
            for lk in GRID_K:
                if lm + lk <= cap and ln + lk <= cap and lm + ln <= cap and lm + ln + lk <= work:
                    pts.append((lm, ln, lk, 0))
    for lb in BATCHES:
        for lm in BATCH_MN:
            for ln in BATCH_MN:
                for lk in BATCH_K:
                    if lb + max(lm + lk, ln + lk, lm + ln) <= cap and lb + lm + ln + lk <= work:
                        pts.append((lm, ln, lk, lb))
    if elem_bytes == 8:
        pts += [(lm, ln, lk, 0) for (lm, ln, lk) in LARGE]
    return pts


def time_candidates(ctx, dtype_code, prec_code, lm, ln, lk, lb, reps):
    """[(kernel, variant, ms)] of every candidate for one shape."""
    cap = 32
    ks = (ctypes.c_int32 * cap)()
    vs = (ctypes.c_int32 * cap)()
    ms = (ctypes.c_float * cap)()
    cnt = ctypes.c_int()
    nat.check(nat.lib().jet_hip_autotune_gemm(
        ctx.handle, dtype_code, prec_code, 1 << lm, 1 << ln, 1 << lk, 1 << lb, reps, cap, ks, vs,
        ms, ctypes.byref(cnt)))
    return [(int(ks[i]), int(vs[i]), float(ms[i])) for i in range(min(cnt.value, cap))]


def rule_choice(dtype_code, prec_code, lm, ln, lk, lb):
    """The kernel the library's built-in rule picks (table cleared)."""
    k, v, src = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    nat.check(nat.lib().jet_hip_gemm_choice(dtype_code, prec_code, 1 << lm, 1 << ln, 1 << lk,
                                            1 << lb, ctypes.byref(k), ctypes.byref(v),
                                            ctypes.byref(src)))
    return int(k.value), int(v.value)


def sweep(modes, log=print, on_mode=None):
    """Entries [mode, lm, ln, lk, lb, kernel, variant, best_ms, rule_ms] and
    the full timing records; on_mode(entries, records) after each mode."""
    ctx = nat.default_context()
    # The built-in rule's choice is recorded beside each winner: clear the
    # table while sweeping.
    nat.check(nat.lib().jet_hip_tuning_set(0, None))
    entries, records = [], []
    t0 = time.time()
    last = t0
    try:
        for name in modes:
            dtype_code, prec_code, mode_id = MODES[name]
            pts = grid_points(8 if dtype_code == nat.C64 else 16)
            for i, (lm, ln, lk, lb) in enumerate(pts):
                t_point = time.time()
                work = lm + ln + lk + lb
                reps = 5 if work <= 30 else 3 if work <= 36 else 1
                cands = time_candidates(ctx, dtype_code, prec_code, lm, ln, lk, lb, reps)
                kern, var, best = min(cands, key=lambda c: c[2])
                rk, _ = rule_choice(dtype_code, prec_code, lm, ln, lk, lb)
                # the rule's kernel at its own variant choice (variant 0) is
                # the launcher default; report the best variant of that kernel
                rule_ms = min((c[2] for c in cands if c[0] == rk), default=None)
                entries.append([mode_id, lm, ln, lk, lb, kern, var, round(best, 5),
                                None if rule_ms is None else round(rule_ms, 5)])
                records.append({"mode": name, "shape": [lm, ln, lk, lb],
                                "candidates": [[k, v, round(t, 5)] for (k, v, t) in cands]})
                now = time.time()
                if now - t_point > 5.0:
                    log(f"autotune {name}: shape {(lm, ln, lk, lb)} took {now - t_point:.1f} s")
                if (i + 1) % 50 == 0 or i + 1 == len(pts) or now - last > 30.0:
                    log(f"autotune {name}: {i + 1}/{len(pts)} shapes ({now - t0:.0f} s)")
                    last = now
            if on_mode:
                on_mode(entries, records)
    finally:
        nat.check(nat.lib().jet_hip_tuning_set(-1, None))
    return entries, records


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=TABLE)
    ap.add_argument("--records", default=None, help="also write every candidate's time here")
    ap.add_argument("--modes", nargs="*", default=["fp32", "f16x3", "fp64", "bf16x3_4m"],
                    choices=sorted(MODES))
    ap.add_argument("--merge", action="store_true",
                    help="keep the --out file's entries of the modes not swept now")
    args = ap.parse_args(argv)
    kept = []
    if args.merge and os.path.exists(args.out):
        with open(args.out) as f:
            old = json.load(f)
        swept = {MODES[name][2] for name in args.modes}
        kept = [e for e in old["entries"] if e[0] not in swept]

    def write(entries, records):
        entries = kept + entries
        doc = {"device": "gfx950 (MI355X)",
               "columns": ["mode", "log2_m", "log2_n", "log2_k", "log2_batch", "kernel",
                           "variant", "best_ms", "rule_ms"],
               "rule_ms": "the built-in rule's kernel family at its fastest variant",
               "modes": {name: MODES[name][2] for name in MODES},
               "made_by": "python -m jet_amd.autotune (jet_hip_autotune_gemm over the shape grid)",
               "entries": entries}
        with open(args.out, "w") as f:
            json.dump(doc, f, separators=(",", ":"))
        if args.records:
            with open(args.records, "w") as f:
                json.dump(records, f, separators=(",", ":"))

    entries, records = sweep(args.modes, log=lambda m: print(m, flush=True), on_mode=write)
    wins = sum(1 for e in entries if e[8] is not None and e[7] < 0.95 * e[8])
    print(f"wrote {args.out}: {len(entries)} entries; the tuned choice beats the built-in rule "
          f"by more than 5% on {wins}", flush=True)


if __name__ == "__main__":
    sys.exit(main())

"""The autotuned GEMM table and its lookup (host only, no device).

* with the table cleared, plans take the built-in rule;
* an entry for a shape's own class decides it; otherwise the nearest entry
  of the same mode within distance 4 (log2 coordinates) whose choice can run
  the shape; entries of other modes and invalid choices are skipped;
* malformed entries are refused, and the compiled-in table comes back;
* the committed table: every entry names a valid kernel / variant for its
  shape and mode, the library holds exactly those entries, and every entry
  of the table decides its own shape.
"""
import ctypes
import json
import os

import numpy as np
import pytest

from jet_amd import _native as nat

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TABLE = os.path.join(ROOT, "jet_amd", "data", "gemm_tuning_gfx950.json")

F32, SMALL, SPLITK, F16 = 2, 0, 6, 7


def choice(m, n, k, batch=1, prec=nat.PREC_FP32, dtype=nat.C64):
    kern, var, src = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    nat.check(nat.lib().jet_hip_gemm_choice(dtype, prec, m, n, k, batch, ctypes.byref(kern),
                                            ctypes.byref(var), ctypes.byref(src)))
    return kern.value, var.value, src.value


def set_table(rows):
    flat = [int(x) for r in rows for x in r]
    arr = (ctypes.c_int32 * max(1, len(flat)))(*flat)
    return nat.lib().jet_hip_tuning_set(len(rows), arr)


@pytest.fixture
def table():
    yield
    nat.check(nat.lib().jet_hip_tuning_set(-1, None))


def test_rule_when_cleared(table):
    nat.check(nat.lib().jet_hip_tuning_set(0, None))
    assert choice(4096, 4096, 4096) == (F32, 0, 0)
    assert choice(4096, 4096, 4096, prec=nat.PREC_F16X3_4M) == (F16, 0, 0)
    assert choice(8, 8, 4) == (SMALL, 0, 0)


def test_exact_nearest_and_skipped_entries(table):
    rows = [
        [0, 12, 12, 12, 0, SMALL, 1],    # fp32 (4096, 4096, 4096): VALU one thread per output
        [0, 8, 8, 4, 0, F32, 3],         # fp32 (256, 256, 16): FP32 MFMA 128x64
        [4, 12, 12, 12, 0, F32, 2],      # split-FP16 mode entry: not used for fp32 queries
        [0, 20, 20, 3, 0, SMALL, 2],     # short-K columns: invalid for K = 32
        [0, 20, 20, 5, 0, F32, 1],
    ]
    nat.check(set_table(rows))
    assert choice(4096, 4096, 4096) == (SMALL, 1, 1)
    assert choice(256, 256, 16) == (F32, 3, 1)
    # nearest: (512, 256, 16) is at distance 1 from (8, 8, 4)
    assert choice(512, 256, 16) == (F32, 3, 2)
    # split-FP16 query takes the mode-4 entry
    assert choice(4096, 4096, 4096, prec=nat.PREC_F16X3_4M) == (F32, 2, 1)
    # (2^20, 2^20, 32): the K = 8 short-K entry cannot run K = 32; the
    # K = 32 entry decides it exactly
    assert choice(1 << 20, 1 << 20, 32) == (F32, 1, 1)
    # (2^20, 2^20, 16): the short-K entry (distance 1) wins over the K = 32
    # one (distance 1, later in the table)
    assert choice(1 << 20, 1 << 20, 16) == (SMALL, 2, 2)
    # far from every entry: the rule
    assert choice(2, 2, 1 << 16) == (SPLITK, 0, 0)
    # a fractional log2 (non-power-of-two extents) still finds its neighbour
    assert choice(4000, 4100, 4096)[:2] == (SMALL, 1)


def test_split_kernels_not_chosen_past_their_k_limit(table):
    """The 4M split kernels keep 32-bit per-lane load offsets, valid for a
    padded k up to 2^21: beyond it neither the rule nor a table entry that
    names them may pick them."""
    f16 = dict(prec=nat.PREC_F16X3_4M)
    nat.check(nat.lib().jet_hip_tuning_set(0, None))
    assert choice(256, 256, 1 << 21, **f16)[0] == F16
    assert choice(256, 256, (1 << 21) + 1, **f16)[0] == F32
    assert choice(256, 256, 1 << 22, **f16)[0] == F32
    nat.check(set_table([[4, 8, 8, 22, 0, F16, 2], [4, 8, 8, 21, 0, F16, 1]]))
    assert choice(256, 256, 1 << 21, **f16) == (F16, 1, 1)
    assert choice(256, 256, 1 << 22, **f16)[0] != F16


def test_malformed_entries_refused(table):
    assert set_table([[9, 1, 1, 1, 0, F32, 1]]) != 0
    assert "precision mode" in nat.lib().jet_hip_last_error().decode()
    assert set_table([[0, 1, 1, 1, 0, F32, 4]]) != 0
    assert set_table([[0, 1, 1, 1, 0, 12, 1]]) != 0


def test_reset_restores_builtin_table(table):
    n0 = ctypes.c_int()
    nat.check(nat.lib().jet_hip_tuning_get(0, None, ctypes.byref(n0)))
    nat.check(set_table([[0, 1, 1, 1, 0, F32, 1]]))
    n1 = ctypes.c_int()
    nat.check(nat.lib().jet_hip_tuning_get(0, None, ctypes.byref(n1)))
    assert n1.value == 1
    nat.check(nat.lib().jet_hip_tuning_set(-1, None))
    n2 = ctypes.c_int()
    nat.check(nat.lib().jet_hip_tuning_get(0, None, ctypes.byref(n2)))
    assert n2.value == n0.value


def test_committed_table_is_compiled_in_and_decides_its_shapes():
    if not os.path.exists(TABLE):
        pytest.skip("no committed tuning table")
    with open(TABLE) as f:
        doc = json.load(f)
    entries = doc["entries"]
    assert entries
    n = ctypes.c_int()
    nat.check(nat.lib().jet_hip_tuning_get(0, None, ctypes.byref(n)))
    buf = (ctypes.c_int32 * (7 * n.value))()
    nat.check(nat.lib().jet_hip_tuning_get(n.value, buf, ctypes.byref(n)))
    lib_rows = np.array(buf[:], dtype=np.int64).reshape(n.value, 7)
    assert np.array_equal(lib_rows, np.array([e[:7] for e in entries], dtype=np.int64))
    prec = {0: (nat.C64, nat.PREC_FP32), 2: (nat.C128, nat.PREC_FP64),
            3: (nat.C64, nat.PREC_BF16X3_4M), 4: (nat.C64, nat.PREC_F16X3_4M)}
    for e in entries:
        mode, lm, ln, lk, lb, kern, var = e[:7]
        dt, pc = prec[mode]
        got = choice(1 << lm, 1 << ln, 1 << lk, 1 << lb, prec=pc, dtype=dt)
        assert got == (kern, var, 1), (e, got)

"""The fused small-contraction pass (plan_fused_small, fused_small_kernel),
checked on the CPU before anything runs on a GPU: an emulation of the
kernel's offset arithmetic (tests/native/network_check.cpp:
fused_small_emulate) on random bond-2 and power-of-two contractions with
scrambled mode orders, batch modes and permuted results must equal numpy's
einsum; the planner fuses small permuted steps (no workspace) and leaves
GEMM-sized ones, unpermuted ones and non-power-of-two extents alone."""
import ctypes

import numpy as np
import pytest

from jet_amd import _native as nat
from test_circuits_paths import _check_lib

LETTERS = "abcdefghijklmnopqrstuvwxyz"


def _random_case(rng):
    nb, nl, nr, nk = (int(x) for x in rng.integers(0, 3, 4))
    nl, nr, nk = nl + 1, nr + 1, nk + 1
    ids = list(range(nb + nl + nr + nk))
    rng.shuffle(ids)
    bat, lef = ids[:nb], ids[nb:nb + nl]
    rig, com = ids[nb + nl:nb + nl + nr], ids[nb + nl + nr:]
    ext = {m: int(2 ** rng.integers(0, 3)) for m in ids}
    ma = [int(x) for x in rng.permutation(bat + lef + com)]
    mb = [int(x) for x in rng.permutation(bat + rig + com)]
    mc = [int(x) for x in rng.permutation(bat + lef + rig)]
    return ma, mb, mc, ext


def _emulate(L, ma, mb, mc, ext, a, b):
    ea, eb = [ext[m] for m in ma], [ext[m] for m in mb]
    c = np.zeros(int(np.prod([ext[m] for m in mc])), np.complex128)
    rc = L.fused_small_emulate(len(ma), nat.i32(ma), nat.i64(ea), len(mb), nat.i32(mb),
                               nat.i64(eb), len(mc), nat.i32(mc),
                               a.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                               b.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                               c.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    return rc, c


@pytest.mark.parametrize("seed", range(40))
def test_fused_offsets_match_einsum(seed):
    L = _check_lib()
    rng = np.random.default_rng(seed)
    ma, mb, mc, ext = _random_case(rng)
    na, nb_ = int(np.prod([ext[m] for m in ma])), int(np.prod([ext[m] for m in mb]))
    a = rng.standard_normal(na) + 1j * rng.standard_normal(na)
    b = rng.standard_normal(nb_) + 1j * rng.standard_normal(nb_)
    rc, c = _emulate(L, ma, mb, mc, ext, a, b)
    assert rc == 0
    sub = lambda ms: "".join(LETTERS[m] for m in ms)
    want = np.einsum(f"{sub(ma)},{sub(mb)}->{sub(mc)}",
                     a.reshape([ext[m] for m in ma]), b.reshape([ext[m] for m in mb]))
    np.testing.assert_allclose(c, want.reshape(-1), rtol=1e-12, atol=1e-12)


def test_fused_declines_non_power_of_two():
    L = _check_lib()
    ext = {0: 3, 1: 2, 2: 2}
    a = np.ones(6, np.complex128)
    b = np.ones(4, np.complex128)
    rc, _ = _emulate(L, [0, 1], [1, 2], [2, 0], ext, a, b)
    assert rc == 1


def _is_fused(L, ma, ea, mb, eb, mc):
    ws = ctypes.c_size_t()
    r = L.contraction_is_fused(len(ma), nat.i32(ma), nat.i64(ea), len(mb), nat.i32(mb),
                               nat.i64(eb), len(mc), nat.i32(mc), ctypes.byref(ws))
    return r, ws.value


def test_planner_fuses_small_permuted_steps():
    L = _check_lib()
    # few outputs, long K (the one-block-per-output kernel): 2 x 4 outputs,
    # K = 2^10, scrambled
    r, ws = _is_fused(L, [3, 10, 0, 5, 11, 1, 7, 9, 2, 8, 4, 6], [2] * 12,
                      [6, 12, 2, 9, 4, 13, 8, 1, 3, 7, 5, 0], [2] * 12, [11, 13, 12, 10])
    assert r == 1 and ws == 0
    # bond-2 rank-6 x rank-5, two shared modes, scrambled orders: tiny, permuted
    r, ws = _is_fused(L, [0, 3, 1, 4, 2, 5], [2] * 6, [5, 6, 4, 7, 8], [2] * 5,
                      [8, 0, 6, 1, 2, 3, 7])
    assert r == 1 and ws == 0
    # already in GEMM order: nothing to permute, no fusion
    r, _ = _is_fused(L, [0, 1, 2], [2] * 3, [3, 2], [2] * 2, [0, 1, 3])
    assert r == 0
    # a 2^10-cube GEMM is not a small step
    r, _ = _is_fused(L, [1, 0], [1024, 1024], [2, 1], [1024, 1024], [2, 0])
    assert r == 0

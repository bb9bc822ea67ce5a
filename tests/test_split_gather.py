"""The split-FP16 gather plan (a permutation folded into the split pass),
checked on the CPU before anything runs on a GPU: an emulation of the
kernel's index mapping (tests/native/network_check.cpp:split_gather_targets)
must send every source element to the row and k of the permuted operand
[rows][K] that numpy's transpose gives, each target exactly once; and the
gather declines sources whose lowest address bits do not fit in a tile."""
import ctypes

import numpy as np
import pytest

from jet_amd import _native as nat
from test_circuits_paths import _check_lib


def _targets(L, ext, perm, K):
    """Gather emulation for source extents `ext` (modes 0..r-1, row-major)
    and target order `perm` (modes); None when the gather does not apply."""
    n = int(np.prod(ext))
    out = np.full(n, -1, np.int64)
    info = (ctypes.c_int * 3)()
    rc = L.split_gather_targets(len(ext), nat.i32(list(range(len(ext)))), nat.i64(ext),
                                len(perm), nat.i32(perm), ctypes.c_int64(K),
                                out.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), info)
    return (out, tuple(info)) if rc == 0 else None


@pytest.mark.parametrize("seed", range(12))
def test_gather_mapping_matches_transpose(seed):
    L = _check_lib()
    rng = np.random.default_rng(seed)
    r = int(rng.integers(10, 17))
    ext = [2] * r
    if seed % 3 == 0:                      # a few wider power-of-two modes
        ext[int(rng.integers(r))] = 4
        ext[int(rng.integers(r))] = 8
    nb = int(np.log2(np.prod(ext)))
    # K modes: a random subset of modes whose extents multiply to 32..2048
    for _ in range(100):
        perm = [int(x) for x in rng.permutation(r)]
        kb, nk = 0, 0
        while nk < r and kb + int(np.log2(ext[perm[-1 - nk]])) <= 10:
            kb += int(np.log2(ext[perm[-1 - nk]]))
            nk += 1
            if kb >= 5 and rng.random() < 0.3:
                break
        if kb >= 5 and nb - kb >= 4:
            break
    K = 2 ** kb
    # the source's lowest bits go to row (not K) modes half of the time
    res = _targets(L, ext, perm, K)
    # numpy: target position j holds source element want[j]
    want = np.arange(2 ** nb).reshape(ext).transpose(perm).reshape(-1)
    if res is None:
        # declined: the source's four lowest bits must not fit in a tile
        # with the K bits (2048 elements)
        kmodes = set(perm[len(perm) - nk:])
        bit_mode = [m for m in reversed(range(r)) for _ in range(int(np.log2(ext[m])))]
        budget, covered = 11 - kb, 0
        for m in bit_mode:
            if m not in kmodes:
                if budget == 0:
                    break
                budget -= 1
            covered += 1
        assert covered < 4
        return
    out, (tb, xr, nw) = res
    assert tb == kb + xr and tb + nw == nb and tb <= 11
    got = np.empty_like(out)
    got[out] = np.arange(out.size)          # target -> source
    assert np.array_equal(np.sort(out), np.arange(out.size))
    assert np.array_equal(got, want)


def test_gather_declines_scattered_sources():
    """When the source's four lowest bits are row bits that cannot all join
    a 2048-element tile with the K bits (K = 512 leaves room for two), the
    plan falls back to a permuted copy."""
    L = _check_lib()
    ext = [2] * 16
    # source modes 12..15 are its lowest bits; put them first in the target
    perm = [12, 13, 14, 15] + list(range(12))
    assert _targets(L, ext, perm, 512) is None
    # with K = 128 the tile has room for four row bits: it applies
    perm = [12, 13, 14, 15] + list(range(12))
    res = _targets(L, ext, perm, 128)
    assert res is not None and res[1][1] == 4

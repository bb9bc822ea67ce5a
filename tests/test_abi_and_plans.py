"""CPU checks of the C-ABI library and of the host-side plans.

* libjet_hip.so loads and exports every function include/jet_hip.h declares
  (no compute calls: this container has no GPU);
* permutation plans, replayed on the CPU with the kernels' own index
  helpers (tests/native/permute_emulation.cpp), stay in bounds, write every
  output element exactly once and match the oracle.
"""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

import oracle
from jet_amd import _native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EMU_SRC = os.path.join(ROOT, "tests", "native", "permute_emulation.cpp")
EMU_LIB = os.path.join(ROOT, "tests", "native", "_build", "libemu.so")


def _declared_functions():
    text = open(os.path.join(ROOT, "include", "jet_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(jet_hip_\w+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = _native.lib()
    declared = _declared_functions()
    assert len(declared) >= 20
    for name in declared:
        assert hasattr(lib, name), name
    assert sorted(_native.EXPORTS) == declared
    out = subprocess.run(["nm", "-D", "--defined-only", _native.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = set(re.findall(r"\bT (jet_hip_\w+)", out))
    assert set(declared) <= exported


def test_version_and_error_without_gpu():
    lib = _native.lib()
    a, b, c = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    assert lib.jet_hip_get_version(ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)) == 0
    assert (a.value, b.value) == (0, 1)


def _emu_lib():
    if not os.path.exists(EMU_LIB) or os.path.getmtime(EMU_LIB) < max(
            os.path.getmtime(EMU_SRC),
            os.path.getmtime(os.path.join(ROOT, "jet_amd", "csrc", "permute_plan.cpp")),
            os.path.getmtime(os.path.join(ROOT, "jet_amd", "csrc", "permute_plan.hpp"))):
        os.makedirs(os.path.dirname(EMU_LIB), exist_ok=True)
        subprocess.run(["g++", "-std=c++17", "-O2", "-fPIC", "-shared", "-I" + os.path.join(ROOT, "include"),
                        "-o", EMU_LIB, EMU_SRC,
                        os.path.join(ROOT, "jet_amd", "csrc", "permute_plan.cpp")], check=True)
    return ctypes.CDLL(EMU_LIB)


def _emulate(L, x, shape, perm):
    out = np.zeros_like(x)
    k, tb = ctypes.c_int(), ctypes.c_int()
    err = ctypes.create_string_buffer(256)
    rc = L.emu_permute(x.itemsize, len(shape), _native.i64(shape), _native.i32(perm),
                       ctypes.c_void_p(x.ctypes.data), ctypes.c_void_p(out.ctypes.data),
                       ctypes.byref(k), ctypes.byref(tb), err, 256)
    return rc, out, k.value, tb.value


@pytest.mark.parametrize("dtype", [np.complex64, np.complex128])
def test_permute_plans_in_bounds_and_exact(dtype):
    L = _emu_lib()
    rng = np.random.default_rng(1 if dtype == np.complex64 else 2)
    kinds = set()
    for _ in range(300):
        r = int(rng.integers(1, 22))
        shape = [int(rng.choice([1, 2, 2, 2, 2, 4, 8, 3, 5])) for _ in range(r)]
        while np.prod(shape) > 2 ** 17:
            shape.pop()
        perm = [int(p) for p in rng.permutation(len(shape))]
        n = int(np.prod(shape)) if shape else 1
        x = (rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(dtype)
        rc, out, kind, tb = _emulate(L, x, shape, perm)
        assert rc == 0, (shape, perm, rc)
        want = oracle.permute_default(x, shape, perm) if shape else x
        assert np.array_equal(out, want), (shape, perm)
        kinds.add(kind)
    assert kinds == {0, 1, 2}


def test_permute_plan_rank28_bond2():
    L = _emu_lib()
    rng = np.random.default_rng(9)
    shape = [2] * 22
    perm = [int(p) for p in rng.permutation(22)]
    x = rng.standard_normal(2 ** 22).astype(np.complex64)
    rc, out, kind, tb = _emulate(L, x, shape, perm)
    assert rc == 0 and kind == 1 and tb == 11
    assert np.array_equal(out, oracle.permute_default(x, shape, perm))


def _bits_plan(L, shape, perm):
    sc = (ctypes.c_int * 4)()
    arrs = [(ctypes.c_int8 * n)() for n in (16, 16, 16, 64, 64)]
    kind = L.emu_bits_plan(len(shape), _native.i64(shape), _native.i32(perm), sc, *arrs)
    nbits, tb, ti, nr = list(sc)
    tile_in, tile_out, v2u, outer_in, outer_out = [list(a) for a in arrs]
    return kind, nbits, tb, ti, nr, tile_in[:tb], tile_out[:tb], v2u[:tb], outer_in[:nr], \
        outer_out[:nr]


@pytest.mark.parametrize("rank", [24, 28, 33, 36, 40, 48, 62])
def test_bit_plans_of_large_bond2_permutations(rank):
    """For bond-2 tensors of rank up to 62 (a rank-40 complex64 tensor is
    8 TiB, beyond one GPU; rank 33, 64 GiB, runs in the GPU tests), the bit
    plan the kernels execute is exactly the permutation: input bit
    rank-1-perm[i] goes to output bit rank-1-i.  Tile and outer bits
    partition the input bits, outer bits and tile bits (through v2u) map by
    that permutation, the tile's low input bits are contiguous and its size
    fits LDS -- so every load and store is in bounds and every output
    element is written once, as the kernels' index helpers compute them."""
    L = _emu_lib()
    rng = np.random.default_rng(rank)
    for trial in range(20):
        perm = [int(p) for p in rng.permutation(rank)]
        if trial == 0:
            perm = list(reversed(range(rank)))
        kind, nbits, tb, ti, nr, t_in, t_out, v2u, o_in, o_out = _bits_plan(L, [2] * rank, perm)
        assert kind == 1 and nbits == rank and tb == 11 and tb + nr == nbits
        sigma = {rank - 1 - perm[i]: rank - 1 - i for i in range(rank)}
        assert sorted(t_in + o_in) == list(range(nbits))
        assert t_in == sorted(t_in) and t_in[:ti] == list(range(ti)) and ti >= 1
        assert o_out == [sigma[b] for b in o_in]
        assert sorted(v2u) == list(range(tb))
        assert t_out == [sigma[t_in[u]] for u in v2u]
        assert t_out == sorted(t_out)

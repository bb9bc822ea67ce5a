"""The single large pairwise contraction (rank 28 x rank 28, 14 shared indices).

A and B are complex64 tensors of rank 28 (all extents 2, 2^28 elements
each) filled on the device by the counter-based generator (re, im ~ N(0,
1/2)); the 14 shared labels sit at seeded random positions in both index
orders (bench.c1_problem).  The contraction (Tensor.hpp:708-752) is one
2^14 x 2^14 x 2^14 complex GEMM behind the operand permutations.  Checked:

  * split-FP16, FP32 MFMA and complex128 on FP64 MFMA against exact FP64 dot
    products of 64 sampled output entries, computed on the host from the
    unpermuted operands by index arithmetic alone (normwise over the
    samples: 1e-5 for complex64, 1e-12 for complex128);
  * the permutation of A to [left, common] order (QFlex.hpp's job at this
    size) for exact equality on 2^20 sampled output positions;
  * a rank-33 bond-2 permutation (2^33 elements, 64 GiB in and out) for
    exact equality on sampled output runs, against the generator's values
    recomputed on the host (uniform [-1, 1): exact in float32).
"""
import ctypes

import numpy as np
import pytest

import bench
from jet_amd import _native as nat
from jet_amd.tensor import _relabel

from hiputil import say

pytestmark = pytest.mark.gpu

N28 = 1 << 28


def _counter_uniform(seed, counter):
    """The library's documented counter generator (splitmix64, 53 bits)."""
    M = np.uint64(0xFFFFFFFFFFFFFFFF)

    def sm(z):
        z = (z + np.uint64(0x9E3779B97F4A7C15)) & M
        z = ((z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)) & M
        z = ((z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)) & M
        return z ^ (z >> np.uint64(31))

    with np.errstate(over="ignore"):
        x = sm(np.uint64(seed) ^ sm(counter.astype(np.uint64)))
    return (x >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def _uniform_c64(seed, idx):
    idx = np.asarray(idx, dtype=np.uint64)
    re = 2.0 * _counter_uniform(seed, 2 * idx) - 1.0
    im = 2.0 * _counter_uniform(seed, 2 * idx + 1) - 1.0
    return (re.astype(np.float32) + 1j * im.astype(np.float32)).astype(np.complex64)


def _src_index(out_idx, rank, perm):
    """Input linear index of output linear index (bond 2, row-major):
    output axis i is input axis perm[i]."""
    src = np.zeros_like(out_idx)
    for i in range(rank):
        bit = (out_idx >> np.int64(rank - 1 - i)) & 1
        src |= bit << np.int64(rank - 1 - perm[i])
    return src


def _plan(ctx, dtype, prec, ma, mb, mc):
    ext = [2] * 28
    p = ctypes.c_void_p()
    nat.check(nat.lib().jet_hip_contraction_plan_create(
        ctx.handle, dtype, prec, nat.ROW_MAJOR, 28, nat.i32(ma), nat.i64(ext), 28, nat.i32(mb),
        nat.i64(ext), 28, nat.i32(mc), ctypes.byref(p)))
    ws = ctypes.c_size_t()
    nat.check(nat.lib().jet_hip_plan_workspace_size(p, ctypes.byref(ws)))
    return p, ws.value


@pytest.fixture(scope="module")
def c1(ctx):
    la, lb = bench.c1_problem()
    left = [x for x in la if x not in lb]
    right = [x for x in lb if x not in la]
    ma, mb, mc = _relabel(la, lb, left + right)
    A = nat.DeviceBuffer(N28 * 8, ctx)
    B = nat.DeviceBuffer(N28 * 8, ctx)
    say("c1: fill A, B (2^28 complex64 each)")
    nat.check(nat.lib().jet_hip_fill_random(ctx.handle, nat.C64, N28, 1, 2, A.ptr, None))
    nat.check(nat.lib().jet_hip_fill_random(ctx.handle, nat.C64, N28, 2, 2, B.ptr, None))
    a_host = np.empty(N28, np.complex64)
    b_host = np.empty(N28, np.complex64)
    A.download(a_host)
    B.download(b_host)
    rng = np.random.default_rng(11)
    rows = rng.integers(0, 1 << 14, 64)
    cols = rng.integers(0, 1 << 14, 64)
    exact = bench.c1_sampled_reference(a_host, b_host, la, lb, rows, cols)
    yield dict(la=la, lb=lb, left=left, ma=ma, mb=mb, mc=mc, A=A, B=B, a_host=a_host,
               rows=rows, cols=cols, exact=exact, b_host=b_host)
    A.free()
    B.free()


@pytest.mark.parametrize("prec,tol", [(nat.PREC_F16X3_4M, 1e-5), (nat.PREC_FP32, 1e-5),
                                      (nat.PREC_BF16X3_4M, 5e-5)])
def test_c1_complex64(ctx, c1, prec, tol):
    say(f"c1: contraction, precision {prec}")
    p, wsz = _plan(ctx, nat.C64, prec, c1["ma"], c1["mb"], c1["mc"])
    try:
        C = nat.DeviceBuffer(N28 * 8, ctx)
        ws = nat.DeviceBuffer(max(wsz, 256), ctx)
        nat.check(nat.lib().jet_hip_contract(p, c1["A"].ptr, c1["B"].ptr, C.ptr, ws.ptr, wsz,
                                             None))
        c = np.empty(N28, np.complex64)
        C.download(c)
    finally:
        nat.lib().jet_hip_plan_destroy(p)
    got = c.reshape(1 << 14, 1 << 14)[c1["rows"], c1["cols"]].astype(np.complex128)
    err = np.linalg.norm(got - c1["exact"]) / np.linalg.norm(c1["exact"])
    say(f"  sampled normwise error {err:.2e}")
    assert err < tol


def test_c1_complex128(ctx, c1):
    say("c1: complex128 on FP64 MFMA")
    A64 = nat.DeviceBuffer(N28 * 16, ctx)
    B64 = nat.DeviceBuffer(N28 * 16, ctx)
    nat.check(nat.lib().jet_hip_convert(ctx.handle, nat.C64, nat.C128, N28, c1["A"].ptr,
                                        A64.ptr, None))
    nat.check(nat.lib().jet_hip_convert(ctx.handle, nat.C64, nat.C128, N28, c1["B"].ptr,
                                        B64.ptr, None))
    p, wsz = _plan(ctx, nat.C128, nat.PREC_FP64, c1["ma"], c1["mb"], c1["mc"])
    try:
        C = nat.DeviceBuffer(N28 * 16, ctx)
        ws = nat.DeviceBuffer(max(wsz, 256), ctx)
        nat.check(nat.lib().jet_hip_contract(p, A64.ptr, B64.ptr, C.ptr, ws.ptr, wsz, None))
        c = np.empty(N28, np.complex128)
        C.download(c)
    finally:
        nat.lib().jet_hip_plan_destroy(p)
    got = c.reshape(1 << 14, 1 << 14)[c1["rows"], c1["cols"]]
    err = np.linalg.norm(got - c1["exact"]) / np.linalg.norm(c1["exact"])
    say(f"  sampled normwise error {err:.2e}")
    assert err < 1e-12


def test_c1_operand_permutation_exact(ctx, c1):
    la = c1["la"]
    common = [x for x in la if x in c1["lb"]]
    new = c1["left"] + common
    perm = [la.index(x) for x in new]
    say("c1: permute A (rank 28) to [left, common]")
    out = nat.DeviceBuffer(N28 * 8, ctx)
    nat.check(nat.lib().jet_hip_permute(ctx.handle, nat.C64, nat.ROW_MAJOR, 28,
                                        nat.i32(range(28)), nat.i64([2] * 28), nat.i32(perm),
                                        c1["A"].ptr, out.ptr, None))
    got = np.empty(N28, np.complex64)
    out.download(got)
    rng = np.random.default_rng(5)
    idx = rng.integers(0, N28, 1 << 20).astype(np.int64)
    src = _src_index(idx, 28, perm)
    assert np.array_equal(got[idx], c1["a_host"][src])


def test_rank33_permutation_exact(ctx):
    r = 33
    n = 1 << r
    rng = np.random.default_rng(33)
    perm = [int(p) for p in rng.permutation(r)]
    say(f"rank-33 bond-2 permutation (64 GiB in, 64 GiB out), perm {perm}")
    x = nat.DeviceBuffer(n * 8, ctx)
    y = nat.DeviceBuffer(n * 8, ctx)
    try:
        nat.check(nat.lib().jet_hip_fill_random(ctx.handle, nat.C64, n, 77, 1, x.ptr, None))
        nat.check(nat.lib().jet_hip_permute(ctx.handle, nat.C64, nat.ROW_MAJOR, r,
                                            nat.i32(range(r)), nat.i64([2] * r), nat.i32(perm),
                                            x.ptr, y.ptr, None))
        ctx.synchronize()
        run = 4096
        starts = [0, n - run] + [int(s) * run for s in rng.integers(0, n // run, 62)]
        host = np.empty(run, np.complex64)
        for s in starts:
            nat.check(nat.lib().jet_hip_memcpy(ctx.handle, host.ctypes.data, y.ptr + s * 8,
                                               run * 8, 1, None, 0))
            out_idx = np.arange(s, s + run, dtype=np.int64)
            want = _uniform_c64(77, _src_index(out_idx, r, perm))
            assert np.array_equal(host, want), s
    finally:
        x.free()
        y.free()

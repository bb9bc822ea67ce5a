"""Parity of the HIP kernels with the oracle, through the C ABI.

Bit-exact for permutations, slices, copies and adds; FP contractions within
the tolerances stated per test (normwise relative error against the FP64
oracle): 1e-5 for complex64 on FP32 MFMA and on split-BF16, 1e-12 for
complex128 on FP64 MFMA.
"""
import ctypes
import string

import numpy as np
import pytest

import oracle
from jet_amd import _native as nat
from jet_amd import Permuter, Tensor

from hiputil import code, contract, from_device, permute, rand_c, rel_err, say

pytestmark = pytest.mark.gpu

TOL = {nat.PREC_FP32: 1e-5, nat.PREC_BF16X3: 1e-5, nat.PREC_FP64: 1e-12,
       nat.PREC_DEFAULT: 1e-5, nat.PREC_BF16X3_4M: 1e-5, nat.PREC_F16X3_4M: 2e-6}


# --------------------------------------------------------------------------
# permutations
# --------------------------------------------------------------------------

@pytest.mark.parametrize("dtype", [np.complex64, np.complex128])
def test_permuter_reference_fixtures(ctx, golden_permuter, dtype):
    p = Permuter(ctx)
    for case in golden_permuter["cases"]:
        say(f"permuter fixture {case['shape']} {case['old_labels']} {np.dtype(dtype)}")
        data = np.array([complex(a, b) for a, b in golden_permuter["arrays"][case["input"]]],
                        dtype=dtype)
        exp = np.array([complex(a, b) for a, b in case["expected"]], dtype=dtype)
        got = p.transpose(data, case["shape"], case["old_labels"], case["new_labels"])
        assert np.array_equal(got, exp), case
        out = np.zeros_like(data)
        p.transpose(data, case["shape"], case["old_labels"], case["new_labels"], out)
        assert np.array_equal(out, exp)


@pytest.mark.parametrize("dtype", [np.complex64, np.complex128])
def test_permute_random_shapes(ctx, dtype):
    rng = np.random.default_rng(21)
    for t in range(120):
        r = int(rng.integers(1, 20))
        shape = [int(rng.choice([1, 2, 2, 2, 2, 4, 8, 3, 5])) for _ in range(r)]
        while np.prod(shape) > 2 ** 18:
            shape.pop()
        perm = [int(p) for p in rng.permutation(len(shape))]
        n = int(np.prod(shape))
        x = rand_c(rng, n, dtype)
        say(f"permute {shape} {perm}")
        got = permute(ctx, x, shape, list(range(len(shape))), perm)
        assert np.array_equal(got, oracle.permute_default(x, shape, perm)), (shape, perm)


def test_permute_column_major_is_reversed_row_major(ctx):
    rng = np.random.default_rng(4)
    shape = [2, 3, 4, 5]
    x = rand_c(rng, 120, np.complex64)
    modes_in, modes_out = [10, 11, 12, 13], [12, 10, 13, 11]
    say("permute col-major")
    got = permute(ctx, x, shape, modes_in, modes_out, layout=nat.COL_MAJOR)
    # column-major (i0 fastest) == row-major with modes and shape reversed
    rm_in = modes_in[::-1]
    want = oracle.permute_default(x, shape[::-1], [rm_in.index(m) for m in modes_out[::-1]])
    assert np.array_equal(got, want)


@pytest.mark.parametrize("dtype", [np.complex64, np.complex128])
def test_identity_permutation_copy(ctx, dtype):
    """The identity permutation is the library's streaming copy: every size
    (16-byte vectors plus an odd complex64 tail), 16-byte-aligned and
    8-byte-offset pointers, copied exactly."""
    rng = np.random.default_rng(5)
    L = nat.lib()
    for n in [1, 2, 3, 255, 1023, 1025, 4097, (1 << 20) + 1]:
        x = rand_c(rng, n, dtype)
        for offset in ([0, 8] if dtype == np.complex64 else [0]):
            src = nat.DeviceBuffer(x.nbytes + 16, ctx)
            dst = nat.DeviceBuffer(x.nbytes + 16, ctx)
            nat.check(L.jet_hip_memset(ctx.handle, dst.ptr, 0x7f, x.nbytes + 16, None))
            nat.check(L.jet_hip_memcpy(ctx.handle, src.ptr + offset, x.ctypes.data, x.nbytes, 0,
                                       None, 0))
            say(f"identity copy n={n} {np.dtype(dtype).name} offset {offset}")
            nat.check(L.jet_hip_permute(ctx.handle, code(dtype), nat.ROW_MAJOR, 1, nat.i32([0]),
                                        nat.i64([n]), nat.i32([0]), src.ptr + offset,
                                        dst.ptr + offset, None))
            ctx.synchronize()
            host = np.empty(x.nbytes + 16, np.uint8)
            dst.download(host)
            got = host[offset:offset + x.nbytes].view(dtype)
            assert np.array_equal(got, x), (n, offset)
            # nothing written past the copy
            assert np.all(host[offset + x.nbytes:] == 0x7f) and np.all(host[:offset] == 0x7f)
            src.free()
            dst.free()


def test_permute_large_bond2(ctx):
    rng = np.random.default_rng(8)
    r = 24
    perm = [int(p) for p in rng.permutation(r)]
    x = rand_c(rng, 2 ** r, np.complex64)
    say("permute rank-24 bond-2")
    got = permute(ctx, x, [2] * r, list(range(r)), perm)
    want = np.ascontiguousarray(np.transpose(x.reshape([2] * r), perm)).reshape(-1)
    assert np.array_equal(got, want)


# --------------------------------------------------------------------------
# contractions
# --------------------------------------------------------------------------

@pytest.mark.parametrize("dtype", [np.complex64, np.complex128])
def test_contract_reference_literals(ctx, dtype):
    say("contract literals")
    # Test_Tensor.cpp:477-563
    t1 = Tensor(["a", "b"], [2, 2], [0, 1, 1, 0], dtype=dtype)
    t2 = Tensor(["b"], [2], [1, 0], dtype=dtype)
    assert Tensor.ContractTensors(t1, t2) == Tensor(["a"], [2], [0, 1], dtype=dtype)
    t1 = Tensor(["a"], [2], [0, 1], dtype=dtype)
    t2 = Tensor(["a", "b"], [2, 2], [0, 1, 2, 3], dtype=dtype)
    assert t1.contract_with_tensor(t2) == Tensor(["b"], [2], [2, 3], dtype=dtype)
    t1 = Tensor(["a", "b", "c"], [2, 3, 4], np.full(24, 0.5 + 0.25j), dtype=dtype)
    t2 = Tensor(["b", "c", "d"], [3, 4, 2], np.full(24, 0.5 + 0.25j), dtype=dtype)
    assert t1.contract_with_tensor(t2) == Tensor(["a", "d"], [2, 2], [2.25 + 3j] * 4, dtype=dtype)
    t1 = Tensor(["a", "b"], [2, 2], np.ones(4), dtype=dtype)
    t2 = Tensor(["a", "b"], [2, 2], np.ones(4), dtype=dtype)
    t3 = Tensor.ContractTensors(t1, t2)
    assert t3 == Tensor([], [], [4], dtype=dtype) and t3.is_scalar()


def test_contract_cuda_layout_mapping(ctx):
    """Test_CudaTensor.cpp:561-625: column-major A(a,b,c) . B(c,b,d) equals the
    row-major contraction of the reversed tensors, with C's modes reversed."""
    rng = np.random.default_rng(7)
    a = rand_c(rng, 30, np.complex64)
    b = rand_c(rng, 60, np.complex64)
    say("contract col-major mapping")
    # column-major: modes (a,b,c) extents (2,3,5) and (c,b,d) extents (5,3,4); C = (a,d)
    got, _ = contract(ctx, a, [0, 1, 2], [2, 3, 5], b, [2, 1, 3], [5, 3, 4], [0, 3],
                      layout=nat.COL_MAJOR)
    lc, sc, want = oracle.contract(["c", "b", "a"], [5, 3, 2], a, ["d", "b", "c"], [4, 3, 5], b)
    assert lc == ["a", "d"] and sc == [2, 4]
    # The column-major (a, d) result is the row-major (d, a) array, i.e. the
    # transpose of the host result ("2x4 RowMaj to ColMaj mapping",
    # Test_CudaTensor.cpp:602-624).
    assert rel_err(got, want.reshape(2, 4).T.reshape(-1)) < 1e-6


def _random_problem(rng, max_rank=6, dims=(1, 2, 3, 4)):
    pool = list(string.ascii_letters[:12])
    na, nb = int(rng.integers(0, max_rank + 1)), int(rng.integers(0, max_rank + 1))
    la = list(rng.choice(pool, size=na, replace=False))
    lb = list(rng.choice(pool, size=nb, replace=False))
    d = {k: int(rng.choice(dims)) for k in pool}
    return la, [d[x] for x in la], lb, [d[x] for x in lb]


@pytest.mark.parametrize("prec,dtype", [(nat.PREC_FP32, np.complex64),
                                        (nat.PREC_BF16X3, np.complex64),
                                        (nat.PREC_BF16X3_4M, np.complex64),
                                        (nat.PREC_F16X3_4M, np.complex64),
                                        (nat.PREC_FP64, np.complex128)])
def test_contract_random_small(ctx, prec, dtype):
    rng = np.random.default_rng(100 + prec)
    for t in range(80):
        la, sa, lb, sb = _random_problem(rng)
        a = rand_c(rng, int(np.prod(sa)), dtype)
        b = rand_c(rng, int(np.prod(sb)), dtype)
        lc, sc, want = oracle.contract(la, sa, a, lb, sb, b)
        ids = {s: i for i, s in enumerate(string.ascii_letters)}
        say(f"contract {la}{sa} x {lb}{sb} prec={prec}")
        got, info = contract(ctx, a, [ids[x] for x in la], sa, b, [ids[x] for x in lb], sb,
                             [ids[x] for x in lc], precision=prec, workspace=(t % 2 == 0))
        assert rel_err(got, want) < TOL[prec], (la, sa, lb, sb, info)


@pytest.mark.parametrize("prec,dtype,shape", [
    (nat.PREC_FP32, np.complex64, (256, 256, 256)),
    (nat.PREC_FP32, np.complex64, (300, 70, 170)),
    (nat.PREC_FP32, np.complex64, (33, 17, 16)),
    (nat.PREC_FP32, np.complex64, (50, 300, 40)),       # 64x128 tiles
    (nat.PREC_FP32, np.complex64, (700, 50, 33)),       # 128x64 tiles
    (nat.PREC_BF16X3, np.complex64, (512, 512, 512)),
    (nat.PREC_BF16X3, np.complex64, (300, 200, 170)),
    (nat.PREC_BF16X3_4M, np.complex64, (512, 512, 512)),
    (nat.PREC_BF16X3_4M, np.complex64, (300, 200, 170)),
    (nat.PREC_BF16X3_4M, np.complex64, (1000, 260, 1100)),
    (nat.PREC_F16X3_4M, np.complex64, (512, 512, 512)),
    (nat.PREC_F16X3_4M, np.complex64, (300, 200, 170)),
    (nat.PREC_F16X3_4M, np.complex64, (1000, 260, 1100)),
    (nat.PREC_F16X3_4M, np.complex64, (300, 4100, 200)),   # 128x192 tiles, ragged edges
    (nat.PREC_F16X3_4M, np.complex64, (300, 260, 64)),     # short K on the split kernel
    (nat.PREC_F16X3_4M, np.complex64, (4096, 4100, 32)),
    (nat.PREC_BF16X3_4M, np.complex64, (130, 4500, 160)),
    (nat.PREC_FP64, np.complex128, (256, 128, 96)),
    (nat.PREC_FP64, np.complex128, (130, 70, 50)),
    (nat.PREC_FP32, np.complex64, (8, 4, 4000)),
    (nat.PREC_FP32, np.complex64, (300, 1000, 5)),      # short K, column form
    (nat.PREC_FP32, np.complex64, (20, 70000, 16)),
    (nat.PREC_FP32, np.complex64, (5000, 7, 3)),        # short K, row form
    (nat.PREC_FP64, np.complex128, (70, 3000, 12)),
    (nat.PREC_FP64, np.complex128, (3000, 40, 1)),
    (nat.PREC_FP32, np.complex64, (1, 1, 1 << 20)),
    (nat.PREC_FP32, np.complex64, (3, 5, 100000)),
    (nat.PREC_FP64, np.complex128, (2, 1, 1 << 18)),
    (nat.PREC_FP64, np.complex128, (3, 5, 3000)),
])
def test_gemm_kernels(ctx, builtin_rule, prec, dtype, shape):
    """Matrix-shaped contractions that land on each GEMM kernel under the
    built-in rule (the autotuned choices are covered by test_gpu_tuning.py
    and by every network test)."""
    M, N, K = shape
    rng = np.random.default_rng(M * 7 + N * 3 + K)
    a = rand_c(rng, M * K, dtype)
    b = rand_c(rng, K * N, dtype)
    # A(m,k) . B(k,n): B is stored [k][n] like the reference's Bt.
    say(f"gemm {shape} prec={prec}")
    got, info = contract(ctx, a, [0, 1], [M, K], b, [1, 2], [K, N], [0, 2], precision=prec)
    want = oracle.gemm(a.reshape(M, K), b.reshape(K, N)).reshape(-1)
    err = rel_err(got, want)
    say(f"  kernel={info['kernel']} err={err:.3e}")
    assert err < TOL[prec], (shape, info, err)
    if prec == nat.PREC_BF16X3 and min(shape) >= 128:
        assert info["kernel"] == "bf16x3_mfma"
    if prec == nat.PREC_BF16X3_4M and min(shape) >= 128:
        assert info["kernel"] == "bf16x3_4m_mfma"
    if prec == nat.PREC_F16X3_4M and min(shape[:2]) >= 128 and (
            shape[2] >= 64 or (shape[2] >= 32 and shape[0] * shape[1] >= 1 << 24)):
        assert info["kernel"] == "f16x3_4m_mfma"


def test_gemm_more_than_2_23_columns(ctx, builtin_rule):
    """An FP32 MFMA GEMM whose N needs more than 65535 column tiles (the
    tile index is flattened): a short, very wide product, with a ragged last
    tile; sampled columns against numpy."""
    M, K, N = 32, 16, (1 << 23) + 77
    rng = np.random.default_rng(31)
    a = rand_c(rng, M * K, np.complex64)
    b = rand_c(rng, K * N, np.complex64)
    say(f"gemm wide N={N}")
    got, info = contract(ctx, a, [0, 1], [M, K], b, [1, 2], [K, N], [0, 2])
    assert info["kernel"] == "fp32_mfma", info
    got = got.reshape(M, N)
    cols = np.concatenate([np.arange(64), rng.integers(0, N, 256), np.arange(N - 200, N)])
    want = a.reshape(M, K).astype(np.complex128) @ b.reshape(K, N)[:, cols].astype(np.complex128)
    assert rel_err(got[:, cols], want) < 1e-5


def test_batch_modes(ctx):
    """Modes present in A, B and C are batch modes (used by the slice loop)."""
    rng = np.random.default_rng(12)
    # A(s, m, k), B(k, s, n) -> C(s, m, n)
    S, M, K, N = 4, 40, 24, 20
    a = rand_c(rng, S * M * K, np.complex64)
    b = rand_c(rng, K * S * N, np.complex64)
    say("batch modes")
    got, info = contract(ctx, a, [0, 1, 2], [S, M, K], b, [2, 0, 3], [K, S, N], [0, 1, 3])
    want = np.einsum("smk,ksn->smn", a.reshape(S, M, K).astype(np.complex128),
                     b.reshape(K, S, N).astype(np.complex128)).reshape(-1)
    assert info["batch"] == S
    assert rel_err(got, want) < 1e-5


def test_contract_rejects_bad_modes(ctx):
    with pytest.raises(nat.JetHipError, match="neither contracted nor kept"):
        contract(ctx, np.ones(4, np.complex64), [0, 1], [2, 2], np.ones(2, np.complex64), [1],
                 [2], [])
    with pytest.raises(nat.JetHipError, match="different extents"):
        contract(ctx, np.ones(4, np.complex64), [0, 1], [2, 2], np.ones(3, np.complex64), [1],
                 [3], [0])


# --------------------------------------------------------------------------
# add / slice / fill / reduce
# --------------------------------------------------------------------------

def test_add_tensors_reference_values(ctx):
    say("add")
    # Test_Tensor.cpp AddTensors sections and Test_CudaTensor.cpp:627-678
    lhs = Tensor([], [], [1])
    rhs = Tensor([], [], [2 + 4j])
    assert Tensor.AddTensors(lhs, rhs) == Tensor([], [], [3 + 4j])
    lhs = Tensor(["i"], [3], [1, 2j, 3])
    rhs = Tensor(["i"], [3], [1, 2j, 3j])
    assert lhs.add_tensor(rhs) == Tensor(["i"], [3], [2, 4j, 3 + 3j])
    lhs = Tensor(["i", "j"], [2, 2], [1, 2, 3, 4])
    rhs = Tensor(["j", "i"], [2, 2], [1j, 2j, 3j, 4j])
    # row-major: rhs[j][i]; want[i][j] = lhs[i][j] + rhs[j][i]
    assert Tensor.AddTensors(lhs, rhs) == Tensor(["i", "j"], [2, 2], [1 + 1j, 2 + 3j, 3 + 2j, 4 + 4j])
    zero = Tensor()
    assert Tensor.AddTensors(zero, lhs) == lhs


def test_add_column_major_swapped(ctx):
    """Test_CudaTensor.cpp:665-677 through the column-major ABI."""
    say("add col-major")
    a = np.array([1, 2, 3, 4], np.complex64)
    b = np.array([1j, 2j, 3j, 4j], np.complex64)
    from hiputil import to_device, from_device
    da, db = to_device(ctx, a), to_device(ctx, b)
    dc = nat.DeviceBuffer(32, ctx)
    nat.check(nat.lib().jet_hip_add(ctx.handle, nat.C64, nat.COL_MAJOR, 2, nat.i32([0, 1]),
                                    nat.i64([2, 2]), nat.i32([1, 0]), da.ptr, db.ptr, dc.ptr,
                                    None))
    assert np.array_equal(from_device(dc, 4, np.complex64), [1 + 1j, 2 + 3j, 3 + 2j, 4 + 4j])


def test_slice_index_reference_values(ctx):
    say("slice")
    # Test_Tensor.cpp:579-599
    t = Tensor(["x", "y"], [2, 3], [1, 2, 3, 4, 5, 6])
    assert t.slice_index("x", 0) == Tensor(["y"], [3], [1, 2, 3])
    assert Tensor.SliceIndex(t, "x", 1) == Tensor(["y"], [3], [4, 5, 6])
    assert t.slice_index("y", 0) == Tensor(["x"], [2], [1, 4])
    assert t.slice_index("y", 1) == Tensor(["x"], [2], [2, 5])
    assert t.slice_index("y", 2) == Tensor(["x"], [2], [3, 6])
    rng = np.random.default_rng(3)
    shape = [3, 2, 4, 5]
    x = rand_c(rng, 120, np.complex128)
    t = Tensor(["a", "b", "c", "d"], shape, x, dtype=np.complex128)
    for ax, lab in enumerate("abcd"):
        for v in range(shape[ax]):
            l, s, want = oracle.slice_index("abcd", shape, x, lab, v)
            got = t.slice_index(lab, v)
            assert got.indices == l and got.shape == s
            assert np.array_equal(got.get_data(), want)


def _counter_uniform(seed, counter):
    """numpy restatement of the documented counter RNG (splitmix64)."""
    M = np.uint64(0xFFFFFFFFFFFFFFFF)

    def sm(z):
        z = (z + np.uint64(0x9E3779B97F4A7C15)) & M
        z = ((z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)) & M
        z = ((z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)) & M
        return z ^ (z >> np.uint64(31))

    with np.errstate(over="ignore"):
        x = sm(np.uint64(seed) ^ sm(counter.astype(np.uint64)))
    return (x >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def test_fill_random_is_counter_based(ctx):
    say("fill")
    n = 5000
    from hiputil import from_device
    buf = nat.DeviceBuffer(n * 16, ctx)
    nat.check(nat.lib().jet_hip_fill_random(ctx.handle, nat.C128, n, 2026, 1, buf.ptr, None))
    got = from_device(buf, n, np.complex128)
    i = np.arange(n, dtype=np.uint64)
    want = (2 * _counter_uniform(2026, 2 * i) - 1) + 1j * (2 * _counter_uniform(2026, 2 * i + 1) - 1)
    assert np.array_equal(got, want)
    nat.check(nat.lib().jet_hip_fill_random(ctx.handle, nat.C64, n, 7, 2, buf.ptr, None))
    g = from_device(buf, n, np.complex64)
    assert abs(np.var(g.real) - 0.5) < 0.05 and abs(np.var(g.imag) - 0.5) < 0.05


def test_reduce_compensated(ctx):
    say("reduce")
    rng = np.random.default_rng(1)
    from hiputil import to_device, from_device
    big = rng.standard_normal(3000) * 1e15
    x = np.concatenate([big, rng.standard_normal(3000), -big]) * (1 + 0.5j)
    rng.shuffle(x)
    d = to_device(ctx, x)
    out = nat.DeviceBuffer(16, ctx)
    nat.check(nat.lib().jet_hip_reduce_compensated(ctx.handle, nat.C128, x.size, d.ptr, out.ptr,
                                                   None))
    got = from_device(out, 2, np.float64)
    want = oracle.sum_compensated(x)
    assert abs(got[0] - want.real) < 1e-6 and abs(got[1] - want.imag) < 1e-6


def test_batch_modes_short_k(ctx, builtin_rule):
    """Batched short-K products on the VALU kernels (column and row forms)."""
    rng = np.random.default_rng(14)
    for (S, M, K, N) in [(3, 100, 4, 700), (2, 3000, 6, 9)]:
        a = rand_c(rng, S * M * K, np.complex64)
        b = rand_c(rng, K * S * N, np.complex64)
        say(f"batch modes short K {(S, M, K, N)}")
        got, info = contract(ctx, a, [0, 1, 2], [S, M, K], b, [2, 0, 3], [K, S, N], [0, 1, 3])
        want = np.einsum("smk,ksn->smn", a.reshape(S, M, K).astype(np.complex128),
                         b.reshape(K, S, N).astype(np.complex128)).reshape(-1)
        assert info["batch"] == S and info["kernel"] == "valu", info
        assert rel_err(got, want) < 1e-5, (info, rel_err(got, want))


@pytest.mark.parametrize("prec", [nat.PREC_BF16X3, nat.PREC_BF16X3_4M, nat.PREC_F16X3_4M,
                                  nat.PREC_FP32])
def test_batch_modes_mfma(ctx, builtin_rule, prec):
    """Batched MFMA GEMMs (batch modes in A, B and C) at tile-sized shapes."""
    rng = np.random.default_rng(13)
    S, M, K, N = 3, 160, 256, 144
    a = rand_c(rng, S * M * K, np.complex64)
    b = rand_c(rng, K * S * N, np.complex64)
    say(f"batch modes mfma prec={prec}")
    got, info = contract(ctx, a, [0, 1, 2], [S, M, K], b, [2, 0, 3], [K, S, N], [0, 1, 3],
                         precision=prec)
    want = np.einsum("smk,ksn->smn", a.reshape(S, M, K).astype(np.complex128),
                     b.reshape(K, S, N).astype(np.complex128)).reshape(-1)
    assert info["batch"] == S
    assert rel_err(got, want) < 1e-5, (info, rel_err(got, want))


def test_split_f16_row_scales(ctx, builtin_rule):
    """Split-FP16 scales every operand row by its own power of two: rows whose
    magnitudes differ by up to 10^24 keep FP32-level relative accuracy
    (error measured per output against |A row| * |B row| * sqrt(K)), where an
    unscaled FP16 split would overflow or flush them."""
    rng = np.random.default_rng(21)
    M, N, K = 256, 192, 320
    sa = 10.0 ** rng.uniform(-12, 12, M)
    sb = 10.0 ** rng.uniform(-12, 12, N)
    a = (rand_c(rng, M * K, np.complex128).reshape(M, K) * sa[:, None]).astype(np.complex64)
    b = (rand_c(rng, N * K, np.complex128).reshape(N, K) * sb[:, None]).astype(np.complex64)
    say("split-f16 row scales")
    # A(m,k) . B(n,k) -> C(m,n)
    got, info = contract(ctx, a.reshape(-1), [0, 1], [M, K], b.reshape(-1), [2, 1], [N, K],
                         [0, 2], precision=nat.PREC_F16X3_4M)
    assert info["kernel"] == "f16x3_4m_mfma", info
    want = a.astype(np.complex128) @ b.astype(np.complex128).T
    norm = np.abs(a).max(1)[:, None] * np.abs(b).max(1)[None, :] * np.sqrt(K)
    err = np.abs(got.reshape(M, N) - want) / norm
    assert err.max() < 1e-6, err.max()


@pytest.mark.parametrize("M,N,K", [(333, 150, 128), (200, 131, 130), (129, 260, 256),
                                   (150, 140, 300), (140, 300, 520), (130, 129, 2048),
                                   (131, 140, 2050), (130, 129, 4096), (129, 131, 5000),
                                   (140, 150, 12000), (128, 130, 16384), (130, 128, 16390)])
def test_split_f16_one_pass_groups(ctx, builtin_rule, M, N, K):
    """The one-pass splits: rows of at most 2048 padded values are packed
    several to a wave (16 lanes per row at K = 128, 32 at K <= 256), rows of
    up to 16384 take one work-group each (4 or 8 chunks per thread), both
    with 16-byte loads when K is a multiple of 8; longer rows take the
    two-pass path.  Neighbouring rows of very different magnitudes must each
    get their own scale: error per output against |A row| |B row| sqrt(K),
    below 1e-6 or, where FP32 accumulation over long K dominates, within 4x
    of the FP32-MFMA contraction's own error on the same operands (a CPU
    emulation of this split with sequential FP32 accumulation per 32-wide
    k step gives 5.2e-7 at K = 12000)."""
    rng = np.random.default_rng(M + 3 * N + 7 * K)
    sa = 10.0 ** rng.uniform(-8, 8, M)
    sb = 10.0 ** rng.uniform(-8, 8, N)
    a = (rand_c(rng, M * K, np.complex128).reshape(M, K) * sa[:, None]).astype(np.complex64)
    b = (rand_c(rng, N * K, np.complex128).reshape(N, K) * sb[:, None]).astype(np.complex64)
    say(f"split-f16 one-pass groups M={M} N={N} K={K}")
    got, info = contract(ctx, a.reshape(-1), [0, 1], [M, K], b.reshape(-1), [2, 1], [N, K],
                         [0, 2], precision=nat.PREC_F16X3_4M)
    assert info["kernel"] == "f16x3_4m_mfma", info
    want = a.astype(np.complex128) @ b.astype(np.complex128).T
    norm = np.abs(a).max(1)[:, None] * np.abs(b).max(1)[None, :] * np.sqrt(K)
    err = (np.abs(got.reshape(M, N) - want) / norm).max()
    if err >= 1e-6:
        got32, _ = contract(ctx, a.reshape(-1), [0, 1], [M, K], b.reshape(-1), [2, 1], [N, K],
                            [0, 2], precision=nat.PREC_FP32)
        err32 = (np.abs(got32.reshape(M, N) - want) / norm).max()
        assert err < 4 * err32, (err, err32)


@pytest.mark.parametrize("K", [300, 2050, 5000, 12000, 16384, 16390])
def test_split_f16_exact_integers(ctx, builtin_rule, K):
    """Operands whose split is exact (small integers times a power of two per
    row, so every lo piece is zero and every sum below 2^24): every split
    path (packed rows, one work-group per row, two passes) must give the
    exact product, bit for bit, whatever the accumulation order."""
    rng = np.random.default_rng(K)
    M, N = 133, 140
    pa = np.ldexp(1.0, rng.integers(-20, 21, M))
    pb = np.ldexp(1.0, rng.integers(-20, 21, N))
    ints = lambda n, m: (rng.integers(-8, 9, (n, m)) + 1j * rng.integers(-8, 9, (n, m)))
    a = (ints(M, K) * pa[:, None]).astype(np.complex64)
    b = (ints(N, K) * pb[:, None]).astype(np.complex64)
    say(f"split-f16 exact integers K={K}")
    got, info = contract(ctx, a.reshape(-1), [0, 1], [M, K], b.reshape(-1), [2, 1], [N, K],
                         [0, 2], precision=nat.PREC_F16X3_4M)
    assert info["kernel"] == "f16x3_4m_mfma", info
    want = (a.astype(np.complex128) @ b.astype(np.complex128).T).astype(np.complex64)
    assert np.array_equal(got.reshape(M, N), want)


@pytest.mark.parametrize("kbits", [6, 8, 9, 10])
def test_split_f16_gathered_exact_integers(ctx, builtin_rule, kbits):
    """Exact-split operands (small integers times a power of two per row) in
    a scrambled mode order whose lowest source bits are K bits, so that the
    permutation is folded into the split (gathered tiles): for K = 64, 256
    and 512 the row maxima come from shuffles within a wave, for K = 1024
    from LDS atomics across waves.  The product is exact, bit for bit."""
    rng = np.random.default_rng(300 + kbits)
    common = list(range(kbits))
    left = list(range(20, 28))
    right = list(range(40, 47))
    M, N, K = 2 ** len(left), 2 ** len(right), 2 ** kbits
    ints = lambda n, m: (rng.integers(-8, 9, (n, m)) + 1j * rng.integers(-8, 9, (n, m)))
    at = (ints(M, K) * np.ldexp(1.0, rng.integers(-20, 21, M))[:, None]).astype(np.complex64)
    bt = (ints(N, K) * np.ldexp(1.0, rng.integers(-20, 21, N))[:, None]).astype(np.complex64)
    lp = [int(x) for x in rng.permutation(left)]
    ma = lp[:4] + [common[0]] + lp[4:] + common[1:]     # lowest source bits: K modes
    a = at.reshape([2] * (len(left) + kbits)).transpose(
        [(left + common).index(m) for m in ma]).reshape(-1)
    mb = right + common
    say(f"split-f16 gathered exact integers K={K} A{ma}")
    got, info = contract(ctx, np.ascontiguousarray(a), ma, [2] * len(ma), bt.reshape(-1), mb,
                         [2] * len(mb), left + right, precision=nat.PREC_F16X3_4M)
    assert info["kernel"] == "f16x3_4m_mfma", info
    want = (at.astype(np.complex128) @ bt.astype(np.complex128).T).astype(np.complex64)
    assert np.array_equal(got.reshape(M, N), want)


@pytest.mark.parametrize("seed", [0, 1, 2, 3])
def test_split_f16_gathered_operands(ctx, builtin_rule, seed):
    """Split-FP16 contractions of bond-2 tensors in scrambled mode orders:
    the operand permutations are folded into the split pass (gathered LDS
    tiles) wherever the plan allows it; results against numpy in FP64, and
    against the same contraction with the operands permuted first (equal up
    to the order of the K sum, which the plan takes from one operand)."""
    rng = np.random.default_rng(100 + seed)
    common = list(range(7 + seed % 2))
    left = list(range(20, 28))
    right = list(range(40, 47 + seed % 3))
    ma = [int(x) for x in rng.permutation(common + left)]
    mb = [int(x) for x in rng.permutation(common + right)]
    a = rand_c(rng, 2 ** len(ma), np.complex64)
    b = rand_c(rng, 2 ** len(mb), np.complex64)
    mc = left + right
    say(f"split-f16 gathered operands seed={seed} A{ma} B{mb}")
    got, info = contract(ctx, a, ma, [2] * len(ma), b, mb, [2] * len(mb), mc,
                         precision=nat.PREC_F16X3_4M)
    assert info["kernel"] == "f16x3_4m_mfma", info
    at = a.reshape([2] * len(ma)).transpose([ma.index(m) for m in left + common])
    bt = b.reshape([2] * len(mb)).transpose([mb.index(m) for m in right + common])
    A2 = at.reshape(2 ** len(left), -1).astype(np.complex128)
    B2 = bt.reshape(2 ** len(right), -1).astype(np.complex128)
    want = (A2 @ B2.T).reshape(-1)
    assert rel_err(got, want) < 1e-6
    # the same product from pre-permuted operands (no gather)
    got2, _ = contract(ctx, np.ascontiguousarray(at).reshape(-1), left + common,
                       [2] * len(ma), np.ascontiguousarray(bt).reshape(-1), right + common,
                       [2] * len(mb), mc, precision=nat.PREC_F16X3_4M)
    assert rel_err(got, got2) < 1e-6


@pytest.mark.parametrize("dtype,prec", [(np.complex64, nat.PREC_FP32),
                                        (np.complex64, nat.PREC_F16X3_4M),
                                        (np.complex128, nat.PREC_FP64)])
def test_split_k_scrambled_operands(ctx, builtin_rule, dtype, prec):
    """Few-output, long-K contractions (split-K GEMM, FP64 partial sums) of
    bond-2 tensors in scrambled mode orders, with a batch mode (the last
    step of a batched amplitude network); results against numpy in FP64."""
    rng = np.random.default_rng(7)
    common = list(range(16))
    ma = [int(x) for x in rng.permutation(common + [30, 40])]    # 30: left, 40: batch
    mb = [int(x) for x in rng.permutation(common + [40])]
    a = rand_c(rng, 2 ** len(ma), dtype)
    b = rand_c(rng, 2 ** len(mb), dtype)
    mc = [40, 30]
    say(f"split-k scrambled operands {np.dtype(dtype).name} prec={prec}")
    got, info = contract(ctx, a, ma, [2] * len(ma), b, mb, [2] * len(mb), mc, precision=prec)
    assert info["kernel"] == "splitk", info
    at = a.reshape([2] * len(ma)).transpose([ma.index(m) for m in mc + common])
    bt = b.reshape([2] * len(mb)).transpose([mb.index(m) for m in [40] + common])
    A3 = at.reshape(2, 2, -1).astype(np.complex128)
    B2 = bt.reshape(2, -1).astype(np.complex128)
    want = np.einsum("bmk,bk->bm", A3, B2).reshape(-1)
    tol = 1e-12 if dtype == np.complex128 else 1e-6
    assert rel_err(got, want) < tol


def test_split_f16_unaligned_operands(ctx, builtin_rule):
    """Operands at 8-byte (not 16-byte) aligned addresses, as a caller's
    complex64 sub-array may be: the split passes fall back from 16-byte
    loads, in the gathered split (scrambled B) and in the one-pass split
    (A in order); results equal those from aligned copies."""
    rng = np.random.default_rng(77)
    common = list(range(8))
    left = list(range(20, 28))
    right = list(range(40, 48))
    ma = left + common
    mb = [int(x) for x in rng.permutation(common + right)]
    a = rand_c(rng, 2 ** len(ma), np.complex64)
    b = rand_c(rng, 2 ** len(mb), np.complex64)
    mc = left + right
    say("split-f16 unaligned operands")
    want, info = contract(ctx, a, ma, [2] * len(ma), b, mb, [2] * len(mb), mc,
                          precision=nat.PREC_F16X3_4M)
    assert info["kernel"] == "f16x3_4m_mfma", info
    plan = ctypes.c_void_p()
    nat.check(nat.lib().jet_hip_contraction_plan_create(
        ctx.handle, nat.C64, nat.PREC_F16X3_4M, nat.ROW_MAJOR, len(ma), nat.i32(ma),
        nat.i64([2] * len(ma)), len(mb), nat.i32(mb), nat.i64([2] * len(mb)), len(mc),
        nat.i32(mc), ctypes.byref(plan)))
    try:
        da = nat.DeviceBuffer(a.nbytes + 16, ctx)
        db = nat.DeviceBuffer(b.nbytes + 16, ctx)
        nat.check(nat.lib().jet_hip_memcpy(ctx.handle, da.ptr + 8, a.ctypes.data, a.nbytes, 0,
                                           None, 0))
        nat.check(nat.lib().jet_hip_memcpy(ctx.handle, db.ptr + 8, b.ctypes.data, b.nbytes, 0,
                                           None, 0))
        dc = nat.DeviceBuffer(8 * 2 ** len(mc), ctx)
        nat.check(nat.lib().jet_hip_contract(plan, da.ptr + 8, db.ptr + 8, dc.ptr, None, 0, None))
        got = from_device(dc, 2 ** len(mc), np.complex64)
    finally:
        nat.lib().jet_hip_plan_destroy(plan)
    assert np.array_equal(got, want)


def test_mfma_ceiling_loop(ctx):
    """The FP16 MFMA rate loop reports a positive rate below the nominal
    2.5 PFLOP/s dense FP16 peak and a shader clock in the chip's range, and
    refuses round counts that are not positive multiples of 4."""
    tf, ghz = ctypes.c_double(), ctypes.c_double()
    say("mfma ceiling: 4000 rounds, 1 rep")
    nat.check(nat.lib().jet_hip_mfma_ceiling(ctx.handle, 4000, 1, ctypes.byref(tf),
                                             ctypes.byref(ghz)))
    assert 100.0 < tf.value < 2600.0, tf.value
    assert 0.3 < ghz.value < 3.0, ghz.value
    assert nat.lib().jet_hip_mfma_ceiling(ctx.handle, 6, 1, ctypes.byref(tf),
                                          ctypes.byref(ghz)) != 0
    assert "multiple of 4" in nat.lib().jet_hip_last_error().decode()

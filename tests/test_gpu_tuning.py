"""Every GEMM kernel and tile variant the autotuner can choose, against the
oracle, and the autotuner's own timing call.

For each precision mode, a handful of contractions (permuted operands,
non-power-of-two extents, a batch mode, short K, matrix-vector and long-K
shapes) run once per candidate (kernel, variant) that can run their GEMM
shape: a one-entry tuning table forces the candidate, the plan reports it,
and the result must match the oracle's FP64 contraction (1e-5 normwise for
complex64 (1e-4 for split-BF16), 1e-12 for complex128).  ``jet_hip_autotune_gemm`` must time every
valid candidate of a shape.
"""
import ctypes

import numpy as np
import pytest

import oracle
from jet_amd import _native as nat

from hiputil import contract, rand_c, rel_err, say

pytestmark = pytest.mark.gpu

MODES = {"fp32": (np.complex64, nat.PREC_FP32, 0, 1e-5),
         "f16x3": (np.complex64, nat.PREC_F16X3_4M, 4, 1e-5),
         "bf16x3_4m": (np.complex64, nat.PREC_BF16X3_4M, 3, 1e-4),
         "fp64": (np.complex128, nat.PREC_FP64, 2, 1e-12)}

# (A modes, A extents, B modes, B extents, C modes): GEMM M x N x K (batch)
CASES = [
    ([0, 1, 2], [24, 20, 40], [2, 3, 1], [40, 36, 20], [0, 3]),          # 24 x 36 x 800
    ([0, 1, 2, 3], [4, 50, 3, 7], [3, 4, 2, 0], [7, 33, 3, 4], [0, 1, 4]),  # batch 4, 50 x 33 x 21
    ([0, 1], [130, 12], [1, 2], [12, 200], [0, 2]),                      # K = 12 (short K)
    ([0, 1], [1, 3000], [1, 2], [3000, 70], [0, 2]),                     # M = 1 (gemv-like)
    ([0, 1], [5, 40000], [1, 2], [40000, 3], [0, 2]),                    # long K, 15 outputs
]


_cand_cache = {}


def _candidates(mode_name, M, N, K, batch):
    key = (mode_name, M, N, K, batch)
    if key not in _cand_cache:
        ctx = nat.default_context()
        dt, pc, _, _ = MODES[mode_name]
        cap = 32
        ks, vs = (ctypes.c_int32 * cap)(), (ctypes.c_int32 * cap)()
        ms, cnt = (ctypes.c_float * cap)(), ctypes.c_int()
        code = nat.C64 if dt == np.complex64 else nat.C128
        nat.check(nat.lib().jet_hip_autotune_gemm(ctx.handle, code, pc, M, N, K, batch, 1, cap,
                                                  ks, vs, ms, ctypes.byref(cnt)))
        assert all(ms[i] > 0 for i in range(cnt.value))
        _cand_cache[key] = {(int(ks[i]), int(vs[i])) for i in range(cnt.value)}
    return _cand_cache[key]


def _shape(ma, sa, mb, sb, mc):
    ext = dict(zip(ma, sa))
    ext.update(zip(mb, sb))
    batch = [m for m in mc if m in ma and m in mb]
    left = [m for m in ma if m not in mb]
    right = [m for m in mb if m not in ma]
    common = [m for m in ma if m in mb and m not in mc]
    p = lambda ms: int(np.prod([ext[m] for m in ms])) if ms else 1
    return p(left), p(right), p(common), p(batch)


@pytest.fixture
def restore_table():
    yield
    nat.check(nat.lib().jet_hip_tuning_set(-1, None))


@pytest.mark.parametrize("mode", sorted(MODES))
def test_every_candidate_matches_oracle(ctx, mode, restore_table):
    dtype, prec, mode_id, tol = MODES[mode]
    rng = np.random.default_rng(11)
    for ci, (ma, sa, mb, sb, mc) in enumerate(CASES):
        M, N, K, B = _shape(ma, sa, mb, sb, mc)
        a = rand_c(rng, int(np.prod(sa)), dtype)
        b = rand_c(rng, int(np.prod(sb)), dtype)
        la, lb = [f"i{m}" for m in ma], [f"i{m}" for m in mb]
        lc_ref, sc_ref, ref = (oracle.contract(la, sa, a, lb, sb, b) if B == 1
                               else (None, None, None))
        # the oracle's output order is left ++ right; reorder to mc
        perm = [lc_ref.index(f"i{m}") for m in mc] if B == 1 else None
        cands = _candidates(mode, M, N, K, B)
        assert cands, (mode, ci)
        for (kern, var) in sorted(cands):
            lm, ln, lk, lbt = (int(round(np.log2(x))) for x in (M, N, K, B))
            row = (ctypes.c_int32 * 7)(mode_id, lm, ln, lk, lbt, kern, var)
            nat.check(nat.lib().jet_hip_tuning_set(1, row))
            say(f"tuning candidate {mode} case {ci} ({M}x{N}x{K}, batch {B}): "
                f"kernel {kern} variant {var}")
            got, info = contract(ctx, a, ma, sa, b, mb, sb, mc, precision=prec)
            assert info["kernel"] == nat.GEMM_KERNEL_NAMES[kern], (info, kern)
            if B == 1:
                want = ref.reshape(sc_ref).transpose(perm).reshape(-1)
                err = rel_err(got, want)
                assert err < tol, (mode, ci, kern, var, err)
            else:
                _check_batched(got, a, sa, b, sb, ma, mb, mc, tol, (mode, ci, kern, var))


def _check_batched(got, a, sa, b, sb, ma, mb, mc, tol, tag):
    """The batched case against numpy einsum in complex128 (the oracle has
    no batch modes; einsum is the textbook definition)."""
    letters = "abcdefghij"
    sub = lambda ms: "".join(letters[m] for m in ms)
    want = np.einsum(f"{sub(ma)},{sub(mb)}->{sub(mc)}",
                     a.astype(np.complex128).reshape(sa), b.astype(np.complex128).reshape(sb),
                     optimize=True)
    err = rel_err(got, want.reshape(-1))
    assert err < tol, (tag, err)


@pytest.mark.parametrize("mode", ["f16x3", "bf16x3_4m"])
def test_persistent_split_walks_many_tiles(ctx, mode, restore_table):
    """The persistent split kernel (variant 3) with more output tiles than
    work-groups, so each work-group carries its next tile's first k step
    across its epilogue: ragged M and N edges, and a batch mode, against
    numpy in complex128."""
    dtype, prec, mode_id, tol = MODES[mode]
    rng = np.random.default_rng(17)
    # (A modes, extents, B modes, extents, C modes): 200 x 40000 x 96 (626
    # tiles), and batch 8 of 256 x 4096 x 64 (512 tiles)
    cases = [([0, 1], [200, 96], [2, 1], [40000, 96], [0, 2]),
             ([3, 0, 1], [8, 256, 64], [3, 2, 1], [8, 4096, 64], [3, 0, 2])]
    for ma, sa, mb, sb, mc in cases:
        M, N, K, B = _shape(ma, sa, mb, sb, mc)
        lm, ln, lk, lbt = (int(round(np.log2(x))) for x in (M, N, K, B))
        row = (ctypes.c_int32 * 7)(mode_id, lm, ln, lk, lbt, 7 if mode == "f16x3" else 5, 3)
        nat.check(nat.lib().jet_hip_tuning_set(1, row))
        a = rand_c(rng, int(np.prod(sa)), dtype)
        b = rand_c(rng, int(np.prod(sb)), dtype)
        say(f"persistent split {mode}: {M}x{N}x{K} batch {B}")
        got, info = contract(ctx, a, ma, sa, b, mb, sb, mc, precision=prec)
        assert info["kernel"] == ("f16x3_4m_mfma" if mode == "f16x3" else "bf16x3_4m_mfma")
        _check_batched(got, a, sa, b, sb, ma, mb, mc, tol, (mode, M, N, K, B))


@pytest.mark.parametrize("mode", ["f16x3", "bf16x3_4m"])
def test_tall_split_tiles(ctx, mode, restore_table):
    """The 192x128 split tile (variant 4, long side along M) over many tiles
    with ragged M and N edges, and with a batch mode, against numpy in
    complex128."""
    dtype, prec, mode_id, tol = MODES[mode]
    rng = np.random.default_rng(23)
    # 5000 x 300 x 96 (27 x 3 tiles, both edges partial) and batch 4 of
    # 1000 x 130 x 64
    cases = [([0, 1], [5000, 96], [2, 1], [300, 96], [0, 2]),
             ([3, 0, 1], [4, 1000, 64], [3, 2, 1], [4, 130, 64], [3, 0, 2])]
    for ma, sa, mb, sb, mc in cases:
        M, N, K, B = _shape(ma, sa, mb, sb, mc)
        lm, ln, lk, lbt = (int(round(np.log2(x))) for x in (M, N, K, B))
        row = (ctypes.c_int32 * 7)(mode_id, lm, ln, lk, lbt, 7 if mode == "f16x3" else 5, 4)
        nat.check(nat.lib().jet_hip_tuning_set(1, row))
        a = rand_c(rng, int(np.prod(sa)), dtype)
        b = rand_c(rng, int(np.prod(sb)), dtype)
        say(f"tall split tile {mode}: {M}x{N}x{K} batch {B}")
        got, info = contract(ctx, a, ma, sa, b, mb, sb, mc, precision=prec)
        assert info["kernel"] == ("f16x3_4m_mfma" if mode == "f16x3" else "bf16x3_4m_mfma")
        _check_batched(got, a, sa, b, sb, ma, mb, mc, tol, (mode, M, N, K, B))

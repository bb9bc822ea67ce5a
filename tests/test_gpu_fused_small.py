"""The fused small-contraction pass on the GPU: random small contractions of
power-of-two extents with scrambled operand and result mode orders and batch
modes (the planner runs them as one fused VALU pass, no permuted copies and
no workspace), in complex64 (FP32) and complex128 (FP64), against numpy's
einsum in complex128: 1e-5 normwise for complex64, 1e-12 for complex128."""
import numpy as np
import pytest

from jet_amd import _native as nat

from hiputil import contract, rand_c, rel_err, say
from test_circuits_paths import _check_lib
from test_fused_small import LETTERS, _is_fused, _random_case

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dtype,prec,tol", [(np.complex64, nat.PREC_FP32, 1e-5),
                                            (np.complex128, nat.PREC_FP64, 1e-12)])
def test_fused_small_contractions(ctx, dtype, prec, tol):
    L = _check_lib()
    rng = np.random.default_rng(31)
    fused = 0
    for case in range(24):
        ma, mb, mc, ext = _random_case(rng)
        sa, sb = [ext[m] for m in ma], [ext[m] for m in mb]
        if dtype == np.complex64:
            fused += _is_fused(L, ma, sa, mb, sb, mc)[0]
        a = rand_c(rng, int(np.prod(sa)), dtype)
        b = rand_c(rng, int(np.prod(sb)), dtype)
        say(f"fused small {np.dtype(dtype).name} case {case}: {ma} x {mb} -> {mc}")
        got, info = contract(ctx, a, ma, sa, b, mb, sb, mc, precision=prec, workspace=True)
        sub = lambda ms: "".join(LETTERS[m] for m in ms)
        want = np.einsum(f"{sub(ma)},{sub(mb)}->{sub(mc)}",
                         a.astype(np.complex128).reshape(sa),
                         b.astype(np.complex128).reshape(sb)).reshape(-1)
        assert rel_err(got, want) < tol, (case, info, rel_err(got, want))
    if dtype == np.complex64:
        assert fused >= 6, fused   # a good share of these small steps is fused


@pytest.mark.parametrize("dtype,prec,tol", [(np.complex64, nat.PREC_FP32, 1e-5),
                                            (np.complex128, nat.PREC_FP64, 1e-12)])
def test_fused_small_reductions(ctx, dtype, prec, tol):
    """Few outputs and long K (2^9-2^12, the one-work-group-per-output form),
    scrambled orders, against einsum."""
    L = _check_lib()
    rng = np.random.default_rng(37)
    fused = 0
    for case in range(6):
        nk = int(rng.integers(9, 13))
        nl, nr, nb = int(rng.integers(1, 3)), int(rng.integers(0, 3)), int(rng.integers(0, 2))
        ids = [int(x) for x in rng.permutation(nk + nl + nr + nb)]
        com, lef = ids[:nk], ids[nk:nk + nl]
        rig, bat = ids[nk + nl:nk + nl + nr], ids[nk + nl + nr:]
        ma = [int(x) for x in rng.permutation(bat + lef + com)]
        mb = [int(x) for x in rng.permutation(bat + rig + com)]
        mc = [int(x) for x in rng.permutation(bat + lef + rig)]
        sa, sb = [2] * len(ma), [2] * len(mb)
        if dtype == np.complex64:
            fused += _is_fused(L, ma, sa, mb, sb, mc)[0]
        a = rand_c(rng, 2 ** len(ma), dtype)
        b = rand_c(rng, 2 ** len(mb), dtype)
        say(f"fused reduction {np.dtype(dtype).name} case {case}: K = 2^{nk}")
        got, info = contract(ctx, a, ma, sa, b, mb, sb, mc, precision=prec, workspace=True)
        sub = lambda ms: "".join(LETTERS[m] for m in ms)
        want = np.einsum(f"{sub(ma)},{sub(mb)}->{sub(mc)}",
                         a.astype(np.complex128).reshape(sa),
                         b.astype(np.complex128).reshape(sb)).reshape(-1)
        assert rel_err(got, want) < tol, (case, info, rel_err(got, want))
    if dtype == np.complex64:
        assert fused >= 3, fused


@pytest.mark.parametrize("dtype,prec,mode,tol", [(np.complex64, nat.PREC_FP32, 0, 1e-5),
                                                 (np.complex128, nat.PREC_FP64, 2, 1e-12)])
def test_fused_small_reductions_many_outputs(ctx, dtype, prec, mode, tol):
    """The one-work-group-per-output form with more outputs (2^12-2^14)
    than work-groups, so that each work-group walks several outputs with
    one set of k offset tables: scrambled orders, the reduction kernel
    forced by a one-entry tuning table, against einsum."""
    import ctypes
    rng = np.random.default_rng(41)
    try:
        for case in range(4):
            nk = int(rng.integers(9, 11))
            nl, nr, nb = int(rng.integers(4, 7)), int(rng.integers(4, 7)), int(rng.integers(0, 3))
            ids = [int(x) for x in rng.permutation(nk + nl + nr + nb)]
            com, lef = ids[:nk], ids[nk:nk + nl]
            rig, bat = ids[nk + nl:nk + nl + nr], ids[nk + nl + nr:]
            ma = [int(x) for x in rng.permutation(bat + lef + com)]
            mb = [int(x) for x in rng.permutation(bat + rig + com)]
            mc = [int(x) for x in rng.permutation(bat + lef + rig)]
            sa, sb = [2] * len(ma), [2] * len(mb)
            row = (ctypes.c_int32 * 7)(mode, nl, nr, nk, nb, 1, 0)
            nat.check(nat.lib().jet_hip_tuning_set(1, row))
            a = rand_c(rng, 2 ** len(ma), dtype)
            b = rand_c(rng, 2 ** len(mb), dtype)
            say(f"fused reduction, many outputs {np.dtype(dtype).name} case {case}: "
                f"2^{nl + nr + nb} outputs, K = 2^{nk}")
            got, info = contract(ctx, a, ma, sa, b, mb, sb, mc, precision=prec, workspace=True)
            assert info["kernel"] == "valu_reduce", info
            sub = lambda ms: "".join(LETTERS[m] for m in ms)
            want = np.einsum(f"{sub(ma)},{sub(mb)}->{sub(mc)}",
                             a.astype(np.complex128).reshape(sa),
                             b.astype(np.complex128).reshape(sb)).reshape(-1)
            assert rel_err(got, want) < tol, (case, info, rel_err(got, want))
    finally:
        nat.check(nat.lib().jet_hip_tuning_set(-1, None))

"""K split into batched chunks on the GPU: contractions with many outputs
and a long K whose one-entry tuning table names the VALU split-K kernel run
as chunked MFMA GEMMs plus a chunk sum; results against numpy in complex128
(1e-5 normwise for complex64, 1e-12 for complex128), with scrambled mode
orders and a batch mode."""
import ctypes

import numpy as np
import pytest

from jet_amd import _native as nat

from hiputil import contract, rand_c, rel_err, say

pytestmark = pytest.mark.gpu

LETTERS = "abcdefghijklmnopqrstuvwxyzABCDEFGHIJKLMNOPQRSTUVWXYZ"


@pytest.mark.parametrize("dtype,prec,mode,kern,tol", [
    (np.complex64, nat.PREC_F16X3_4M, 4, "f16x3_4m_mfma", 1e-5),
    (np.complex64, nat.PREC_FP32, 0, "fp32_mfma", 1e-5),
    (np.complex128, nat.PREC_FP64, 2, "fp64_mfma", 1e-12)])
def test_chunked_long_k(ctx, dtype, prec, mode, kern, tol):
    rng = np.random.default_rng(61)
    try:
        for case, (nl, nr, nk, nb) in enumerate([(7, 7, 16, 0), (8, 7, 17, 0), (7, 7, 15, 1)]):
            ids = [int(x) for x in rng.permutation(nl + nr + nk + nb)]
            lef, rig = ids[:nl], ids[nl:nl + nr]
            com, bat = ids[nl + nr:nl + nr + nk], ids[nl + nr + nk:]
            ma = bat + lef + [int(x) for x in rng.permutation(com)]
            mb = [int(x) for x in rng.permutation(bat + rig + com)]
            mc = bat + lef + rig
            row = (ctypes.c_int32 * 7)(mode, nl, nr, nk, nb, 6, 0)   # split-K named
            nat.check(nat.lib().jet_hip_tuning_set(1, row))
            a = rand_c(rng, 2 ** len(ma), dtype)
            b = rand_c(rng, 2 ** len(mb), dtype)
            say(f"chunked K {np.dtype(dtype).name} case {case}: M=2^{nl} N=2^{nr} K=2^{nk} "
                f"batch=2^{nb}")
            got, info = contract(ctx, a, ma, [2] * len(ma), b, mb, [2] * len(mb), mc,
                                 precision=prec, workspace=True)
            assert info["kernel"] == kern, info
            assert info["K"] == 2 ** nk and info["batch"] == 2 ** nb, info
            sub = lambda ms: "".join(LETTERS[m] for m in ms)
            want = np.einsum(f"{sub(ma)},{sub(mb)}->{sub(mc)}",
                             a.astype(np.complex128).reshape([2] * len(ma)),
                             b.astype(np.complex128).reshape([2] * len(mb)),
                             optimize=True).reshape(-1)
            assert rel_err(got, want) < tol, (case, rel_err(got, want))
    finally:
        nat.check(nat.lib().jet_hip_tuning_set(-1, None))

"""K split into batched chunks (planner, host only).

A contraction with many outputs and a long K whose tuned kernel is a VALU
one (the split-K and one-block-per-output forms, built for few outputs) is
planned as a batched MFMA GEMM over chunks of K -- the leading common modes
become leading batch modes of both operands -- followed by a sum over the
chunks.  Checked on the plan replayed on the host
(tests/native/network_check.cpp:contraction_gemm_plan):

* 128 x 128 outputs, K = 2^19, table entry split-K: the FP16-split, FP32
  and FP64 modes chunk into 256 batched GEMMs of K = 2048 on their MFMA
  kernel, with a chunk buffer of chunks * outputs elements inside the
  workspace; the 64 x 256 x 2^20 step of the 6x6 depth-12 amplitude too;
* few outputs (the gemv-like root steps), short K, or an MFMA table choice:
  no chunks;
* the chunked contraction is the unchunked one: numpy einsum over the
  chunk-major reshaping sums to the same tensor (the identity the executor
  relies on).
"""
import ctypes

import numpy as np
import pytest

from test_circuits_paths import _check_lib

SMALL, SMALLRED, F32, F64, SPLITK, F16 = 0, 1, 2, 3, 6, 7


def plan(mode, left, right, common, table=None, mc=None):
    """Bond-2 modes: A = left ++ common, B = common ++ right (scrambled K
    order in B), C = left ++ right."""
    L = _check_lib()
    ma = left + common
    mb = common[::-1] + right
    mc = left + right if mc is None else mc
    ea, eb = [2] * len(ma), [2] * len(mb)
    i32 = lambda v: (ctypes.c_int32 * max(1, len(v)))(*v)
    i64 = lambda v: (ctypes.c_int64 * max(1, len(v)))(*v)
    n = -1 if table is None else len(table)
    flat = [x for row in (table or []) for x in row]
    out = (ctypes.c_int64 * 7)()
    assert L.contraction_gemm_plan(mode, n, i32(flat), len(ma), i32(ma), i64(ea), len(mb),
                                   i32(mb), i64(eb), len(mc), i32(mc), out) == 0
    return dict(kernel=out[0], chunks=out[1], batch=out[2], K=out[3], workspace=out[4],
                off_chunks=out[5], size_c=out[6])


LEFT = list(range(100, 107))     # M = 128
RIGHT = list(range(200, 207))    # N = 128
COMMON = list(range(19))         # K = 2^19


@pytest.mark.parametrize("mode,kernel,elem", [(4, F16, 8), (0, F32, 8), (2, F64, 16)])
def test_many_outputs_long_k_is_chunked(mode, kernel, elem):
    table = [[mode, 7, 7, 19, 0, SPLITK, 0]]
    p = plan(mode, LEFT, RIGHT, COMMON, table)
    assert p["kernel"] == kernel, p
    assert p["chunks"] == 256 and p["K"] == 2048 and p["batch"] == 256, p
    assert p["size_c"] == 128 * 128
    # the chunk buffer (chunks * outputs elements) lies inside the workspace
    assert p["off_chunks"] % 256 == 0
    assert p["off_chunks"] + elem * p["chunks"] * p["size_c"] <= p["workspace"]


def test_uneven_batch_root_shape_is_chunked():
    """The 6x6 depth-12 amplitude's 64 x 256 x 2^20 step (its table entry
    names split-K): chunked onto an MFMA kernel."""
    p = plan(4, list(range(100, 106)), list(range(200, 208)), list(range(20)),
             [[4, 6, 8, 20, 0, SPLITK, 1]])
    assert p["kernel"] in (F16, F32), p
    assert p["chunks"] >= 64 and p["K"] >= 2048 and p["K"] * p["chunks"] == 1 << 20, p


def test_one_block_per_output_choice_is_chunked_too():
    p = plan(4, LEFT, RIGHT, COMMON, [[4, 7, 7, 19, 0, SMALLRED, 0]])
    assert p["kernel"] == F16 and p["chunks"] == 256, p


@pytest.mark.parametrize("left,right,common,entry", [
    ([100], [], list(range(19)), [4, 1, 0, 19, 0, SPLITK, 0]),           # gemv-like root
    (LEFT, RIGHT, list(range(12)), [4, 7, 7, 12, 0, SMALLRED, 0]),       # K < 2^15
    (LEFT, RIGHT, COMMON, [4, 7, 7, 19, 0, F16, 1]),                     # MFMA chosen
])
def test_not_chunked(left, right, common, entry):
    p = plan(4, left, right, common, [entry])
    assert p["chunks"] == 1, p
    assert p["off_chunks"] == 0


def test_chunk_sum_identity():
    """C = sum over the leading chunk index of the batched products: the
    chunk modes are the leading common modes of the chosen K order."""
    rng = np.random.default_rng(3)
    M, N, K, chunks = 8, 6, 64, 4
    a = rng.standard_normal((M, K)) + 1j * rng.standard_normal((M, K))
    b = rng.standard_normal((N, K)) + 1j * rng.standard_normal((N, K))
    want = a @ b.T
    ac = a.reshape(M, chunks, K // chunks).transpose(1, 0, 2)      # [chunk][M][K']
    bc = b.reshape(N, chunks, K // chunks).transpose(1, 0, 2)
    part = np.einsum("cmk,cnk->cmn", ac, bc)
    assert np.allclose(part.sum(axis=0), want, rtol=1e-12, atol=1e-12)

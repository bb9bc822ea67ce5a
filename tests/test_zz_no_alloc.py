"""jet_hip_contract_noalloc against jet_hip_contract, through the C ABI.

The no-allocation entry point runs a contraction plan's kernels with the
library's fixed per-device scratch array as the workspace.  Each case
contracts the same device operands through both entry points and requires
bit-identical results (same kernels, same workspace layout), and checks both
against the oracle's FP64 contraction.  Two calls in a row, the second on a
smaller input, so the second runs on the scratch contents the first left.
A plan whose workspace exceeds the scratch size is refused before anything
is queued.
"""
import ctypes

import numpy as np
import pytest

import oracle
from jet_amd import _native as nat

from hiputil import code, from_device, rand_c, rel_err, say, to_device

pytestmark = pytest.mark.gpu

SCRATCH = 64 * 1024 * 1024


def _plan(ctx, dtype, prec, ma, sa, mb, sb, mc):
    plan = ctypes.c_void_p()
    nat.check(nat.lib().jet_hip_contraction_plan_create(
        ctx.handle, code(dtype), prec, nat.ROW_MAJOR, len(ma), nat.i32(ma), nat.i64(sa),
        len(mb), nat.i32(mb), nat.i64(sb), len(mc), nat.i32(mc), ctypes.byref(plan)))
    return plan


def _workspace(plan):
    sz = ctypes.c_size_t()
    nat.check(nat.lib().jet_hip_plan_workspace_size(plan, ctypes.byref(sz)))
    return sz.value


# (A modes, A extents, B modes, B extents): modes are ints, the C modes are
# ContractTensors' left ++ right.  The first case of each pair is the larger.
# Every case permutes an operand through the workspace: the small ones carry
# an extent of 12, so the planner does not fuse them into one in-place pass.
CASES = {
    "fp32": [([0, 1, 2, 3], [64, 32, 64, 16], [2, 4, 1], [64, 128, 32]),
             ([0, 1, 2, 3], [16, 8, 32, 12], [2, 4, 1], [32, 128, 8])],
    "f16x3": [([0, 1, 2, 3], [64, 32, 64, 16], [2, 4, 1], [64, 128, 32]),
              ([5, 0, 1, 2], [8, 32, 16, 16], [1, 2, 7], [16, 16, 256])],
    "fp64": [([0, 1, 2], [128, 64, 32], [2, 3, 1], [32, 96, 64]),
             ([0, 1, 2], [12, 64, 8], [2, 3, 1], [8, 32, 64])],
}
PREC = {"fp32": nat.PREC_FP32, "f16x3": nat.PREC_F16X3_4M, "fp64": nat.PREC_FP64}
TOL = {"fp32": 1e-5, "f16x3": 1e-5, "fp64": 1e-12}


def _out_modes(ma, mb):
    return [m for m in ma if m not in mb] + [m for m in mb if m not in ma]


@pytest.mark.parametrize("mode", sorted(CASES))
def test_noalloc_matches_contract_twice(ctx, mode):
    dtype = np.complex128 if mode == "fp64" else np.complex64
    rng = np.random.default_rng(7 + len(mode))
    for call, (ma, sa, mb, sb) in enumerate(CASES[mode]):
        mc = _out_modes(ma, mb)
        ext = dict(zip(ma, sa))
        ext.update(zip(mb, sb))
        nc = int(np.prod([ext[m] for m in mc]))
        a = rand_c(rng, int(np.prod(sa)), dtype)
        b = rand_c(rng, int(np.prod(sb)), dtype)
        plan = _plan(ctx, dtype, PREC[mode], ma, sa, mb, sb, mc)
        try:
            ws = _workspace(plan)
            say(f"noalloc {mode} call {call}: A{sa} B{sb} -> {nc} outputs, workspace {ws} B")
            assert 0 < ws <= SCRATCH
            da, db = to_device(ctx, a), to_device(ctx, b)
            dc_ref = nat.DeviceBuffer(nc * a.itemsize, ctx)
            dc = nat.DeviceBuffer(nc * a.itemsize, ctx)
            nat.check(nat.lib().jet_hip_contract(plan, da.ptr, db.ptr, dc_ref.ptr, None, 0, None))
            nat.check(nat.lib().jet_hip_contract_noalloc(plan, da.ptr, db.ptr, dc.ptr, None))
            ctx.synchronize()
            want = from_device(dc_ref, nc, dtype)
            got = from_device(dc, nc, dtype)
        finally:
            nat.lib().jet_hip_plan_destroy(plan)
        assert np.array_equal(got.view(np.float32 if dtype == np.complex64 else np.float64),
                              want.view(np.float32 if dtype == np.complex64 else np.float64))
        la = [f"m{m}" for m in ma]
        lb = [f"m{m}" for m in mb]
        lc, _, ref = oracle.contract(la, sa, a, lb, sb, b)
        assert lc == [f"m{m}" for m in mc]
        assert rel_err(got, ref) < TOL[mode]


def test_noalloc_refuses_workspace_beyond_scratch(ctx):
    # A[0..3] x B: permuting a 2^24-element complex64 operand needs 128 MiB.
    ma, sa = [0, 1, 2], [256, 256, 256]
    mb, sb = [0, 3], [256, 4]
    mc = _out_modes(ma, mb)
    plan = _plan(ctx, np.complex64, nat.PREC_FP32, ma, sa, mb, sb, mc)
    try:
        ws = _workspace(plan)
        say(f"noalloc refusal: workspace {ws} B")
        assert ws > SCRATCH
        rc = nat.lib().jet_hip_contract_noalloc(plan, ctypes.c_void_p(16), ctypes.c_void_p(16),
                                                ctypes.c_void_p(16), None)
        assert rc == 1
        msg = nat.lib().jet_hip_last_error().decode()
        assert "workspace" in msg and str(SCRATCH) in msg
    finally:
        nat.lib().jet_hip_plan_destroy(plan)

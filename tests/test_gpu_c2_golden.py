"""C2 slices on the GPU against the CPU oracle.

tests/golden/c2_slices_fp64.json holds complex128 values of 12 slices of the
sliced 7x7 depth-14 amplitude network (slices 0-7, 300, 1337, 2600, 4095),
contracted on the CPU by the oracle (tests/golden/make_c2_slices.py; slice
order UnravelIndex over the 12 sliced indices, TensorNetwork.hpp:230).  Both
committed plans run those slices through SlicedContraction (4 sliced indices
batched, the other 8 looped, so slices past 15 exercise the loop indices) in
split-FP16, FP32 MFMA and complex128 on FP64 MFMA.  Tolerances: 1e-5
normwise over the 12 slices for complex64, 1e-10 for complex128.  The whole
4096-slice amplitude in split-FP16 is checked against the complex128
amplitude of tests/golden/c2_amplitude_fp64.json (1e-5 relative).
"""
import json
import os

import numpy as np
import pytest

from jet_amd import circuits
from jet_amd.network import SlicedContraction

from hiputil import say

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PLANS = ["grid_7x7x14_seed2026_s12_refined.json", "grid_7x7x14_seed2026_s12.json"]


@pytest.fixture(scope="module")
def c2():
    with open(os.path.join(ROOT, "tests", "golden", "c2_slices_fp64.json")) as f:
        fx = json.load(f)
    circ = circuits.random_grid_circuit(7, 7, 14, seed=2026)
    net = circuits.amplitude_network(circ, dtype=np.complex64)
    want = np.array([complex(*v) for v in fx["values"]])
    return fx, net, want


@pytest.mark.parametrize("plan_name", PLANS)
@pytest.mark.parametrize("prec,dtype,tol", [("f16x3", np.complex64, 1e-5),
                                            ("fp32", np.complex64, 1e-5),
                                            ("fp64", np.complex128, 1e-10)])
def test_c2_slices_match_cpu_oracle(ctx, c2, plan_name, prec, dtype, tol):
    fx, net, want = c2
    with open(os.path.join(ROOT, "jet_amd", "data", plan_name)) as f:
        plan = json.load(f)
    sliced = plan["sliced"]
    assert sliced == fx["sliced"]
    data = [d.astype(dtype) for d in net.data]
    sc = SlicedContraction(net.indices, net.shapes, data, [tuple(p) for p in plan["path"]],
                           sliced, batch=sliced[-4:], dtype=dtype, precision=prec, ctx=ctx,
                           use_graph=False)
    execs = sorted({s // sc.n_batch for s in fx["slices"]})
    say(f"c2 golden {plan_name} {prec}: executions {execs}")
    sc.run(execs)
    got = sc.slice_results()[np.array(fx["slices"])].astype(np.complex128)
    sc.net.close()
    err = float(np.linalg.norm(got - want) / np.linalg.norm(want))
    say(f"  normwise error {err:.2e}")
    assert err < tol, err


def test_c2_whole_amplitude_split_fp16(ctx, c2):
    """All 4096 slices along the headline plan (256 per execution, the bench's
    setting) in split-FP16, summed in FP64 by the compensated reduction,
    against tests/golden/c2_amplitude_fp64.json (complex128 on FP64 MFMA over
    every slice, accepted only after its slices matched the CPU oracle's to
    1e-10; tests/golden/make_c2_amplitude.py): relative error <= 1e-5."""
    _, net, _ = c2
    with open(os.path.join(ROOT, "tests", "golden", "c2_amplitude_fp64.json")) as f:
        ref = json.load(f)
    with open(os.path.join(ROOT, ref["plan"])) as f:
        plan = json.load(f)
    assert ref["slices"] == 4096 and ref["checked_slices_rel_err"] < 1e-10
    sliced = plan["sliced"]
    sc = SlicedContraction(net.indices, net.shapes, list(net.data),
                           [tuple(p) for p in plan["path"]], sliced, batch=sliced[-8:],
                           dtype=np.complex64, precision="f16x3", ctx=ctx, use_graph=False)
    say(f"c2 whole amplitude split-FP16: {sc.n_loop} executions of {sc.n_batch} slices")
    sc.run()
    amp = sc.reduce()
    sc.net.close()
    want = complex(*ref["amplitude"])
    err = abs(amp - want) / abs(want)
    say(f"  amplitude {amp:.6e}, relative error {err:.2e}")
    assert err <= 1e-5, err

"""Parity of the device contraction sequences and the batched slice loop.

Oracles: the reference's literal network result (Test_TensorNetwork.cpp:
529-579), the statevector amplitude of seeded grid circuits (closed form,
complex128), and per-slice FP64 contractions of the sliced leaves on the CPU
(oracle.contract_path).  Tolerances: 1e-5 relative for complex64 (FP32 and
split-BF16), 1e-12 for complex128 (FP64).
"""
import numpy as np
import pytest

import oracle
from jet_amd import _native as nat
from jet_amd import circuits, paths
from jet_amd.network import DeviceNetwork, SlicedContraction, _slice_leaf

from hiputil import say, to_device, from_device

pytestmark = pytest.mark.gpu


def _make_tensor_data(shape):
    n = int(np.prod(shape))
    i = np.arange(n)
    return (i + 2j * i).astype(np.complex64)


def test_network_reference_literal(ctx):
    say("network literal")
    idx = [["A0", "B1"], ["C2", "B1"], ["C2", "D3"]]
    shp = [[2, 3], [2, 3], [2, 2]]
    bufs = [to_device(ctx, _make_tensor_data(s)) for s in shp]
    net = DeviceNetwork(idx, shp, [(1, 2), (0, 3)], ["A0", "D3"], ctx=ctx)
    out = nat.DeviceBuffer(64, ctx)
    net.execute([b.ptr for b in bufs], out.ptr)
    got = from_device(out, 4, np.complex64)
    assert np.array_equal(got, np.array([-308 - 56j, -517 - 94j, -1100 - 200j, -1804 - 328j],
                                        np.complex64))


def _circuit(r, c, d, seed=2026):
    circ = circuits.random_grid_circuit(r, c, d, seed=seed)
    net = circuits.amplitude_network(circ)
    dims = {l: 2 for ix in net.indices for l in ix}
    path = paths.best_greedy_path(net.indices, dims, trials=4)
    return circ, net, dims, path


@pytest.mark.parametrize("dtype,prec,tol", [(np.complex64, "fp32", 1e-5),
                                            (np.complex64, "bf16x3", 1e-5),
                                            (np.complex128, "fp64", 1e-12)])
def test_amplitude_matches_statevector(ctx, dtype, prec, tol):
    circ, net, dims, path = _circuit(4, 4, 8)
    want = circuits.amplitude_statevector(circ)
    say(f"amplitude 4x4x8 {prec}")
    bufs = [to_device(ctx, d.astype(dtype)) for d in net.data]
    dn = DeviceNetwork(net.indices, net.shapes, path, [], dtype=dtype, precision=prec, ctx=ctx)
    out = nat.DeviceBuffer(16, ctx)
    dn.execute([b.ptr for b in bufs], out.ptr)
    got = complex(from_device(out, 1, dtype)[0])
    assert abs(got - want) / abs(want) < tol, (got, want)
    # graph replay equals the eager result on every replay
    for rep in range(3):
        say(f"  graph replay {rep}")
        out2 = nat.DeviceBuffer(16, ctx)
        dn.launch([b.ptr for b in bufs], out2.ptr)
        dn.execute([b.ptr for b in bufs], out.ptr)
        ctx.synchronize()
        assert from_device(out2, 1, dtype)[0] == from_device(out, 1, dtype)[0]


def _per_slice_oracle(net, path, sliced, n):
    dims = {l: 2 for ix in net.indices for l in ix}
    res = []
    for i in range(n):
        vals = oracle.unravel_index(i, [dims[l] for l in sliced])
        fixed = dict(zip(sliced, vals))
        tensors = [_slice_leaf(ix, sh, d, fixed) for ix, sh, d in
                   zip(net.indices, net.shapes, net.data)]
        _, _, c = oracle.contract_path(tensors, path)
        res.append(c[0])
    return np.array(res)


@pytest.mark.parametrize("mode", ["all_batch", "all_loop", "mixed"])
def test_sliced_contraction(ctx, mode):
    circ, net, dims, path = _circuit(4, 4, 8, seed=7)
    sliced = paths.choose_slices(net.indices, dims, path, 5)
    batch = {"all_batch": sliced, "all_loop": [], "mixed": sliced[1::2]}[mode]
    say(f"sliced 4x4x8 {mode} {sliced} batch={batch}")
    sc = SlicedContraction(net.indices, net.shapes, net.data, path, sliced, batch=batch,
                           precision="fp32", ctx=ctx, use_graph=(mode != "all_loop"))
    sc.run()
    got = sc.slice_results().astype(np.complex128)
    want = _per_slice_oracle(net, path, sc.sliced, sc.n_slices)
    err = np.linalg.norm(got - want) / np.linalg.norm(want)
    assert err < 1e-5, err
    total = sc.reduce()
    amp = circuits.amplitude_statevector(circ)
    assert abs(total - amp) / abs(amp) < 1e-5
    assert abs(total - oracle.sum_compensated(got)) < 1e-6 * max(1.0, abs(total))


def test_c2_plans_agree_in_fp64(ctx):
    """The two committed plans of the 7x7 depth-14 amplitude (min-fill, and
    refined for batched execution) are contraction orders of the same
    sliced network: their first 16 slices agree in complex128 on FP64 MFMA
    to 1e-10, and the split-FP16 run of the refined plan is within 1e-5 of
    them (per-slice normwise).  Four sliced indices batched keep it small."""
    import json
    import os
    root = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                        "jet_amd", "data")
    circ = circuits.random_grid_circuit(7, 7, 14, seed=2026)
    net = circuits.amplitude_network(circ)
    res = {}
    for name, prec, dtype in (("grid_7x7x14_seed2026_s12.json", "fp64", np.complex128),
                              ("grid_7x7x14_seed2026_s12_refined.json", "fp64", np.complex128),
                              ("grid_7x7x14_seed2026_s12_refined.json", "f16x3", np.complex64)):
        with open(os.path.join(root, name)) as f:
            plan = json.load(f)
        sliced = plan["sliced"]
        say(f"c2 plan {name} {prec}: 16 slices")
        data = [d.astype(dtype) for d in net.data]
        sc = SlicedContraction(net.indices, net.shapes, data, [tuple(p) for p in plan["path"]],
                               sliced, batch=sliced[-4:], dtype=dtype, precision=prec, ctx=ctx,
                               use_graph=False)
        sc.run([0])
        res[(name, prec)] = sc.slice_results(16).astype(np.complex128)
        sc.net.close()
    ref = res[("grid_7x7x14_seed2026_s12.json", "fp64")]
    got64 = res[("grid_7x7x14_seed2026_s12_refined.json", "fp64")]
    got16 = res[("grid_7x7x14_seed2026_s12_refined.json", "f16x3")]
    assert np.linalg.norm(got64 - ref) / np.linalg.norm(ref) < 1e-10
    assert np.linalg.norm(got16 - ref) / np.linalg.norm(ref) < 1e-5

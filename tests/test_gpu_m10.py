"""The reference's m10 network (a TensorNetworkSerializer file) on the GPU.

tests/golden/m10.json is examples/paper_benchmarks/data_files/m10.json of the
Jet tree (322 rank-3 bond-2 tensors, a stored 321-step SSA path);
tests/golden/m10_values.json holds its complex128 CPU-oracle values: the
whole network along the stored path, and the 16 slices over p7, s7, h4, m1
(the first four sliced indices of examples/paper_benchmarks/CPU/
jet_cpu_m10/jet_sliced.cpp:53-54), both from tests/golden/
make_m10_fixture.py.  The file is loaded through TensorNetworkSerializer
(TensorNetworkIO.hpp:94-293) and contracted along its stored path on the
GPU: through TensorNetwork::Contract(path) (one ContractTensors per step),
through one planned device network per precision, and as the sliced
contraction of jet_sliced.cpp (TensorNetwork::SliceIndices + a reduction)
with the slices looped or batched.  Tolerances: 1e-5 relative for complex64,
1e-10 for complex128.
"""
import json
import os

import numpy as np
import pytest

from jet_amd import _native as nat
from jet_amd.network import DeviceNetwork, SlicedContraction
from jet_amd.tensor_network_io import TensorNetworkSerializer

from hiputil import say, to_device, from_device

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module")
def m10():
    with open(os.path.join(GOLDEN, "m10.json")) as f:
        text = f.read()
    with open(os.path.join(GOLDEN, "m10_values.json")) as f:
        vals = json.load(f)
    doc = json.loads(text)
    idx = [list(t[1]) for t in doc["tensors"]]
    shp = [list(t[2]) for t in doc["tensors"]]
    data = [np.array([complex(re, im) for re, im in t[3]], np.complex64) for t in doc["tensors"]]
    path = [tuple(p) for p in doc["path"]]
    return text, idx, shp, data, path, vals


def _rel(got, want):
    return abs(complex(got) - want) / abs(want)


@pytest.mark.parametrize("dtype,tol", [(np.complex64, 1e-5), (np.complex128, 1e-10)])
def test_m10_serializer_and_contract_path(ctx, m10, dtype, tol):
    text, _, _, _, _, vals = m10
    say(f"m10 serializer + TensorNetwork.contract(path) {np.dtype(dtype).name}")
    f = TensorNetworkSerializer(dtype=dtype)(text)
    assert f.tensors.num_tensors == 322 and len(f.path.path) == 321
    result = f.tensors.contract(f.path.path)
    assert result.indices == []
    # A complex128 serializer keeps the file's decimal values unrounded.
    want = complex(*vals["whole" if dtype == np.complex64 else "whole_unrounded"])
    assert _rel(result.get_data()[0], want) < tol


@pytest.mark.parametrize("prec,dtype,tol", [("fp32", np.complex64, 1e-5),
                                            ("f16x3", np.complex64, 1e-5),
                                            ("fp64", np.complex128, 1e-10)])
def test_m10_device_network(ctx, m10, prec, dtype, tol):
    _, idx, shp, data, path, vals = m10
    say(f"m10 device network {prec}")
    bufs = [to_device(ctx, d.astype(dtype)) for d in data]
    net = DeviceNetwork(idx, shp, path, [], dtype=dtype, precision=prec, ctx=ctx)
    out = nat.DeviceBuffer(16, ctx)
    net.execute([b.ptr for b in bufs], out.ptr)
    got = from_device(out, 1, dtype)[0]
    net.close()
    assert _rel(got, complex(*vals["whole"])) < tol


@pytest.mark.parametrize("batch", [0, 2, 4])
@pytest.mark.parametrize("prec,dtype,tol", [("fp32", np.complex64, 1e-5),
                                            ("fp64", np.complex128, 1e-10)])
def test_m10_sliced(ctx, m10, batch, prec, dtype, tol):
    _, idx, shp, data, path, vals = m10
    sliced = vals["sliced"]
    say(f"m10 sliced on {sliced}, {batch} batched, {prec}")
    sc = SlicedContraction(idx, shp, [d.astype(dtype) for d in data], path, sliced,
                           batch=sliced[len(sliced) - batch:] if batch else [], dtype=dtype,
                           precision=prec, ctx=ctx, use_graph=batch == 2)
    sc.run()
    got = sc.slice_results().astype(np.complex128)
    want = np.array([complex(*v) for v in vals["slice_values"]])
    assert np.linalg.norm(got - want) / np.linalg.norm(want) < tol
    total = sc.reduce()
    sc.net.close()
    assert _rel(total, complex(*vals["whole"])) < tol

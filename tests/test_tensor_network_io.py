"""TensorNetworkSerializer / TensorNetworkFile (jet_amd.tensor_network_io).

Cases and literal strings from test/Test_TensorNetworkIO.cpp:12-130 and
python/tests/test_tensor_network_io.py:7-120, run with the numpy tensor of
the network-class tests (CPU) and with jet_amd.Tensor (GPU).  The reference
network files m10.json / m12.json load, re-serialize to the same JSON and
give PathInfo flop counts computed independently from their paths.
"""
import json
import os
import re

import numpy as np
import pytest

import oracle
from hosttensor import HostTensor
from jet_amd import PathInfo, TensorNetwork
from jet_amd.tensor_network_io import (TensorFileException, TensorNetworkFile,
                                       TensorNetworkSerializer)

DATA = "/root/reference/examples/paper_benchmarks/data_files"


@pytest.fixture(params=["host", pytest.param("gpu", marks=pytest.mark.gpu)])
def TT(request):
    if request.param == "host":
        return HostTensor
    from jet_amd import Tensor
    return Tensor


def _compact(s):
    return json.dumps(json.loads(s), separators=(",", ":"), sort_keys=True)


FOUR = """{"tensors": [
        [["I0"], ["a"], [2], [[1.0, 0.0], [0.0,0.0]]],
        [["I1"], ["b"], [2], [[1.0, 0.0], [0.0,0.0]]],
        [["I2"], ["c"], [2], [[1.0, 0.0], [0.0,0.0]]],
        [["I3"], ["d"], [2], [[1.0, 0.0], [0.0,0.0]]]]}"""


@pytest.mark.parametrize("dtype", [np.complex64, np.complex128])
def test_errors(TT, dtype):
    ser = TensorNetworkSerializer(tensor_type=TT, dtype=dtype)
    with pytest.raises(ValueError):
        ser("")
    for bad in ("[]", "{}", '{"path": [[0,1]]}'):
        with pytest.raises(TensorFileException):
            ser(bad)
    bad = """{"tensors": [
        [["I0"], ["a"], [2], [[1.0], [0.0,0.0]]],
        [["I1"], ["b"], [2], [[1.0, 0.0], [0.0,0.0]]]]}"""
    with pytest.raises(TensorFileException, match=re.escape("[1.0]")):
        ser(bad)


@pytest.mark.parametrize("dtype", [np.complex64, np.complex128])
def test_round_trips(TT, dtype):
    ser = TensorNetworkSerializer(tensor_type=TT, dtype=dtype)
    js = _compact(FOUR)
    f = ser(js)
    assert f.path is None
    assert f.tensors.num_indices == 4 and f.tensors.num_tensors == 4
    assert ser(f.tensors) == js

    js = _compact('{"path": [[0, 1], [1, 2]], ' + FOUR[1:])
    f = ser(js)
    assert f.path.path == [(0, 1), (1, 2)]
    assert ser(f.tensors, f.path) == js

    js = _compact(FOUR.replace('["c"]', '["a"]').replace('["d"]', '["b"]'))
    f = ser(js)
    assert f.tensors.num_indices == 2 and f.tensors.num_tensors == 4
    assert ser(f.tensors) == js


PY_NET = """{"tensors": [
    [["A", "Hermitian"], ["i", "j"], [2, 2], [[1.0, 0.0], [0.0, 1.0], [-0.0, -1.0], [1.0, 0.0]]],
    [["B", "Identity", "Real"], ["j", "k"], [2, 2], [[1.0, 0.0], [0.0, 0.0], [0.0, 0.0], [1.0, 0.0]]],
    [["C", "Vector", "Real"], ["k"], [2], [[1.0, 0.0], [0.0, 0.0]]]]}"""


@pytest.mark.parametrize("dtype", [np.complex64, np.complex128])
def test_python_fixture_network(TT, dtype):
    tn = TensorNetwork(TT)
    tn.add_tensor(TT(["i", "j"], [2, 2], np.array([1, 1j, -1j, 1]), dtype=dtype),
                  ["A", "Hermitian"])
    tn.add_tensor(TT(["j", "k"], [2, 2], np.array([1, 0, 0, 1]), dtype=dtype),
                  ["B", "Identity", "Real"])
    tn.add_tensor(TT(["k"], [2], np.array([1, 0]), dtype=dtype), ["C", "Vector", "Real"])
    ser = TensorNetworkSerializer(tensor_type=TT, dtype=dtype)
    assert ser(tn) == re.sub(r"\s+", "", PY_NET)
    path = PathInfo(tn, [[0, 2], [2, 1]])
    want = re.sub(r"\s+", "", '{"path": [[0, 2], [2, 1]], ' + PY_NET[1:])
    assert ser(tn, path) == want
    f = ser(want)
    assert f.path.path == [(0, 2), (2, 1)] and len(f.tensors.nodes) == 3
    assert TensorNetworkFile().path is None and len(TensorNetworkFile().tensors.nodes) == 0


def test_column_major_load():
    js = _compact('{"tensors": [[["T"], ["a", "b"], [2, 3], '
                  '[[0,0],[1,0],[2,0],[3,0],[4,0],[5,0]]]]}')
    f = TensorNetworkSerializer(tensor_type=HostTensor)(js, True)
    t = f.tensors.nodes[0].tensor
    assert t.indices == ["b", "a"] and t.shape == [3, 2]


@pytest.mark.parametrize("name,n_tensors,n_steps,pathinfo_flops,max_rank",
                         [("m10.json", 322, 321, 1.279e10, 26),
                          ("m12.json", 410, 409, 2.027e13, 36)])
def test_reference_network_files(name, n_tensors, n_steps, pathinfo_flops, max_rank):
    path = os.path.join(DATA, name)
    if not os.path.exists(path):
        pytest.skip("reference data files not present")
    text = open(path).read()
    f = TensorNetworkSerializer(tensor_type=HostTensor, dtype=np.complex64)(text)
    tn, info = f.tensors, f.path
    assert tn.num_tensors == n_tensors and len(info.path) == n_steps
    assert abs(info.total_flops() / pathinfo_flops - 1) < 1e-3
    assert max(len(s.tensor_indices) for s in info.steps) == max_rank
    # float32 data written back as the doubles they are: same values
    again = json.loads(TensorNetworkSerializer(tensor_type=HostTensor, dtype=np.complex64)(tn, info))
    orig = json.loads(text)
    assert again["path"] == orig["path"]
    for a, b in zip(again["tensors"], orig["tensors"]):
        assert a[:3] == b[:3]
        np.testing.assert_allclose(np.array(a[3]), np.array(b[3]), rtol=1e-7, atol=1e-7)


def test_m10_fixture_slices_sum_to_whole():
    """The committed m10 oracle values are consistent: summing the 16
    slices over p7, s7, h4, m1 restores the unsliced contraction
    (TensorNetwork::SliceIndices keeps every term of the sum once)."""
    import json
    golden = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
    with open(os.path.join(golden, "m10_values.json")) as f:
        vals = json.load(f)
    total = oracle.sum_compensated(np.array([complex(*v) for v in vals["slice_values"]]))
    whole = complex(*vals["whole"])
    assert abs(total - whole) / abs(whole) < 1e-12
    with open(os.path.join(golden, "m10.json")) as f:
        doc = json.load(f)
    assert len(doc["tensors"]) == 322 and len(doc["path"]) == 321

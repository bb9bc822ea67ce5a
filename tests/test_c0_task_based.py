"""Config C0: one amplitude of a seeded 4x4-qubit grid circuit of depth 8
through TensorNetwork + PathInfo + TaskBasedContractor with a greedy path,
checked against the statevector (closed form); and the same amplitude as
the reduction of sliced networks, where slice-independent contractions are
shared tasks (the reference's slice loop,
examples/paper_benchmarks/CPU/jet_cpu_m10/jet_sliced.cpp:53-93).

CPU: the numpy tensor of the network-class tests (oracle contractions).
GPU: jet_amd.Tensor (libjet_hip kernels), complex64, error <= 1e-5.
"""
import numpy as np
import pytest

from hosttensor import HostTensor
from jet_amd import PathInfo, TaskBasedContractor, TensorNetwork, circuits, paths


@pytest.fixture(params=["host", pytest.param("gpu", marks=pytest.mark.gpu)])
def TT(request):
    if request.param == "host":
        return HostTensor
    from jet_amd import Tensor
    return Tensor


def _c0(TT):
    circ = circuits.random_grid_circuit(4, 4, 8, seed=2026)
    net = circuits.amplitude_network(circ)
    tn = TensorNetwork(TT)
    for ix, sh, d, tags in zip(net.indices, net.shapes, net.data, net.tags):
        tn.add_tensor(TT(ix, sh, d.astype(np.complex64), dtype=np.complex64), tags)
    dims = {l: 2 for ix in net.indices for l in ix}
    path = paths.best_greedy_path(net.indices, dims, trials=4, seed=0)
    return circ, net, tn, path, dims


def test_c0_amplitude_task_based(TT):
    circ, net, tn, path, _ = _c0(TT)
    want = circuits.amplitude_statevector(circ)
    info = PathInfo(tn, path)
    tbc = TaskBasedContractor(tensor_type=TT, dtype=np.complex64)
    assert tbc.add_contraction_tasks(tn, info) == 0
    tbc.add_deletion_tasks()
    tbc.add_reduction_task()
    tbc.contract()
    got = complex(np.asarray(tbc.results[0].get_data()).reshape(-1)[0])
    assert abs(got - want) / abs(want) < 1e-5, (got, want)
    assert tbc.flops == info.total_flops()


def test_c0_sliced_reduction_shares_tasks(TT):
    circ, net, tn, path, dims = _c0(TT)
    want = circuits.amplitude_statevector(circ)
    sliced = paths.choose_slices(net.indices, dims, path, 3)
    tbc = TaskBasedContractor(tensor_type=TT, dtype=np.complex64)
    shared = 0
    for value in range(2 ** len(sliced)):
        t = tn.copy()
        t.slice_indices(sliced, value)
        shared += tbc.add_contraction_tasks(t, PathInfo(t, path))
    assert tbc.add_reduction_task() == 1
    tbc.contract()
    assert len(tbc.results) == 8 and shared > 0
    got = complex(np.asarray(tbc.reduction_result.get_data()).reshape(-1)[0])
    assert abs(got - want) / abs(want) < 1e-5, (got, want)

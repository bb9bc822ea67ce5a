"""CPU checks of the circuit generator, the path finders and network plans.

* gate matrices: unitarity, sqrt(U)^2 == U (closed forms);
* amplitude networks: the FP64 contraction along a found path equals the
  statevector amplitude (a closed-form simulation independent of the
  tensor-network code);
* paths are valid SSA paths and path_stats follows PathInfo's flop
  convention (|out| * 2 * prod contracted, PathInfo.hpp:157-183; pinned by
  Test_PathInfo.cpp:316-329);
* network plans (tests/native/network_check.cpp): arena bounds and overlap
  of live intermediates, before anything runs on a GPU.
"""
import ctypes
import os
import subprocess

import numpy as np
import pytest

import oracle
from jet_amd import _native as nat
from jet_amd import circuits, paths
from jet_amd.network import choose_batch, estimate, LabelIds, _flatten

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CHECK_SRC = os.path.join(ROOT, "tests", "native", "network_check.cpp")
CHECK_LIB = os.path.join(ROOT, "tests", "native", "_build", "libnetcheck.so")


def test_gate_matrices():
    for name, u in circuits.SINGLE_QUBIT_GATES:
        assert np.allclose(u.conj().T @ u, np.eye(2))
    assert np.allclose(circuits.SQRT_X @ circuits.SQRT_X, [[0, 1], [1, 0]])
    assert np.allclose(circuits.SQRT_Y @ circuits.SQRT_Y, [[0, -1j], [1j, 0]])
    w = (np.array([[0, 1], [1, 0]]) + np.array([[0, -1j], [1j, 0]])) / np.sqrt(2)
    assert np.allclose(circuits.SQRT_W @ circuits.SQRT_W, w)
    f = circuits.fsim(np.pi / 2, np.pi / 6)
    assert np.allclose(f.conj().T @ f, np.eye(4))
    assert np.isclose(f[3, 3], np.exp(-1j * np.pi / 6)) and np.isclose(f[1, 2], -1j)


def test_grid_patterns_cover_every_edge_once():
    pats = circuits.grid_edges(7, 7)
    edges = [e for p in pats for e in p]
    assert len(edges) == len(set(edges)) == 2 * 7 * 6
    for p in pats:  # no qubit twice in one pattern
        qs = [q for e in p for q in e]
        assert len(qs) == len(set(qs))


@pytest.mark.parametrize("finder", ["greedy", "elimination"])
def test_amplitude_network_matches_statevector(finder):
    circ = circuits.random_grid_circuit(3, 4, 6, seed=11)
    net = circuits.amplitude_network(circ)
    dims = {l: 2 for ix in net.indices for l in ix}
    if finder == "greedy":
        path = paths.best_greedy_path(net.indices, dims, trials=2)
    else:
        path = paths.elimination_path(net.indices, dims, seed=1, noise=0.5)
    assert len(path) == len(net) - 1
    used = [x for p in path for x in p]
    assert len(used) == len(set(used))
    tensors = list(zip(net.indices, net.shapes, net.data))
    labels, shape, amp = oracle.contract_path(tensors, path)
    assert labels == [] and shape == []
    want = circuits.amplitude_statevector(circ)
    assert abs(amp[0] - want) < 1e-12


def test_path_stats_pathinfo_convention():
    # Test_PathInfo.cpp:316-329: contracting (a:4, b:6) with (b:6, c:5)
    # gives |out| = 4*5 and 2*6 flops per element.
    idx = [["a", "b"], ["b", "c"]]
    dims = {"a": 4, "b": 6, "c": 5}
    st = paths.path_stats(idx, dims, [(0, 1)])
    assert st.total_flops == (4 * 5) * (6 + 6)


def _check_lib():
    srcs = [CHECK_SRC] + [os.path.join(ROOT, "jet_amd", "csrc", f) for f in
                          ("network_plan.cpp", "contract_plan.cpp", "permute_plan.cpp")]
    if not os.path.exists(CHECK_LIB) or os.path.getmtime(CHECK_LIB) < max(map(os.path.getmtime, srcs)):
        os.makedirs(os.path.dirname(CHECK_LIB), exist_ok=True)
        subprocess.run(["g++", "-std=c++17", "-O2", "-fPIC", "-shared",
                        "-I" + os.path.join(ROOT, "include"), "-o", CHECK_LIB] + srcs, check=True)
    return ctypes.CDLL(CHECK_LIB)


def _check(L, indices, shapes, path, out, mode=0, elem=8):
    lab = LabelIds()
    ranks, modes, exts = _flatten(indices, shapes, lab)
    o = [lab(l) for l in out]
    fl = ctypes.c_double()
    ar, ws = ctypes.c_size_t(), ctypes.c_size_t()
    err = ctypes.create_string_buffer(256)
    flat = [int(x) for p in path for x in p]
    rc = L.check_network(elem, mode, len(indices), nat.i32(ranks), nat.i32(modes), nat.i64(exts),
                         len(path), nat.i32(flat), len(o), nat.i32(o), ctypes.byref(fl),
                         ctypes.byref(ar), ctypes.byref(ws), err, 256)
    return rc, fl.value, ar.value, ws.value, err.value.decode()


@pytest.mark.parametrize("shape,n_sliced", [((4, 4, 8), 0), ((4, 4, 8), 5), ((5, 5, 10), 8)])
def test_network_plans_arena_and_flops(shape, n_sliced):
    L = _check_lib()
    circ = circuits.random_grid_circuit(*shape, seed=5)
    net = circuits.amplitude_network(circ)
    dims = {l: 2 for ix in net.indices for l in ix}
    path = paths.best_greedy_path(net.indices, dims, trials=2)
    sliced = paths.choose_slices(net.indices, dims, path, n_sliced) if n_sliced else []
    rc, fl, ar, ws, err = _check(L, net.indices, net.shapes, path, sliced)
    assert rc == 0, (rc, err)
    if not sliced:
        st = paths.path_stats(net.indices, dims, path)
        assert fl == pytest.approx(4 * st.total_flops)
    # the library's own estimate agrees with the checked plan
    fl2, ar2, ws2 = estimate(net.indices, net.shapes, path, sliced)
    assert (fl2, ar2, ws2) == (fl, ar, ws)


def test_network_plan_errors():
    L = _check_lib()
    rc, *_, err = _check(L, [["a"], ["a"]], [[2], [2]], [(0, 0)], [])
    assert rc == 3 and "invalid" in err
    rc, *_, err = _check(L, [["a", "b"], ["a"]], [[2, 2], [2]], [(0, 1)], [])
    assert rc == 3 and "Output indices" in err


def test_choose_batch_respects_budget():
    circ = circuits.random_grid_circuit(4, 4, 8, seed=3)
    net = circuits.amplitude_network(circ)
    dims = {l: 2 for ix in net.indices for l in ix}
    path = paths.best_greedy_path(net.indices, dims, trials=2)
    sliced = paths.choose_slices(net.indices, dims, path, 6)
    assert choose_batch(net.indices, net.shapes, path, sliced, 1e12) == sliced
    small = choose_batch(net.indices, net.shapes, path, sliced, 64 * 1024)
    assert len(small) < len(sliced)


@pytest.mark.parametrize("mode", [0, 4])
def test_stacked_amplitude_plan(mode):
    """The stacked form of an amplitude batch (a bitstring label on every
    output vector, kept open): the plan stays inside its arena and its
    flops are at least those of the single-amplitude plan."""
    L = _check_lib()
    circ = circuits.random_grid_circuit(4, 4, 8, seed=4)
    net = circuits.amplitude_network(circ)
    dims = {l: 2 for ix in net.indices for l in ix}
    path = paths.best_minfill_path(net.indices, dims, trials=2, seed=1)
    idx = [list(ix) for ix in net.indices]
    shp = [list(s) for s in net.shapes]
    nq = circ.n_qubits
    for q in range(nq):
        k = len(idx) - nq + q
        idx[k] = ["%amplitude"] + idx[k]
        shp[k] = [6] + shp[k]
    rc, fl, ar, ws, err = _check(L, idx, shp, path, ["%amplitude"], mode=mode)
    assert rc == 0, err
    rc1, fl1, *_ = _check(L, net.indices, net.shapes, path, [], mode=mode)
    assert rc1 == 0 and fl >= fl1

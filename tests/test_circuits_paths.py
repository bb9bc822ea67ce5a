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
import string
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
    from jet_amd import build as jb
    jb.write_tuning_table()   # the compiled-in GEMM table the planner reads
    srcs = [CHECK_SRC] + [os.path.join(ROOT, "jet_amd", "csrc", f) for f in
                          ("network_plan.cpp", "contract_plan.cpp", "permute_plan.cpp",
                           "tuning.cpp")]
    deps = srcs + [jb.TUNING_INC]
    if not os.path.exists(CHECK_LIB) or os.path.getmtime(CHECK_LIB) < max(map(os.path.getmtime, deps)):
        os.makedirs(os.path.dirname(CHECK_LIB), exist_ok=True)
        # built under a per-process name and renamed into place, so that
        # parallel test workers never load a half-written library
        tmp = f"{CHECK_LIB}.{os.getpid()}.tmp"
        subprocess.run(["g++", "-std=c++17", "-O2", "-fcx-limited-range", "-fPIC", "-shared",
                        "-I" + os.path.join(ROOT, "include"), "-I" + jb.BUILD, "-o", tmp]
                       + srcs, check=True)
        os.replace(tmp, CHECK_LIB)
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


def _emu(L, indices, shapes, data, path, out, mode):
    lab = LabelIds()
    ranks, modes, exts = _flatten(indices, shapes, lab)
    o = [lab(l) for l in out]
    flat = [int(x) for p in path for x in p]
    arrs = [np.ascontiguousarray(np.asarray(d, np.complex128).reshape(-1)) for d in data]
    ptrs = (ctypes.c_void_p * len(arrs))(*[a.ctypes.data for a in arrs])
    dims = dict(zip([l for ix in indices for l in ix], [s for sh in shapes for s in sh]))
    res = np.zeros(int(np.prod([dims[l] for l in out])) if out else 1, np.complex128)
    perm = ctypes.c_double()
    err = ctypes.create_string_buffer(256)
    L.emu_network.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                              ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_int,
                              ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                              ctypes.POINTER(ctypes.c_double), ctypes.c_char_p, ctypes.c_int]
    r32, m32, e64, p32, o32 = (nat.i32(ranks), nat.i32(modes), nat.i64(exts), nat.i32(flat),
                               nat.i32(o))
    rc = L.emu_network(mode, len(indices), r32, m32, e64, len(path), p32, len(o), o32, ptrs,
                       res.ctypes.data, ctypes.byref(perm), err, 256)
    assert rc == 0, (rc, err.value)
    return res, perm.value


def test_network_plan_emulation_matches_numpy():
    """A network plan replayed on the host (tests/native/network_check.cpp:
    the plan's own permutation index arithmetic, the batched NT products,
    the result permutations) equals numpy along the same path, for a network
    with scrambled leaf mode orders, an open batch mode, and steps of the
    shapes that run on the 128x192 split kernel (FP32 and split-FP16 plans)."""
    L = _check_lib()
    rng = np.random.default_rng(17)
    # a(7 bits) c(7) n(12) g(7) h(1, open batch) labels; bond 2
    a = [f"a{i}" for i in range(7)]
    c = [f"c{i}" for i in range(7)]
    nn = [f"n{i}" for i in range(12)]
    g = [f"g{i}" for i in range(7)]
    hb = ["h0"]
    order = lambda ls: [ls[i] for i in rng.permutation(len(ls))]
    indices = [order(a + c + hb), order(c + nn + hb), order(nn[:6] + g + hb),
               order(nn[6:] + a + g + hb)]
    shapes = [[2] * len(ix) for ix in indices]
    data = [(rng.standard_normal(2 ** len(ix)) + 1j * rng.standard_normal(2 ** len(ix))) / 4
            for ix in indices]
    path = [(0, 1), (4, 2), (5, 3)]
    out = ["h0"]
    for mode in (0, 4):
        got, permuted = _emu(L, indices, shapes, data, path, out, mode)
        # numpy along the same path, pairwise (the open batch labels kept)
        nodes = [(list(ix), d.reshape([2] * len(ix))) for ix, d in zip(indices, data)]
        for (x, y) in path:
            (lx, tx), (ly, ty) = nodes[x], nodes[y]
            keep = [l for l in lx if l not in ly or l in out] + [l for l in ly if l not in lx]
            letters = {l: string.ascii_letters[i] for i, l in enumerate(dict.fromkeys(lx + ly))}
            spec = "".join(letters[l] for l in lx) + "," + "".join(letters[l] for l in ly) + \
                "->" + "".join(letters[l] for l in keep)
            nodes.append((keep, np.einsum(spec, tx, ty)))
        lr, tr = nodes[-1]
        want = np.transpose(tr, [lr.index(l) for l in out])
        assert np.allclose(got.reshape(want.shape), want, rtol=1e-10, atol=1e-10), mode
        assert permuted > 0      # the scrambled leaves are permuted

"""Subtree reconfiguration of contraction paths (jet_amd.paths).

* On a 6-tensor network, reconfiguring the whole tree (subtree size >= the
  number of tensors) reaches the cheapest of all contraction trees, found
  here by brute force over every pairwise order.
* On a 4x4 depth-8 circuit amplitude: the refined path is a valid SSA path
  of the same network, never costs more than the input path under the
  cost it minimises, and respects the intermediate-size limit.
* The MI355X step-time model charges the split GEMM for large M, N, K and
  memory traffic elsewhere.
"""
import itertools
import math

import numpy as np
import pytest

from jet_amd import circuits, paths


def _cost_of(indices, dims, path):
    return paths.path_stats(indices, dims, path).total_flops


def _brute_force_best(indices, dims, open_labels=frozenset()):
    """Cheapest PathInfo flops over all contraction orders (tiny networks);
    open labels stay on the result of every contraction that holds them."""
    best = math.inf

    def rec(nodes, cost):
        nonlocal best
        if cost >= best:
            return
        if len(nodes) == 1:
            best = cost
            return
        for i, j in itertools.combinations(range(len(nodes)), 2):
            a, b = nodes[i], nodes[j]
            out = (a ^ b) | (a & b & open_labels)
            c = math.prod(dims[l] for l in out) * 2 * math.prod(dims[l] for l in (a & b) - out)
            rest = [n for k, n in enumerate(nodes) if k not in (i, j)] + [out]
            rec(rest, cost + c)

    rec([frozenset(ix) for ix in indices], 0)
    return best


def test_whole_tree_reconfiguration_is_optimal():
    rng = np.random.default_rng(7)
    # a ring of 6 tensors with a chord, random bond dimensions
    indices = [["a", "b", "g"], ["b", "c"], ["c", "d", "h"], ["d", "e"], ["e", "f", "g"],
               ["f", "a", "h"]]
    dims = {l: int(rng.integers(2, 6)) for l in "abcdefgh"}
    path = [(0, 1), (2, 3), (4, 5), (6, 7), (8, 9)]
    got = paths.reconfigure_path(indices, dims, path, subtree_size=8, passes=4)
    assert _cost_of(indices, dims, got) == pytest.approx(_brute_force_best(indices, dims))


def test_open_label_reconfiguration_is_optimal():
    """Batch labels kept open (on every tensor of a batched execution that
    holds one): the whole-tree reconfiguration still reaches the brute-force
    optimum of the batched cost, and a time-model cost receives the batch
    extent as its fourth argument."""
    rng = np.random.default_rng(11)
    indices = [["a", "b", "x"], ["b", "c", "y"], ["c", "d"], ["d", "e", "x"], ["e", "f", "y"],
               ["f", "a", "x"]]
    dims = {l: int(rng.integers(2, 6)) for l in "abcdefxy"}
    path = [(0, 1), (2, 3), (4, 5), (6, 7), (8, 9)]
    open_labels = {"x", "y"}
    got = paths.reconfigure_path(indices, dims, path, subtree_size=8, passes=4,
                                 open_labels=open_labels)
    tree = paths._Tree(indices, dims, got, open_labels=open_labels)
    assert tree.mask[tree.root] == tree.open       # only the open labels remain
    assert tree.total() == pytest.approx(_brute_force_best(indices, dims, frozenset(open_labels)))
    seen = []

    def cost(lm, ln, lk, lb):
        seen.append(lb)
        return 2.0 ** (lm + ln + lk + lb)

    paths.reconfigure_path(indices, dims, path, subtree_size=8, passes=1, step_cost=cost,
                           open_labels=open_labels)
    assert max(seen) > 0


def _valid_ssa(n, path):
    used = set()
    for s, (a, b) in enumerate(path):
        assert a != b and a < n + s and b < n + s and a not in used and b not in used
        used.update((a, b))
    assert len(path) == n - 1


def test_refined_circuit_path():
    circ = circuits.random_grid_circuit(4, 4, 8, seed=3)
    net = circuits.amplitude_network(circ)
    dims = {l: 2 for ix in net.indices for l in ix}
    path = paths.best_greedy_path(net.indices, dims, trials=1)
    sliced = paths.choose_slices_for_flops(net.indices, dims, path, 3, 12)
    before = paths.path_stats(net.indices, dims, path, sliced=sliced)
    for cost in (None, lambda lm, ln, lk: paths.mi355x_step_seconds(lm, ln, lk)):
        p2 = paths.reconfigure_path(net.indices, dims, path, sliced, max_log2_size=12,
                                    subtree_size=8, passes=3, step_cost=cost)
        _valid_ssa(len(net.indices), p2)
        after = paths.path_stats(net.indices, dims, p2, sliced=sliced)
        assert after.max_log2_size <= max(12, before.max_log2_size)
        t0 = paths._Tree(net.indices, dims, path, sliced, 0.0, cost).total()
        t1 = paths._Tree(net.indices, dims, p2, sliced, 0.0, cost).total()
        assert t1 <= t0 * (1 + 1e-12)
        if cost is None:
            assert after.total_flops <= before.total_flops


def test_step_time_model():
    big = paths.mi355x_step_seconds(14, 14, 14)
    gemm = 8 * 2.0 ** 42 / 470e12           # the MFMA time of a 2^14-cube GEMM
    assert gemm < big < 1.2 * gemm           # plus operand preparation
    # a K=2 outer product of the same output size is memory-bound, far cheaper
    # in flops but not free
    outer = paths.mi355x_step_seconds(14, 14, 1)
    assert outer < big and outer > 8 * 2.0 ** 28 / 3e12


def test_committed_c2_plans_are_consistent():
    """The committed C2 plans (min-fill, and refined for batched execution)
    are valid SSA paths of the seeded 7x7 depth-14 amplitude network over
    the same sliced indices; their recorded per-slice PathInfo flops and
    sizes match path_stats; the refined one is cheaper per slice and in the
    modelled batched execution it was refined for, whose largest batched
    intermediate grew by at most the allowed extra_log2."""
    import json
    import os
    root = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                        "jet_amd", "data")
    circ = circuits.random_grid_circuit(7, 7, 14, seed=2026)
    net = circuits.amplitude_network(circ)
    dims = {l: 2 for ix in net.indices for l in ix}
    per, plans = {}, {}
    for name in ("grid_7x7x14_seed2026_s12.json", "grid_7x7x14_seed2026_s12_refined.json"):
        with open(os.path.join(root, name)) as f:
            plan = json.load(f)
        path = [tuple(p) for p in plan["path"]]
        _valid_ssa(len(net.indices), path)
        assert len(plan["sliced"]) == 12 and len(set(plan["sliced"])) == 12
        st = paths.path_stats(net.indices, dims, path, sliced=plan["sliced"])
        assert st.total_flops == plan["per_slice_pathinfo_flops"]
        assert st.max_log2_size == plan["per_slice_max_log2_size"] <= 28
        per[name] = st.total_flops
        plans[name] = (path, plan)
    base_path, base = plans["grid_7x7x14_seed2026_s12.json"]
    ref_path, ref = plans["grid_7x7x14_seed2026_s12_refined.json"]
    assert ref["sliced"] == base["sliced"]
    assert per["grid_7x7x14_seed2026_s12_refined.json"] < per["grid_7x7x14_seed2026_s12.json"] / 3
    info = ref["refined"]
    nb = info["batch_log2"]
    t0 = paths.batched_tree(net.indices, dims, base_path, base["sliced"], nb)
    t1 = paths.batched_tree(net.indices, dims, ref_path, ref["sliced"], nb)
    assert [t0.total(), t1.total()] == pytest.approx(info["modelled_execution_s"], rel=1e-12)
    assert t1.total() < 0.5 * t0.total()
    assert t1.max_log2size() <= t0.max_log2size() + info["extra_log2"]


def test_refine_batched_path_small_circuit():
    """refine_batched_path on a 5x5 depth-10 amplitude sliced on 6 indices,
    3 of them batched: a valid SSA path of the same network whose modelled
    batched execution is no slower than the input's, within the allowed
    growth of the largest batched intermediate."""
    circ = circuits.random_grid_circuit(5, 5, 10, seed=4)
    net = circuits.amplitude_network(circ)
    dims = {l: 2 for ix in net.indices for l in ix}
    path = paths.best_greedy_path(net.indices, dims, trials=1)
    sliced = paths.choose_slices_for_flops(net.indices, dims, path, 6, 14)
    t0 = paths.batched_tree(net.indices, dims, path, sliced, 3)
    p2 = paths.refine_batched_path(net.indices, dims, path, sliced, 3, extra_log2=1.0,
                                   subtree_size=6, rounds=2)
    _valid_ssa(len(net.indices), p2)
    t1 = paths.batched_tree(net.indices, dims, p2, sliced, 3)
    assert t1.total() <= t0.total() * (1 + 1e-12)
    assert t1.max_log2size() <= t0.max_log2size() + 1.0
    assert t1.mask[t1.root] == t1.open        # only the batched indices remain open


def test_reconfigured_minfill_path_amplitude():
    """reconfigured_minfill_path on a 4x4 depth-8 amplitude: a valid SSA
    path, no larger intermediate than the min-fill path it starts from and
    no more PathInfo flops after the flops stage alone, and the oracle
    contracted along it gives the statevector amplitude."""
    import oracle
    circ = circuits.random_grid_circuit(4, 4, 8, seed=9)
    net = circuits.amplitude_network(circ)
    dims = {l: 2 for ix in net.indices for l in ix}
    base = paths.best_minfill_path(net.indices, dims, trials=4, seed=1)
    flops_only = paths.reconfigured_minfill_path(net.indices, dims, trials=4, seed=1,
                                                 subtree_sizes=(6, 8), passes=3, step_cost=None)
    b = paths.path_stats(net.indices, dims, base)
    assert paths.path_stats(net.indices, dims, flops_only).total_flops <= b.total_flops
    path = paths.reconfigured_minfill_path(net.indices, dims, trials=4, seed=1,
                                           subtree_sizes=(6, 8), passes=3)
    _valid_ssa(len(net.indices), path)
    assert paths.path_stats(net.indices, dims, path).max_log2_size <= b.max_log2_size
    tensors = list(zip(net.indices, net.shapes, [d.astype(np.complex128) for d in net.data]))
    _, _, amp = oracle.contract_path(tensors, path)
    want = circuits.amplitude_statevector(circ)
    assert abs(complex(np.asarray(amp).reshape(-1)[0]) - want) < 1e-12 * max(1.0, abs(want))


def test_committed_amplitude_paths_are_consistent():
    """jet_amd/data/amplitude_paths.json (the uneven-batch and gradient
    circuits): each path is a valid SSA path of its seeded amplitude
    network, its recorded PathInfo flops and largest intermediate match
    path_stats, and its modelled time is below the min-fill path's."""
    import json
    import os
    name = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                        "jet_amd", "data", paths.AMPLITUDE_PATHS)
    with open(name) as f:
        table = json.load(f)
    assert set(table) == {"5x5x14_seed51", "6x6x12_seed52", "6x7x10_seed53",
                          "6x6x10_seed2026_angles"}
    for key, e in table.items():
        dims_txt, seed_txt = key.split("_seed")
        r, c, d = (int(x) for x in dims_txt.split("x"))
        seed = int(seed_txt.split("_")[0])
        circ = circuits.random_grid_circuit(r, c, d, seed=seed,
                                            random_angles=key.endswith("_angles"))
        net = circuits.amplitude_network(circ)
        dims = {l: 2 for ix in net.indices for l in ix}
        path = [tuple(p) for p in e["path"]]
        _valid_ssa(len(net.indices), path)
        st = paths.path_stats(net.indices, dims, path)
        assert st.total_flops == e["pathinfo_flops"] and st.max_log2_size == e["max_log2_size"]
        assert e["modelled_s"] == pytest.approx(paths.modelled_amplitude_seconds(
            net.indices, dims, path), rel=1e-12)
        assert e["modelled_s"] < e["minfill_modelled_s"]
        assert e["max_log2_size"] <= e["minfill_max_log2_size"]
        assert paths.load_or_make_amplitude_path(r, c, d, seed, key.endswith("_angles")) == path

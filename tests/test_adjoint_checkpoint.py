"""Checkpointed recomputation in the adjoint contraction DAG (CPU).

``adjoint_dag_steps`` with ``checkpoint_every`` k > 0 keeps every k-th
forward intermediate and recomputes the others inside the DAG.  Checked here
without a device:

* the DAG's steps, evaluated step by step with the oracle's ContractTensors
  in complex128, give the root value and every fSim environment of a 4x4
  depth-8 circuit equal to ``oracle.environments`` (itself pinned to exact
  statevector derivatives), for k = 0, 1, 3 and 7;
* each step only reads nodes that exist before it, and each dropped forward
  node is recomputed at most once, so the extra flops stay below the forward
  pass's;
* on the 6x6 depth-10 circuit the planned arena (``jet_hip_dag_estimate``,
  the library's own planner) shrinks with checkpointing.
"""
import numpy as np
import pytest

import oracle
from jet_amd import circuits, paths
from jet_amd.gradient import adjoint_dag_steps
from jet_amd.network import dag_estimate


def _circuit(rows, cols, depth, seed):
    circ = circuits.random_grid_circuit(rows, cols, depth, seed=seed, random_angles=True)
    net = circuits.amplitude_network(circ)
    dims = {l: 2 for ix in net.indices for l in ix}
    path = paths.best_minfill_path(net.indices, dims, trials=4, seed=1)
    wanted = [circ.n_qubits + gi for gi, g in enumerate(circ.gates) if g.param_index is not None]
    return net, path, wanted


def _evaluate(net, plan):
    """Run the DAG on the CPU: node values as (labels, shape, complex128)."""
    nodes = [(list(l), list(s), np.asarray(d, np.complex128).reshape(-1))
             for l, s, d in zip(net.indices, net.shapes, net.data)]
    nodes.append(([], [], np.ones(1, np.complex128)))
    for (a, b) in plan["steps"]:
        assert a < len(nodes) and b < len(nodes)
        la, sa, da = nodes[a]
        lb, sb, db = nodes[b]
        nodes.append(oracle.contract(la, sa, da, lb, sb, db))
    return nodes


@pytest.mark.parametrize("k", [0, 1, 3, 7])
def test_checkpointed_dag_matches_oracle_environments(k):
    net, path, wanted = _circuit(4, 4, 8, seed=5)
    plan = adjoint_dag_steps(net.indices, path, wanted, checkpoint_every=k)
    a_ref, env_ref = oracle.environments(list(zip(net.indices, net.shapes, net.data)), path,
                                         wanted)
    nodes = _evaluate(net, plan)
    root = nodes[plan["root_node"]][2]
    assert abs(complex(root[0]) - a_ref) <= 1e-12 * abs(a_ref)
    for w in wanted:
        labels, shape, data = nodes[plan["env_node"][w]]
        want_labels = list(net.indices[w])
        got = data.reshape(shape).transpose([labels.index(l) for l in want_labels]).reshape(-1)
        assert np.allclose(got, env_ref[w], rtol=1e-11, atol=1e-14 * np.abs(env_ref[w]).max())
    n_fwd = len(path)
    if k in (0, 1):
        assert plan["recomputed"] == {}
    else:
        assert plan["recomputed"], "no forward node was recomputed"
        assert len(plan["recomputed"]) < n_fwd
        assert all(len(net.indices) <= u < len(net.indices) + n_fwd for u in plan["recomputed"])


def test_recomputation_bounded_and_arena_smaller():
    net, path, wanted = _circuit(6, 6, 10, seed=2026)
    full = adjoint_dag_steps(net.indices, path, wanted, checkpoint_every=0)
    ckpt = adjoint_dag_steps(net.indices, path, wanted, checkpoint_every=4)
    idx = [list(ix) for ix in net.indices] + [[]]
    shp = [list(s) for s in net.shapes] + [[]]
    keep_full = [full["root_node"]] + [full["env_node"][w] for w in wanted]
    keep_ckpt = [ckpt["root_node"]] + [ckpt["env_node"][w] for w in wanted]
    fl0, ar0, _ = dag_estimate(idx, shp, full["steps"], keep_full)
    fl4, ar4, _ = dag_estimate(idx, shp, ckpt["steps"], keep_ckpt)
    fwd, _, _ = dag_estimate([list(ix) for ix in net.indices], [list(s) for s in net.shapes],
                             path, [len(net.indices) + len(path) - 1])
    assert len(ckpt["steps"]) == len(full["steps"]) + len(ckpt["recomputed"])
    assert fl0 < fl4 <= fl0 + fwd
    assert ar4 < ar0, (ar4, ar0)

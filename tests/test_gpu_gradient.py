"""Adjoint gradients of |<b|C(theta)|0>|^2 on the GPU (jet_amd.gradient).

* 4x4 grid, depth 8: GPU environments (complex64) against the complex128 CPU
  reference gradient (oracle.environments, itself pinned against exact
  statevector derivatives), for every fSim angle, with both methods: the
  captured contraction DAG ("graph") and per-step calls ("eager");
  checkpointed recomputation (eager, and inside the DAG) gives the same
  gradient as keeping every intermediate.
* A triangle network run as a contraction DAG: every kept value against
  numpy.
* 6x6 grid, depth 10, ~300 angles ~ U(0, 2 pi), every 4th forward
  intermediate kept and the others recomputed inside the DAG: the
  torch.autograd.Function gradient against central finite differences (h = 1e-3, complex128 on FP64
  MFMA) for every angle, normwise relative error <= 1e-4, and against the
  CPU reference gradient.
"""
import numpy as np
import pytest

import oracle
from jet_amd import circuits, paths
from jet_amd.gradient import AmplitudeLoss, ParameterizedAmplitude

from hiputil import say

pytestmark = pytest.mark.gpu


def _setup(rows, cols, depth, seed, trials=8):
    circ = circuits.random_grid_circuit(rows, cols, depth, seed=seed, random_angles=True)
    net = circuits.amplitude_network(circ)
    dims = {l: 2 for ix in net.indices for l in ix}
    path = paths.best_minfill_path(net.indices, dims, trials=trials, seed=1)
    return circ, net, path


def _reference_gradient(circ, net, path):
    n0 = circ.n_qubits
    pg = [gi for gi, g in enumerate(circ.gates) if g.param_index is not None]
    a, envs = oracle.environments(list(zip(net.indices, net.shapes, net.data)), path,
                                  [n0 + gi for gi in pg])
    grad = np.zeros(2 * len(pg))
    for gi in pg:
        g = circ.gates[gi]
        th, ph = g.params
        c, s = np.cos(th), np.sin(th)
        d_th = np.zeros((4, 4), np.complex128)
        d_th[1, 1] = d_th[2, 2] = -s
        d_th[1, 2] = d_th[2, 1] = -1j * c
        d_ph = np.zeros((4, 4), np.complex128)
        d_ph[3, 3] = -1j * np.exp(-1j * ph)
        e = envs[n0 + gi]
        grad[g.param_index] = 2 * np.real(np.conj(a) * np.sum(e * d_th.reshape(-1)))
        grad[g.param_index + 1] = 2 * np.real(np.conj(a) * np.sum(e * d_ph.reshape(-1)))
    return a, grad


def _rel(x, y):
    return float(np.linalg.norm(x - y) / np.linalg.norm(y))


def test_gradient_4x4_against_cpu_reference():
    circ, net, path = _setup(4, 4, 8, seed=5)
    say("gradient 4x4 d8")
    a_ref, g_ref = _reference_gradient(circ, net, path)
    prob = ParameterizedAmplitude(circ, path, checkpoint_every=1, method="eager")
    loss, g = prob.loss_and_gradient()
    assert abs(prob.amplitude - a_ref) / abs(a_ref) < 1e-5
    assert abs(loss - abs(a_ref) ** 2) / abs(a_ref) ** 2 < 2e-5
    assert _rel(g, g_ref) < 1e-4, _rel(g, g_ref)

    prob3 = ParameterizedAmplitude(circ, path, checkpoint_every=3, method="eager")
    loss3, g3 = prob3.loss_and_gradient()
    assert prob3.adjoint.recomputed_steps > 0
    assert loss3 == loss and np.array_equal(g3, g)

    say("gradient 4x4 d8, contraction DAG")
    pg = ParameterizedAmplitude(circ, path)
    assert pg.method == "graph" and pg.program.recomputed_steps > 0
    loss_g, g_g = pg.loss_and_gradient()
    assert abs(pg.amplitude - a_ref) / abs(a_ref) < 1e-5
    assert _rel(g_g, g_ref) < 1e-4, _rel(g_g, g_ref)
    assert _rel(g_g, g) < 1e-5, _rel(g_g, g)
    # The DAG without checkpoints: a recomputed node is the same contraction
    # of the same values, so the gradient is the same to the bit.
    say("gradient 4x4 d8, contraction DAG keeping every intermediate")
    p0 = ParameterizedAmplitude(circ, path, checkpoint_every=0)
    assert p0.program.recomputed_steps == 0
    assert p0.program.dag.arena_bytes > pg.program.dag.arena_bytes
    loss_0, g_0 = p0.loss_and_gradient()
    assert loss_0 == loss_g and np.array_equal(g_0, g_g)
    # New angles: one upload, the same captured graph.
    theta = pg.theta + 0.1
    loss_e, g_e = prob.loss_and_gradient(theta)
    loss_g2, g_g2 = pg.loss_and_gradient(theta)
    assert abs(loss_g2 - loss_e) / abs(loss_e) < 2e-5
    assert _rel(g_g2, g_e) < 1e-5, _rel(g_g2, g_e)


def test_contraction_dag_triangle():
    """A(a,b) B(b,c) C(c,a): forward to the scalar, then the environments of
    every leaf; each kept value against numpy (complex128)."""
    from jet_amd import _native as nat
    from jet_amd.network import DeviceDag
    rng = np.random.default_rng(3)
    da, db, dc = 2, 3, 4
    A = (rng.normal(size=(da, db)) + 1j * rng.normal(size=(da, db)))
    B = (rng.normal(size=(db, dc)) + 1j * rng.normal(size=(db, dc)))
    C = (rng.normal(size=(dc, da)) + 1j * rng.normal(size=(dc, da)))
    indices = [["a", "b"], ["b", "c"], ["c", "a"], []]
    shapes = [[da, db], [db, dc], [dc, da], []]
    # nodes: 0 A, 1 B, 2 C, 3 one; 4 = AB, 5 = ABC, 6 = E_AB, 7 = E_C, 8 = E_A, 9 = E_B
    steps = [(0, 1), (4, 2), (3, 2), (3, 4), (6, 1), (6, 0)]
    keep = [5, 8, 9, 7]
    say("contraction DAG triangle")
    dag = DeviceDag(indices, shapes, steps, keep, dtype=np.complex128)
    bufs = []
    for x in (A, B, C, np.ones(1)):
        b = nat.DeviceBuffer(max(16, x.size * 16), dag.ctx)
        b.upload(np.ascontiguousarray(x.reshape(-1), dtype=np.complex128))
        bufs.append(b)
    dag.launch([b.ptr for b in bufs])
    vals = dag.kept_values()
    want = {5: ([], np.einsum("ab,bc,ca->", A, B, C)),
            8: (["a", "b"], np.einsum("bc,ca->ab", B, C)),
            9: (["b", "c"], np.einsum("ab,ca->bc", A, C)),
            7: (["c", "a"], np.einsum("ab,bc->ca", A, B))}
    for node, (labels, ref) in want.items():
        got_labels, got = vals[node]
        assert sorted(got_labels) == sorted(labels), (node, got_labels)
        if labels:
            got = np.transpose(got, [got_labels.index(l) for l in labels])
        assert np.allclose(got, ref, rtol=1e-12, atol=1e-12), node


def test_gradient_6x6_torch_against_finite_differences():
    torch = pytest.importorskip("torch")
    circ, net, path = _setup(6, 6, 10, seed=2026)
    prob = ParameterizedAmplitude(circ, path, checkpoint_every=4)
    assert prob.method == "graph" and prob.program.recomputed_steps > 0
    theta = torch.tensor(prob.theta, dtype=torch.float64, requires_grad=True)
    say(f"gradient 6x6 d10: {prob.n_params} angles, {len(path)} steps, "
        f"{prob.program.recomputed_steps} recomputed in the DAG")
    loss = AmplitudeLoss.apply(theta, prob)
    loss.backward()
    g = theta.grad.numpy().copy()
    assert g.shape == (prob.n_params,) and prob.n_params >= 250

    base = prob.theta.copy()
    l0 = prob.loss_fp64(base)
    assert abs(float(loss) - l0) / l0 < 1e-4
    h = 1e-3
    fd = np.zeros_like(g)
    for k in range(prob.n_params):
        tp, tm = base.copy(), base.copy()
        tp[k] += h
        tm[k] -= h
        fd[k] = (prob.loss_fp64(tp) - prob.loss_fp64(tm)) / (2 * h)
    err_fd = _rel(g, fd)
    say(f"  |g - fd| / |fd| = {err_fd:.2e}")
    assert err_fd <= 1e-4, err_fd

    _, g_ref = _reference_gradient(circ, net, path)
    err_ref = _rel(g, g_ref)
    say(f"  |g - cpu reference| / |cpu reference| = {err_ref:.2e}")
    assert err_ref <= 1e-4, err_ref

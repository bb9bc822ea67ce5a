"""Adjoint gradients of |<b|C(theta)|0>|^2 on the GPU (jet_amd.gradient).

* 4x4 grid, depth 8: GPU environments (complex64) against the complex128 CPU
  reference gradient (oracle.environments, itself pinned against exact
  statevector derivatives), for every fSim angle; checkpointed recomputation
  gives the same gradient as keeping every intermediate.
* 6x6 grid, depth 10, ~300 angles ~ U(0, 2 pi): the torch.autograd.Function
  gradient against central finite differences (h = 1e-3, complex128 on FP64
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
    prob = ParameterizedAmplitude(circ, path, checkpoint_every=1)
    loss, g = prob.loss_and_gradient()
    assert abs(prob.amplitude - a_ref) / abs(a_ref) < 1e-5
    assert abs(loss - abs(a_ref) ** 2) / abs(a_ref) ** 2 < 2e-5
    assert _rel(g, g_ref) < 1e-4, _rel(g, g_ref)

    prob3 = ParameterizedAmplitude(circ, path, checkpoint_every=3)
    loss3, g3 = prob3.loss_and_gradient()
    assert prob3.adjoint.recomputed_steps > 0
    assert loss3 == loss and np.array_equal(g3, g)


def test_gradient_6x6_torch_against_finite_differences():
    torch = pytest.importorskip("torch")
    circ, net, path = _setup(6, 6, 10, seed=2026)
    prob = ParameterizedAmplitude(circ, path, checkpoint_every=4)
    theta = torch.tensor(prob.theta, dtype=torch.float64, requires_grad=True)
    say(f"gradient 6x6 d10: {prob.n_params} angles, {len(path)} steps")
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

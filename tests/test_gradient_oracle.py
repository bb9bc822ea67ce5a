"""The CPU reference gradient (oracle.environments) against closed forms.

The derivative of an amplitude with respect to one gate entry is exact in a
statevector simulation: A is linear in each gate, so dA/dtheta is the
amplitude of the circuit with that gate replaced by dG/dtheta.  Both the
environments (dA/dG for every fSim gate) and the angle gradients of
|A|^2 are checked against it on a 3x3 grid, and against central differences.
"""
import numpy as np
import pytest

import oracle
from jet_amd import circuits, paths


def _problem(rows=3, cols=3, depth=6, seed=11):
    circ = circuits.random_grid_circuit(rows, cols, depth, seed=seed, random_angles=True)
    net = circuits.amplitude_network(circ)
    dims = {l: 2 for ix in net.indices for l in ix}
    path = paths.best_minfill_path(net.indices, dims, trials=4, seed=3)
    return circ, net, path


def _fsim_derivatives(theta, phi):
    c, s = np.cos(theta), np.sin(theta)
    d_th = np.zeros((4, 4), np.complex128)
    d_th[1, 1] = d_th[2, 2] = -s
    d_th[1, 2] = d_th[2, 1] = -1j * c
    d_ph = np.zeros((4, 4), np.complex128)
    d_ph[3, 3] = -1j * np.exp(-1j * phi)
    return d_th, d_ph


def test_environments_match_statevector_derivatives():
    circ, net, path = _problem()
    n0 = circ.n_qubits
    pgates = [gi for gi, g in enumerate(circ.gates) if g.param_index is not None]
    wanted = [n0 + gi for gi in pgates]
    a, envs = oracle.environments(list(zip(net.indices, net.shapes, net.data)), path, wanted)
    assert abs(a - circuits.amplitude_statevector(circ)) < 1e-12
    for gi in pgates[::3]:
        g = circ.gates[gi]
        e = envs[n0 + gi]
        for m in _fsim_derivatives(*g.params):
            saved = g.matrix
            g.matrix = m
            want = circuits.amplitude_statevector(circ)
            g.matrix = saved
            got = np.sum(e * m.reshape(-1))
            assert abs(got - want) < 1e-12, (gi, got, want)


def test_loss_gradient_matches_central_differences():
    circ, net, path = _problem(seed=12)
    n0 = circ.n_qubits
    pgates = [gi for gi, g in enumerate(circ.gates) if g.param_index is not None]
    a, envs = oracle.environments(list(zip(net.indices, net.shapes, net.data)), path,
                                  [n0 + gi for gi in pgates])
    h = 1e-5
    for gi in pgates[:6]:
        g = circ.gates[gi]
        th, ph = g.params
        d_th, d_ph = _fsim_derivatives(th, ph)
        for which, dm in ((0, d_th), (1, d_ph)):
            grad = 2 * np.real(np.conj(a) * np.sum(envs[n0 + gi] * dm.reshape(-1)))
            vals = []
            for sgn in (1, -1):
                p = [th, ph]
                p[which] += sgn * h
                g.matrix = circuits.fsim(*p)
                vals.append(abs(circuits.amplitude_statevector(circ)) ** 2)
            g.matrix = circuits.fsim(th, ph)
            fd = (vals[0] - vals[1]) / (2 * h)
            assert abs(grad - fd) < 1e-8 * max(1.0, abs(fd)), (gi, which, grad, fd)

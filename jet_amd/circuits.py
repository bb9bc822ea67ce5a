"""Seeded grid random circuits as amplitude tensor networks.

The networks follow the reference's circuit conventions: the wire label of
qubit q after d operations is "q-d" (python/jet/circuit.py:46-49), a gate's
tensor has indices outputs ++ inputs (circuit.py:111-119), qubits start in
|0> (python/jet/state.py:121-124).  An amplitude <b|C|0...0> is the full
contraction of the network closed by <b_q| on every output wire.

Gate set (the reference has no sqrt(Y), sqrt(W) or fSim, python/jet/gate.py:
20-61, so they are defined here from their textbook matrices):
  sqrt(X), sqrt(Y), sqrt(W) with W = (X + Y)/sqrt(2) -- principal square roots
  of involutions, sqrt(U) = (1+i)/2 I + (1-i)/2 U;
  fSim(theta, phi) = [[1,0,0,0],[0,cos t,-i sin t,0],[0,-i sin t,cos t,0],
                      [0,0,0,exp(-i phi)]].
Each cycle applies a single-qubit gate drawn uniformly from {sqrt X, sqrt Y,
sqrt W} to every qubit, then fSim on one of four alternating nearest-
neighbour edge patterns (horizontal even, horizontal odd, vertical even,
vertical odd columns/rows, in that rotation).
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

_I2 = np.eye(2, dtype=np.complex128)
_X = np.array([[0, 1], [1, 0]], dtype=np.complex128)
_Y = np.array([[0, -1j], [1j, 0]], dtype=np.complex128)
_W = (_X + _Y) / np.sqrt(2.0)


def sqrt_involution(u: np.ndarray) -> np.ndarray:
    """Principal square root of a unitary with u @ u == I."""
    return 0.5 * (1 + 1j) * _I2 + 0.5 * (1 - 1j) * u


SQRT_X = sqrt_involution(_X)
SQRT_Y = sqrt_involution(_Y)
SQRT_W = sqrt_involution(_W)
SINGLE_QUBIT_GATES = (("sqrt_x", SQRT_X), ("sqrt_y", SQRT_Y), ("sqrt_w", SQRT_W))


def fsim(theta: float, phi: float) -> np.ndarray:
    """fSim(theta, phi) in the basis |00>, |01>, |10>, |11>."""
    c, s = np.cos(theta), np.sin(theta)
    return np.array([[1, 0, 0, 0], [0, c, -1j * s, 0], [0, -1j * s, c, 0],
                     [0, 0, 0, np.exp(-1j * phi)]], dtype=np.complex128)


def grid_edges(rows: int, cols: int) -> List[List[Tuple[int, int]]]:
    """The four nearest-neighbour edge patterns of a rows x cols grid."""
    q = lambda r, c: r * cols + c
    pats: List[List[Tuple[int, int]]] = [[], [], [], []]
    for r in range(rows):
        for c in range(cols - 1):
            pats[c % 2].append((q(r, c), q(r, c + 1)))
    for r in range(rows - 1):
        for c in range(cols):
            pats[2 + r % 2].append((q(r, c), q(r + 1, c)))
    return pats


@dataclass
class Gate:
    name: str
    qubits: Tuple[int, ...]
    matrix: np.ndarray          # 2^k x 2^k, rows = outputs
    params: Tuple[float, ...] = ()
    param_index: Optional[int] = None  # first parameter slot (fSim gradients)


@dataclass
class Circuit:
    rows: int
    cols: int
    depth: int
    gates: List[Gate] = field(default_factory=list)
    bitstring: Tuple[int, ...] = ()

    @property
    def n_qubits(self) -> int:
        return self.rows * self.cols


def random_grid_circuit(rows: int, cols: int, depth: int, seed: int,
                        theta: float = np.pi / 2, phi: float = np.pi / 6,
                        random_angles: bool = False) -> Circuit:
    """Seeded random circuit; with random_angles every fSim gets its own
    (theta_k, phi_k) ~ U(0, 2 pi) (parameter slots 2k, 2k+1)."""
    rng = np.random.default_rng(seed)
    n = rows * cols
    pats = grid_edges(rows, cols)
    circ = Circuit(rows, cols, depth)
    n_params = 0
    for t in range(depth):
        choice = rng.integers(0, 3, size=n)
        for qb in range(n):
            name, m = SINGLE_QUBIT_GATES[int(choice[qb])]
            circ.gates.append(Gate(name, (qb,), m))
        for (a, b) in pats[t % 4]:
            if random_angles:
                th, ph = rng.uniform(0, 2 * np.pi, size=2)
                circ.gates.append(Gate("fsim", (a, b), fsim(th, ph), (float(th), float(ph)),
                                       n_params))
                n_params += 2
            else:
                circ.gates.append(Gate("fsim", (a, b), fsim(theta, phi), (theta, phi)))
    circ.bitstring = tuple(int(x) for x in rng.integers(0, 2, size=n))
    return circ


@dataclass
class NetworkSpec:
    """A tensor network as plain data: per tensor its indices, shape and
    row-major data (complex128), in the order they are added."""
    indices: List[List[str]]
    shapes: List[List[int]]
    data: List[np.ndarray]
    tags: List[List[str]]

    def __len__(self) -> int:
        return len(self.indices)


def amplitude_network(circ: Circuit, bitstring: Optional[Sequence[int]] = None,
                      dtype=np.complex128) -> NetworkSpec:
    """<b| C |0...0> as a closed tensor network (Jet wire conventions)."""
    b = tuple(circ.bitstring if bitstring is None else bitstring)
    depth = [0] * circ.n_qubits
    lab = lambda q: f"{q}-{depth[q]}"
    idx, shp, dat, tags = [], [], [], []
    for q in range(circ.n_qubits):
        idx.append([lab(q)])
        shp.append([2])
        dat.append(np.array([1, 0], dtype=dtype))
        tags.append(["state", f"q{q}"])
    for gi, g in enumerate(circ.gates):
        ins = [lab(q) for q in g.qubits]
        for q in g.qubits:
            depth[q] += 1
        outs = [lab(q) for q in g.qubits]
        k = len(g.qubits)
        idx.append(outs + ins)
        shp.append([2] * (2 * k))
        dat.append(np.asarray(g.matrix, dtype=dtype).reshape(-1))
        tags.append([g.name, f"gate{gi}"])
    for q in range(circ.n_qubits):
        idx.append([lab(q)])
        shp.append([2])
        v = np.zeros(2, dtype=dtype)
        v[b[q]] = 1
        dat.append(v)
        tags.append(["output", f"q{q}"])
    return NetworkSpec(idx, shp, dat, tags)


def statevector(circ: Circuit) -> np.ndarray:
    """C|0...0> in complex128, qubit 0 most significant (closed form for the
    small grids)."""
    n = circ.n_qubits
    psi = np.zeros((2,) * n, dtype=np.complex128)
    psi[(0,) * n] = 1
    for g in circ.gates:
        k = len(g.qubits)
        u = g.matrix.reshape((2,) * (2 * k))
        # contract gate inputs with the state's axes, then move outputs back
        psi = np.tensordot(u, psi, axes=(list(range(k, 2 * k)), list(g.qubits)))
        psi = np.moveaxis(psi, list(range(k)), list(g.qubits))
    return psi


def amplitude_statevector(circ: Circuit, bitstring: Optional[Sequence[int]] = None) -> complex:
    b = tuple(circ.bitstring if bitstring is None else bitstring)
    return complex(statevector(circ)[b])

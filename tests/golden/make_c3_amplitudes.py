"""Complex128 CPU amplitudes of the uneven batch's circuits.

The batch holds 24 amplitudes of a 5x5-qubit depth-14, 16 of a 6x6 depth-12
and 24 of a 6x7 depth-10 seeded grid random circuit
(jet_amd.circuits.random_grid_circuit with seeds 51, 52, 53); the bitstrings
come from numpy.random.default_rng(2026), drawn circuit after circuit, r*c
bits per amplitude.  Values, all in complex128 from the unrounded gates:

  * 5x5: the statevector (closed form), plus the oracle's tensor-network
    contraction of the first bitstring as a cross-check;
  * 6x6 and 6x7: the reference CPU algorithm (Tensor::ContractTensors with
    cblas_zgemm, Tensor.hpp:708-752, TensorHelpers.hpp:47-61, restated in
    oracle.baseline_contract) along a min-fill path; the subtrees that hold
    no output vector are contracted once for all bitstrings.  The first
    bitstring is also contracted by oracle.contract (the oracle's own FP64
    GEMM) and must agree to 1e-12.

Writes tests/golden/c3_amplitudes_fp64.json.  About 15 minutes on 8 cores.

Usage: python tests/golden/make_c3_amplitudes.py
"""
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from jet_amd import circuits, paths  # noqa: E402

OUT = os.path.join(HERE, "c3_amplitudes_fp64.json")
CIRCUITS = [((5, 5, 14), 24, 51), ((6, 6, 12), 16, 52), ((6, 7, 10), 24, 53)]
THREADS = os.cpu_count() or 1


def _blas_contract(la, sa, a, lb, sb, b):
    dims = dict(zip(la, sa))
    dims.update(zip(lb, sb))
    labels, c = oracle.baseline_contract(la, sa, a, lb, sb, b, THREADS)
    return labels, [dims[l] for l in labels], c


def _oracle_contract(la, sa, a, lb, sb, b):
    return oracle.contract(la, sa, a, lb, sb, b, THREADS)


def amplitudes(circ, bitstrings, path, contract, log):
    """<b|C|0> for every b, sharing the subtrees that hold no output vector."""
    nq = circ.n_qubits
    base = circuits.amplitude_network(circ, bitstrings[0], dtype=np.complex128)
    n = len(base.indices)
    dep = [i >= n - nq for i in range(n)]
    for (a, b) in path:
        dep.append(dep[a] or dep[b])
    frontier = {c for k, (a, b) in enumerate(path) if dep[n + k] for c in (a, b) if not dep[c]}
    leaves = [(list(ix), list(sh), np.asarray(d, np.complex128))
              for ix, sh, d in zip(base.indices, base.shapes, base.data)]
    uses = [0] * (n + len(path))
    for k, (a, b) in enumerate(path):
        if not dep[n + k]:
            uses[a] += 1
            uses[b] += 1
    nodes = {i: leaves[i] for i in range(n) if not dep[i]}
    shared = {c: leaves[c] for c in frontier if c < n}
    for k, (a, b) in enumerate(path):
        node = n + k
        if dep[node]:
            continue
        nodes[node] = contract(*nodes[a], *nodes[b])
        if node in frontier:
            shared[node] = nodes[node]
        for c in (a, b):
            uses[c] -= 1
            if uses[c] == 0 and c >= n:
                nodes.pop(c, None)
    del nodes
    uses = [0] * (n + len(path))
    for k, (a, b) in enumerate(path):
        if dep[n + k]:
            uses[a] += 1
            uses[b] += 1
    out = []
    for bits in bitstrings:
        t0 = time.time()
        nodes = dict(shared)
        for q in range(nq):
            v = np.zeros(2, np.complex128)
            v[bits[q]] = 1
            nodes[n - nq + q] = (leaves[n - nq + q][0], [2], v)
        left = list(uses)
        for k, (a, b) in enumerate(path):
            node = n + k
            if not dep[node]:
                continue
            nodes[node] = contract(*nodes[a], *nodes[b])
            for c in (a, b):
                left[c] -= 1
                if left[c] == 0 and c not in shared:
                    nodes.pop(c, None)
        _, _, v = nodes[n + len(path) - 1]
        out.append(complex(v[0]))
        log(f"  {bits[:8]}...: {out[-1]!r} ({time.time() - t0:.0f} s)")
    return out


def main():
    log = lambda m: print(m, flush=True)  # noqa: E731
    oracle.load_blas(THREADS)
    rng = np.random.default_rng(2026)
    doc = {"bitstring_seed": 2026, "circuits": []}
    for (r, c, d), n_amp, seed in CIRCUITS:
        circ = circuits.random_grid_circuit(r, c, d, seed=seed)
        bits = [tuple(int(x) for x in rng.integers(0, 2, r * c)) for _ in range(n_amp)]
        net = circuits.amplitude_network(circ)
        dims = {l: 2 for ix in net.indices for l in ix}
        path = paths.best_minfill_path(net.indices, dims, trials=64, seed=1)
        log(f"{r}x{c} depth {d} (seed {seed}): {n_amp} amplitudes")
        check = amplitudes(circ, bits[:1], path, _oracle_contract, log)[0]
        if r * c <= 25:
            psi = circuits.statevector(circ)
            vals = [complex(psi[b]) for b in bits]
            del psi
        else:
            vals = amplitudes(circ, bits, path, _blas_contract, log)
        rel = abs(vals[0] - check) / abs(check)
        log(f"  cross-check of the first amplitude: {rel:.2e}")
        if not rel < 1e-12:
            raise SystemExit("the two CPU contractions disagree; nothing written")
        doc["circuits"].append({"rows": r, "cols": c, "depth": d, "seed": seed,
                                "bitstrings": ["".join(map(str, b)) for b in bits],
                                "amplitudes": [[v.real, v.imag] for v in vals],
                                "first_by_oracle_contract": [check.real, check.imag]})
    doc["made_by"] = "tests/golden/make_c3_amplitudes.py"
    with open(OUT, "w") as f:
        json.dump(doc, f, indent=1)
    log(f"wrote {OUT}")


if __name__ == "__main__":
    main()

"""Per-slice complex128 CPU values of the sliced 7x7 depth-14 amplitude.

The network is one amplitude <b|C|0> of the seeded 7x7-qubit grid random
circuit of depth 14 (jet_amd.circuits.random_grid_circuit(7, 7, 14,
seed=2026)), sliced on the 12 indices of the committed plan
(jet_amd/data/grid_7x7x14_seed2026_s12_refined.json).  Slice s fixes the
sliced indices to UnravelIndex(s, [2]*12), first sliced index most
significant (Utilities.hpp:463-475, TensorNetwork.hpp:230).  Each slice is
contracted on the CPU by the oracle (oracle.slice_index, Tensor.hpp:493-526,
then oracle.contract, Tensor.hpp:708-752, FP64 accumulation) along the
plan's path.  Subtrees that hold no sliced index are the same in every slice
(the tasks TaskBasedContractor shares by name, TaskBasedContractor.hpp:
218-222): they are contracted once and reused.

Writes tests/golden/c2_slices_fp64.json.  About two minutes of CPU per slice
on 8 cores.  The leaves are the complex64 gate and vector data (widened
to complex128 for the contraction).

Usage: python tests/golden/make_c2_slices.py [slice ...]
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
from jet_amd import circuits  # noqa: E402

PLAN = os.path.join(ROOT, "jet_amd", "data", "grid_7x7x14_seed2026_s12_refined.json")
OUT = os.path.join(HERE, "c2_slices_fp64.json")
# Slices 0-7 lie in the first batched execution; the others need loop
# indices fixed (the last one is slice 4095).
DEFAULT_SLICES = list(range(8)) + [300, 1337, 2600, 4095]


def slice_values(net, path, sliced, slice_ids, log=print):
    n = len(net.indices)
    sl = set(sliced)
    dep = [any(l in sl for l in ix) for ix in net.indices]
    for (a, b) in path:
        dep.append(dep[a] or dep[b])
    # Slice-independent nodes read by slice-dependent steps.
    frontier = {c for k, (a, b) in enumerate(path) if dep[n + k]
                for c in (a, b) if not dep[c]}
    leaves = [(list(ix), list(sh), np.asarray(d, np.complex128))
              for ix, sh, d in zip(net.indices, net.shapes, net.data)]

    # Pre-pass: every slice-independent step once.
    t0 = time.time()
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
        la, sa, da = nodes[a]
        lb, sb, db = nodes[b]
        nodes[node] = oracle.contract(la, sa, da, lb, sb, db)
        if node in frontier:
            shared[node] = nodes[node]
        for c in (a, b):
            uses[c] -= 1
            if uses[c] == 0 and c >= n:
                nodes.pop(c, None)
    del nodes
    log(f"slice-independent steps: {time.time() - t0:.0f} s, {len(shared)} kept")

    uses = [0] * (n + len(path))
    for k, (a, b) in enumerate(path):
        if dep[n + k]:
            uses[a] += 1
            uses[b] += 1
    out = []
    for s in slice_ids:
        t0 = time.time()
        vals = oracle.unravel_index(s, [2] * len(sliced))
        fixed = dict(zip(sliced, vals))
        nodes = dict(shared)
        for i in range(n):
            if not dep[i]:
                continue
            lab, shp, arr = leaves[i]
            for l in list(lab):
                if l in fixed:
                    lab, shp, arr = oracle.slice_index(lab, shp, arr, l, fixed[l])
            nodes[i] = (lab, shp, arr)
        left = list(uses)
        for k, (a, b) in enumerate(path):
            node = n + k
            if not dep[node]:
                continue
            la, sa, da = nodes[a]
            lb, sb, db = nodes[b]
            nodes[node] = oracle.contract(la, sa, da, lb, sb, db)
            for c in (a, b):
                left[c] -= 1
                if left[c] == 0 and c not in shared:
                    nodes.pop(c, None)
        _, _, v = nodes[n + len(path) - 1]
        out.append(complex(v[0]))
        log(f"slice {s}: {out[-1]!r} ({time.time() - t0:.0f} s)")
    return out


def main():
    ids = [int(x) for x in sys.argv[1:]] or DEFAULT_SLICES
    with open(PLAN) as f:
        plan = json.load(f)
    path = [tuple(p) for p in plan["path"]]
    circ = circuits.random_grid_circuit(7, 7, 14, seed=2026)
    # Leaf data rounded to complex64: the inputs every GPU run starts from
    # (complex128 runs widen them), so FP64 runs can be held to 1e-10.
    net = circuits.amplitude_network(circ, dtype=np.complex64)
    vals = slice_values(net, path, plan["sliced"], ids, log=lambda m: print(m, flush=True))
    doc = {"circuit": {"rows": 7, "cols": 7, "depth": 14, "seed": 2026},
           "plan": os.path.relpath(PLAN, ROOT),
           "sliced": plan["sliced"],
           "slices": ids,
           "values": [[v.real, v.imag] for v in vals],
           "made_by": "tests/golden/make_c2_slices.py (oracle.contract in complex128)"}
    with open(OUT, "w") as f:
        json.dump(doc, f, indent=1)
    print("wrote", OUT)


if __name__ == "__main__":
    main()

"""The complex128 amplitude of the sliced 7x7 depth-14 network (all 4096 slices).

Needs the GPU: every slice is contracted by the library in complex128 on
FP64 MFMA along the committed plan (jet_amd/data/
grid_7x7x14_seed2026_s12_refined.json), 64 slices per execution, and the
per-slice results are summed with FP64 compensated summation.  Before the
amplitude is written, the slices that tests/golden/c2_slices_fp64.json holds
(CPU oracle, complex128) are compared with the GPU's: they must agree to
1e-10 relative, or nothing is written.

Writes tests/golden/c2_amplitude_fp64.json (the amplitude, plus the
compensated sum of the CPU-checked slices' GPU values as a cross-check).
About two minutes on one MI355X.

Usage: python tests/golden/make_c2_amplitude.py [--out PATH]
"""
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from jet_amd import _native as nat, circuits  # noqa: E402
from jet_amd.network import SlicedContraction  # noqa: E402

OUT = os.path.join(HERE, "c2_amplitude_fp64.json")


def main():
    out = OUT
    if "--out" in sys.argv:
        out = sys.argv[sys.argv.index("--out") + 1]
    with open(os.path.join(HERE, "c2_slices_fp64.json")) as f:
        fixture = json.load(f)
    with open(os.path.join(ROOT, fixture["plan"])) as f:
        plan = json.load(f)
    circ = circuits.random_grid_circuit(7, 7, 14, seed=2026)
    net = circuits.amplitude_network(circ, dtype=np.complex64)
    data = [d.astype(np.complex128) for d in net.data]
    sliced = plan["sliced"]
    ctx = nat.default_context()
    sc = SlicedContraction(net.indices, net.shapes, data, [tuple(p) for p in plan["path"]],
                           sliced, batch=sliced[-6:], dtype=np.complex128, precision="fp64",
                           ctx=ctx, use_graph=False)
    t0 = time.time()
    for i in range(sc.n_loop):
        sc.run([i])
        ctx.synchronize()
        if i % 8 == 7:
            print(f"{i + 1}/{sc.n_loop} executions ({time.time() - t0:.0f} s)", flush=True)
    res = sc.slice_results()
    want = np.array([complex(*v) for v in fixture["values"]])
    got = res[np.array(fixture["slices"])]
    err = float(np.linalg.norm(got - want) / np.linalg.norm(want))
    print(f"GPU FP64 vs CPU oracle on slices {fixture['slices']}: {err:.2e}", flush=True)
    if not err < 1e-10:
        raise SystemExit("GPU FP64 slices disagree with the CPU oracle; nothing written")
    amp = sc.reduce()
    doc = {"circuit": fixture["circuit"], "plan": fixture["plan"], "slices": int(res.size),
           "amplitude": [amp.real, amp.imag],
           "checked_slices_rel_err": err,
           "made_by": "tests/golden/make_c2_amplitude.py (complex128, FP64 MFMA, "
                      "compensated sum over all slices)"}
    with open(out, "w") as f:
        json.dump(doc, f, indent=1)
    print(f"amplitude {amp!r}; wrote {out}", flush=True)


if __name__ == "__main__":
    main()

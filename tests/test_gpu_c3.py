"""The uneven batch at its configuration, against CPU fixtures.

One hipGraph launch (jet_amd.batch.UnevenBatch) holds 24 amplitudes of a
5x5 depth-14, 16 of a 6x6 depth-12 and 24 of a 6x7 depth-10 seeded grid
circuit (forms chosen by the batch's own timing) plus 256 slices of the 7x7
depth-14 amplitude (the min-fill plan, 128 slices per execution).  Every
amplitude is compared with tests/golden/c3_amplitudes_fp64.json (CPU,
complex128; tests/golden/make_c3_amplitudes.py) and the slices with
tests/golden/c2_slices_fp64.json (slices 0-7).  Tolerances: per amplitude
1e-5 relative in split-FP16 (complex64), 1e-10 in complex128 on FP64 MFMA.
"""
import json
import os

import numpy as np
import pytest

from jet_amd import circuits, paths
from jet_amd.batch import AmplitudeJob, SliceJob, UnevenBatch

from hiputil import say

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")


def _jobs(dtype):
    with open(os.path.join(GOLDEN, "c3_amplitudes_fp64.json")) as f:
        fx = json.load(f)
    jobs, want = [], []
    for c in fx["circuits"]:
        circ = circuits.random_grid_circuit(c["rows"], c["cols"], c["depth"], seed=c["seed"])
        net = circuits.amplitude_network(circ)
        dims = {l: 2 for ix in net.indices for l in ix}
        path = paths.best_minfill_path(net.indices, dims, trials=8, seed=1)
        bits = [tuple(int(x) for x in b) for b in c["bitstrings"]]
        jobs.append(AmplitudeJob(circ, bits, path))
        want.append(np.array([complex(*v) for v in c["amplitudes"]]))
    with open(os.path.join(ROOT, "jet_amd", "data", "grid_7x7x14_seed2026_s12.json")) as f:
        plan = json.load(f)
    circ = circuits.random_grid_circuit(7, 7, 14, seed=2026)
    net = circuits.amplitude_network(circ, dtype=np.complex64)
    sliced = plan["sliced"]
    jobs.append(SliceJob(net.indices, net.shapes, [d.astype(dtype) for d in net.data],
                         [tuple(p) for p in plan["path"]], sliced, sliced[-7:], 0, 2))
    with open(os.path.join(GOLDEN, "c2_slices_fp64.json")) as f:
        c2 = json.load(f)
    first = [i for i, s in enumerate(c2["slices"]) if s < 256]
    slices = (np.array([c2["slices"][i] for i in first]),
              np.array([complex(*c2["values"][i]) for i in first]))
    return jobs, want, slices


@pytest.mark.parametrize("prec,dtype,tol", [("f16x3", np.complex64, 1e-5),
                                            ("fp64", np.complex128, 1e-10)])
def test_uneven_batch_matches_cpu_fixtures(ctx, prec, dtype, tol):
    jobs, want, (s_idx, s_want) = _jobs(dtype)
    say(f"c3 uneven batch {prec}: {sum(len(w) for w in want)} amplitudes + 256 c2 slices")
    ub = UnevenBatch(jobs, dtype=dtype, precision=prec, ctx=ctx)
    say(f"  forms {ub.forms()}")
    ub.launch()
    ctx.synchronize()
    worst = 0.0
    for j, w in enumerate(want):
        got = ub.amplitudes(j).astype(np.complex128)
        rel = np.abs(got - w) / np.abs(w)
        worst = max(worst, float(rel.max()))
        assert rel.max() < tol, (j, rel.max())
    got_s = ub.slice_results(len(want)).astype(np.complex128)[s_idx]
    err_s = np.linalg.norm(got_s - s_want) / np.linalg.norm(s_want)
    ub.close()
    say(f"  worst amplitude error {worst:.2e}, c2 slices {err_s:.2e}")
    assert err_s < tol


def test_uneven_batch_reconfigured_paths(ctx):
    """The amplitude jobs along the committed reconfigured paths
    (jet_amd/data/amplitude_paths.json, the paths the benchmark uses) in
    split-FP16, against the same CPU fixtures: normwise error per circuit
    below 1e-5, and every amplitude within 2e-5 relative (the largest
    relative errors fall on the smallest amplitudes: FP32 MFMA on the same
    batch reaches 1.2e-5 there)."""
    with open(os.path.join(GOLDEN, "c3_amplitudes_fp64.json")) as f:
        fx = json.load(f)
    jobs, want = [], []
    for c in fx["circuits"]:
        circ = circuits.random_grid_circuit(c["rows"], c["cols"], c["depth"], seed=c["seed"])
        path = paths.load_or_make_amplitude_path(c["rows"], c["cols"], c["depth"], c["seed"])
        bits = [tuple(int(x) for x in b) for b in c["bitstrings"]]
        jobs.append(AmplitudeJob(circ, bits, path))
        want.append(np.array([complex(*v) for v in c["amplitudes"]]))
    say(f"c3 amplitudes along reconfigured paths: {sum(len(w) for w in want)} amplitudes")
    ub = UnevenBatch(jobs, dtype=np.complex64, precision="f16x3", ctx=ctx)
    ub.launch()
    ctx.synchronize()
    for j, w in enumerate(want):
        got = ub.amplitudes(j).astype(np.complex128)
        norm = np.linalg.norm(got - w) / np.linalg.norm(w)
        rel = np.abs(got - w) / np.abs(w)
        say(f"  circuit {j}: normwise {norm:.2e}, worst amplitude {rel.max():.2e}")
        assert norm < 1e-5, (j, norm)
        assert rel.max() < 2e-5, (j, rel.max())
    ub.close()

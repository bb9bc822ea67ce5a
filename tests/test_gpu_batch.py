"""Uneven batches (jet_amd.batch): amplitudes of different circuits and a
slice range of a sliced network captured as one hipGraph and launched once.

Every amplitude is checked against the complex128 oracle contraction of the
same network (oracle.contract_path); the complex128 batch (FP64 MFMA) to
1e-10, the complex64 batches (FP32 MFMA, split-FP16) to 1e-5 per amplitude.
The slice job's per-slice results and their FP64 compensated sum are checked
against the oracle's sliced contractions.
"""
import numpy as np
import pytest

import oracle
from jet_amd import circuits, paths
from jet_amd.batch import AmplitudeJob, SliceJob, UnevenBatch

from hiputil import say

pytestmark = pytest.mark.gpu


def _path(circ, trials=4):
    net = circuits.amplitude_network(circ)
    dims = {l: 2 for ix in net.indices for l in ix}
    return net, paths.best_minfill_path(net.indices, dims, trials=trials, seed=1)


def _jobs():
    rng = np.random.default_rng(77)
    jobs, refs = [], []
    for (r, c, d, nb, seed) in [(4, 4, 8, 5, 1), (4, 5, 10, 3, 2), (5, 5, 10, 4, 3)]:
        circ = circuits.random_grid_circuit(r, c, d, seed=seed)
        _, path = _path(circ)
        bits = [tuple(int(x) for x in rng.integers(0, 2, r * c)) for _ in range(nb)]
        jobs.append(AmplitudeJob(circ, bits, path))
        want = []
        for b in bits:
            net = circuits.amplitude_network(circ, b)
            want.append(oracle.contract_path(list(zip(net.indices, net.shapes, net.data)),
                                             path)[2][0])
        refs.append(np.array(want))
    # slices 8..15 of a 4x4 depth-8 amplitude sliced on 5 indices (2 batched)
    circ = circuits.random_grid_circuit(4, 4, 8, seed=9)
    net, path = _path(circ)
    dims = {l: 2 for ix in net.indices for l in ix}
    sliced = paths.choose_slices(net.indices, dims, path, 5)
    sj = SliceJob(net.indices, net.shapes, net.data, path, sliced, sliced[-2:],
                  first_execution=2, executions=2)
    jobs.append(sj)
    return jobs, refs, (net, path, sliced)


def _slice_reference(net, path, sliced, first, count):
    """Per-slice values in the slice order loop indices first, then batch."""
    order = sliced[:-2] + sliced[-2:]
    out = []
    for v in range(first, first + count):
        vals = oracle.unravel_index(v, [2] * len(order))
        fixed = dict(zip(order, vals))
        ts = []
        for ix, sh, d in zip(net.indices, net.shapes, net.data):
            lab, shp, dat = list(ix), list(sh), d
            for l in [x for x in ix if x in fixed]:
                lab, shp, dat = oracle.slice_index(lab, shp, dat, l, fixed[l])
            ts.append((lab, shp, dat))
        out.append(oracle.contract_path(ts, path)[2][0])
    return np.array(out)


@pytest.mark.parametrize("dtype,precision,tol", [(np.complex128, "fp64", 1e-10),
                                                 (np.complex64, None, 1e-5),
                                                 (np.complex64, "f16x3", 1e-5)])
def test_uneven_batch_one_launch(dtype, precision, tol):
    jobs, refs, (net, path, sliced) = _jobs()
    say(f"uneven batch {np.dtype(dtype)} {precision}")
    ub = UnevenBatch(jobs, dtype=dtype, precision=precision)
    say(f"  forms {ub.forms()} tuning {ub.tuning}")
    ub.launch()
    for j, want in enumerate(refs):
        got = ub.amplitudes(j)
        err = np.abs(got - want) / np.abs(want)
        assert err.max() < tol, (j, err)
    want_s = _slice_reference(net, path, sliced, 8, 8)
    got_s = ub.slice_results(3)
    assert np.abs(got_s - want_s).max() / np.abs(want_s).max() < tol
    assert abs(ub.slice_sum(3) - want_s.sum()) / np.abs(want_s).max() < tol
    # a second launch gives the same values (the graph is replayable)
    first = [ub.amplitudes(j).copy() for j in range(3)]
    ub.launch()
    for j in range(3):
        assert np.array_equal(ub.amplitudes(j), first[j])


def test_both_amplitude_forms_agree():
    """Every amplitude form (one execution per bitstring, the stacked batch,
    the stack padded to a power of two) gives the same amplitudes."""
    jobs, refs, _ = _jobs()
    for form in ("each", "stacked", "stacked_pow2"):
        js = [AmplitudeJob(j.circuit, j.bitstrings, j.path, form) for j in jobs[:3]]
        ub = UnevenBatch(js, dtype=np.complex128, precision="fp64")
        assert ub.forms() == [form] * 3
        ub.launch()
        for j, want in enumerate(refs):
            assert np.abs(ub.amplitudes(j) - want).max() / np.abs(want).max() < 1e-10

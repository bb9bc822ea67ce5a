"""The slice loop's iteration rule (CPU) and its device regression (GPU).

Loop iteration i of a sliced contraction runs loop-value combination
execution_slot(i, n_loop) = i mod n_loop and writes its n_batch per-slice
results into that slot, so any number of iterations stays inside the
n_loop * n_batch results vector (the reference keeps one results[i] slot
per slice, TaskBasedContractor.hpp:186-197).
"""
import numpy as np
import pytest

from jet_amd.abort import JetError
from jet_amd.network import execution_slot


def test_execution_slot_wraps():
    assert [execution_slot(i, 16) for i in range(40)] == [i % 16 for i in range(40)]
    assert execution_slot(0, 1) == 0 and execution_slot(7, 1) == 0


def test_execution_slot_rejects_bad_arguments():
    with pytest.raises(JetError, match="negative"):
        execution_slot(-1, 16)
    with pytest.raises(JetError, match="at least one"):
        execution_slot(0, 0)


def test_batch_indices_must_be_sliced():
    """Batch modes are a subset of the sliced indices; the check raises
    before any device work."""
    from jet_amd.network import SlicedContraction
    with pytest.raises(JetError, match="sliced indices"):
        SlicedContraction([["a", "b"], ["b"]], [[2, 2], [2]], [None, None], [1, 0],
                          sliced=["b"], batch=["a"])


def test_execution_slot_writes_stay_in_bounds():
    """Every byte an iteration writes lies in the results vector, for 5
    warm-up + 20 timed iterations (bench.py --warmup 5 --steps 20) on a
    16-iteration loop."""
    n_loop, n_batch, item = 16, 256, 8
    size = n_loop * n_batch * item
    for i in range(25):
        lo = execution_slot(i, n_loop) * n_batch * item
        assert 0 <= lo and lo + n_batch * item <= size


@pytest.mark.gpu
def test_more_iterations_than_loop_values(ctx):
    """More iterations than loop-value combinations (eager, replayed graph
    and one captured batch): results equal the single pass over the loop,
    the compensated sum is the statevector amplitude, and count checks
    reject reads past the results."""
    from hiputil import say
    from jet_amd import circuits, paths
    from jet_amd.network import SlicedContraction

    circ = circuits.random_grid_circuit(4, 4, 8, seed=7)
    net = circuits.amplitude_network(circ)
    dims = {l: 2 for ix in net.indices for l in ix}
    path = paths.best_greedy_path(net.indices, dims, trials=4)
    sliced = paths.choose_slices(net.indices, dims, path, 4)
    amp = circuits.amplitude_statevector(circ)
    ref = None
    for use_graph in (False, True):
        say(f"slice loop past n_loop, use_graph={use_graph}")
        sc = SlicedContraction(net.indices, net.shapes, net.data, path, sliced,
                               batch=sliced[2:], precision="fp32", ctx=ctx,
                               use_graph=use_graph)
        assert sc.n_loop == 4 and sc.n_batch == 4
        sc.run()
        first = sc.slice_results()
        ref = first if ref is None else ref
        np.testing.assert_array_equal(first, ref)
        sc.run(range(3, 3 + 3 * sc.n_loop + 1))
        np.testing.assert_array_equal(sc.slice_results(), first)
        b = sc.capture(range(5, 5 + 2 * sc.n_loop))
        b.launch()
        ctx.synchronize()
        b.close()
        np.testing.assert_array_equal(sc.slice_results(), first)
        total = sc.reduce()
        assert abs(total - amp) / abs(amp) < 1e-5
        with pytest.raises(JetError):
            sc.slice_results(sc.n_slices + 1)
        with pytest.raises(JetError):
            sc.reduce(sc.n_slices + 1)
        sc.net.close()

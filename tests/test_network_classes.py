"""TensorNetwork, PathInfo and TaskBasedContractor (jet_amd host classes).

Every case and literal value below is one of the reference's own:
test/Test_TensorNetwork.cpp:77-613, test/Test_PathInfo.cpp:46-515,
test/Test_TaskBasedContractor.cpp:94-544 and
python/tests/test_task_based_contractor.py:18-50.  Each test runs twice: with
a numpy tensor whose contractions are the oracle's (CPU: the bookkeeping of
the classes), and with jet_amd.Tensor (GPU: the same values from the HIP
kernels).
"""
import numpy as np
import pytest

from hosttensor import HostTensor
from jet_amd.abort import JetError
from jet_amd.path_info import MISSING_ID, PathInfo
from jet_amd.task_based_contractor import TaskBasedContractor
from jet_amd.tensor_network import Edge, TensorNetwork


def _gpu_tensor():
    from jet_amd import Tensor
    return Tensor


@pytest.fixture(params=["host", pytest.param("gpu", marks=pytest.mark.gpu)])
def TT(request):
    return HostTensor if request.param == "host" else _gpu_tensor()


def make(TT, indices, shape, imag=False):
    """MakeTensor of the reference tests: element i is i (or i + 2i*j)."""
    n = int(np.prod(shape)) if shape else 1
    data = np.arange(n, dtype=np.float64)
    data = data + 2j * data if imag else data + 0j
    if not shape:
        data = np.zeros(1, np.complex128)
    return TT(indices, shape, data.astype(np.complex64), dtype=np.complex64)


def data_of(t):
    return [complex(x) for x in np.asarray(t.get_data()).reshape(-1)]


# --------------------------------------------------------------------------
# TensorNetwork (Test_TensorNetwork.cpp)
# --------------------------------------------------------------------------

def test_tn_num_indices_and_tensors(TT):
    tn = TensorNetwork(TT)
    assert tn.NumIndices() == 0 and tn.NumTensors() == 0
    tn.AddTensor(make(TT, ["A0", "B1"], [2, 3], True), [])
    assert tn.NumIndices() == 2
    tn.AddTensor(make(TT, ["B1", "C2"], [3, 2], True), [])
    assert tn.NumIndices() == 3 and tn.NumTensors() == 2


def test_tn_add_tensor(TT):
    tn = TensorNetwork(TT)
    tn.AddTensor(make(TT, [], []), [])
    node = tn.GetNodes()[0]
    assert (node.id, node.name, node.indices, node.tags, node.contracted) == (0, "_", [], [], False)
    assert tn.GetIndexToEdgeMap() == {} and tn.GetTagToNodesMap() == {}

    tn = TensorNetwork(TT)
    tn.AddTensor(make(TT, ["A0", "B1"], [2, 3], True), ["chaotic", "neutral"])
    node = tn.GetNodes()[0]
    assert (node.name, node.indices, node.tags) == ("A0B1", ["A0", "B1"], ["chaotic", "neutral"])
    assert tn.GetIndexToEdgeMap() == {"A0": Edge(2, [0]), "B1": Edge(3, [0])}
    assert tn.GetTagToNodesMap() == {"chaotic": [0], "neutral": [0]}

    tn = TensorNetwork(TT)
    tn.AddTensor(make(TT, ["C2", "D3"], [1, 4], True), ["bot", "mid"])
    tn.AddTensor(make(TT, ["D3", "E4"], [4, 2], True), ["mid", "top"])
    assert [n.name for n in tn.GetNodes()] == ["C2D3", "D3E4"]
    assert [n.tags for n in tn.GetNodes()] == [["bot", "mid"], ["mid", "top"]]
    # The extent-1 index C2 is not an edge.
    assert tn.GetIndexToEdgeMap() == {"D3": Edge(4, [0, 1]), "E4": Edge(2, [1])}
    assert tn.GetTagToNodesMap() == {"bot": [0], "mid": [0, 1], "top": [1]}


@pytest.mark.parametrize("sliced,value,name,node_idx,t_idx,t_shape,want", [
    (["D3"], 0, "A0B1C2", ["A0", "B1", "C2"], ["A0", "B1", "C2"], [2, 3, 4], list(range(24))),
    (["A0"], 0, "A0(0)B1C2", ["A0(0)", "B1", "C2"], ["B1", "C2"], [3, 4], list(range(12))),
    (["A0"], 1, "A0(1)B1C2", ["A0(1)", "B1", "C2"], ["B1", "C2"], [3, 4], list(range(12, 24))),
    (["A0", "C2"], 1 * 4 + 2, "A0(1)B1C2(2)", ["A0(1)", "B1", "C2(2)"], ["B1"], [3],
     [14, 18, 22]),
    (["A0", "B1", "C2"], 1 * 3 * 4 + 2 * 4 + 3, "A0(1)B1(2)C2(3)", ["A0(1)", "B1(2)", "C2(3)"],
     [], [], [23]),
])
def test_tn_slice_indices(TT, sliced, value, name, node_idx, t_idx, t_shape, want):
    tn = TensorNetwork(TT)
    tn.AddTensor(make(TT, ["A0", "B1", "C2"], [2, 3, 4], True), [])
    tn.AddTensor(make(TT, ["D3"], [2], True), [])
    tn.SliceIndices(sliced, value)
    node = tn.GetNodes()[0]
    assert node.name == name and node.indices == node_idx
    assert list(node.tensor.indices) == t_idx and list(node.tensor.shape) == t_shape
    assert data_of(node.tensor) == [complex(v, 2 * v) for v in want]
    for label in sliced:
        assert label not in tn.GetIndexToEdgeMap()


def test_tn_slice_missing_index(TT):
    tn = TensorNetwork(TT)
    tn.AddTensor(make(TT, ["A0", "B1", "C2"], [2, 3, 4], True), [])
    with pytest.raises(JetError, match="Sliced index does not exist."):
        tn.SliceIndices(["E0", "B0"], 0)


def test_tn_contract_implicit(TT):
    # [2]
    tn = TensorNetwork(TT)
    tn.AddTensor(make(TT, ["A0"], [2], True), [])
    result = tn.Contract()
    assert tn.GetPath() == [] and data_of(result) == [0, 1 + 2j]
    assert tn.GetIndexToEdgeMap() == {"A0": Edge(2, [0])}

    # [3] -(0)- [3]
    tn = TensorNetwork(TT)
    tn.AddTensor(make(TT, ["A0"], [3], True), [])
    tn.AddTensor(make(TT, ["A0"], [3], True), [])
    result = tn.Contract()
    assert tn.GetPath() == [(0, 1)] and data_of(result) == [-15 + 20j]
    nodes = tn.GetNodes()
    assert nodes[0].contracted and nodes[1].contracted
    assert (nodes[2].id, nodes[2].name, nodes[2].indices, nodes[2].contracted) == (2, "_", [], False)
    assert tn.GetIndexToEdgeMap() == {}

    # [3, 2] -(0)- [3]
    tn = TensorNetwork(TT)
    tn.AddTensor(make(TT, ["A0", "B1"], [3, 2], True), [])
    tn.AddTensor(make(TT, ["A0"], [3], True), [])
    result = tn.Contract()
    assert data_of(result) == [-30 + 40j, -39 + 52j]
    assert tn.GetNodes()[2].name == "B1"
    assert tn.GetIndexToEdgeMap() == {"B1": Edge(2, [2])}

    # [3, 2] -(0, 1)- [3, 2]
    tn = TensorNetwork(TT)
    tn.AddTensor(make(TT, ["A0", "B1"], [3, 2], True), [])
    tn.AddTensor(make(TT, ["A0", "B1"], [3, 2], True), [])
    result = tn.Contract()
    assert data_of(result) == [-165 + 220j] and tn.GetIndexToEdgeMap() == {}

    # [2, 3] -(1)- [3, 3]
    tn = TensorNetwork(TT)
    tn.AddTensor(make(TT, ["A0", "B1"], [2, 3], True), [])
    tn.AddTensor(make(TT, ["C2", "B1"], [3, 3], True), [])
    result = tn.Contract()
    assert list(result.shape) == [2, 3]
    assert data_of(result) == [-15 + 20j, -42 + 56j, -69 + 92j, -42 + 56j, -150 + 200j,
                               -258 + 344j]
    assert tn.GetNodes()[2].name == "A0C2"
    assert tn.GetIndexToEdgeMap() == {"A0": Edge(2, [2]), "C2": Edge(3, [2])}


def test_tn_contract_explicit(TT):
    tn = TensorNetwork(TT)
    tn.AddTensor(make(TT, ["A0", "B1"], [2, 3], True), [])
    tn.AddTensor(make(TT, ["C2", "B1"], [2, 3], True), [])
    tn.AddTensor(make(TT, ["C2", "D3"], [2, 2], True), [])
    result = tn.Contract([(1, 2), (0, 3)])
    assert tn.GetPath() == [(1, 2), (0, 3)]
    assert list(result.shape) == [2, 2]
    assert data_of(result) == [-308 - 56j, -517 - 94j, -1100 - 200j, -1804 - 328j]
    nodes = tn.GetNodes()
    assert [n.contracted for n in nodes] == [True, True, True, True, False]
    assert (nodes[3].name, nodes[3].indices) == ("B1D3", ["B1", "D3"])
    assert (nodes[4].name, nodes[4].indices) == ("A0D3", ["A0", "D3"])
    assert tn.GetIndexToEdgeMap() == {"D3": Edge(2, [4]), "A0": Edge(2, [4])}


def test_tn_contract_errors(TT):
    with pytest.raises(JetError, match="An empty tensor network cannot be contracted."):
        TensorNetwork(TT).Contract()
    tn = TensorNetwork(TT)
    tn.AddTensor(make(TT, ["A0", "B1"], [2, 3], True), [])
    tn.AddTensor(make(TT, ["C2", "B1"], [2, 3], True), [])
    tn.AddTensor(make(TT, ["C2", "D3"], [2, 2], True), [])
    with pytest.raises(JetError, match="Node ID 1 in contraction pair is invalid."):
        tn.Contract([(10, 2), (0, 3)])
    tn = TensorNetwork(TT)
    tn.AddTensor(make(TT, ["A0", "B1"], [2, 3], True), [])
    tn.AddTensor(make(TT, ["C2", "B1"], [2, 3], True), [])
    tn.AddTensor(make(TT, ["C2", "D3"], [2, 2], True), [])
    with pytest.raises(JetError, match="Node ID 2 in contraction pair is invalid."):
        tn.Contract([(1, 2), (0, 4)])


# --------------------------------------------------------------------------
# PathInfo (Test_PathInfo.cpp)
# --------------------------------------------------------------------------

def test_path_info_default():
    p = PathInfo()
    assert p.GetNumLeaves() == 0 and p.GetPath() == [] and p.GetIndexSizes() == {}
    assert p.GetSteps() == [] and p.total_flops() == 0 and p.total_memory() == 0


def test_path_info_steps(TT):
    tn = TensorNetwork(TT)
    p = PathInfo(tn, [])
    assert p.GetNumLeaves() == 0 and p.GetSteps() == []

    tn = TensorNetwork(TT)
    tn.AddTensor(make(TT, ["A0"], [3]), [])
    p = PathInfo(tn, [])
    assert p.GetNumLeaves() == 1 and p.GetIndexSizes() == {"A0": 3}
    s = p.GetSteps()[0]
    assert (s.id, s.parent, s.children) == (0, MISSING_ID, (MISSING_ID, MISSING_ID))
    assert (s.node_indices, s.tensor_indices, s.contracted_indices, s.tags) == (
        ["A0"], ["A0"], [], [])

    tn = TensorNetwork(TT)
    tn.AddTensor(make(TT, ["A0", "B1"], [3, 2]), ["apple"])
    tn.AddTensor(make(TT, ["B1", "C2"], [2, 4]), ["banana"])
    tn.AddTensor(make(TT, ["D3"], [5]), ["cherry"])
    p = PathInfo(tn, [(0, 1)])
    assert p.GetNumLeaves() == 3 and p.GetPath() == [(0, 1)]
    assert p.GetIndexSizes() == {"A0": 3, "B1": 2, "C2": 4, "D3": 5}
    steps = p.GetSteps()
    assert len(steps) == 4
    assert [s.parent for s in steps] == [3, 3, MISSING_ID, MISSING_ID]
    assert steps[3].children == (0, 1)
    assert steps[3].node_indices == ["A0", "C2"] and steps[3].tensor_indices == ["A0", "C2"]
    assert steps[3].contracted_indices == ["B1"] and steps[3].tags == ["apple", "banana"]
    assert steps[2].tags == ["cherry"] and steps[2].node_indices == ["D3"]


def test_path_info_sliced(TT):
    tn = TensorNetwork(TT)
    tn.AddTensor(make(TT, ["A0", "B1"], [3, 2]), ["apple"])
    tn.AddTensor(make(TT, ["B1", "C2"], [2, 4]), ["banana"])
    tn.SliceIndices(["A0", "C2"], 0)
    p = PathInfo(tn, [(0, 1)])
    assert p.GetNumLeaves() == 2 and p.GetIndexSizes() == {"B1": 2}
    steps = p.GetSteps()
    assert steps[0].node_indices == ["A0(0)", "B1"] and steps[0].tensor_indices == ["B1"]
    assert steps[1].node_indices == ["B1", "C2(0)"] and steps[1].tensor_indices == ["B1"]
    assert steps[2].node_indices == ["A0(0)", "C2(0)"] and steps[2].tensor_indices == []
    assert steps[2].contracted_indices == ["B1"] and steps[2].parent == MISSING_ID


@pytest.mark.parametrize("a,b,want", [
    (([], []), ([], []), 2),
    ((["A0"], [2]), (["B1"], [3]), 12),
    ((["A0"], [3]), (["A0"], [3]), 3 + 3),
    ((["A0", "B1", "C2"], [2, 3, 4]), (["A0", "B1", "C3"], [2, 3, 5]), (4 * 5) * (6 + 6)),
])
def test_path_info_step_flops(TT, a, b, want):
    tn = TensorNetwork(TT)
    tn.AddTensor(make(TT, *a), [])
    tn.AddTensor(make(TT, *b), [])
    p = PathInfo(tn, [(0, 1)])
    assert p.GetPathStepFlops(0) == 0
    assert p.GetPathStepFlops(2) == want
    with pytest.raises(JetError, match="Step ID is invalid."):
        p.GetPathStepFlops(20)


def test_path_info_totals(TT):
    tn = TensorNetwork(TT)
    assert PathInfo(tn, []).GetTotalFlops() == 0 and PathInfo(tn, []).GetTotalMemory() == 0
    tn.AddTensor(make(TT, ["A0", "B1"], [4, 2]), [])
    tn.AddTensor(make(TT, ["B1", "C2"], [2, 3]), [])
    tn.AddTensor(make(TT, ["C2", "D3"], [3, 5]), [])
    p = PathInfo(tn, [(0, 1), (2, 3)])
    assert p.GetTotalFlops() == 12 * (2 + 2) + 20 * (3 + 3)
    for bad, msg in (([(10, 1), (2, 3)], "Node ID 1"), ([(0, 10), (2, 3)], "Node ID 2"),
                     ([(0, 1), (20, 3)], "Node ID 1"), ([(0, 1), (2, 30)], "Node ID 2")):
        with pytest.raises(JetError, match=msg + " in contraction path pair is invalid."):
            PathInfo(tn, bad)

    tn = TensorNetwork(TT)
    tn.AddTensor(make(TT, ["A0", "B1", "C2", "D3", "E4"], [1, 2, 3, 4, 5]), [])
    p = PathInfo(tn, [])
    assert p.GetPathStepMemory(0) == 120
    with pytest.raises(JetError, match="Step ID is invalid."):
        p.GetPathStepMemory(10)

    tn = TensorNetwork(TT)
    tn.AddTensor(make(TT, ["A0", "B1"], [3, 2]), [])
    tn.AddTensor(make(TT, ["B1", "C2"], [2, 4]), [])
    tn.AddTensor(make(TT, ["D3"], [5]), [])
    assert PathInfo(tn, [(0, 1)]).GetTotalMemory() == 6 + 8 + 5 + 12


# --------------------------------------------------------------------------
# TaskBasedContractor (Test_TaskBasedContractor.cpp)
# --------------------------------------------------------------------------

def tensor_map(tbc):
    return {k: (None if v is None else data_of(v)) for k, v in tbc.GetNameToTensorMap().items()}


def test_tbc_empty(TT):
    tbc = TaskBasedContractor(tensor_type=TT, dtype=np.complex64)
    assert tbc.AddContractionTasks(TensorNetwork(TT), PathInfo()) == 0
    assert tbc.GetFlops() == 0 and tbc.GetMemory() == 0
    assert tbc.GetNameToTaskMap() == {} and tbc.GetNameToTensorMap() == {}
    assert tbc.GetNameToParentsMap() == {}
    tn = TensorNetwork(TT)
    tn.AddTensor(make(TT, ["A0", "C2"], [2, 4]), [])
    tn.AddTensor(make(TT, ["A0", "B1"], [2, 3]), [])
    assert tbc.AddContractionTasks(tn, PathInfo(tn, [])) == 0
    assert tbc.GetNameToTaskMap() == {} and tbc.GetFlops() == 0
    tbc.contract()
    assert tbc.GetResults() == []


def test_tbc_path_not_empty(TT):
    tbc = TaskBasedContractor(tensor_type=TT, dtype=np.complex64)
    tn = TensorNetwork(TT)
    tn.AddTensor(make(TT, ["A0", "C2"], [2, 4]), [])
    tn.AddTensor(make(TT, ["A0", "B1"], [2, 3]), [])
    tn.AddTensor(make(TT, ["B1", "C2"], [3, 4]), [])
    assert tbc.AddContractionTasks(tn, PathInfo(tn, [(0, 1), (1, 2), (3, 4)])) == 0
    assert tbc.GetFlops() == (3 * 4) * 4 + (2 * 4) * 6 + (2 * 3) * 8
    assert tbc.GetMemory() == (3 * 4) + (2 * 4) + (2 * 3)
    assert {k: t.name() for k, t in tbc.GetNameToTaskMap().items()} == {
        "3:C2B1": "3:C2B1", "4:A0C2": "4:A0C2", "5:B1A0:results[0]": "5:B1A0:results[0]"}
    assert tensor_map(tbc) == {
        "0:A0C2": list(map(complex, range(8))), "1:A0B1": list(map(complex, range(6))),
        "2:B1C2": list(map(complex, range(12))), "3:C2B1": None, "4:A0C2": None,
        "5:B1A0:results[0]": None}
    assert tbc.GetNameToParentsMap() == {
        "0:A0C2": {"3:C2B1"}, "1:A0B1": {"3:C2B1", "4:A0C2"}, "2:B1C2": {"4:A0C2"},
        "3:C2B1": {"5:B1A0:results[0]"}, "4:A0C2": {"5:B1A0:results[0]"}}


def test_tbc_shared_tasks(TT):
    tbc = TaskBasedContractor(tensor_type=TT, dtype=np.complex64)
    tn = TensorNetwork(TT)
    tn.AddTensor(make(TT, ["A0", "B1"], [2, 3]), [])
    tn.AddTensor(make(TT, ["A0"], [2]), [])
    tn.AddTensor(make(TT, ["B1"], [3]), [])
    p = PathInfo(tn, [(0, 1), (2, 3)])
    assert tbc.AddContractionTasks(tn, p) == 0
    assert tbc.AddContractionTasks(tn, p) == 1
    assert tbc.GetFlops() == 3 * 4 + 2 * 6 and tbc.GetMemory() == 3 + 2 * 1
    assert set(tbc.GetNameToTaskMap()) == {"3:B1", "4:_:results[0]", "4:_:results[1]"}
    assert tbc.GetNameToParentsMap() == {
        "0:A0B1": {"3:B1"}, "1:A0": {"3:B1"}, "2:B1": {"4:_:results[0]", "4:_:results[1]"},
        "3:B1": {"4:_:results[0]", "4:_:results[1]"}}
    tbc.contract()
    assert [data_of(r) for r in tbc.GetResults()] == [[14], [14]]


def _scalar_pair_network(TT, n_a=3, n_b=4):
    tn = TensorNetwork(TT)
    tn.AddTensor(make(TT, ["A0"], [n_a]), [])
    tn.AddTensor(make(TT, ["A0"], [n_a]), [])
    tn.AddTensor(make(TT, ["B1"], [n_b]), [])
    tn.AddTensor(make(TT, ["B1"], [n_b]), [])
    return tn


def test_tbc_contract(TT):
    tbc = TaskBasedContractor(tensor_type=TT, dtype=np.complex64)
    tn = TensorNetwork(TT)
    tn.AddTensor(make(TT, ["A0"], [2]), [])
    tn.AddTensor(make(TT, ["A0", "B1"], [2, 3]), [])
    tbc.AddContractionTasks(tn, PathInfo(tn, [(0, 1)]))
    tbc.Contract().result()
    (r,) = tbc.GetResults()
    assert list(r.indices) == ["B1"] and data_of(r) == [3, 4, 5]

    tbc = TaskBasedContractor(tensor_type=TT, dtype=np.complex64)
    tn = _scalar_pair_network(TT)
    tbc.AddContractionTasks(tn, PathInfo(tn, [(0, 1)]))
    tbc.AddContractionTasks(tn, PathInfo(tn, [(2, 3)]))
    tbc.contract()
    assert [data_of(r) for r in tbc.GetResults()] == [[5], [14]]


def test_tbc_deletion(TT):
    tbc = TaskBasedContractor(tensor_type=TT, dtype=np.complex64)
    tn = TensorNetwork(TT)
    assert tbc.AddContractionTasks(tn, PathInfo()) == 0 and tbc.AddDeletionTasks() == 0

    tbc = TaskBasedContractor(tensor_type=TT, dtype=np.complex64)
    tn = TensorNetwork(TT)
    tn.AddTensor(make(TT, ["A0"], [3]), [])
    tn.AddTensor(make(TT, ["A0"], [3]), [])
    tbc.AddContractionTasks(tn, PathInfo(tn, [(0, 1)]))
    assert tbc.AddDeletionTasks() == 2
    tbc.contract()
    assert tensor_map(tbc) == {"0:A0": None, "1:A0": None, "2:_:results[0]": [5]}

    tbc = TaskBasedContractor(tensor_type=TT, dtype=np.complex64)
    tn = TensorNetwork(TT)
    tn.AddTensor(make(TT, ["A0", "B1"], [2, 3]), [])
    tn.AddTensor(make(TT, ["A0"], [2]), [])
    tn.AddTensor(make(TT, ["B1"], [3]), [])
    tbc.AddContractionTasks(tn, PathInfo(tn, [(0, 1), (2, 3)]))
    assert tbc.AddDeletionTasks() == 4
    tbc.contract()
    assert tensor_map(tbc) == {"0:A0B1": None, "1:A0": None, "2:B1": None, "3:B1": None,
                               "4:_:results[0]": [14]}


def test_tbc_reduction(TT):
    tbc = TaskBasedContractor(tensor_type=TT, dtype=np.complex64)
    assert data_of(tbc.GetReductionResult()) == [0] and list(tbc.GetReductionResult().indices) == []

    tbc = TaskBasedContractor(tensor_type=TT, dtype=np.complex64)
    tn = TensorNetwork(TT)
    tn.AddTensor(make(TT, ["A0", "B1"], [2, 3]), [])
    tn.AddTensor(make(TT, ["A0"], [2]), [])
    tbc.AddContractionTasks(tn, PathInfo(tn, [(0, 1)]))
    tbc.AddReductionTask()
    tbc.contract()
    r = tbc.GetReductionResult()
    assert list(r.indices) == ["B1"] and data_of(r) == [3, 4, 5]

    tbc = TaskBasedContractor(tensor_type=TT, dtype=np.complex64)
    tn = _scalar_pair_network(TT)
    tbc.AddContractionTasks(tn, PathInfo(tn, [(0, 1)]))
    tbc.AddContractionTasks(tn, PathInfo(tn, [(2, 3)]))
    tbc.AddReductionTask()
    tbc.contract()
    assert data_of(tbc.GetReductionResult()) == [5 + 14]

    tbc = TaskBasedContractor(tensor_type=TT, dtype=np.complex64)
    tn = TensorNetwork(TT)
    tn.AddTensor(make(TT, ["A0"], [3]), [])
    tn.AddTensor(make(TT, ["A0"], [3]), [])
    tbc.AddContractionTasks(tn, PathInfo(tn, [(0, 1)]))
    assert [tbc.AddReductionTask() for _ in range(3)] == [1, 0, 0]
    tbc.contract()
    assert data_of(tbc.GetReductionResult()) == [5]


@pytest.mark.parametrize("dtype", [np.complex64, np.complex128])
def test_tbc_python_contract(TT, dtype):
    # python/tests/test_task_based_contractor.py:18-50 with its tensor_network
    # fixture (python/tests/conftest.py:6-20).
    tn = TensorNetwork(TT)
    tn.add_tensor(TT(["i", "j"], [2, 2], np.array([1, 1j, -1j, 1]), dtype=dtype), ["A", "Hermitian"])
    tn.add_tensor(TT(["j", "k"], [2, 2], np.array([1, 0, 0, 1]), dtype=dtype),
                  ["B", "Identity", "Real"])
    tn.add_tensor(TT(["k"], [2], np.array([1, 0]), dtype=dtype), ["C", "Vector", "Real"])
    path = PathInfo(tn, [[0, 1], [2, 3]])
    tbc = TaskBasedContractor(tensor_type=TT, dtype=dtype)
    assert tbc.add_contraction_tasks(tn, path) == 0
    assert tbc.add_deletion_tasks() == 4
    assert tbc.add_reduction_task() == 1
    tbc.contract()
    m = tbc.name_to_tensor_map
    assert {k for k, v in m.items() if v is None} == {"0:ij", "1:jk", "2:k", "3:ik"}
    want = [1, -1j]
    assert data_of(m["4:i:results[0]"]) == want
    assert tbc.name_to_parents_map == {"0:ij": {"3:ik"}, "1:jk": {"3:ik"},
                                       "2:k": {"4:i:results[0]"}, "3:ik": {"4:i:results[0]"}}
    assert data_of(tbc.results[0]) == want and data_of(tbc.reduction_result) == want
    assert tbc.flops == 2 * 2 * 4 + 2 * 4 and tbc.memory == 6

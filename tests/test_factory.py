"""The reference's Python class names and dtype-dispatching factories.

python/jet/factory.py:40-151 dispatches on ``dtype`` (default complex128) to
the C64/C128 classes of python/src/Type.hpp:10-18 and raises
``TypeError("Data type '...' is not supported.")`` for anything else.  The
GPU test rebuilds python/tests/test_task_based_contractor.py:18-50 (network
of python/tests/conftest.py:6-20; result [1, -1j], flops 2*2*4 + 2*4,
memory 6) through these names.
"""
import numpy as np
import pytest

from jet_amd import factory


@pytest.mark.parametrize("make", [factory.Tensor, factory.TensorNetwork,
                                  factory.TaskBasedContractor, factory.TensorNetworkFile,
                                  factory.TensorNetworkSerializer])
@pytest.mark.parametrize("dtype", [np.float32, "float64", np.int64, "complex256"])
def test_unsupported_dtype_raises_type_error(make, dtype):
    with pytest.raises(TypeError, match=r"Data type '.*' is not supported\."):
        make(dtype=dtype)


def test_factories_default_to_complex128():
    assert factory.TensorNetwork().dtype == np.complex128
    assert factory.TensorNetworkC64().dtype == np.complex64


def _network(dtype):
    # python/tests/conftest.py:9-18
    A = factory.Tensor(shape=[2, 2], indices=["i", "j"], data=[1, 1j, -1j, 1], dtype=dtype)
    B = factory.Tensor(shape=[2, 2], indices=["j", "k"], data=[1, 0, 0, 1], dtype=dtype)
    C = factory.Tensor(shape=[2], indices=["k"], data=[1, 0], dtype=dtype)
    tn = factory.TensorNetwork(dtype=dtype)
    tn.add_tensor(A, ["A", "Hermitian"])
    tn.add_tensor(B, ["B", "Identity", "Real"])
    tn.add_tensor(C, ["C", "Vector", "Real"])
    return tn


@pytest.mark.gpu
@pytest.mark.parametrize("dtype", ["complex64", "complex128"])
def test_task_based_contractor_through_factories(ctx, dtype):
    from hiputil import say
    from jet_amd.path_info import PathInfo

    say(f"factory TBC {dtype}")
    tn = _network(dtype)
    path = PathInfo(tn, [[0, 1], [2, 3]])
    tbc = factory.TaskBasedContractor(dtype=dtype)
    assert tbc.add_contraction_tasks(tn, path) == 0
    assert tbc.add_deletion_tasks() == 4
    assert tbc.add_reduction_task() == 1
    tbc.contract()
    want = factory.Tensor(shape=[2], indices=["i"], data=[1, -1j], dtype=dtype)
    assert tbc.results == [want]
    assert tbc.reduction_result == want
    assert tbc.flops == 2 * 2 * 4 + 2 * 4
    assert tbc.memory == 6
    assert sorted(tbc.name_to_tensor_map) == ["0:ij", "1:jk", "2:k", "3:ik", "4:i:results[0]"]


@pytest.mark.gpu
def test_typed_names_build_the_typed_classes(ctx):
    from hiputil import say

    say("factory typed names")
    a = factory.TensorC64(["a", "b"], [2, 2], [1, 2, 3, 4])
    b = factory.TensorC128(["a", "b"], [2, 2], [1, 2, 3, 4])
    assert a.get_data().dtype == np.complex64 and b.get_data().dtype == np.complex128
    assert factory.Tensor(["a"], [2], [1, 2], dtype=np.complex64).get_data().dtype == np.complex64
    c = factory.TensorC64(["b", "c"], [2, 2], [1, 0, 0, 1])
    from jet_amd import contract_tensors
    r = contract_tensors(a, c)
    assert r.indices == ["a", "c"] and np.array_equal(r.get_data(), [1, 2, 3, 4])
    tbc = factory.TaskBasedContractorC64()
    assert tbc.flops == 0 and tbc.results == []
    tn = factory.TensorNetworkC128()
    tn.add_tensor(b)
    assert tn.num_tensors == 1

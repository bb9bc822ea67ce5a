"""Permuter front-end validation against the reference's exception tests.

Cases and expected messages are those of test/Test_Permuter.cpp:390-456
("QFlexPermuter<>::Transpose exceptional data conditions"), whose input is
the 30-element array data_non_pow2_cab of that file; the messages the tests
match (Contains(...)) come from include/jet/permute/Permuter.hpp:59-79 and
102-117.  The checks run on the host before any device work, so these
tests need no GPU.  The QFlex backend's own "Fast transpose expects power-
of-2 data" error (Test_Permuter.cpp:350-386) has no counterpart: the GPU
permutation accepts every shape, and tests/test_gpu_kernels.py checks those
non-power-of-2 cases' outputs against the same file's expected arrays.
"""
import numpy as np
import pytest

from jet_amd.abort import JetError
from jet_amd.permuter import Permuter


@pytest.fixture
def data_non_pow2_cab(golden_permuter):
    vals = golden_permuter["arrays"]["data_non_pow2_cab"]
    return np.array([complex(re, im) for re, im in vals], dtype=np.complex64)


def _raises(msg, fn, *args):
    with pytest.raises(JetError) as e:
        fn(*args)
    assert msg in str(e.value), str(e.value)
    assert str(e.value).startswith("Error in Jet: ")


def test_fixture_is_the_30_element_array(data_non_pow2_cab):
    assert data_non_pow2_cab.size == 30


def test_incorrect_shape_indices_match(data_non_pow2_cab):
    # Test_Permuter.cpp:399-412
    p = Permuter()
    _raises("Tensor shape does not match number of indices.", p.transpose,
            data_non_pow2_cab, [2, 3, 5, 2], ["c", "a", "b"], ["a", "b", "c"])


def test_incorrect_shape_data_match(data_non_pow2_cab):
    # Test_Permuter.cpp:413-424
    _raises("Tensor shape does not match given tensor data.", Permuter().transpose,
            data_non_pow2_cab, [2, 3, 5, 2], ["c", "a", "b", "d"], ["a", "b", "c", "d"])


def test_empty_shape(data_non_pow2_cab):
    # Test_Permuter.cpp:425-436
    _raises("Tensor shape does not match number of indices", Permuter().transpose,
            data_non_pow2_cab, [], ["c", "a", "b", "d"], ["a", "b", "c"])


def test_empty_indices(data_non_pow2_cab):
    # Test_Permuter.cpp:437-455
    p = Permuter()
    _raises("Tensor shape does not match number of indices", p.transpose,
            data_non_pow2_cab, [2, 3, 5], ["c", "a", "b", "d"], [])
    with pytest.raises(JetError):
        p.transpose(data_non_pow2_cab, [2, 3, 5], [], [])
    _raises("Tensor shape does not match given tensor data", p.transpose,
            data_non_pow2_cab, [], [], [])


def test_front_end_checks_in_order(data_non_pow2_cab):
    """The remaining checks of Permuter.hpp:59-79 and 102-117, in the order
    the front end makes them (the first failing one names the error)."""
    p = Permuter()
    x = data_non_pow2_cab
    _raises("Duplicate existing indices found. Please ensure indices are unique.",
            p.transpose, x, [2, 3, 5], ["a", "a", "b"], ["a", "b", "c"])
    _raises("Duplicate transpose indices found. Please ensure indices are unique.",
            p.transpose, x, [2, 3, 5], ["a", "b", "c"], ["a", "b", "b"])
    _raises("New indices are an invalid permutation of the existing indices",
            p.transpose, x, [2, 3, 5], ["a", "b", "c"], ["a", "b", "d"])
    # duplicates are reported before a shape mismatch
    _raises("Duplicate existing indices", p.transpose, x, [2, 3], ["a", "a", "b"], ["a", "b", "c"])
    # the in-place form names the input and output buffers (Permuter.hpp:65-71)
    out = np.zeros(29, np.complex64)
    _raises("Tensor shape does not match given output tensor data.", p.transpose,
            x, [2, 3, 5], ["c", "a", "b"], ["a", "b", "c"], out)
    _raises("Tensor shape does not match given input tensor data.", p.transpose,
            x[:29], [2, 3, 5], ["c", "a", "b"], ["a", "b", "c"], np.zeros(30, np.complex64))

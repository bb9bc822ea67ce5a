"""Type-specialised constructors with the reference's Python names.

python/jet/factory.py:40-151 and python/src/Type.hpp:10-18 expose
``TensorC64``/``TensorC128``, ``TensorNetworkC64``/``C128`` and
``TaskBasedContractorC64``/``C128``, ``TensorNetworkFileC64``/``C128`` and
``TensorNetworkSerializerC64``/``C128``, plus dtype-dispatching factories of
the same names without the suffix, whose default dtype is complex128.  The same names here build the GPU-backed classes.
"""
from __future__ import annotations

import numpy as np

from .task_based_contractor import TaskBasedContractor as _TBC
from .tensor import Tensor as _Tensor
from .tensor_network import TensorNetwork as _TN
from .tensor_network_io import TensorNetworkFile as _TNF
from .tensor_network_io import TensorNetworkSerializer as _TNS

__all__ = ["TensorC64", "TensorC128", "TensorNetworkC64", "TensorNetworkC128",
           "TaskBasedContractorC64", "TaskBasedContractorC128", "Tensor", "TensorNetwork",
           "TaskBasedContractor", "TensorNetworkFileC64", "TensorNetworkFileC128",
           "TensorNetworkSerializerC64", "TensorNetworkSerializerC128", "TensorNetworkFile",
           "TensorNetworkSerializer"]


def _check(dtype) -> np.dtype:
    dt = np.dtype(dtype)
    if dt not in (np.dtype(np.complex64), np.dtype(np.complex128)):
        raise TypeError(f"Data type '{dtype}' is not supported.")
    return dt


def TensorC64(*args, **kwargs) -> _Tensor:
    return _Tensor(*args, dtype=np.complex64, **kwargs)


def TensorC128(*args, **kwargs) -> _Tensor:
    return _Tensor(*args, dtype=np.complex128, **kwargs)


def TensorNetworkC64() -> _TN:
    return _TN(_Tensor, np.complex64)


def TensorNetworkC128() -> _TN:
    return _TN(_Tensor, np.complex128)


def TaskBasedContractorC64(*args, **kwargs) -> _TBC:
    return _TBC(*args, tensor_type=_Tensor, dtype=np.complex64, **kwargs)


def TaskBasedContractorC128(*args, **kwargs) -> _TBC:
    return _TBC(*args, tensor_type=_Tensor, dtype=np.complex128, **kwargs)


def Tensor(*args, dtype=np.complex128, **kwargs) -> _Tensor:
    return _Tensor(*args, dtype=_check(dtype), **kwargs)


def TensorNetwork(dtype=np.complex128) -> _TN:
    return _TN(_Tensor, _check(dtype))


def TaskBasedContractor(*args, dtype=np.complex128, **kwargs) -> _TBC:
    return _TBC(*args, tensor_type=_Tensor, dtype=_check(dtype), **kwargs)


def TensorNetworkFileC64() -> _TNF:
    return _TNF(_TN(_Tensor, np.complex64))


def TensorNetworkFileC128() -> _TNF:
    return _TNF(_TN(_Tensor, np.complex128))


def TensorNetworkSerializerC64(indent: int = -1) -> _TNS:
    return _TNS(indent, _Tensor, np.complex64)


def TensorNetworkSerializerC128(indent: int = -1) -> _TNS:
    return _TNS(indent, _Tensor, np.complex128)


def TensorNetworkFile(dtype=np.complex128) -> _TNF:
    return _TNF(_TN(_Tensor, _check(dtype)))


def TensorNetworkSerializer(indent: int = -1, dtype=np.complex128) -> _TNS:
    return _TNS(indent, _Tensor, _check(dtype))

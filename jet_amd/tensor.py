"""Device-resident tensors with the interface of Jet::Tensor.

``Tensor`` mirrors ``Jet::Tensor<T>`` (include/jet/Tensor.hpp:30-776): labelled
indices, row-major data (last index fastest), and the operations of the
contraction path -- ``ContractTensors`` (Tensor.hpp:708-752), ``Transpose``
(578-660), ``AddTensors`` (412-464), ``SliceIndex`` (493-536), ``Reshape``
(546-568), ``Conj`` (668-682).  The data lives in MI355X HBM and every
operation runs through libjet_hip.so; host copies are made only by
``get_data`` / ``get_value`` / ``==``.

Method names follow the reference's Python binding (python/src/Tensor.hpp:
``contract_with_tensor``, ``add_tensor``, ``slice_index``, ``transpose``, ...)
with the C++ names as aliases.
"""
from __future__ import annotations

import ctypes
from typing import Dict, List, Optional, Sequence, Tuple, Union

import numpy as np

from . import _native as nat
from .utilities import (generate_string_index, shape_to_size, vector_disjunctive_union,
                        vector_intersection, vector_subtraction, vector_union, ravel_index)
from .abort import JetError

_DTYPES = {np.dtype(np.complex64): nat.C64, np.dtype(np.complex128): nat.C128}

_PRECISIONS = {
    None: nat.PREC_DEFAULT, "default": nat.PREC_DEFAULT, "fp32": nat.PREC_FP32,
    "bf16x3": nat.PREC_BF16X3, "fp64": nat.PREC_FP64,
    "bf16x3_4m": nat.PREC_BF16X3_4M, "f16x3": nat.PREC_F16X3_4M,
}


def _dtype_code(dtype) -> int:
    dt = np.dtype(dtype)
    if dt not in _DTYPES:
        raise JetError("Tensor data type must be one of complex64, complex128")
    return _DTYPES[dt]


class _PlanCache:
    """Contraction plans keyed by the relabelled problem (so contractions of
    the same shapes with different label names share one plan)."""

    def __init__(self):
        self._plans: Dict[tuple, ctypes.c_void_p] = {}

    def get(self, ctx, dtype_code, prec, a_modes, a_ext, b_modes, b_ext, c_modes):
        key = (ctx.device, dtype_code, prec, tuple(a_modes), tuple(a_ext), tuple(b_modes),
               tuple(b_ext), tuple(c_modes))
        p = self._plans.get(key)
        if p is None:
            p = ctypes.c_void_p()
            nat.check(nat.lib().jet_hip_contraction_plan_create(
                ctx.handle, dtype_code, prec, nat.ROW_MAJOR, len(a_modes), nat.i32(a_modes),
                nat.i64(a_ext), len(b_modes), nat.i32(b_modes), nat.i64(b_ext), len(c_modes),
                nat.i32(c_modes), ctypes.byref(p)))
            self._plans[key] = p
        return p


_plan_cache = _PlanCache()


def _relabel(*label_lists):
    ids: Dict[str, int] = {}
    out = []
    for labels in label_lists:
        row = []
        for s in labels:
            if s not in ids:
                ids[s] = len(ids)
            row.append(ids[s])
        out.append(row)
    return out


class Tensor:
    """n-rank complex tensor on the GPU (row-major).

    Tensor()                       -- default tensor: no indices, one zero
    Tensor(shape)                  -- indices "?a", "?b", ..., zero-filled
    Tensor(indices, shape)         -- zero-filled
    Tensor(indices, shape, data)   -- data in row-major order
    """

    def __init__(self, indices=None, shape=None, data=None, dtype=np.complex64, ctx=None):
        self._ctx = ctx or nat.default_context()
        self.dtype = np.dtype(dtype)
        self._code = _dtype_code(self.dtype)
        if indices is not None and shape is None and all(isinstance(x, (int, np.integer)) for x in indices):
            shape = list(indices)
            indices = ["?" + generate_string_index(i) for i in range(len(shape))]
        if indices is None and shape is None:
            self._default = True
            self._init_indices_and_shape([], [])
            self._alloc(1)
            self._zero()
            return
        self._default = False
        indices = list(indices or [])
        shape = [int(s) for s in (shape or [])]
        self._init_indices_and_shape(indices, shape)
        self._alloc(shape_to_size(shape))
        if data is None:
            self._zero()
        else:
            self.set_data(data)

    # -- construction helpers ----------------------------------------------
    def _init_indices_and_shape(self, indices, shape):
        self._indices = list(indices)
        self._shape = [int(s) for s in shape]
        self._index_to_dimension = {i: s for i, s in zip(self._indices, self._shape)}

    def _alloc(self, n: int):
        self._size = int(n)
        self._buf = nat.DeviceBuffer(max(1, self._size) * self.dtype.itemsize, self._ctx)

    def _zero(self):
        nat.check(nat.lib().jet_hip_memset(self._ctx.handle, self._buf.ptr, 0,
                                           self._size * self.dtype.itemsize, None))

    @classmethod
    def _empty(cls, indices, shape, dtype, ctx) -> "Tensor":
        t = cls.__new__(cls)
        t._ctx = ctx
        t.dtype = np.dtype(dtype)
        t._code = _dtype_code(t.dtype)
        t._default = False
        t._init_indices_and_shape(indices, shape)
        t._alloc(shape_to_size(shape))
        return t

    def init_indices_and_shape(self, indices, shape):
        """InitIndicesAndShape (Tensor.hpp:144-153): relabel/reshape metadata
        only; the data is untouched."""
        self._init_indices_and_shape(indices, shape)
        self._default = False

    # -- accessors ---------------------------------------------------------
    @property
    def indices(self) -> List[str]:
        return list(self._indices)

    @property
    def shape(self) -> List[int]:
        return list(self._shape)

    @property
    def index_to_dimension_map(self) -> Dict[str, int]:
        return dict(self._index_to_dimension)

    @property
    def size(self) -> int:
        return self._size

    @property
    def data_ptr(self) -> int:
        return self._buf.ptr

    @property
    def context(self):
        return self._ctx

    def get_indices(self):
        return self.indices

    def get_shape(self):
        return self.shape

    def get_size(self):
        return self.size

    def get_index_to_dimension(self):
        return self.index_to_dimension_map

    def set_shape(self, shape):
        self._shape = [int(s) for s in shape]

    def get_data(self) -> np.ndarray:
        host = np.empty(self._size, dtype=self.dtype)
        self._buf.download(host)
        return host

    @property
    def data(self) -> np.ndarray:
        return self.get_data()

    def set_data(self, data) -> None:
        arr = np.ascontiguousarray(np.asarray(data, dtype=self.dtype).reshape(-1))
        if arr.size != self._size:
            raise JetError("Size of data and tensor do not match.")
        self._buf.upload(arr)

    def get_value(self, index: Sequence[int]):
        pos = ravel_index(index, self._shape)
        host = np.empty(1, dtype=self.dtype)
        nat.check(nat.lib().jet_hip_memcpy(self._ctx.handle, host.ctypes.data,
                                           self._buf.ptr + pos * self.dtype.itemsize,
                                           self.dtype.itemsize, 1, None, 0))
        return host[0]

    def set_value(self, index: Sequence[int], value) -> None:
        pos = ravel_index(index, self._shape)
        host = np.array([value], dtype=self.dtype)
        nat.check(nat.lib().jet_hip_memcpy(self._ctx.handle,
                                           self._buf.ptr + pos * self.dtype.itemsize,
                                           host.ctypes.data, self.dtype.itemsize, 0, None, 0))

    def get_scalar(self):
        return self.get_value([0] * 0) if False else self.get_data()[0]

    @property
    def scalar(self):
        return self.get_scalar()

    def is_scalar(self) -> bool:
        return self._size == 1

    def rename_index(self, pos: int, new_label: str) -> None:
        old = self._indices[pos]
        dim = self._index_to_dimension.pop(old)
        self._indices[pos] = new_label
        self._index_to_dimension[new_label] = dim

    def fill_random(self, seed: Optional[int] = None, distribution: int = 1) -> None:
        """Seeded fill; re and im uniform in [-1, 1) by default (the range of
        Tensor::FillRandom, Tensor.hpp:361-386) from a counter-based generator
        (values differ from std::mt19937's)."""
        if seed is None:
            seed = int(np.random.SeedSequence().entropy & 0xFFFFFFFFFFFFFFFF)
        nat.check(nat.lib().jet_hip_fill_random(self._ctx.handle, self._code, self._size,
                                                int(seed) & 0xFFFFFFFFFFFFFFFF, distribution,
                                                self._buf.ptr, None))

    def copy(self) -> "Tensor":
        t = Tensor._empty(self._indices, self._shape, self.dtype, self._ctx)
        t._default = self._default
        nat.check(nat.lib().jet_hip_memcpy(self._ctx.handle, t._buf.ptr, self._buf.ptr,
                                           self._size * self.dtype.itemsize, 2, None, 1))
        return t

    __copy__ = copy

    # -- comparison --------------------------------------------------------
    def __eq__(self, other) -> bool:
        if not isinstance(other, Tensor):
            return NotImplemented
        return (self._shape == other._shape and self._indices == other._indices
                and self._index_to_dimension == other._index_to_dimension
                and self.dtype == other.dtype
                and np.array_equal(self.get_data(), other.get_data()))

    def __ne__(self, other) -> bool:
        r = self.__eq__(other)
        return r if r is NotImplemented else not r

    def _is_default_zero(self) -> bool:
        return (not self._indices and not self._shape and self._size == 1
                and self.get_data()[0] == 0)

    def __repr__(self) -> str:
        return f"Tensor(indices={self._indices}, shape={self._shape}, dtype={self.dtype})"

    # -- operations --------------------------------------------------------
    def transpose(self, new_order: Union[Sequence[str], Sequence[int]]) -> "Tensor":
        return Tensor.Transpose(self, new_order)

    @staticmethod
    def Transpose(A: "Tensor", new_order) -> "Tensor":
        """Tensor::Transpose (Tensor.hpp:578-642) on the GPU permutation kernel."""
        new_order = list(new_order)
        if new_order and all(isinstance(x, (int, np.integer)) for x in new_order) and \
                not all(isinstance(x, str) for x in new_order):
            if len(A._indices) != len(new_order):
                raise JetError("Size of ordering must match number of tensor indices.")
            new_order = [A._indices[i] for i in new_order]
        if A._indices == new_order:
            return A.copy()
        if len(A._indices) == 0:
            raise JetError("Number of indices cannot be zero.")
        new_shape = [A._index_to_dimension[i] for i in new_order]
        out = Tensor._empty(new_order, new_shape, A.dtype, A._ctx)
        modes_in, modes_out = _relabel(A._indices, new_order)
        nat.check(nat.lib().jet_hip_permute(A._ctx.handle, A._code, nat.ROW_MAJOR,
                                            len(modes_in), nat.i32(modes_in), nat.i64(A._shape),
                                            nat.i32(modes_out), A._buf.ptr, out._buf.ptr, None))
        return out

    def conj(self) -> "Tensor":
        """Tensor::Conj (Tensor.hpp:668-682)."""
        host = np.conj(self.get_data())
        return Tensor(self._indices, self._shape, host, dtype=self.dtype, ctx=self._ctx)

    def reshape(self, new_shape: Sequence[int]) -> "Tensor":
        return Tensor.Reshape(self, new_shape)

    @staticmethod
    def Reshape(old: "Tensor", new_shape: Sequence[int]) -> "Tensor":
        """Tensor::Reshape (Tensor.hpp:546-558): same data, default labels."""
        if old._size != shape_to_size(new_shape):
            raise JetError("Size is inconsistent between tensors.")
        labels = ["?" + generate_string_index(i) for i in range(len(new_shape))]
        out = Tensor._empty(labels, list(new_shape), old.dtype, old._ctx)
        nat.check(nat.lib().jet_hip_memcpy(old._ctx.handle, out._buf.ptr, old._buf.ptr,
                                           old._size * old.dtype.itemsize, 2, None, 1))
        return out

    def slice_index(self, index: str, value: int) -> "Tensor":
        return Tensor.SliceIndex(self, index, value)

    @staticmethod
    def SliceIndex(tensor: "Tensor", index: str, value: int) -> "Tensor":
        """Tensor::SliceIndex (Tensor.hpp:493-526): drop `index`, keep block
        `value`; the remaining indices keep their order."""
        if index not in tensor._indices:
            raise JetError("Sliced index does not exist.")
        ax = tensor._indices.index(index)
        labels = [l for i, l in enumerate(tensor._indices) if i != ax]
        shape = [s for i, s in enumerate(tensor._shape) if i != ax]
        out = Tensor._empty(labels, shape, tensor.dtype, tensor._ctx)
        nat.check(nat.lib().jet_hip_slice(tensor._ctx.handle, tensor._code, nat.ROW_MAJOR,
                                          len(tensor._shape), nat.i64(tensor._shape), ax,
                                          int(value), tensor._buf.ptr, out._buf.ptr, None))
        return out

    def add_tensor(self, other: "Tensor") -> "Tensor":
        return Tensor.AddTensors(self, other)

    @staticmethod
    def AddTensors(A: "Tensor", B: "Tensor") -> "Tensor":
        """Tensor::AddTensors (Tensor.hpp:412-454).  A default tensor (no
        indices, one zero) is the additive identity used by reductions."""
        if A._is_default_zero():
            return B.copy()
        if B._is_default_zero():
            return A.copy()
        if vector_disjunctive_union(A._indices, B._indices):
            raise JetError("Tensor addition with disjoint indices is not supported.")
        out = Tensor._empty(A._indices, A._shape, A.dtype, A._ctx)
        ma, mb = _relabel(A._indices, B._indices)
        nat.check(nat.lib().jet_hip_add(A._ctx.handle, A._code, nat.ROW_MAJOR, len(ma),
                                        nat.i32(ma), nat.i64(A._shape), nat.i32(mb), A._buf.ptr,
                                        B._buf.ptr, out._buf.ptr, None))
        return out

    def contract_with_tensor(self, other: "Tensor", precision: Optional[str] = None) -> "Tensor":
        return Tensor.ContractTensors(self, other, precision)

    @staticmethod
    def ContractTensors(A: "Tensor", B: "Tensor", precision: Optional[str] = None) -> "Tensor":
        """Tensor::ContractTensors (Tensor.hpp:708-752): sum over the shared
        indices; the result's indices are (A \\ B) ++ (B \\ A)."""
        if A.dtype != B.dtype:
            raise JetError("Tensor data types do not match.")
        left = vector_subtraction(A._indices, B._indices)
        right = vector_subtraction(B._indices, A._indices)
        c_indices = vector_union(left, right)
        c_shape = [A._index_to_dimension[i] for i in left] + \
                  [B._index_to_dimension[i] for i in right]
        C = Tensor._empty(c_indices, c_shape, A.dtype, A._ctx)
        ma, mb, mc = _relabel(A._indices, B._indices, c_indices)
        plan = _plan_cache.get(A._ctx, A._code, _PRECISIONS[precision], ma, A._shape, mb,
                               B._shape, mc)
        nat.check(nat.lib().jet_hip_contract(plan, A._buf.ptr, B._buf.ptr, C._buf.ptr, None, 0,
                                             None))
        return C

    # C++-style aliases
    ContractWithTensor = contract_with_tensor
    AddTensor = add_tensor
    GetData = get_data
    GetIndices = get_indices
    GetShape = get_shape
    GetSize = get_size


def contract_tensors(A: Tensor, B: Tensor, precision: Optional[str] = None) -> Tensor:
    """jet.contract_tensors (python/jet/factory.py) on the GPU."""
    return Tensor.ContractTensors(A, B, precision)


def add_tensors(A: Tensor, B: Tensor) -> Tensor:
    return Tensor.AddTensors(A, B)


def slice_index(A: Tensor, index: str, value: int) -> Tensor:
    return Tensor.SliceIndex(A, index, value)


def transpose(A: Tensor, order) -> Tensor:
    return Tensor.Transpose(A, order)


def reshape(A: Tensor, shape) -> Tensor:
    return Tensor.Reshape(A, shape)


def conj(A: Tensor) -> Tensor:
    return A.conj()

"""A numpy tensor type for CPU checks of the host-side network classes.

It has the interface the TensorNetwork / PathInfo / TaskBasedContractor
mirrors use (the TensorType concept of TaskBasedContractor.hpp:23-29 and
TensorNetwork.hpp:23-37), with contraction and slicing done by the oracle, so
that the bookkeeping of those classes (names, edges, parents, flops, results)
is checked on the CPU; the same checks run on the GPU with jet_amd.Tensor.
"""
from __future__ import annotations

import numpy as np

import oracle
from jet_amd.abort import JetError


class HostTensor:
    def __init__(self, indices=None, shape=None, data=None, dtype=np.complex64):
        self.dtype = np.dtype(dtype)
        self._default = indices is None and shape is None
        self.indices = list(indices or [])
        self.shape = [int(s) for s in (shape or [])]
        n = int(np.prod(self.shape)) if self.shape else 1
        if data is None:
            self._data = np.zeros(n, self.dtype)
        else:
            self._data = np.asarray(data, self.dtype).reshape(-1).copy()
            if self._data.size != n:
                raise JetError("Size of data and tensor do not match.")

    # --- TensorType interface ------------------------------------------------
    def init_indices_and_shape(self, indices, shape):
        self.indices, self.shape = list(indices), [int(s) for s in shape]
        self._default = False

    def get_data(self):
        return self._data.copy()

    @property
    def data(self):
        return self._data.copy()

    def copy(self):
        t = HostTensor(self.indices, self.shape, self._data, self.dtype)
        t._default = self._default
        return t

    def __eq__(self, other):
        return (self.indices == other.indices and self.shape == other.shape
                and np.array_equal(self._data, other.get_data()))

    def _is_default_zero(self):
        return not self.indices and self._data.size == 1 and self._data[0] == 0

    @staticmethod
    def ContractTensors(A, B):
        lc, sc, c = oracle.contract(A.indices, A.shape, A._data, B.indices, B.shape, B._data)
        return HostTensor(lc, sc, np.asarray(c).astype(A.dtype), A.dtype)

    @staticmethod
    def SliceIndex(t, label, value):
        if label not in t.indices:
            raise JetError("Sliced index does not exist.")
        lab, shp, out = oracle.slice_index(t.indices, t.shape, t._data, label, value)
        return HostTensor(lab, shp, out, t.dtype)

    @staticmethod
    def Transpose(t, order):
        perm = [t.indices.index(l) for l in order]
        arr = t._data.reshape(t.shape).transpose(perm) if t.shape else t._data
        return HostTensor(list(order), [t.shape[p] for p in perm], arr.reshape(-1), t.dtype)

    @staticmethod
    def Reshape(t, shape):
        if int(np.prod(shape)) != t._data.size:
            raise JetError("Size is inconsistent between tensors.")
        return HostTensor(["?" + "abcdefghijklmnopqrstuvwxyz"[i] for i in range(len(shape))],
                          shape, t._data, t.dtype)

    def add_tensor(self, other):
        if self._is_default_zero():
            return other
        if other._is_default_zero():
            return self
        if sorted(self.indices) != sorted(other.indices):
            raise JetError("Tensor addition with disjoint indices is not supported.")
        b = HostTensor.Transpose(other, self.indices)
        return HostTensor(self.indices, self.shape, self._data + b._data, self.dtype)

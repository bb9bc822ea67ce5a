"""Permuter front end (include/jet/permute/Permuter.hpp:49-119) on the GPU.

``Permuter().transpose(data, shape, current_order, new_order)`` validates its
arguments exactly as the reference front end does (same checks, same order,
same messages) and then runs the permutation on the LDS-tiled HIP kernel.
Unlike the reference's QFlex backend (QFlex.hpp:333-336, 517-519) the GPU
backend accepts every shape, so no backend-specific errors exist.
"""
from __future__ import annotations

from typing import Optional, Sequence

import numpy as np

from . import _native as nat
from .abort import JetError
from .utilities import shape_to_size


class Permuter:
    def __init__(self, ctx: Optional[nat.Context] = None):
        self._ctx = ctx

    @staticmethod
    def _validate(n_in, shape, current_order, new_order, n_out=None):
        if len(set(current_order)) != len(current_order):
            raise JetError("Duplicate existing indices found. Please ensure indices are unique.")
        if len(set(new_order)) != len(new_order):
            raise JetError("Duplicate transpose indices found. Please ensure indices are unique.")
        if len(shape) != len(new_order):
            raise JetError("Tensor shape does not match number of indices.")
        if n_out is None:
            if shape_to_size(shape) != n_in:
                raise JetError("Tensor shape does not match given tensor data.")
        else:
            if shape_to_size(shape) != n_in:
                raise JetError("Tensor shape does not match given input tensor data.")
            if shape_to_size(shape) != n_out:
                raise JetError("Tensor shape does not match given output tensor data.")
        if set(current_order) != set(new_order):
            raise JetError("New indices are an invalid permutation of the existing indices")
        if n_out is None:
            if len(shape) == 0:
                raise JetError("Tensor shape cannot be empty.")
            if len(new_order) == 0:
                raise JetError("Tensor indices cannot be empty.")

    def transpose(self, data, shape: Sequence[int], current_order: Sequence[str],
                  new_order: Sequence[str], data_out: Optional[np.ndarray] = None):
        """Return-based form (Permuter.hpp:92-119); with ``data_out`` the
        in-place form (Permuter.hpp:49-79) fills it and returns it."""
        data = np.ascontiguousarray(data)
        shape = [int(s) for s in shape]
        current_order = list(current_order)
        new_order = list(new_order)
        self._validate(data.size, shape, current_order, new_order,
                       None if data_out is None else data_out.size)
        if data.dtype not in (np.complex64, np.complex128):
            raise JetError("Tensor data type must be one of complex64, complex128")
        ctx = self._ctx or nat.default_context()
        code = nat.C64 if data.dtype == np.complex64 else nat.C128
        out = np.empty_like(data)
        if current_order == new_order or len(shape) == 0:
            out[...] = data
        else:
            ids = {s: i for i, s in enumerate(current_order)}
            src = nat.DeviceBuffer(data.nbytes, ctx)
            dst = nat.DeviceBuffer(data.nbytes, ctx)
            src.upload(data)
            nat.check(nat.lib().jet_hip_permute(
                ctx.handle, code, nat.ROW_MAJOR, len(shape), nat.i32(range(len(shape))),
                nat.i64(shape), nat.i32([ids[s] for s in new_order]), src.ptr, dst.ptr, None))
            dst.download(out)
        if data_out is not None:
            data_out[...] = out.reshape(data_out.shape)
            return data_out
        return out

    Transpose = transpose

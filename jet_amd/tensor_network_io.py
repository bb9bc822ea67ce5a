"""Tensor network files: the JSON format of Jet's TensorNetworkSerializer.

``Jet::TensorNetworkSerializer`` / ``TensorNetworkFile``
(include/jet/TensorNetworkIO.hpp:24-293; schema
docs/_static/schema/tensor_network.json).  A file is a JSON object with a
``"tensors"`` array -- one ``[tags, indices, shape, data]`` entry per tensor,
data as ``[re, im]`` pairs in row-major order -- and an optional ``"path"``
of SSA pairs.  Serialization writes keys in sorted order with no whitespace
(indent -1), as the reference's JSON library does; ``col_major=True`` on
loading reverses each tensor's indices and shape (files written from
column-major tensors).  The reference networks
``examples/paper_benchmarks/data_files/m10.json`` and ``m12.json`` load with
this reader.
"""
from __future__ import annotations

import json
from typing import Optional

import numpy as np

from .abort import JetError
from .path_info import PathInfo
from .tensor_network import TensorNetwork


class TensorFileException(JetError):
    """TensorFileException (TensorNetworkIO.hpp:37-46)."""

    def __init__(self, what: str):
        super().__init__("Error parsing tensor network file: " + what)


class TensorNetworkFile:
    """TensorNetworkFile (TensorNetworkIO.hpp:24-30)."""

    def __init__(self, tensors: Optional[TensorNetwork] = None,
                 path: Optional[PathInfo] = None):
        self.tensors = tensors if tensors is not None else TensorNetwork()
        self.path = path


def _dump(obj, indent: int) -> str:
    if indent is None or indent < 0:
        return json.dumps(obj, separators=(",", ":"), sort_keys=True)
    return json.dumps(obj, indent=indent, sort_keys=True)


class TensorNetworkSerializer:
    """serializer(tn[, path]) -> str;  serializer(str[, col_major]) -> file."""

    def __init__(self, indent: int = -1, tensor_type=None, dtype=np.complex128):
        if tensor_type is None:
            from .tensor import Tensor
            tensor_type = Tensor
        self.indent = indent
        self.tensor_type = tensor_type
        self.dtype = np.dtype(dtype)

    def __call__(self, arg, second=None):
        if isinstance(arg, str):
            return self.load(arg, bool(second) if second is not None else False)
        return self.dump(arg, second)

    # -- writing ------------------------------------------------------------
    def dump(self, tn: TensorNetwork, path: Optional[PathInfo] = None) -> str:
        js = {}
        if path is not None:
            js["path"] = [[int(a), int(b)] for (a, b) in path.path]
        js["tensors"] = []
        for node in tn.nodes:
            t = node.tensor
            data = np.asarray(t.get_data()).reshape(-1)
            js["tensors"].append([list(node.tags), list(t.indices), [int(s) for s in t.shape],
                                  [[float(x.real), float(x.imag)] for x in data]])
        return _dump(js, self.indent)

    # -- reading --------------------------------------------------------------
    def load(self, text: str, col_major: bool = False) -> TensorNetworkFile:
        js = json.loads(text)  # invalid JSON raises ValueError, as json::parse throws
        if not isinstance(js, dict):
            raise TensorFileException("root element must be an object.")
        if "tensors" not in js:
            raise TensorFileException("root object must contain 'tensors' key")
        tn = TensorNetwork(self.tensor_type, self.dtype)
        for ti, entry in enumerate(js["tensors"]):
            tags, indices, shape, jdata = entry[0], entry[1], entry[2], entry[3]
            data = np.empty(len(jdata), dtype=self.dtype)
            for i, x in enumerate(jdata):
                try:
                    if not isinstance(x, list) or len(x) < 2:
                        raise IndexError
                    data[i] = complex(float(x[0]), float(x[1]))
                except (IndexError, TypeError, ValueError):
                    raise TensorFileException(
                        f"Invalid element at index {i} of tensor {ti}: Could not parse "
                        f"{json.dumps(x, separators=(',', ':'))} as complex.") from None
            if col_major:
                indices, shape = list(reversed(indices)), list(reversed(shape))
            tn.add_tensor(self.tensor_type(list(indices), [int(s) for s in shape], data,
                                           dtype=self.dtype), list(tags))
        path = PathInfo(tn, js["path"]) if "path" in js else None
        return TensorNetworkFile(tn, path)

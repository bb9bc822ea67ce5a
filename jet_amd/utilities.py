"""Index algebra of the contraction path (include/jet/Utilities.hpp).

Small-vector set operations keep the order of their first argument
(Utilities.hpp:316-384); ravel/unravel are row-major with the first entry most
significant (Utilities.hpp:463-501).
"""
from __future__ import annotations

from typing import List, Sequence

from .abort import JetError

_ALPHABET = "abcdefghijklmnopqrstuvwxyzABCDEFGHIJKLMNOPQRSTUVWXYZ"


def is_pow_2(value: int) -> bool:
    """Utilities.hpp:20-23."""
    return bool(value) and not (value & (value - 1))


def fast_log2(value: int) -> int:
    """Utilities.hpp:37-41 (floor of log2)."""
    return int(value).bit_length() - 1


def generate_string_index(id_: int) -> str:
    """GenerateStringIndex (Utilities.hpp:86-97): a..zA..Z then a0..Z0, a1, ..."""
    div = id_ // len(_ALPHABET)
    prefix = _ALPHABET[id_ % len(_ALPHABET)]
    return prefix + ("" if div == 0 else str(div - 1))


def vector_intersection(v1: Sequence, v2: Sequence) -> List:
    return [x for x in v1 if x in v2]


def vector_union(v1: Sequence, v2: Sequence) -> List:
    return list(v1) + [x for x in v2 if x not in v1]


def vector_subtraction(v1: Sequence, v2: Sequence) -> List:
    return [x for x in v1 if x not in v2]


def vector_disjunctive_union(v1: Sequence, v2: Sequence) -> List:
    return vector_subtraction(vector_union(v1, v2), vector_intersection(v1, v2))


def vector_concatenation(v1: Sequence, v2: Sequence) -> List:
    return list(v1) + list(v2)


def join_string_vector(v: Sequence[str]) -> str:
    return "".join(v)


def shape_to_size(shape: Sequence[int]) -> int:
    size = 1
    for d in shape:
        size *= int(d)
    return size


def unravel_index(index: int, shape: Sequence[int]) -> List[int]:
    """UnravelIndex (Utilities.hpp:463-475)."""
    if shape_to_size(shape) <= index:
        raise JetError("Linear index does not fit in the shape.")
    out = [0] * len(shape)
    for i in range(len(shape) - 1, -1, -1):
        out[i] = index % shape[i]
        index //= shape[i]
    return out


def ravel_index(index: Sequence[int], shape: Sequence[int]) -> int:
    """RavelIndex (Utilities.hpp:486-501)."""
    if len(index) != len(shape):
        raise JetError("Number of index and shape dimensions must match.")
    mult, lin = 1, 0
    for i in range(len(index) - 1, -1, -1):
        if index[i] >= shape[i]:
            raise JetError("Index does not fit in the shape.")
        lin += index[i] * mult
        mult *= shape[i]
    return lin

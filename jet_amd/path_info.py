"""PathInfo: the steps, names, flops and memory of a contraction path.

``Jet::PathInfo`` (include/jet/PathInfo.hpp:19-298) and its Python binding
(python/src/PathInfo.hpp).  A path is a list of SSA pairs: leaf nodes are
steps ``0 .. num_leaves-1`` and the i-th pair creates step ``num_leaves + i``
(PathInfo.hpp:262-297).  A step's contracted indices are the intersection of
its children's tensor indices, its tensor indices their symmetric difference,
and its name the concatenation of its node indices (``"_"`` if none).  Flops
of a step are ``|tensor| * 2 * prod(contracted)``; memory is ``|tensor|``;
indices without an edge (extent 1, or sliced away) count as 1.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence, Tuple

from .abort import JetError
from .utilities import (vector_concatenation, vector_disjunctive_union, vector_intersection,
                        vector_subtraction, vector_union)

MISSING_ID = 2 ** 64 - 1  # PathStepInfo::MISSING_ID (size_t max)


@dataclass
class PathStepInfo:
    """PathStepInfo (PathInfo.hpp:19-47)."""
    id: int
    parent: int = MISSING_ID
    children: Tuple[int, int] = (MISSING_ID, MISSING_ID)
    name: str = ""
    node_indices: List[str] = field(default_factory=list)
    tensor_indices: List[str] = field(default_factory=list)
    tags: List[str] = field(default_factory=list)
    contracted_indices: List[str] = field(default_factory=list)


class PathInfo:
    def __init__(self, tn=None, path: Optional[Sequence[Sequence[int]]] = None):
        self._path: List[Tuple[int, int]] = [(int(a), int(b)) for (a, b) in (path or [])]
        self._steps: List[PathStepInfo] = []
        self._num_leaves = 0
        self._sizes: Dict[str, int] = {}
        if tn is None:
            if self._path:
                raise JetError("Node ID 1 in contraction path pair is invalid.")
            return
        nodes = tn.nodes
        self._num_leaves = len(nodes)
        for node in nodes:
            self._steps.append(PathStepInfo(
                id=node.id, name=node.name, node_indices=list(node.indices),
                tensor_indices=list(node.tensor.indices), tags=list(node.tags)))
        for label, edge in tn.index_to_edge_map.items():
            self._sizes[label] = edge.dim
        for (a, b) in self._path:
            if not 0 <= a < len(self._steps):
                raise JetError("Node ID 1 in contraction path pair is invalid.")
            if not 0 <= b < len(self._steps):
                raise JetError("Node ID 2 in contraction path pair is invalid.")
            self._contract_steps(a, b)

    def _contract_steps(self, a: int, b: int) -> None:
        """PathInfo.hpp:262-297."""
        s1, s2 = self._steps[a], self._steps[b]
        new_id = len(self._steps)
        contracted = vector_intersection(s1.tensor_indices, s2.tensor_indices)
        node_indices = vector_subtraction(
            vector_concatenation(s1.node_indices, s2.node_indices), contracted)
        s1.parent = new_id
        s2.parent = new_id
        self._steps.append(PathStepInfo(
            id=new_id, children=(a, b),
            name="".join(node_indices) if node_indices else "_",
            node_indices=node_indices,
            tensor_indices=vector_disjunctive_union(s1.tensor_indices, s2.tensor_indices),
            tags=vector_union(s1.tags, s2.tags),
            contracted_indices=contracted))

    # --- accessors ---------------------------------------------------------
    @property
    def num_leaves(self) -> int:
        return self._num_leaves

    @property
    def path(self) -> List[Tuple[int, int]]:
        return list(self._path)

    @property
    def steps(self) -> List[PathStepInfo]:
        return self._steps

    @property
    def index_to_size_map(self) -> Dict[str, int]:
        return dict(self._sizes)

    GetNumLeaves = lambda self: self._num_leaves  # noqa: E731
    GetPath = lambda self: list(self._path)  # noqa: E731
    GetSteps = lambda self: self._steps  # noqa: E731
    GetIndexSizes = lambda self: dict(self._sizes)  # noqa: E731

    def _size(self, labels: Sequence[str]) -> float:
        out = 1.0
        for l in labels:
            out *= self._sizes.get(l, 1)
        return out

    def GetPathStepFlops(self, step_id: int) -> float:
        """PathInfo.hpp:157-183."""
        if not 0 <= step_id < len(self._steps):
            raise JetError("Step ID is invalid.")
        if step_id < self._num_leaves:
            return 0.0
        step = self._steps[step_id]
        muls = self._size(step.contracted_indices)
        return self._size(step.tensor_indices) * (muls + muls)

    def GetTotalFlops(self) -> float:
        return float(sum(self.GetPathStepFlops(i)
                         for i in range(self._num_leaves, len(self._steps))))

    def GetPathStepMemory(self, step_id: int) -> float:
        """PathInfo.hpp:211-224."""
        if not 0 <= step_id < len(self._steps):
            raise JetError("Step ID is invalid.")
        return self._size(self._steps[step_id].tensor_indices)

    def GetTotalMemory(self) -> float:
        return float(sum(self.GetPathStepMemory(i) for i in range(len(self._steps))))

    # Python binding names (python/src/PathInfo.hpp).
    step_flops = GetPathStepFlops
    step_memory = GetPathStepMemory
    total_flops = GetTotalFlops
    total_memory = GetTotalMemory

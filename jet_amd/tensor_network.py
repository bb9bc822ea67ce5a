"""TensorNetwork: nodes, edges, tags, slicing and path contraction.

The interface and behaviour of ``Jet::TensorNetwork``
(include/jet/TensorNetwork.hpp:38-535) and of its Python binding
(python/src/TensorNetwork.hpp): node names are the concatenated node
indices or ``"_"``; edges record indices of extent >= 2 and the nodes holding
them; ``SliceIndices`` fixes indices by a raveled value (first index most
significant) and renames them ``label(value)`` in the node; ``Contract``
follows an SSA path (step i creates node ``num_leaves + i``) or, without a
path, contracts every edge held by exactly two nodes and then the scalars.

The network is generic over its tensor type (``jet_amd.Tensor`` by
default, data in device memory): it needs ``indices``/``shape``, the static
``ContractTensors``/``SliceIndex``/``Transpose``/``Reshape`` and
``init_indices_and_shape``.  Whole paths are best run with
``jet_amd.network.DeviceNetwork``, which plans every step and the
intermediate memory at once; ``Contract`` here keeps every intermediate
tensor in its node, as the reference does.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence, Tuple

from .abort import JetError
from .utilities import (unravel_index, vector_disjunctive_union, vector_intersection,
                        vector_union)

Path = List[Tuple[int, int]]


def derive_node_name(indices: Sequence[str]) -> str:
    """TensorNetwork.hpp:522-527."""
    return "".join(indices) if len(indices) else "_"


@dataclass
class Node:
    """TensorNetwork::Node (TensorNetwork.hpp:45-63)."""
    id: int
    name: str
    indices: List[str]
    tags: List[str]
    contracted: bool
    tensor: object


@dataclass
class Edge:
    """TensorNetwork::Edge (TensorNetwork.hpp:68-91): equal when the extent
    and the set of node ids are equal."""
    dim: int
    node_ids: List[int] = field(default_factory=list)

    def __eq__(self, other) -> bool:
        if isinstance(other, tuple):
            other = Edge(other[0], list(other[1]))
        return self.dim == other.dim and set(self.node_ids) == set(other.node_ids)


class TensorNetwork:
    def __init__(self, tensor_type=None, dtype=None):
        if tensor_type is None:
            from .tensor import Tensor
            tensor_type = Tensor
        self.tensor_type = tensor_type
        self.dtype = dtype
        self._nodes: List[Node] = []
        self._edges: Dict[str, Edge] = {}
        self._tags: Dict[str, List[int]] = {}
        self._path: Path = []

    # --- accessors (TensorNetwork.hpp:98-150) ------------------------------
    @property
    def nodes(self) -> List[Node]:
        return self._nodes

    @property
    def index_to_edge_map(self) -> Dict[str, Edge]:
        return self._edges

    @property
    def tag_to_node_id_map(self) -> Dict[str, List[int]]:
        return self._tags

    @property
    def path(self) -> Path:
        return list(self._path)

    @property
    def num_indices(self) -> int:
        return len(self._edges)

    @property
    def num_tensors(self) -> int:
        return len(self._nodes)

    GetNodes = lambda self: self._nodes  # noqa: E731
    GetIndexToEdgeMap = lambda self: self._edges  # noqa: E731
    GetTagToNodesMap = lambda self: self._tags  # noqa: E731
    GetPath = lambda self: list(self._path)  # noqa: E731
    NumIndices = lambda self: len(self._edges)  # noqa: E731
    NumTensors = lambda self: len(self._nodes)  # noqa: E731

    def copy(self) -> "TensorNetwork":
        """The copy constructor: node, edge and tag records are copied, the
        tensors shared (slicing and contraction replace a node's tensor,
        they never modify it)."""
        out = TensorNetwork(self.tensor_type, self.dtype)
        out._nodes = [Node(n.id, n.name, list(n.indices), list(n.tags), n.contracted, n.tensor)
                      for n in self._nodes]
        out._edges = {k: Edge(e.dim, list(e.node_ids)) for k, e in self._edges.items()}
        out._tags = {k: list(v) for k, v in self._tags.items()}
        out._path = list(self._path)
        return out

    __copy__ = copy

    # --- building ------------------------------------------------------------
    def add_tensor(self, tensor, tags: Optional[Sequence[str]] = None) -> int:
        """TensorNetwork::AddTensor (TensorNetwork.hpp:164-180)."""
        tags = list(tags or [])
        node_id = len(self._nodes)
        indices = list(tensor.indices)
        node = Node(node_id, derive_node_name(indices), indices, tags, False, tensor)
        self._nodes.append(node)
        for label, dim in zip(indices, tensor.shape):
            if dim < 2:
                continue
            edge = self._edges.get(label)
            if edge is None:
                self._edges[label] = Edge(int(dim), [node_id])
            else:
                edge.node_ids.append(node_id)
        for tag in tags:
            self._tags.setdefault(tag, []).append(node_id)
        return node_id

    AddTensor = add_tensor

    def slice_indices(self, indices: Sequence[str], value: int) -> None:
        """TensorNetwork::SliceIndices (TensorNetwork.hpp:210-284)."""
        indices = list(indices)
        sizes = []
        node_slices: Dict[int, List[int]] = {}
        for i, label in enumerate(indices):
            edge = self._edges.get(label)
            if edge is None:
                raise JetError("Sliced index does not exist.")
            sizes.append(edge.dim)
            for nid in edge.node_ids:
                node_slices.setdefault(nid, []).append(i)
        values = unravel_index(value, sizes)
        T = self.tensor_type
        for nid, which in node_slices.items():
            node = self._nodes[nid]
            tensor = node.tensor
            for i in which:
                label, v = indices[i], values[i]
                keep = [l for l in tensor.indices if l != label]
                keep_shape = [s for l, s in zip(tensor.indices, tensor.shape) if l != label]
                tensor = T.SliceIndex(tensor, label, v)
                if list(tensor.indices) != keep:
                    tensor = T.Transpose(tensor, keep)
                if keep and list(tensor.shape) != keep_shape:
                    tensor = T.Reshape(tensor, keep_shape)
                # relabels in place: `tensor` is a new tensor made above
                tensor.init_indices_and_shape(keep, keep_shape)
                node.indices = [f"{l}({v})" if l == label else l for l in node.indices]
            node.tensor = tensor
            node.name = derive_node_name(node.indices)
        for label in indices:
            self._edges.pop(label, None)

    SliceIndices = slice_indices

    # --- contraction -----------------------------------------------------------
    def contract(self, path: Optional[Sequence[Sequence[int]]] = None):
        """TensorNetwork::Contract (TensorNetwork.hpp:301-328)."""
        if not self._nodes:
            raise JetError("An empty tensor network cannot be contracted.")
        if path:
            for (a, b) in path:
                if not 0 <= a < len(self._nodes):
                    raise JetError("Node ID 1 in contraction pair is invalid.")
                if not 0 <= b < len(self._nodes):
                    raise JetError("Node ID 2 in contraction pair is invalid.")
                c = self._contract_nodes(a, b)
                self._update_edges(self._nodes[a], self._nodes[b], self._nodes[c])
            self._path = [(int(a), int(b)) for (a, b) in path]
        else:
            self._contract_edges()
            self._contract_scalars()
        return self._nodes[-1].tensor

    Contract = contract

    def _contract_nodes(self, a: int, b: int) -> int:
        """TensorNetwork.hpp:379-409."""
        n1, n2 = self._nodes[a], self._nodes[b]
        t3 = self.tensor_type.ContractTensors(n1.tensor, n2.tensor)
        n1.contracted = True
        n2.contracted = True
        indices = vector_disjunctive_union(n1.indices, n2.indices)
        node = Node(len(self._nodes), derive_node_name(indices), indices,
                    vector_union(n1.tags, n2.tags), False, t3)
        self._nodes.append(node)
        return node.id

    def _update_edges(self, n1: Node, n2: Node, n3: Node) -> None:
        """TensorNetwork.hpp:418-445."""
        for label in n3.indices:
            edge = self._edges.get(label)
            if edge is None:
                continue
            edge.node_ids = [n3.id if x in (n1.id, n2.id) else x for x in edge.node_ids]
        for label in vector_intersection(n1.tensor.indices, n2.tensor.indices):
            self._edges.pop(label, None)

    def _contract_edges(self) -> None:
        """TensorNetwork.hpp:450-492 (edges in insertion order)."""
        for label in list(self._edges.keys()):
            edge = self._edges.get(label)
            if edge is None or len(edge.node_ids) != 2:
                continue
            a, b = edge.node_ids
            c = self._contract_nodes(a, b)
            self._update_edges(self._nodes[a], self._nodes[b], self._nodes[c])
            self._path.append((a, b))

    def _contract_scalars(self) -> None:
        """TensorNetwork.hpp:497-514."""
        ids = [n.id for n in self._nodes if len(n.tensor.indices) == 0]
        if len(ids) >= 2:
            acc = ids[0]
            for nid in ids[1:]:
                self._path.append((acc, nid))
                acc = self._contract_nodes(acc, nid)

    def __str__(self) -> str:
        """operator<< (TensorNetwork.hpp:538-557)."""
        lines = ["Printing Nodes"]
        lines += [f"{n.id} {n.name} {' '.join(n.tags)}" for n in self._nodes]
        lines.append("Printing Edges")
        lines += [f"{k} {e.dim} {' '.join(map(str, e.node_ids))}" for k, e in self._edges.items()]
        return "\n".join(lines)

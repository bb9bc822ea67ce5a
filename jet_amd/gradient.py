"""Adjoint (environment) contraction and gate-angle gradients on the GPU.

For a closed tensor network contracted along an SSA path to the scalar A,
the environment of a tensor T is E_T = dA/dT: the network with T removed,
contracted, its free indices those of T, so that A = sum(E_T * T).  All
environments come from one reverse pass over the contraction tree: a step
z = contract(x, y) passes E_x = contract(E_z, y) (reordered to x's indices)
and E_y = contract(E_z, x) down, with E_root = 1.  Every step is the same
pairwise contraction as the forward pass (permute, complex GEMM, permute),
run by libjet_hip's kernels; only subtrees that hold a wanted tensor are
visited.

Checkpointing: the forward pass keeps every ``checkpoint_every``-th
intermediate (by step) and drops the rest once used; the reverse pass
recomputes a dropped intermediate from its nearest kept descendants when it
needs it.

``AdjointProgram`` plans the same forward and reverse steps as one
contraction DAG (``jet_hip_dag_*``) replayed as a hipGraph; the environments
that are asked for end in one contiguous region read back with a single
copy.  With ``checkpoint_every = 0`` every forward intermediate stays in the
device arena until its reverse-pass use; with ``checkpoint_every = k`` only
every k-th forward intermediate (and the root) is kept, and the DAG itself
recomputes each dropped one, once, from its nearest kept descendants just
before the reverse pass first needs it -- the arena then holds the kept
intermediates plus the recomputed ones still in use.  Either way the
per-step host work of ``AdjointPath`` (planning lookups, pool allocations,
one launch per kernel) is paid once, at capture.

``ParameterizedAmplitude`` applies this to the fSim gates of a grid circuit:
dA/dtheta_k = sum(E_g * dG_g/dtheta_k), and the loss L = |A|^2 has
dL/dtheta_k = 2 Re(conj(A) dA/dtheta_k).  ``AmplitudeLoss`` is the
``torch.autograd.Function`` around it (torch is used for autograd only; the
contractions run in libjet_hip).
"""
from __future__ import annotations

from typing import Dict, List, Optional, Sequence, Set

import numpy as np

from . import _native as nat
from .abort import JetError
from .circuits import Circuit, amplitude_network, fsim
from .network import DeviceDag, DeviceNetwork
from .tensor import Tensor
from .utilities import vector_disjunctive_union


class AdjointPath:
    """Forward contraction and environments of chosen leaves along a path."""

    def __init__(self, indices: Sequence[Sequence[str]], path: Sequence[Sequence[int]],
                 wanted: Sequence[int], checkpoint_every: int = 4,
                 precision: Optional[str] = None):
        if checkpoint_every < 1:
            raise JetError("checkpoint_every must be at least 1.")
        counts: Dict[str, int] = {}
        for ix in indices:
            for l in ix:
                counts[l] = counts.get(l, 0) + 1
        if any(c != 2 for c in counts.values()):
            raise JetError("Environments need a closed network: every index on two tensors.")
        self.n = len(indices)
        self.path = [(int(a), int(b)) for (a, b) in path]
        if len(self.path) != self.n - 1:
            raise JetError("The path must contract the network to one tensor.")
        self.labels: List[List[str]] = [list(ix) for ix in indices]
        self.children: Dict[int, tuple] = {}
        for s, (a, b) in enumerate(self.path):
            z = self.n + s
            self.children[z] = (a, b)
            self.labels.append(vector_disjunctive_union(self.labels[a], self.labels[b]))
        self.root = self.n + len(self.path) - 1
        self.wanted = [int(w) for w in wanted]
        # nodes whose subtree holds a wanted leaf
        self.need: Set[int] = set(self.wanted)
        for s, (a, b) in enumerate(self.path):
            if a in self.need or b in self.need:
                self.need.add(self.n + s)
        self.checkpoint_every = int(checkpoint_every)
        self.precision = precision
        self._leaves: List[Tensor] = []
        self._kept: Dict[int, Tensor] = {}
        self.recomputed_steps = 0

    # -- forward ------------------------------------------------------------
    def forward(self, leaves: Sequence[Tensor]) -> Tensor:
        """Contracts the network; keeps the checkpointed intermediates."""
        if len(leaves) != self.n:
            raise JetError("Wrong number of leaf tensors.")
        self._leaves = list(leaves)
        self._kept = {}
        live: Dict[int, Tensor] = {}
        value = None
        for s, (a, b) in enumerate(self.path):
            va = live.pop(a) if a in live else self._node(a)
            vb = live.pop(b) if b in live else self._node(b)
            value = Tensor.ContractTensors(va, vb, self.precision)
            z = self.n + s
            live[z] = value
            if s % self.checkpoint_every == 0 or z == self.root:
                self._kept[z] = value
        return value

    def _node(self, i: int) -> Tensor:
        if i < self.n:
            return self._leaves[i]
        return self._kept[i]

    def _value(self, i: int) -> Tensor:
        """Value of a node: a leaf, a kept intermediate, or recomputed from
        kept descendants."""
        if i < self.n:
            return self._leaves[i]
        v = self._kept.get(i)
        if v is not None:
            return v
        a, b = self.children[i]
        self.recomputed_steps += 1
        return Tensor.ContractTensors(self._value(a), self._value(b), self.precision)

    # -- reverse --------------------------------------------------------------
    def environments(self) -> Dict[int, Tensor]:
        """Environments of the wanted leaves, each with its leaf's indices."""
        if not self._leaves:
            raise JetError("Run forward() before environments().")
        root = self._kept[self.root]
        env: Dict[int, Tensor] = {self.root: Tensor([], [], np.ones(1, root.dtype),
                                                     dtype=root.dtype, ctx=root.context)}
        out: Dict[int, Tensor] = {}
        for s in range(len(self.path) - 1, -1, -1):
            z = self.n + s
            ez = env.pop(z, None)
            if ez is None:
                continue
            x, y = self.path[s]
            for (u, other) in ((x, y), (y, x)):
                if u not in self.need:
                    continue
                e = Tensor.ContractTensors(ez, self._value(other), self.precision)
                if list(e.indices) != self.labels[u]:
                    e = Tensor.Transpose(e, self.labels[u])
                if u < self.n:
                    out[u] = e
                else:
                    env[u] = e
        return out


def adjoint_dag_steps(indices: Sequence[Sequence[str]], path: Sequence[Sequence[int]],
                      wanted: Sequence[int], checkpoint_every: int = 0) -> Dict[str, object]:
    """Steps of the contraction DAG that gives the root value and the
    environments of ``wanted`` leaves of a closed network along an SSA path
    (no device work).

    DAG leaves are the network's n leaves plus one scalar 1 (node n, the
    root's environment); step s gives DAG node n + 1 + s.  Steps: the forward
    path, then for each forward step z = (x, y) taken in reverse whose
    subtree holds a wanted leaf, E_x = contract(E_z, y) and
    E_y = contract(E_z, x) where needed (the reverse pass of
    ``AdjointPath.environments``).  With ``checkpoint_every`` k > 0 only
    forward steps s with s % k == 0 (and the root) are reused by the reverse
    pass; the value of any other forward node it needs is recomputed by
    extra steps, appended just before its first use, from its nearest kept
    descendants, and each dropped node is recomputed at most once.

    Returns a dict: steps (list of DAG node pairs), root_node, env_node
    ({leaf: DAG node of its environment}), kept (forward node ids reused),
    recomputed ({forward node: DAG node of its recomputation})."""
    if checkpoint_every < 0:
        raise JetError("checkpoint_every must be 0 (keep every intermediate) or positive.")
    counts: Dict[str, int] = {}
    for ix in indices:
        for l in ix:
            counts[l] = counts.get(l, 0) + 1
    if any(c != 2 for c in counts.values()):
        raise JetError("Environments need a closed network: every index on two tensors.")
    n = len(indices)
    path = [(int(a), int(b)) for (a, b) in path]
    if len(path) != n - 1:
        raise JetError("The path must contract the network to one tensor.")
    one = n                              # the scalar leaf
    n_dag_leaves = n + 1
    steps: List[tuple] = []
    dag_id: Dict[int, int] = {i: i for i in range(n)}   # forward node -> DAG node
    for s, (a, b) in enumerate(path):
        steps.append((dag_id[a], dag_id[b]))
        dag_id[n + s] = n_dag_leaves + len(steps) - 1
    root = n + len(path) - 1
    need: Set[int] = set(int(w) for w in wanted)
    for s, (a, b) in enumerate(path):
        if a in need or b in need:
            need.add(n + s)
    kept = {n + s for s in range(len(path)) if checkpoint_every == 0 or s % checkpoint_every == 0}
    kept.add(root)
    recomputed: Dict[int, int] = {}

    def reused(u: int) -> bool:
        return u < n or u in kept

    def value(u: int) -> int:
        """DAG node holding forward node u's value in the reverse pass: the
        leaf, the kept forward result, or its recomputation (appended here,
        children first)."""
        if reused(u):
            return dag_id[u]
        todo = [u]
        while todo:
            v = todo[-1]
            if v in recomputed:
                todo.pop()
                continue
            a, b = path[v - n]
            pending = [c for c in (a, b) if not reused(c) and c not in recomputed]
            if pending:
                todo.extend(pending)
                continue
            todo.pop()
            ra = dag_id[a] if reused(a) else recomputed[a]
            rb = dag_id[b] if reused(b) else recomputed[b]
            steps.append((ra, rb))
            recomputed[v] = n_dag_leaves + len(steps) - 1
        return recomputed[u]

    env: Dict[int, int] = {root: one}
    env_node: Dict[int, int] = {}
    for s in range(len(path) - 1, -1, -1):
        ez = env.pop(n + s, None)
        if ez is None:
            continue
        x, y = path[s]
        for (u, other) in ((x, y), (y, x)):
            if u not in need:
                continue
            v = value(other)
            steps.append((ez, v))
            e = n_dag_leaves + len(steps) - 1
            if u < n:
                env_node[u] = e
            else:
                env[u] = e
    missing = [int(w) for w in wanted if int(w) not in env_node]
    if missing:
        raise JetError("No environment for leaves %s." % missing)
    return {"steps": steps, "root_node": dag_id[root], "env_node": env_node, "kept": kept,
            "recomputed": recomputed}


class AdjointProgram:
    """The root value and the environments of ``wanted`` leaves of a closed
    network along an SSA path, as one planned contraction DAG on the GPU
    (the steps of ``adjoint_dag_steps``, replayed as a hipGraph)."""

    def __init__(self, indices: Sequence[Sequence[str]], shapes, path: Sequence[Sequence[int]],
                 wanted: Sequence[int], dtype=np.complex64, precision: Optional[str] = None,
                 ctx=None, checkpoint_every: int = 0):
        plan = adjoint_dag_steps(indices, path, wanted, checkpoint_every)
        self.n = len(indices)
        self.one = self.n
        self.wanted = [int(w) for w in wanted]
        self.checkpoint_every = int(checkpoint_every)
        self.steps = plan["steps"]
        self.root_node = plan["root_node"]
        self.env_node = plan["env_node"]
        self.recomputed_steps = len(plan["recomputed"])
        self.leaf_labels = [list(ix) for ix in indices]
        self.dag = DeviceDag(self.leaf_labels + [[]], [list(sh) for sh in shapes] + [[]],
                             self.steps, [self.root_node] + [self.env_node[w] for w in self.wanted],
                             dtype=dtype, precision=precision, ctx=ctx)
        self.ctx = self.dag.ctx
        self.dtype = self.dag.dtype
        self._one = nat.DeviceBuffer(16, self.ctx)
        self._one.upload(np.ones(1, self.dtype))

    def run(self, leaf_ptrs: Sequence[int]):
        """(root value, {leaf: environment in the leaf's label order})."""
        if len(leaf_ptrs) != self.n:
            raise JetError("Wrong number of leaf tensors.")
        self.dag.launch(list(leaf_ptrs) + [self._one.ptr])
        vals = self.dag.kept_values()
        root = complex(vals[self.root_node][1].reshape(-1)[0])
        envs = {}
        for w in self.wanted:
            labels, data = vals[self.env_node[w]]
            want = self.leaf_labels[w]
            envs[w] = np.transpose(data, [labels.index(l) for l in want]) if want else data
        return root, envs


class ParameterizedAmplitude:
    """<b| C(theta) |0> of a grid circuit whose fSim gates carry their own
    angles (theta_k, phi_k), with the GPU adjoint gradient of |A|^2."""

    def __init__(self, circ: Circuit, path: Sequence[Sequence[int]], dtype=np.complex64,
                 precision: Optional[str] = None, checkpoint_every: int = 4, ctx=None,
                 method: str = "graph"):
        """method "graph": ``AdjointProgram`` (one captured DAG; with
        ``checkpoint_every`` k > 0 it keeps every k-th forward intermediate
        and recomputes the others inside the DAG, with 0 it keeps them all);
        "eager": ``AdjointPath`` (per-step calls, checkpointed intermediates
        recomputed; k >= 1)."""
        if method not in ("graph", "eager"):
            raise JetError(f"Unknown gradient method '{method}'.")
        self.method = method
        self.circ = circ
        self.ctx = ctx or nat.default_context()
        self.dtype = np.dtype(dtype)
        net = amplitude_network(circ, dtype=np.complex128)
        self.net = net
        self.path = [tuple(p) for p in path]
        n_state = circ.n_qubits
        self.param_leaf: List[int] = []       # leaf id of the k-th parameterised gate
        self.param_slot: List[int] = []       # first angle slot of that gate
        for gi, g in enumerate(circ.gates):
            if g.param_index is not None:
                self.param_leaf.append(n_state + gi)
                self.param_slot.append(g.param_index)
        if not self.param_leaf:
            raise JetError("The circuit has no parameterised gates.")
        self.n_params = 2 * len(self.param_leaf)
        self.theta = np.zeros(self.n_params)
        for gi, g in enumerate(circ.gates):
            if g.param_index is not None:
                self.theta[g.param_index:g.param_index + 2] = g.params
        self.amplitude: Optional[complex] = None
        self._fp64 = None
        self._envs: Optional[Dict[int, np.ndarray]] = None
        if method == "eager":
            self.leaves = [Tensor(ix, sh, d.astype(self.dtype), dtype=self.dtype, ctx=self.ctx)
                           for ix, sh, d in zip(net.indices, net.shapes, net.data)]
            self.adjoint = AdjointPath(net.indices, self.path, self.param_leaf,
                                       checkpoint_every=checkpoint_every, precision=precision)
            return
        self.adjoint = None
        self.program = AdjointProgram(net.indices, net.shapes, self.path, self.param_leaf,
                                      dtype=self.dtype, precision=precision, ctx=self.ctx,
                                      checkpoint_every=checkpoint_every)
        # Every fSim gate in one device buffer (one upload per angle update).
        item = self.dtype.itemsize
        is_param = set(self.param_leaf)
        self._bufs = []
        self._leaf_ptrs = []
        self._param_buf = nat.DeviceBuffer(len(self.param_leaf) * 16 * item, self.ctx)
        pos = {leaf: j for j, leaf in enumerate(self.param_leaf)}
        for i, d in enumerate(net.data):
            if i in is_param:
                self._leaf_ptrs.append(self._param_buf.ptr + pos[i] * 16 * item)
            else:
                b = nat.DeviceBuffer(max(1, d.size) * item, self.ctx)
                b.upload(np.ascontiguousarray(d.reshape(-1), dtype=self.dtype))
                self._bufs.append(b)
                self._leaf_ptrs.append(b.ptr)
        self._upload_params()

    def _upload_params(self) -> None:
        gates = np.stack([fsim(self.theta[slot], self.theta[slot + 1]).reshape(-1)
                          for slot in self.param_slot]).astype(self.dtype)
        self._param_buf.upload(np.ascontiguousarray(gates.reshape(-1)))

    def set_angles(self, theta: Sequence[float]) -> None:
        theta = np.asarray(theta, dtype=np.float64).reshape(-1)
        if theta.size != self.n_params:
            raise JetError("Wrong number of angles.")
        if self.method == "graph":
            if not np.array_equal(theta, self.theta):
                self.theta = theta.copy()
                self._upload_params()
            return
        for leaf, slot in zip(self.param_leaf, self.param_slot):
            if theta[slot] != self.theta[slot] or theta[slot + 1] != self.theta[slot + 1]:
                self.leaves[leaf].set_data(fsim(theta[slot], theta[slot + 1])
                                           .astype(self.dtype).reshape(-1))
        self.theta = theta.copy()

    def forward(self, theta: Optional[Sequence[float]] = None) -> complex:
        if theta is not None:
            self.set_angles(theta)
        if self.method == "graph":
            self.amplitude, self._envs = self.program.run(self._leaf_ptrs)
            return self.amplitude
        root = self.adjoint.forward(self.leaves)
        self.amplitude = complex(root.get_data().reshape(-1)[0])
        return self.amplitude

    def amplitude_gradient(self) -> np.ndarray:
        """dA/dtheta (complex, one per angle) from the environments of the
        fSim gates; needs forward() first."""
        k = len(self.param_leaf)
        if self.method == "graph":
            if self._envs is None:
                raise JetError("Run forward() before amplitude_gradient().")
            E = np.stack([self._envs[leaf].reshape(-1) for leaf in self.param_leaf]
                         ).astype(np.complex128)
        else:
            envs = self.adjoint.environments()
            buf = nat.DeviceBuffer(k * 16 * self.dtype.itemsize, self.ctx)
            L = nat.lib()
            for j, leaf in enumerate(self.param_leaf):
                e = envs[leaf]
                nat.check(L.jet_hip_memcpy(self.ctx.handle,
                                           buf.ptr + j * 16 * self.dtype.itemsize,
                                           e.data_ptr, 16 * self.dtype.itemsize, 2, None, 1))
            host = np.empty(k * 16, self.dtype)
            buf.download(host)
            E = host.astype(np.complex128).reshape(k, 16)
        grad = np.zeros(self.n_params, dtype=np.complex128)
        for j, slot in enumerate(self.param_slot):
            th, ph = self.theta[slot], self.theta[slot + 1]
            c, s = np.cos(th), np.sin(th)
            d_th = np.zeros((4, 4), np.complex128)
            d_th[1, 1] = d_th[2, 2] = -s
            d_th[1, 2] = d_th[2, 1] = -1j * c
            d_ph = np.zeros((4, 4), np.complex128)
            d_ph[3, 3] = -1j * np.exp(-1j * ph)
            grad[slot] = np.sum(E[j] * d_th.reshape(-1))
            grad[slot + 1] = np.sum(E[j] * d_ph.reshape(-1))
        return grad

    def loss_and_gradient(self, theta: Optional[Sequence[float]] = None):
        """L = |A|^2 and dL/dtheta = 2 Re(conj(A) dA/dtheta)."""
        a = self.forward(theta)
        g = self.amplitude_gradient()
        return abs(a) ** 2, 2.0 * np.real(np.conj(a) * g)

    # -- complex128 evaluation of the whole network (FP64 MFMA) ---------------
    def loss_fp64(self, theta: Sequence[float]) -> float:
        """|A|^2 in complex128 through one planned device network (replayed as
        a hipGraph), for finite-difference checks."""
        if self._fp64 is None:
            net = self.net
            dn = DeviceNetwork(net.indices, net.shapes, self.path, [], dtype=np.complex128,
                               precision="fp64", ctx=self.ctx)
            bufs = []
            for d in net.data:
                b = nat.DeviceBuffer(max(1, d.size) * 16, self.ctx)
                b.upload(np.ascontiguousarray(d, dtype=np.complex128))
                bufs.append(b)
            out = nat.DeviceBuffer(16, self.ctx)
            self._fp64 = (dn, bufs, out, np.array(self.theta))
        dn, bufs, out, cur = self._fp64
        theta = np.asarray(theta, dtype=np.float64)
        for leaf, slot in zip(self.param_leaf, self.param_slot):
            if theta[slot] != cur[slot] or theta[slot + 1] != cur[slot + 1]:
                bufs[leaf].upload(fsim(theta[slot], theta[slot + 1]).reshape(-1))
                cur[slot:slot + 2] = theta[slot:slot + 2]
        dn.launch([b.ptr for b in bufs], out.ptr)
        host = np.empty(1, np.complex128)
        out.download(host)
        return float(abs(host[0]) ** 2)


def _torch():
    import torch
    return torch


class AmplitudeLoss:
    """``AmplitudeLoss.apply(theta, problem)``: L = |<b|C(theta)|0>|^2 as a
    torch.autograd.Function (theta a float64 CPU tensor of all angles; the
    forward pass and the adjoint gradient run on the GPU)."""

    _fn = None

    @classmethod
    def function(cls):
        if cls._fn is None:
            torch = _torch()

            class _Fn(torch.autograd.Function):
                @staticmethod
                def forward(ctx, theta, problem):
                    a = problem.forward(theta.detach().cpu().numpy())
                    ctx.problem = problem
                    return torch.tensor(abs(a) ** 2, dtype=torch.float64)

                @staticmethod
                def backward(ctx, grad_out):
                    p = ctx.problem
                    g = 2.0 * np.real(np.conj(p.amplitude) * p.amplitude_gradient())
                    return grad_out * torch.from_numpy(g), None

            cls._fn = _Fn
        return cls._fn

    @classmethod
    def apply(cls, theta, problem: ParameterizedAmplitude):
        return cls.function().apply(theta, problem)

"""Device contraction sequences and the batched slice loop.

``DeviceNetwork`` wraps ``jet_hip_network_*``: a whole contraction path over
device tensors planned once (one ContractionDesc per step, one arena, one
workspace) and executed in order on one stream or replayed as a hipGraph.

``SlicedContraction`` is the MI355X form of the reference's slice loop
(examples/paper_benchmarks/CPU/jet_cpu_m10/jet_sliced.cpp:53-93 driving
TensorNetwork::SliceIndices, TensorNetwork.hpp:210-284, and
TaskBasedContractor::AddContractionTasks / AddReductionTask,
TaskBasedContractor.hpp:162-280).  The sliced indices split in two groups:

  * batch indices stay in the network as modes that no step sums (they become
    batch modes of the GEMMs where both operands hold them), so one execution
    computes every combination of their values, and subtrees that hold none
    of them are computed once -- the sharing that TaskBasedContractor gets
    from identical task names;
  * loop indices are fixed per execution by slicing the leaves that hold them
    (Tensor::SliceIndex, Tensor.hpp:493-526), one execution per value
    combination.

The per-slice results land in one device vector in the reference's slice
order (UnravelIndex over the sliced indices, first index most significant,
Utilities.hpp:463-475; loop indices first, then batch indices) and are
summed by the FP64 compensated reduction.
"""
from __future__ import annotations

import ctypes
import itertools
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

from . import _native as nat
from .abort import JetError

_PREC = {None: nat.PREC_DEFAULT, "default": nat.PREC_DEFAULT, "fp32": nat.PREC_FP32,
         "bf16x3": nat.PREC_BF16X3, "fp64": nat.PREC_FP64, "bf16x3_4m": nat.PREC_BF16X3_4M,
         "f16x3": nat.PREC_F16X3_4M}


class LabelIds:
    def __init__(self):
        self.ids: Dict[str, int] = {}

    def __call__(self, s: str) -> int:
        if s not in self.ids:
            self.ids[s] = len(self.ids)
        return self.ids[s]


def _flatten(indices, shapes, lab):
    ranks = [len(ix) for ix in indices]
    modes = [lab(l) for ix in indices for l in ix]
    exts = [int(s) for sh in shapes for s in sh]
    return ranks, modes, exts


def estimate(indices: Sequence[Sequence[str]], shapes, path, out_labels, dtype=np.complex64,
             precision=None) -> Tuple[float, int, int]:
    """(flops, arena bytes, workspace bytes) of a network plan; no device."""
    lab = LabelIds()
    ranks, modes, exts = _flatten(indices, shapes, lab)
    out = [lab(l) for l in out_labels]
    flat_path = [int(x) for p in path for x in p]
    fl = ctypes.c_double()
    ar, ws = ctypes.c_size_t(), ctypes.c_size_t()
    code = nat.C64 if np.dtype(dtype) == np.complex64 else nat.C128
    nat.check(nat.lib().jet_hip_network_estimate(
        code, _PREC[precision], len(indices), nat.i32(ranks), nat.i32(modes), nat.i64(exts),
        len(path), nat.i32(flat_path), len(out), nat.i32(out), ctypes.byref(fl),
        ctypes.byref(ar), ctypes.byref(ws)))
    return fl.value, ar.value, ws.value


class DeviceNetwork:
    """A planned contraction path over device tensors."""

    def __init__(self, indices: Sequence[Sequence[str]], shapes, path, out_labels,
                 dtype=np.complex64, precision: Optional[str] = None, ctx=None):
        self.ctx = ctx or nat.default_context()
        self.dtype = np.dtype(dtype)
        self.code = nat.C64 if self.dtype == np.complex64 else nat.C128
        lab = LabelIds()
        ranks, modes, exts = _flatten(indices, shapes, lab)
        self.out_labels = list(out_labels)
        out = [lab(l) for l in self.out_labels]
        flat_path = [int(x) for p in path for x in p]
        self.n_leaves = len(indices)
        h = ctypes.c_void_p()
        nat.check(nat.lib().jet_hip_network_create(
            self.ctx.handle, self.code, _PREC[precision], len(indices), nat.i32(ranks),
            nat.i32(modes), nat.i64(exts), len(path), nat.i32(flat_path), len(out), nat.i32(out),
            ctypes.byref(h)))
        self.handle = h
        fl = ctypes.c_double()
        ar, ws = ctypes.c_size_t(), ctypes.c_size_t()
        osz = ctypes.c_int64()
        ns = ctypes.c_int()
        nat.check(nat.lib().jet_hip_network_info(h, ctypes.byref(fl), ctypes.byref(ar),
                                                 ctypes.byref(ws), ctypes.byref(osz),
                                                 ctypes.byref(ns)))
        self.flops, self.arena_bytes, self.workspace_bytes = fl.value, ar.value, ws.value
        self.out_size, self.n_steps = osz.value, ns.value
        self._graph_key = None

    def step_info(self, s: int) -> dict:
        m, n, k, b = (ctypes.c_int64() for _ in range(4))
        kern, perm = ctypes.c_int(), ctypes.c_int()
        nat.check(nat.lib().jet_hip_network_step_info(self.handle, s, ctypes.byref(m),
                                                      ctypes.byref(n), ctypes.byref(k),
                                                      ctypes.byref(b), ctypes.byref(kern),
                                                      ctypes.byref(perm)))
        return dict(M=m.value, N=n.value, K=k.value, batch=b.value,
                    kernel=nat.GEMM_KERNEL_NAMES[kern.value], permutes=perm.value)

    def _ptrs(self, leaves):
        if len(leaves) != self.n_leaves:
            raise JetError(f"Expected {self.n_leaves} leaf pointers, got {len(leaves)}.")
        return (ctypes.c_void_p * self.n_leaves)(*[int(x) for x in leaves])

    def execute(self, leaf_ptrs: Sequence[int], out_ptr: int, stream=None) -> None:
        nat.check(nat.lib().jet_hip_network_execute(self.handle, self._ptrs(leaf_ptrs), out_ptr,
                                                    stream))

    def launch(self, leaf_ptrs: Sequence[int], out_ptr: int, stream=None) -> None:
        """Replay as a hipGraph (captured on first use for these pointers)."""
        key = (tuple(int(x) for x in leaf_ptrs), int(out_ptr))
        if key != self._graph_key:
            nat.check(nat.lib().jet_hip_network_capture(self.handle, self._ptrs(leaf_ptrs),
                                                        out_ptr, stream))
            self._graph_key = key
        nat.check(nat.lib().jet_hip_network_launch(self.handle, stream))

    def close(self) -> None:
        if self.handle:
            nat.check(nat.lib().jet_hip_network_destroy(self.handle))
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DeviceDag:
    """A planned contraction DAG (``jet_hip_dag_*``): steps whose results may
    feed several later steps, e.g. a forward contraction and its adjoint
    pass.  Step s contracts nodes ``steps[s]`` (leaves 0..n-1, step s gives
    node n + s) over all their shared labels; its result holds the first
    operand's other labels, then the second's.  The results of ``keep`` stay
    in one contiguous device region after an execution."""

    def __init__(self, indices: Sequence[Sequence[str]], shapes, steps, keep: Sequence[int],
                 dtype=np.complex64, precision: Optional[str] = None, ctx=None):
        self.ctx = ctx or nat.default_context()
        self.dtype = np.dtype(dtype)
        code = nat.C64 if self.dtype == np.complex64 else nat.C128
        self.labels = LabelIds()
        ranks, modes, exts = _flatten(indices, shapes, self.labels)
        flat = [int(x) for p in steps for x in p]
        self.keep = [int(k) for k in keep]
        self.n_leaves = len(indices)
        h = ctypes.c_void_p()
        nat.check(nat.lib().jet_hip_dag_create(
            self.ctx.handle, code, _PREC[precision], len(indices), nat.i32(ranks), nat.i32(modes),
            nat.i64(exts), len(steps), nat.i32(flat), len(self.keep), nat.i32(self.keep),
            ctypes.byref(h)))
        self.handle = h
        fl = ctypes.c_double()
        ar, ws = ctypes.c_size_t(), ctypes.c_size_t()
        nat.check(nat.lib().jet_hip_network_info(h, ctypes.byref(fl), ctypes.byref(ar),
                                                 ctypes.byref(ws), None, None))
        self.flops, self.arena_bytes, self.workspace_bytes = fl.value, ar.value, ws.value
        names = {v: k for k, v in self.labels.ids.items()}
        # kept values: (device pointer, labels, shape)
        self.values = {}
        for k in self.keep:
            ptr, rank = ctypes.c_void_p(), ctypes.c_int()
            m = (ctypes.c_int32 * 64)()
            e = (ctypes.c_int64 * 64)()
            nat.check(nat.lib().jet_hip_dag_value(h, k, ctypes.byref(ptr), ctypes.byref(rank),
                                                  m, e))
            r = rank.value
            self.values[k] = (int(ptr.value or 0), [names[m[i]] for i in range(r)],
                              [int(e[i]) for i in range(r)])
        first = min(v[0] for v in self.values.values()) if self.values else 0
        last = max(v[0] + max(1, int(np.prod(v[2]))) * self.dtype.itemsize
                   for v in self.values.values()) if self.values else 0
        self._region = (first, last - first)
        self._graph_key = None

    def _ptrs(self, leaves):
        if len(leaves) != self.n_leaves:
            raise JetError(f"Expected {self.n_leaves} leaf pointers, got {len(leaves)}.")
        return (ctypes.c_void_p * self.n_leaves)(*[int(x) for x in leaves])

    def execute(self, leaf_ptrs: Sequence[int], stream=None) -> None:
        nat.check(nat.lib().jet_hip_network_execute(self.handle, self._ptrs(leaf_ptrs), None,
                                                    stream))

    def launch(self, leaf_ptrs: Sequence[int], stream=None) -> None:
        """Replay as a hipGraph (captured on first use for these pointers)."""
        key = tuple(int(x) for x in leaf_ptrs)
        if key != self._graph_key:
            nat.check(nat.lib().jet_hip_network_capture(self.handle, self._ptrs(leaf_ptrs), None,
                                                        stream))
            self._graph_key = key
        nat.check(nat.lib().jet_hip_network_launch(self.handle, stream))

    def kept_values(self) -> Dict[int, Tuple[List[str], np.ndarray]]:
        """{node: (labels, row-major data)} of the kept results, read back in
        one copy of their region."""
        first, size = self._region
        raw = np.empty(size // self.dtype.itemsize, self.dtype)
        nat.check(nat.lib().jet_hip_memcpy(self.ctx.handle, raw.ctypes.data, first, size, 1,
                                           None, 0))
        out = {}
        for k, (ptr, labels, shape) in self.values.items():
            o = (ptr - first) // self.dtype.itemsize
            n = max(1, int(np.prod(shape)))
            out[k] = (labels, raw[o:o + n].reshape(shape) if shape else raw[o:o + 1].reshape(()))
        return out

    def close(self) -> None:
        if self.handle:
            nat.check(nat.lib().jet_hip_network_destroy(self.handle))
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def dag_estimate(indices, shapes, steps, keep, dtype=np.complex64,
                 precision=None) -> Tuple[float, int, int]:
    """(flops, arena bytes, workspace bytes) of a contraction DAG; no device."""
    lab = LabelIds()
    ranks, modes, exts = _flatten(indices, shapes, lab)
    flat = [int(x) for p in steps for x in p]
    fl = ctypes.c_double()
    ar, ws = ctypes.c_size_t(), ctypes.c_size_t()
    code = nat.C64 if np.dtype(dtype) == np.complex64 else nat.C128
    nat.check(nat.lib().jet_hip_dag_estimate(
        code, _PREC[precision], len(indices), nat.i32(ranks), nat.i32(modes), nat.i64(exts),
        len(steps), nat.i32(flat), len(keep), nat.i32([int(k) for k in keep]),
        ctypes.byref(fl), ctypes.byref(ar), ctypes.byref(ws)))
    return fl.value, ar.value, ws.value


def execution_slot(i: int, n_loop: int) -> int:
    """Loop-value combination (and result slot) of loop iteration i of a
    sliced contraction with n_loop combinations: iterations past the last
    combination start the slice loop again, so slot and leaves always agree
    and every write stays inside the n_loop * n_batch results."""
    i, n_loop = int(i), int(n_loop)
    if n_loop < 1:
        raise JetError("A sliced contraction has at least one loop iteration.")
    if i < 0:
        raise JetError(f"Loop iteration {i} is negative.")
    return i % n_loop


def _slice_leaf(labels, shape, data, fixed: Dict[str, int]):
    arr = np.asarray(data).reshape(shape) if shape else np.asarray(data).reshape(())
    keep_l, keep_s, idx = [], [], []
    for l, s in zip(labels, shape):
        if l in fixed:
            idx.append(fixed[l])
        else:
            idx.append(slice(None))
            keep_l.append(l)
            keep_s.append(s)
    out = arr[tuple(idx)] if idx else arr
    return keep_l, keep_s, np.ascontiguousarray(out).reshape(-1)


class SlicedContraction:
    """All slices of a sliced network on one GPU (see the module docstring)."""

    def __init__(self, indices, shapes, data, path, sliced: Sequence[str],
                 batch: Optional[Sequence[str]] = None, dtype=np.complex64,
                 precision: Optional[str] = None, ctx=None, use_graph: bool = True):
        sliced = list(sliced)
        batch = list(sliced if batch is None else batch)
        if not set(batch) <= set(sliced):
            raise JetError("Batch indices must be sliced indices.")
        self.ctx = ctx or nat.default_context()
        self.dtype = np.dtype(dtype)
        self.code = nat.C64 if self.dtype == np.complex64 else nat.C128
        self.loop = [l for l in sliced if l not in batch]
        self.batch = [l for l in sliced if l in batch]
        # Slice order of the results: loop indices first, then batch indices.
        self.sliced = self.loop + self.batch
        self.dims = {}
        for ix, sh in zip(indices, shapes):
            for l, s in zip(ix, sh):
                self.dims[l] = int(s)
        self.n_loop = int(np.prod([self.dims[l] for l in self.loop])) if self.loop else 1
        self.n_batch = int(np.prod([self.dims[l] for l in self.batch])) if self.batch else 1
        self.n_slices = self.n_loop * self.n_batch
        self.use_graph = use_graph

        # Leaves with the loop indices removed (their shapes for the plan).
        loop_set = set(self.loop)
        self.leaf_labels = [[l for l in ix if l not in loop_set] for ix in indices]
        self.leaf_shapes = [[s for l, s in zip(ix, sh) if l not in loop_set]
                            for ix, sh in zip(indices, shapes)]
        self.net = DeviceNetwork(self.leaf_labels, self.leaf_shapes, path, self.batch,
                                 dtype=self.dtype, precision=precision, ctx=self.ctx)
        # Device copies of every leaf variant (one per value combination of
        # the loop indices the leaf holds).
        self._bufs: List[nat.DeviceBuffer] = []
        self._variants: List[Dict[Tuple[int, ...], int]] = []
        self._leaf_loop: List[List[str]] = []
        for ix, sh, d in zip(indices, shapes, data):
            held = [l for l in self.loop if l in ix]
            self._leaf_loop.append(held)
            var = {}
            for vals in itertools.product(*[range(self.dims[l]) for l in held]):
                _, _, arr = _slice_leaf(ix, sh, d, dict(zip(held, vals)))
                arr = arr.astype(self.dtype)
                buf = nat.DeviceBuffer(max(arr.nbytes, 16), self.ctx)
                buf.upload(arr)
                self._bufs.append(buf)
                var[tuple(vals)] = buf.ptr
            self._variants.append(var)
        self.results = nat.DeviceBuffer(max(16, self.n_slices * self.dtype.itemsize), self.ctx)
        self._sum = nat.DeviceBuffer(16, self.ctx)
        self.flops_per_execution = self.net.flops

    def _leaf_ptrs(self, loop_values: Dict[str, int]) -> List[int]:
        return [var[tuple(loop_values[l] for l in held)]
                for var, held in zip(self._variants, self._leaf_loop)]

    def loop_values(self, i: int) -> Dict[str, int]:
        """Values of the loop indices at loop iteration i (last loop index
        fastest, as UnravelIndex orders them)."""
        vals, rem = {}, execution_slot(i, self.n_loop)
        for l in reversed(self.loop):
            vals[l] = rem % self.dims[l]
            rem //= self.dims[l]
        return vals

    def result_ptr(self, i: int) -> int:
        """Device address of the n_batch per-slice results of loop iteration i."""
        return self.results.ptr + execution_slot(i, self.n_loop) * self.n_batch * \
            self.dtype.itemsize

    def run(self, loop_iterations: Optional[Sequence[int]] = None) -> None:
        """Execute the given loop iterations (all by default); iteration i
        runs loop-value combination ``execution_slot(i, n_loop)`` and writes
        its n_batch per-slice results into that slot of the results vector,
        so iterations past the last combination repeat the slice loop."""
        its = range(self.n_loop) if loop_iterations is None else loop_iterations
        for i in its:
            out = self.result_ptr(i)
            ptrs = self._leaf_ptrs(self.loop_values(i))
            if self.use_graph:
                self.net.launch(ptrs, out)
            else:
                self.net.execute(ptrs, out)

    def capture(self, loop_iterations: Sequence[int]):
        """One hipGraph over the given loop iterations (jet_amd.batch.Batch):
        a single launch then runs all of them, each writing its n_batch
        per-slice results (slots as in ``run``)."""
        from .batch import Batch
        ex = [(self.net, self._leaf_ptrs(self.loop_values(i)), self.result_ptr(i))
              for i in loop_iterations]
        return Batch(ex, self.ctx)

    def _count(self, count: Optional[int]) -> int:
        n = self.n_slices if count is None else int(count)
        if not 0 <= n <= self.n_slices:
            raise JetError(f"Slice count {n} is outside [0, {self.n_slices}].")
        return n

    def slice_results(self, count: Optional[int] = None) -> np.ndarray:
        """The first `count` per-slice results (all by default), in slice order."""
        n = self._count(count)
        host = np.empty(n, dtype=self.dtype)
        self.results.download(host)
        return host

    def reduce_enqueue(self, count: Optional[int] = None, stream=None) -> None:
        """Queue the FP64 compensated sum of the first `count` per-slice
        results on the stream (no wait); ``reduced()`` reads it back."""
        n = self._count(count)
        nat.check(nat.lib().jet_hip_reduce_compensated(self.ctx.handle, self.code, n,
                                                       self.results.ptr, self._sum.ptr, stream))

    def reduced(self) -> complex:
        """The last queued compensated sum (waits for it)."""
        host = np.empty(2, dtype=np.float64)
        self._sum.download(host)
        return complex(host[0], host[1])

    def reduce(self, count: Optional[int] = None) -> complex:
        """FP64 compensated sum of the first `count` per-slice results."""
        self.reduce_enqueue(count)
        return self.reduced()


def choose_batch(indices, shapes, path, sliced: Sequence[str], budget_bytes: float,
                 dtype=np.complex64, precision=None) -> List[str]:
    """Largest set of sliced indices to batch whose plan fits in the budget:
    start from all of them and move to the loop, one at a time, the index
    whose removal shrinks the plan most."""
    dims = {l: int(s) for ix, sh in zip(indices, shapes) for l, s in zip(ix, sh)}

    def cost(batch):
        loop = set(sliced) - set(batch)
        li = [[l for l in ix if l not in loop] for ix in indices]
        ls = [[s for l, s in zip(ix, sh) if l not in loop] for ix, sh in zip(indices, shapes)]
        fl, ar, ws = estimate(li, ls, path, [l for l in sliced if l in batch], dtype, precision)
        return ar + ws

    batch = list(sliced)
    while batch and cost(batch) > budget_bytes:
        best = min(batch, key=lambda l: cost([x for x in batch if x != l]))
        batch.remove(best)
    return [l for l in sliced if l in batch]

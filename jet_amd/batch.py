"""Uneven batches: many contractions of different networks in one launch.

The reference runs independent contractions (amplitudes of different
circuits, slices of a sliced network) as tasks of one Taskflow graph
(TaskBasedContractor::AddContractionTasks per network, AddTaskflow to
compose graphs, TaskBasedContractor.hpp:147-248).  Here every job is a
planned device network (``DeviceNetwork``: per-step permutation and GEMM
plans, its own arena), and all executions of all jobs are captured into one
hipGraph (``jet_hip_batch_capture``) that one ``launch()`` replays.

Jobs:
  * ``AmplitudeJob``: amplitudes <b|C|0> of one circuit for several
    bitstrings.  Two forms:
      - "each": one network, one execution per bitstring (only the output
        vectors differ between executions);
      - "stacked": the bitstring index is a batch label on every output
        vector, so one execution computes all of them and every step's GEMM
        is batched over the bitstrings (the better form for small circuits,
        whose single-amplitude GEMMs are too small for the MFMA tiles);
      - "stacked_pow2": the same with the stack padded to a power of two
        (the last bitstring repeated, its extra results dropped), so that
        every extent is a power of two: bit permutations and the fused
        small-step pass instead of the generic odometer permutation.
  * ``SliceJob``: a contiguous range of slices of a sliced network
    (``SlicedContraction``'s executions), summed afterwards by the FP64
    compensated reduction.

Autotuning: for each amplitude job whose stacked plan fits the memory
budget, the forms are captured alone and timed (median of a few replays);
the fastest one is kept.  ``tuning`` records the measured times.
"""
from __future__ import annotations

import ctypes
import time
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

from . import _native as nat
from .abort import JetError
from .circuits import Circuit, amplitude_network
from .network import DeviceNetwork, SlicedContraction, estimate


class Batch:
    """One hipGraph over (network, leaf pointers, output pointer) executions."""

    def __init__(self, executions: Sequence[Tuple[DeviceNetwork, Sequence[int], int]], ctx=None):
        if not executions:
            raise JetError("A batch needs at least one execution.")
        self.ctx = ctx or nat.default_context()
        n = len(executions)
        nets = (ctypes.c_void_p * n)(*[e[0].handle.value for e in executions])
        self._leaf_arrays = [(ctypes.c_void_p * len(e[1]))(*[int(p) for p in e[1]])
                             for e in executions]
        leaves = (ctypes.POINTER(ctypes.c_void_p) * n)(
            *[ctypes.cast(a, ctypes.POINTER(ctypes.c_void_p)) for a in self._leaf_arrays])
        outs = (ctypes.c_void_p * n)(*[int(e[2]) for e in executions])
        h = ctypes.c_void_p()
        nat.check(nat.lib().jet_hip_batch_capture(self.ctx.handle, n, nets, leaves, outs, None,
                                                  ctypes.byref(h)))
        self.handle = h
        self._keep = [e[0] for e in executions]  # networks outlive the graph

    def launch(self, stream=None) -> None:
        nat.check(nat.lib().jet_hip_batch_launch(self.handle, stream))

    def close(self) -> None:
        if self.handle:
            nat.check(nat.lib().jet_hip_batch_destroy(self.handle))
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _upload(arr: np.ndarray, ctx) -> nat.DeviceBuffer:
    buf = nat.DeviceBuffer(max(16, arr.nbytes), ctx)
    buf.upload(np.ascontiguousarray(arr))
    return buf


@dataclass
class AmplitudeJob:
    circuit: Circuit
    bitstrings: Sequence[Sequence[int]]
    path: Sequence[Tuple[int, int]]
    form: str = "auto"      # "auto" (tuned), "each", "stacked" or "stacked_pow2"


class _Amplitudes:
    """Device state of one amplitude job in one form."""

    def __init__(self, job: AmplitudeJob, form: str, dtype, precision, ctx):
        self.form = form
        self.dtype = np.dtype(dtype)
        self.ctx = ctx
        circ = job.circuit
        nq = circ.n_qubits
        bits = [tuple(int(x) for x in b) for b in job.bitstrings]
        self.n = len(bits)
        if form == "stacked_pow2":
            n2 = 1 << max(0, (self.n - 1).bit_length())
            bits = bits + [bits[-1]] * (n2 - self.n)
            form = "stacked"
        self.n_stack = len(bits)
        net = amplitude_network(circ, bits[0], dtype=np.complex128)
        n_leaves = len(net.indices)
        first_out = n_leaves - nq           # output vectors are the last nq leaves
        self.bufs: List[nat.DeviceBuffer] = [_upload(d.astype(self.dtype), ctx)
                                             for d in net.data[:first_out]]
        self.out = nat.DeviceBuffer(max(16, self.n_stack * self.dtype.itemsize), ctx)
        if form == "each":
            self.net = DeviceNetwork(net.indices, net.shapes, job.path, [], dtype=self.dtype,
                                     precision=precision, ctx=ctx)
            self.execs = []
            for i, b in enumerate(bits):
                outs = []
                for q in range(nq):
                    v = np.zeros(2, self.dtype)
                    v[b[q]] = 1
                    outs.append(_upload(v, ctx))
                self.bufs.extend(outs)
                ptrs = [x.ptr for x in self.bufs[:first_out]] + [x.ptr for x in outs]
                self.execs.append((self.net, ptrs, self.out.ptr + i * self.dtype.itemsize))
        elif form == "stacked":
            lab = "%amplitude"
            idx = [list(ix) for ix in net.indices]
            shp = [list(s) for s in net.shapes]
            outs = []
            for q in range(nq):
                k = first_out + q
                idx[k] = [lab] + idx[k]
                shp[k] = [self.n_stack] + shp[k]
                v = np.zeros((self.n_stack, 2), self.dtype)
                for i, b in enumerate(bits):
                    v[i, b[q]] = 1
                outs.append(_upload(v.reshape(-1), ctx))
            self.bufs.extend(outs)
            self.net = DeviceNetwork(idx, shp, job.path, [lab], dtype=self.dtype,
                                     precision=precision, ctx=ctx)
            self.execs = [(self.net, [x.ptr for x in self.bufs], self.out.ptr)]
        else:
            raise JetError(f"Unknown amplitude form '{form}'.")

    def results(self) -> np.ndarray:
        host = np.empty(self.n_stack, self.dtype)
        self.out.download(host)
        return host[:self.n]


def stacked_bytes(job: AmplitudeJob, dtype=np.complex64, precision=None) -> int:
    """Arena + workspace bytes of the stacked form (no device needed)."""
    circ = job.circuit
    nq = circ.n_qubits
    net = amplitude_network(circ, job.bitstrings[0], dtype=np.complex128)
    idx = [list(ix) for ix in net.indices]
    shp = [list(s) for s in net.shapes]
    n = len(job.bitstrings)
    for q in range(nq):
        k = len(idx) - nq + q
        idx[k] = ["%amplitude"] + idx[k]
        shp[k] = [n] + shp[k]
    _, ar, ws = estimate(idx, shp, job.path, ["%amplitude"], dtype, precision)
    return ar + ws


@dataclass
class SliceJob:
    indices: Sequence[Sequence[str]]
    shapes: Sequence[Sequence[int]]
    data: Sequence[np.ndarray]
    path: Sequence[Tuple[int, int]]
    sliced: Sequence[str]
    batch: Sequence[str]
    first_execution: int = 0
    executions: int = 1


class UnevenBatch:
    """All jobs in one hipGraph; ``launch()`` runs every one of them."""

    def __init__(self, jobs: Sequence, dtype=np.complex64, precision: Optional[str] = None,
                 ctx=None, budget_bytes: float = 64e9, tune_reps: int = 3):
        self.ctx = ctx or nat.default_context()
        self.dtype = np.dtype(dtype)
        self.precision = precision
        self.jobs = list(jobs)
        self.parts: List = []
        self.tuning: List[Dict[str, float]] = []
        for job in self.jobs:
            if isinstance(job, AmplitudeJob):
                self.parts.append(self._amplitude_part(job, budget_bytes, tune_reps))
            elif isinstance(job, SliceJob):
                sc = SlicedContraction(job.indices, job.shapes, job.data, job.path, job.sliced,
                                       batch=job.batch, dtype=self.dtype,
                                       precision=precision, ctx=self.ctx, use_graph=False)
                if job.first_execution + job.executions > sc.n_loop:
                    raise JetError("Slice range beyond the network's slices.")
                self.parts.append(sc)
                self.tuning.append({})
            else:
                raise JetError("Unknown job type.")
        self.batch = Batch(self._executions(), self.ctx)

    def _amplitude_part(self, job: AmplitudeJob, budget: float, reps: int):
        n = len(job.bitstrings)
        pow2 = n & (n - 1) == 0
        forms = [job.form] if job.form != "auto" else \
            (["each", "stacked"] + ([] if pow2 else ["stacked_pow2"]))
        if job.form == "auto":
            padded = AmplitudeJob(job.circuit,
                                  list(job.bitstrings) +
                                  [job.bitstrings[-1]] * ((1 << max(0, (n - 1).bit_length())) - n),
                                  job.path)
            if stacked_bytes(job, self.dtype, self.precision) > budget:
                forms = [f for f in forms if f == "each"]
            elif "stacked_pow2" in forms and \
                    stacked_bytes(padded, self.dtype, self.precision) > budget:
                forms.remove("stacked_pow2")
        if len(forms) == 1:
            self.tuning.append({})
            return _Amplitudes(job, forms[0], self.dtype, self.precision, self.ctx)
        times, best, best_t = {}, None, None
        for f in forms:
            part = _Amplitudes(job, f, self.dtype, self.precision, self.ctx)
            b = Batch(part.execs, self.ctx)
            samples = []
            for _ in range(reps + 1):
                self.ctx.synchronize()
                t0 = time.perf_counter()
                b.launch()
                self.ctx.synchronize()
                samples.append(time.perf_counter() - t0)
            b.close()
            t = float(np.median(samples[1:]))
            times[f] = t * 1e3
            if best_t is None or t < best_t:
                best, best_t = part, t
        self.tuning.append(times)
        return best

    def _executions(self):
        ex = []
        for job, part in zip(self.jobs, self.parts):
            if isinstance(part, _Amplitudes):
                ex.extend(part.execs)
            else:
                sc = part
                for i in range(job.first_execution, job.first_execution + job.executions):
                    ex.append((sc.net, sc._leaf_ptrs(sc.loop_values(i)), sc.result_ptr(i)))
        return ex

    def launch(self) -> None:
        self.batch.launch()

    def amplitudes(self, j: int) -> np.ndarray:
        part = self.parts[j]
        if not isinstance(part, _Amplitudes):
            raise JetError("Job is not an amplitude job.")
        return part.results()

    def slice_sum(self, j: int) -> complex:
        """FP64 compensated sum of the job's executed slices."""
        job, sc = self.jobs[j], self.parts[j]
        first = job.first_execution * sc.n_batch
        count = job.executions * sc.n_batch
        out = nat.DeviceBuffer(16, self.ctx)
        nat.check(nat.lib().jet_hip_reduce_compensated(
            self.ctx.handle, sc.code, count, sc.results.ptr + first * self.dtype.itemsize,
            out.ptr, None))
        host = np.empty(2, np.float64)
        out.download(host)
        return complex(host[0], host[1])

    def slice_results(self, j: int) -> np.ndarray:
        job, sc = self.jobs[j], self.parts[j]
        first = job.first_execution * sc.n_batch
        return sc.slice_results(first + job.executions * sc.n_batch)[first:]

    def close(self) -> None:
        """Frees the graph and every job's device memory now."""
        if self.batch is not None:
            self.batch.close()
            self.batch = None
        for part in self.parts:
            part.net.close()
        self.parts = []

    def forms(self) -> List[str]:
        return [p.form if isinstance(p, _Amplitudes) else "slices" for p in self.parts]

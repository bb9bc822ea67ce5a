"""TaskBasedContractor: named contraction tasks with sharing and reduction.

``Jet::TaskBasedContractor`` (include/jet/TaskBasedContractor.hpp:30-427) and
its Python binding (python/src/TaskBasedContractor.hpp), over GPU tensors.

* Every contraction step of every added path is a task named
  ``"{step id}:{step name}"`` (TaskBasedContractor.hpp:367-370); the last step
  of the n-th added path is named ``...:results[n]`` (192-197).
* A task whose name already exists is not added again and counts as shared
  (218-222): the slices of a sliced network share every step that does not
  involve a sliced index, because such steps keep their name.
* Flops and memory are those of PathInfo, summed over the tasks actually
  added (224-225).
* ``AddReductionTask`` sums all results with ``AddTensor`` starting from the
  zero tensor (258-280); ``AddDeletionTasks`` frees every tensor once all
  tasks that read it have run (290-315).

Execution.  The reference runs the task graph on a Taskflow thread pool and
each task runs its contraction to completion.  Here every task only enqueues
GPU work (kernels on the context's stream, which orders them), so one host
thread walks the graph in dependency order; deletions return memory to the
stream-ordered pool, which is safe once the reading kernels are enqueued.
``Contract()`` returns a ``concurrent.futures.Future`` that completes when
the GPU has finished (``contract()`` waits for it).
"""
from __future__ import annotations

import concurrent.futures
import threading
from typing import Callable, Dict, List, Optional, Set

import numpy as np

from .path_info import PathInfo


class Task:
    """A node of the task graph (the role of tf::Task)."""

    __slots__ = ("_name", "fn", "successors", "n_predecessors")

    def __init__(self, name: str, fn: Callable[[], None]):
        self._name = name
        self.fn = fn
        self.successors: List["Task"] = []
        self.n_predecessors = 0

    def name(self) -> str:
        return self._name

    def precede(self, other: "Task") -> None:
        self.successors.append(other)
        other.n_predecessors += 1

    def __repr__(self) -> str:
        return f"Task({self._name!r})"


class TaskBasedContractor:
    def __init__(self, num_threads: Optional[int] = None, tensor_type=None,
                 dtype=np.complex128):
        if tensor_type is None:
            from .tensor import Tensor
            tensor_type = Tensor
        self.tensor_type = tensor_type
        self.dtype = np.dtype(dtype)
        self._tasks: List[Task] = []
        self._name_to_task: Dict[str, Task] = {}
        self._name_to_tensor: Dict[str, object] = {}
        self._name_to_parents: Dict[str, Set[str]] = {}
        self._result_tasks: List[Task] = []
        self._results: List[object] = []
        self._reduction_result = tensor_type(dtype=self.dtype)
        self._flops = 0.0
        self._memory = 0.0
        self._reduced = False

    # --- accessors (TaskBasedContractor.hpp:64-147) --------------------------
    @property
    def name_to_task_map(self) -> Dict[str, Task]:
        return self._name_to_task

    @property
    def name_to_tensor_map(self) -> Dict[str, object]:
        return self._name_to_tensor

    @property
    def name_to_parents_map(self) -> Dict[str, Set[str]]:
        return self._name_to_parents

    @property
    def results(self) -> List[object]:
        return self._results

    @property
    def reduction_result(self):
        return self._reduction_result

    @property
    def flops(self) -> float:
        return self._flops

    @property
    def memory(self) -> float:
        return self._memory

    GetNameToTaskMap = lambda self: self._name_to_task  # noqa: E731
    GetNameToTensorMap = lambda self: self._name_to_tensor  # noqa: E731
    GetNameToParentsMap = lambda self: self._name_to_parents  # noqa: E731
    GetResults = lambda self: self._results  # noqa: E731
    GetReductionResult = lambda self: self._reduction_result  # noqa: E731
    GetFlops = lambda self: self._flops  # noqa: E731
    GetMemory = lambda self: self._memory  # noqa: E731

    # --- building the task graph -------------------------------------------
    def _emplace(self, name: str, fn: Callable[[], None]) -> Task:
        task = Task(name, fn)
        self._tasks.append(task)
        return task

    def add_contraction_tasks(self, tn, path_info: PathInfo) -> int:
        """AddContractionTasks (TaskBasedContractor.hpp:162-248): returns the
        number of shared (already present) tasks."""
        path = path_info.path
        steps = path_info.steps
        if not path:
            return 0
        nodes = tn.nodes
        num_leaves = len(nodes)
        result_id = len(self._results)
        self._results.append(None)
        shared = 0
        for i, (id1, id2) in enumerate(path):
            s1, s2, s3 = steps[id1], steps[id2], steps[num_leaves + i]
            n1, n2 = f"{s1.id}:{s1.name}", f"{s2.id}:{s2.name}"
            n3 = f"{s3.id}:{s3.name}"
            last = i == len(path) - 1
            if last:
                n3 += f":results[{result_id}]"
            self._name_to_parents.setdefault(n1, set()).add(n3)
            self._name_to_parents.setdefault(n2, set()).add(n3)
            if id1 < num_leaves and n1 not in self._name_to_tensor:
                self._name_to_tensor[n1] = nodes[id1].tensor
            if id2 < num_leaves and n2 not in self._name_to_tensor:
                self._name_to_tensor[n2] = nodes[id2].tensor
            self._name_to_tensor.setdefault(n3, None)
            if n3 in self._name_to_task:
                shared += 1
                continue
            self._flops += path_info.GetPathStepFlops(s3.id)
            self._memory += path_info.GetPathStepMemory(s3.id)
            task3 = self._add_contraction_task(n1, n2, n3)
            if id1 >= num_leaves:
                self._name_to_task[n1].precede(task3)
            if id2 >= num_leaves:
                self._name_to_task[n2].precede(task3)
            if last:
                self._add_storage_task(n3, result_id)
        return shared

    def _add_contraction_task(self, n1: str, n2: str, n3: str) -> Task:
        T = self.tensor_type

        def run():
            self._name_to_tensor[n3] = T.ContractTensors(self._name_to_tensor[n1],
                                                         self._name_to_tensor[n2])

        task = self._emplace(n3, run)
        self._name_to_task[n3] = task
        return task

    def _add_storage_task(self, name: str, result_id: int) -> None:
        """TaskBasedContractor.hpp:407-426."""
        def run():
            self._results[result_id] = self._name_to_tensor[name]

        task = self._emplace(f"{name}:storage[{result_id}]", run)
        self._name_to_task[name].precede(task)
        self._result_tasks.append(task)

    def add_reduction_task(self) -> int:
        """AddReductionTask (TaskBasedContractor.hpp:258-280): one per
        contractor; later calls return 0."""
        if self._reduced:
            return 0
        self._reduced = True

        def run():
            acc = self._reduction_result
            for r in self._results:
                acc = acc.add_tensor(r)
            self._reduction_result = acc

        task = self._emplace("reduce", run)
        for t in self._result_tasks:
            t.precede(task)
        return 1

    def add_deletion_tasks(self) -> int:
        """AddDeletionTasks (TaskBasedContractor.hpp:290-315)."""
        count = 0
        for name, parents in self._name_to_parents.items():
            if not parents:
                continue

            def run(name=name):
                self._name_to_tensor[name] = None

            task = self._emplace(f"{name}:delete", run)
            count += 1
            for p in parents:
                pt = self._name_to_task.get(p)
                if pt is not None:
                    pt.precede(task)
        return count

    # --- execution ---------------------------------------------------------------
    def _run(self) -> None:
        pending = {id(t): t.n_predecessors for t in self._tasks}
        ready = [t for t in self._tasks if t.n_predecessors == 0]
        done = 0
        while ready:
            nxt = []
            for t in ready:
                t.fn()
                done += 1
                for s in t.successors:
                    pending[id(s)] -= 1
                    if pending[id(s)] == 0:
                        nxt.append(s)
            ready = nxt
        if done != len(self._tasks):
            raise RuntimeError("Task graph has a cycle.")
        for r in self._results:
            ctx = getattr(r, "context", None)
            if ctx is not None:
                ctx.synchronize()
                break

    def Contract(self) -> concurrent.futures.Future:
        """Contract (TaskBasedContractor.hpp:322): runs the task graph on a
        thread of its own and returns a future that completes when every
        result is ready."""
        fut: concurrent.futures.Future = concurrent.futures.Future()

        def body():
            if not fut.set_running_or_notify_cancel():
                return
            try:
                self._run()
            except BaseException as e:  # handed to the waiting caller
                fut.set_exception(e)
            else:
                fut.set_result(None)

        threading.Thread(target=body, name="jet-task-based-contractor", daemon=True).start()
        return fut

    def contract(self) -> None:
        self.Contract().result()

    AddContractionTasks = add_contraction_tasks
    AddReductionTask = add_reduction_task
    AddDeletionTasks = add_deletion_tasks

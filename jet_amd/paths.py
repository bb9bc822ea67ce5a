"""Contraction paths and slice selection for tensor networks.

Jet itself has no path finder in C++: its Python layer samples random paths
(python/jet/interpreter.py:533-618) and the paper benchmarks load cotengra
paths from JSON.  This module provides

  * greedy_path: greedy pairwise contraction order (at each step contract the
    connected pair whose result minimises size(out) - size(a) - size(b)),
    optionally randomised with Gumbel noise and repeated, keeping the path
    with the fewest flops;
  * path_stats: per-step flops / sizes in the PathInfo conventions
    (PathInfo.hpp:157-224);
  * choose_slices: greedy choice of the indices to slice
    (TensorNetwork::SliceIndices, TensorNetwork.hpp:210-284): repeatedly the
    index whose removal most reduces the largest intermediate, then the flops.

Paths use SSA ids like PathInfo: leaves are 0..n-1 and step i creates n+i.
"""
from __future__ import annotations

import heapq
import math
from dataclasses import dataclass
from typing import Dict, List, Optional, Sequence, Set, Tuple

import numpy as np

Path = List[Tuple[int, int]]


def _log2size(labels, dims) -> float:
    return sum(math.log2(dims[l]) for l in labels)


def greedy_path(indices: Sequence[Sequence[str]], dims: Dict[str, int],
                open_indices: Sequence[str] = (), seed: Optional[int] = None,
                temperature: float = 0.0) -> Path:
    """One greedy contraction path (SSA ids)."""
    rng = np.random.default_rng(seed)
    n = len(indices)
    open_set = set(open_indices)
    alive: Dict[int, frozenset] = {i: frozenset(ix) for i, ix in enumerate(indices)}
    where: Dict[str, Set[int]] = {}
    for i, ix in alive.items():
        for l in ix:
            where.setdefault(l, set()).add(i)
    size = {i: _log2size(ix, dims) for i, ix in alive.items()}

    def result_labels(a: int, b: int) -> frozenset:
        la, lb = alive[a], alive[b]
        keep = set()
        for l in la | lb:
            others = len(where[l] - {a, b})
            if l in open_set or others > 0 or not (l in la and l in lb):
                keep.add(l)
        return frozenset(keep)

    def cost(a: int, b: int) -> float:
        out = result_labels(a, b)
        c = 2.0 ** _log2size(out, dims) - 2.0 ** size[a] - 2.0 ** size[b]
        if temperature > 0:
            c -= temperature * math.log(-math.log(rng.uniform(1e-12, 1.0))) * (
                2.0 ** max(size[a], size[b]))
        return c

    heap: List[Tuple[float, int, int, int]] = []
    counter = 0

    def push_neighbours(a: int):
        nonlocal counter
        seen = set()
        for l in alive[a]:
            for b in where[l]:
                if b != a and b not in seen:
                    seen.add(b)
                    x, y = (a, b) if a < b else (b, a)
                    heapq.heappush(heap, (cost(x, y), counter, x, y))
                    counter += 1

    for i in list(alive):
        push_neighbours(i)

    path: Path = []
    next_id = n
    while len(alive) > 1:
        pair = None
        while heap:
            _, _, a, b = heapq.heappop(heap)
            if a in alive and b in alive:
                pair = (a, b)
                break
        if pair is None:
            # Disconnected components: contract the two smallest tensors.
            a, b = sorted(alive, key=lambda i: (size[i], i))[:2]
            pair = (a, b)
        a, b = pair
        out = result_labels(a, b)
        for l in alive[a] | alive[b]:
            where[l].discard(a)
            where[l].discard(b)
            if l in out:
                where[l].add(next_id)
        del alive[a], alive[b]
        alive[next_id] = out
        size[next_id] = _log2size(out, dims)
        path.append((a, b))
        push_neighbours(next_id)
        next_id += 1
    return path


@dataclass
class StepStats:
    out: List[frozenset]          # labels of each step's result
    contracted: List[frozenset]   # labels summed at each step
    log2_size: np.ndarray         # log2 |result| per step
    log2_flops: np.ndarray        # log2 (|result| * prod contracted) per step
    total_flops: float            # PathInfo convention: sum |out| * 2 * prod contracted
    max_log2_size: float


def path_stats(indices: Sequence[Sequence[str]], dims: Dict[str, int], path: Path,
               sliced: Sequence[str] = (), open_indices: Sequence[str] = ()) -> StepStats:
    """Per-step results of contracting along `path` with `sliced` indices
    removed (PathInfo.hpp:262-297: result = symmetric difference)."""
    sl = set(sliced)
    nodes: List[frozenset] = [frozenset(l for l in ix if l not in sl) for ix in indices]
    outs, cons, ls, lf = [], [], [], []
    total = 0
    for (a, b) in path:
        la, lb = nodes[a], nodes[b]
        c = (la & lb) - set(open_indices)
        o = (la | lb) - c
        nodes.append(o)
        outs.append(o)
        cons.append(c)
        size_o = math.prod(dims[l] for l in o)
        size_c = math.prod(dims[l] for l in c)
        ls.append(math.log2(size_o))
        lf.append(math.log2(size_o * size_c))
        total += size_o * 2 * size_c
    ls = np.array(ls)
    lf = np.array(lf)
    total = float(total)
    return StepStats(outs, cons, ls, lf, total, float(ls.max()) if len(ls) else 0.0)


def best_greedy_path(indices, dims, trials: int = 16, seed: int = 0,
                     temperature: float = 0.02, open_indices=()) -> Path:
    """Deterministic greedy plus `trials` randomised runs; fewest flops wins."""
    best, best_cost = None, None
    for t in range(trials + 1):
        p = greedy_path(indices, dims, open_indices, seed=seed + t,
                        temperature=0.0 if t == 0 else temperature)
        st = path_stats(indices, dims, p, open_indices=open_indices)
        key = (st.total_flops, st.max_log2_size)
        if best_cost is None or key < best_cost:
            best, best_cost = p, key
    return best


def choose_slices(indices, dims, path: Path, n_slices: int,
                  candidates: Optional[Sequence[str]] = None) -> List[str]:
    """Greedy slice selection: each round removes the index that minimises
    (largest intermediate, total flops) of the per-slice contraction."""
    st = path_stats(indices, dims, path)
    labels = sorted(candidates if candidates is not None else
                    {l for ix in indices for l in ix})
    col = {l: j for j, l in enumerate(labels)}
    nstep = len(path)
    O = np.zeros((nstep, len(labels)), dtype=np.float64)   # label in result
    F = np.zeros((nstep, len(labels)), dtype=np.float64)   # label in result or contracted
    for s in range(nstep):
        for l in st.out[s]:
            if l in col:
                O[s, col[l]] = math.log2(dims[l])
                F[s, col[l]] = math.log2(dims[l])
        for l in st.contracted[s]:
            if l in col:
                F[s, col[l]] = math.log2(dims[l])
    ls = st.log2_size.copy()
    lf = st.log2_flops.copy()
    chosen: List[str] = []
    avail = np.ones(len(labels), dtype=bool)
    for _ in range(n_slices):
        new_max = np.max(ls[:, None] - O, axis=0)
        new_flops = np.sum(2.0 ** (lf[:, None] - F), axis=0)
        key = np.where(avail, new_max * 1e6 + np.log2(new_flops + 1.0), np.inf)
        j = int(np.argmin(key))
        chosen.append(labels[j])
        avail[j] = False
        ls = ls - O[:, j]
        lf = lf - F[:, j]
    return chosen


def subtree_leaves(n_leaves: int, path: Path) -> List[Set[int]]:
    """Leaf ids below each SSA node."""
    leaves: List[Set[int]] = [{i} for i in range(n_leaves)]
    for (a, b) in path:
        leaves.append(leaves[a] | leaves[b])
    return leaves


def elimination_path(indices: Sequence[Sequence[str]], dims: Dict[str, int],
                     seed: Optional[int] = None, open_indices: Sequence[str] = (),
                     noise: float = 0.0) -> Path:
    """Contraction path from a greedy index-elimination order on the line
    graph (indices adjacent when they share a tensor): repeatedly eliminate
    the index whose neighbourhood is smallest (min-degree, log2-weighted;
    `noise` adds random tie-breaking), contracting the tensors that hold it.
    For circuit networks this follows the treewidth heuristics of variable
    elimination and gives far cheaper paths than pairwise greedy."""
    rng = np.random.default_rng(seed)
    open_set = set(open_indices)
    w = {l: math.log2(dims[l]) for ix in indices for l in ix}
    adj: Dict[str, Set[str]] = {l: set() for l in w}
    for ix in indices:
        s = set(ix)
        for a in s:
            adj[a] |= s - {a}
    alive: Dict[int, Set[str]] = {i: set(ix) for i, ix in enumerate(indices)}
    where: Dict[str, Set[int]] = {l: set() for l in w}
    for i, ix in alive.items():
        for l in ix:
            where[l].add(i)
    path: Path = []
    next_id = len(indices)

    def contract(a: int, b: int) -> int:
        nonlocal next_id
        la, lb = alive.pop(a), alive.pop(b)
        out = set()
        for l in la | lb:
            where[l].discard(a)
            where[l].discard(b)
            if l in open_set or where[l] or not (l in la and l in lb):
                out.add(l)
        for l in out:
            where[l].add(next_id)
        alive[next_id] = out
        path.append((a, b))
        next_id += 1
        return next_id - 1

    def degree(v: str) -> float:
        return sum(w[u] for u in adj[v])

    heap = [(degree(v) + (noise * rng.random() if noise else 0.0), v) for v in adj
            if v not in open_set]
    heapq.heapify(heap)
    eliminated: Set[str] = set()
    while heap:
        d, v = heapq.heappop(heap)
        if v in eliminated:
            continue
        cur = degree(v) + (noise * rng.random() if noise else 0.0)
        if cur > d + 1e-9:
            heapq.heappush(heap, (cur, v))
            continue
        eliminated.add(v)
        ts = sorted(where[v], key=lambda i: (len(alive[i]), i))
        while len(ts) > 1:
            a, b = ts[0], ts[1]
            c = contract(a, b)
            ts = sorted([c] + ts[2:], key=lambda i: (len(alive[i]), i))
        nbrs = adj.pop(v)
        for u in nbrs:
            adj[u].discard(v)
            adj[u] |= nbrs - {u}
        for u in nbrs:
            if u not in eliminated and u not in open_set:
                heapq.heappush(heap, (degree(u) + (noise * rng.random() if noise else 0.0), u))
    rest = sorted(alive, key=lambda i: (len(alive[i]), i))
    while len(rest) > 1:
        c = contract(rest[0], rest[1])
        rest = sorted([c] + rest[2:], key=lambda i: (len(alive[i]), i))
    return path


def _absorb_small(indices: Sequence[Sequence[str]], open_set: Set[str]):
    """Contract every tensor of rank <= 2 into the neighbour it shares the
    most indices with (gates on one wire, input and output states)."""
    alive: Dict[int, Set[str]] = {i: set(ix) for i, ix in enumerate(indices)}
    where: Dict[str, Set[int]] = {}
    for i, ix in alive.items():
        for l in ix:
            where.setdefault(l, set()).add(i)
    path: Path = []
    nid = len(indices)
    changed = True
    while changed:
        changed = False
        for i in sorted(alive, key=lambda i: (len(alive[i]), i)):
            if i not in alive or len(alive[i]) > 2:
                continue
            nbrs = {j for l in alive[i] for j in where[l] if j != i}
            if not nbrs:
                continue
            j = max(sorted(nbrs), key=lambda j: (len(alive[i] & alive[j]), -len(alive[j])))
            la, lb = alive.pop(i), alive.pop(j)
            out = set()
            for l in la | lb:
                where[l].discard(i)
                where[l].discard(j)
                if l in open_set or where[l] or not (l in la and l in lb):
                    out.add(l)
            for l in out:
                where[l].add(nid)
            alive[nid] = out
            path.append((i, j))
            nid += 1
            changed = True
    return path, alive, where, nid


def _minfill_order(tensors: Sequence[Sequence[str]], rng, noise: float) -> List[str]:
    """Greedy min-fill elimination order of the line graph (ties by degree,
    randomised by `noise`)."""
    adj: Dict[str, Set[str]] = {}
    for ix in tensors:
        s = set(ix)
        for a in s:
            adj.setdefault(a, set()).update(s - {a})

    def fill(v):
        nb = list(adj[v])
        f = 0
        for i in range(len(nb)):
            ai = adj[nb[i]]
            for j in range(i + 1, len(nb)):
                if nb[j] not in ai:
                    f += 1
        return f

    score = {v: fill(v) for v in adj}
    order: List[str] = []
    alive = set(adj)
    while alive:
        best = min(sorted(alive), key=lambda v: (score[v] + (noise * rng.random() if noise else 0.0),
                                                 len(adj[v])))
        nb = adj.pop(best)
        order.append(best)
        alive.discard(best)
        for u in nb:
            adj[u].discard(best)
            adj[u] |= nb - {u}
        touched = set(nb)
        for u in nb:
            touched |= adj[u]
        for u in touched:
            if u in alive:
                score[u] = fill(u)
    return order


def minfill_path(indices: Sequence[Sequence[str]], dims: Dict[str, int], seed: int = 0,
                 noise: float = 0.0, open_indices: Sequence[str] = ()) -> Path:
    """Absorb rank<=2 tensors, then contract along a min-fill elimination
    order of the remaining network's indices."""
    open_set = set(open_indices)
    path, alive, where, nid = _absorb_small(indices, open_set)
    rng = np.random.default_rng(seed)
    ids = sorted(alive)
    order = _minfill_order([sorted(alive[i]) for i in ids], rng, noise)

    def contract(a, b):
        nonlocal nid
        la, lb = alive.pop(a), alive.pop(b)
        out = set()
        for l in la | lb:
            where[l].discard(a)
            where[l].discard(b)
            if l in open_set or where[l] or not (l in la and l in lb):
                out.add(l)
        for l in out:
            where[l].add(nid)
        alive[nid] = out
        path.append((a, b))
        nid += 1
        return nid - 1

    for v in order:
        if v in open_set:
            continue
        ts = sorted(where.get(v, ()), key=lambda i: (len(alive[i]), i))
        while len(ts) > 1:
            c = contract(ts[0], ts[1])
            ts = sorted([c] + ts[2:], key=lambda i: (len(alive[i]), i))
    rest = sorted(alive, key=lambda i: (len(alive[i]), i))
    while len(rest) > 1:
        c = contract(rest[0], rest[1])
        rest = sorted([c] + rest[2:], key=lambda i: (len(alive[i]), i))
    return path


def minfill_trial(indices, dims, t: int, seed: int = 0, open_indices=()) -> Path:
    """Trial t of the min-fill search: t = 0 deterministic, then randomised
    elimination orders with noise levels cycling through 0.3 .. 1.2."""
    return minfill_path(indices, dims, seed=seed + t,
                        noise=0.0 if t == 0 else 0.3 + 0.1 * (t % 10),
                        open_indices=open_indices)


def best_minfill_path(indices, dims, trials: int = 16, seed: int = 0,
                      open_indices=()) -> Path:
    """Deterministic min-fill plus randomised trials; fewest flops wins."""
    best, best_key = None, None
    for t in range(trials):
        p = minfill_trial(indices, dims, t, seed, open_indices)
        st = path_stats(indices, dims, p, open_indices=open_indices)
        key = (st.total_flops, st.max_log2_size)
        if best_key is None or key < best_key:
            best, best_key = p, key
    return best


def choose_slices_for_flops(indices, dims, path: Path, n_slices: int,
                            max_log2_size: Optional[float] = None) -> List[str]:
    """Greedy slice selection minimising the per-slice flops, while keeping
    the largest per-slice intermediate reachable below max_log2_size."""
    st = path_stats(indices, dims, path)
    labels = sorted({l for ix in indices for l in ix})
    col = {l: j for j, l in enumerate(labels)}
    ns = len(path)
    O = np.zeros((ns, len(labels)))
    F = np.zeros((ns, len(labels)))
    for s in range(ns):
        for l in st.out[s]:
            O[s, col[l]] = math.log2(dims[l])
            F[s, col[l]] = math.log2(dims[l])
        for l in st.contracted[s]:
            F[s, col[l]] = math.log2(dims[l])
    ls, lf = st.log2_size.copy(), st.log2_flops.copy()
    chosen: List[str] = []
    avail = np.ones(len(labels), dtype=bool)
    for i in range(n_slices):
        new_max = np.max(ls[:, None] - O, axis=0)
        key = np.log2(np.sum(2.0 ** (lf[:, None] - F), axis=0))
        if max_log2_size is not None:
            key = key + 100.0 * np.maximum(0.0, new_max - (max_log2_size + (n_slices - i - 1)))
        key[~avail] = np.inf
        j = int(np.argmin(key))
        chosen.append(labels[j])
        avail[j] = False
        ls = ls - O[:, j]
        lf = lf - F[:, j]
    return chosen


class _Tree:
    """A contraction tree over label bitmasks (labels of a network whose
    every index is on at most two tensors: a contraction's result is the
    symmetric difference of its operands, except for open labels, which are
    never summed -- batch modes where both operands hold them)."""

    def __init__(self, indices, dims, path: Path, sliced=(), memory_weight: float = 0.0,
                 step_cost=None, open_labels=()):
        self.memory_weight = float(memory_weight)
        self.step_cost = step_cost
        sl = set(sliced)
        labels = sorted({l for ix in indices for l in ix if l not in sl})
        bit = {l: i for i, l in enumerate(labels)}
        self.open = 0
        for l in open_labels:
            if l in bit:
                self.open |= 1 << bit[l]
        self.log2dim = [math.log2(dims[l]) for l in labels]
        self.unit = all(dims[l] == 2 for l in labels)
        self.n = len(indices)
        self.mask: Dict[int, int] = {}
        for i, ix in enumerate(indices):
            m = 0
            for l in ix:
                if l not in sl:
                    m |= 1 << bit[l]
            self.mask[i] = m
        self.kids: Dict[int, Tuple[int, int]] = {}
        self.next_id = self.n
        for (a, b) in path:
            z = self.next_id
            self.next_id += 1
            self.kids[z] = (a, b)
            self.mask[z] = self.result(self.mask[a], self.mask[b])
        self.root = self.next_id - 1

    def result(self, a_mask: int, b_mask: int) -> int:
        return (a_mask ^ b_mask) | (a_mask & b_mask & self.open)

    def log2size(self, m: int) -> float:
        if self.unit:
            return float(m.bit_count())
        return sum(self.log2dim[i] for i in range(m.bit_length()) if m >> i & 1)

    def cost(self, a_mask: int, b_mask: int) -> float:
        """PathInfo flops of one contraction, |out| * 2 * prod(contracted),
        plus memory_weight flops per element read or written (|a| + |b| +
        |out|): the exchange rate between the two on the target device."""
        if self.step_cost is not None:
            lm = self.log2size(a_mask & ~b_mask)
            ln = self.log2size(b_mask & ~a_mask)
            if not self.open:
                return self.step_cost(lm, ln, self.log2size(a_mask & b_mask))
            batch = a_mask & b_mask & self.open
            return self.step_cost(lm, ln, self.log2size(a_mask & b_mask & ~self.open),
                                  self.log2size(batch))
        c = 2.0 * 2.0 ** self.log2size(a_mask | b_mask)
        if self.memory_weight:
            c += self.memory_weight * (2.0 ** self.log2size(a_mask) + 2.0 ** self.log2size(b_mask)
                                       + 2.0 ** self.log2size(self.result(a_mask, b_mask)))
        return c

    def node_cost(self, z: int) -> float:
        a, b = self.kids[z]
        return self.cost(self.mask[a], self.mask[b])

    def total(self) -> float:
        return sum(self.node_cost(z) for z in self.kids)

    def max_log2size(self) -> float:
        return max(self.log2size(self.mask[z]) for z in self.kids)

    def to_path(self) -> Path:
        """SSA path by post-order traversal (larger-cost child first)."""
        order: List[int] = []
        stack = [(self.root, False)]
        while stack:
            z, done = stack.pop()
            if z < self.n:
                continue
            if done:
                order.append(z)
                continue
            stack.append((z, True))
            a, b = self.kids[z]
            ca = self.subtree_cost(a)
            cb = self.subtree_cost(b)
            first, second = (a, b) if ca >= cb else (b, a)
            stack.append((second, False))
            stack.append((first, False))
        ssa = {i: i for i in range(self.n)}
        path: Path = []
        for z in order:
            a, b = self.kids[z]
            path.append((ssa[a], ssa[b]))
            ssa[z] = self.n + len(path) - 1
        return path

    def subtree_cost(self, z: int) -> float:
        if z < self.n:
            return 0.0
        total, stack = 0.0, [z]
        while stack:
            x = stack.pop()
            if x < self.n:
                continue
            total += self.node_cost(x)
            stack.extend(self.kids[x])
        return total


def _optimal_subtree(masks: List[int], tree: _Tree, max_log2: float):
    """Cheapest contraction tree over `masks` (dynamic programming over
    subsets); intermediates above max_log2 are not allowed.  Returns
    (cost, split) where split[S] = the first part of subset S's best split."""
    k = len(masks)
    full = (1 << k) - 1
    res = [0] * (1 << k)
    union = [0] * (1 << k)
    for S in range(1, 1 << k):
        low = S & -S
        m = masks[low.bit_length() - 1]
        union[S] = union[S ^ low] | m
        res[S] = res[S ^ low] ^ m
    if tree.open:
        for S in range(1, 1 << k):
            res[S] |= union[S] & tree.open
    inf = float("inf")
    best = [inf] * (1 << k)
    split = [0] * (1 << k)
    for i in range(k):
        best[1 << i] = 0.0
    for S in range(1, 1 << k):
        if S & (S - 1) == 0:
            continue
        if S != full and tree.log2size(res[S]) > max_log2:
            continue
        low = S & -S
        rest = S ^ low
        # A holds S's lowest element: each unordered split once
        sub = rest
        b_best, b_split = inf, 0
        while True:
            A = sub | low
            B = S ^ A
            if B and best[A] < inf and best[B] < inf:
                c = best[A] + best[B] + tree.cost(res[A], res[B])
                if c < b_best:
                    b_best, b_split = c, A
            if sub == 0:
                break
            sub = (sub - 1) & rest
        best[S] = b_best
        split[S] = b_split
    return best[full], split


def reconfigure_path(indices, dims, path: Path, sliced: Sequence[str] = (),
                     max_log2_size: float = float("inf"), subtree_size: int = 8,
                     passes: int = 4, max_nodes: Optional[int] = None,
                     memory_weight: float = 0.0, step_cost=None, open_labels=()) -> Path:
    """Subtree reconfiguration of a contraction path, for the network with
    `sliced` indices removed: for every internal node, the subtree below it
    cut at `subtree_size` frontier tensors (expanding the costliest
    contractions first) is replaced by the cheapest contraction order over
    that frontier when that is cheaper (exact dynamic programming over
    subsets), keeping every intermediate within max_log2_size.  Repeated for
    `passes` passes or until nothing improves; with max_nodes, each pass
    visits only that many of the costliest nodes.  The cost of a contraction
    is its PathInfo flops plus memory_weight per element of its operands and
    result (0: flops only; on a GPU whose memory-bound steps run at B bytes/s
    and MFMA steps at F flop/s, about 8 * F / B / 4 for complex64).  A
    step_cost(log2 M, log2 N, log2 K) callable replaces that cost (e.g. a
    time model of the kernels that will run each step).  open_labels are
    never summed (batch modes where both operands hold them), e.g. the
    sliced indices that a batched execution keeps; with them step_cost takes
    a fourth argument, log2 of the step's batch extent.  Needs every other
    index on at most two tensors (circuit networks)."""
    counts: Dict[str, int] = {}
    for ix in indices:
        for l in ix:
            counts[l] = counts.get(l, 0) + 1
    if any(c > 2 and l not in open_labels for l, c in counts.items()):
        raise ValueError("reconfigure_path needs every index that is not open on at most two "
                         "tensors.")
    tree = _Tree(indices, dims, path, sliced, memory_weight, step_cost, open_labels)
    for _ in range(passes):
        improved = False
        nodes = sorted(tree.kids, key=tree.node_cost, reverse=True)
        if max_nodes is not None:
            nodes = nodes[:max_nodes]
        for v in nodes:
            if v not in tree.kids:
                continue
            # frontier: expand the costliest internal node until subtree_size
            frontier = list(tree.kids[v])
            inner = [v]
            while len(frontier) < subtree_size:
                cand = [x for x in frontier if x >= tree.n]
                if not cand:
                    break
                x = max(cand, key=tree.node_cost)
                frontier.remove(x)
                frontier.extend(tree.kids[x])
                inner.append(x)
            if len(inner) < 2:
                continue
            old = sum(tree.node_cost(z) for z in inner)
            masks = [tree.mask[x] for x in frontier]
            new, split = _optimal_subtree(masks, tree, max_log2_size)
            if not new < old * (1 - 1e-12):
                continue
            # rebuild: v keeps its id (same result), fresh ids for the rest
            for z in inner[1:]:
                del tree.kids[z]
                del tree.mask[z]

            def build(S, node_id=None):
                if S & (S - 1) == 0:
                    return frontier[S.bit_length() - 1]
                A = split[S]
                a = build(A)
                b = build(S ^ A)
                z = node_id if node_id is not None else tree.next_id
                if node_id is None:
                    tree.next_id += 1
                tree.kids[z] = (a, b)
                tree.mask[z] = tree.result(tree.mask[a], tree.mask[b])
                return z

            build((1 << len(frontier)) - 1, v)
            improved = True
        if not improved:
            break
    return tree.to_path()


def mi355x_step_seconds(log2_m: float, log2_n: float, log2_k: float,
                        operand_bytes: float = 40.0, output_bytes: float = 16.0,
                        launch_s: float = 8e-6, batch_bits: float = 6.0,
                        log2_batch: float = 0.0) -> float:
    """Roofline time model of one complex64 pairwise contraction (one slice)
    on this library's MI355X kernels, for path refinement:

    * M, N, K >= 128: the split-FP16 4M GEMM at ~470 TF/s algorithmic, plus
      operand preparation (permutation, row maximum, split: about
      `operand_bytes` per operand element) and the output write (and its
      permutation, `output_bytes` per element) at ~3 TB/s;
    * other MFMA shapes (M >= 32, N, K >= 16): the FP32 MFMA kernel with
      128x128 tiles (tile waste counted) at ~110 TF/s, or its operand and
      output traffic at ~3 TB/s (x1.5 for permutations), whichever is larger;
    * the rest: VALU kernels at ~20 TF/s or 2.5 TB/s;
    * times 2**log2_batch for a batched step (batch modes of the GEMM);
    * plus `launch_s` per step, shared by the 2**batch_bits slices of one
      batched execution."""
    lm, ln, lk = log2_m, log2_n, log2_k
    fl = 8.0 * 2.0 ** (lm + ln + lk)
    a, b, c = 2.0 ** (lm + lk), 2.0 ** (ln + lk), 2.0 ** (lm + ln)
    if lm >= 7 and ln >= 7 and lk >= 7:
        t = fl / 470e12 + (operand_bytes * (a + b) + output_bytes * c) / 3e12
    elif lm >= 5 and ln >= 4 and lk >= 4:
        padded = 8.0 * 2.0 ** (max(lm, 7) + max(ln, 7) + lk)
        t = max(padded / 110e12, 1.5 * 8.0 * (a + b + c) / 3e12)
    else:
        t = max(fl / 20e12, 1.5 * 8.0 * (a + b + c) / 2.5e12)
    return t * 2.0 ** log2_batch + launch_s / 2.0 ** batch_bits


# Refinement stages, as (operand_bytes, output_bytes) of mi355x_step_seconds:
# an output-light model first, then the full operand preparation cost (the
# second stage alone stops in a worse local optimum).
REFINE_STAGES = ((12.0, 0.0), (40.0, 16.0))


def batched_tree(indices, dims, path: Path, sliced: Sequence[str], batch_log2: int,
                 operand_bytes: float = 40.0, output_bytes: float = 16.0) -> "_Tree":
    """The contraction tree of one batched execution: the last `batch_log2`
    sliced indices stay open (batch modes), the others are fixed; each
    step's cost is mi355x_step_seconds of its batched GEMM (one launch per
    execution), so total() models the seconds of one execution."""
    nb = int(batch_log2)
    loop, batch = list(sliced[:len(sliced) - nb]), list(sliced[len(sliced) - nb:])

    def cost(lm, ln, lk, lb):
        return mi355x_step_seconds(lm, ln, lk, operand_bytes, output_bytes, batch_bits=0.0,
                                   log2_batch=lb)
    return _Tree(indices, dims, path, loop, step_cost=cost, open_labels=batch)


def refine_batched_path(indices, dims, path: Path, sliced: Sequence[str], batch_log2: int,
                        extra_log2: float = 1.0, subtree_size: int = 10, rounds: int = 2,
                        stages=REFINE_STAGES) -> Path:
    """Subtree reconfiguration of a sliced path for batched execution: the
    last `batch_log2` sliced indices are batch modes of every step that
    holds them (open labels), the others are sliced away, and each step
    costs the modelled seconds of its batched launch (batched_tree).  The
    largest batched intermediate may grow by `extra_log2` over the input
    path's.  Each round runs the stages in order; rounds stop early when one
    does not lower the modelled execution time."""
    nb = int(batch_log2)
    loop, batch = list(sliced[:len(sliced) - nb]), list(sliced[len(sliced) - nb:])
    limit = batched_tree(indices, dims, path, sliced, nb).max_log2size() + extra_log2
    best = batched_tree(indices, dims, path, sliced, nb).total()
    for _ in range(rounds):
        cand = path
        for ob, cb in stages:
            def cost(lm, ln, lk, lb, ob=ob, cb=cb):
                return mi355x_step_seconds(lm, ln, lk, ob, cb, batch_bits=0.0, log2_batch=lb)
            cand = reconfigure_path(indices, dims, cand, loop, max_log2_size=limit,
                                    subtree_size=subtree_size, passes=20, step_cost=cost,
                                    open_labels=batch)
        t = batched_tree(indices, dims, cand, sliced, nb).total()
        if t >= best:
            break
        path, best = cand, t
    return path


def amplitude_step_seconds(log2_m: float, log2_n: float, log2_k: float) -> float:
    """mi355x_step_seconds of one unbatched contraction (its launch cost
    not shared with other slices)."""
    return mi355x_step_seconds(log2_m, log2_n, log2_k, batch_bits=0.0)


def reconfigured_minfill_path(indices, dims, trials: int = 8, seed: int = 0,
                              subtree_sizes: Sequence[int] = (8, 10), passes: int = 6,
                              step_cost=amplitude_step_seconds) -> Path:
    """best_minfill_path, then subtree reconfiguration (reconfigure_path) at
    each subtree size in turn, first against PathInfo flops, then against
    step_cost (the time model of these kernels; None skips that stage), the
    largest intermediate never above the min-fill path's.  The flops stage
    alone removes most of the min-fill path's flops on grid-circuit
    amplitudes (6x6 depth 12: 4.1e12 to 3.1e11) but leaves memory-bound
    steps on 2^30-element intermediates; the time stage trades some of
    those flops back for smaller, MFMA-friendly steps (modelled 69 -> 44 ->
    19 ms, largest intermediate 2^28)."""
    path = best_minfill_path(indices, dims, trials=trials, seed=seed)
    limit = path_stats(indices, dims, path).max_log2_size
    for cost in (None, step_cost) if step_cost is not None else (None,):
        for size in subtree_sizes:
            path = reconfigure_path(indices, dims, path, subtree_size=size, passes=passes,
                                    max_log2_size=limit, step_cost=cost)
    return path


def modelled_amplitude_seconds(indices, dims, path: Path) -> float:
    """Modelled seconds of one amplitude contraction along `path`
    (amplitude_step_seconds summed over its steps)."""
    return _Tree(indices, dims, path, [], 0.0, amplitude_step_seconds).total()


AMPLITUDE_PATHS = "amplitude_paths.json"


def _amplitude_key(rows, cols, depth, seed, random_angles):
    return f"{rows}x{cols}x{depth}_seed{seed}" + ("_angles" if random_angles else "")


def amplitude_path_entry(rows, cols, depth, seed, random_angles=False, trials=8):
    """{"path", "pathinfo_flops", "max_log2_size", "modelled_s" and the same
    of the min-fill path it starts from} of one amplitude of a seeded grid
    circuit (jet_amd.circuits) along reconfigured_minfill_path."""
    from . import circuits
    circ = circuits.random_grid_circuit(rows, cols, depth, seed=seed,
                                        random_angles=random_angles)
    net = circuits.amplitude_network(circ)
    dims = {l: 2 for ix in net.indices for l in ix}
    base = best_minfill_path(net.indices, dims, trials=trials, seed=1)
    path = reconfigured_minfill_path(net.indices, dims, trials=trials, seed=1)
    st, sb = path_stats(net.indices, dims, path), path_stats(net.indices, dims, base)
    return {"path": [list(p) for p in path], "pathinfo_flops": st.total_flops,
            "max_log2_size": st.max_log2_size,
            "modelled_s": modelled_amplitude_seconds(net.indices, dims, path),
            "minfill_pathinfo_flops": sb.total_flops, "minfill_max_log2_size": sb.max_log2_size,
            "minfill_modelled_s": modelled_amplitude_seconds(net.indices, dims, base)}


def load_or_make_amplitude_path(rows, cols, depth, seed, random_angles=False,
                                directory: Optional[str] = None) -> Path:
    """The path of one amplitude of a seeded grid circuit from
    jet_amd/data/amplitude_paths.json (written by `python -m jet_amd.paths
    --amplitude-paths`), else computed here (amplitude_path_entry)."""
    import json
    import os
    directory = directory or os.path.join(os.path.dirname(os.path.abspath(__file__)), "data")
    name = os.path.join(directory, AMPLITUDE_PATHS)
    key = _amplitude_key(rows, cols, depth, seed, random_angles)
    if os.path.exists(name):
        with open(name) as f:
            table = json.load(f)
        if key in table:
            return [tuple(p) for p in table[key]["path"]]
    return [tuple(p) for p in
            amplitude_path_entry(rows, cols, depth, seed, random_angles)["path"]]


def _sliced_trial(args):
    indices, dims, t, seed, n_slices, max_log2_size = args
    p = minfill_trial(indices, dims, t, seed)
    sl = choose_slices_for_flops(indices, dims, p, n_slices, max_log2_size)
    per = path_stats(indices, dims, p, sliced=sl)
    return t, per.total_flops, per.max_log2_size, p, sl


def best_sliced_minfill_plan(indices, dims, n_slices: int, max_log2_size: float,
                             trials: int = 64, seed: int = 0, workers: int = 1):
    """Min-fill trials judged by their sliced cost: each trial's path is
    sliced on `n_slices` indices (choose_slices_for_flops) and the trial
    with the fewest per-slice flops whose largest per-slice intermediate
    fits max_log2_size wins (the unsliced flops of a path predict its sliced
    cost poorly).  Returns (path, sliced indices)."""
    jobs = [(indices, dims, t, seed, n_slices, max_log2_size) for t in range(trials)]
    if workers > 1:
        from multiprocessing import Pool
        with Pool(workers) as pool:
            results = pool.map(_sliced_trial, jobs)
    else:
        results = [_sliced_trial(j) for j in jobs]
    ok = [r for r in results if r[2] <= max_log2_size] or results
    best = min(ok, key=lambda r: (r[1], r[0]))
    return best[3], best[4]


def grid_circuit_plan(rows: int, cols: int, depth: int, seed: int, n_slices: int,
                      trials: int = 16, max_log2_size: float = 29.0, workers: int = 1) -> dict:
    """Contraction path and sliced indices for one amplitude of a seeded grid
    random circuit (jet_amd.circuits): of `trials` min-fill paths, the one
    with the fewest per-slice flops after choosing `n_slices` indices."""
    from . import circuits
    circ = circuits.random_grid_circuit(rows, cols, depth, seed=seed)
    net = circuits.amplitude_network(circ)
    dims = {l: 2 for ix in net.indices for l in ix}
    path, sliced = best_sliced_minfill_plan(net.indices, dims, n_slices, max_log2_size,
                                            trials, seed, workers)
    full = path_stats(net.indices, dims, path)
    per = path_stats(net.indices, dims, path, sliced=sliced)
    return {"rows": rows, "cols": cols, "depth": depth, "seed": seed, "trials": trials,
            "n_tensors": len(net), "path": [list(p) for p in path], "sliced": sliced,
            "pathinfo_flops": full.total_flops, "max_log2_size": full.max_log2_size,
            "per_slice_pathinfo_flops": per.total_flops,
            "per_slice_max_log2_size": per.max_log2_size}


def load_or_make_grid_plan(rows, cols, depth, seed, n_slices, trials=16, max_log2_size=29.0,
                           directory: Optional[str] = None) -> dict:
    """Plan from jet_amd/data (written by `python -m jet_amd.paths`) when one
    with the same parameters exists, else computed here."""
    import json
    import os
    directory = directory or os.path.join(os.path.dirname(os.path.abspath(__file__)), "data")
    name = os.path.join(directory, f"grid_{rows}x{cols}x{depth}_seed{seed}_s{n_slices}.json")
    if os.path.exists(name):
        with open(name) as f:
            plan = json.load(f)
        if (plan["rows"], plan["cols"], plan["depth"], plan["seed"], len(plan["sliced"])) == \
                (rows, cols, depth, seed, n_slices):
            plan["path"] = [tuple(p) for p in plan["path"]]
            return plan
    plan = grid_circuit_plan(rows, cols, depth, seed, n_slices, trials, max_log2_size)
    plan["path"] = [tuple(p) for p in plan["path"]]
    return plan


if __name__ == "__main__":
    import argparse
    import json
    import os
    ap = argparse.ArgumentParser(description="write a grid-circuit contraction plan to jet_amd/data")
    ap.add_argument("--rows", type=int, default=7)
    ap.add_argument("--cols", type=int, default=7)
    ap.add_argument("--depth", type=int, default=14)
    ap.add_argument("--seed", type=int, default=2026)
    ap.add_argument("--slices", type=int, default=12)
    ap.add_argument("--trials", type=int, default=64)
    ap.add_argument("--max-log2-size", type=float, default=29.0)
    ap.add_argument("--workers", type=int, default=1)
    ap.add_argument("--refine", metavar="PLAN_JSON",
                    help="refine an existing plan's path for batched execution "
                         "(refine_batched_path) and write it to --out")
    ap.add_argument("--batch-log2", type=int, default=7,
                    help="--refine: sliced indices kept as batch modes (the last ones)")
    ap.add_argument("--extra-log2", type=float, default=1.0,
                    help="--refine: growth allowed for the largest batched intermediate")
    ap.add_argument("--rounds", type=int, default=2)
    ap.add_argument("--out", help="where --refine writes (default: over its input)")
    ap.add_argument("--amplitude-paths", action="store_true",
                    help="write jet_amd/data/amplitude_paths.json: reconfigured min-fill "
                         "paths of the uneven-batch circuits (5x5 depth 14 seed 51, 6x6 "
                         "depth 12 seed 52, 6x7 depth 10 seed 53) and of the gradient "
                         "circuit (6x6 depth 10 seed 2026, random angles)")
    a = ap.parse_args()
    if a.amplitude_paths:
        d = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data")
        table = {}
        for (r, c, dp, sd, ra) in [(5, 5, 14, 51, False), (6, 6, 12, 52, False),
                                    (6, 7, 10, 53, False), (6, 6, 10, 2026, True)]:
            e = amplitude_path_entry(r, c, dp, sd, ra)
            table[_amplitude_key(r, c, dp, sd, ra)] = e
            print(_amplitude_key(r, c, dp, sd, ra),
                  "pathinfo flops %.3e (min-fill %.3e) max %d, modelled %.2f ms (min-fill %.2f)" % (
                      e["pathinfo_flops"], e["minfill_pathinfo_flops"], e["max_log2_size"],
                      e["modelled_s"] * 1e3, e["minfill_modelled_s"] * 1e3), flush=True)
        with open(os.path.join(d, AMPLITUDE_PATHS), "w") as f:
            json.dump(table, f)
        raise SystemExit(0)
    if a.refine:
        from . import circuits
        with open(a.refine) as f:
            plan = json.load(f)
        circ = circuits.random_grid_circuit(plan["rows"], plan["cols"], plan["depth"],
                                            seed=plan["seed"])
        net = circuits.amplitude_network(circ)
        dims = {l: 2 for ix in net.indices for l in ix}
        path0 = [tuple(p) for p in plan["path"]]
        sliced = plan["sliced"]
        path = refine_batched_path(net.indices, dims, path0, sliced, a.batch_log2, a.extra_log2,
                                   rounds=a.rounds)
        t0 = batched_tree(net.indices, dims, path0, sliced, a.batch_log2).total()
        t1 = batched_tree(net.indices, dims, path, sliced, a.batch_log2).total()
        full = path_stats(net.indices, dims, path)
        per = path_stats(net.indices, dims, path, sliced=sliced)
        plan.update({"path": [list(p) for p in path],
                     "refined": {"method": "refine_batched_path (reconfigure_path with the "
                                           "batched indices open, against mi355x_step_seconds)",
                                 "stages": [list(x) for x in REFINE_STAGES], "subtree_size": 10,
                                 "rounds": a.rounds, "batch_log2": a.batch_log2,
                                 "extra_log2": a.extra_log2,
                                 "modelled_execution_s": [t0, t1],
                                 "unrefined_per_slice_pathinfo_flops":
                                     plan["per_slice_pathinfo_flops"]},
                     "pathinfo_flops": full.total_flops, "max_log2_size": full.max_log2_size,
                     "per_slice_pathinfo_flops": per.total_flops,
                     "per_slice_max_log2_size": per.max_log2_size})
        out = a.out or a.refine
        with open(out, "w") as f:
            json.dump(plan, f)
        print(out, "per slice %.3e max %d; modelled execution %.1f -> %.1f ms" % (
            per.total_flops, per.max_log2_size, t0 * 1e3, t1 * 1e3))
        raise SystemExit(0)
    plan = grid_circuit_plan(a.rows, a.cols, a.depth, a.seed, a.slices, a.trials, a.max_log2_size,
                             a.workers)
    d = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data")
    os.makedirs(d, exist_ok=True)
    out = os.path.join(d, f"grid_{a.rows}x{a.cols}x{a.depth}_seed{a.seed}_s{a.slices}.json")
    with open(out, "w") as f:
        json.dump(plan, f)
    print(out, "pathinfo flops %.3e max %d, per slice %.3e max %d" % (
        plan["pathinfo_flops"], plan["max_log2_size"], plan["per_slice_pathinfo_flops"],
        plan["per_slice_max_log2_size"]))

"""Oracle: CPU restatement of Jet's pairwise-contraction path.

TEST INFRASTRUCTURE ONLY.  This package is the checker the HIP product is
compared against.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import it; the product package
``jet_amd`` never does.

Layers:
  * ``oracle/_build/libjet_oracle.so`` -- own C restatement (jet_oracle.c):
    Default/QFlex permutations, FP64-accumulating GEMM, ContractTensors.
  * ``oracle/_ref/libjet_ref.so`` -- the reference's own permuter headers
    compiled where they lie (ref_permuters.cpp); present only where the
    reference tree was available at build time.
  * pure-Python/numpy restatements of the host-side algebra used by the
    checks (index algebra, slicing, network contraction along a path).

Each function cites the reference lines it restates; ``tests/test_oracle.py``
pins each one against the reference's fixtures and literal values, the
reference's own permuter build, or closed forms (numpy transpose and einsum,
exact rational sums, statevector simulation in tests/test_circuits_paths.py).
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "libjet_oracle.so")
_REF_PATH = os.path.join(_HERE, "_ref", "libjet_ref.so")

_ERRORS = {
    1: "Fast transpose expects power-of-2 data.",
    2: "New ordering is too small to be used at this point. "
       "Consider increasing the BLOCKSIZE.",
    3: "New ordering is too small to be used at this point.",
    4: "Number of minimum dimensions must be greater than one.",
    5: "Invalid argument.",
    6: "Allocation failed.",
}

_lib = None
_ref = None


class OracleError(RuntimeError):
    pass


def build(ref: bool = True) -> None:
    """Compile the oracle (and, when the reference tree exists, oracle/_ref)."""
    subprocess.run(["make", "-s", "-C", _HERE, "all"], check=True)
    if ref and os.path.isdir("/root/reference/include/jet/permute"):
        subprocess.run(["make", "-s", "-C", _HERE, "ref"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build(ref=False)
        L = ctypes.CDLL(_LIB_PATH)
        i64p = ctypes.POINTER(ctypes.c_int64)
        i32p = ctypes.POINTER(ctypes.c_int32)
        L.jo_permute_default.argtypes = [ctypes.c_size_t, ctypes.c_int, i64p, i32p,
                                         ctypes.c_void_p, ctypes.c_void_p]
        L.jo_permute_qflex.argtypes = [ctypes.c_size_t, ctypes.c_int, i64p, i32p,
                                       ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t,
                                       ctypes.c_size_t, ctypes.c_int]
        L.jo_qflex_plan_kinds.argtypes = [ctypes.c_int, i64p, i32p, ctypes.c_size_t,
                                          ctypes.c_size_t, ctypes.c_char_p,
                                          ctypes.POINTER(ctypes.c_int)]
        L.jo_gemm.argtypes = [ctypes.c_size_t, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64,
                              ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int]
        L.jo_gemm.restype = None
        L.jo_gemm_entries.argtypes = [ctypes.c_size_t, ctypes.c_int64, ctypes.c_void_p,
                                      ctypes.c_int64, ctypes.c_void_p, ctypes.c_int64,
                                      ctypes.c_int64, i64p, i64p, ctypes.c_void_p, ctypes.c_int]
        L.jo_gemm_entries.restype = None
        L.jo_contract.argtypes = [ctypes.c_size_t, ctypes.c_int, i32p, i64p, ctypes.c_void_p,
                                  ctypes.c_int, i32p, i64p, ctypes.c_void_p,
                                  ctypes.POINTER(ctypes.c_int), i32p, i64p, ctypes.c_void_p,
                                  ctypes.c_int]
        L.jo_sum_compensated.argtypes = [ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p]
        L.jo_sum_compensated.restype = None
        L.jb_load_blas.argtypes = [ctypes.c_char_p, ctypes.c_char_p]
        L.jb_set_blas_threads.argtypes = [ctypes.c_int]
        L.jb_set_blas_threads.restype = None
        L.jb_contract.argtypes = [ctypes.c_size_t, ctypes.c_int, i32p, i64p, ctypes.c_void_p,
                                  ctypes.c_int, i32p, i64p, ctypes.c_void_p, ctypes.c_void_p,
                                  ctypes.c_int]
        _lib = L
    return _lib


def ref_lib():
    """The reference's own permuters, or None where oracle/_ref was not built."""
    global _ref
    if _ref is None and os.path.exists(_REF_PATH):
        R = ctypes.CDLL(_REF_PATH)
        R.jr_permute.argtypes = [ctypes.c_int, ctypes.c_size_t, ctypes.c_size_t, ctypes.c_int,
                                 ctypes.POINTER(ctypes.c_int64),
                                 ctypes.POINTER(ctypes.c_char_p),
                                 ctypes.POINTER(ctypes.c_char_p), ctypes.c_void_p,
                                 ctypes.c_size_t, ctypes.c_void_p, ctypes.c_char_p,
                                 ctypes.c_size_t]
        _ref = R
    return _ref


def _i64(a) -> ctypes.Array:
    return (ctypes.c_int64 * max(1, len(a)))(*[int(x) for x in a])


def _i32(a) -> ctypes.Array:
    return (ctypes.c_int32 * max(1, len(a)))(*[int(x) for x in a])


def _nthreads(n: Optional[int]) -> int:
    return int(n if n else (os.cpu_count() or 1))


# ---------------------------------------------------------------------------
# Permutations (Default.hpp:20-133, QFlex.hpp:20-539)
# ---------------------------------------------------------------------------

def permute_default(data: np.ndarray, shape: Sequence[int], new_order: Sequence[int]) -> np.ndarray:
    """Row-major permutation: output axis i is input axis new_order[i]."""
    data = np.ascontiguousarray(data)
    out = np.empty_like(data)
    rc = lib().jo_permute_default(data.itemsize, len(shape), _i64(shape), _i32(new_order),
                                  data.ctypes.data, out.ctypes.data)
    if rc:
        raise OracleError(_ERRORS.get(rc, str(rc)))
    return out


def permute_qflex(data: np.ndarray, shape: Sequence[int], new_order: Sequence[int],
                  blocksize: int = 1024, min_dims: int = 32,
                  nthreads: Optional[int] = 1) -> np.ndarray:
    data = np.ascontiguousarray(data)
    out = np.empty_like(data)
    rc = lib().jo_permute_qflex(data.itemsize, len(shape), _i64(shape), _i32(new_order),
                                data.ctypes.data, out.ctypes.data, blocksize, min_dims,
                                _nthreads(nthreads))
    if rc:
        raise OracleError(_ERRORS.get(rc, str(rc)))
    return out


def qflex_plan_kinds(shape: Sequence[int], new_order: Sequence[int], blocksize: int = 1024,
                     min_dims: int = 32) -> str:
    buf = ctypes.create_string_buffer(4)
    n = ctypes.c_int(0)
    rc = lib().jo_qflex_plan_kinds(len(shape), _i64(shape), _i32(new_order), blocksize,
                                   min_dims, buf, ctypes.byref(n))
    if rc:
        raise OracleError(_ERRORS.get(rc, str(rc)))
    return buf.raw[: n.value].decode()


def ref_permute(data: np.ndarray, shape: Sequence[int], old_labels: Sequence[str],
                new_labels: Sequence[str], backend: str = "qflex", blocksize: int = 1024,
                min_dims: int = 32) -> np.ndarray:
    """Run the reference's own Permuter<Backend>::Transpose (Permuter.hpp:92-119).

    Raises OracleError with the reference's exception text on failure."""
    R = ref_lib()
    if R is None:
        raise FileNotFoundError("oracle/_ref/libjet_ref.so was not built")
    data = np.ascontiguousarray(data, dtype=np.complex64)
    out = np.zeros_like(data)
    olds = (ctypes.c_char_p * max(1, len(old_labels)))(*[s.encode() for s in old_labels])
    news = (ctypes.c_char_p * (len(new_labels) + 1))(*([s.encode() for s in new_labels] + [None]))
    err = ctypes.create_string_buffer(1024)
    rc = R.jr_permute(0 if backend == "default" else 1, blocksize, min_dims, len(shape),
                      _i64(shape), olds, news, data.ctypes.data, data.size, out.ctypes.data,
                      err, 1024)
    if rc == 2:
        raise ValueError("template configuration not instantiated in oracle/_ref")
    if rc:
        raise OracleError(err.value.decode())
    return out


# ---------------------------------------------------------------------------
# Index algebra (Utilities.hpp:316-384) and ContractTensors (Tensor.hpp:708-752)
# ---------------------------------------------------------------------------

def vector_intersection(a: Sequence, b: Sequence) -> list:
    return [x for x in a if x in b]


def vector_subtraction(a: Sequence, b: Sequence) -> list:
    return [x for x in a if x not in b]


def vector_union(a: Sequence, b: Sequence) -> list:
    return list(a) + [x for x in b if x not in a]


def vector_disjunctive_union(a: Sequence, b: Sequence) -> list:
    return vector_subtraction(vector_union(a, b), vector_intersection(a, b))


def gemm(A: np.ndarray, B: np.ndarray, nthreads: Optional[int] = None) -> np.ndarray:
    """C = A @ B for row-major complex matrices with FP64 accumulation."""
    A = np.ascontiguousarray(A)
    B = np.ascontiguousarray(B)
    m, k = A.shape
    k2, n = B.shape
    assert k == k2 and A.dtype == B.dtype
    C = np.empty((m, n), dtype=np.complex128)
    lib().jo_gemm(A.itemsize, m, n, k, A.ctypes.data, B.ctypes.data, C.ctypes.data,
                  _nthreads(nthreads))
    return C


def gemm_entries(A: np.ndarray, B: np.ndarray, rows, cols, nthreads: Optional[int] = None):
    """Selected entries of A @ B, each an FP64 dot product."""
    A = np.ascontiguousarray(A)
    B = np.ascontiguousarray(B)
    rows = np.ascontiguousarray(rows, dtype=np.int64)
    cols = np.ascontiguousarray(cols, dtype=np.int64)
    out = np.empty(rows.size, dtype=np.complex128)
    k = A.shape[1]
    lib().jo_gemm_entries(A.itemsize, k, A.ctypes.data, A.shape[1], B.ctypes.data, B.shape[1],
                          rows.size, rows.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                          cols.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), out.ctypes.data,
                          _nthreads(nthreads))
    return out


class LabelTable:
    """Maps string labels to the int32 ids the C oracle works on."""

    def __init__(self):
        self.ids: Dict[str, int] = {}
        self.names: List[str] = []

    def id(self, s: str) -> int:
        if s not in self.ids:
            self.ids[s] = len(self.names)
            self.names.append(s)
        return self.ids[s]


def contract(a_labels: Sequence[str], a_shape: Sequence[int], a: np.ndarray,
             b_labels: Sequence[str], b_shape: Sequence[int], b: np.ndarray,
             nthreads: Optional[int] = None) -> Tuple[List[str], List[int], np.ndarray]:
    """Tensor::ContractTensors in FP64: returns (labels, shape, complex128 data).

    Output labels are left ++ right (Tensor.hpp:714-745)."""
    assert a.dtype == b.dtype and a.dtype in (np.complex64, np.complex128)
    tab = LabelTable()
    la = [tab.id(s) for s in a_labels]
    lb = [tab.id(s) for s in b_labels]
    ra, rb = len(la), len(lb)
    lc = (ctypes.c_int32 * max(1, ra + rb))()
    sc = (ctypes.c_int64 * max(1, ra + rb))()
    rc = ctypes.c_int(0)
    left = vector_subtraction(a_labels, b_labels)
    right = vector_subtraction(b_labels, a_labels)
    size = 1
    for s in left:
        size *= a_shape[list(a_labels).index(s)]
    for s in right:
        size *= b_shape[list(b_labels).index(s)]
    C = np.empty(size, dtype=np.complex128)
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b)
    err = lib().jo_contract(a.itemsize, ra, _i32(la), _i64(a_shape), a.ctypes.data, rb,
                            _i32(lb), _i64(b_shape), b.ctypes.data, ctypes.byref(rc), lc, sc,
                            C.ctypes.data, _nthreads(nthreads))
    if err:
        raise OracleError(_ERRORS.get(err, str(err)))
    labels = [tab.names[lc[i]] for i in range(rc.value)]
    shape = [int(sc[i]) for i in range(rc.value)]
    return labels, shape, C


def sum_compensated(x: np.ndarray) -> complex:
    x = np.ascontiguousarray(x, dtype=np.complex128)
    out = np.zeros(2)
    lib().jo_sum_compensated(x.size, x.ctypes.data, out.ctypes.data)
    return complex(out[0], out[1])


# ---------------------------------------------------------------------------
# Slicing (Tensor.hpp:493-526, TensorNetwork.hpp:210-284) and sequential
# network contraction (TensorNetwork.hpp:301-328, 394-421).
# ---------------------------------------------------------------------------

def unravel_index(index: int, shape: Sequence[int]) -> List[int]:
    """Utilities.hpp:463-475: row-major, first entry most significant."""
    size = 1
    for s in shape:
        size *= s
    if index >= size:
        raise OracleError("Linear index does not fit in the shape.")
    out = [0] * len(shape)
    for i in range(len(shape) - 1, -1, -1):
        out[i] = index % shape[i]
        index //= shape[i]
    return out


def slice_index(labels: Sequence[str], shape: Sequence[int], data: np.ndarray, label: str,
                value: int) -> Tuple[List[str], List[int], np.ndarray]:
    """Tensor::SliceIndex: move `label` to the front, take block `value`."""
    pos = list(labels).index(label)
    arr = np.asarray(data).reshape(shape)
    out = np.take(arr, value, axis=pos)
    new_labels = [l for i, l in enumerate(labels) if i != pos]
    new_shape = [s for i, s in enumerate(shape) if i != pos]
    return new_labels, new_shape, np.ascontiguousarray(out).reshape(-1)


def contract_path(tensors: List[Tuple[Sequence[str], Sequence[int], np.ndarray]],
                  path: Sequence[Tuple[int, int]], nthreads: Optional[int] = None):
    """TensorNetwork::Contract(path) in FP64 (SSA ids: step i makes node n+i)."""
    nodes = [(list(l), list(s), np.asarray(d, dtype=np.complex128).reshape(-1))
             for (l, s, d) in tensors]
    for (i, j) in path:
        la, sa, a = nodes[i]
        lb, sb, b = nodes[j]
        nodes.append(contract(la, sa, a, lb, sb, b, nthreads))
    return nodes[-1]


# ---------------------------------------------------------------------------
# CPU baseline: the reference CPU path (QFlex transposes + OpenBLAS CBLAS),
# restated in cpu_baseline.c.  Used only by bench.py's cpu_baseline leg.
# ---------------------------------------------------------------------------

_blas_path = None


def find_openblas() -> Optional[Tuple[str, str]]:
    """(path, symbol prefix) of an OpenBLAS with CBLAS symbols, or None."""
    import glob
    cands = []
    try:
        import scipy
        cands += [(p, "scipy_") for p in glob.glob(os.path.join(
            os.path.dirname(os.path.dirname(scipy.__file__)), "scipy.libs", "libscipy_openblas*.so"))]
    except ImportError:
        pass
    cands += [(p, "") for p in glob.glob("/usr/lib/x86_64-linux-gnu/libopenblas*.so*")]
    return cands[0] if cands else None


def load_blas(threads: int) -> str:
    global _blas_path
    if _blas_path is None:
        found = find_openblas()
        if found is None:
            raise OracleError("no OpenBLAS with CBLAS symbols found")
        rc = lib().jb_load_blas(found[0].encode(), found[1].encode())
        if rc:
            raise OracleError(f"could not load CBLAS from {found[0]} ({rc})")
        _blas_path = found[0]
    lib().jb_set_blas_threads(int(threads))
    return _blas_path


def baseline_contract(a_labels, a_shape, a, b_labels, b_shape, b, nthreads: int = 1):
    """Reference CPU ContractTensors restated (output in the input dtype)."""
    tab = LabelTable()
    la = [tab.id(s) for s in a_labels]
    lb = [tab.id(s) for s in b_labels]
    left = vector_subtraction(a_labels, b_labels)
    right = vector_subtraction(b_labels, a_labels)
    size = 1
    for s in left:
        size *= a_shape[list(a_labels).index(s)]
    for s in right:
        size *= b_shape[list(b_labels).index(s)]
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b)
    C = np.empty(size, dtype=a.dtype)
    err = lib().jb_contract(a.itemsize, len(la), _i32(la), _i64(a_shape), a.ctypes.data,
                            len(lb), _i32(lb), _i64(b_shape), b.ctypes.data, C.ctypes.data,
                            int(nthreads))
    if err:
        raise OracleError(_ERRORS.get(err, str(err)))
    return left + right, C


# ---------------------------------------------------------------------------
# Environments (adjoint contraction) in complex128.  The reference has no
# gradient code; this is the CPU reference gradient the GPU adjoint is checked
# against.  It is pinned in tests by the exact derivative of a statevector
# simulation (the circuit with one gate replaced by its derivative matrix).
# ---------------------------------------------------------------------------

def environments(tensors: List[Tuple[Sequence[str], Sequence[int], np.ndarray]],
                 path: Sequence[Tuple[int, int]], wanted: Sequence[int],
                 nthreads: Optional[int] = None):
    """Returns (A, {leaf: E_leaf}) for a closed network: A the full
    contraction along the SSA path, E_leaf = dA/d(leaf) as complex128 data in
    the leaf's own index order."""
    n = len(tensors)
    nodes = [(list(l), list(s), np.asarray(d, dtype=np.complex128).reshape(-1))
             for (l, s, d) in tensors]
    children = {}
    for i, (a, b) in enumerate(path):
        children[n + i] = (a, b)
        la, sa, da = nodes[a]
        lb, sb, db = nodes[b]
        nodes.append(contract(la, sa, da, lb, sb, db, nthreads))
    root = n + len(path) - 1
    need = set(wanted)
    for i, (a, b) in enumerate(path):
        if a in need or b in need:
            need.add(n + i)
    env = {root: ([], [], np.ones(1, np.complex128))}
    out = {}
    for i in range(len(path) - 1, -1, -1):
        z = n + i
        if z not in env:
            continue
        lz, sz, ez = env.pop(z)
        for (u, v) in (children[z], children[z][::-1]):
            if u not in need:
                continue
            lv, sv, dv = nodes[v]
            le, se, de = contract(lz, sz, ez, lv, sv, dv, nthreads)
            lu, su, _ = nodes[u]
            perm = [le.index(l) for l in lu]
            de = np.ascontiguousarray(de.reshape(se).transpose(perm)).reshape(-1) if se else de
            if u < n:
                out[u] = de
            else:
                env[u] = (lu, su, de)
    return complex(nodes[root][2][0]), out

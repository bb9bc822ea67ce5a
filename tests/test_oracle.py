"""Pins the oracle (CPU restatement) before it is trusted as the checker.

What pins what:
  * permutations: the reference's literal fixtures (test/Test_Permuter.cpp,
    extracted to tests/golden/permuter_fixtures.json), the reference's own
    Default/QFlex permuters compiled from their headers (oracle/_ref, when
    built), and numpy.transpose (a textbook routine) on random shapes;
  * QFlex error conditions: the reference's exception texts
    (Test_Permuter.cpp:344-387, 492-519) and oracle/_ref;
  * ContractTensors: literal values of test/Test_Tensor.cpp:348-564 and
    numpy.einsum on random tensors (a transposed operand or a dropped term
    fails it);
  * compensated sum: exact rational arithmetic (fractions) on adversarial
    cancellations.
"""
import fractions
import string

import numpy as np
import pytest

import oracle


def _case_arrays(golden, case):
    data = np.array([complex(a, b) for a, b in golden["arrays"][case["input"]]], dtype=np.complex64)
    exp = np.array([complex(a, b) for a, b in case["expected"]], dtype=np.complex64)
    old, new = case["old_labels"], case["new_labels"]
    return data, exp, old, new, [old.index(l) for l in new]


def test_default_permuter_matches_reference_fixtures(golden_permuter):
    for case in golden_permuter["cases"]:
        data, exp, _, _, order = _case_arrays(golden_permuter, case)
        assert np.array_equal(oracle.permute_default(data, case["shape"], order), exp), case


@pytest.mark.parametrize("blocksize,min_dims", [(1024, 32), (64, 32), (128, 32), (256, 32),
                                                (512, 32), (2048, 32), (1024, 64), (1024, 128),
                                                (1024, 256), (1024, 512), (2048, 1024),
                                                (1024, 2)])
def test_qflex_restatement_matches_reference_fixtures(golden_permuter, blocksize, min_dims):
    for case in golden_permuter["cases"]:
        data, exp, _, _, order = _case_arrays(golden_permuter, case)
        pow2 = all(s & (s - 1) == 0 for s in case["shape"])
        if not pow2:
            with pytest.raises(oracle.OracleError, match="Fast transpose expects power-of-2 data"):
                oracle.permute_qflex(data, case["shape"], order, blocksize, min_dims)
            continue
        if blocksize == 64 and len(case["shape"]) == 9:
            # Test_Permuter.cpp:492-519
            with pytest.raises(oracle.OracleError, match="Consider increasing the BLOCKSIZE"):
                oracle.permute_qflex(data, case["shape"], order, blocksize, min_dims)
            continue
        assert np.array_equal(oracle.permute_qflex(data, case["shape"], order, blocksize,
                                                   min_dims), exp), case


def test_permutations_match_numpy_transpose():
    rng = np.random.default_rng(11)
    for _ in range(200):
        r = int(rng.integers(1, 9))
        shape = [int(x) for x in rng.integers(1, 5, size=r)]
        perm = [int(p) for p in rng.permutation(r)]
        x = (rng.standard_normal(int(np.prod(shape))) + 1j * rng.standard_normal(int(np.prod(shape))))
        x = x.astype(np.complex128)
        want = np.ascontiguousarray(np.transpose(x.reshape(shape), perm)).reshape(-1)
        assert np.array_equal(oracle.permute_default(x, shape, perm), want)
        if all(s & (s - 1) == 0 for s in shape):
            assert np.array_equal(oracle.permute_qflex(x, shape, perm, 64, 2), want)


def test_qflex_restatement_matches_reference_build():
    if oracle.ref_lib() is None:
        pytest.skip("oracle/_ref not built (reference tree absent)")
    rng = np.random.default_rng(5)
    qflex_params = [(64, 32), (128, 32), (256, 32), (512, 32), (1024, 32), (2048, 32), (1024, 2),
               (1024, 64)]
    kinds = set()
    for t in range(120):
        r = int(rng.integers(2, 16))
        shape = [2 if rng.random() < 0.8 else 4 for _ in range(r)]
        while np.prod(shape) > 2 ** 15:
            shape.pop()
        r = len(shape)
        perm = [int(p) for p in rng.permutation(r)]
        labels = [string.ascii_letters[i] for i in range(r)]
        new = [labels[p] for p in perm]
        bs, md = qflex_params[t % len(qflex_params)]
        n = int(np.prod(shape))
        x = (rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(np.complex64)
        try:
            ref = oracle.ref_permute(x, shape, labels, new, "qflex", bs, md)
        except oracle.OracleError as e:
            with pytest.raises(oracle.OracleError) as got:
                oracle.permute_qflex(x, shape, perm, bs, md)
            assert str(got.value).split(".")[0] in str(e)
            kinds.add("error")
            continue
        got = oracle.permute_qflex(x, shape, perm, bs, md)
        kinds.add(oracle.qflex_plan_kinds(shape, perm, bs, md))
        assert np.array_equal(got, ref)
        assert np.array_equal(oracle.ref_permute(x, shape, labels, new, "default", 1024), ref)
    # The sample exercises the single-pass, matched-suffix and 3-pass plans.
    assert {"R", "LRL"} <= kinds


# --- ContractTensors (Tensor.hpp:708-752) ---------------------------------

@pytest.mark.parametrize("dtype", [np.complex64, np.complex128])
def test_contract_reference_literals(dtype):
    # Test_Tensor.cpp:477-498: T0(a,b) . T1(b) -> T2(a)
    l, s, c = oracle.contract(["a", "b"], [2, 2], np.array([0, 1, 1, 0], dtype),
                              ["b"], [2], np.array([1, 0], dtype))
    assert l == ["a"] and s == [2] and np.array_equal(c, [0, 1])
    # Test_Tensor.cpp:500-521: T0(a) . T1(a,b) -> T2(b)
    l, s, c = oracle.contract(["a"], [2], np.array([0, 1], dtype),
                              ["a", "b"], [2, 2], np.array([0, 1, 2, 3], dtype))
    assert l == ["b"] and np.array_equal(c, [2, 3])
    # Test_Tensor.cpp:523-543: T0(a,b,c) . T1(b,c,d) -> T2(a,d) = 2.25+3i
    l, s, c = oracle.contract(["a", "b", "c"], [2, 3, 4], np.full(24, 0.5 + 0.25j, dtype),
                              ["b", "c", "d"], [3, 4, 2], np.full(24, 0.5 + 0.25j, dtype))
    assert l == ["a", "d"] and s == [2, 2] and np.array_equal(c, [2.25 + 3j] * 4)
    # Test_Tensor.cpp:544-563: full contraction -> scalar 4
    l, s, c = oracle.contract(["a", "b"], [2, 2], np.ones(4, dtype), ["a", "b"], [2, 2],
                              np.ones(4, dtype))
    assert l == [] and s == [] and np.array_equal(c, [4])
    # Test_Tensor.cpp:348-364 (gemv): [[.5,.5],[.5,-.5]] . [1,0] = [.5,.5]
    l, s, c = oracle.contract(["i", "j"], [2, 2], np.array([.5, .5, .5, -.5], dtype),
                              ["j"], [2], np.array([1, 0], dtype))
    assert np.array_equal(c, [0.5, 0.5])


def test_contract_matches_einsum():
    rng = np.random.default_rng(3)
    letters = string.ascii_letters
    for _ in range(150):
        na, nb = int(rng.integers(0, 6)), int(rng.integers(0, 6))
        pool = list(letters[:10])
        la = list(rng.choice(pool, size=na, replace=False))
        lb = list(rng.choice(pool, size=nb, replace=False))
        dims = {k: int(rng.integers(1, 4)) for k in pool}
        sa, sb = [dims[x] for x in la], [dims[x] for x in lb]
        a = (rng.standard_normal(int(np.prod(sa))) + 1j * rng.standard_normal(int(np.prod(sa))))
        b = (rng.standard_normal(int(np.prod(sb))) + 1j * rng.standard_normal(int(np.prod(sb))))
        lc, sc, c = oracle.contract(la, sa, a, lb, sb, b)
        left = [x for x in la if x not in lb]
        right = [x for x in lb if x not in la]
        assert lc == left + right
        want = np.einsum(f"{''.join(la)},{''.join(lb)}->{''.join(lc)}", a.reshape(sa),
                         b.reshape(sb)).reshape(-1)
        np.testing.assert_allclose(c, want, rtol=1e-12, atol=1e-12)


def test_gemm_entries_match_gemm():
    rng = np.random.default_rng(0)
    A = (rng.standard_normal((37, 53)) + 1j * rng.standard_normal((37, 53))).astype(np.complex64)
    B = (rng.standard_normal((53, 29)) + 1j * rng.standard_normal((53, 29))).astype(np.complex64)
    full = oracle.gemm(A, B)
    np.testing.assert_allclose(full, A.astype(np.complex128) @ B.astype(np.complex128),
                               rtol=1e-13)
    rows = rng.integers(0, 37, 40)
    cols = rng.integers(0, 29, 40)
    assert np.array_equal(oracle.gemm_entries(A, B, rows, cols), full[rows, cols])


def test_sum_compensated_exact():
    rng = np.random.default_rng(2)
    for _ in range(20):
        big = rng.standard_normal(500) * 1e16
        small = rng.standard_normal(500)
        x = np.concatenate([big, small, -big]) + 1j * np.concatenate([small, big, -big])
        rng.shuffle(x)
        exact_re = sum(fractions.Fraction(float(v)) for v in x.real)
        exact_im = sum(fractions.Fraction(float(v)) for v in x.imag)
        got = oracle.sum_compensated(x)
        assert abs(got.real - float(exact_re)) <= 1e-9 * max(1.0, abs(float(exact_re)))
        assert abs(got.imag - float(exact_im)) <= 1e-9 * max(1.0, abs(float(exact_im)))


def test_unravel_index_reference_literals():
    # Test_Utilities.cpp:453-495.
    assert oracle.unravel_index(0, []) == []
    want = {0: [0, 0], 1: [0, 1], 2: [0, 2], 3: [1, 0], 4: [1, 1], 5: [1, 2]}
    for i, w in want.items():
        assert oracle.unravel_index(i, [2, 3]) == w
    for idx, shape in ((6, [2, 3]), (0, [1, 0, 1])):
        with pytest.raises(oracle.OracleError, match="Linear index does not fit in the shape."):
            oracle.unravel_index(idx, shape)
    # Brute force: every index of a mixed-radix shape, in numpy's C order.
    shape = [3, 1, 4, 2]
    for i in range(24):
        assert oracle.unravel_index(i, shape) == list(np.unravel_index(i, shape))


def test_slice_index_reference_literals():
    # Test_Tensor.cpp:579-599.
    data = np.arange(1, 7).astype(np.complex64)
    cases = {("x", 0): (["y"], [1, 2, 3]), ("x", 1): (["y"], [4, 5, 6]),
             ("y", 0): (["x"], [1, 4]), ("y", 1): (["x"], [2, 5]), ("y", 2): (["x"], [3, 6])}
    for (label, value), (labels, vals) in cases.items():
        lab, shp, out = oracle.slice_index(["x", "y"], [2, 3], data, label, value)
        assert lab == labels and shp == [len(vals)]
        assert np.array_equal(out, np.array(vals, np.complex64))


def test_cpu_baseline_contract_matches_oracle():
    if oracle.find_openblas() is None:
        pytest.skip("no OpenBLAS with CBLAS symbols found")
    oracle.load_blas(2)
    rng = np.random.default_rng(5)
    cases = [(["a", "b", "c", "d"], ["c", "e", "a"]),   # gemm
             (["a", "b"], ["b", "a"]),                  # dot
             (["a", "b", "c"], ["c"]),                  # gemv, right empty
             (["c"], ["c", "b", "a"])]                  # gemv, left empty
    for la, lb in cases:
        sa, sb = [2] * len(la), [2] * len(lb)
        a = (rng.standard_normal(2 ** len(la)) + 1j * rng.standard_normal(2 ** len(la))
             ).astype(np.complex64)
        b = (rng.standard_normal(2 ** len(lb)) + 1j * rng.standard_normal(2 ** len(lb))
             ).astype(np.complex64)
        got_labels, got = oracle.baseline_contract(la, sa, a, lb, sb, b, 1)
        want_labels, _, want = oracle.contract(la, sa, a, lb, sb, b)
        assert got_labels == want_labels
        np.testing.assert_allclose(got, want, rtol=1e-5, atol=1e-5)

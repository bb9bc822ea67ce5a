"""Extract the literal permutation fixtures of the reference's permuter tests.

Reads test/Test_Permuter.cpp of the Jet source tree (as text; nothing is run)
and writes tests/golden/permuter_fixtures.json: every
``static const std::vector<Data> NAME{...}`` array (complex values) plus the
(shape, old labels, new labels, input array) cases of the test sections
(Test_Permuter.cpp:194-342, 469-519).

Usage: python tests/golden/make_permuter_fixtures.py [/path/to/jet]
"""
import json
import os
import re
import sys

ROOT = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
SRC = os.path.join(ROOT, "test", "Test_Permuter.cpp")
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "permuter_fixtures.json")


def main():
    text = open(SRC).read()
    arrays = {}
    for m in re.finditer(r"static const std::vector<Data>\s+(\w+)\{(.*?)\};", text, re.S):
        name, body = m.group(1), m.group(2)
        vals = [[float(a), float(b)] for a, b in re.findall(r"\{\s*(-?[\d.]+)\s*,\s*(-?[\d.]+)\s*\}", body)]
        arrays[name] = vals
    # Sections: (shape, input array, old labels, expected = FillArray(n)) as
    # written in the tests; new order is always alphabetical a,b,c,...
    cases = [
        # 2x2x2x2 (Test_Permuter.cpp:194-229)
        ([2, 2, 2, 2], "data_pow2_dbca", ["d", "b", "c", "a"]),
        ([2, 2, 2, 2], "data_pow2_cbad", ["c", "b", "a", "d"]),
        ([2, 2, 2, 2], "data_pow2_bacd", ["b", "a", "c", "d"]),
        ([2, 2, 2, 2], "data_pow2_badc", ["b", "a", "d", "c"]),
        # 4x4 (230-241)
        ([4, 4], "data_pow2_4x4_ba", ["b", "a"]),
        # 2x4x2 (242-266)
        ([2, 4, 2], "data_pow2_2x4x2_bac", ["b", "a", "c"]),
        ([2, 4, 2], "data_pow2_2x4x2_acb", ["a", "c", "b"]),
        ([2, 4, 2], "data_pow2_dbca", ["c", "b", "a"]),
        # 4x2x2 (267-291)
        ([4, 2, 2], "data_pow2_4x2x2_bac", ["b", "a", "c"]),
        ([4, 2, 2], "data_pow2_4x2x2_acb", ["a", "c", "b"]),
        ([4, 2, 2], "data_pow2_4x2x2_cba", ["c", "b", "a"]),
        # 2x2x4 (292-316)
        ([2, 2, 4], "data_pow2_bacd", ["b", "a", "c"]),
        ([2, 2, 4], "data_pow2_2x2x4_acb", ["a", "c", "b"]),
        ([2, 2, 4], "data_pow2_2x2x4_cba", ["c", "b", "a"]),
        # non power of 2 (319-342); the shape passed is the input's shape
        ([3, 5, 2], "data_non_pow2_bca", ["b", "c", "a"]),
        ([5, 3, 2], "data_non_pow2_cba", ["c", "b", "a"]),
        ([5, 2, 3], "data_non_pow2_cab", ["c", "a", "b"]),
        # rank 9 (469-489)
        ([2] * 9, "large_example", ["d", "g", "h", "b", "i", "a", "f", "c", "e"]),
    ]
    # The 'shape' used by the tests is the shape vector passed to Transpose,
    # i.e. the dimensions of the INPUT in its old label order.  For the
    # pow-2 sections the tests pass the expected (output) shape; for the
    # 2x2x2x2 cases all dims are equal so it makes no difference, but for
    # 2x4x2 etc. the test passes `shape` = output shape {2,4,2} together with
    # old labels; the reference permuter only uses it as per-old-label dims.
    out_cases = []
    for shape, arr, old in cases:
        n = len(arrays[arr])
        out_cases.append({
            "shape": shape,
            "input": arr,
            "old_labels": old,
            "new_labels": sorted(old),
            "expected": [[float(i), float(i)] for i in range(n)],
        })
    json.dump({"source": "test/Test_Permuter.cpp", "arrays": arrays, "cases": out_cases},
              open(OUT, "w"), indent=0)
    print("wrote", OUT, len(arrays), "arrays", len(out_cases), "cases")


if __name__ == "__main__":
    main()

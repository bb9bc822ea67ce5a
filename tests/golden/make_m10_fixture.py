"""The reference's m10 network as a fixture, with complex128 CPU values.

Copies examples/paper_benchmarks/data_files/m10.json of the Jet source tree
(a TensorNetworkSerializer file: 322 rank-3 bond-2 tensors of a Sycamore
m=10 circuit and a 321-step SSA contraction path, read as text) to
tests/golden/m10.json, and writes tests/golden/m10_values.json:

  * the whole network contracted along the stored path (a scalar), from the
    complex64-rounded data and from the file's values read in complex128,
    and
  * the 16 slices over the first four entries of the benchmark's sliced-
    index list (examples/paper_benchmarks/CPU/jet_cpu_m10/jet_sliced.cpp:
    53-54: p7, s7, h4, m1), slice s fixing them to UnravelIndex(s, [2]*4),
    first index most significant (TensorNetwork.hpp:230),

each contracted by the oracle in complex128 (oracle.contract along the
path, FP64 accumulation) from the file's data widened from complex64, as
the benchmark loads it (Tensor<c_fp32>, jet_sliced.cpp:33).  A few minutes
on 8 cores.

Usage: python tests/golden/make_m10_fixture.py [/path/to/jet]
"""
import json
import os
import shutil
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, HERE)

import oracle  # noqa: E402
from make_c2_slices import slice_values  # noqa: E402

SRC_ROOT = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
SRC = os.path.join(SRC_ROOT, "examples", "paper_benchmarks", "data_files", "m10.json")
NET = os.path.join(HERE, "m10.json")
OUT = os.path.join(HERE, "m10_values.json")
SLICED = ["p7", "s7", "h4", "m1"]


class _Net:
    def __init__(self, doc):
        self.indices = [list(t[1]) for t in doc["tensors"]]
        self.shapes = [list(t[2]) for t in doc["tensors"]]
        self.data = [np.array([complex(re, im) for re, im in t[3]], np.complex64)
                     for t in doc["tensors"]]


def main():
    shutil.copyfile(SRC, NET)
    with open(NET) as f:
        doc = json.load(f)
    net = _Net(doc)
    path = [tuple(p) for p in doc["path"]]
    tensors = [(ix, sh, d.astype(np.complex128)) for ix, sh, d in
               zip(net.indices, net.shapes, net.data)]
    labels, _, full = oracle.contract_path(tensors, path)
    assert labels == [], labels
    print(f"whole network: {complex(full[0])!r}", flush=True)
    # The file's decimal values read straight into complex128 (what a
    # complex128 TensorNetworkSerializer holds), not rounded to complex64.
    exact = [(list(t[1]), list(t[2]), np.array([complex(re, im) for re, im in t[3]]))
             for t in doc["tensors"]]
    _, _, full128 = oracle.contract_path(exact, path)
    print(f"whole network from unrounded data: {complex(full128[0])!r}", flush=True)
    vals = slice_values(net, path, SLICED, list(range(16)), log=lambda m: print(m, flush=True))
    total = oracle.sum_compensated(np.array(vals))
    print(f"sum of slices {total!r}", flush=True)
    with open(OUT, "w") as f:
        json.dump({"network": "tests/golden/m10.json",
                   "whole": [full[0].real, full[0].imag],
                   "whole_unrounded": [full128[0].real, full128[0].imag],
                   "sliced": SLICED,
                   "slice_values": [[v.real, v.imag] for v in vals],
                   "made_by": "tests/golden/make_m10_fixture.py (oracle.contract in complex128)"},
                  f, indent=1)
    print("wrote", OUT)


if __name__ == "__main__":
    main()

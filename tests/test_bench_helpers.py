"""bench.py's host-side helpers (no device): the traffic figure it reads
from the committed rocprofv3 summaries, the fields it derives from the
measured MFMA loop ceiling, and its refusal of more than one GPU."""
import os
import re
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402


def _table_rows(summary, names):
    """(dispatches, scaled GB) of the traffic table rows of `names`, read
    independently of bench.profiled_traffic."""
    rows, in_traffic = [], False
    for line in open(os.path.join(ROOT, "profiles", summary)):
        if line.startswith("## HBM traffic"):
            in_traffic = True
        elif line.startswith("## "):
            in_traffic = False
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if in_traffic and len(cells) == 5 and cells[0] in names:
            rows.append((int(cells[1]), float(cells[4])))
    return rows


@pytest.mark.parametrize("summary", ["c2_sliced_amplitude.md", "c1_split_contraction.md",
                                     "c2_min_fill_path.md"])
def test_profiled_traffic_is_the_dispatch_weighted_table_mean(summary):
    family = "gemm_f16x3_4m_mfma"
    traffic, commit = bench.profiled_traffic(summary, family)
    rows = _table_rows(summary, bench.ROCPROF_NAME[family])
    assert rows, summary
    want = sum(n * gb for n, gb in rows) / sum(n for n, _ in rows) * 1e9
    assert traffic == pytest.approx(want, rel=1e-12)
    assert re.fullmatch(r"[0-9a-f]{7,40}", commit), commit
    # the counters' own commit line is what the bench prints
    text = open(os.path.join(ROOT, "profiles", summary)).read()
    assert f"Measured at commit {commit}." in text


def test_profiled_traffic_absent():
    assert bench.profiled_traffic("c2_sliced_amplitude.md", "no_such_family") == (None, None)
    assert bench.profiled_traffic("no_such_summary.md", "gemm_f16x3_4m_mfma") == (None, None)


def test_mfma_loop_fields():
    f = bench.mfma_loop_fields(1.5e15, 1.68e15, 1.76)
    assert f["fp16_tflops"] == 1680.0
    assert f["clock_ghz"] == 1.76
    assert f["hw_16bit_frac_of_loop"] == round(1.5 / 1.68, 4)


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 2
    assert "one GPU only" in r.stderr
    assert r.stdout == ""

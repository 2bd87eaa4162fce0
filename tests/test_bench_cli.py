"""bench.py's host-side parts without a GPU: the refusal of more than one GPU, the profiled per-launch figures it
reports (traffic and VALU instructions, keyed by workload and tile), and the FP64 issue rate they are divided by."""
import os
import subprocess
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, REPO)
import bench  # noqa: E402


@pytest.mark.parametrize("n", [2, 8])
def test_more_than_one_gpu_is_refused(n):
    r = subprocess.run([sys.executable, os.path.join(REPO, "bench.py"), "--gpus", str(n)], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and r.stdout == ""


def test_profiled_counts_match_the_committed_profile():
    traffic, valu = bench.profiled_counts(256, 5, True, (8, 8, 4))
    assert traffic == 10041548420 and valu == 2003060736
    traffic, valu = bench.profiled_counts(256, 5, True, (8, 8, 8))
    assert traffic == 10378527865 and valu == 1841119232
    assert bench.profiled_counts(128, 5, True, (8, 8, 4)) == (None, None)
    assert bench.profiled_counts(256, 6, True, (8, 8, 4)) == (None, None)


def test_valu_issue_rate():
    # 256 CUs x 4 SIMDs, one wave64 FP64 instruction per 4 cycles each, at 2.4 GHz
    assert bench.VALU_ISSUE_PER_S == 256 * 4 * 2.4e9 / 4


def test_algorithmic_bytes():
    # 8 (3 nprim + G): G = 12 face + 3 dl + 1 vol (+ 9 metric viscous, + 1 yn RANS)
    assert bench.algorithmic_bytes_per_cell_stage(5, True, False) == 320
    assert bench.algorithmic_bytes_per_cell_stage(5, False, False) == 8 * (15 + 16)
    assert bench.algorithmic_bytes_per_cell_stage(6, True, True) == 8 * (18 + 26)


def test_chemistry_instruction_count_from_the_profile():
    assert bench.chem_valu_per_attempt() == 3921.0

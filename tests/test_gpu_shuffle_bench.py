"""The wavefront-shuffle column-pass variants (tests/shuffle_bench): their
DPP lane FFT (N = N1 x 16, N1-point DFT in registers, 16-point DFT across the
lanes of a DPP row), with direct and with LDS-staged global access, must
reproduce the library's x pass on 128^3 and 256^3 FP32 grids (relative L2
difference <= 1e-5; the binary exits 1 otherwise).
The binary also prints both passes' times; those are recorded, not asserted."""
import json
import os
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "shuffle_bench", "col_shuffle")


def test_lane_shuffle_column_pass_matches_library_x_pass():
    assert os.path.exists(EXE), "build first: make -C tests/shuffle_bench (or __graft_entry__.build())"
    print("case shuffle bench 128^3 and 256^3 FP32 x pass", file=sys.stderr, flush=True)
    out = subprocess.run([EXE, "3"], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    res = json.loads(out.stdout.strip().splitlines()[-1])
    assert [g["n"] for g in res["grids"]] == [128, 256]
    for g in res["grids"]:
        assert g["rel_l2_diff"] <= 1e-5 and g["rel_l2_diff_lds_io"] <= 1e-5, g

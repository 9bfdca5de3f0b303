"""The C++ harness (tests/harness/gpu_fft_test.cpp), built against the
drop-in header include/GPU_FFT/GPU_FFT.h, run as the reference's test.cu is
run: analytic field, R2C, 1/N, peaks, C2R, max error."""
import os
import re
import subprocess

import pytest

pytestmark = pytest.mark.gpu
EXE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "harness", "gpu_fft_test")


@pytest.mark.parametrize("args", [["128", "f64"], ["64", "48", "96", "f32"], ["128", "128", "128", "f32"],
                                  ["32", "40", "30", "f64"]])
def test_harness(args):
    if not os.path.exists(EXE):
        pytest.fail("harness not built (run __graft_entry__.build())")
    r = subprocess.run([EXE] + args, capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert len(re.findall(r"rank = 0 , data at", r.stdout)) == 8
    assert "PASS" in r.stdout

"""The full-size accuracy measurement (tests/full_size_accuracy.py) on small
grids: its R2C and round-trip errors against the FP64 oracle stay within the
precision's bound (the full-size run takes minutes of oracle time)."""
import sys

import pytest

pytestmark = pytest.mark.gpu


def test_accuracy_script_on_small_grids():
    import full_size_accuracy as A
    cases = [((16, 16, 16), "f32"), ((12, 10, 8), "f64")]
    print(f"case accuracy script {cases}", file=sys.stderr, flush=True)
    out = A.accuracy(cases)
    assert set(out) == {"16x16x16_f32", "12x10x8_f64"}
    for k, v in out.items():
        assert v["tolerance"] == (1e-6 if k.endswith("f32") else 1e-13)
        assert v["r2c_rel_l2"] <= v["tolerance"] and v["round_trip_rel_l2"] <= v["tolerance"], (k, v)

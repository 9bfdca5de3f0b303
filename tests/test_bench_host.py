"""Host checks of bench.py: the one-GPU refusal and the synthetic workloads
its result lines describe (the batched mix's counts and bytes, the served
request mix).  No GPU is touched."""
import os
import subprocess
import sys
from collections import Counter

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)


def test_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120, cwd=ROOT)
    assert r.returncode == 2
    assert "one GPU only" in r.stderr
    assert r.stdout == ""


def test_batch_mix_is_the_documented_one():
    import bench
    dims = bench.batch_dims(256, seed=0)
    assert len(dims) == 256 and all(d[0] == d[1] == d[2] for d in dims)
    assert Counter(d[0] for d in dims) == {32: 159, 48: 52, 64: 27, 96: 10, 128: 4, 160: 3, 192: 1}
    # the sum of half-spectrum bytes the batched lines report (sum_S_bytes)
    assert sum(n * n * (n // 2 + 1) * 8 for n, _, _ in dims) == 223879168


def test_served_request_mix():
    import bench
    from gpu_fft_amd.serving import SERVE_SIZES
    assert len(SERVE_SIZES) == 42
    assert all(64 <= n <= 512 for n in SERVE_SIZES)
    for n in SERVE_SIZES:  # 2^a 3^b 5^c
        m = n
        for p in (2, 3, 5):
            while m % p == 0:
                m //= p
        assert m == 1, n
    reqs = bench.serve_dims(10000, seed=0)
    assert {n for n, _, _ in reqs} <= set(SERVE_SIZES)
    f64 = sum(p == "f64" for _, p, _ in reqs) / len(reqs)
    c2r = sum(d for _, _, d in reqs) / len(reqs)
    assert 0.18 < f64 < 0.22 and 0.47 < c2r < 0.53
    # log-uniform edges: about a third of the requests above 256
    big = sum(n > 256 for n, _, _ in reqs) / len(reqs)
    assert 0.28 < big < 0.38

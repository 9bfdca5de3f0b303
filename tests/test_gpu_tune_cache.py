"""Persistent autotuning: a tuned plan's choices exported as text and
imported into a cleared cache are applied by the next plan's tune without
timing (same variants, same results), and the text names this device."""
import time

import numpy as np
import pytest

import gpu_fft_amd as G
from oracle import gfft_oracle as O

pytestmark = pytest.mark.gpu


def test_exported_cache_reproduces_the_tuned_plan(tmp_path):
    import torch
    shape = (96, 128, 128)
    G.clear_tune_cache()
    buf = torch.zeros((shape[0], shape[1], 130), device="cuda:0")
    with G.Plan(*shape) as p:
        p.tune(buf, reps=2)
        d1 = p.describe()
    text = G.tune_cache_text()
    lines = text.splitlines()
    assert len(lines) >= 3 and all("|" in l and " " in l for l in lines)
    arch = torch.cuda.get_device_properties(0).gcnArchName.split(":")[0]
    assert all(l.startswith(arch) for l in lines)
    path = tmp_path / "tune.txt"
    G.save_tune_cache(path)
    G.clear_tune_cache()
    G.load_tune_cache(path)
    assert sorted(G.tune_cache_text().splitlines()) == sorted(lines)
    fp = O.pad_real(O.random_field(*shape, seed=4)).astype(np.float32)
    x = torch.from_numpy(fp).to("cuda:0")
    with G.Plan(*shape) as p:
        t0 = time.perf_counter()
        p.tune(buf, reps=20)  # served from the imported cache: no timing runs
        dt = time.perf_counter() - t0
        d2 = p.describe()
        p.r2c(x)
        torch.cuda.synchronize()
    for ax in "zyx":
        assert d2[ax]["variant"] == d1[ax]["variant"], ax
    assert d2["zy_fused"] == d1["zy_fused"]
    assert dt < 0.5, dt
    got = G.complex_view(x, *shape).cpu().numpy()
    assert O.rel_l2(got, O.r2c_p1(fp.astype(np.float64), shape[2])) <= 1e-6
    G.clear_tune_cache()


def test_new_plans_start_with_the_cached_choices():
    """A plan created after a tune of its geometry (or an imported cache)
    starts with the tuned variants, whatever the tuner chose (with few timing
    repetitions the choice may be the default); with the cache cleared it
    starts with the defaults."""
    import torch
    n = 512
    G.clear_tune_cache()
    with G.Plan(n, n, n) as p:
        fresh = {a: p.describe()[a]["variant"] for a in "zyx"}
    assert fresh == {"z": 0, "y": 0, "x": 0}
    buf = torch.zeros((n, n, 2 * (n // 2 + 1)), device="cuda:0")
    with G.Plan(n, n, n) as p:
        p.tune(buf, reps=2)
        tuned = {a: p.describe()[a]["variant"] for a in "zyx"}
    assert len([l for l in G.tune_cache_text().splitlines() if "512" in l]) >= 3  # one entry per pass
    with G.Plan(n, n, n) as p:
        assert {a: p.describe()[a]["variant"] for a in "zyx"} == tuned
    G.clear_tune_cache()
    with G.Plan(n, n, n) as p:
        assert {a: p.describe()[a]["variant"] for a in "zyx"} == fresh

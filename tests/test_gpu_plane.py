"""Fused z + y plane kernels (one LDS-resident x-plane per workgroup): the
full R2C and C2R with the fused kernels agree with the separate z and y
passes to rounding (same schedules and twiddles; the compiler may contract
multiply-adds differently in the two kernels: rel L2 <= 2e-7 FP32, 1e-15
FP64) and match the FP64 oracle (FP32 rel L2 <= 1e-6, FP64 <= 1e-13).
Includes odd Nx and Nx = 1 (the y pass then carries the R2C scale)."""
import sys

import numpy as np
import pytest

import gpu_fft_amd as G
from oracle import gfft_oracle as O

pytestmark = pytest.mark.gpu
TOL = {"f32": 1e-6, "f64": 1e-13}
AGREE = {"f32": 2e-7, "f64": 1e-15}
# every compiled plane geometry, plus odd Nx and Nx = 1
CASES = [((64, 64, 64), "f32"), ((96, 96, 96), "f32"), ((128, 128, 128), "f32"), ((160, 160, 160), "f32"),
         ((64, 64, 64), "f64"), ((96, 96, 96), "f64"), ((128, 128, 128), "f64"), ((45, 128, 128), "f64"),
         ((192, 192, 192), "f32"), ((1, 96, 96), "f32"), ((5, 64, 64), "f64"), ((27, 160, 160), "f32"),
         ((7, 192, 192), "f32"), ((144, 144, 144), "f32"), ((9, 144, 144), "f32")]


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def run(torch, shape, prec, fused, scale_r2c=1.0, scale_c2r=1.0, seed=0):
    nx, ny, nz = shape
    rdt = np.float32 if prec == "f32" else np.float64
    fp = O.pad_real(O.random_field(nx, ny, nz, seed=seed)).astype(rdt)
    buf = torch.from_numpy(fp).to("cuda:0")
    with G.Plan(nx, ny, nz, precision=prec) as p:
        d = p.describe()
        assert d["zy_plane_kernel"] == 1, shape
        p.set_fused(fused)
        assert p.describe()["zy_fused"] == int(fused)
        p.set_scale(scale_r2c, scale_c2r)
        p.r2c(buf)
        torch.cuda.synchronize()
        spec = buf.clone()
        p.c2r(buf)
        torch.cuda.synchronize()
    return fp, spec, buf


@pytest.mark.parametrize("shape,prec", CASES, ids=lambda v: "x".join(map(str, v)) if isinstance(v, tuple) else v)
def test_fused_equals_separate_passes_and_oracle(shape, prec):
    import torch
    nx, ny, nz = shape
    say(f"case plane {shape} {prec}")
    n = nx * ny * nz
    fp, spec_f, back_f = run(torch, shape, prec, True, scale_c2r=1.0 / n)
    _, spec_s, back_s = run(torch, shape, prec, False, scale_c2r=1.0 / n)
    assert O.rel_l2(spec_f.cpu().numpy(), spec_s.cpu().numpy()) <= AGREE[prec]
    assert O.rel_l2(back_f.cpu().numpy()[..., :nz], back_s.cpu().numpy()[..., :nz]) <= AGREE[prec]
    got = G.complex_view(spec_f, nx, ny, nz).cpu().numpy()
    assert O.rel_l2(got, O.r2c_p1(fp.astype(np.float64), nz)) <= TOL[prec]
    assert O.rel_l2(back_f.cpu().numpy()[..., :nz], fp[..., :nz]) <= TOL[prec]


def test_r2c_scale_lands_on_the_y_pass_when_nx_is_1():
    import torch
    shape = (1, 128, 128)
    fp, spec_f, _ = run(torch, shape, "f64", True, scale_r2c=0.25)
    _, spec_s, _ = run(torch, shape, "f64", False, scale_r2c=0.25)
    assert O.rel_l2(spec_f.cpu().numpy(), spec_s.cpu().numpy()) <= 1e-15
    got = G.complex_view(spec_f, *shape).cpu().numpy()
    assert O.rel_l2(got, 0.25 * O.r2c_p1(fp.astype(np.float64), 128)) <= 1e-13


def test_tune_picks_a_mode_and_records_both_times():
    import torch
    buf = torch.zeros((128, 128, 130), dtype=torch.float64, device="cuda:0")
    with G.Plan(128, 128, 128, precision="f64") as p:
        p.tune(buf, reps=2)
        d = p.describe()
        assert d["fused_ms"] > 0 and d["unfused_ms"] > 0
        assert d["zy_fused"] == int(d["fused_ms"] < d["unfused_ms"])


def test_set_fused_rejected_without_plane_kernel():
    with G.Plan(64, 64, 72) as p:
        assert p.describe()["zy_plane_kernel"] == 0
        with pytest.raises(G.GfftError):
            p.set_fused(True)
        p.set_fused(False)


def test_axis_3_is_the_fused_pass():
    """gfft_exec_axis(3): the fused pass alone; R2C = (3, x), C2R = (x, 3)
    give the same result as r2c / c2r; axis 3 is rejected while unfused."""
    import torch
    nx = ny = nz = 96
    fp = O.pad_real(O.random_field(nx, ny, nz, seed=9)).astype(np.float32)
    a = torch.from_numpy(fp).to("cuda:0")
    b = a.clone()
    with G.Plan(nx, ny, nz) as p:
        assert p.describe()["zy_fused"] == 1
        p.r2c(a)
        p.axis(3, False, b)
        p.axis(2, False, b)
        torch.cuda.synchronize()
        assert torch.equal(a, b)
        p.c2r(a)
        p.axis(2, True, b)
        p.axis(3, True, b)
        torch.cuda.synchronize()
        assert torch.equal(a[..., :nz], b[..., :nz])
        p.set_fused(False)
        with pytest.raises(G.GfftError):
            p.axis(3, False, b)

"""Every serving size (the 42 edges 2^a 3^b 5^c in [64, 512]) in both
precisions, per axis pass: z R2C / C2R against numpy's 1D rfft / irfft, y and
x forward / inverse against numpy's 1D fft / ifft, all evaluated in FP64 on
the pass's own input (FP32 rel L2 <= 2e-6, FP64 <= 1e-13).  Shapes are
chosen so that each pass has several tiles and a partial last tile.  A few
full cubes are checked against the FP64 oracle end to end."""
import sys

import numpy as np
import pytest

import gpu_fft_amd as G
from gpu_fft_amd.serving import SERVE_SIZES
from oracle import gfft_oracle as O

pytestmark = pytest.mark.gpu
TOL = {"f32": 2e-6, "f64": 1e-13}


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def axis_error(torch, shape, prec, axis, inverse, seed):
    nx, ny, nz = shape
    K = O.kdim(nz)
    rdt = np.float32 if prec == "f32" else np.float64
    cdt = np.complex64 if prec == "f32" else np.complex128
    host = O.splitmix64_uniform(nx * ny * 2 * K, seed).reshape(nx, ny, 2 * K)
    if axis == 0 and not inverse:
        host[:, :, nz:] = 0
    host = host.astype(rdt)
    buf = torch.from_numpy(host).to("cuda:0")
    with G.Plan(nx, ny, nz, precision=prec) as p:
        eng = p.describe()["zyx"[axis]]["engine"]
        p.axis(axis, inverse, buf)
        torch.cuda.synchronize()
    out = buf.cpu().numpy()
    x64 = host.astype(np.float64)
    if axis == 0 and not inverse:
        ref, got = np.fft.rfft(x64[:, :, :nz], axis=2), out.view(cdt)
    elif axis == 0:
        ref, got = np.fft.irfft(x64.view(np.complex128), n=nz, axis=2) * nz, out[:, :, :nz]
    else:
        a = x64.view(np.complex128)
        ax = 1 if axis == 1 else 0
        ref = np.fft.ifft(a, axis=ax) * a.shape[ax] if inverse else np.fft.fft(a, axis=ax)
        got = out.view(cdt)
    return O.rel_l2(got, ref), eng


@pytest.mark.parametrize("prec", ["f32", "f64"])
@pytest.mark.parametrize("n", SERVE_SIZES)
def test_serving_size_every_pass(n, prec):
    import torch
    say(f"case size {n} {prec}")
    cases = [((9, 12, n), 0), ((3, n, 10), 1), ((n, 5, 6), 2)]
    for shape, axis in cases:
        for inverse in (False, True):
            err, eng = axis_error(torch, shape, prec, axis, inverse, seed=n + axis)
            assert eng == "register", (n, prec, axis, eng)
            assert err <= TOL[prec], (n, prec, shape, axis, inverse, err)


@pytest.mark.parametrize("n,prec", [(486, "f32"), (405, "f64"), (250, "f32"), (135, "f64"), (75, "f32")])
def test_serving_cube_against_oracle(n, prec):
    import torch
    say(f"case cube {n} {prec}")
    rdt = np.float32 if prec == "f32" else np.float64
    fp = O.pad_real(O.random_field(n, n, n, seed=n)).astype(rdt)
    x64 = fp.astype(np.float64)
    buf = torch.from_numpy(fp).to("cuda:0")
    with G.Plan(n, n, n, precision=prec) as p:
        p.r2c(buf)
        torch.cuda.synchronize()
        got = G.complex_view(buf, n, n, n).cpu().numpy()
        want = O.r2c_p1(x64, n, workers=8)
        err = O.rel_l2(got, want)
        assert err <= (1e-6 if prec == "f32" else 1e-13), err
        p.set_scale(1.0, 1.0 / n ** 3)
        p.c2r(buf)
        torch.cuda.synchronize()
    rt = O.rel_l2(buf.cpu().numpy()[..., :n], x64[..., :n])
    assert rt <= (1e-6 if prec == "f32" else 1e-13), rt

"""Per-pass GPU parity: each axis pass (gfft_exec_axis) against the 1D DFT
along that axis, evaluated in FP64 on the pass's own input.  A failure here
names the kernel (axis, direction, size, precision) directly."""
import sys

import numpy as np
import pytest

import gpu_fft_amd as G
from oracle import gfft_oracle as O

pytestmark = pytest.mark.gpu
TOL = {"f32": 2e-6, "f64": 1e-13}


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def _run_axis(shape, prec, axis, inverse, seed=1, variant=0):
    import torch
    nx, ny, nz = shape
    K = O.kdim(nz)
    rdt = np.float32 if prec == "f32" else np.float64
    cdt = np.complex64 if prec == "f32" else np.complex128
    rng_field = O.splitmix64_uniform(nx * ny * 2 * K, seed).reshape(nx, ny, 2 * K)
    if axis == 0 and not inverse:
        host = rng_field.copy()
        host[:, :, nz:] = 0
        host = host.astype(rdt)
    else:
        host = rng_field.astype(rdt)  # arbitrary complex (nx, ny, K) bits
    buf = torch.from_numpy(host).to("cuda:0")
    with G.Plan(nx, ny, nz, precision=prec) as p:
        if variant:
            p.set_variant(axis, variant)
        p.axis(axis, inverse, buf)
        torch.cuda.synchronize()
    out = buf.cpu().numpy()
    x64 = host.astype(np.float64)
    if axis == 0 and not inverse:
        ref = np.fft.rfft(x64[:, :, :nz], axis=2)
        got = out.view(cdt)
    elif axis == 0:
        spec = x64.view(np.complex128)
        ref = np.fft.irfft(spec, n=nz, axis=2) * nz
        got = out[:, :, :nz]
    else:
        a = x64.view(np.complex128)
        ax = 1 if axis == 1 else 0
        ref = np.fft.ifft(a, axis=ax) * a.shape[ax] if inverse else np.fft.fft(a, axis=ax)
        got = out.view(cdt)
    return O.rel_l2(got, ref)


SHAPES = [(16, 768, 24), (8, 768, 16), (768, 8, 16), (4, 4, 768), (32, 512, 24), (512, 16, 16), (8, 8, 512),
          (24, 24, 24), (16, 384, 8), (8, 1024, 8), (1024, 8, 8), (4, 4, 1024), (4, 2048, 4), (4, 4, 2048),
          (4, 448, 16), (448, 4, 16), (4, 4, 448), (4, 896, 8), (896, 4, 8), (4, 4, 896)]


@pytest.mark.parametrize("prec", ["f32", "f64"])
@pytest.mark.parametrize("shape", SHAPES, ids=lambda s: "x".join(map(str, s)))
def test_each_axis_pass(shape, prec):
    bad = []
    d = G.query_plan(*shape, precision=prec)
    for axis, name in ((0, "z"), (1, "y"), (2, "x")):
        for variant in range(d[name].get("candidates", 1)):
            for inverse in (False, True):
                err = _run_axis(shape, prec, axis, inverse, variant=variant)
                say(f"axis pass {shape} {prec} axis={axis} variant={variant} inverse={inverse}: rel L2 {err:.2e}")
                if not err <= TOL[prec]:
                    bad.append((axis, variant, inverse, err))
    assert not bad, bad

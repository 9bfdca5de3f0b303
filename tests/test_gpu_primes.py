"""Edges with prime factors 7, 11 and 13 (the odd-prime butterflies): every
pass in both precisions against numpy's 1D transforms evaluated in FP64,
the smallest cases also against the oracle's O(N^2) DFT, and whole 3D R2C /
C2R round trips against the FP64 oracle (FP32 rel L2 <= 2e-6 per pass and
1e-6 end to end, FP64 <= 1e-13)."""
import sys

import numpy as np
import pytest

import gpu_fft_amd as G
from oracle import gfft_oracle as O

from test_gpu_sizes import TOL, axis_error

pytestmark = pytest.mark.gpu
EDGES = [7, 11, 13, 14, 21, 49, 56, 77, 84, 88, 91, 104, 112, 143, 168, 169, 176, 208, 224, 336, 343, 352, 416,
         448, 672, 704, 832, 896, 1001]
# register-engine variants: 7 x 2^k, 3 x 7 x 2^k, 11 x 2^k, 13 x 2^k
COMPILED = {56, 112, 224, 448, 896, 84, 168, 336, 672, 88, 176, 352, 704, 104, 208, 416, 832}


def say(*a):
    print(*a, file=sys.stderr, flush=True)


@pytest.mark.parametrize("prec", ["f32", "f64"])
@pytest.mark.parametrize("n", EDGES)
def test_prime_edge_every_pass(n, prec):
    import torch
    say(f"case prime edge {n} {prec}")
    cases = [((6, 5, n), 0), ((3, n, 10), 1), ((n, 4, 6), 2)]
    for shape, axis in cases:
        for inverse in (False, True):
            err, eng = axis_error(torch, shape, prec, axis, inverse, seed=3 * n + axis)
            assert err <= TOL[prec], (n, prec, shape, axis, inverse, eng, err)
            assert (eng == "register") == (n in COMPILED), (n, prec, axis, eng)


@pytest.mark.parametrize("shape", [(7, 11, 13), (13, 7, 22), (11, 13, 14)])
def test_small_prime_cube_against_brute_dft(shape):
    """The oracle's direct O(N^2) DFT (no FFT library) as the reference."""
    import torch
    nx, ny, nz = shape
    f = O.random_field(nx, ny, nz, seed=nx * ny * nz)
    want = O.brute_rfftn(f)
    buf = torch.from_numpy(O.pad_real(f)).to("cuda:0")
    with G.Plan(nx, ny, nz, precision="f64") as p:
        p.r2c(buf)
        torch.cuda.synchronize()
    got = G.complex_view(buf, nx, ny, nz).cpu().numpy()
    assert O.rel_l2(got, want) <= 1e-13


@pytest.mark.parametrize("shape,prec", [((7, 14, 49), "f32"), ((77, 91, 143), "f32"), ((49, 26, 91), "f64"),
                                        ((448, 448, 448), "f32"), ((143, 169, 98), "f64")])
def test_prime_cube_round_trip_against_oracle(shape, prec):
    import torch
    nx, ny, nz = shape
    say(f"case prime cube {shape} {prec}")
    rdt = np.float32 if prec == "f32" else np.float64
    tol = 1e-6 if prec == "f32" else 1e-13
    fp = O.pad_real(O.random_field(nx, ny, nz, seed=nx + ny + nz)).astype(rdt)
    buf = torch.from_numpy(fp).to("cuda:0")
    with G.Plan(nx, ny, nz, precision=prec) as p:
        p.r2c(buf)
        torch.cuda.synchronize()
        got = G.complex_view(buf, nx, ny, nz).cpu().numpy()
        err = O.rel_l2(got, O.r2c_p1(fp.astype(np.float64), nz, workers=8))
        assert err <= tol, err
        p.set_scale(1.0, 1.0 / (nx * ny * nz))
        p.c2r(buf)
        torch.cuda.synchronize()
    back = buf.cpu().numpy()[:, :, :nz]
    assert O.rel_l2(back, fp[:, :, :nz]) <= tol

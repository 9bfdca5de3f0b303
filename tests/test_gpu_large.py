"""Large grids of the benchmark configurations (FP32): 1024^3 (4 GiB real),
1024 x 512 x 2048 (K = 1025) and 768^3 (radix-3), and the largest grid
tested, 2048^3 (34 GB, more than 2^32 elements).  A full FP64 oracle at these
sizes is minutes of CPU time, so parity uses size-independent checks:
- a separable field f = a(x) b(y) c(z) whose spectrum is the closed form
  A(kx) B(ky) C(kz) (three 1D FP64 DFTs), for both directions;
- the round trip C2R(R2C f) / N = f on a random field;
- the reference plan geometry on hipFFT (the reference's arithmetic) on the
  same random field."""
import ctypes
import os
import sys

import numpy as np
import pytest

import gpu_fft_amd as G
from oracle import gfft_oracle as O

pytestmark = [pytest.mark.gpu, pytest.mark.slow]
HERE = os.path.dirname(os.path.abspath(__file__))
SHAPES = [(1024, 1024, 1024), (1024, 512, 2048), (768, 768, 768)]


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def rel(torch, a, b):
    return float(torch.linalg.vector_norm((a - b).reshape(-1)) / torch.linalg.vector_norm(b.reshape(-1)))


# 2048^3 FP32: 34 GB, 8.6e9 reals (> 2^32): 64-bit offsets end to end
@pytest.mark.parametrize("shape", SHAPES + [(2048, 2048, 2048)], ids=lambda s: "x".join(map(str, s)))
def test_separable_closed_form(shape):
    import torch
    nx, ny, nz = shape
    K = O.kdim(nz)
    say(f"case separable {shape}")
    a = O.splitmix64_uniform(nx, 1)
    b = O.splitmix64_uniform(ny, 2)
    c = O.splitmix64_uniform(nz, 3)
    dev = "cuda:0"
    ta, tb, tc = (torch.from_numpy(v).to(dev) for v in (a, b, c))
    buf = torch.zeros((nx, ny, 2 * K), dtype=torch.float32, device=dev)
    buf[:, :, :nz] = (ta[:, None, None] * tb[None, :, None] * tc[None, None, :]).float()
    A, B, C = (torch.from_numpy(np.fft.fft(v)).to(dev) for v in (a, b, c))
    with G.Plan(nx, ny, nz) as p:
        p.r2c(buf)
        torch.cuda.synchronize()
        got = G.complex_view(buf, nx, ny, nz)
        err = 0.0
        for x0 in range(0, nx, 128):  # compare in x blocks (the FP64 outer product is large)
            want = A[x0:x0 + 128, None, None] * B[None, :, None] * C[None, None, :K]
            err = max(err, rel(torch, got[x0:x0 + 128].to(torch.complex128), want))
        say(f"  r2c rel L2 vs closed form {err:.2e}")
        assert err <= 1e-6
        # inverse of the exact spectrum gives N f
        for x0 in range(0, nx, 128):
            got[x0:x0 + 128] = (A[x0:x0 + 128, None, None] * B[None, :, None] * C[None, None, :K]).to(torch.complex64)
        p.c2r(buf)
        torch.cuda.synchronize()
        n = nx * ny * nz
        err2 = 0.0
        for x0 in range(0, nx, 128):
            want = ta[x0:x0 + 128, None, None] * tb[None, :, None] * tc[None, None, :] * n
            err2 = max(err2, rel(torch, buf[x0:x0 + 128, :, :nz].double(), want))
        say(f"  c2r rel L2 vs closed form {err2:.2e}")
        assert err2 <= 1e-6
    del buf
    torch.cuda.empty_cache()


@pytest.mark.parametrize("shape", SHAPES, ids=lambda s: "x".join(map(str, s)))
def test_roundtrip_and_hipfft_geometry(shape):
    import torch
    nx, ny, nz = shape
    K = O.kdim(nz)
    say(f"case roundtrip+hipfft {shape}")
    g = torch.Generator(device="cuda:0")
    g.manual_seed(7)
    x = torch.rand((nx, ny, 2 * K), dtype=torch.float32, device="cuda:0", generator=g) * 2 - 1
    x[:, :, nz:] = 0
    y = x.clone()
    with G.Plan(nx, ny, nz) as p:
        p.r2c(y)
        path = os.path.join(HERE, "hipfft_ref", "libgfft_hipfft_ref.so")
        if os.path.exists(path):
            L = ctypes.CDLL(path)
            L.gref_r2c.argtypes = [ctypes.c_int64] * 3 + [ctypes.c_int, ctypes.c_void_p]
            z = x.clone()
            torch.cuda.synchronize()
            assert L.gref_r2c(nx, ny, nz, 0, z.data_ptr()) == 0
            err = rel(torch, G.complex_view(y, nx, ny, nz), G.complex_view(z, nx, ny, nz))
            say(f"  r2c rel L2 vs hipFFT geometry {err:.2e}")
            assert err <= 1e-6
            del z
            torch.cuda.empty_cache()
        p.c2r(y)
        torch.cuda.synchronize()
    err2 = rel(torch, y[:, :, :nz] / (nx * ny * nz), x[:, :, :nz])
    say(f"  round trip rel L2 {err2:.2e}")
    assert err2 <= 1e-6
    del x, y
    torch.cuda.empty_cache()

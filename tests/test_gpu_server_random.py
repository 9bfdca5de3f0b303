"""Serving runtime over random geometries: 13-smooth non-cubic grids with
edges up to 256 (odd edges, factors 7 / 11 / 13), FP32 and FP64, R2C and
C2R, device and host-staged requests in one mix on a staging server.  Every
result is bitwise the one the same transform gives alone through a plan and
agrees with the FP64 oracle (FP32 rel L2 <= 1e-6, FP64 <= 1e-13)."""
import sys

import numpy as np
import pytest

import gpu_fft_amd as G
from gpu_fft_amd import serving as S
from oracle import gfft_oracle as O

from test_gpu_server import alone, make_input

pytestmark = pytest.mark.gpu


def smooth(n):
    for p in (2, 3, 5, 7, 11, 13):
        while n % p == 0:
            n //= p
    return n == 1


EDGES = [n for n in range(2, 257) if smooth(n)]


def test_random_geometries_served_equal_alone_and_oracle():
    import torch
    rng = np.random.default_rng(77)
    reqs = []
    while len(reqs) < 40:
        d = tuple(int(rng.choice(EDGES)) for _ in range(3))
        if np.prod(d) <= 2_000_000:
            reqs.append((d, "f32" if rng.random() < 0.7 else "f64", int(rng.integers(2)), rng.random() < 0.3))
    print(f"case served random geometries, {len(reqs)} requests", file=sys.stderr, flush=True)
    inputs = [make_input(torch, *d, p, dr, seed=500 + i) for i, (d, p, dr, _) in enumerate(reqs)]
    work, host_in, host_out = [], [], []
    for x, (_, _, _, staged) in zip(inputs, reqs):
        if staged:
            host_in.append(x.cpu().pin_memory())
            host_out.append(torch.empty_like(host_in[-1]).pin_memory())
            work.append(None)
        else:
            host_in.append(None)
            host_out.append(None)
            work.append(x.clone())
    with S.Server(budget_bytes=4 << 30, nstreams=4, staging=True) as srv:
        for i, (d, p, dr, staged) in enumerate(reqs):
            if staged:
                srv.submit(d, p, dr, host_in[i], host_out=host_out[i], device=False, tag=i)
            else:
                srv.submit(d, p, dr, work[i], tag=i)
        res = srv.drain(timeout_s=120)
    assert len(res) == len(reqs) and all(r["status"] == 0 for r in res)
    for i, ((d, p, dr, staged), x) in enumerate(zip(reqs, inputs)):
        got = host_out[i] if staged else work[i].cpu()
        ref = alone(torch, d, p, dr, x).cpu()
        assert torch.equal(got, ref), (i, d, p, dr, staged)
        nx, ny, nz = d
        tol = 1e-6 if p == "f32" else 1e-13
        x64 = x.cpu().numpy().astype(np.float64)
        if dr == S.R2C:
            want = O.r2c_p1(x64, nz)
            err = O.rel_l2(G.complex_view(got, nx, ny, nz).numpy(), want)
        else:
            xc = x64.reshape(nx, ny, -1, 2)
            want = O.c2r_p1(xc[..., 0] + 1j * xc[..., 1], nz)
            err = O.rel_l2(got.numpy()[..., :nz], want)
        assert err <= tol, (i, d, p, dr, err)

"""Random 13-smooth grids with edges 1..256 (odd edges, factors 7 / 11 / 13,
length-1 axes), both precisions, fused and separate plane passes, some plans
autotuned first (which runs every candidate kernel of the geometry): R2C
against the FP64 oracle (numpy's pocketfft) and the normalised round trip
against the input, rel L2 <= 1e-6 (FP32) / 1e-13 (FP64).  A fixed seed."""
import sys

import numpy as np
import pytest

import gpu_fft_amd as G
from oracle import gfft_oracle as O

pytestmark = pytest.mark.gpu


def smooth(n):
    for p in (2, 3, 5, 7, 11, 13):
        while n % p == 0:
            n //= p
    return n == 1


EDGES = [n for n in range(1, 257) if smooth(n)]


def cases(count, seed):
    rng = np.random.default_rng(seed)
    out = []
    while len(out) < count:
        nx, ny, nz = (int(rng.choice(EDGES)) for _ in range(3))
        if nz < 2 or nx * ny * nz > 2_000_000:
            continue
        out.append(((nx, ny, nz), "f32" if rng.random() < 0.6 else "f64", rng.random() < 0.1, rng.random() < 0.5))
    return out


def test_random_grids_against_oracle():
    import torch
    failures = []
    for i, (dims, prec, tune, unfuse) in enumerate(cases(150, seed=2024)):
        nx, ny, nz = dims
        print(f"case {i} {dims} {prec} tune={tune}", file=sys.stderr, flush=True)
        tol = 1e-6 if prec == "f32" else 1e-13
        rdt = np.float32 if prec == "f32" else np.float64
        fp = O.pad_real(O.random_field(nx, ny, nz, seed=i)).astype(rdt)
        buf = torch.from_numpy(fp).to("cuda:0")
        with G.Plan(nx, ny, nz, precision=prec) as p:
            if tune:
                p.tune(buf.clone(), reps=1)
            if unfuse and p.describe().get("zy_plane_kernel"):
                p.set_fused(False)
            p.set_scale(1.0, 1.0 / (nx * ny * nz))
            p.r2c(buf)
            torch.cuda.synchronize()
            got = G.complex_view(buf, nx, ny, nz).cpu().numpy()
            e1 = O.rel_l2(got, O.r2c_p1(fp.astype(np.float64), nz, workers=8))
            p.c2r(buf)
            torch.cuda.synchronize()
        e2 = O.rel_l2(buf.cpu().numpy()[..., :nz], fp[..., :nz])
        if not (e1 <= tol and e2 <= tol):
            failures.append((dims, prec, tune, e1, e2))
    assert not failures, failures

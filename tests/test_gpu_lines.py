"""GPU parity of the line-aligned strided pass (gfft_lines.h, the variants
gfft_plan_describe reports with "straight": 2).

The kernel runs the same register engine, schedule and twiddles as the
persistent column kernel of the same points per thread, so its outputs agree
with that kernel's to rounding (the compiler may contract multiply-adds
differently in the two kernels); both are also checked against numpy's FP64
FFT along the axis.  Covered: both walk directions, a base address one complex
element past a cache line (every row's chunk offset shifts), row lengths
below, at and above a multiple of the chunk, and the x-pass geometry.
"""
import sys

import numpy as np
import pytest

import gpu_fft_amd as G
from oracle import gfft_oracle as O

pytestmark = pytest.mark.gpu
TOL = {"f32": 2e-6, "f64": 1e-13}
TWIN_TOL = {"f32": 3e-7, "f64": 2e-15}  # relative L2 between the two kernels


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def _variants(shape, prec, axis):
    """(lines variant, persistent variant with the same P and radix order) pairs."""
    name = "zyx"[axis]
    out = []
    with G.Plan(*shape, precision=prec) as p:
        n = p.describe()[name]["candidates"]
        desc = []
        for v in range(n):
            p.set_variant(axis, v)
            desc.append(p.describe()[name])
    for v, d in enumerate(desc):
        if d.get("straight") != 2:
            continue
        twin = [u for u, e in enumerate(desc) if e.get("straight") == 0 and e.get("persist", 0) > 0
                and e["points_per_thread"] == d["points_per_thread"] and e["tile"] == d["tile"]
                and e["radices"] == d["radices"] and e.get("readdr", 0) == 0]
        out.append((v, twin[0] if twin else None))
    return out


def _run(shape, prec, axis, inverse, variant, offset_elems=0, walk_mask=None, seed=3):
    import torch
    nx, ny, nz = shape
    K = O.kdim(nz)
    rdt = np.float32 if prec == "f32" else np.float64
    n = nx * ny * 2 * K
    host = O.splitmix64_uniform(n + 2 * offset_elems, seed).astype(rdt)
    buf = torch.from_numpy(host).to("cuda:0")
    view = buf[2 * offset_elems:]
    with G.Plan(nx, ny, nz, precision=prec) as p:
        p.set_variant(axis, variant)
        if walk_mask is not None:
            p.set_walk_mask(walk_mask)
        p.axis(axis, inverse, view)
        torch.cuda.synchronize()
    out = buf.cpu().numpy()
    assert np.array_equal(out[:2 * offset_elems], host[:2 * offset_elems]), "wrote before the array"
    return out[2 * offset_elems:], host[2 * offset_elems:]


def _check_axis(shape, prec, axis, inverse, got, src):
    nx, ny, nz = shape
    K = O.kdim(nz)
    a = src.astype(np.float64).view(np.complex128).reshape(nx, ny, K)
    ax = 1 if axis == 1 else 0
    ref = np.fft.ifft(a, axis=ax) * a.shape[ax] if inverse else np.fft.fft(a, axis=ax)
    cdt = np.complex64 if prec == "f32" else np.complex128
    return O.rel_l2(got.view(cdt).reshape(nx, ny, K), ref)


CASES = [
    ((16, 512, 512, "f32"), 1),   # y pass, K = 257 = 16 * 16 + 1
    ((8, 512, 512, "f64"), 1),    # K = 257 = 8 * 32 + 1
    ((24, 512, 30, "f32"), 1),    # K = 16: rows exactly one chunk
    ((40, 512, 20, "f32"), 1),    # K = 11 < 16
    ((6, 512, 66, "f64"), 1),     # K = 34
    ((512, 4, 36, "f32"), 2),     # x pass, one plane of 4 * 19 columns
    ((512, 3, 16, "f64"), 2),     # x pass, 27 columns
]


@pytest.mark.parametrize("case", CASES, ids=lambda c: "x".join(map(str, c[0][:3])) + f"_{c[0][3]}_axis{c[1]}")
def test_lines_pass_matches_fft_and_persistent_twin(case):
    (nx, ny, nz, prec), axis = case
    shape = (nx, ny, nz)
    pairs = _variants(shape, prec, axis)
    assert pairs, "no line-aligned variant registered for this size"
    for v, twin in pairs:
        for inverse in (False, True):
            for off in (0, 1, 5):
                for mask in (0, 63):
                    say(f"case lines {shape} {prec} axis={axis} variant={v} inverse={inverse} off={off} mask={mask}")
                    got, src = _run(shape, prec, axis, inverse, v, off, mask)
                    err = _check_axis(shape, prec, axis, inverse, got, src)
                    assert err <= TOL[prec], (v, inverse, off, mask, err)
                    if twin is not None:
                        ref, _ = _run(shape, prec, axis, inverse, twin, off, mask)
                        d = O.rel_l2(got.astype(np.float64), ref.astype(np.float64))
                        assert d <= TWIN_TOL[prec], (v, twin, inverse, off, mask, d)

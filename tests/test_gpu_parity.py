"""GPU parity of the in-place 3D R2C/C2R path against the FP64 oracle, the
reference's recorded output and the reference plan geometry on hipFFT.

Tolerances: relative L2 <= 1e-6 for FP32, <= 1e-13 for FP64,
against the FP64 oracle on the same seeded inputs.  Every transform goes
through libgpufft.so's C ABI.
"""
import ctypes
import json
import os
import sys

import numpy as np
import pytest

import gpu_fft_amd as G
from oracle import gfft_oracle as O

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
TOL = {"f32": 1e-6, "f64": 1e-13}


def _torch():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    return torch


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def upload(a):
    torch = _torch()
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda:0")


def gpu_r2c(f_padded, nx, ny, nz, prec):
    torch = _torch()
    dt = np.float32 if prec == "f32" else np.float64
    buf = upload(f_padded.astype(dt))
    with G.Plan(nx, ny, nz, precision=prec) as p:
        p.r2c(buf)
        torch.cuda.synchronize()
    return G.complex_view(buf, nx, ny, nz).cpu().numpy()


def gpu_c2r(spec, nx, ny, nz, prec):
    torch = _torch()
    ct = np.complex64 if prec == "f32" else np.complex128
    K = O.kdim(nz)
    c = np.ascontiguousarray(spec.astype(ct))
    buf = upload(c.view(np.float32 if prec == "f32" else np.float64).reshape(nx, ny, 2 * K))
    with G.Plan(nx, ny, nz, precision=prec) as p:
        p.c2r(buf)
        torch.cuda.synchronize()
    return buf.cpu().numpy()[:, :, :nz]


# ---------------------------------------------------------------------------
# Config 1: 128^3 FP64 analytic harness field (test.cu:86-178)
# ---------------------------------------------------------------------------

def test_analytic_128_fp64_matches_reference_output():
    torch = _torch()
    with open(os.path.join(HERE, "golden", "analytic_128_reference.json")) as fh:
        doc = json.load(fh)
    n = 128
    say("case analytic 128^3 f64")
    f = O.analytic_field(n, n, n)
    buf = upload(f)
    with G.Plan(n, n, n, precision="f64") as p:
        p.r2c(buf)
        torch.cuda.synchronize()
        spec = G.complex_view(buf, n, n, n).cpu().numpy() / n ** 3
        got = O.peaks(spec, doc["threshold"])
        want = sorted((b["x"], b["y"], b["z"], b["re"], b["im"]) for b in doc["printed_bins"])
        assert [g[:3] for g in got] == [w[:3] for w in want]
        for g, w in zip(got, want):
            assert abs(g[3] - w[3]) < 1e-12 and abs(g[4] - w[4]) < 1e-12
        mask = np.ones(spec.shape, bool)
        for x, y, z, _, _ in got:
            mask[x, y, z] = False
        assert np.abs(spec[mask]).max() < 1e-12
        # normalise on the device like test.cu:149, then invert
        buf /= n ** 3
        p.c2r(buf)
        torch.cuda.synchronize()
    err = float(np.max(np.abs(buf.cpu().numpy()[:, :, :n] - f[:, :, :n])))
    say(f"  round-trip max error {err:.3e} (reference: {doc['roundtrip_max_error_per_rank']})")
    assert err < 3 * max(doc["roundtrip_max_error_per_rank"])


# ---------------------------------------------------------------------------
# Random inputs vs the FP64 oracle, small and mid grids
# ---------------------------------------------------------------------------

SHAPES = [
    (8, 8, 8), (16, 16, 16), (32, 32, 32), (64, 48, 96), (128, 128, 128), (96, 160, 192),
    (256, 64, 32), (32, 384, 48), (24, 20, 30), (27, 25, 75), (10, 12, 15), (3, 5, 16),
    (1, 1, 16), (1, 8, 8), (8, 1, 8), (6, 10, 1), (5, 6, 2), (12, 18, 81), (20, 15, 125),
    (64, 64, 243), (40, 36, 405), (16, 768, 24), (8, 1024, 16), (4, 8, 2048), (2, 2048, 8),
]


@pytest.mark.parametrize("prec", ["f32", "f64"])
@pytest.mark.parametrize("shape", SHAPES, ids=lambda s: "x".join(map(str, s)))
def test_r2c_matches_oracle(shape, prec):
    nx, ny, nz = shape
    say(f"case r2c {shape} {prec}")
    fp = O.pad_real(O.random_field(nx, ny, nz, seed=nx * 7 + ny * 3 + nz))
    got = gpu_r2c(fp, nx, ny, nz, prec)
    ref = O.r2c(fp, nz)
    err = O.rel_l2(got, ref)
    assert err <= TOL[prec], err


@pytest.mark.parametrize("prec", ["f32", "f64"])
@pytest.mark.parametrize("shape", SHAPES, ids=lambda s: "x".join(map(str, s)))
def test_c2r_matches_oracle(shape, prec):
    nx, ny, nz = shape
    say(f"case c2r {shape} {prec}")
    spec = O.hermitian_spectrum(nx, ny, nz, seed=nx + ny + nz)
    got = gpu_c2r(spec, nx, ny, nz, prec)
    ref = O.c2r(spec, nz)[:, :, :nz]
    err = O.rel_l2(got, ref)
    assert err <= TOL[prec], err


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_c2r_ignores_imaginary_dc_and_nyquist_planes(prec):
    nx, ny, nz = 8, 12, 16
    spec = O.hermitian_spectrum(nx, ny, nz, seed=5)
    bad = spec.copy()
    pert = np.zeros((nx, ny), complex)
    pert[0, 0] = 2j  # stays purely imaginary through the inverse x, y passes
    bad[:, :, 0] += pert
    bad[:, :, nz // 2] += pert
    got = gpu_c2r(bad, nx, ny, nz, prec)
    ref = O.c2r(spec, nz)[:, :, :nz]
    assert O.rel_l2(got, ref) <= TOL[prec]


def test_plan_describe_and_bytes():
    _torch()
    with G.Plan(512, 512, 512, precision="f32") as p:
        d = p.describe()
        assert d["z"]["engine"] == "register"
        assert p.data_bytes == 512 * 512 * 257 * 8
        assert 0 < p.device_bytes < 1 << 20


def test_exec_rejects_procs_gt_1_full_transform():
    torch = _torch()
    buf = torch.zeros(16 * 16 * 18, device="cuda:0")
    with G.Plan(16, 16, 16, procs=2, rank=0) as p:
        with pytest.raises(G.GfftError) as e:
            p.r2c(buf)
        assert e.value.status == 4


def test_stream_ordering_on_side_stream():
    """Execution is stream-ordered on the caller's stream (no hidden syncs)."""
    torch = _torch()
    nx, ny, nz = 64, 64, 64
    fp = O.pad_real(O.random_field(nx, ny, nz, seed=9))
    s = torch.cuda.Stream()
    buf = upload(fp.astype(np.float32))
    torch.cuda.synchronize()
    with G.Plan(nx, ny, nz) as p, torch.cuda.stream(s):
        p.r2c(buf, stream=s)
        p.c2r(buf, stream=s)
        buf /= nx * ny * nz
    s.synchronize()
    assert np.max(np.abs(buf.cpu().numpy()[:, :, :nz] - fp[:, :, :nz])) < 1e-5


# ---------------------------------------------------------------------------
# Config 2: 512^3 FP32 and FP64, seeded uniform[-1, 1), vs the FP64 oracle
# ---------------------------------------------------------------------------

@pytest.mark.slow
@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_512_cube_vs_oracle(prec):
    torch = _torch()
    n = 512
    say(f"case 512^3 {prec}: input")
    fp = O.pad_real(O.random_field(n, n, n, seed=0))
    dt = np.float32 if prec == "f32" else np.float64
    buf = upload(fp.astype(dt))
    with G.Plan(n, n, n, precision=prec) as p:
        p.r2c(buf)
        torch.cuda.synchronize()
        got = G.complex_view(buf, n, n, n).cpu().numpy()
        say("  oracle")
        ref = O.r2c_p1(fp, n, workers=16)
        err = O.rel_l2(got, ref)
        say(f"  r2c rel L2 {err:.3e}")
        assert err <= TOL[prec], err
        del got
        p.c2r(buf)
        torch.cuda.synchronize()
    back = buf.cpu().numpy()[:, :, :n].astype(np.float64) / n ** 3
    err2 = O.rel_l2(back, fp[:, :, :n])
    say(f"  round trip rel L2 {err2:.3e}")
    assert err2 <= TOL[prec]
    # C2R alone against the oracle's inverse of the oracle spectrum
    buf2 = upload(ref.astype(np.complex64 if prec == "f32" else np.complex128).view(dt).reshape(n, n, 2 * O.kdim(n)))
    with G.Plan(n, n, n, precision=prec) as p:
        p.c2r(buf2)
        torch.cuda.synchronize()
    err3 = O.rel_l2(buf2.cpu().numpy()[:, :, :n], O.c2r_p1(ref, n, workers=16))
    say(f"  c2r rel L2 {err3:.3e}")
    assert err3 <= TOL[prec]
    del buf, buf2
    torch.cuda.empty_cache()


# ---------------------------------------------------------------------------
# The reference plan geometry on hipFFT (the reference's own arithmetic)
# ---------------------------------------------------------------------------

def _hipfft_ref():
    path = os.path.join(HERE, "hipfft_ref", "libgfft_hipfft_ref.so")
    if not os.path.exists(path):
        pytest.skip("hipFFT comparator not built")
    L = ctypes.CDLL(path)
    for f in (L.gref_r2c, L.gref_c2r):
        f.restype = ctypes.c_int
        f.argtypes = [ctypes.c_int64, ctypes.c_int64, ctypes.c_int64, ctypes.c_int, ctypes.c_void_p]
    return L


@pytest.mark.parametrize("shape,prec", [((128, 128, 128), "f64"), ((64, 48, 96), "f32"), ((512, 512, 512), "f32"),
                                        ((512, 512, 512), "f64"), ((27, 25, 75), "f64")])
def test_matches_reference_plan_geometry_on_hipfft(shape, prec):
    torch = _torch()
    L = _hipfft_ref()
    nx, ny, nz = shape
    say(f"case hipfft-geometry {shape} {prec}")
    fp = O.pad_real(O.random_field(nx, ny, nz, seed=123))
    dt = np.float32 if prec == "f32" else np.float64
    a = upload(fp.astype(dt))
    b = a.clone()
    with G.Plan(nx, ny, nz, precision=prec) as p:
        p.r2c(a)
        torch.cuda.synchronize()
        assert L.gref_r2c(nx, ny, nz, int(prec == "f64"), b.data_ptr()) == 0
        ca, cb = G.complex_view(a, nx, ny, nz), G.complex_view(b, nx, ny, nz)
        err = float(torch.linalg.vector_norm(ca - cb) / torch.linalg.vector_norm(cb))
        say(f"  r2c rel L2 vs hipFFT {err:.3e}")
        assert err <= TOL[prec]
        p.c2r(a)
        torch.cuda.synchronize()
        assert L.gref_c2r(nx, ny, nz, int(prec == "f64"), b.data_ptr()) == 0
        ra, rb = a[:, :, :nz], b[:, :, :nz]
        err2 = float(torch.linalg.vector_norm(ra - rb) / torch.linalg.vector_norm(rb))
        say(f"  c2r rel L2 vs hipFFT {err2:.3e}")
        assert err2 <= TOL[prec]
    del a, b
    torch.cuda.empty_cache()


@pytest.mark.parametrize("shape,prec", [((64, 48, 96), "f32"), ((75, 81, 125), "f64"), ((256, 256, 256), "f32"),
                                        ((36, 1, 40), "f32")])
def test_walk_order_does_not_change_results(shape, prec):
    """Alternating tile walks (the default) reorder tiles only: R2C and C2R
    outputs are bitwise those of the all-forward walk."""
    torch = _torch()
    nx, ny, nz = shape
    say(f"case walk order {shape} {prec}")
    fp = O.pad_real(O.random_field(nx, ny, nz, seed=77))
    a = upload(fp.astype(np.float32 if prec == "f32" else np.float64))
    b = a.clone()
    with G.Plan(nx, ny, nz, precision=prec) as p:
        p.r2c(a)
        p.set_walk(0)
        p.r2c(b)
        torch.cuda.synchronize()
        assert torch.equal(a, b)
        p.c2r(b)
        p.set_walk(1)
        p.c2r(a)
        torch.cuda.synchronize()
        assert torch.equal(a[..., :nz], b[..., :nz])


@pytest.mark.parametrize("shape,prec", [((48, 64, 40), "f32"), ((30, 36, 50), "f64"), ((128, 128, 128), "f32")])
def test_every_walk_mask_gives_the_same_bits(shape, prec):
    """Per-pass walk directions (gfft_plan_set_walk_mask, all 64 masks on the
    smallest shape, a spread of them otherwise) reorder tiles only: R2C and
    normalised C2R outputs are bitwise those of mask 0."""
    torch = _torch()
    nx, ny, nz = shape
    say(f"case walk masks {shape} {prec}")
    fp = O.pad_real(O.random_field(nx, ny, nz, seed=78)).astype(np.float32 if prec == "f32" else np.float64)
    masks = range(64) if nx * ny * nz < 150000 else (1, 2, 4, 8, 16, 32, 42, 49, 63)
    with G.Plan(nx, ny, nz, precision=prec) as p:
        if p.describe().get("zy_plane_kernel"):
            p.set_fused(False)
        p.set_scale(1.0, 1.0 / (nx * ny * nz))
        p.set_walk_mask(0)
        ref = upload(fp)
        p.r2c(ref)
        ref_spec = ref.clone()
        p.c2r(ref)
        for m in masks:
            a = upload(fp)
            p.set_walk_mask(m)
            p.r2c(a)
            torch.cuda.synchronize()
            assert torch.equal(a, ref_spec), m
            p.c2r(a)
            torch.cuda.synchronize()
            assert torch.equal(a[..., :nz], ref[..., :nz]), m


@pytest.mark.parametrize("shape,prec", [((48, 64, 40), "f32"), ((30, 36, 50), "f64"), ((256, 128, 96), "f32")])
def test_pass_orders_match_oracle(shape, prec):
    """The four strided pass orders (gfft_plan_set_pass_order: R2C z,y,x or
    z,x,y; C2R x,y,z or y,x,z) compute the same transform: each R2C against
    the FP64 oracle and the default order (rel L2 within the precision's
    bound), each normalised round trip against the input."""
    torch = _torch()
    nx, ny, nz = shape
    say(f"case pass orders {shape} {prec}")
    fp = O.pad_real(O.random_field(nx, ny, nz, seed=79)).astype(np.float32 if prec == "f32" else np.float64)
    want = O.r2c_p1(fp.astype(np.float64), nz)
    with G.Plan(nx, ny, nz, precision=prec) as p:
        if p.describe().get("zy_plane_kernel"):
            p.set_fused(False)
        p.set_scale(1.0, 1.0 / (nx * ny * nz))
        for order in range(4):
            p.set_pass_order(order)
            a = upload(fp)
            p.r2c(a)
            torch.cuda.synchronize()
            got = G.complex_view(a, nx, ny, nz).cpu().numpy()
            err = O.rel_l2(got, want)
            assert err <= TOL[prec], (order, err)
            p.c2r(a)
            torch.cuda.synchronize()
            err = O.rel_l2(a.cpu().numpy()[..., :nz], fp[..., :nz])
            assert err <= TOL[prec], (order, err)


def test_misaligned_data_pointers_rejected_before_any_launch():
    """Every kernel moves whole complex elements: a pointer off by one real
    is a layout error and is rejected (GFFT_ERR_INVALID_ARG) with the array
    untouched."""
    import torch
    for prec, dt, off in (("f32", torch.float32, 4), ("f64", torch.float64, 8)):
        buf = torch.zeros((4, 8, 18), dtype=dt, device="cuda:0")
        keep = buf.clone()
        with G.Plan(4, 8, 16, precision=prec) as p:
            bad = buf.data_ptr() + off
            for call in (lambda: p.r2c(bad), lambda: p.c2r(bad), lambda: p.axis(1, False, bad)):
                with pytest.raises(G.GfftError) as e:
                    call()
                assert e.value.status == 1
        torch.cuda.synchronize()
        assert torch.equal(buf, keep)

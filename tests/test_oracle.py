"""Pins the FP64 oracle (CPU only).

Each check is chosen so that a plausible mistake in the oracle fails it:
known-answer vectors of the input generator, the reference's recorded output
(image.png), brute-force DFT sums, the closed form of the analytic field, the
multi-rank gather identity (pins the transpose.cu index maps), and planted
mistakes that must be caught.
"""
import json
import os

import numpy as np
import pytest

from oracle import gfft_oracle as O

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def test_splitmix64_known_answers():
    # Published splitmix64 outputs for state 0 (Vigna's reference implementation).
    want = [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F, 0xF88BB8A8724C81EC, 0x1B39896A51A8749B]
    vals = O.splitmix64_uniform(5, 0)
    assert np.array_equal(vals, np.array([(w >> 40) * 2.0 ** -23 - 1.0 for w in want]))
    assert vals.min() >= -1.0 and vals.max() < 1.0
    # offset continues the same stream
    assert np.array_equal(O.splitmix64_uniform(3, 0, offset=2), vals[2:5])
    # FP32-exact
    assert np.array_equal(vals.astype(np.float32).astype(np.float64), vals)


def test_brute_fixtures_match_oracle():
    d = np.load(os.path.join(GOLD, "brute_small.npz"))
    for i in range(int(d["count"])):
        nx, ny, nz = (int(v) for v in d[f"shape_{i}"])
        f = d[f"in_{i}"]
        assert np.array_equal(f, O.random_field(nx, ny, nz, seed=1000 + i))
        got = O.r2c(O.pad_real(f), nz)
        np.testing.assert_allclose(got, d[f"out_{i}"], rtol=0, atol=1e-12)


@pytest.mark.parametrize("n", [1, 2, 3, 5, 7, 8, 12, 15])
def test_brute_dft_is_the_dft(n):
    # DFT of a unit impulse at m is exp(-2 pi i k m / n); of a constant, n at k=0.
    for m in range(n):
        e = np.zeros(n)
        e[m] = 1
        k = np.arange(n)
        np.testing.assert_allclose(O.brute_dft(e, 0), np.exp(-2j * np.pi * k * m / n), atol=1e-14)
    np.testing.assert_allclose(O.brute_dft(np.ones(n), 0), np.eye(1, n, 0)[0] * n, atol=1e-13)
    # inverse sign
    x = O.splitmix64_uniform(n, 5)
    np.testing.assert_allclose(O.brute_dft(O.brute_dft(x, 0, -1), 0, +1) / n, x, atol=1e-14)


@pytest.mark.parametrize("shape,procs", [((8, 8, 8), 1), ((8, 8, 8), 2), ((16, 8, 12), 4),
                                         ((12, 6, 10), 2), ((8, 16, 6), 8), ((6, 4, 5), 2), ((8, 8, 9), 4)])
def test_multirank_gather_is_rfftn(shape, procs):
    """FFT_R2C.cu:41-70 with all p ranks: the y-slabs, gathered along y,
    equal the 3D rfftn -- which holds only if the restated index maps of
    transpose.cu:41-91 and the exchange are right."""
    nx, ny, nz = shape
    f = O.random_field(nx, ny, nz, seed=7)
    nxl, nyl = nx // procs, ny // procs
    slabs = [O.pad_real(f[r * nxl:(r + 1) * nxl]) for r in range(procs)]
    out = O.gpu_fft_r2c_ranks(slabs, nx, ny, nz)
    full = O.brute_rfftn(f)
    for r in range(procs):
        assert out[r].shape == (nx, nyl, O.kdim(nz))
        np.testing.assert_allclose(out[r], full[:, r * nyl:(r + 1) * nyl, :], atol=1e-11)
    back = O.gpu_fft_c2r_ranks(out, nx, ny, nz)
    for r in range(procs):
        np.testing.assert_allclose(back[r][:, :, :nz] / (nx * ny * nz), f[r * nxl:(r + 1) * nxl], atol=1e-14)


def test_planted_index_mistake_is_caught(monkeypatch):
    """A transposed chunk map (swapping the local x and y roles) must break
    the gather identity for p > 1."""
    nx, ny, nz, procs = 8, 16, 6, 2
    f = O.random_field(nx, ny, nz, seed=3)
    good = O.chunk_transpose

    def wrong(src, nx_, ny_, nz_, p_):
        K = O.kdim(nz_)
        a = src.reshape(p_, ny_ // p_, nx_ // p_, K)
        return np.ascontiguousarray(a.transpose(0, 1, 2, 3)).reshape(-1)  # forgets the (y, x) swap

    monkeypatch.setattr(O, "chunk_transpose", wrong)
    slabs = [O.pad_real(f[r * nx // procs:(r + 1) * nx // procs]) for r in range(procs)]
    out = O.gpu_fft_r2c_ranks(slabs, nx, ny, nz)
    full = np.fft.rfftn(f)
    assert not np.allclose(out[0], full[:, :ny // procs, :], atol=1e-6)
    monkeypatch.setattr(O, "chunk_transpose", good)


def test_planted_sign_mistake_is_caught():
    """Conjugating the forward spectrum (wrong sign convention) changes the
    analytic peaks: (1,2,3) would read -i instead of the recorded +i."""
    n = 16
    f = O.analytic_field(n, n, n)
    spec = O.normalize(O.r2c(f, n), n, n, n)
    assert abs(spec[1, 2, 3] - 1j) < 1e-12
    assert abs(np.conj(spec)[1, 2, 3] - 1j) > 1.0


def test_analytic_field_matches_reference_output():
    """The reference's recorded output (image.png) at 128^3, p = 2: the eight
    printed bins; the oracle at p = 1 gives the same bins with global y."""
    with open(os.path.join(GOLD, "analytic_128_reference.json")) as fh:
        doc = json.load(fh)
    n = doc["nx"]
    f = O.analytic_field(n, n, n)
    spec = O.normalize(O.r2c(f, n), n, n, n)
    got = O.peaks(spec, doc["threshold"])
    want = sorted((b["x"], b["y"], b["z"], b["re"], b["im"]) for b in doc["printed_bins"])
    assert [g[:3] for g in got] == [w[:3] for w in want]
    for g, w in zip(got, want):
        # printed with %f: six decimals
        assert abs(g[3] - w[3]) < 5e-7 and abs(g[4] - w[4]) < 5e-7
    # every other bin is at rounding level
    mask = np.ones(spec.shape, bool)
    for x, y, z, _, _ in got:
        mask[x, y, z] = False
    assert np.abs(spec[mask]).max() < 1e-12
    # and the same list is the closed form
    assert [g[:3] for g in got] == [a[:3] for a in O.analytic_peaks(n, n, n)]
    # p = 2 rank split reproduces the per-rank printout (global y = r*Ny/p + y)
    slabs = [O.analytic_field(n, n, n, rank=r, procs=2) for r in range(2)]
    out = O.gpu_fft_r2c_ranks(slabs, n, n, n)
    printed = []
    for r in range(2):
        s = O.normalize(out[r], n, n, n)
        for x, y, z, re, im in O.peaks(s, doc["threshold"]):
            printed.append((r, x, r * (n // 2) + y, z))
    assert sorted(printed) == sorted((b["rank"], b["x"], b["y"], b["z"]) for b in doc["printed_bins"])
    # round trip: same order as the reference's 1.07e-14 / 1.42e-14
    back = O.c2r(O.r2c(f, n), n)
    err = O.roundtrip_max_error(f, back, n, n ** 3)
    assert err < 3 * max(doc["roundtrip_max_error_per_rank"])


@pytest.mark.parametrize("shape", [(4, 6, 9), (6, 10, 15), (8, 9, 25), (3, 3, 1), (1, 1, 2)])
def test_c2r_of_r2c_is_n_times_f(shape):
    nx, ny, nz = shape
    f = O.random_field(nx, ny, nz, seed=11)
    back = O.c2r(O.r2c(O.pad_real(f), nz), nz)
    np.testing.assert_allclose(back[:, :, :nz], f * (nx * ny * nz), atol=1e-12 * nx * ny * nz)


def test_c2r_drops_imaginary_dc_and_nyquist():
    """The C2R convention: Im of the kz = 0 and kz = Nz/2 bins after the x, y
    passes does not matter (numpy irfft)."""
    nx, ny, nz = 4, 6, 8
    spec = O.hermitian_spectrum(nx, ny, nz, seed=2)
    ref = O.c2r(spec, nz)
    bad = spec.copy()
    # a perturbation that stays imaginary after the inverse x,y passes
    pert = np.zeros((nx, ny), complex)
    pert[0, 0] = 3j
    bad[:, :, 0] += pert
    bad[:, :, nz // 2] += pert
    np.testing.assert_allclose(O.c2r(bad, nz), ref, atol=1e-12)


def test_parseval():
    nx, ny, nz = 6, 8, 10
    f = O.random_field(nx, ny, nz, seed=4)
    s = O.r2c(O.pad_real(f), nz)
    w = np.full(O.kdim(nz), 2.0)
    w[0] = 1.0
    w[-1] = 1.0  # even nz: Nyquist bin appears once
    energy = np.sum(w * np.abs(s) ** 2) / (nx * ny * nz)
    assert abs(energy - np.sum(f ** 2)) < 1e-10 * np.sum(f ** 2)


@pytest.mark.parametrize("shape", [(4, 6, 9), (8, 8, 8), (5, 3, 10)])
def test_p1_shortcut_equals_pipeline(shape):
    nx, ny, nz = shape
    fp = O.pad_real(O.random_field(nx, ny, nz, seed=21))
    np.testing.assert_allclose(O.r2c_p1(fp, nz), O.r2c(fp, nz), atol=1e-12)
    s = O.r2c(fp, nz)
    np.testing.assert_allclose(O.c2r_p1(s, nz), O.c2r(s, nz)[:, :, :nz], atol=1e-12)

"""Host check of the plane pass's packed DC / Nyquist column
(gpu_fft_amd/csrc/gfft_plane.h, PlanePass::y_pass): after the z pass of an
x-plane, columns kz = 0 and kz = Nz/2 are transformed along y together as
one complex column.  The pairing loops and formulas of the kernel are
restated here in numpy (pairs k, (Ny - k) mod Ny for k = 0 .. Ny/2) and
checked against separate transforms of the two columns (numpy.fft), for even
and odd Ny:

  R2C: c = x0 + i xN (x0, xN real) -> C = FFT(c); X0[k] = (C[k] + conj
       C[-k]) / 2, XN[k] = (C[k] - conj C[-k]) / 2i, times the pass's scale;
  C2R: the z pass reads only the real parts of the inverse-transformed
       columns (numpy irfft convention), so the kernel packs their Hermitian
       parts, h0 + i hN with h[k] = (a[k] + conj a[-k]) / 2, and the inverse
       transform's real and imaginary parts must equal Re IFFT(a0) and
       Re IFFT(aN) (unnormalised, as the kernel's transforms are).

A planted mistake (the conjugate dropped in the split) fails the check."""
import numpy as np
import pytest


def split_r2c(C, scale=1.0, conj=True):
    """The kernel's split of the packed column's forward transform."""
    ny = C.shape[0]
    x0 = np.empty(ny, complex)
    xn = np.empty(ny, complex)
    f = 0.5 * scale
    for p in range(ny // 2 + 1):
        k1, k2 = p, (ny - p) % ny
        c1, c2 = C[k1], C[k2]
        c2c = np.conj(c2) if conj else c2
        a = (c1 + c2c) * f          # X0[k1]
        b = (c1 - c2c) * f / 1j     # XN[k1]
        x0[k1], xn[k1] = a, b
        if k2 != k1:
            x0[k2], xn[k2] = np.conj(a), np.conj(b)
    return x0, xn


def pack_c2r(a0, an):
    """The kernel's packing of the Hermitian parts before the inverse."""
    ny = a0.shape[0]
    out = np.empty(ny, complex)
    for p in range(ny // 2 + 1):
        k1, k2 = p, (ny - p) % ny
        h0 = (a0[k1] + np.conj(a0[k2])) * 0.5
        hn = (an[k1] + np.conj(an[k2])) * 0.5
        out[k1] = h0 + 1j * hn
        if k2 != k1:
            out[k2] = np.conj(h0) + 1j * np.conj(hn)
    return out


@pytest.mark.parametrize("ny", [1, 2, 3, 8, 45, 64, 96, 128, 144])
def test_packed_dc_nyquist_column_r2c(ny):
    rng = np.random.default_rng(ny)
    x0, xn = rng.uniform(-1, 1, ny), rng.uniform(-1, 1, ny)
    C = np.fft.fft(x0 + 1j * xn)
    for scale in (1.0, 0.25):
        got0, gotn = split_r2c(C, scale)
        assert np.allclose(got0, scale * np.fft.fft(x0), atol=1e-12)
        assert np.allclose(gotn, scale * np.fft.fft(xn), atol=1e-12)


@pytest.mark.parametrize("ny", [1, 2, 3, 8, 45, 64, 96, 128, 144])
def test_packed_dc_nyquist_column_c2r(ny):
    rng = np.random.default_rng(100 + ny)
    # arbitrary (not Hermitian) spectra: only the real parts of the inverse
    # transforms are used downstream
    a0 = rng.uniform(-1, 1, ny) + 1j * rng.uniform(-1, 1, ny)
    an = rng.uniform(-1, 1, ny) + 1j * rng.uniform(-1, 1, ny)
    z = np.fft.ifft(pack_c2r(a0, an)) * ny  # unnormalised inverse, as the kernel's
    assert np.allclose(z.real, (np.fft.ifft(a0) * ny).real, atol=1e-12)
    assert np.allclose(z.imag, (np.fft.ifft(an) * ny).real, atol=1e-12)


def test_planted_missing_conjugate_fails():
    ny = 64
    rng = np.random.default_rng(7)
    x0, xn = rng.uniform(-1, 1, ny), rng.uniform(-1, 1, ny)
    got0, _ = split_r2c(np.fft.fft(x0 + 1j * xn), conj=False)
    assert not np.allclose(got0, np.fft.fft(x0), atol=1e-6)

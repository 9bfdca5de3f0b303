"""FP64 CPU restatement of GPU_FFT's in-place 3D R2C / C2R path.

TEST INFRASTRUCTURE ONLY: this module is the checker for the HIP engine.
Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s
``cpu_baseline`` leg may import it; the product never does.

What it restates (reference paths relative to the reference repository root):

* ``GPU_FFT_R2C`` -- GPU_FFT/src/FFT_implementation/FFT_R2C.cu:41-70:
  2D R2C over (y, z) per local x-plane  (FFT_R2C.cu:45, plan geometry
  GPU_FFT/src/initialize_GPU_FFT/GPU_FFT_init.cu:61-64,90), slab transpose
  (FFT_R2C.cu:48), chunk exchange with self-copy (FFT_R2C.cu:51-63), chunk
  transpose (FFT_R2C.cu:66) and the batched strided 1D C2C along x, forward
  (FFT_R2C.cu:69, plan GPU_FFT_init.cu:67-73,89).
* ``GPU_FFT_C2R`` -- FFT_C2R.cu:41-70, the mirror image, unnormalised.
* The three permutation kernels of
  GPU_FFT/src/matrix_transpose/transpose.cu:41-91, restated as explicit
  integer index maps (not as numpy reshapes) so that tests can pin them.
* The harness field and normalisation of GPU_FFT_TEST_SCRIPT/test.cu:53-65,
  105-136.

The FFT arithmetic itself lives in hipFFT/rocFFT in the reference
(FFT_definations.cu:42-78), a third-party library that is not vendored
(the reference Makefile default points at ROCm 5.0.2, GPU_FFT/Makefile:7).
Its published definition is the unnormalised DFT with sign -1 forward and +1
backward (transitions.h:107-108 map FFT_FORWARD/FFT_INVERSE to
HIPFFT_FORWARD/HIPFFT_BACKWARD); here that definition is evaluated by
pocketfft (numpy / scipy.fft) in FP64, and independently by ``brute_dft``.
For C2R input that is not exactly Hermitian, the convention is numpy's
``irfft``: the imaginary parts of the kz = 0 and (even Nz) kz = Nz/2 bins
are ignored.

Multi-rank (p > 1) behaviour is simulated in one process: every rank's
pipeline runs in turn and the MPI exchange becomes chunk copies, so the index
maps of transpose.cu can be pinned against the closed form (the gathered
y-slabs equal ``rfftn``).
"""
from __future__ import annotations

import numpy as np

try:  # scipy.fft is pocketfft with a thread pool; numpy.fft is the fallback.
    import scipy.fft as _sfft
except Exception:  # pragma: no cover
    _sfft = None


# ----------------------------------------------------------------------------
# Inputs
# ----------------------------------------------------------------------------

_MASK64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def splitmix64_uniform(n: int, seed: int, offset: int = 0) -> np.ndarray:
    """Deterministic i.i.d. values in [-1, 1) (FP64 array of length ``n``).

    Element j is ``(splitmix64(seed, offset + j) >> 40) * 2**-23 - 1``: a
    multiple of 2**-23, exactly representable in FP32 and FP64, so the FP32
    and FP64 transforms see identical inputs.  splitmix64 (Steele, Lea and
    Flood, OOPSLA 2014): z = seed + (j+1) * 0x9E3779B97F4A7C15, then
    z ^= z >> 30; z *= 0xBF58476D1CE4E5B9; z ^= z >> 27;
    z *= 0x94D049BB133111EB; z ^= z >> 31.
    """
    with np.errstate(over="ignore"):
        j = np.arange(offset + 1, offset + n + 1, dtype=np.uint64)
        z = (np.uint64(seed) + j * np.uint64(0x9E3779B97F4A7C15)) & _MASK64
        z = ((z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)) & _MASK64
        z = ((z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)) & _MASK64
        z = z ^ (z >> np.uint64(31))
    return (z >> np.uint64(40)).astype(np.float64) * (2.0 ** -23) - 1.0


def kdim(nz: int) -> int:
    """Number of half-spectrum bins along z, K = floor(Nz/2) + 1."""
    return nz // 2 + 1


def pad_real(f: np.ndarray) -> np.ndarray:
    """(nx, ny, nz) real -> in-place padded layout (nx, ny, 2K), pads = 0.

    This is the layout the reference harness builds (test.cu:124,133-134) and
    the one the in-place R2C plan assumes (NULL embed, GPU_FFT_init.cu:61-64).
    """
    nx, ny, nz = f.shape
    out = np.zeros((nx, ny, 2 * kdim(nz)), dtype=f.dtype)
    out[:, :, :nz] = f
    return out


def unpad_real(fp: np.ndarray, nz: int) -> np.ndarray:
    return fp[:, :, :nz]


def random_field(nx: int, ny: int, nz: int, seed: int) -> np.ndarray:
    """Unpadded (nx, ny, nz) FP64 field from ``splitmix64_uniform``, row-major."""
    return splitmix64_uniform(nx * ny * nz, seed).reshape(nx, ny, nz)


def analytic_field(nx: int, ny: int, nz: int, rank: int = 0, procs: int = 1) -> np.ndarray:
    """Rank ``rank``'s padded x-slab of the harness field (test.cu:105-136).

    f = 8 (sin x sin 2y sin 3z + sin 4x sin 5y sin 6z) on the uniform grid
    x_l = l * 2pi/Nx (global l), y_j = j * 2pi/Ny, z_k = k * 2pi/Nz, with the
    two pad slots per row set to zero.  Unlike test.cu:88-90 the three edges
    may differ.
    """
    nxl = nx // procs
    l = np.arange(nxl * rank, nxl * (rank + 1), dtype=np.float64)
    j = np.arange(ny, dtype=np.float64)
    k = np.arange(nz, dtype=np.float64)
    dx, dy, dz = 2 * np.pi / nx, 2 * np.pi / ny, 2 * np.pi / nz
    X = l[:, None, None] * dx
    Y = j[None, :, None] * dy
    Z = k[None, None, :] * dz
    f = 8.0 * (np.sin(1.0 * X) * np.sin(2.0 * Y) * np.sin(3.0 * Z)
               + np.sin(4.0 * X) * np.sin(5.0 * Y) * np.sin(6.0 * Z))
    return pad_real(f)


# ----------------------------------------------------------------------------
# FFT arithmetic (hipFFT's definition, FP64)
# ----------------------------------------------------------------------------

def _rfft2(a, workers):
    if _sfft is not None:
        return _sfft.rfft2(a, axes=(-2, -1), workers=workers)
    return np.fft.rfft2(a, axes=(-2, -1))


def _irfft2_unnorm(a, ny, nz, workers):
    if _sfft is not None:
        return _sfft.irfft2(a, s=(ny, nz), axes=(-2, -1), norm="forward", workers=workers)
    return np.fft.irfft2(a, s=(ny, nz), axes=(-2, -1)) * (ny * nz)


def _fft_axis0(a, inverse, workers):
    if _sfft is not None:
        if inverse:
            return _sfft.ifft(a, axis=0, norm="forward", workers=workers)
        return _sfft.fft(a, axis=0, workers=workers)
    if inverse:
        return np.fft.ifft(a, axis=0) * a.shape[0]
    return np.fft.fft(a, axis=0)


def brute_dft(a: np.ndarray, axis: int, sign: int = -1) -> np.ndarray:
    """Unnormalised DFT along ``axis`` by an explicit O(N^2) sum in FP64.

    X[k] = sum_n a[n] exp(sign * 2 pi i n k / N); the phase n*k is reduced
    mod N in integers before the exponential so the matrix is accurate to a
    few ulp.  Independent of pocketfft; used to pin it on small grids.
    """
    n = a.shape[axis]
    idx = np.arange(n)
    nk = np.outer(idx, idx) % n
    w = np.exp(sign * 2j * np.pi * nk / n)
    moved = np.moveaxis(np.asarray(a, dtype=np.complex128), axis, -1)
    return np.moveaxis(moved @ w.T, -1, axis)


def brute_rfftn(f: np.ndarray) -> np.ndarray:
    """3D R2C by three brute-force DFTs, kz truncated to K bins."""
    nz = f.shape[2]
    s = brute_dft(f, 2, -1)[:, :, :kdim(nz)]
    s = brute_dft(s, 1, -1)
    return brute_dft(s, 0, -1)


# ----------------------------------------------------------------------------
# Permutation kernels, restated as index maps (transpose.cu:41-91)
# ----------------------------------------------------------------------------

def transpose_slab(src: np.ndarray, ny_cs: int, nx_cs: int, nz: int) -> np.ndarray:
    """transpose.cu:41-54 -- ``out[x*K + y*K*Nx_cs + z] = in[i]``.

    i runs over Nx_cs*Ny_cs*K with z = i % K, y = (i / K) % Ny_cs,
    x = i / (K * Ny_cs): the outer two axes of a (Nx_cs, Ny_cs, K) array are
    swapped, kz stays innermost.  64-bit indices (the reference's 32-bit
    ``int`` locals are not replicated).
    """
    K = kdim(nz)
    total = nx_cs * ny_cs * K
    i = np.arange(total, dtype=np.int64)
    z = i % K
    y = (i // K) % ny_cs
    x = i // (K * ny_cs)
    put = x * K + y * K * nx_cs + z
    out = np.empty(total, dtype=src.dtype)
    out[put] = src.reshape(-1)[:total]
    return out


def chunk_transpose(src: np.ndarray, nx: int, ny: int, nz: int, procs: int) -> np.ndarray:
    """transpose.cu:56-73 -- unpack (p, Ny/p, Nx/p, K) -> (Nx, Ny/p, K)."""
    K = kdim(nz)
    nxl, nyl = nx // procs, ny // procs
    total = nxl * ny * K
    i = np.arange(total, dtype=np.int64)
    zz = i % K                      # Nx_no
    xl = (i // K) % nxl             # Ny_no
    row = i // (K * nxl)            # Nz_no
    s = row // nyl                  # odd_even
    yl = row % nyl                  # put_odd_even
    put = s * nxl * K * nyl + yl * K + xl * (K * ny // procs) + zz
    out = np.empty(total, dtype=src.dtype)
    out[put] = src.reshape(-1)[:total]
    return out


def chunk_transpose_inverse(src: np.ndarray, nx: int, ny: int, nz: int, procs: int) -> np.ndarray:
    """transpose.cu:75-91 -- pack (Nx, Ny/p, K) -> (p, Ny/p, Nx/p, K)."""
    K = kdim(nz)
    nxl, nyl = nx // procs, ny // procs
    total = nx * nyl * K
    i = np.arange(total, dtype=np.int64)
    zz = i % K
    yl = (i // K) % nyl
    xg = i // (K * nyl)
    s = xg // nxl
    xl = xg % nxl
    put = s * nxl * K * nyl + xl * K + yl * K * nxl + zz
    out = np.empty(total, dtype=src.dtype)
    out[put] = src.reshape(-1)[:total]
    return out


# ----------------------------------------------------------------------------
# The two drivers, all ranks simulated (FFT_R2C.cu:41-70, FFT_C2R.cu:41-70)
# ----------------------------------------------------------------------------

def _exchange(send, procs, chunk):
    """FFT_R2C.cu:51-63 / FFT_C2R.cu:51-63: rank r's send chunk i lands in
    rank i's receive chunk r (the self chunk by the D2D copy)."""
    recv = [np.empty_like(b) for b in send]
    for r in range(procs):
        for i in range(procs):
            recv[i][r * chunk:(r + 1) * chunk] = send[r][i * chunk:(i + 1) * chunk]
    return recv


def gpu_fft_r2c_ranks(slabs, nx: int, ny: int, nz: int, workers: int = 1):
    """Forward transform of the p x-slabs ``slabs[r]`` (padded real,
    (Nx/p, Ny, 2K)).  Returns p complex y-slabs, each (Nx, Ny/p, K)."""
    procs = len(slabs)
    K = kdim(nz)
    nxl, nyl = nx // procs, ny // procs
    chunk = nxl * nyl * K
    buffers = []
    for r in range(procs):
        real = np.asarray(slabs[r], dtype=np.float64).reshape(nxl, ny, 2 * K)[:, :, :nz]
        spec = _rfft2(real, workers)                                   # FFT_R2C.cu:45
        buffers.append(transpose_slab(spec.reshape(-1), ny, nxl, nz))  # FFT_R2C.cu:48
    recv = _exchange(buffers, procs, chunk)                            # FFT_R2C.cu:51-63
    out = []
    for r in range(procs):
        buf = chunk_transpose(recv[r], nx, ny, nz, procs)              # FFT_R2C.cu:66
        res = _fft_axis0(buf.reshape(nx, nyl * K), False, workers)     # FFT_R2C.cu:69
        out.append(res.reshape(nx, nyl, K))
    return out


def gpu_fft_c2r_ranks(yslabs, nx: int, ny: int, nz: int, workers: int = 1):
    """Inverse of ``gpu_fft_r2c_ranks`` without normalisation: each rank's
    (Nx, Ny/p, K) complex y-slab -> its padded real x-slab (Nx/p, Ny, 2K).
    Pad slots are returned as zero (the reference leaves them unspecified)."""
    procs = len(yslabs)
    K = kdim(nz)
    nxl, nyl = nx // procs, ny // procs
    chunk = nxl * nyl * K
    packed = []
    for r in range(procs):
        d = np.asarray(yslabs[r], dtype=np.complex128).reshape(nx, nyl * K)
        buf = _fft_axis0(d, True, workers).reshape(-1)                 # FFT_C2R.cu:45
        packed.append(chunk_transpose_inverse(buf, nx, ny, nz, procs))  # FFT_C2R.cu:48
    recv = _exchange(packed, procs, chunk)                             # FFT_C2R.cu:51-63
    out = []
    for r in range(procs):
        data = transpose_slab(recv[r], nxl, ny, nz)                    # FFT_C2R.cu:66
        spec = data.reshape(nxl, ny, K)
        real = _irfft2_unnorm(spec, ny, nz, workers)                   # FFT_C2R.cu:69
        out.append(pad_real(real))
    return out


def r2c(f_padded: np.ndarray, nz: int, workers: int = 1) -> np.ndarray:
    """p = 1 forward: padded real (Nx, Ny, 2K) -> complex (Nx, Ny, K)."""
    nx, ny = f_padded.shape[0], f_padded.shape[1]
    return gpu_fft_r2c_ranks([f_padded], nx, ny, nz, workers)[0]


def c2r(spec: np.ndarray, nz: int, workers: int = 1) -> np.ndarray:
    """p = 1 inverse, unnormalised: complex (Nx, Ny, K) -> padded real."""
    nx, ny = spec.shape[0], spec.shape[1]
    return gpu_fft_c2r_ranks([spec], nx, ny, nz, workers)[0]


def r2c_p1(f_padded: np.ndarray, nz: int, workers: int = 1, dtype=np.float64) -> np.ndarray:
    """p = 1 forward without the identity permutations: at p = 1
    transpose_slab, the self-copy and chunk_transpose compose to the identity
    (pinned by test_multirank_gather_is_rfftn with procs = 1), so the
    pipeline is rfftn of the unpadded field.  For large grids.  dtype
    float32 runs pocketfft in single precision (the CPU baseline of an FP32
    GPU run); the checker always uses float64."""
    real = np.asarray(f_padded, dtype=dtype)[..., :nz]
    if _sfft is not None:
        return _sfft.rfftn(real, workers=workers)
    return np.fft.rfftn(real)


def c2r_p1(spec: np.ndarray, nz: int, workers: int = 1) -> np.ndarray:
    """p = 1 inverse, unnormalised, unpadded (Nx, Ny, Nz) FP64 result."""
    s = tuple(spec.shape[:2]) + (nz,)
    if _sfft is not None:
        return _sfft.irfftn(spec, s=s, norm="forward", workers=workers)
    return np.fft.irfftn(spec, s=s) * float(np.prod(s))


def hermitian_spectrum(nx: int, ny: int, nz: int, seed: int) -> np.ndarray:
    """A valid C2R input: the R2C spectrum of a seeded real field."""
    return r2c(pad_real(random_field(nx, ny, nz, seed)), nz)


# ----------------------------------------------------------------------------
# Harness helpers (test.cu:53-65, 67-84, 162-174)
# ----------------------------------------------------------------------------

def normalize(spec: np.ndarray, nx: int, ny: int, nz: int) -> np.ndarray:
    """test.cu:53-65: multiply the half spectrum by 1/(Nx Ny Nz)."""
    return spec / float(nx * ny * nz)


def peaks(spec_norm: np.ndarray, threshold: float = 0.1):
    """test.cu:67-84: bins whose |re| or |im| exceeds ``threshold``, as
    sorted (x, y, z, re, im) tuples (global y at p = 1)."""
    mask = (np.abs(spec_norm.real) > threshold) | (np.abs(spec_norm.imag) > threshold)
    idx = np.argwhere(mask)
    return sorted((int(x), int(y), int(z), float(spec_norm[x, y, z].real),
                   float(spec_norm[x, y, z].imag)) for x, y, z in idx)


def analytic_peaks(nx: int, ny: int, nz: int):
    """Closed form of the normalised spectrum of ``analytic_field``.

    8 sin(ax) sin(by) sin(cz) = 8 * (e^{iax}-e^{-iax})(...)(...)/(2i)^3, so
    after 1/N each of the 8 sign combinations (sx a, sy b, sz c) carries
    i * sx * sy * sz; the half spectrum keeps sz = +1.  Returns sorted
    (x, y, z, re, im) for both modes.
    """
    out = []
    for a, b, c in ((1, 2, 3), (4, 5, 6)):
        for sx in (1, -1):
            for sy in (1, -1):
                out.append(((sx * a) % nx, (sy * b) % ny, c, 0.0, float(sx * sy)))
    return sorted(out)


def roundtrip_max_error(f_in_padded: np.ndarray, f_out_padded: np.ndarray,
                        nz: int, n_total: int) -> float:
    """test.cu:162-174, corrected: max |f - C2R(R2C f)/N| over z < Nz only
    (the reference loop also reads pad slots, test.cu:166)."""
    a = f_in_padded[..., :nz].astype(np.float64)
    b = f_out_padded[..., :nz].astype(np.float64) / float(n_total)
    return float(np.max(np.abs(a - b)))


def rel_l2(a: np.ndarray, b: np.ndarray) -> float:
    """||a - b||_2 / ||b||_2 in FP64."""
    a = np.asarray(a, dtype=np.complex128 if np.iscomplexobj(a) or np.iscomplexobj(b) else np.float64)
    b = np.asarray(b, dtype=a.dtype)
    den = np.linalg.norm(b.ravel())
    return float(np.linalg.norm((a - b).ravel()) / (den if den > 0 else 1.0))

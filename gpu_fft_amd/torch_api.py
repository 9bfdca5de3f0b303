"""torch-native functional API over the C ABI.

The reference announces a Python wrapper but ships none (Readme.md:68-70).
These functions take and return torch tensors on a ROCm device and run on
torch's current stream:

* ``rfftn(x)`` / ``irfftn(X, s)``: out of place, the semantics of
  ``torch.fft.rfftn`` / ``torch.fft.irfftn`` over the last three dimensions
  (norm "backward": forward unnormalised, inverse scaled by 1/N);
* ``rfftn_(buf, nz)`` / ``irfftn_(buf, nz)``: in place on the reference's
  padded layout (Nx, Ny, 2K) reals, returning views — the zero-copy path.

Plans (twiddle tables) are cached per (shape, precision, device).  Only 3D
float32 / float64 tensors on a GPU are accepted; there is no CPU fallback.
"""
from __future__ import annotations

import threading

from . import Plan, complex_view, kdim

__all__ = ["rfftn", "irfftn", "rfftn_", "irfftn_", "clear_plan_cache"]

_lock = threading.Lock()
_plans: dict = {}


def _prec(dtype):
    import torch
    if dtype in (torch.float32, torch.complex64):
        return "f32"
    if dtype in (torch.float64, torch.complex128):
        return "f64"
    raise TypeError(f"unsupported dtype {dtype}: float32/float64 (complex64/complex128) only")


def _check_device(t):
    if not t.is_cuda:
        raise ValueError("gpu_fft_amd transforms need a GPU tensor (no CPU path)")


def _plan(nx, ny, nz, prec, device, scale_c2r):
    import torch
    key = (nx, ny, nz, prec, torch.device(device).index, scale_c2r)
    with _lock:
        p = _plans.get(key)
        if p is None:
            with torch.cuda.device(device):
                p = Plan(nx, ny, nz, precision=prec)
                if scale_c2r != 1.0:
                    p.set_scale(1.0, scale_c2r)
            _plans[key] = p
        return p


def clear_plan_cache() -> None:
    with _lock:
        for p in _plans.values():
            p.close()
        _plans.clear()


def rfftn_(buf, nz: int):
    """In-place forward R2C of a padded (Nx, Ny, 2K) real tensor (K = nz//2+1).
    Returns the complex (Nx, Ny, K) view of the same memory."""
    _check_device(buf)
    if buf.dim() != 3 or buf.shape[2] != 2 * kdim(nz) or not buf.is_contiguous():
        raise ValueError(f"expected a contiguous (Nx, Ny, {2 * kdim(nz)}) tensor")
    nx, ny = buf.shape[0], buf.shape[1]
    _plan(nx, ny, nz, _prec(buf.dtype), buf.device, 1.0).r2c(buf)
    return complex_view(buf, nx, ny, nz)


def irfftn_(buf, nz: int, normalize: bool = False):
    """In-place inverse C2R of the padded buffer holding the (Nx, Ny, K)
    spectrum (as reals (Nx, Ny, 2K) or the complex view).  Returns the real
    (Nx, Ny, nz) view; unnormalised unless ``normalize``."""
    import torch
    _check_device(buf)
    if buf.is_complex():
        buf = torch.view_as_real(buf).reshape(buf.shape[0], buf.shape[1], -1)
    if buf.dim() != 3 or buf.shape[2] != 2 * kdim(nz) or not buf.is_contiguous():
        raise ValueError(f"expected a contiguous (Nx, Ny, {2 * kdim(nz)}) tensor")
    nx, ny = buf.shape[0], buf.shape[1]
    scale = 1.0 / (nx * ny * nz) if normalize else 1.0
    _plan(nx, ny, nz, _prec(buf.dtype), buf.device, scale).c2r(buf)
    return buf[..., :nz]


def rfftn(x):
    """torch.fft.rfftn(x) for a real (Nx, Ny, Nz) GPU tensor: a new complex
    (Nx, Ny, Nz//2+1) tensor, unnormalised."""
    import torch
    _check_device(x)
    if x.dim() != 3 or x.is_complex():
        raise ValueError("expected a real 3D tensor (Nx, Ny, Nz)")
    _prec(x.dtype)
    nx, ny, nz = x.shape
    buf = torch.empty((nx, ny, 2 * kdim(nz)), dtype=x.dtype, device=x.device)
    buf[..., :nz].copy_(x)
    return rfftn_(buf, nz)


def irfftn(X, s):
    """torch.fft.irfftn(X, s=s) for a complex (Nx, Ny, Nz//2+1) GPU tensor:
    a new real (Nx, Ny, Nz) tensor scaled by 1/(Nx Ny Nz) (norm "backward").
    Imaginary parts of the kz = 0 and Nyquist planes are ignored, as in
    numpy/torch irfftn."""
    import torch
    _check_device(X)
    nx, ny, nz = (int(v) for v in s)
    if X.dim() != 3 or not X.is_complex() or tuple(X.shape) != (nx, ny, kdim(nz)):
        raise ValueError(f"expected a complex ({nx}, {ny}, {kdim(nz)}) tensor")
    rdt = torch.float32 if X.dtype == torch.complex64 else torch.float64
    _prec(rdt)
    buf = torch.empty((nx, ny, 2 * kdim(nz)), dtype=rdt, device=X.device)
    torch.view_as_complex(buf.view(nx, ny, kdim(nz), 2)).copy_(X)
    return irfftn_(buf, nz, normalize=True).contiguous()

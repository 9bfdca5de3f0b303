"""gpu_fft_amd -- MI355X-native in-place 3D R2C/C2R FFT engine.

The compute path is ``libgpufft.so`` (hand-written HIP kernels for gfx950,
built in-tree from ``gpu_fft_amd/csrc``), reached through its C ABI
(``include/gpufft/gpufft.h``) with ctypes.  PyTorch only provides device
memory and streams.  There is no CPU fallback: if the library is missing or
fails to load, importing the transform API raises.

The reference exposes C++ templates (GPU_FFT/src/GPU_FFT/GPU_FFT.h:114-121);
``Plan`` mirrors ``INIT_GPU_FFT`` (constructor), ``GPU_FFT_R2C`` (``r2c``) and
``GPU_FFT_C2R`` (``c2r``) with the same layouts: the real view is
(Nx/p, Ny, 2K) with K = Nz//2 + 1, the complex view (Nx, Ny/p, K), in place.
"""
from __future__ import annotations

import ctypes
import json
import os

__all__ = [
    "GfftError", "Plan", "lib", "lib_path", "query_plan", "kdim", "alloc_inplace",
    "real_view", "complex_view", "F32", "F64", "tune_cache_text", "load_tune_cache_text", "clear_tune_cache",
    "save_tune_cache", "load_tune_cache", "ERR_TIMEOUT", "ERR_ABORTED", "ERR_SCRATCH_TOO_SMALL",
]

F32 = 0
F64 = 1
ERR_TIMEOUT = 11  # GFFT_ERR_TIMEOUT
ERR_ABORTED = 12  # GFFT_ERR_ABORTED
ERR_SCRATCH_TOO_SMALL = 13  # GFFT_ERR_SCRATCH_TOO_SMALL

_HERE = os.path.dirname(os.path.abspath(__file__))


def lib_path() -> str:
    """The in-tree library; GPU_FFT_AMD_LIB may point at another build of
    the same sources (developer A/B comparisons)."""
    return os.environ.get("GPU_FFT_AMD_LIB") or os.path.join(_HERE, "libgpufft.so")


class GfftError(RuntimeError):
    def __init__(self, status: int, what: str = ""):
        self.status = status
        msg = _strerror(status)
        super().__init__(f"{what}: gfft status {status} ({msg})" if what else f"gfft status {status} ({msg})")


_lib = None


def lib() -> ctypes.CDLL:
    """Load libgpufft.so (raises if it is not built)."""
    global _lib
    if _lib is not None:
        return _lib
    path = lib_path()
    if not os.path.exists(path):
        raise ImportError(f"{path} is not built; run __graft_entry__.build() or "
                          f"make -C gpu_fft_amd/csrc")
    L = ctypes.CDLL(path, mode=ctypes.RTLD_GLOBAL)
    i64, vp, sz = ctypes.c_int64, ctypes.c_void_p, ctypes.c_size_t
    L.gfft_version.restype = ctypes.c_int
    L.gfft_strerror.restype = ctypes.c_char_p
    L.gfft_strerror.argtypes = [ctypes.c_int]
    L.gfft_plan_create.restype = ctypes.c_int
    L.gfft_plan_create.argtypes = [ctypes.POINTER(vp), i64, i64, i64, ctypes.c_int, ctypes.c_int, ctypes.c_int]
    L.gfft_plan_destroy.restype = ctypes.c_int
    L.gfft_plan_destroy.argtypes = [vp]
    for name in ("gfft_exec_r2c", "gfft_exec_c2r"):
        f = getattr(L, name)
        f.restype = ctypes.c_int
        f.argtypes = [vp, vp, vp]
    for name in ("gfft_exec_r2c_local", "gfft_exec_r2c_finish", "gfft_exec_c2r_start", "gfft_exec_c2r_local"):
        f = getattr(L, name)
        f.restype = ctypes.c_int
        f.argtypes = [vp, vp, vp, vp]
    L.gfft_plan_tune.restype = ctypes.c_int
    L.gfft_plan_tune.argtypes = [vp, vp, vp, ctypes.c_int]
    L.gfft_plan_set_variant.restype = ctypes.c_int
    L.gfft_plan_set_variant.argtypes = [vp, ctypes.c_int, ctypes.c_int]
    L.gfft_plan_set_fused.restype = ctypes.c_int
    L.gfft_plan_set_fused.argtypes = [vp, ctypes.c_int]
    L.gfft_tune_cache_export.restype = ctypes.c_int
    L.gfft_tune_cache_export.argtypes = [ctypes.c_char_p, sz, ctypes.POINTER(sz)]
    L.gfft_tune_cache_import.restype = ctypes.c_int
    L.gfft_tune_cache_import.argtypes = [ctypes.c_char_p]
    L.gfft_tune_cache_clear.restype = ctypes.c_int
    L.gfft_tune_cache_clear.argtypes = []
    L.gfft_plan_set_chunking.restype = ctypes.c_int
    L.gfft_plan_set_chunking.argtypes = [vp, ctypes.c_int64]
    L.gfft_plan_set_walk.restype = ctypes.c_int
    L.gfft_plan_set_walk.argtypes = [vp, ctypes.c_int]
    L.gfft_plan_set_walk_mask.restype = ctypes.c_int
    L.gfft_plan_set_walk_mask.argtypes = [vp, ctypes.c_int]
    L.gfft_plan_set_pass_order.restype = ctypes.c_int
    L.gfft_plan_set_pass_order.argtypes = [vp, ctypes.c_int]
    L.gfft_plan_set_scale.restype = ctypes.c_int
    L.gfft_plan_set_scale.argtypes = [vp, ctypes.c_double, ctypes.c_double]
    L.gfft_exec_axis.restype = ctypes.c_int
    L.gfft_exec_axis.argtypes = [vp, ctypes.c_int, ctypes.c_int, vp, vp]
    for name in ("gfft_plan_device_bytes", "gfft_plan_data_bytes"):
        f = getattr(L, name)
        f.restype = ctypes.c_int
        f.argtypes = [vp, ctypes.POINTER(sz)]
    L.gfft_plan_describe.restype = ctypes.c_int
    L.gfft_plan_describe.argtypes = [vp, ctypes.c_char_p, sz]
    L.gfft_query_plan.restype = ctypes.c_int
    L.gfft_query_plan.argtypes = [i64, i64, i64, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_char_p, sz]
    _lib = L
    return L


def _strerror(status: int) -> str:
    try:
        return lib().gfft_strerror(status).decode()
    except Exception:  # pragma: no cover - only when the library is absent
        return "?"


def _check(status: int, what: str) -> None:
    if status != 0:
        raise GfftError(status, what)


def _prec(precision) -> int:
    if precision in (F32, "f32", "float32", "float"):
        return F32
    if precision in (F64, "f64", "float64", "double"):
        return F64
    raise ValueError(f"unknown precision {precision!r}")


def kdim(nz: int) -> int:
    """Half-spectrum length along z, K = Nz//2 + 1."""
    return nz // 2 + 1


def query_plan(nx, ny, nz, procs=1, rank=0, precision="f32") -> dict:
    """Host-only planning (no GPU needed): the plan JSON, or GfftError."""
    buf = ctypes.create_string_buffer(8192)
    _check(lib().gfft_query_plan(nx, ny, nz, procs, rank, _prec(precision), buf, len(buf)), "gfft_query_plan")
    return json.loads(buf.value.decode())


def tune_cache_text() -> str:
    """The process-wide autotuning cache (what gfft_plan_tune has chosen so
    far), one "<device>|<geometry> <kernel>" line per tuned pass."""
    need = ctypes.c_size_t(0)
    _check(lib().gfft_tune_cache_export(None, 0, ctypes.byref(need)), "gfft_tune_cache_export")
    buf = ctypes.create_string_buffer(max(1, need.value))
    _check(lib().gfft_tune_cache_export(buf, len(buf), None), "gfft_tune_cache_export")
    return buf.value.decode()


def load_tune_cache_text(text: str) -> None:
    """Add the entries of an exported cache (later plans skip their timing)."""
    _check(lib().gfft_tune_cache_import(text.encode()), "gfft_tune_cache_import")


def clear_tune_cache() -> None:
    _check(lib().gfft_tune_cache_clear(), "gfft_tune_cache_clear")


def save_tune_cache(path) -> None:
    with open(path, "w") as fh:
        fh.write("# gpu_fft_amd tuning cache\n" + tune_cache_text())


def load_tune_cache(path) -> None:
    with open(path) as fh:
        load_tune_cache_text(fh.read())


def _stream_handle(stream):
    if stream is None:
        import torch
        return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    if isinstance(stream, int):
        return ctypes.c_void_p(stream)
    return ctypes.c_void_p(stream.cuda_stream)


def _ptr(x) -> ctypes.c_void_p:
    if isinstance(x, int):
        return ctypes.c_void_p(x)
    return ctypes.c_void_p(x.data_ptr())


class Plan:
    """A transform plan on the current HIP device (mirror of INIT_GPU_FFT).

    ``r2c``/``c2r`` take a torch tensor (any view of the in-place array) or a
    raw device address and run on ``stream`` (default: torch's current
    stream), asynchronously.
    """

    def __init__(self, nx: int, ny: int, nz: int, procs: int = 1, rank: int = 0, precision="f32"):
        self.nx, self.ny, self.nz = int(nx), int(ny), int(nz)
        self.procs, self.rank = int(procs), int(rank)
        self.prec = _prec(precision)
        self.K = kdim(self.nz)
        h = ctypes.c_void_p()
        _check(lib().gfft_plan_create(ctypes.byref(h), self.nx, self.ny, self.nz, self.procs, self.rank, self.prec),
               "gfft_plan_create")
        self._h = h

    @property
    def handle(self) -> ctypes.c_void_p:
        if self._h is None:
            raise ValueError("plan destroyed")
        return self._h

    def r2c(self, data, stream=None) -> None:
        _check(lib().gfft_exec_r2c(self.handle, _ptr(data), _stream_handle(stream)), "gfft_exec_r2c")

    def c2r(self, data, stream=None) -> None:
        _check(lib().gfft_exec_c2r(self.handle, _ptr(data), _stream_handle(stream)), "gfft_exec_c2r")

    # local halves around a slab exchange (see gpufft.h)
    def r2c_local(self, data, sendbuf, stream=None) -> None:
        _check(lib().gfft_exec_r2c_local(self.handle, _ptr(data), _ptr(sendbuf), _stream_handle(stream)),
               "gfft_exec_r2c_local")

    def r2c_finish(self, recvbuf, out, stream=None) -> None:
        _check(lib().gfft_exec_r2c_finish(self.handle, _ptr(recvbuf), _ptr(out), _stream_handle(stream)),
               "gfft_exec_r2c_finish")

    def c2r_start(self, data, sendbuf, stream=None) -> None:
        _check(lib().gfft_exec_c2r_start(self.handle, _ptr(data), _ptr(sendbuf), _stream_handle(stream)),
               "gfft_exec_c2r_start")

    def c2r_local(self, recvbuf, data, stream=None) -> None:
        _check(lib().gfft_exec_c2r_local(self.handle, _ptr(recvbuf), _ptr(data), _stream_handle(stream)),
               "gfft_exec_c2r_local")

    def tune(self, scratch, stream=None, reps: int = 5) -> dict:
        """Autotune every axis on ``scratch`` (contents destroyed); returns
        the plan description with the chosen variants."""
        _check(lib().gfft_plan_tune(self.handle, _ptr(scratch), _stream_handle(stream), int(reps)), "gfft_plan_tune")
        return self.describe()

    def set_variant(self, axis: int, variant: int) -> None:
        _check(lib().gfft_plan_set_variant(self.handle, int(axis), int(variant)), "gfft_plan_set_variant")

    def set_fused(self, on: bool) -> None:
        """Use (True) or bypass (False) the fused z + y plane kernels."""
        _check(lib().gfft_plan_set_fused(self.handle, 1 if on else 0), "gfft_plan_set_fused")

    def set_chunking(self, planes: int) -> None:
        """Alternate z and y passes per block of ``planes`` x-planes (0 = off)."""
        _check(lib().gfft_plan_set_chunking(self.handle, int(planes)), "gfft_plan_set_chunking")

    def set_walk(self, mode: int) -> None:
        """0: every pass walks its tiles in order; 1: passes alternate the
        order so each starts where the previous one ended (cache reuse)."""
        _check(lib().gfft_plan_set_walk(self.handle, int(mode)), "gfft_plan_set_walk")

    def set_walk_mask(self, mask: int) -> None:
        """Per-pass tile order: bit 3*direction + axis set = that pass walks
        its tiles last to first (direction 0 = R2C, 1 = C2R; axis 0 = z,
        1 = y, 2 = x)."""
        _check(lib().gfft_plan_set_walk_mask(self.handle, int(mask)), "gfft_plan_set_walk_mask")

    def set_pass_order(self, order: int) -> None:
        """Strided pass order: bit 0 = R2C runs x before y, bit 1 = C2R runs
        y before x (0 = the reference's order)."""
        _check(lib().gfft_plan_set_pass_order(self.handle, int(order)), "gfft_plan_set_pass_order")

    def set_scale(self, r2c_scale: float = 1.0, c2r_scale: float = 1.0) -> None:
        """Output scales fused into the last pass of each direction."""
        _check(lib().gfft_plan_set_scale(self.handle, float(r2c_scale), float(c2r_scale)), "gfft_plan_set_scale")

    def axis(self, axis: int, inverse: bool, data, stream=None) -> None:
        """One in-place axis pass (0 = z, 1 = y, 2 = x; 3 = the fused z + y
        plane pass when the plan uses it)."""
        _check(lib().gfft_exec_axis(self.handle, int(axis), int(bool(inverse)), _ptr(data), _stream_handle(stream)),
               "gfft_exec_axis")

    def describe(self) -> dict:
        buf = ctypes.create_string_buffer(8192)
        _check(lib().gfft_plan_describe(self.handle, buf, len(buf)), "gfft_plan_describe")
        return json.loads(buf.value.decode())

    @property
    def device_bytes(self) -> int:
        v = ctypes.c_size_t()
        _check(lib().gfft_plan_device_bytes(self.handle, ctypes.byref(v)), "gfft_plan_device_bytes")
        return v.value

    @property
    def data_bytes(self) -> int:
        v = ctypes.c_size_t()
        _check(lib().gfft_plan_data_bytes(self.handle, ctypes.byref(v)), "gfft_plan_data_bytes")
        return v.value

    def close(self) -> None:
        if getattr(self, "_h", None) is not None:
            lib().gfft_plan_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def alloc_inplace(nx, ny, nz, precision="f32", procs=1, device="cuda"):
    """A zeroed torch tensor holding one in-place array: real view
    (Nx/p, Ny, 2K) of float32/float64."""
    import torch
    dt = torch.float32 if _prec(precision) == F32 else torch.float64
    return torch.zeros((nx // procs, ny, 2 * kdim(nz)), dtype=dt, device=device)


def real_view(buf, nz):
    """Unpadded real view (..., Nz) of an in-place buffer (no copy)."""
    return buf[..., :nz]


def complex_view(buf, nx_out, ny_out, nz):
    """Complex (Nx, Ny/p, K) view of an in-place buffer (no copy)."""
    import torch
    return torch.view_as_complex(buf.view(-1, 2)).view(nx_out, ny_out, kdim(nz))


def _bind_batch(L):
    if getattr(L, "_gfft_batch_bound", False):
        return L
    i64, vp, sz = ctypes.c_int64, ctypes.c_void_p, ctypes.c_size_t
    L.gfft_batch_query.restype = ctypes.c_int
    L.gfft_batch_query.argtypes = [i64, i64, i64, ctypes.c_int]
    L.gfft_batch_create.restype = ctypes.c_int
    L.gfft_batch_create.argtypes = [ctypes.POINTER(vp), ctypes.c_int, ctypes.POINTER(i64), ctypes.c_int]
    L.gfft_batch_set_data.restype = ctypes.c_int
    L.gfft_batch_set_data.argtypes = [vp, ctypes.POINTER(vp)]
    L.gfft_batch_set_scales.restype = ctypes.c_int
    L.gfft_batch_set_scales.argtypes = [vp, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]
    for name in ("gfft_batch_exec", "gfft_batch_exec_graph", "gfft_batch_exec_noalloc"):
        f = getattr(L, name)
        f.restype = ctypes.c_int
        f.argtypes = [vp, ctypes.c_int, vp]
    L.gfft_batch_data_bytes.restype = ctypes.c_int
    L.gfft_batch_data_bytes.argtypes = [vp, ctypes.c_int, ctypes.POINTER(sz)]
    L.gfft_batch_tiles.restype = ctypes.c_int
    L.gfft_batch_tiles.argtypes = [vp, ctypes.c_int, ctypes.POINTER(ctypes.c_uint32),
                                   ctypes.POINTER(ctypes.c_uint32)]
    L.gfft_batch_destroy.restype = ctypes.c_int
    L.gfft_batch_destroy.argtypes = [vp]
    L.gfft_batch_tune.restype = ctypes.c_int
    L.gfft_batch_tune.argtypes = [vp, ctypes.c_int, ctypes.c_int, vp]
    L.gfft_batch_set_variant.restype = ctypes.c_int
    L.gfft_batch_set_variant.argtypes = [vp, i64, ctypes.c_int]
    L.gfft_batch_variant.restype = ctypes.c_int
    L.gfft_batch_variant.argtypes = [vp, i64, ctypes.POINTER(ctypes.c_int)]
    ip = ctypes.POINTER(ctypes.c_int)
    for name, args in (("gfft_batch_set_scheduler", [vp, ctypes.c_int]), ("gfft_batch_scheduler", [vp, ip]),
                       ("gfft_batch_set_persist", [vp, ctypes.c_int, ctypes.c_int]),
                       ("gfft_batch_persist", [vp, ip, ip, ip]),
                       ("gfft_batch_set_span", [vp, ctypes.c_int]), ("gfft_batch_span", [vp, ip]),
                       ("gfft_batch_set_wait_limit", [vp, ctypes.c_double]),
                       ("gfft_batch_set_abort", [vp, ctypes.c_int]), ("gfft_batch_status", [vp, ip]),
                       ("gfft_batch_plan_info", [vp, ctypes.c_int] + [ctypes.POINTER(ctypes.c_uint32)] * 3),
                       ("gfft_batch_tuned_ms", [vp, ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float)])):
        f = getattr(L, name)
        f.restype = ctypes.c_int
        f.argtypes = args
    L._gfft_batch_bound = True
    return L


R2C, C2R, ROUND_TRIP = 0, 1, 2


def batch_supported(nx, ny, nz, precision="f32") -> bool:
    return _bind_batch(lib()).gfft_batch_query(nx, ny, nz, _prec(precision)) == 0


class Batch:
    """Many independent in-place transforms of uneven sizes executed as one
    unit, eagerly or as one hipGraph (gfft_batch_*).  ``dims`` is a list of
    (nx, ny, nz); all edges of a transform are one supported size."""

    def __init__(self, dims, precision="f32"):
        L = _bind_batch(lib())
        self.dims = [tuple(int(v) for v in d) for d in dims]
        flat = (ctypes.c_int64 * (3 * len(self.dims)))(*[v for d in self.dims for v in d])
        h = ctypes.c_void_p()
        _check(L.gfft_batch_create(ctypes.byref(h), len(self.dims), flat, _prec(precision)), "gfft_batch_create")
        self._h = h
        self._keep = None

    @property
    def handle(self):
        if self._h is None:
            raise ValueError("batch destroyed")
        return self._h

    def set_data(self, arrays) -> None:
        ptrs = (ctypes.c_void_p * len(arrays))(*[a if isinstance(a, int) else a.data_ptr() for a in arrays])
        _check(lib().gfft_batch_set_data(self.handle, ptrs), "gfft_batch_set_data")
        self._keep = list(arrays)  # the device memory must outlive later launches

    def set_scales(self, r2c=None, c2r=None) -> None:
        n = len(self.dims)
        a = (ctypes.c_double * n)(*r2c) if r2c is not None else None
        b = (ctypes.c_double * n)(*c2r) if c2r is not None else None
        _check(lib().gfft_batch_set_scales(self.handle, a, b), "gfft_batch_set_scales")

    def exec(self, mode=ROUND_TRIP, stream=None) -> None:
        _check(lib().gfft_batch_exec(self.handle, int(mode), _stream_handle(stream)), "gfft_batch_exec")

    def exec_graph(self, mode=ROUND_TRIP, stream=None) -> None:
        _check(lib().gfft_batch_exec_graph(self.handle, int(mode), _stream_handle(stream)), "gfft_batch_exec_graph")

    def exec_noalloc(self, mode=ROUND_TRIP, stream=None) -> None:
        """The persistent scheduler's run of ``mode`` with its counters in the
        library's static per-GPU array: allocates nothing.  Calls that overlap
        on one GPU race on that array."""
        _check(lib().gfft_batch_exec_noalloc(self.handle, int(mode), _stream_handle(stream)),
               "gfft_batch_exec_noalloc")

    def tune(self, mode=ROUND_TRIP, reps: int = 3, stream=None) -> dict:
        """Autotune both schedulers on the current data (destroyed) and keep
        the faster; returns describe() with both tuned times."""
        _check(lib().gfft_batch_tune(self.handle, int(mode), int(reps), _stream_handle(stream)), "gfft_batch_tune")
        p, l = ctypes.c_float(), ctypes.c_float()
        _check(lib().gfft_batch_tuned_ms(self.handle, ctypes.byref(p), ctypes.byref(l)), "gfft_batch_tuned_ms")
        d = self.describe()
        d["tuned_ms"] = {"persistent": round(p.value, 5), "levels": round(l.value, 5)}
        d["persist"] = self.persist()
        d["variants"] = self.variants()
        return d

    PLAIN = 3  # variant: the class's transforms through single-transform plans
    PERSISTENT, LEVELS = 0, 1  # schedulers

    def set_scheduler(self, scheduler: int) -> None:
        """PERSISTENT (default): one persistent launch per execution;
        LEVELS: one launch per (size, axis pass)."""
        _check(lib().gfft_batch_set_scheduler(self.handle, int(scheduler)), "gfft_batch_set_scheduler")

    def scheduler(self) -> int:
        v = ctypes.c_int()
        _check(lib().gfft_batch_scheduler(self.handle, ctypes.byref(v)), "gfft_batch_scheduler")
        return v.value

    def set_persist(self, set_index: int, team_factor: int) -> None:
        """Persistent scheduler: variant set (0..1) and team factor (a
        transform gets a team of workgroups in one XCD above that many slot
        shares; 0 = the global-level plan, every workgroup in every
        transform's axis passes)."""
        _check(lib().gfft_batch_set_persist(self.handle, int(set_index), int(team_factor)), "gfft_batch_set_persist")

    def set_span(self, parts: int) -> None:
        """Team plans: the most XCDs one team may span (1..8, default 2);
        a transform whose share of workgroups exceeds one XCD's slots gets a
        team over that many XCDs."""
        _check(lib().gfft_batch_set_span(self.handle, int(parts)), "gfft_batch_set_span")

    def span(self) -> int:
        v = ctypes.c_int()
        _check(lib().gfft_batch_span(self.handle, ctypes.byref(v)), "gfft_batch_span")
        return v.value

    def persist(self) -> dict:
        s, c = ctypes.c_int(), ctypes.c_int()
        v = (ctypes.c_int * 8)()
        _check(lib().gfft_batch_persist(self.handle, ctypes.byref(s), ctypes.byref(c), v), "gfft_batch_persist")
        sizes = [32, 48, 64, 96, 128, 160, 192, 256]
        present = {d[2] for d in self.dims}
        return {"set": s.value, "team_factor": c.value,
                "variants": {n: v[i] for i, n in enumerate(sizes) if n in present}}

    def plan_info(self, mode=ROUND_TRIP) -> dict:
        """Persistent plan of ``mode``: slots (workgroups), teams, members."""
        a, b, c = ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_uint32()
        _check(lib().gfft_batch_plan_info(self.handle, int(mode), ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)),
               "gfft_batch_plan_info")
        return {"slots": a.value, "teams": b.value, "members": c.value}

    def describe(self) -> dict:
        if self.scheduler() == self.PERSISTENT:
            return {"scheduler": "persistent", **self.persist()}
        return {"scheduler": "levels", "variants": self.variants()}

    def set_wait_limit(self, seconds: float) -> None:
        _check(lib().gfft_batch_set_wait_limit(self.handle, float(seconds)), "gfft_batch_set_wait_limit")

    def set_abort(self, abort: bool) -> None:
        _check(lib().gfft_batch_set_abort(self.handle, 1 if abort else 0), "gfft_batch_set_abort")

    def status(self) -> int:
        """Synchronise, then the sticky outcome of the persistent launches
        since the last call (0 = ok, 11 = a wait timed out, 12 = aborted),
        which is cleared."""
        v = ctypes.c_int()
        _check(lib().gfft_batch_status(self.handle, ctypes.byref(v)), "gfft_batch_status")
        return v.value

    def set_variant(self, edge: int, variant: int) -> None:
        """Variant of a size class: 0-2 = level kernels (radix order, points
        per thread), 3 = per-transform plans."""
        _check(lib().gfft_batch_set_variant(self.handle, int(edge), int(variant)), "gfft_batch_set_variant")

    def variants(self) -> dict:
        out = {}
        for n in sorted({d[2] for d in self.dims}):
            v = ctypes.c_int()
            _check(lib().gfft_batch_variant(self.handle, int(n), ctypes.byref(v)), "gfft_batch_variant")
            out[n] = v.value
        return out

    def tiles(self, mode=ROUND_TRIP):
        """(tiles, kernel launches) of one execution of ``mode``."""
        t, n = ctypes.c_uint32(), ctypes.c_uint32()
        _check(lib().gfft_batch_tiles(self.handle, int(mode), ctypes.byref(t), ctypes.byref(n)),
               "gfft_batch_tiles")
        return t.value, n.value

    def close(self) -> None:
        if getattr(self, "_h", None) is not None:
            lib().gfft_batch_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

"""Serving runtime binding (gfft_server_* of include/gpufft/gpufft.h).

Requests of mixed sizes, precisions and directions are submitted one by one;
the native runtime admits them in arrival order under an HBM byte budget,
runs them on several internal streams and stamps arrival / admission /
completion on one monotonic clock (``clock()``).  Plans and twiddle tables
stay resident per geometry.

Two request kinds:
  * device: ``data`` is a device tensor (or pointer) transformed in place;
  * host-staged: ``data`` is a host tensor (pinned for overlap), copied into
    the server's arena, transformed, and copied to ``host_out``.
"""
from __future__ import annotations

import ctypes

from . import F32, F64, GfftError, _check, _prec, lib

__all__ = ["Server", "Request", "Result", "clock", "request_bytes", "R2C", "C2R", "SERVE_SIZES"]

R2C, C2R = 0, 1
STAGING = 1
SHARED_QUEUES = 2
REQ_DEVICE = 1
REQ_PRODUCER = 2

def _is_235(n: int) -> bool:
    for p in (2, 3, 5):
        while n % p == 0:
            n //= p
    return n == 1


# every edge 2^a 3^b 5^c in [64, 512]: the 42 serving sizes
SERVE_SIZES = [n for n in range(64, 513) if _is_235(n)]


class Request(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int64), ("ny", ctypes.c_int64), ("nz", ctypes.c_int64),
                ("prec", ctypes.c_int), ("direction", ctypes.c_int), ("flags", ctypes.c_int),
                ("reserved", ctypes.c_int), ("data", ctypes.c_void_p), ("host_out", ctypes.c_void_p),
                ("scale", ctypes.c_double), ("arrival", ctypes.c_double), ("tag", ctypes.c_uint64),
                ("producer", ctypes.c_void_p)]


class Result(ctypes.Structure):
    _fields_ = [("id", ctypes.c_uint64), ("tag", ctypes.c_uint64), ("status", ctypes.c_int),
                ("stream", ctypes.c_int), ("arrival", ctypes.c_double), ("admitted", ctypes.c_double),
                ("done", ctypes.c_double), ("gpu_ms", ctypes.c_double), ("staged_ms", ctypes.c_double),
                ("bytes", ctypes.c_size_t), ("started", ctypes.c_double)]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


class Stats(ctypes.Structure):
    _fields_ = [(k, ctypes.c_uint64) for k in ("submitted", "completed", "rejected", "waiting", "running", "plans")] + \
               [(k, ctypes.c_size_t) for k in ("budget", "in_flight_bytes", "peak_in_flight_bytes", "arena_bytes",
                                               "arena_used", "plan_bytes")]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


def _bind(L):
    if getattr(L, "_gfft_server_bound", False):
        return L
    vp, i64, sz = ctypes.c_void_p, ctypes.c_int64, ctypes.c_size_t
    L.gfft_server_clock.restype = ctypes.c_double
    L.gfft_server_clock.argtypes = []
    L.gfft_request_bytes.restype = ctypes.c_int
    L.gfft_request_bytes.argtypes = [i64, i64, i64, ctypes.c_int, ctypes.POINTER(sz)]
    L.gfft_server_create.restype = ctypes.c_int
    L.gfft_server_create.argtypes = [ctypes.POINTER(vp), sz, ctypes.c_int, ctypes.c_int]
    L.gfft_server_set_small_lane.restype = ctypes.c_int
    L.gfft_server_set_small_lane.argtypes = [vp, sz]
    L.gfft_server_set_poll_interval.restype = ctypes.c_int
    L.gfft_server_set_poll_interval.argtypes = [vp, ctypes.c_int]
    L.gfft_server_set_stream_depth.restype = ctypes.c_int
    L.gfft_server_set_stream_depth.argtypes = [vp, ctypes.c_int]
    L.gfft_server_prepare.restype = ctypes.c_int
    L.gfft_server_prepare.argtypes = [vp, i64, i64, i64, ctypes.c_int]
    L.gfft_server_submit.restype = ctypes.c_int
    L.gfft_server_submit.argtypes = [vp, ctypes.POINTER(Request), ctypes.POINTER(ctypes.c_uint64)]
    L.gfft_server_poll.restype = ctypes.c_int
    L.gfft_server_poll.argtypes = [vp, ctypes.POINTER(Result), ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
    L.gfft_server_wait.restype = ctypes.c_int
    L.gfft_server_wait.argtypes = [vp, ctypes.c_double]
    L.gfft_server_stats.restype = ctypes.c_int
    L.gfft_server_stats.argtypes = [vp, ctypes.POINTER(Stats)]
    L.gfft_server_destroy.restype = ctypes.c_int
    L.gfft_server_destroy.argtypes = [vp]
    L._gfft_server_bound = True
    return L


def clock() -> float:
    """The runtime's monotonic clock, seconds."""
    return _bind(lib()).gfft_server_clock()


def request_bytes(nx, ny, nz, precision="f32") -> int:
    """Bytes of a request's in-place array (raises GfftError on bad geometry)."""
    v = ctypes.c_size_t()
    _check(_bind(lib()).gfft_request_bytes(nx, ny, nz, _prec(precision), ctypes.byref(v)), "gfft_request_bytes")
    return v.value


def _ptr(x) -> int:
    return x if isinstance(x, int) else x.data_ptr()


class Server:
    """A serving runtime on the current HIP device.

    nstreams = 3 by default: two streams for large requests and the
    small-request lane.  On the mixed-size Poisson workload at 80 % load it
    gave a lower p99 latency than 4 streams (6.1-6.2 ms against 6.4-6.5)
    at 1 % less capacity; 2 streams lost 14 % of capacity."""

    def __init__(self, budget_bytes: int, nstreams: int = 3, staging: bool = False, own_queues: bool = True):
        L = _bind(lib())
        h = ctypes.c_void_p()
        flags = (STAGING if staging else 0) | (0 if own_queues else SHARED_QUEUES)
        _check(L.gfft_server_create(ctypes.byref(h), int(budget_bytes), int(nstreams), flags), "gfft_server_create")
        self._h = h
        self._keep = {}  # request id -> tensors that must outlive the request

    @property
    def handle(self):
        if self._h is None:
            raise ValueError("server destroyed")
        return self._h

    def set_small_lane(self, max_bytes: int) -> None:
        """Reserve the last stream for requests of at most max_bytes (0 = off)."""
        _check(lib().gfft_server_set_small_lane(self.handle, int(max_bytes)), "gfft_server_set_small_lane")

    def set_stream_depth(self, depth: int) -> None:
        """At most ``depth`` admitted, unfinished requests per stream (0 = no cap)."""
        _check(lib().gfft_server_set_stream_depth(self.handle, int(depth)), "gfft_server_set_stream_depth")

    def set_poll_interval(self, microseconds: int) -> None:
        """Sleep between completion queries of the stream waiters (0 = block
        in hipEventSynchronize)."""
        _check(lib().gfft_server_set_poll_interval(self.handle, int(microseconds)), "gfft_server_set_poll_interval")

    def prepare(self, nx, ny, nz, precision="f32") -> None:
        _check(lib().gfft_server_prepare(self.handle, nx, ny, nz, _prec(precision)), "gfft_server_prepare")

    def submit(self, dims, precision, direction, data, host_out=None, device=None, scale=0.0, arrival=-1.0,
               tag=0, producer="current") -> int:
        """Queue one request; returns its id.  ``device`` defaults to
        ``data.is_cuda`` for tensors.  A device request starts after the
        work queued so far on ``producer``: a torch stream, a raw stream
        handle, ``"current"`` (default: torch's current stream when ``data``
        is a CUDA tensor) or None (the caller guarantees ``data`` is ready)."""
        if device is None:
            device = bool(getattr(data, "is_cuda", False))
        flags = REQ_DEVICE if device else 0
        handle = None
        if device and producer is not None:
            if isinstance(producer, str):
                if producer != "current":
                    raise ValueError("producer must be 'current', a stream, a handle or None")
                if getattr(data, "is_cuda", False):
                    import torch
                    handle = torch.cuda.current_stream(data.device).cuda_stream
            else:
                handle = int(getattr(producer, "cuda_stream", producer))
        if handle is not None:
            flags |= REQ_PRODUCER
        r = Request(int(dims[0]), int(dims[1]), int(dims[2]), _prec(precision), int(direction),
                    flags, 0, _ptr(data),
                    0 if device else _ptr(host_out if host_out is not None else data),
                    float(scale), float(arrival), int(tag), handle or 0)
        rid = ctypes.c_uint64()
        _check(lib().gfft_server_submit(self.handle, ctypes.byref(r), ctypes.byref(rid)), "gfft_server_submit")
        if not isinstance(data, int):
            self._keep[rid.value] = (data, host_out)
        return rid.value

    def poll(self, max_results: int = 1024) -> list:
        buf = (Result * max_results)()
        n = ctypes.c_int()
        _check(lib().gfft_server_poll(self.handle, buf, max_results, ctypes.byref(n)), "gfft_server_poll")
        out = [buf[i].as_dict() for i in range(n.value)]
        for r in out:
            self._keep.pop(r["id"], None)
        return out

    def wait(self, timeout_s: float = 0.0) -> None:
        _check(lib().gfft_server_wait(self.handle, float(timeout_s)), "gfft_server_wait")

    def drain(self, timeout_s: float = 0.0) -> list:
        """wait() then every remaining result."""
        self.wait(timeout_s)
        out = []
        while True:
            part = self.poll()
            if not part:
                return out
            out += part

    def stats(self) -> dict:
        st = Stats()
        _check(lib().gfft_server_stats(self.handle, ctypes.byref(st)), "gfft_server_stats")
        return st.as_dict()

    def close(self) -> None:
        if getattr(self, "_h", None) is not None:
            lib().gfft_server_destroy(self._h)
            self._h = None
            self._keep.clear()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


__all__ += ["F32", "F64", "GfftError", "STAGING", "SHARED_QUEUES", "REQ_DEVICE", "REQ_PRODUCER"]

"""Benchmark of the in-place 3D R2C + C2R round trip on one MI355X.

Workload (default): 512^3 FP32, input uniform[-1, 1) generated on the device
from a fixed seed, p = 1.  One step = GPU_FFT_R2C followed by GPU_FFT_C2R on
the same in-place array (six axis passes).  The inverse carries the fused
1/(Nx Ny Nz) scale (gfft_plan_set_scale), the normalisation the reference
harness applies between the two calls (GPU_FFT_TEST_SCRIPT/test.cu:149), so
repeated steps keep the data at the input's magnitude; the scale is one
multiply inside the z pass's epilogue and adds no memory traffic.

Prints one JSON line (stdout).  Throughput is GFLOP/s with the 5 N log2 N
convention per transform (2 transforms per step).  The roofline object is for
the dominant kernel (the slowest of the six passes): algorithmic bytes per
launch = 2 S (the half-spectrum S = Nx*Ny*(Nz/2+1)*2*sizeof(real) is read
once and written once), divided by that pass's mean duration measured with
HIP events on the exec stream.

Two regions of K steps each, on the same stream and array:
- the timed region (ms_per_step, value): each step is one gfft_exec_r2c and
  one gfft_exec_c2r call, the entry points GPU_FFT_R2C / GPU_FFT_C2R map to;
- the per-pass region (per_pass, roofline): each step runs the same passes
  one gfft_exec_axis call at a time with a HIP event pair around each.  The
  events and the extra host calls add gaps between kernels (about 10 us per
  pass, visible on small grids), which is why they stay out of the timed
  region; per-pass durations are unaffected.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import platform
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0
METRIC = "512^3 FP32 R2C+C2R GFLOP/s & % HBM roofline; mixed-batch p50/p99 latency"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def cpu_baseline(shape, prec: str, reps: int, workers: int) -> dict:
    """The oracle's p = 1 transform (pocketfft via scipy.fft, multithreaded)
    in the GPU line's precision: R2C + C2R of the seeded field, median of
    ``reps`` runs.  FFTW3 is not installed on these machines."""
    import numpy as np
    from oracle import gfft_oracle as O
    nx, ny, nz = shape
    dt = np.float32 if prec == "f32" else np.float64
    f = O.pad_real(O.random_field(nx, ny, nz, seed=0)).astype(dt)
    times = []
    for _ in range(reps):
        t0 = time.perf_counter()
        s = O.r2c_p1(f, nz, workers=workers, dtype=dt)
        O.c2r_p1(s, nz, workers=workers)
        times.append(time.perf_counter() - t0)
        del s
    t = statistics.median(times)
    npts = nx * ny * nz
    flops = 2 * 5.0 * npts * math.log2(npts)
    return {
        "value": round(flops / t / 1e9, 3),
        "unit": "GFLOP/s",
        "cores": workers,
        "kind": "port",
        "sample": (f"pocketfft (scipy), {workers} threads -- not FFTW3: {prec.upper()} R2C+C2R of "
                   f"{nx}x{ny}x{nz} (the oracle's p=1 transform in the GPU line's precision), median of {reps} "
                   f"runs ({t * 1e3:.1f} ms each); CPU {cpu_model()}, {os.cpu_count()} logical CPUs"),
    }


def measured_traffic(pass_name: str, n: int, prec: str):
    """HBM bytes per launch of the pass's kernel family from the committed
    rocprofv3 FETCH_SIZE/WRITE_SIZE summary (calibrated against a plain copy
    of known size); None when no summary covers this workload."""
    path = os.path.join(ROOT, "profiles", f"hbm_traffic_{n}cube_{prec}_passes.json")
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        doc = json.load(fh)
    v = doc.get("passes", {}).get(pass_name)
    return int(v["read_bytes"] + v["write_bytes"]) if v else None


def measure_slab_share(nx, ny, nz, procs, prec, steps, warmup, tune_reps=3) -> dict:
    """One rank's share of a slab decomposition: the local R2C half (z R2C +
    y C2C with the slab-transpose pack fused into the y stores), on the
    (Nx/p, Ny, 2K) block.  Data moved: 4 S_local (z read + write, y read +
    write into the p destination chunks); the unfused model (z, y, pack
    each read + write) would move 6 S_local."""
    import torch
    import gpu_fft_amd as G
    K = nz // 2 + 1
    nxl = nx // procs
    rb = 4 if prec == "f32" else 8
    S = nxl * ny * K * 2 * rb
    dev = torch.device("cuda:0")
    dt = torch.float32 if prec == "f32" else torch.float64
    gen = torch.Generator(device=dev)
    gen.manual_seed(0)
    data = torch.rand((nxl, ny, 2 * K), dtype=dt, device=dev, generator=gen) * 2 - 1
    send = torch.empty_like(data)
    plan = G.Plan(nx, ny, nz, procs=procs, rank=0, precision=prec)
    if tune_reps > 0:
        plan.tune(send, reps=tune_reps)
    stream = torch.cuda.current_stream()
    for _ in range(warmup):
        plan.r2c_local(data, send, stream)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(steps):
        plan.r2c_local(data, send, stream)
    e1.record(stream)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / steps
    variants = {ax: plan.describe()[ax] for ax in ("z", "y")}
    plan.close()
    del data, send
    torch.cuda.empty_cache()
    return {"ms": ms, "S_local": S, "variants": variants,
            "data_moved_4S_GBps": round(4 * S / (ms * 1e-3) / 1e9, 1),
            "frac_4S": round(4 * S / (ms * 1e-3) / 1e9 / PEAK_HBM_GBS, 4),
            "frac_unfused_6S_model": round(6 * S / (ms * 1e-3) / 1e9 / PEAK_HBM_GBS, 4)}


def slab_share(args, nx, ny, nz, procs):
    m = measure_slab_share(nx, ny, nz, procs, args.prec, args.steps, args.warmup, 0 if args.no_tune else 3)
    ms = m["ms"]
    out = {"metric": "p-share local R2C half (z R2C + y C2C + pack) time", "value": round(ms, 5), "unit": "ms",
           "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms, 5),
           "higher_is_better": False, "dtype": args.prec,
           "config": {"workload": f"rank share of {nx}x{ny}x{nz} at p={procs}: {nx // procs}x{ny}x{nz} local block",
                      "S_local": m["S_local"]},
           "roofline": {"bound": "hbm", "data_moved_4S_GBps": m["data_moved_4S_GBps"], "frac_4S": m["frac_4S"],
                        "frac_unfused_6S_model": m["frac_unfused_6S_model"]},
           "variants": m["variants"]}
    print(json.dumps(out), flush=True)


# Config grids measured beside the headline in the default line (every one
# tuned outside its timed region; 10 timed round trips each).
CONFIG_GRIDS = [((128, 128, 128), "f64"), ((512, 512, 512), "f64"), ((1024, 1024, 1024), "f32"),
                ((1024, 512, 2048), "f32"), ((768, 768, 768), "f32")]


def config_results(steps: int = 10, warmup: int = 2) -> dict:
    """ms per round trip, the 12 S model fraction and the dominant pass of
    every config grid, plus the p = 8 share's local R2C half of 1024^3."""
    out = {}
    for (nx, ny, nz), prec in CONFIG_GRIDS:
        log(f"config {nx}x{ny}x{nz} {prec}")
        m = measure_round_trip(nx, ny, nz, prec, steps, warmup, tune_reps=3)
        S, ms = m["S"], m["ms_step"]
        dom = m["dom"]
        out[f"{nx}x{ny}x{nz}_{prec}"] = {
            "ms_per_round_trip": round(ms, 5),
            "frac_12S": round(12 * S / (ms * 1e-3) / 1e9 / PEAK_HBM_GBS, 4),
            "dominant_pass": dom,
            "dominant_GBps": m["per_pass"][dom]["GBps"],
            "per_pass_GBps": {k: v["GBps"] for k, v in m["per_pass"].items()},
        }
    for prec in ("f32", "f64"):
        log(f"config p=8 share of 1024^3 {prec}")
        m = measure_slab_share(1024, 1024, 1024, 8, prec, 20, 3, tune_reps=3)
        out[f"p8_share_1024cube_local_r2c_{prec}"] = {
            "ms": round(m["ms"], 5), "frac_4S": m["frac_4S"], "data_moved_4S_GBps": m["data_moved_4S_GBps"],
            "model": "4 S_local moved (z read+write, y read+write into the 8 chunks)"}
    return out


BATCH_SIZES = [32, 48, 64, 96, 128, 160, 192, 256]


def batch_dims(count: int, seed: int):
    """Cube edges drawn from BATCH_SIZES with weights proportional to 1/N^3."""
    import numpy as np
    rng = np.random.default_rng(seed)
    w = np.array([1.0 / s ** 3 for s in BATCH_SIZES])
    pick = rng.choice(BATCH_SIZES, size=count, p=w / w.sum())
    return [(int(n), int(n), int(n)) for n in pick]


def batch_bench(args):
    """256 independent FP32 R2C+C2R round trips of uneven cube sizes: the
    batched level kernels replayed as one hipGraph, against sequential
    plain-method launches (one prebuilt plan per size, R2C then C2R per
    transform, same stream, same fused 1/N scale)."""
    import torch
    import gpu_fft_amd as G
    dims = batch_dims(args.batch, seed=0)
    dev = torch.device("cuda:0")
    gen = torch.Generator(device=dev)
    gen.manual_seed(0)
    bufs = []
    for nx, ny, nz in dims:
        b = torch.rand((nx, ny, 2 * (nz // 2 + 1)), device=dev, generator=gen) * 2 - 1
        b[:, :, nz:] = 0
        bufs.append(b)
    scales = [1.0 / (nx * ny * nz) for nx, ny, nz in dims]
    s_bytes = sum(nx * ny * (nz // 2 + 1) * 8 for nx, ny, nz in dims)
    flops = sum(2 * 5.0 * nx * ny * nz * math.log2(nx * ny * nz) for nx, ny, nz in dims)
    stream = torch.cuda.current_stream()

    plans = {}
    for d in set(dims):
        p = G.Plan(*d)
        p.set_scale(1.0, 1.0 / (d[0] * d[1] * d[2]))
        plans[d] = p

    def plain():
        for d, b in zip(dims, bufs):
            plans[d].r2c(b, stream)
            plans[d].c2r(b, stream)

    batch = G.Batch(dims)
    batch.set_data(bufs)
    batch.set_scales(c2r=scales)
    # autotune the per-size kernel variants outside the timed region (the
    # normalised round trips keep the data at its magnitude)
    chosen = batch.tune(G.ROUND_TRIP, reps=3, stream=stream)
    tuned_span = batch.span()

    def timed(fn):
        for _ in range(args.warmup):
            fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(args.steps):
            fn()
        e1.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / args.steps

    def settle(fn):
        """Untimed steps after a scheduler or plan switch, before both of its
        launch forms are timed (graph first, then eager), so that neither
        is timed while the GPU is still ramping up from the host-side switch."""
        for _ in range(args.warmup):
            fn()
        torch.cuda.synchronize()

    ref = [b.clone() for b in bufs]
    print(f"batch: {len(dims)} transforms, plain path", file=sys.stderr, flush=True)
    ms_plain = timed(plain)
    print("batch: graph path", file=sys.stderr, flush=True)
    ms_batch = timed(lambda: batch.exec_graph(G.ROUND_TRIP, stream))
    print("batch: eager path", file=sys.stderr, flush=True)
    ms_batch_eager = timed(lambda: batch.exec(G.ROUND_TRIP, stream))
    # every execution above was a normalised round trip: the fields came back
    torch.cuda.synchronize()
    status = batch.status()
    rt_err = max(float((b - r)[..., :d[2]].abs().max() / r[..., :d[2]].abs().max())
                 for b, r, d in zip(bufs, ref, dims))
    tiles, launches = batch.tiles(G.ROUND_TRIP)
    # each scheduler with its tuned choices, as one graph launch
    sched_ms = {}
    for name, sch in (("persistent", G.Batch.PERSISTENT), ("levels", G.Batch.LEVELS)):
        print(f"batch: {name} scheduler, graph path", file=sys.stderr, flush=True)
        batch.set_scheduler(sch)
        settle(lambda: batch.exec(G.ROUND_TRIP, stream))
        sched_ms[name] = {"graph_ms": round(timed(lambda: batch.exec_graph(G.ROUND_TRIP, stream)), 5),
                          "eager_ms": round(timed(lambda: batch.exec(G.ROUND_TRIP, stream)), 5),
                          "launches": batch.tiles(G.ROUND_TRIP)[1]}
    # the persistent kernel's global-level plan (team factor 0: every
    # workgroup in every transform's passes), whatever the tuner kept
    print("batch: persistent global-level plan, graph path", file=sys.stderr, flush=True)
    batch.set_scheduler(G.Batch.PERSISTENT)
    batch.set_persist(chosen["persist"]["set"], 0)
    settle(lambda: batch.exec(G.ROUND_TRIP, stream))
    sched_ms["persistent_global_levels"] = {"graph_ms": round(timed(lambda: batch.exec_graph(G.ROUND_TRIP, stream)), 5),
                                            "eager_ms": round(timed(lambda: batch.exec(G.ROUND_TRIP, stream)), 5),
                                            "launches": batch.tiles(G.ROUND_TRIP)[1]}
    batch.set_persist(chosen["persist"]["set"], chosen["persist"]["team_factor"])
    batch.set_scheduler(G.Batch.PERSISTENT if chosen["scheduler"] == "persistent" else G.Batch.LEVELS)
    status = max(status, batch.status())
    counts = {n: sum(1 for d in dims if d[0] == n) for n in BATCH_SIZES}
    out = {"metric": "batched 256 mixed-size FP32 R2C+C2R throughput vs sequential plain-method launches",
           "value": round(args.batch / (ms_batch * 1e-3), 1), "unit": "transforms/s (round trips)",
           "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_batch, 5),
           "higher_is_better": True, "dtype": "f32", "data": "synthetic uniform[-1,1) on device",
           "config": {"workload": f"{args.batch} cubes from {BATCH_SIZES}, weights 1/N^3, seed 0",
                      "counts": counts, "sum_S_bytes": s_bytes, "tiles": tiles, "launches": launches},
           "plain_ms": round(ms_plain, 5), "batch_graph_ms": round(ms_batch, 5),
           "batch_eager_ms": round(ms_batch_eager, 5),
           "speedup_vs_plain": round(ms_plain / ms_batch, 3),
           "gflops_batch": round(flops / (ms_batch * 1e-3) / 1e9, 1),
           "alg_GBps_12S": round(12 * s_bytes / (ms_batch * 1e-3) / 1e9, 1),
           "roundtrip_max_rel_err": rt_err, "persistent_status": status,
           "plan": batch.plan_info(G.ROUND_TRIP),
           "tuned": {"scheduler": chosen["scheduler"], "tuned_eager_ms": chosen["tuned_ms"],
                     "persistent": {"set": chosen["persist"]["set"], "team_factor": chosen["persist"]["team_factor"],
                                    "span": tuned_span,
                                    "variants": {str(k): v for k, v in chosen["persist"]["variants"].items()}},
                     "level_variants": {str(k): v for k, v in chosen["variants"].items()}},
           "schedulers": sched_ms}
    batch.close()
    for p in plans.values():
        p.close()
    return out


def serve_dims(count: int, seed: int):
    """Config-5 request mix: edge log-uniform in [64, 512] rounded to the
    nearest 2^a 3^b 5^c size, 80% FP32 / 20% FP64, R2C or C2R with equal
    probability."""
    import numpy as np
    from gpu_fft_amd.serving import SERVE_SIZES
    rng = np.random.default_rng(seed)
    sizes = np.array(SERVE_SIZES)
    edges = np.exp(rng.uniform(math.log(64), math.log(512), size=count))
    pick = sizes[np.abs(sizes[None, :] - edges[:, None]).argmin(axis=1)]
    prec = np.where(rng.random(count) < 0.8, "f32", "f64")
    dirs = rng.integers(0, 2, size=count)
    return [(int(n), str(p), int(d)) for n, p, d in zip(pick, prec, dirs)]


def pct(xs, q):
    xs = sorted(xs)
    if not xs:
        return None
    k = min(len(xs) - 1, max(0, int(math.ceil(q / 100.0 * len(xs))) - 1))
    return xs[k]


def serve_bench(n_requests: int, staged: bool, load: float = 0.8, budget_gb: float = 64.0, nstreams: int = 3,
                seed: int = 0, small_lane: int = -1, own_queues: bool = True, stream_depth: int = -1,
                poll_us: int = -1) -> dict:
    """Poisson-arrival serving run of the config-5 mix.

    1. Plans of all 84 (size, precision) geometries are built up front
       (plan / twiddle residency; untimed).
    2. Capacity: a closed burst of requests submitted at once; capacity =
       requests / (last completion - first arrival).
    3. Open loop: arrivals at rate load * capacity (exponential gaps); each
       request is submitted at its scheduled arrival time, latency = its
       completion - that arrival.
    Device requests are transformed in place in one resident array per
    geometry (requests of one geometry share it; each request carries the
    unitary scale 1/sqrt(N) so repeated transforms keep the values bounded).
    Host-staged requests copy from / to one pinned host array per precision
    through the server's budget-sized arena.  small_lane >= 0 sets the
    reserved small-request stream's threshold in bytes (0 = no reservation);
    -1 keeps the server's default."""
    import torch
    from gpu_fft_amd import serving as SV
    dev = torch.device("cuda:0")
    reqs_cap = serve_dims(max(64, n_requests // 2), seed + 1)
    reqs = serve_dims(n_requests, seed)
    keys = sorted({(n, p) for n, p, _ in reqs + reqs_cap})
    srv = SV.Server(int(budget_gb * (1 << 30)), nstreams=nstreams, staging=staged, own_queues=own_queues)
    if small_lane >= 0:
        srv.set_small_lane(small_lane)
    if stream_depth >= 0:
        srv.set_stream_depth(stream_depth)
    if poll_us >= 0:
        srv.set_poll_interval(poll_us)
    for n, p in keys:
        srv.prepare(n, n, n, p)
    bufs, hin, hout = {}, {}, {}
    if staged:
        for p in ("f32", "f64"):
            mx = max((SV.request_bytes(n, n, n, p) for n, q in keys if q == p), default=0)
            if mx:
                dt = torch.float32 if p == "f32" else torch.float64
                esz = 4 if p == "f32" else 8
                hin[p] = (torch.rand(mx // esz, dtype=dt) * 2 - 1).pin_memory()
                hout[p] = torch.empty(mx // esz, dtype=dt).pin_memory()
    else:
        gen = torch.Generator(device=dev)
        gen.manual_seed(seed)
        for n, p in keys:
            dt = torch.float32 if p == "f32" else torch.float64
            bufs[(n, p)] = torch.rand((n, n, 2 * (n // 2 + 1)), dtype=dt, device=dev, generator=gen) * 2 - 1
        torch.cuda.synchronize()

    def submit(n, p, d, arrival, tag=0):
        scale = 1.0 / math.sqrt(n ** 3)
        if staged:
            return srv.submit((n, n, n), p, d, hin[p], host_out=hout[p], device=False, scale=scale,
                              arrival=arrival, tag=tag)
        return srv.submit((n, n, n), p, d, bufs[(n, p)], scale=scale, arrival=arrival, tag=tag)

    # warm every plan once (first launches load code objects)
    for n, p in keys:
        submit(n, p, 0, -1.0)
    srv.drain(timeout_s=300)

    # capacity: closed burst
    t0 = SV.clock()
    for n, p, d in reqs_cap:
        submit(n, p, d, t0)
    res = srv.drain(timeout_s=300)
    span = max(r["done"] for r in res) - t0
    capacity = len(reqs_cap) / span
    rate = load * capacity

    # open loop, Poisson arrivals
    import numpy as np
    gaps = np.random.default_rng(seed + 2).exponential(1.0 / rate, size=len(reqs))
    sched = SV.clock() + 0.01 + np.cumsum(gaps)
    late = 0.0
    for i, ((n, p, d), ta) in enumerate(zip(reqs, sched)):
        while SV.clock() < ta:
            pass
        late = max(late, SV.clock() - ta)
        submit(n, p, d, float(ta), tag=i)
    res = srv.drain(timeout_s=600)
    st = srv.stats()
    srv.close()
    ok = [r for r in res if r["status"] == 0]
    lat = [(r["done"] - r["arrival"]) * 1e3 for r in ok]
    queue = [(r["admitted"] - r["arrival"]) * 1e3 for r in ok]
    squeue = [(r["started"] - r["admitted"]) * 1e3 for r in ok]
    service = [(r["done"] - r["started"]) * 1e3 for r in ok]
    span = max(r["done"] for r in ok) - float(sched[0])
    bytes_total = sum(r["bytes"] for r in ok)
    flops = sum(5.0 * n ** 3 * math.log2(n ** 3) for n, _, _ in reqs)
    by_prec = {p: [(r["done"] - r["arrival"]) * 1e3 for r in ok if reqs[r["tag"]][1] == p] for p in ("f32", "f64")}
    buckets = (("<=128", 0, 128), ("129-256", 129, 256), (">256", 257, 1 << 30))
    by_edge = {name: [(r["done"] - r["arrival"]) * 1e3 for r in ok if lo <= reqs[r["tag"]][0] <= hi]
               for name, lo, hi in buckets}
    return {
        "mode": "host-staged (H2D + transform + D2H)" if staged else "device-resident, in place",
        "requests": len(reqs), "failed": len(res) - len(ok),
        "capacity_req_per_s": round(capacity, 1), "offered_req_per_s": round(rate, 1), "load": load,
        "throughput_req_per_s": round(len(ok) / span, 1),
        "throughput_GFLOPs": round(flops / span / 1e9, 1),
        "throughput_GBps_6S": round(6 * bytes_total / span / 1e9, 1) if not staged else None,
        "latency_ms": {"p50": round(pct(lat, 50), 4), "p99": round(pct(lat, 99), 4), "mean": round(sum(lat) / len(lat), 4),
                       "max": round(max(lat), 4)},
        "latency_ms_f32": {"p50": round(pct(by_prec["f32"], 50), 4), "p99": round(pct(by_prec["f32"], 99), 4)}
        if by_prec["f32"] else None,
        "latency_ms_f64": {"p50": round(pct(by_prec["f64"], 50), 4), "p99": round(pct(by_prec["f64"], 99), 4)}
        if by_prec["f64"] else None,
        "latency_ms_by_edge": {k: {"n": len(v), "p50": round(pct(v, 50), 4), "p99": round(pct(v, 99), 4)}
                               for k, v in by_edge.items() if v},
        "queue_ms": {"p50": round(pct(queue, 50), 4), "p99": round(pct(queue, 99), 4)},
        "stream_queue_ms": {"p50": round(pct(squeue, 50), 4), "p99": round(pct(squeue, 99), 4)},
        "service_ms": {"p50": round(pct(service, 50), 4), "p99": round(pct(service, 99), 4)},
        "latency_parts": "queue = admitted - arrival (server FIFO), stream_queue = device start - admitted "
                         "(behind earlier requests on its stream), service = done - device start",
        "gpu_ms_mean": round(sum(r["gpu_ms"] for r in ok) / len(ok), 4),
        "staged_ms_mean": round(sum(r["staged_ms"] for r in ok) / len(ok), 4),
        "budget_bytes": st["budget"], "peak_in_flight_bytes": st["peak_in_flight_bytes"],
        "plans_resident": st["plans"], "plan_bytes": st["plan_bytes"], "streams": nstreams,
        "small_lane_bytes": small_lane if small_lane >= 0 else (32 << 20 if nstreams >= 2 else 0),
        "stream_depth": stream_depth if stream_depth >= 0 else 1,
        "waiter_poll_us": poll_us if poll_us >= 0 else 20,
        "max_submit_lateness_ms": round(late * 1e3, 4),
        "workload": "edges log-uniform in [64,512] -> nearest of 42 sizes 2^a3^b5^c, 80/20 FP32/FP64, "
                    "R2C/C2R 50/50, Poisson arrivals, seed 0",
    }


def serve_main(args) -> dict:
    dev_res = serve_bench(args.serve, staged=False, load=args.serve_load, small_lane=args.serve_small_lane,
                          stream_depth=args.serve_depth, poll_us=args.serve_poll_us, nstreams=args.serve_streams)
    if args.serve_device_only:
        return {"device": dev_res}
    host_res = serve_bench(max(100, args.serve // 5), staged=True, load=args.serve_load,
                           small_lane=args.serve_small_lane, stream_depth=args.serve_depth,
                           poll_us=args.serve_poll_us, nstreams=args.serve_streams)
    return {"metric": "served mixed-size R2C/C2R requests: p50/p99 arrival-to-completion latency at "
                      f"{args.serve_load:.0%} of capacity",
            "value": dev_res["latency_ms"]["p99"], "unit": "ms (p99 latency, device-resident requests)",
            "n_gpus": 1, "steps": args.serve, "warmup": 0, "higher_is_better": False,
            "dtype": "f32/f64 (80/20)", "data": "synthetic, resident per geometry",
            "config": {"workload": dev_res["workload"], "budget_GB": 64},
            "device": dev_res, "host_staged": host_res}


def copy_calibration(nx, ny, nz, prec, steps):
    """A plain device copy of S bytes, for HBM counter calibration under rocprofv3."""
    import torch
    K = nz // 2 + 1
    dt = torch.float32 if prec == "f32" else torch.float64
    buf = torch.rand((nx, ny, 2 * K), dtype=dt, device="cuda:0")
    dst = torch.empty_like(buf)
    for _ in range(steps):
        dst.copy_(buf)
    torch.cuda.synchronize()
    S = buf.numel() * buf.element_size()
    print(json.dumps({"copy_calibration_bytes_read": S, "copy_calibration_bytes_written": S,
                      "launches": steps}), flush=True)


def measure_round_trip(nx, ny, nz, prec, steps, warmup, tune_reps=10):
    """R2C + C2R round trips of the seeded uniform field on one plan (tuned
    outside the timed region when tune_reps > 0).  Two regions of `steps`
    steps: the timed one (gfft_exec_r2c + gfft_exec_c2r per step, HIP events
    around the region) and the per-pass one (one event pair around each
    pass)."""
    import torch
    import gpu_fft_amd as G
    K = nz // 2 + 1
    real_bytes = 4 if prec == "f32" else 8
    S = nx * ny * K * 2 * real_bytes
    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    dt = torch.float32 if prec == "f32" else torch.float64
    gen = torch.Generator(device=dev)
    gen.manual_seed(0)
    buf = torch.rand((nx, ny, 2 * K), dtype=dt, device=dev, generator=gen) * 2 - 1
    buf[:, :, nz:] = 0
    stream = torch.cuda.current_stream()

    plan = G.Plan(nx, ny, nz, precision=prec)
    plan.set_scale(1.0, 1.0 / (nx * ny * nz))
    if tune_reps > 0:
        # plan-time autotuning (outside the timed region): it overwrites the
        # buffer, which is regenerated from the same seed afterwards
        plan.tune(buf, reps=tune_reps)
        torch.cuda.synchronize()
        gen.manual_seed(0)
        buf.copy_(torch.rand((nx, ny, 2 * K), dtype=dt, device=dev, generator=gen) * 2 - 1)
        buf[:, :, nz:] = 0
    desc = plan.describe()
    log("plan", json.dumps(desc))

    if desc.get("zy_fused"):
        # small planes: z and y run as one fused LDS-resident plane pass
        # (axis 3); its HBM bytes per launch are 2 S like every other pass
        passes = [(3, False), (2, False), (2, True), (3, True)]
        names = ["zy_r2c", "x_fwd", "x_inv", "zy_c2r"]
    else:
        passes = [(0, False), (1, False), (2, False), (2, True), (1, True), (0, True)]
        names = ["z_r2c", "y_fwd", "x_fwd", "x_inv", "y_inv", "z_c2r"]

    def step():
        plan.r2c(buf, stream)
        plan.c2r(buf, stream)

    def step_passes(ev):
        for j, (ax, inv) in enumerate(passes):
            ev[j][0].record(stream)
            plan.axis(ax, inv, buf, stream)
            ev[j][1].record(stream)

    for _ in range(warmup):
        step()
    torch.cuda.synchronize()

    # timed region: exactly K steps through the R2C / C2R entry points
    t_start = torch.cuda.Event(enable_timing=True)
    t_end = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    w0 = time.perf_counter()
    t_start.record(stream)
    for k in range(steps):
        step()
    t_end.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - w0
    ms_total = t_start.elapsed_time(t_end)
    ms_step = ms_total / steps

    # per-pass region: K more steps, one event pair around each pass
    evs = [[[torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)] for _ in passes]
           for _ in range(steps)]
    p_start = torch.cuda.Event(enable_timing=True)
    p_end = torch.cuda.Event(enable_timing=True)
    p_start.record(stream)
    for k in range(steps):
        step_passes(evs[k])
    p_end.record(stream)
    torch.cuda.synchronize()
    ms_step_passes = p_start.elapsed_time(p_end) / steps

    per_pass = {}
    for j, nm in enumerate(names):
        ts = [evs[k][j][0].elapsed_time(evs[k][j][1]) for k in range(steps)]
        mean = sum(ts) / len(ts)
        per_pass[nm] = {"ms_mean": round(mean, 5), "ms_median": round(statistics.median(ts), 5),
                        "ms_min": round(min(ts), 5), "GBps": round(2 * S / (mean * 1e-3) / 1e9, 1)}
    # sanity: the array must still hold the (normalised) input's magnitude
    rms = float(buf[:, :, :nz].float().pow(2).mean().sqrt())
    finite = bool(torch.isfinite(buf[:, :, :nz]).all())
    log(f"ms/step {ms_step:.4f}  wall/step {wall / steps * 1e3:.4f}  rms {rms:.4f} finite {finite}")

    plan.close()
    del buf
    torch.cuda.empty_cache()
    dom = max(per_pass, key=lambda k: per_pass[k]["ms_mean"])
    return {"S": S, "ms_step": ms_step, "ms_step_passes": ms_step_passes, "per_pass": per_pass, "names": names,
            "dom": dom, "desc": desc, "wall_s": wall, "rms": rms, "finite": finite}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=512, help="cube edge")
    ap.add_argument("--shape", type=int, nargs=3, metavar=("NX", "NY", "NZ"), help="non-cubic grid")
    ap.add_argument("--slab-share", type=int, default=0, metavar="P",
                    help="time the local R2C half (z R2C + y C2C + pack) of rank 0's share at p = P")
    ap.add_argument("--batch", type=int, default=0, metavar="COUNT",
                    help="batched mode: COUNT mixed-size cube round trips as one graph launch")
    ap.add_argument("--serve", type=int, default=0, metavar="REQUESTS",
                    help="serving mode: REQUESTS Poisson arrivals of the mixed-size request stream")
    ap.add_argument("--serve-load", type=float, default=0.8, help="offered load as a fraction of capacity")
    ap.add_argument("--serve-small-lane", type=int, default=-1, metavar="BYTES",
                    help="reserved small-request stream threshold (0 = off, -1 = server default)")
    ap.add_argument("--serve-depth", type=int, default=-1, metavar="D",
                    help="unfinished requests per server stream (0 = no cap, -1 = server default 1)")
    ap.add_argument("--serve-poll-us", type=int, default=-1, metavar="US",
                    help="server waiters' sleep between completion queries (0 = blocking, -1 = default 20)")
    ap.add_argument("--serve-device-only", action="store_true", help="serving: device requests only")
    ap.add_argument("--serve-streams", type=int, default=3, help="server streams")
    ap.add_argument("--no-extras", action="store_true",
                    help="default run without the batched and served sub-results")
    ap.add_argument("--prec", default="f32", choices=["f32", "f64"])
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-reps", type=int, default=40,
                    help="CPU baseline repetitions (40 x ~0.2 s on 16 threads: ~8 s wall, ~2 CPU-minutes)")
    ap.add_argument("--no-tune", action="store_true", help="use the default kernel variants")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain device copy of S bytes (HBM counter calibration under rocprofv3)")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: the transform path is single-GPU (p = 1); "
              "multi-GPU runs are not part of this build", file=sys.stderr)
        sys.exit(2)

    import torch
    import gpu_fft_amd as G

    n = args.n
    nx, ny, nz = args.shape if args.shape else (n, n, n)
    if args.slab_share:
        return slab_share(args, nx, ny, nz, args.slab_share)
    if args.batch:
        print(json.dumps(batch_bench(args)), flush=True)
        return
    if args.serve:
        print(json.dumps(serve_main(args)), flush=True)
        return
    if args.copy_calibration:
        copy_calibration(nx, ny, nz, args.prec, args.steps)
        return
    m = measure_round_trip(nx, ny, nz, args.prec, args.steps, args.warmup, tune_reps=0 if args.no_tune else 10)
    S, ms_step, per_pass, desc = m["S"], m["ms_step"], m["per_pass"], m["desc"]
    names, dom = m["names"], m["dom"]
    achieved = 2 * S / (per_pass[dom]["ms_mean"] * 1e-3) / 1e9
    traffic = measured_traffic(dom, n, args.prec) if (nx, ny, nz) == (n, n, n) else None
    ms_step_passes, wall, rms, finite = m["ms_step_passes"], m["wall_s"], m["rms"], m["finite"]
    flops_step = 2 * 5.0 * (nx * ny * nz) * math.log2(nx * ny * nz)
    gflops = flops_step / (ms_step * 1e-3) / 1e9
    total_alg = 12 * S
    out = {
        "metric": METRIC,
        "value": round(gflops, 2),
        "unit": "GFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if args.prec == "f32" else "f64",
        "data": "synthetic uniform[-1,1) generated on device (seed 0), pads zero",
        "config": {"workload": f"{nx}x{ny}x{nz} {args.prec.upper()} in-place R2C+C2R round trip, p=1",
                   "nx": nx, "ny": ny, "nz": nz, "procs": 1, "precision": args.prec,
                   "inverse_scale": "1/(Nx*Ny*Nz) fused"},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1), "peak": PEAK_HBM_GBS,
                     "unit": "GB/s", "frac": round(achieved / PEAK_HBM_GBS, 4), "traffic": traffic,
                     "alg_bytes_per_launch": 2 * S},
        "roofline_round_trip": {"alg_bytes": total_alg, "model": "12*S (3 passes x read+write per direction)",
                                "achieved_GBps": round(total_alg / (ms_step * 1e-3) / 1e9, 1),
                                "frac": round(total_alg / (ms_step * 1e-3) / 1e9 / PEAK_HBM_GBS, 4),
                                "compulsory_frac": round(4 * S / (ms_step * 1e-3) / 1e9 / PEAK_HBM_GBS, 4)},
        "per_pass": per_pass,
        "ms_per_step_per_pass_region": round(ms_step_passes, 5),
        # the passes of each direction (R2C: the first half of `passes`)
        "r2c_ms": round(sum(per_pass[nm]["ms_mean"] for nm in names[:len(names) // 2]), 5),
        "c2r_ms": round(sum(per_pass[nm]["ms_mean"] for nm in names[len(names) // 2:]), 5),
        "variants": {ax: {k: desc[ax][k] for k in ("engine", "radices", "points_per_thread", "tile", "threads",
                                                   "persist", "order", "readdr", "straight", "variant")} for ax in ("z", "y", "x")},
        "zy_fused": int(desc.get("zy_fused", 0)),
        "wall_ms_per_step": round(wall / args.steps * 1e3, 5),
        "output_rms": round(rms, 5),
        "output_finite": finite,
    }
    if not args.no_cpu:
        workers = max(1, min(16, os.cpu_count() or 1))
        log(f"cpu baseline: {workers} threads")
        out["cpu_baseline"] = cpu_baseline((nx, ny, nz), args.prec, args.cpu_reps, workers)
    else:
        out["cpu_baseline"] = None
    if not args.no_extras:
        out["configs"] = config_results()
        log("batched mode: 256 mixed cubes")
        # 10 untimed steps per measurement: the first replays after a scheduler
        # switch and a graph instantiation run slower than steady state
        b = batch_bench(argparse.Namespace(batch=256, steps=20, warmup=10))
        out["batch"] = {k: b[k] for k in ("plain_ms", "batch_graph_ms", "batch_eager_ms", "speedup_vs_plain",
                                          "gflops_batch", "roundtrip_max_rel_err", "persistent_status", "tuned",
                                          "schedulers", "plan")}
        out["batch"]["launches"] = b["config"]["launches"]
        out["batch"]["workload"] = b["config"]["workload"]
        log("serving: device-resident requests")
        # 10000 requests: the p99 then rests on the slowest 100 of them
        out["serving"] = serve_bench(10000, staged=False, load=0.8)
        log("serving: host-staged requests")
        out["serving_host_staged"] = serve_bench(300, staged=True, load=0.8)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

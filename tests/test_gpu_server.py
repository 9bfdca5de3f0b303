"""Serving runtime on the GPU: served results are bitwise the same as the
same request executed alone through the plan API, agree with the FP64
oracle (FP32 rel L2 <= 1e-6, FP64 <= 1e-13), host-staged equals in-place
device requests, admission keeps the in-flight bytes under the budget in
strict arrival order, and bad requests are rejected before anything runs."""
import sys

import numpy as np
import pytest

import gpu_fft_amd as G
from gpu_fft_amd import serving as S
from oracle import gfft_oracle as O

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def make_input(torch, nx, ny, nz, prec, direction, seed, device="cuda:0"):
    """R2C: padded random real field; C2R: the R2C spectrum of one (Hermitian)."""
    fp = O.pad_real(O.random_field(nx, ny, nz, seed))
    if direction == S.C2R:
        spec = O.r2c_p1(fp, nz)
        fp = np.ascontiguousarray(np.stack([spec.real, spec.imag], axis=-1).reshape(nx, ny, -1))
    dt = np.float32 if prec == "f32" else np.float64
    return torch.from_numpy(fp.astype(dt)).to(device)


def mix(count, seed, max_edge):
    rng = np.random.default_rng(seed)
    sizes = [n for n in S.SERVE_SIZES if n <= max_edge]
    out = []
    for i in range(count):
        n = int(rng.choice(sizes))
        out.append(((n, n, n), "f32" if rng.random() < 0.8 else "f64", int(rng.integers(2))))
    # the odd sizes and non-cubic grids are served too
    out += [((75, 75, 75), "f64", S.C2R), ((81, 64, 75), "f32", S.R2C), ((64, 80, 81), "f64", S.R2C)]
    return out


def alone(torch, dims, prec, direction, x):
    y = x.clone()
    with G.Plan(*dims, precision=prec) as p:
        (p.r2c if direction == S.R2C else p.c2r)(y)
    torch.cuda.synchronize()
    return y


def test_served_mix_equals_alone_bitwise_and_oracle():
    import torch
    reqs = mix(36, seed=1, max_edge=128)
    say(f"case served mix of {len(reqs)} device requests")
    inputs = [make_input(torch, *d, p, dr, seed=10 + i) for i, (d, p, dr) in enumerate(reqs)]
    work = [x.clone() for x in inputs]
    with S.Server(budget_bytes=8 << 30, nstreams=4) as srv:
        ids = [srv.submit(d, p, dr, w, tag=i) for i, ((d, p, dr), w) in enumerate(zip(reqs, work))]
        res = srv.drain(timeout_s=120)
        st = srv.stats()
    assert len(res) == len(reqs) and all(r["status"] == 0 for r in res)
    assert sorted(r["tag"] for r in res) == list(range(len(reqs)))
    assert st["completed"] == len(reqs) and st["in_flight_bytes"] == 0
    assert len(set(ids)) == len(ids)
    for i, ((d, p, dr), x, w) in enumerate(zip(reqs, inputs, work)):
        ref = alone(torch, d, p, dr, x)
        assert torch.equal(w, ref), (i, d, p, dr)
        nx, ny, nz = d
        tol = 1e-6 if p == "f32" else 1e-13
        if dr == S.R2C:
            want = O.r2c_p1(x.cpu().numpy().astype(np.float64), nz)
            got = G.complex_view(w, nx, ny, nz).cpu().numpy()
        else:
            xc = x.cpu().numpy().astype(np.float64).reshape(nx, ny, -1, 2)
            want = O.c2r_p1(xc[..., 0] + 1j * xc[..., 1], nz)
            got = w.cpu().numpy()[..., :nz]
        err = O.rel_l2(got, want)
        assert err <= tol, (i, d, p, dr, err)


def test_host_staged_equals_device():
    import torch
    reqs = mix(10, seed=2, max_edge=96)
    say(f"case host-staged vs device, {len(reqs)} requests")
    dev_in = [make_input(torch, *d, p, dr, seed=40 + i) for i, (d, p, dr) in enumerate(reqs)]
    host_in = [x.cpu().pin_memory() for x in dev_in]
    host_out = [torch.empty_like(h).pin_memory() for h in host_in]
    with S.Server(budget_bytes=1 << 30, nstreams=3, staging=True) as srv:
        for (d, p, dr), hi, ho in zip(reqs, host_in, host_out):
            srv.submit(d, p, dr, hi, host_out=ho, device=False)
        for (d, p, dr), x in zip(reqs, dev_in):
            srv.submit(d, p, dr, x)
        res = srv.drain(timeout_s=120)
    assert all(r["status"] == 0 for r in res) and len(res) == 2 * len(reqs)
    for x, ho in zip(dev_in, host_out):
        assert torch.equal(x.cpu(), ho)
    staged = [r for r in res if r["staged_ms"] >= r["gpu_ms"] > 0]
    assert len(staged) == len(res)


def test_admission_respects_budget_in_arrival_order():
    import torch
    dims, prec = (256, 256, 256), "f32"
    nbytes = S.request_bytes(*dims, prec)
    budget = 3 * nbytes + nbytes // 2  # room for three at a time
    base = make_input(torch, *dims, prec, S.R2C, seed=70)
    bufs = [base.clone() for _ in range(12)]
    say("case admission budget, 12 requests of 256^3 f32")
    with S.Server(budget_bytes=budget, nstreams=4) as srv:
        for i, b in enumerate(bufs):
            srv.submit(dims, prec, S.R2C, b, tag=i)
        res = srv.drain(timeout_s=120)
        st = srv.stats()
    assert all(r["status"] == 0 for r in res)
    assert st["peak_in_flight_bytes"] <= budget
    assert st["peak_in_flight_bytes"] >= 2 * nbytes  # several ran side by side
    by_tag = sorted(res, key=lambda r: r["tag"])
    adm = [r["admitted"] for r in by_tag]
    assert adm == sorted(adm)  # strict FIFO admission
    for r in res:
        assert r["arrival"] <= r["admitted"] <= r["done"]


def test_over_budget_and_bad_requests_rejected():
    import torch
    x = make_input(torch, 64, 64, 64, "f32", S.R2C, seed=3)
    with S.Server(budget_bytes=1 << 20, nstreams=1) as srv:
        with pytest.raises(G.GfftError) as e:
            srv.submit((64, 64, 64), "f32", S.R2C, x)  # 1.1 MB > 1 MiB budget
        assert e.value.status == 10
        with pytest.raises(G.GfftError) as e:
            srv.submit((64, 64, 0), "f32", S.R2C, x)
        assert e.value.status == 2
        with pytest.raises(G.GfftError) as e:  # prime factor 17: no plan
            srv.submit((17, 17, 34), "f32", S.R2C, x)
        assert e.value.status == 3
        with pytest.raises(G.GfftError) as e:  # not aligned to a complex element
            srv.submit((8, 8, 8), "f32", S.R2C, x.data_ptr() + 4, device=True)
        assert e.value.status == 1
        host = torch.zeros(8, 8, 10)
        with pytest.raises(G.GfftError) as e:
            srv.submit((8, 8, 8), "f32", S.R2C, host, device=True)  # not device memory
        assert e.value.status == 1
        with pytest.raises(G.GfftError) as e:  # host-staged needs a staging server
            srv.submit((8, 8, 8), "f32", S.R2C, host, device=False)
        assert e.value.status == 1
        st = srv.stats()
    assert st["rejected"] == 1 and st["submitted"] == 0


def test_request_scale_normalises_round_trip():
    import torch
    dims = (96, 80, 75)
    x = make_input(torch, *dims, "f64", S.R2C, seed=5)
    y = x.clone()
    with S.Server(budget_bytes=1 << 30, nstreams=2) as srv:
        srv.submit(dims, "f64", S.R2C, y)
        srv.wait(60)
        srv.submit(dims, "f64", S.C2R, y, scale=1.0 / (96 * 80 * 75))
        res = srv.drain(timeout_s=60)
    assert all(r["status"] == 0 for r in res)
    err = O.rel_l2(y.cpu().numpy()[..., :75], x.cpu().numpy()[..., :75])
    assert err <= 1e-13, err


def test_small_lane_keeps_large_requests_off_the_last_stream():
    """The reserved stream only ever runs requests at or under the threshold;
    results stay bitwise those of the plan API; threshold 0 lifts it."""
    import torch
    big, small = ((256, 256, 256), "f32"), ((64, 64, 64), "f64")
    limit = S.request_bytes(*small[0], small[1])
    assert S.request_bytes(*big[0], big[1]) > limit
    reqs = [big if i % 3 == 0 else small for i in range(18)]
    say(f"case small lane, {len(reqs)} requests, limit {limit} bytes")
    inputs = [make_input(torch, *d, p, S.R2C, seed=90 + i) for i, (d, p) in enumerate(reqs)]
    work = [x.clone() for x in inputs]
    with S.Server(budget_bytes=8 << 30, nstreams=3) as srv:
        srv.set_small_lane(limit)
        for i, ((d, p), w) in enumerate(zip(reqs, work)):
            srv.submit(d, p, S.R2C, w, tag=i)
        res = srv.drain(timeout_s=120)
        srv.set_small_lane(0)
    assert len(res) == len(reqs) and all(r["status"] == 0 for r in res)
    for r in res:
        if r["bytes"] > limit:
            assert r["stream"] in (0, 1), r
    assert any(r["stream"] == 2 for r in res)
    for (d, p), x, w in zip(reqs, inputs, work):
        assert torch.equal(w, alone(torch, d, p, S.R2C, x)), (d, p)


def test_shared_queue_streams_give_the_same_results():
    """GFFT_SERVER_SHARED_QUEUES (plain streams from the runtime's pool)
    against the default CU-masked streams with queues of their own."""
    import torch
    reqs = mix(12, seed=4, max_edge=128)
    say(f"case shared vs own queues, {len(reqs)} requests")
    inputs = [make_input(torch, *d, p, dr, seed=120 + i) for i, (d, p, dr) in enumerate(reqs)]
    out = {}
    for own in (True, False):
        work = [x.clone() for x in inputs]
        torch.cuda.synchronize()  # the copies are complete before any server stream reads them
        with S.Server(budget_bytes=2 << 30, nstreams=4, own_queues=own) as srv:
            for (d, p, dr), w in zip(reqs, work):
                srv.submit(d, p, dr, w)
            res = srv.drain(timeout_s=120)
        assert len(res) == len(reqs) and all(r["status"] == 0 for r in res)
        out[own] = work
    for a, b in zip(out[True], out[False]):
        assert torch.equal(a, b)


def test_producer_stream_orders_the_request_after_its_data():
    """A device request names the stream that writes its array: the server's
    stream waits for that work (queued before submit, not yet run) without
    any host synchronisation, on non-blocking server streams."""
    import torch
    dims, prec = (128, 128, 128), "f32"
    x = make_input(torch, *dims, prec, S.R2C, seed=150)
    ref = alone(torch, dims, prec, S.R2C, x)
    side = torch.cuda.Stream()
    work = torch.zeros_like(x)
    a = torch.randn(4096, 4096, device="cuda:0")
    torch.cuda.synchronize()
    say("case producer stream, 128^3 f32 written on a side stream")
    with S.Server(budget_bytes=1 << 30, nstreams=2, own_queues=False) as srv:
        with torch.cuda.stream(side):
            for _ in range(8):          # several ms of work ahead of the copy
                a = a @ a
                a = a / a.norm()
            work.copy_(x)
        srv.submit(dims, prec, S.R2C, work, producer=side)
        res = srv.drain(timeout_s=60)
    assert len(res) == 1 and res[0]["status"] == 0
    assert torch.equal(work, ref)


def test_stream_depth_caps_requests_per_stream():
    """With depth 1 each stream holds one unfinished request at a time, so
    the requests of one stream never overlap in [admitted, done] and later
    requests wait in the FIFO (admitted - arrival grows); results are those
    of the plan API.  The started stamp lies between admission and done."""
    import torch
    dims, prec = (192, 192, 192), "f32"
    base = make_input(torch, *dims, prec, S.R2C, seed=160)
    ref = alone(torch, dims, prec, S.R2C, base)
    bufs = [base.clone() for _ in range(10)]
    torch.cuda.synchronize()
    say("case stream depth 1, 10 requests of 192^3 f32 on 2 streams")
    with S.Server(budget_bytes=8 << 30, nstreams=2) as srv:
        srv.set_small_lane(0)
        srv.set_stream_depth(1)
        for i, b in enumerate(bufs):
            srv.submit(dims, prec, S.R2C, b, tag=i)
        res = srv.drain(timeout_s=120)
    assert len(res) == 10 and all(r["status"] == 0 for r in res)
    for st in (0, 1):
        rs = sorted((r for r in res if r["stream"] == st), key=lambda r: r["admitted"])
        for a, b in zip(rs, rs[1:]):
            assert b["admitted"] >= a["done"], (a, b)
    assert max(r["admitted"] - r["arrival"] for r in res) > 0
    for r in res:
        assert r["admitted"] <= r["started"] <= r["done"], r
    for b in bufs:
        assert torch.equal(b, ref)


def test_staged_request_that_no_arena_block_can_hold_is_rejected():
    """Host-staged 5^3 FP32 is 600 B (768 B in 256-B blocks); a 700-B budget
    gives a 512-B arena: rejected at submission instead of blocking the
    FIFO for good."""
    import torch
    host = torch.zeros(5, 5, 6).pin_memory()
    out = torch.zeros_like(host).pin_memory()
    assert S.request_bytes(5, 5, 5, "f32") == 600
    with S.Server(budget_bytes=700, nstreams=1, staging=True) as srv:
        with pytest.raises(G.GfftError) as e:
            srv.submit((5, 5, 5), "f32", S.R2C, host, host_out=out, device=False)
        assert e.value.status == 10
        srv.wait(10)
        assert srv.stats()["rejected"] == 1

"""Batched mode: many transforms of uneven sizes executed as one unit.
Each result is checked against the FP64 oracle (FP32 tolerance 1e-6), is
bitwise the same alone and in the mix and the same as the single-transform
plan's, and graph replays (with eager work between them) reproduce the eager
result."""
import sys

import numpy as np
import pytest

import gpu_fft_amd as G
from oracle import gfft_oracle as O

pytestmark = pytest.mark.gpu
SIZES = [32, 48, 64, 96, 128, 160, 192, 256]


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def mix(count, seed, max_edge=256):
    rng = np.random.default_rng(seed)
    sizes = [s for s in SIZES if s <= max_edge]
    w = np.array([1.0 / s ** 3 for s in sizes])
    pick = rng.choice(sizes, size=count, p=w / w.sum())
    dims = [(int(n), int(n), int(n)) for n in pick]
    # a few degenerate axes too (length-1 x or y)
    dims += [(1, 64, 64), (96, 1, 96), (1, 1, 160)]
    return dims


def inputs(torch, dims, seed):
    out, host = [], []
    for i, (nx, ny, nz) in enumerate(dims):
        fp = O.pad_real(O.random_field(nx, ny, nz, seed=seed + i)).astype(np.float32)
        host.append(fp)
        out.append(torch.from_numpy(fp).to("cuda:0"))
    return out, host


def test_batch_r2c_then_c2r_matches_oracle():
    import torch
    dims = mix(40, seed=1, max_edge=128) + [(160, 160, 160), (192, 192, 192), (256, 256, 256)]
    bufs, host = inputs(torch, dims, 100)
    say(f"case batch {len(dims)} transforms")
    with G.Batch(dims) as b:
        b.set_data(bufs)
        b.exec(G.R2C)
        torch.cuda.synchronize()
        for (nx, ny, nz), buf, fp in zip(dims, bufs, host):
            got = G.complex_view(buf, nx, ny, nz).cpu().numpy()
            err = O.rel_l2(got, O.r2c_p1(fp.astype(np.float64), nz))
            assert err <= 1e-6, ((nx, ny, nz), err)
        b.set_scales(c2r=[1.0 / (nx * ny * nz) for nx, ny, nz in dims])
        b.exec(G.C2R)
        torch.cuda.synchronize()
        for (nx, ny, nz), buf, fp in zip(dims, bufs, host):
            err = O.rel_l2(buf.cpu().numpy()[:, :, :nz], fp[:, :, :nz])
            assert err <= 1e-6, ((nx, ny, nz), err)


def test_alone_equals_in_mix_bitwise():
    import torch
    dims = mix(60, seed=2)
    bufs, _ = inputs(torch, dims, 7)
    ref = [b.clone() for b in bufs]
    with G.Batch(dims) as b:
        b.set_data(bufs)
        b.exec(G.ROUND_TRIP)
    torch.cuda.synchronize()
    for i in (0, 7, len(dims) - 1, len(dims) // 2):
        with G.Batch([dims[i]]) as one:
            one.set_data([ref[i]])
            one.exec(G.ROUND_TRIP)
        torch.cuda.synchronize()
        assert torch.equal(ref[i], bufs[i]), dims[i]


def test_graph_replays_match_eager():
    import torch
    dims = mix(30, seed=3, max_edge=128)
    bufs, host = inputs(torch, dims, 11)
    eager = [x.clone() for x in bufs]
    scales = [1.0 / (nx * ny * nz) for nx, ny, nz in dims]
    with G.Batch(dims) as be:
        be.set_data(eager)
        be.set_scales(c2r=scales)
        be.exec(G.ROUND_TRIP)
    torch.cuda.synchronize()
    with G.Batch(dims) as bg:
        bg.set_scales(c2r=scales)
        for rep in range(4):
            work = [torch.from_numpy(h).to("cuda:0") for h in host]
            bg.set_data(work)
            bg.exec_graph(G.ROUND_TRIP)
            torch.cuda.synchronize()
            _ = torch.ones(1000, device="cuda:0").sum().item()  # eager work between replays
            for w, e in zip(work, eager):
                assert torch.equal(w, e), rep


def test_batch_matches_single_plan_bitwise():
    """The batched kernels and the single-transform plan (which may use other
    radix schedules and tile shapes) agree to FP32 rounding."""
    import torch
    dims = [(64, 64, 64), (32, 32, 32), (256, 256, 256), (160, 160, 160)]
    bufs, host = inputs(torch, dims, 21)
    with G.Batch(dims) as b:
        b.set_data(bufs)
        b.exec(G.R2C)
    torch.cuda.synchronize()
    for (nx, ny, nz), buf, fp in zip(dims, bufs, host):
        other = torch.from_numpy(fp).to("cuda:0")
        with G.Plan(nx, ny, nz) as p:
            p.r2c(other)
        torch.cuda.synchronize()
        err = O.rel_l2(buf.cpu().numpy(), other.cpu().numpy())
        assert err <= 1e-6, ((nx, ny, nz), err)


def test_eager_on_side_stream_orders_with_caller():
    """exec on a non-default stream: work queued on that stream afterwards
    sees the finished batch (fork/join of the internal streams)."""
    import torch
    dims = mix(12, seed=5, max_edge=128)
    bufs, host = inputs(torch, dims, 31)
    scales = [1.0 / (nx * ny * nz) for nx, ny, nz in dims]
    side = torch.cuda.Stream()
    with G.Batch(dims) as b:
        b.set_data(bufs)
        b.set_scales(c2r=scales)
        torch.cuda.synchronize()
        with torch.cuda.stream(side):
            b.exec(G.ROUND_TRIP, side)
            sums = [x[..., :d[2]].double().sum() for x, d in zip(bufs, dims)]
        side.synchronize()
        for s, fp, (nx, ny, nz) in zip(sums, host, dims):
            ref = float(fp[..., :nz].astype(np.float64).sum())
            assert abs(float(s) - ref) <= 1e-4 * max(1.0, np.abs(fp).sum() / 1e3), ((nx, ny, nz), float(s), ref)


def test_tile_and_launch_counts():
    dims = [(64, 64, 64), (1, 64, 64), (32, 32, 32)]
    with G.Batch(dims) as b:
        tiles, launches = b.tiles(G.R2C)
        # class 32: P=8 -> 4 threads/row -> 64 pairs or columns per tile
        # class 64: P=8 -> 8 threads/row -> 32 pairs or columns per tile
        z = -(-(32 * 32 // 2) // 64) + -(-(64 * 64 // 2) // 32) + -(-(64 // 2) // 32)
        y = -(-(32 * 17) // 64) + -(-(64 * 33) // 32) + -(-(1 * 33) // 32)
        x = -(-(32 * 17) // 64) + -(-(64 * 33) // 32)
        assert tiles == z + y + x
        assert launches == 6
        assert b.tiles(G.ROUND_TRIP) == (2 * tiles, 12)


def test_batch_rejects_unsupported_sizes():
    with pytest.raises(G.GfftError):
        G.Batch([(32, 32, 32), (40, 40, 40)])
    with pytest.raises(G.GfftError):
        G.Batch([(32, 64, 32)])


def test_batch_rejects_misaligned_arrays():
    import torch
    buf = torch.zeros((32, 32, 34), device="cuda:0")
    with G.Batch([(32, 32, 32)]) as b:
        with pytest.raises(G.GfftError) as e:
            b.set_data([buf.data_ptr() + 4])
        assert e.value.status == 1


@pytest.mark.parametrize("variant", [0, 1, 2, 3])
def test_every_variant_matches_oracle(variant):
    """Each compiled variant (radix order, pencils per tile) of every size
    class and the per-transform-plan variant (3), forced for all classes,
    against the FP64 oracle."""
    import torch
    dims = [(n, n, n) for n in SIZES if n <= 192] + [(1, 256, 256), (256, 1, 256)]
    bufs, host = inputs(torch, dims, 300 + variant)
    say(f"case batch variant {variant}, {len(dims)} transforms")
    with G.Batch(dims) as b:
        for n in SIZES:
            b.set_variant(n, variant)
        assert set(b.variants().values()) == {variant}
        b.set_data(bufs)
        b.exec(G.R2C)
        torch.cuda.synchronize()
        for (nx, ny, nz), buf, fp in zip(dims, bufs, host):
            got = G.complex_view(buf, nx, ny, nz).cpu().numpy()
            err = O.rel_l2(got, O.r2c_p1(fp.astype(np.float64), nz))
            assert err <= 1e-6, ((nx, ny, nz), variant, err)


def test_tuned_batch_matches_oracle():
    import torch
    dims = mix(24, seed=4, max_edge=128) + [(160, 160, 160), (1, 256, 256)]
    scratch, _ = inputs(torch, dims, 500)
    with G.Batch(dims) as b:
        b.set_data(scratch)
        chosen = b.tune(G.ROUND_TRIP, reps=2)
        assert set(chosen) == {d[2] for d in dims} and all(0 <= v <= 3 for v in chosen.values())
        bufs, host = inputs(torch, dims, 600)
        b.set_data(bufs)
        b.exec_graph(G.R2C)
        torch.cuda.synchronize()
        for (nx, ny, nz), buf, fp in zip(dims, bufs, host):
            got = G.complex_view(buf, nx, ny, nz).cpu().numpy()
            err = O.rel_l2(got, O.r2c_p1(fp.astype(np.float64), nz))
            assert err <= 1e-6, ((nx, ny, nz), err)


def test_plain_variant_in_graph_follows_new_data_and_scales():
    """Classes on the per-transform-plan variant are captured into the graph
    with their data pointers and scales: set_data / set_scales rebuild it,
    so replays on new arrays and scales are right (vs the eager result)."""
    import torch
    dims = mix(10, seed=8, max_edge=128) + [(160, 160, 160)]
    with G.Batch(dims) as b:
        for n in SIZES:
            b.set_variant(n, G.Batch.PLAIN)
        for rep, sc in enumerate((1.0, 0.5)):
            bufs, host = inputs(torch, dims, 700 + rep)
            eager = [x.clone() for x in bufs]
            scales = [sc / (nx * ny * nz) for nx, ny, nz in dims]
            b.set_scales(c2r=scales)
            b.set_data(eager)
            b.exec(G.ROUND_TRIP)
            torch.cuda.synchronize()
            b.set_data(bufs)
            b.exec_graph(G.ROUND_TRIP)
            torch.cuda.synchronize()
            for w, e, fp, (nx, ny, nz) in zip(bufs, eager, host, dims):
                assert torch.equal(w, e)
                err = O.rel_l2(w.cpu().numpy()[..., :nz], sc * fp[..., :nz])
                assert err <= 1e-6, ((nx, ny, nz), rep, err)


def test_bench_mix_256_seed0_matches_oracle():
    """The benchmark's exact batched workload: 256 cubes drawn from SIZES
    with weights 1/N^3 (seed 0), inputs generated on the device as the
    benchmark does, autotuned as the benchmark does, executed as one graph
    launch: every R2C against the FP64 oracle (rel L2 <= 1e-6), then the
    normalised C2R against the input."""
    import torch
    from bench import batch_dims
    dims = batch_dims(256, seed=0)
    dev = torch.device("cuda:0")

    def gen_inputs():
        g = torch.Generator(device=dev)
        g.manual_seed(0)
        out = []
        for nx, ny, nz in dims:
            b = torch.rand((nx, ny, 2 * (nz // 2 + 1)), device=dev, generator=g) * 2 - 1
            b[:, :, nz:] = 0
            out.append(b)
        return out

    say(f"case bench mix, {len(dims)} transforms, sizes {sorted(set(d[0] for d in dims))}")
    bufs = gen_inputs()
    with G.Batch(dims) as b:
        b.set_data(bufs)
        b.set_scales(c2r=[1.0 / (nx * ny * nz) for nx, ny, nz in dims])
        b.tune(G.ROUND_TRIP, reps=3)
        bufs = gen_inputs()
        host = [x.cpu().numpy() for x in bufs]
        b.set_data(bufs)
        b.exec_graph(G.R2C)
        torch.cuda.synchronize()
        worst = 0.0
        for i, ((nx, ny, nz), buf, fp) in enumerate(zip(dims, bufs, host)):
            got = G.complex_view(buf, nx, ny, nz).cpu().numpy()
            err = O.rel_l2(got, O.r2c_p1(fp.astype(np.float64), nz))
            worst = max(worst, err)
            assert err <= 1e-6, (i, (nx, ny, nz), err)
        b.exec_graph(G.C2R)
        torch.cuda.synchronize()
        for i, ((nx, ny, nz), buf, fp) in enumerate(zip(dims, bufs, host)):
            err = O.rel_l2(buf.cpu().numpy()[:, :, :nz], fp[:, :, :nz])
            assert err <= 1e-6, (i, (nx, ny, nz), err)
    say(f"  worst R2C rel L2 {worst:.2e}")

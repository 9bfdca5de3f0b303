"""Batched mode: many transforms of uneven sizes executed as one unit, by
the persistent scheduler (default: one launch over a work queue with
per-pass completion counters) or the level scheduler (one launch per size
and axis pass).  Each result is checked against the FP64 oracle (FP32
tolerance 1e-6), is bitwise the same alone and in the mix and the same as
the single-transform plan's, and graph replays (with eager work between
them) reproduce the eager result.  The persistent kernel's bounded waits are
exercised from the host: an abort request and a zero wait bound make every
launch drain and report, and the next launch is correct again."""
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
        assert b.scheduler() == G.Batch.PERSISTENT
        assert b.tiles(G.ROUND_TRIP)[1] == 1  # one persistent launch
        b.set_scheduler(G.Batch.LEVELS)
        tiles, launches = b.tiles(G.R2C)
        # classes up to 128 run z and y fused: one tile per x-plane
        planes = 64 + 1 + 32
        # x columns, class 32: P=8 -> 4 threads per column -> 64 columns per
        # tile; class 64: P=8 -> 8 threads -> 32 columns (length-1 x: none)
        x = -(-(32 * 17) // 64) + -(-(64 * 33) // 32)
        assert tiles == planes + x
        assert launches == 4  # per class: the plane level and the x level
        assert b.tiles(G.ROUND_TRIP) == (2 * tiles, 8)


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
        b.set_scheduler(G.Batch.LEVELS)
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
        assert chosen["scheduler"] in ("persistent", "levels")
        assert chosen["tuned_ms"]["persistent"] > 0 and chosen["tuned_ms"]["levels"] > 0
        faster = "persistent" if chosen["tuned_ms"]["persistent"] <= chosen["tuned_ms"]["levels"] else "levels"
        assert chosen["scheduler"] == faster
        p = chosen["persist"]
        assert 0 <= p["set"] <= 1 and p["team_factor"] in (0, 1, 2, 4) and b.span() in (1, 2)
        assert set(chosen["variants"]) == {d[2] for d in dims}
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
        b.set_scheduler(G.Batch.LEVELS)
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


@pytest.mark.parametrize("set_index,chunk", [(0, 1), (0, 3), (1, 2), (1, 8), (0, 0), (1, 0)])
def test_persistent_sets_and_chunks_match_oracle(set_index, chunk):
    """Every variant set of the persistent kernel, with several team factors
    (teams of many members down to all-solo plans; 0 = the global-level plan,
    every workgroup in every transform), R2C against the FP64 oracle and the
    normalised round trip against the input (all sizes, length-1 axes)."""
    import torch
    dims = mix(30, seed=9) + [(256, 256, 256), (192, 192, 192), (1, 1, 48)]
    bufs, host = inputs(torch, dims, 800 + set_index)
    say(f"case persistent set {set_index} chunk {chunk}, {len(dims)} transforms")
    with G.Batch(dims) as b:
        b.set_persist(set_index, chunk)
        assert b.persist()["set"] == set_index and b.persist()["team_factor"] == chunk
        info = b.plan_info(G.R2C)
        say(f"  plan {info}")
        if chunk == 0:
            assert info["teams"] == 0 and info["members"] == info["slots"] >= 1
        else:
            assert info["slots"] >= 1 and (info["teams"] == 0) == (info["members"] == 0)
        b.set_data(bufs)
        b.exec(G.R2C)
        torch.cuda.synchronize()
        assert b.status() == 0
        for (nx, ny, nz), buf, fp in zip(dims, bufs, host):
            got = G.complex_view(buf, nx, ny, nz).cpu().numpy()
            err = O.rel_l2(got, O.r2c_p1(fp.astype(np.float64), nz))
            assert err <= 1e-6, ((nx, ny, nz), err)
        b.set_scales(c2r=[1.0 / (nx * ny * nz) for nx, ny, nz in dims])
        b.exec_graph(G.C2R)
        torch.cuda.synchronize()
        assert b.status() == 0
        for (nx, ny, nz), buf, fp in zip(dims, bufs, host):
            err = O.rel_l2(buf.cpu().numpy()[:, :, :nz], fp[:, :, :nz])
            assert err <= 1e-6, ((nx, ny, nz), err)


def test_persistent_matches_level_scheduler():
    """Set 1 runs variant 1 of every size; the level scheduler with variant 1
    forced for every size runs the same butterflies and radix order (the
    persistent kernel pipelines the tiles; multiply-add contraction may
    differ between the two instantiations): R2C and round trip agree to FP32
    rounding (rel L2 <= 1e-6 on the transform region), and the R2C matches
    the FP64 oracle."""
    import torch
    dims = mix(40, seed=10) + [(128, 1, 128)]
    bufs, host = inputs(torch, dims, 900)
    for mode in (G.R2C, G.ROUND_TRIP):
        a = [x.clone() for x in bufs]
        c = [x.clone() for x in bufs]
        with G.Batch(dims) as b:
            b.set_persist(1, 4)
            b.set_data(a)
            b.exec_graph(mode)
            torch.cuda.synchronize()
            assert b.status() == 0
            b.set_scheduler(G.Batch.LEVELS)
            for n in SIZES:
                b.set_variant(n, 1)
            b.set_data(c)
            b.exec(mode)
            torch.cuda.synchronize()
        for x, y, d, fp in zip(a, c, dims, host):
            nx, ny, nz = d
            if mode == G.R2C:
                gx = G.complex_view(x, nx, ny, nz).cpu().numpy()
                gy = G.complex_view(y, nx, ny, nz).cpu().numpy()
                assert O.rel_l2(gx, gy) <= 1e-6, (mode, d)
                assert O.rel_l2(gx, O.r2c_p1(fp.astype(np.float64), nz)) <= 1e-6, (mode, d)
            else:
                assert O.rel_l2(x.cpu().numpy()[..., :nz], y.cpu().numpy()[..., :nz]) <= 1e-6, (mode, d)


def test_persistent_abort_drains_and_reports():
    """With the host's abort word set, every team wait gives up at once:
    the launch still ends, reports GFFT_ERR_ABORTED once, and after the word
    is cleared the next launch is right."""
    import torch
    dims = mix(20, seed=11, max_edge=128)
    bufs, host = inputs(torch, dims, 950)
    with G.Batch(dims) as b:
        b.set_persist(0, 1)  # team factor 1: the larger transforms get teams, whose levels wait
        assert b.plan_info(G.ROUND_TRIP)["teams"] > 0
        b.set_data(bufs)
        b.set_abort(True)
        b.exec(G.ROUND_TRIP)
        torch.cuda.synchronize()
        assert b.status() == G.ERR_ABORTED
        assert b.status() == 0  # cleared by the first read
        b.set_abort(False)
        work = [torch.from_numpy(h).to("cuda:0") for h in host]
        b.set_data(work)
        b.exec(G.R2C)
        torch.cuda.synchronize()
        assert b.status() == 0
        for (nx, ny, nz), buf, fp in zip(dims, work, host):
            got = G.complex_view(buf, nx, ny, nz).cpu().numpy()
            assert O.rel_l2(got, O.r2c_p1(fp.astype(np.float64), nz)) <= 1e-6


def test_persistent_zero_wait_bound_drains():
    """A zero wait bound: a wait that is not satisfied at once gives up, the
    launch drains and reports a timeout (or OK if no wait had to wait); with
    the default bound back, results are right."""
    import torch
    dims = mix(30, seed=12) + [(256, 256, 256)]
    bufs, host = inputs(torch, dims, 970)
    with G.Batch(dims) as b:
        b.set_data(bufs)
        b.set_wait_limit(0.0)
        for _ in range(3):
            b.exec_graph(G.ROUND_TRIP)
        torch.cuda.synchronize()
        assert b.status() in (0, G.ERR_TIMEOUT)
        b.set_wait_limit(2.0)
        work = [torch.from_numpy(h).to("cuda:0") for h in host]
        b.set_data(work)
        b.exec_graph(G.R2C)
        torch.cuda.synchronize()
        assert b.status() == 0
        for (nx, ny, nz), buf, fp in zip(dims, work, host):
            got = G.complex_view(buf, nx, ny, nz).cpu().numpy()
            assert O.rel_l2(got, O.r2c_p1(fp.astype(np.float64), nz)) <= 1e-6


@pytest.mark.parametrize("dims", [[(192, 192, 192)], [(256, 256, 256), (160, 160, 160)] + [(32, 32, 32)] * 12])
def test_span_teams_match_one_xcd_teams_bitwise(dims):
    """A transform whose share of workgroups exceeds one XCD's slots runs as
    a team spanning two XCDs (span 2, the default), each level handed over
    through one L2 write-back per XCD part; with span 1 the same transform
    runs as a team in one XCD.  The same tile code runs either way, so the
    round trips are bitwise equal, and they match the FP64 oracle; three
    repetitions (a part reading lines another XCD had not written back would
    show as large errors)."""
    import torch
    for rep in range(3):
        say(f"case span teams {dims[0]} +{len(dims) - 1}, rep {rep}")
        out = {}
        for span in (2, 1):
            bufs, host = inputs(torch, dims, 1500 + rep)
            with G.Batch(dims) as b:
                b.set_persist(1, 2)
                b.set_span(span)
                assert b.span() == span
                info = b.plan_info(G.ROUND_TRIP)
                if len(dims) == 1:  # one team: more members than one XCD's slots only when it spans
                    assert (info["members"] > info["slots"] // 8) == (span == 2), info
                b.set_scales(c2r=[1.0 / (nx * ny * nz) for nx, ny, nz in dims])
                b.set_data(bufs)
                b.exec(G.ROUND_TRIP)
                torch.cuda.synchronize()
                assert b.status() == 0
            out[span] = bufs
            for (nx, ny, nz), buf, fp in zip(dims, bufs, host):
                err = O.rel_l2(buf.cpu().numpy()[:, :, :nz], fp[:, :, :nz])
                assert err <= 1e-6, ((nx, ny, nz), span, rep, err)
        for x, y, d in zip(out[2], out[1], dims):
            assert torch.equal(x, y), d


def test_single_transform_teams_repeated_match_oracle():
    """A single 48^3 or 64^3 transform runs as one large team (19-32 members
    in one XCD) whose levels hand data over through the L2; each member must
    read the previous level's data, not lines it kept in its L1.  Repeated
    round trips against the FP64 oracle (a stale read left errors of 3e-2 to
    1.6e-1 in about one run in three)."""
    import torch
    for n in (48, 64):
        dims = [(n, n, n)]
        say(f"case single-transform team {n}^3")
        for rep in range(4):
            bufs, host = inputs(torch, dims, 500 + 10 * n + rep)
            with G.Batch(dims) as b:
                assert b.scheduler() == G.Batch.PERSISTENT
                b.set_scales(c2r=[1.0 / n ** 3])
                b.set_data(bufs)
                assert b.plan_info(G.ROUND_TRIP)["teams"] == 1
                b.exec(G.ROUND_TRIP)
                torch.cuda.synchronize()
                assert b.status() == 0
            err = O.rel_l2(bufs[0].cpu().numpy()[:, :, :n], host[0][:, :, :n])
            assert err <= 1e-6, (n, rep, err)


def test_global_level_plan_alone_mix_graph_abort():
    """The global-level plan (team factor 0): every workgroup takes shares of
    every transform's level and the level hand-over crosses XCDs (one agent-
    scope release per workgroup and level, then an acquire).  Each result is
    bitwise the same alone and in the mix, graph replays (with eager work
    between them) equal the eager launch, an abort request and a zero wait
    bound drain and report, and the next launch is right."""
    import torch
    dims = mix(40, seed=21) + [(192, 192, 192), (160, 160, 160)]
    bufs, host = inputs(torch, dims, 1100)
    scales = [1.0 / (nx * ny * nz) for nx, ny, nz in dims]
    say(f"case global-level plan, {len(dims)} transforms")
    with G.Batch(dims) as b:
        b.set_persist(0, 0)
        b.set_scales(c2r=scales)
        b.set_data(bufs)
        b.exec(G.ROUND_TRIP)
        torch.cuda.synchronize()
        assert b.status() == 0
        eager = [x.clone() for x in bufs]
        for i, ((nx, ny, nz), fp) in enumerate(zip(dims, host)):
            err = O.rel_l2(bufs[i].cpu().numpy()[..., :nz], fp[..., :nz])
            assert err <= 1e-6, ((nx, ny, nz), err)
        for rep in range(3):
            work = [torch.from_numpy(h).to("cuda:0") for h in host]
            b.set_data(work)
            b.exec_graph(G.ROUND_TRIP)
            torch.cuda.synchronize()
            _ = torch.ones(1000, device="cuda:0").sum().item()  # eager work between replays
            assert b.status() == 0
            for w, e in zip(work, eager):
                assert torch.equal(w, e), rep
        for i in (0, 5, len(dims) - 2, len(dims) - 1):
            alone = [torch.from_numpy(host[i]).to("cuda:0")]
            with G.Batch([dims[i]]) as one:
                one.set_persist(0, 0)
                one.set_scales(c2r=[scales[i]])
                one.set_data(alone)
                one.exec(G.ROUND_TRIP)
                torch.cuda.synchronize()
                assert one.status() == 0
            assert torch.equal(alone[0], eager[i]), dims[i]
        b.set_abort(True)
        b.exec(G.ROUND_TRIP)
        torch.cuda.synchronize()
        assert b.status() == G.ERR_ABORTED
        b.set_abort(False)
        b.set_wait_limit(0.0)
        b.exec_graph(G.ROUND_TRIP)
        torch.cuda.synchronize()
        assert b.status() in (0, G.ERR_TIMEOUT)
        b.set_wait_limit(2.0)
        work = [torch.from_numpy(h).to("cuda:0") for h in host]
        b.set_data(work)
        b.exec_graph(G.R2C)
        torch.cuda.synchronize()
        assert b.status() == 0
        for (nx, ny, nz), buf, fp in zip(dims, work, host):
            got = G.complex_view(buf, nx, ny, nz).cpu().numpy()
            assert O.rel_l2(got, O.r2c_p1(fp.astype(np.float64), nz)) <= 1e-6


def test_global_level_single_transforms_repeated_match_oracle():
    """One transform alone under the global-level plan: its levels are cut
    over every workgroup of every XCD, so each level's readers run on other
    XCDs than its writers and must see the written data, not lines kept in
    their L1 or L2.  Repeated round trips against the input (FP32 rel L2
    <= 1e-6)."""
    import torch
    for n in (48, 64, 192):
        dims = [(n, n, n)]
        say(f"case global-level single transform {n}^3")
        for rep in range(3):
            bufs, host = inputs(torch, dims, 1300 + 10 * n + rep)
            with G.Batch(dims) as b:
                b.set_persist(rep % 2, 0)
                b.set_scales(c2r=[1.0 / n ** 3])
                b.set_data(bufs)
                for _ in range(2):
                    b.exec(G.ROUND_TRIP)
                torch.cuda.synchronize()
                assert b.status() == 0
            err = O.rel_l2(bufs[0].cpu().numpy()[:, :, :n], host[0][:, :, :n])
            assert err <= 1e-6, (n, rep, err)

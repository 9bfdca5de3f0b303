"""Allocation-free batched entry point (gfft_batch_exec_noalloc): the
persistent scheduler's launch with its counter words in the library's static
per-GPU array instead of the batch's own allocation.  Checked against
gfft_batch_exec under the persistent scheduler on the same inputs, on two
calls in a row (the second with a smaller batch, so it runs on the counters
the first call left behind), both bitwise (same kernel, same plan) and
against the FP64 oracle (FP32 tolerance 1e-6); no failure bits afterwards.
Both plan kinds: team plans (the default team factor) and the global-level
plan (team factor 0), whose counter words are laid out differently."""
import sys

import numpy as np
import pytest

import gpu_fft_amd as G
from oracle import gfft_oracle as O

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def _inputs(torch, dims, seed):
    host = [O.pad_real(O.random_field(nx, ny, nz, seed=seed + i)).astype(np.float32)
            for i, (nx, ny, nz) in enumerate(dims)]
    return host, [torch.from_numpy(h).to("cuda:0") for h in host]


def _run(torch, dims, seed, mode, team_factor):
    host, a = _inputs(torch, dims, seed)
    _, b = _inputs(torch, dims, seed)
    with G.Batch(dims) as ba, G.Batch(dims) as bb:
        if team_factor is not None:
            ba.set_persist(0, team_factor)
            bb.set_persist(0, team_factor)
        ba.set_scheduler(G.Batch.PERSISTENT)
        ba.set_data(a)
        ba.exec(mode)
        bb.set_scheduler(G.Batch.LEVELS)  # the entry point runs the persistent launch whatever the scheduler
        bb.set_data(b)
        bb.exec_noalloc(mode)
        torch.cuda.synchronize()
        assert ba.status() == 0 and bb.status() == 0
        info = bb.plan_info(mode)
    return host, a, b, info


@pytest.mark.parametrize("team_factor", [None, 0])
@pytest.mark.parametrize("mode", [G.R2C, G.ROUND_TRIP])
def test_noalloc_matches_exec_on_two_calls_second_smaller(mode, team_factor):
    import torch
    # first call: a large transform (a team of workgroups) among small ones;
    # second call: a smaller batch
    first = [(256, 256, 256), (192, 192, 192)] + [(32, 32, 32)] * 20 + [(64, 64, 64)] * 6 + [(1, 64, 64)]
    second = [(128, 128, 128)] + [(48, 48, 48)] * 5 + [(96, 1, 96)]
    for call, dims in enumerate((first, second)):
        say(f"case noalloc mode={mode} team_factor={team_factor} call={call} transforms={len(dims)}")
        host, a, b, info = _run(torch, dims, 40 + 10 * call, mode, team_factor)
        if team_factor == 0:
            assert info["teams"] == 0 and info["members"] == info["slots"]  # every workgroup in every level
        elif call == 0:
            assert info["teams"] > 0  # the shared counter words are exercised
        for (nx, ny, nz), x, y, fp in zip(dims, a, b, host):
            assert torch.equal(x, y), (nx, ny, nz)
            if mode == G.R2C:
                got = G.complex_view(y, nx, ny, nz).cpu().numpy()
                err = O.rel_l2(got, O.r2c_p1(fp.astype(np.float64), nz))
            else:
                err = O.rel_l2(y.cpu().numpy()[:, :, :nz] / (nx * ny * nz), fp[:, :, :nz])
            assert err <= 1e-6, ((nx, ny, nz), err)

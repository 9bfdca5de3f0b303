"""Slab decomposition on one GPU: the local halves (gfft_exec_r2c_local /
_finish, gfft_exec_c2r_start / _local) of p simulated ranks, with the MPI
exchange of FFT_R2C.cu:51-63 replaced by device copies of the chunks,
against the oracle's multi-rank restatement (all ranks, the reference's
transpose.cu index maps).  Also the p = 8 share of a 1024^3 grid at its real
shape (128 x 1024 x 1024 local block): z R2C + y C2C + pack."""
import sys

import numpy as np
import pytest

import gpu_fft_amd as G
from oracle import gfft_oracle as O

pytestmark = pytest.mark.gpu
TOL = {"f32": 1e-6, "f64": 1e-13}


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def _exchange(torch, send, procs):
    """recv[r] chunk s = send[s] chunk r (the all-to-all of the reference)."""
    chunk = send[0].numel() // procs
    recv = [torch.empty_like(b) for b in send]
    for r in range(procs):
        for s in range(procs):
            recv[r].view(-1)[s * chunk:(s + 1) * chunk].copy_(send[s].view(-1)[r * chunk:(r + 1) * chunk])
    return recv


@pytest.mark.parametrize("prec", ["f32", "f64"])
@pytest.mark.parametrize("shape,procs", [((8, 8, 8), 2), ((16, 8, 12), 4), ((12, 6, 10), 2), ((8, 16, 6), 8),
                                         ((32, 32, 24), 4), ((64, 48, 30), 2), ((16, 16, 16), 1),
                                         ((64, 64, 64), 8), ((40, 40, 27), 4)],
                         ids=lambda v: "x".join(map(str, v)) if isinstance(v, tuple) else str(v))
def test_simulated_ranks_match_multirank_oracle(shape, procs, prec):
    import torch
    nx, ny, nz = shape
    say(f"case slab {shape} p={procs} {prec}")
    rdt = np.float32 if prec == "f32" else np.float64
    f = O.random_field(nx, ny, nz, seed=nx + 3 * procs)
    nxl = nx // procs
    slabs = [O.pad_real(f[r * nxl:(r + 1) * nxl]) for r in range(procs)]
    want_fwd = O.gpu_fft_r2c_ranks(slabs, nx, ny, nz)
    want_inv = O.gpu_fft_c2r_ranks(want_fwd, nx, ny, nz)
    plans = [G.Plan(nx, ny, nz, procs=procs, rank=r, precision=prec) for r in range(procs)]
    data = [torch.from_numpy(s.astype(rdt)).to("cuda:0") for s in slabs]
    send = [torch.empty_like(d) for d in data]
    out = [torch.empty_like(d) for d in data]
    for r in range(procs):
        plans[r].r2c_local(data[r], send[r])
    recv = _exchange(torch, send, procs)
    for r in range(procs):
        plans[r].r2c_finish(recv[r], out[r])
    torch.cuda.synchronize()
    K = O.kdim(nz)
    for r in range(procs):
        got = G.complex_view(out[r], nx, ny // procs, nz).cpu().numpy()
        err = O.rel_l2(got, want_fwd[r])
        assert err <= TOL[prec], (r, err)
    # inverse: out (y-slabs) -> x-slabs
    for r in range(procs):
        plans[r].c2r_start(out[r], send[r])
    recv = _exchange(torch, send, procs)
    for r in range(procs):
        plans[r].c2r_local(recv[r], data[r])
    torch.cuda.synchronize()
    for r in range(procs):
        got = data[r].cpu().numpy().reshape(nxl, ny, 2 * K)[:, :, :nz]
        err = O.rel_l2(got, want_inv[r][:, :, :nz])
        assert err <= TOL[prec], (r, err)
    for p in plans:
        p.close()


def test_local_halves_equal_full_transform_at_p1():
    import torch
    nx, ny, nz = 32, 24, 40
    fp = O.pad_real(O.random_field(nx, ny, nz, seed=4)).astype(np.float32)
    a = torch.from_numpy(fp).to("cuda:0")
    b = a.clone()
    tmp = torch.empty_like(a)
    with G.Plan(nx, ny, nz) as p:
        p.r2c(a)
        p.r2c_local(b, tmp)
        p.r2c_finish(tmp, b)
        torch.cuda.synchronize()
        assert torch.equal(a, b)
        p.c2r(a)
        p.c2r_start(b, tmp)
        p.c2r_local(tmp, b)
        torch.cuda.synchronize()
        assert torch.equal(a[:, :, :nz], b[:, :, :nz])


def test_aliasing_rejected():
    import torch
    a = torch.zeros(8 * 8 * 10, device="cuda:0")
    with G.Plan(8, 8, 8) as p:
        with pytest.raises(G.GfftError):
            p.r2c_local(a, a)


@pytest.mark.slow
def test_p8_share_of_1024_cube_real_shape():
    """Rank 3 of 1024^3 at p = 8: 128 x 1024 x 1024 local real block.
    z R2C + y C2C + pack; every x-plane's 2D spectrum must appear in the
    (Ny, Nx/p, K) send layout."""
    import torch
    nx = ny = nz = 1024
    procs, rank = 8, 3
    nxl, K = nx // procs, O.kdim(nz)
    say("case p=8 share 128x1024x1024 f32")
    torch.manual_seed(0)
    data = torch.rand((nxl, ny, 2 * K), device="cuda:0") * 2 - 1
    data[:, :, nz:] = 0
    host_in = data.cpu().numpy()
    send = torch.empty_like(data)
    with G.Plan(nx, ny, nz, procs=procs, rank=rank) as p:
        p.r2c_local(data, send)
        torch.cuda.synchronize()
    got = G.complex_view(send, ny, nxl, nz).cpu().numpy()       # (Ny, Nx/p, K)
    err_max = 0.0
    for x0 in range(0, nxl, 32):
        blk = host_in[x0:x0 + 32, :, :nz].astype(np.float64)
        ref = np.fft.rfft2(blk, axes=(1, 2))                       # (32, Ny, K)
        err = O.rel_l2(np.transpose(got[:, x0:x0 + 32, :], (1, 0, 2)), ref)
        err_max = max(err_max, err)
    say(f"  max block rel L2 {err_max:.2e}")
    assert err_max <= 1e-6
    del data, send
    torch.cuda.empty_cache()


def _rand_chunks(torch, n_complex, dt, seed):
    g = torch.Generator(device="cuda:0")
    g.manual_seed(seed)
    return torch.rand(2 * n_complex, dtype=dt, device="cuda:0", generator=g) * 2 - 1


@pytest.mark.slow
@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_p8_share_finish_start_local_real_shape(prec):
    """The other three local halves of rank 3's share of 1024^3 at p = 8, at
    their real shapes, on synthetic exchange blocks, against FP64 numpy
    block by block:
      r2c_finish: received chunks (p, Ny/p, Nx/p, K) = (8, 128, 128, 513),
        chunk s holding x-planes s*128.. -> x C2C forward -> (1024, 128, 513);
      c2r_start:  a (1024, 128, 513) y-slab -> x C2C inverse -> packed into
        chunk s = x-planes s*128.. in (Ny/p, Nx/p, K) order;
      c2r_local:  received chunks, chunk s holding y rows s*128.. -> y C2C
        inverse + z C2R (numpy irfft convention) -> (128, 1024, 1026) real."""
    import torch
    nx = ny = nz = 1024
    procs, rank = 8, 3
    nxl, nyl, K = nx // procs, ny // procs, O.kdim(nz)
    dt = torch.float32 if prec == "f32" else torch.float64
    ncx = nx * nyl * K                      # complex elements of one exchange buffer
    say(f"case p=8 share finish/start/local {prec}")
    with G.Plan(nx, ny, nz, procs=procs, rank=rank, precision=prec) as p:
        # r2c_finish
        recv = _rand_chunks(torch, ncx, dt, 11)
        out = torch.empty_like(recv)
        p.r2c_finish(recv, out)
        torch.cuda.synchronize()
        r = recv.view(-1, 2).cpu().numpy().astype(np.float64)
        rc = (r[:, 0] + 1j * r[:, 1]).reshape(procs, nyl, nxl, K)
        got = G.complex_view(out, nx, nyl, nz).cpu().numpy()
        a = np.transpose(rc, (0, 2, 1, 3)).reshape(nx, nyl, K)   # x = s*128 + x_l
        err = max(O.rel_l2(got[:, :, k0:k0 + 64], np.fft.fft(a[:, :, k0:k0 + 64], axis=0))
                  for k0 in range(0, K, 64))
        say(f"  r2c_finish max block rel L2 {err:.2e}")
        assert err <= TOL[prec]
        # c2r_start: reuse `out` (a y-slab) as input
        yslab = G.complex_view(out, nx, nyl, nz).cpu().numpy().astype(np.complex128)
        send = torch.empty_like(recv)
        p.c2r_start(out, send)
        torch.cuda.synchronize()
        s_ = send.view(-1, 2).cpu().numpy().astype(np.float64)
        sc = (s_[:, 0] + 1j * s_[:, 1]).reshape(procs, nyl, nxl, K)
        want = np.fft.ifft(yslab, axis=0) * nx                       # unnormalised inverse
        wc = np.transpose(want.reshape(procs, nxl, nyl, K), (0, 2, 1, 3))
        err = max(O.rel_l2(sc[s], wc[s]) for s in range(procs))
        say(f"  c2r_start max chunk rel L2 {err:.2e}")
        assert err <= TOL[prec]
        del yslab, want, wc, sc, s_, a, rc, r, got
        # c2r_local
        recv2 = _rand_chunks(torch, ncx, dt, 12)
        data = torch.empty((nxl, ny, 2 * K), dtype=dt, device="cuda:0")
        r = recv2.view(-1, 2).cpu().numpy().astype(np.float64)
        rc = (r[:, 0] + 1j * r[:, 1]).reshape(procs, nyl, nxl, K)
        b = np.transpose(rc, (2, 0, 1, 3)).reshape(nxl, ny, K)       # y = s*128 + y_l
        p.c2r_local(recv2, data)
        torch.cuda.synchronize()
        got = data.cpu().numpy()[:, :, :nz]
        err = 0.0
        for x0 in range(0, nxl, 16):
            ref = np.fft.irfft2(b[x0:x0 + 16], s=(ny, nz), axes=(1, 2)) * (ny * nz)
            err = max(err, O.rel_l2(got[x0:x0 + 16], ref))
        say(f"  c2r_local max block rel L2 {err:.2e}")
        assert err <= TOL[prec]
    del recv, out, send, recv2, data
    torch.cuda.empty_cache()

"""torch functional API on the GPU: rfftn / irfftn against the FP64 oracle
and torch.fft, round trips, and the in-place variants' views."""
import numpy as np
import pytest

import gpu_fft_amd as G
from gpu_fft_amd import torch_api as T
from oracle import gfft_oracle as O

pytestmark = pytest.mark.gpu
TOL = {"f32": 1e-6, "f64": 1e-13}


@pytest.mark.parametrize("shape,prec", [((64, 48, 96), "f32"), ((75, 81, 125), "f64"), ((128, 128, 128), "f32"),
                                        ((36, 1, 40), "f64")])
def test_rfftn_irfftn_match_oracle_and_torch(shape, prec):
    import torch
    dt = torch.float32 if prec == "f32" else torch.float64
    f = O.random_field(*shape, seed=9).astype(np.float32 if prec == "f32" else np.float64)
    x = torch.from_numpy(f).to("cuda:0")
    X = T.rfftn(x)
    torch.cuda.synchronize()
    assert X.shape == (shape[0], shape[1], shape[2] // 2 + 1) and X.is_complex()
    want = O.r2c_p1(f.astype(np.float64), shape[2])
    assert O.rel_l2(X.cpu().numpy(), want) <= TOL[prec]
    ref = torch.fft.rfftn(x.to(torch.float64))
    assert float(torch.linalg.vector_norm(X.to(torch.complex128) - ref) / torch.linalg.vector_norm(ref)) <= TOL[prec]
    y = T.irfftn(X, shape)
    torch.cuda.synchronize()
    assert y.shape == shape and y.dtype == dt and y.is_contiguous()
    assert O.rel_l2(y.cpu().numpy(), f.astype(np.float64)) <= TOL[prec]
    assert float(x.abs().max()) == float(torch.from_numpy(f).abs().max())  # input untouched


def test_in_place_views_share_memory():
    import torch
    buf = G.alloc_inplace(32, 16, 20)
    buf[..., :20].copy_(torch.rand(32, 16, 20, device="cuda:0"))
    ref = buf[..., :20].clone()
    X = T.rfftn_(buf, 20)
    assert X.data_ptr() == buf.data_ptr() and X.shape == (32, 16, 11)
    y = T.irfftn_(X, 20, normalize=True)
    torch.cuda.synchronize()
    assert y.data_ptr() == buf.data_ptr()
    assert O.rel_l2(y.cpu().numpy(), ref.cpu().numpy()) <= 1e-6

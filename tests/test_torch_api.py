"""torch functional API: CPU-side argument checks (no GPU needed)."""
import pytest

torch = pytest.importorskip("torch")

from gpu_fft_amd import torch_api as T  # noqa: E402


def test_cpu_tensors_are_refused_not_computed():
    x = torch.zeros(4, 4, 4)
    with pytest.raises(ValueError):
        T.rfftn(x)
    with pytest.raises(ValueError):
        T.rfftn_(torch.zeros(4, 4, 6), 4)
    with pytest.raises(ValueError):
        T.irfftn(torch.zeros(4, 4, 3, dtype=torch.complex64), (4, 4, 4))

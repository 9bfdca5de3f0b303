import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (run with -m gpu on the GPU box)")
    config.addinivalue_line("markers", "slow: large-grid GPU cases (still run by -m gpu)")


import pytest  # noqa: E402


@pytest.fixture(autouse=True)
def _fresh_tune_cache():
    """Every test starts with an empty process-wide tune cache, so new plans
    start with the default variants unless the test tunes them itself (plans
    take the cached choices of their geometry)."""
    try:
        import gpu_fft_amd as G
        G.clear_tune_cache()
    except (ImportError, OSError):
        pass
    yield

"""Host-only checks of the serving runtime: the arena / admission bookkeeping
(native unit test built with g++), request sizing and rejection of bad
geometry, and that every serving size plans in both precisions."""
import os
import subprocess

import pytest

import gpu_fft_amd as G
from gpu_fft_amd import serving as S

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_arena_and_admission_native(tmp_path):
    exe = tmp_path / "test_arena"
    src = os.path.join(ROOT, "tests", "native", "test_arena.cpp")
    inc = os.path.join(ROOT, "gpu_fft_amd", "csrc")
    subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-fsanitize=address,undefined", "-I", inc, src,
                    "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    assert out.stdout.strip().endswith("ok")


def test_serving_sizes():
    assert len(S.SERVE_SIZES) == 42
    assert [n for n in S.SERVE_SIZES if n % 2] == [75, 81, 125, 135, 225, 243, 375, 405]


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_every_serving_size_plans(prec):
    for n in S.SERVE_SIZES:
        q = G.query_plan(n, n, n, precision=prec)
        assert q["nz"] == n and q["precision"] == prec
        assert S.request_bytes(n, n, n, prec) == n * n * 2 * (n // 2 + 1) * (4 if prec == "f32" else 8)


def test_request_bytes_matches_config_table():
    # S of the 512^3 FP64 request, the largest of the serving mix
    assert S.request_bytes(512, 512, 512, "f64") == 1_077_936_128


@pytest.mark.parametrize("dims", [(0, 8, 8), (8, -1, 8), (8, 8, 0)])
def test_request_bytes_rejects_bad_geometry(dims):
    with pytest.raises(G.GfftError) as e:
        S.request_bytes(*dims)
    assert e.value.status == 2


def test_request_bytes_rejects_bad_precision():
    with pytest.raises(ValueError):
        S.request_bytes(8, 8, 8, "f16")


def test_clock_is_monotonic():
    a = S.clock()
    b = S.clock()
    assert b >= a > 0

"""Property tests on the CPU (hypothesis, derandomised): the planner's
factorisation of any edge up to 2^20, and the oracle's pocketfft path
against its own O(N^2) DFT on random small grids."""
import math

import numpy as np
import pytest
from hypothesis import given, settings, strategies as st

import gpu_fft_amd as G
from oracle import gfft_oracle as O

RADICES = {2, 3, 4, 5, 7, 8, 11, 13}


def largest_prime_factor(n):
    p, f = 2, 1
    while p * p <= n:
        while n % p == 0:
            f, n = p, n // p
        p += 1
    return max(f, n) if n > 1 else f


@settings(max_examples=200, deadline=None, derandomize=True)
@given(st.integers(min_value=1, max_value=4096))
def test_planner_factorises_or_rejects_any_edge(n):
    """A z edge plans iff its prime factors are <= 13 and it fits the
    LDS-resident limit; a planned edge's radices multiply back to n."""
    try:
        d = G.query_plan(2, 2, n, precision="f64")
    except G.GfftError as e:
        assert e.status == 3
        assert largest_prime_factor(n) > 13 or n > 1024
        return
    assert largest_prime_factor(n) <= 13
    r = d["z"]["radices"]
    assert math.prod(r) == n and set(r) <= RADICES


small_edge = st.sampled_from([1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16])


@settings(max_examples=40, deadline=None, derandomize=True)
@given(small_edge, small_edge, small_edge, st.integers(min_value=0, max_value=2**31))
def test_oracle_rfftn_equals_brute_dft(nx, ny, nz, seed):
    f = O.random_field(nx, ny, nz, seed=seed)
    want = O.brute_rfftn(f)
    got = O.r2c_p1(O.pad_real(f), nz)
    assert got.shape == want.shape
    assert O.rel_l2(got, want) <= 1e-13 or np.abs(want).max() == 0

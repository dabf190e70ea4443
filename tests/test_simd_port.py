"""The SimdBackend-equivalent CPU port (oracle/simd_port.cpp, the bench's
CPU baseline) against the oracle: interpolate, evaluate at a blowup and the
Merkle tree, bit-exact, for the AVX2 build and (where the host has AVX-512)
the AVX-512 build, across the in-cache (<= 2^16) and blocked (> 2^16)
shapes and column counts that are not a multiple of the 16 words of a
Blake2s block."""
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "oracle"))
import oracle  # noqa: E402
import simd_port  # noqa: E402

VARIANTS = ["v3"] + (["v4"] if simd_port.host_has_avx512() else [])


@pytest.mark.parametrize("variant", VARIANTS)
@pytest.mark.parametrize("n,b,C", [(5, 1, 3), (6, 3, 1), (8, 2, 4), (12, 1, 16), (16, 1, 17), (17, 2, 5), (19, 1, 2)])
def test_port_matches_oracle(variant, n, b, C):
    N = n + b
    tw, itw = oracle.precompute_twiddles(oracle.half_odds_initial(N - 1), N - 1)
    port = simd_port.Port(tw, itw, variant)
    evs = [oracle.fill_random(1 << n, 11, c) for c in range(C)]
    if C > 1:
        evs[1][:] = oracle.P - 1  # extreme values
    want = [e.copy() for e in evs]
    oracle.interpolate_columns(want, itw)
    got = [e.copy() for e in evs]
    port.interpolate_columns(got, n)
    for g, w in zip(got, want):
        assert np.array_equal(g, w)
    ew = oracle.evaluate_polynomials(want, b, tw)
    eg = port.evaluate_columns(got, n, b)
    for g, w in zip(eg, ew):
        assert np.array_equal(g, w)
    assert np.array_equal(port.merkle_commit(eg, N), oracle.merkle_commit(ew))


def test_port_rejects_small_domains():
    tw, itw = oracle.precompute_twiddles(oracle.half_odds_initial(3), 3)
    port = simd_port.Port(tw, itw, "v3")
    with pytest.raises(AssertionError):
        port.interpolate_columns([np.zeros(8, np.uint32)], 3)

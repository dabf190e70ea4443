"""Pins the SimdBackend layout restatement (oracle/simd_layout.py).  CPU only.

transpose_vecs (STWO/prover/backend/simd/fft/mod.rs:22-56) is checked
against the bit permutation its doc states -- vector index bits abc -> cba,
|a| = |c|, |b| = log_n_vecs mod 2 -- computed here on binary strings, and
for being its own inverse.
"""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "..", "oracle"))
import simd_layout as S  # noqa: E402


@pytest.mark.parametrize("log_n_vecs", [1, 2, 3, 4, 5, 8, 9])
def test_transpose_vecs_is_the_stated_bit_permutation(log_n_vecs):
    n = 1 << log_n_vecs
    vals = np.arange(16 * n, dtype=np.uint32)
    out = S.transpose_vecs(vals, log_n_vecs).reshape(n, 16)
    half = log_n_vecs // 2
    for i in range(n):
        s = format(i, "0%db" % log_n_vecs)
        a, b, c = s[:half], s[half:log_n_vecs - half], s[log_n_vecs - half:]
        j = int(c + b + a, 2)
        assert out[i].tolist() == list(range(16 * j, 16 * j + 16)), (i, j)
    assert np.array_equal(S.transpose_vecs(out.reshape(-1), log_n_vecs), vals)


def test_small_columns_keep_natural_order():
    v = np.arange(1 << 16, dtype=np.uint32)
    assert np.array_equal(S.simd_coeff_order(v, 16), v)
    w = np.arange(1 << 17, dtype=np.uint32)
    assert not np.array_equal(S.simd_coeff_order(w, 17), w)

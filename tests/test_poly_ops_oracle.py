"""Pins the oracle's restatements of barycentric_weights /
barycentric_eval_at_point, eval_at_point_by_folding, split_at_mid and
FriOps::decompose (oracle/poly_ops.py) against checks from outside them:
the C oracle's eval_at_point of the interpolated polynomial (itself pinned by
brute-force evaluation and the reference KATs), the reference test
test_cpu_eval_at_point_by_folding (coefficients [691, 805673, 5, 435684,
4832, 23876431, 197, 897346068] on CanonicCoset(10) at get_point(348),
get_point(9736524), get_point(13), get_point(346752)), the split identity
p(z) = left(z) + pi^(L-2)(z.x) * right(z), and decompose's lemma: g is in
the FFT space and f = g + lambda * (+1 on the first half, -1 on the second)
(decompose_coeff_out_fft_space_test)."""
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "oracle"))
sys.path.insert(0, HERE)
import oracle as O  # noqa: E402
import poly_ops as PO  # noqa: E402
from prove import double_x, pt_m31, q_add, q_mul  # noqa: E402
from test_quotients_oracle import secure_point  # noqa: E402

P = O.P


def evals_of(coeffs, n):
    tw, itw = O.precompute_twiddles(O.half_odds_initial(n - 1), n - 1)
    return O.evaluate(np.ascontiguousarray(coeffs, dtype=np.uint32), O.half_odds_initial(n - 1), n, tw)


@pytest.mark.parametrize("n", [1, 2, 3, 5, 8])
def test_barycentric_matches_eval_at_point(n):
    rng = np.random.default_rng(n)
    coeffs = rng.integers(0, P, 1 << n, dtype=np.uint64).astype(np.uint32)
    ev = evals_of(coeffs, n)
    for k in (3, 77):
        pt = secure_point(k)
        w = PO.barycentric_weights(n, pt)
        assert PO.barycentric_eval_at_point(ev, w) == O.eval_at_point(coeffs, pt)


def test_barycentric_a_wrong_weight_is_caught():
    n = 5
    coeffs = np.arange(1, 33, dtype=np.uint32)
    ev = evals_of(coeffs, n)
    pt = secure_point(9)
    w = PO.barycentric_weights(n, pt)
    w[7] = q_add(w[7], (1, 0, 0, 0))
    assert PO.barycentric_eval_at_point(ev, w) != O.eval_at_point(coeffs, pt)


def test_eval_at_point_by_folding_reference_kat():
    coeffs = np.array([691, 805673, 5, 435684, 4832, 23876431, 197, 897346068], dtype=np.uint32)
    n = 10
    full = np.zeros(1 << n, dtype=np.uint32)
    full[:8] = coeffs
    ev = evals_of(full, n)
    for idx in (348, 9736524, 13, 346752):
        pt = pt_m31(idx)
        assert PO.eval_at_point_by_folding(ev, n, pt) == O.eval_at_point(coeffs, pt)


@pytest.mark.parametrize("n", [2, 4, 9])
def test_eval_at_point_by_folding_secure_points(n):
    rng = np.random.default_rng(40 + n)
    coeffs = rng.integers(0, P, 1 << n, dtype=np.uint64).astype(np.uint32)
    ev = evals_of(coeffs, n)
    pt = secure_point(1234 + n)
    assert PO.eval_at_point_by_folding(ev, n, pt) == O.eval_at_point(coeffs, pt)


@pytest.mark.parametrize("L", [1, 2, 3, 6])
def test_split_at_mid_identity(L):
    rng = np.random.default_rng(L)
    coeffs = rng.integers(0, P, 1 << L, dtype=np.uint64).astype(np.uint32)
    left, right = PO.split_at_mid(coeffs)
    pt = secure_point(L + 5)
    x = pt[0]
    for _ in range(L - 2):
        x = double_x(x)
    factor = pt[1] if L == 1 else x  # the half's top basis element: y for L = 1, pi^(L-2)(x) otherwise
    got = q_add(O.eval_at_point(left, pt), q_mul(factor, O.eval_at_point(right, pt)))
    assert got == O.eval_at_point(coeffs, pt)


@pytest.mark.parametrize("n", [5, 8, 11])
def test_decompose_lemma(n):
    half_log = n - 1
    # out of the FFT space: the coefficient at 2^(n-1) set, as the reference test
    coeffs = np.zeros(1 << n, dtype=np.uint32)
    coeffs[1 << half_log] = 1
    rng = np.random.default_rng(n)
    coeffs[: 1 << half_log] = rng.integers(0, P, 1 << half_log, dtype=np.uint64).astype(np.uint32)
    ev = evals_of(coeffs, n)
    eval4 = [ev.copy(), ev.copy(), (ev.astype(np.uint64) * 3 % P).astype(np.uint32), ev.copy()]
    g, lam = PO.decompose(eval4)
    assert lam != (0, 0, 0, 0)
    tw, itw = O.precompute_twiddles(O.half_odds_initial(n - 1), n - 1)
    for k in range(4):
        c = O.interpolate(g[k].copy(), O.half_odds_initial(n - 1), n, itw)
        nz = np.nonzero(c)[0]
        assert len(nz) == 0 or nz.max() < (1 << half_log)  # is_in_fft_space(n - 1)
        half = 1 << half_log
        back = np.concatenate([(g[k][:half].astype(np.int64) + lam[k]) % P, (g[k][half:].astype(np.int64) - lam[k]) % P])
        assert np.array_equal(back.astype(np.uint32), eval4[k])

"""GPU parity of barycentric_weights / barycentric_eval_at_point,
eval_at_point_by_folding, split_at_mid and FriOps::decompose against the
oracle's restatements (oracle/poly_ops.py, pinned by
tests/test_poly_ops_oracle.py), bit-exact; and at sizes the Python oracle
cannot reach, against the C oracle's eval_at_point of the interpolated
polynomial (barycentric == eval_at_point is an identity of the field)."""
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "oracle"))
sys.path.insert(0, HERE)
import oracle as O  # noqa: E402
import poly_ops as PO  # noqa: E402
import stwo_hip as sh  # noqa: E402
from prove import pt_m31  # noqa: E402
from test_quotients_oracle import secure_point  # noqa: E402

pytestmark = pytest.mark.gpu
P = O.P


@pytest.fixture(scope="module")
def ctx():
    c = sh.Context(0)
    yield c
    c.close()


def evals_of(coeffs, n):
    tw, itw = O.precompute_twiddles(O.half_odds_initial(n - 1), n - 1)
    return O.evaluate(np.ascontiguousarray(coeffs, dtype=np.uint32), O.half_odds_initial(n - 1), n, tw)


@pytest.mark.parametrize("n", [1, 2, 4, 7, 10])
def test_barycentric_weights_vs_oracle(ctx, n):
    pt = secure_point(100 + n)
    out = [ctx.alloc(1 << n) for _ in range(4)]
    ctx.barycentric_weights(n, pt, out)
    got = [o.download() for o in out]
    want = PO.barycentric_weights(n, pt)
    for i, w in enumerate(want):
        assert tuple(int(g[i]) for g in got) == w, i
    for o in out:
        o.free()


@pytest.mark.parametrize("n,C,offset", [(1, 1, 0), (3, 5, 0), (10, 9, 0), (11, 3, 0), (12, 17, 0), (16, 2, 0),
                                         (20, 3, 0), (13, 70, 0), (12, 11, 1)])
def test_barycentric_eval_matches_eval_at_point(ctx, n, C, offset):
    """Weights and dot products on the device; the values against the C
    oracle's eval_at_point of each column's coefficients (columns 32 to a
    workgroup in steps of 8, with remainders; one column all P - 1; with
    offset 1 the columns start one word into their buffers, off the 16-byte
    loads)."""
    rng = np.random.default_rng(n * 31 + C)
    coeffs = [rng.integers(0, P, 1 << n, dtype=np.uint64).astype(np.uint32) for _ in range(C)]
    if C > 2:
        coeffs[2][:] = P - 1
    evs = [evals_of(c, n) for c in coeffs]
    bufs = [ctx.alloc((1 << n) + offset) for _ in evs]
    for b_, e in zip(bufs, evs):
        b_.upload(e, offset=offset)
    d = [b_.at(offset) for b_ in bufs]
    w = [ctx.alloc(1 << n) for _ in range(4)]
    for k in (5, 0xBEEF):
        pt = secure_point(k)
        ctx.barycentric_weights(n, pt, w)
        got = ctx.barycentric_eval_at_point(d, n, w)
        for c, g in zip(coeffs, got):
            assert tuple(int(v) for v in g) == O.eval_at_point(c, pt)
    for x in bufs + w:
        x.free()


@pytest.mark.parametrize("n", [3, 4, 10, 14])
def test_eval_at_point_by_folding(ctx, n):
    rng = np.random.default_rng(7 + n)
    coeffs = rng.integers(0, P, 1 << n, dtype=np.uint64).astype(np.uint32)
    ev = evals_of(coeffs, n)
    tw = ctx.precompute_twiddles(sh.half_odds(n - 1))
    d = ctx.to_device(ev)
    for pt in (secure_point(3 * n), pt_m31(348)):
        got = ctx.eval_at_point_by_folding(d, n, pt, tw)
        assert got == O.eval_at_point(coeffs, pt)
        if n <= 10:
            assert got == PO.eval_at_point_by_folding(ev, n, pt)
    d.free()
    tw.destroy()


def test_eval_at_point_by_folding_reference_kat(ctx):
    """test_cpu_eval_at_point_by_folding: coefficients [691, 805673, 5,
    435684, 4832, 23876431, 197, 897346068] on CanonicCoset(10), twiddles of
    CanonicCoset(11)'s half coset, at get_point(348, 9736524, 13, 346752)."""
    coeffs = np.array([691, 805673, 5, 435684, 4832, 23876431, 197, 897346068], dtype=np.uint32)
    n = 10
    full = np.zeros(1 << n, dtype=np.uint32)
    full[:8] = coeffs
    d = ctx.to_device(evals_of(full, n))
    tw = ctx.precompute_twiddles(sh.half_odds(10))
    for idx in (348, 9736524, 13, 346752):
        pt = pt_m31(idx)
        assert ctx.eval_at_point_by_folding(d, n, pt, tw) == O.eval_at_point(coeffs, pt)
    d.free()
    tw.destroy()


@pytest.mark.parametrize("L", [1, 5, 17])
def test_split_at_mid(ctx, L):
    rng = np.random.default_rng(L)
    coeffs = rng.integers(0, P, 1 << L, dtype=np.uint64).astype(np.uint32)
    d = ctx.to_device(coeffs)
    left, right = ctx.alloc(1 << (L - 1)), ctx.alloc(1 << (L - 1))
    ctx.split_at_mid(d, L, left, right)
    wl, wr = PO.split_at_mid(coeffs)
    assert np.array_equal(left.download(), wl) and np.array_equal(right.download(), wr)
    # in place: left = the input's first half
    ctx.split_at_mid(d, L, d, right)
    assert np.array_equal(right.download(), wr) and np.array_equal(d.download()[: 1 << (L - 1)], wl)
    for x in (d, left, right):
        x.free()


@pytest.mark.parametrize("n", [1, 5, 11, 21])
def test_fri_decompose(ctx, n):
    rng = np.random.default_rng(50 + n)
    ev4 = [rng.integers(0, P, 1 << n, dtype=np.uint64).astype(np.uint32) for _ in range(4)]
    ev4[1][:] = P - 1
    d = [ctx.to_device(e) for e in ev4]
    g = [ctx.alloc(1 << n) for _ in range(4)]
    lam = ctx.fri_decompose(d, n, g)
    wg, wl = PO.decompose(ev4)
    assert lam == wl
    for x, y in zip(g, wg):
        assert np.array_equal(x.download(), y)
    # in place (g = eval)
    lam2 = ctx.fri_decompose(d, n, d)
    assert lam2 == wl
    for x, y in zip(d, wg):
        assert np.array_equal(x.download(), y)
    for x in d + g:
        x.free()

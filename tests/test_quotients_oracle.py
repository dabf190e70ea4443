"""Pins the FRI-quotient restatement (or_compute_fri_quotients in
oracle/stwo_oracle.c) before it checks the library.  CPU only.

What the mathematics fixes: a quotient built from correct samples of
low-degree columns is itself of low degree (the reference's
test_quotients_are_low_degree, STWO/prover/pcs/quotient_ops.rs:250-287),
and one built from a wrong sample value is not.  Lifted columns (smaller
than the lifting domain) and the periodicity sample that a column with two
samples receives (STWO/core/pcs/quotients.rs:289-336) are covered the same
way: the periodicity sample's value is only right if its point is the
second sample point shifted by the right period generator.

The reference test draws its points and coefficients from SmallRng, which is
not reproducible here; fixed multiples of SECURE_FIELD_CIRCLE_GEN
(STWO/core/circle.rs:211-214) stand in for them.
"""
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "oracle"))
sys.path.insert(0, os.path.join(HERE, "golden"))
import gen_golden as g  # noqa: E402
import oracle as O  # noqa: E402

P = O.P
U = np.uint32
GX = (1, 0, 478637715, 513582971)
GY = (992285211, 649143431, 740191619, 1186584352)


def spt_add(p, q):
    return (g.qm_add(g.qm_mul(p[0], q[0]), g.qm_scale(g.qm_mul(p[1], q[1]), P - 1)),
            g.qm_add(g.qm_mul(p[0], q[1]), g.qm_mul(p[1], q[0])))


def secure_point(k):
    """SECURE_FIELD_CIRCLE_GEN.mul(k)."""
    res, cur = ((1, 0, 0, 0), (0, 0, 0, 0)), (GX, GY)
    while k:
        if k & 1:
            res = spt_add(res, cur)
        cur = spt_add(cur, cur)
        k >>= 1
    return res


def qdouble_x(x):
    """CirclePoint::double_x: 2x^2 - 1 (STWO/core/circle.rs)."""
    return g.qm_add(g.qm_scale(g.qm_mul(x, x), 2), (P - 1, 0, 0, 0))


def qfold(values, factors):
    """fold (STWO/core/poly/utils.rs:27-38) of M31 values with QM31 factors."""
    if len(values) == 1:
        return (int(values[0]) % P, 0, 0, 0)
    h = len(values) // 2
    return g.qm_add(qfold(values[:h], factors[1:]), g.qm_mul(qfold(values[h:], factors[1:]), factors[0]))


def eval_at(coeffs, pt):
    """CpuBackend::eval_at_point at a secure point (CPUB/circle.rs:82-99):
    factors y, x, pi(x), ..., applied from the last."""
    n = len(coeffs).bit_length() - 1
    if n == 0:
        return (int(coeffs[0]), 0, 0, 0)
    x, y = pt
    maps = [y]
    for _ in range(1, n):
        maps.append(x)
        x = qdouble_x(x)
    maps.reverse()
    return qfold([int(c) for c in coeffs], maps)


def domain_eval(coeffs, log_size):
    init = O.half_odds_initial(log_size - 1)
    tw, _ = O.precompute_twiddles(init, log_size - 1)
    padded = np.zeros(1 << log_size, U)
    padded[: len(coeffs)] = coeffs
    return O.evaluate(padded, init, log_size, tw)


def interp(values, log_size):
    init = O.half_odds_initial(log_size - 1)
    _, itw = O.precompute_twiddles(init, log_size - 1)
    return O.interpolate(values, init, log_size, itw)


def lifted_poly(col, L):
    """Coefficients (over CanonicCoset(L)) of column `col` lifted to log size L."""
    k = len(col).bit_length() - 1
    s = L - k
    lifted = np.array([col[((i >> (s + 1)) << 1) + (i & 1)] for i in range(1 << L)], dtype=U)
    return interp(lifted, L)


def high_coeffs_zero(quot, L, bound_log):
    return all(not interp(q, L)[(1 << bound_log) - 1:].any() for q in quot) if bound_log else True


ALPHA = (123456789, 987654321, 55555, 777)


def test_quotients_are_low_degree_reference_shape():
    """LOG_SIZE 3, blowup 4, two samples of the polynomial 0..7."""
    L = 7
    coeffs = np.arange(8, dtype=U)
    ev = domain_eval(coeffs, L)
    pts = [secure_point(k) for k in (0x1234567, 0xABCDEF01)]
    samples = [[(p[0], p[1], eval_at(coeffs, p)) for p in pts]]
    quot = O.compute_fri_quotients([ev], samples, ALPHA, L)
    for q in quot:
        assert not interp(q, L)[(1 << 3) - 1:].any()
    bad = [[(pts[0][0], pts[0][1], samples[0][0][2]), (pts[1][0], pts[1][1], g.qm_add(samples[0][1][2], (1, 0, 0, 0)))]]
    quot = O.compute_fri_quotients([ev], bad, ALPHA, L)
    assert any(interp(q, L)[(1 << 3) - 1:].any() for q in quot)


@pytest.mark.parametrize("seed", range(3))
def test_lifted_columns_and_periodicity_samples(seed):
    """Columns of log 4..8 lifted to 8 (blowup 2), one, two or three
    samples each; sample values are the lifted polynomials' values."""
    rng = np.random.default_rng(seed)
    L, b = 8, 2
    logs = [8, 5, 6, 8, 4, 7]
    cols, samples = [], []
    z = [secure_point(int(k)) for k in rng.integers(1, 1 << 62, 3)]
    for i, lg in enumerate(logs):
        c = rng.integers(0, P, 1 << (lg - b), dtype=np.uint64).astype(U)
        ev = domain_eval(c, lg)
        cols.append(ev)
        lp = lifted_poly(ev, L)
        n = 1 + (i % 3)
        samples.append([(p[0], p[1], eval_at(lp, p)) for p in z[:n]])
    quot = O.compute_fri_quotients(cols, samples, ALPHA, L)
    for q in quot:
        assert not interp(q, L)[1 << (L - b):].any()
    # a wrong value in a lifted column's second sample breaks it
    bad = [list(s) for s in samples]
    x, y, v = bad[1][1]
    bad[1][1] = (x, y, g.qm_add(v, (0, 0, 1, 0)))
    quot = O.compute_fri_quotients(cols, bad, ALPHA, L)
    assert any(interp(q, L)[1 << (L - b):].any() for q in quot)


def test_periodicity_point_matters():
    """A column of log 6 lifted to 8 with two samples: replacing the second
    sample's value by the first's keeps only the periodicity sample wrong
    if the period generator were ignored -- checked through a column whose
    lifted polynomial differs at z2 and z2 + g only if g is wrong."""
    rng = np.random.default_rng(7)
    L, b, lg = 8, 2, 6
    c = rng.integers(0, P, 1 << (lg - b), dtype=np.uint64).astype(U)
    ev = domain_eval(c, lg)
    lp = lifted_poly(ev, L)
    z1, z2 = secure_point(99991), secure_point(31337)
    # the lifted polynomial is periodic with period g = CanonicCoset(L).step()^(2^lg)
    gidx = (1 << (31 - L + lg)) % (1 << 31)
    gp = O.index_to_point(gidx)
    z2g = spt_add(z2, ((gp[0], 0, 0, 0), (gp[1], 0, 0, 0)))
    assert eval_at(lp, z2) == eval_at(lp, z2g)
    quot = O.compute_fri_quotients([ev], [[(z1[0], z1[1], eval_at(lp, z1)), (z2[0], z2[1], eval_at(lp, z2))]],
                                   ALPHA, L)
    for q in quot:
        assert not interp(q, L)[1 << (L - b):].any()


# ---------------------------------------------------------- eval_at_point ---

def test_eval_at_point_reference_kat():
    """test_eval_at_point_with_4_coeffs (CPUB/circle.rs tests): coefficients
    [1, 3, 2, 4] give c0 + c1 y + c2 x + c3 x y."""
    x, y = O.index_to_point(123456)
    got = O.eval_at_point(np.array([1, 3, 2, 4], U), ((x, 0, 0, 0), (y, 0, 0, 0)))
    want = (1 + 3 * y + 2 * x + 4 * x * y) % P
    assert got == (want, 0, 0, 0)


@pytest.mark.parametrize("n", [0, 1, 2, 5, 9])
def test_eval_at_point_matches_fft_on_the_domain(n):
    """At every point of CanonicCoset(n) the value equals the FFT evaluation
    (which brute force pins) at that point's bit-reversed position."""
    rng = np.random.default_rng(n)
    coeffs = rng.integers(0, P, 1 << n, dtype=np.uint64).astype(U)
    if n == 0:
        assert O.eval_at_point(coeffs, ((5, 0, 0, 0), (7, 0, 0, 0))) == (int(coeffs[0]), 0, 0, 0)
        return
    ev = domain_eval(coeffs, n)
    half_init = O.half_odds_initial(n - 1)
    for j in range(1 << n):
        i = int(format(j, "0%db" % n)[::-1], 2) if n else 0
        h = 1 << (n - 1)
        step = (1 << (31 - (n - 1))) % (1 << 31)
        idx = (half_init + step * (i % h)) % (1 << 31)
        if i >= h:
            idx = (-idx) % (1 << 31)
        x, y = O.index_to_point(idx)
        assert O.eval_at_point(coeffs, ((x, 0, 0, 0), (y, 0, 0, 0))) == (int(ev[j]), 0, 0, 0), j


def test_eval_at_point_secure_points_match_fold():
    rng = np.random.default_rng(3)
    for n in (1, 4, 7):
        coeffs = rng.integers(0, P, 1 << n, dtype=np.uint64).astype(U)
        pt = secure_point(int(rng.integers(1, 1 << 60)))
        assert O.eval_at_point(coeffs, pt) == eval_at(coeffs, pt)

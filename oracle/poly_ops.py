"""ORACLE (test infrastructure only; only tests/ load it): plain-integer
restatements of the PolyOps / FriOps methods beside the FFT --
barycentric_weights, barycentric_eval_at_point, eval_at_point_by_folding,
split_at_mid and FriOps::decompose -- for checking the library's kernels.

Sources (vendor/stwo-upstream-dev-62b228e/crates/stwo/src):
  prover/poly/circle/ops.rs:67-88,163 (the trait), CpuBackend
  prover/backend/cpu/circle.rs:101-180,283-291, prover/backend/cpu/fri.rs:41-63,141-156,
  core/constraints.rs:12-93 (coset_vanishing, coset_vanishing_derivative,
  point_vanishing), core/poly/utils.rs:45-70 (get_folding_alphas).
Pinned by tests/test_poly_ops_oracle.py against the C oracle's
eval_at_point (itself pinned by brute force and the reference KATs)."""
import numpy as np

import oracle as O
from prove import ONE, ZERO, double_x, pt_add, q_add, q_inv, q_m31, q_mul, q_sub

P = O.P
IDX_MASK = (1 << 31) - 1


def domain_index(log_size, k):
    """Circle-group index of CanonicCoset(log_size).circle_domain().at(k)
    (CircleDomain::at, STWO/core/poly/circle/domain.rs:62-68): the half coset
    half_odds(log_size - 1) for k < 2^(log_size-1), its negation after."""
    half = 1 << (log_size - 1)
    init = O.half_odds_initial(log_size - 1)
    step = 1 << (31 - (log_size - 1))
    if k < half:
        return (init + k * step) & IDX_MASK
    return (-(init + (k - half) * step)) & IDX_MASK


def m31_point(idx):
    x, y = O.index_to_point(idx & IDX_MASK)
    return int(x), int(y)


def barycentric_weights(log_size, p):
    """CpuBackend::barycentric_weights(CanonicCoset(log_size), p): for each
    bit-reversed domain point q_i, w_i = vn(p) / (s_i * v_i(p)) with
    s_i = -2 q_i.y * coset_vanishing_derivative(q_i) (the derivative of the
    domain's vanishing polynomial: 4^(n-1) * prod_{k=1}^{n-1} x-doubling^(k-1)(q.x)),
    v_i(p) = point_vanishing(q_i, p) = h.y / (1 + h.x) with h = p - q_i, and
    vn(p) = coset_vanishing(CanonicCoset(n).coset, p) = the x coordinate of
    p doubled n - 1 times (the odds coset's shift is zero)."""
    n = log_size
    px, py = p
    x = px
    for _ in range(1, n):
        x = double_x(x)
    vn = x
    four_pow = pow(4, n - 1, P)
    out = []
    for i in range(1 << n):
        k = _bitrev(i, n)
        qx, qy = m31_point(domain_index(n, k))
        # coset_vanishing_derivative over the base field
        prod = 1
        xx = qx
        for _ in range(1, n):
            prod = prod * xx % P
            xx = (2 * xx * xx - 1) % P
        s = (-2 * qy) % P * four_pow % P * prod % P
        # h = p - q = p + (q.x, -q.y)
        hx = q_add(q_mul(px, q_m31(qx)), q_mul(py, q_m31(qy)))
        hy = q_sub(q_mul(py, q_m31(qx)), q_mul(px, q_m31(qy)))
        v = q_mul(hy, q_inv(q_add(ONE, hx)))
        out.append(q_mul(vn, q_inv(q_mul(q_m31(s), v))))
    return out


def _bitrev(i, n):
    r = 0
    for _ in range(n):
        r = (r << 1) | (i & 1)
        i >>= 1
    return r


def barycentric_eval_at_point(evals, weights):
    """sum_i evals[i] * weights[i] (CpuBackend::barycentric_eval_at_point)."""
    acc = [0, 0, 0, 0]
    for e, w in zip(evals, weights):
        for k in range(4):
            acc[k] = (acc[k] + int(e) * w[k]) % P
    return tuple(acc)


def folding_alphas(point, n):
    """get_folding_alphas(point, n): [pi^(n-2)(x), ..., pi(x), x, y] (the last
    is y; each earlier one doubles x once more)."""
    px, py = point
    alphas = [None] * n
    alphas[n - 1] = py
    x = px
    for i in range(n - 2, -1, -1):
        alphas[i] = x
        x = double_x(x)
    return alphas


def eval_at_point_by_folding(evals, log_size, point):
    """CpuBackend::eval_at_point_by_folding: fold the evaluation (as secure
    values) circle -> line with the last folding alpha, then fold_line down to
    one value with the others, divided by 2^log_size.  Uses the C oracle's
    folds (pinned separately)."""
    n = log_size
    alphas = folding_alphas(point, n)
    src = [np.ascontiguousarray(evals, dtype=np.uint32)] + [np.zeros(1 << n, np.uint32) for _ in range(3)]
    dst = [np.zeros(1 << (n - 1), np.uint32) for _ in range(4)]
    O.fold_circle_into_line(dst, src, O.half_odds_initial(n - 1), n, list(alphas.pop()))
    cur, lg = dst, n - 1
    line_init = O.half_odds_initial(n - 1)
    while lg > 0:
        cur = O.fold_line(cur, line_init, lg, list(alphas.pop()), 1)
        line_init = (2 * line_init) & IDX_MASK
        lg -= 1
    v = tuple(int(c[0]) for c in cur)
    inv = pow(pow(2, n, P), P - 2, P)
    return tuple(x * inv % P for x in v)


def split_at_mid(coeffs):
    """CpuBackend::split_at_mid: the coefficient vector's two halves."""
    h = len(coeffs) // 2
    return coeffs[:h].copy(), coeffs[h:].copy()


def decompose(eval4):
    """CpuBackend::decompose (prover/backend/cpu/fri.rs:41-63,141-156):
    lambda = (sum of the first half - sum of the second half) / N over the
    bit-reversed secure evaluation; g = f - lambda on the first half, f +
    lambda on the second.  eval4: 4 coordinate arrays; returns (g4, lambda)."""
    N = len(eval4[0])
    half = N // 2
    a = [int(np.sum(c[:half].astype(np.uint64)) % P) for c in eval4]
    b = [int(np.sum(c[half:].astype(np.uint64)) % P) for c in eval4]
    ninv = pow(N % P, P - 2, P)
    lam = tuple((x - y) * ninv % P for x, y in zip(a, b))
    g = []
    for k in range(4):
        c = eval4[k].astype(np.int64)
        r = np.empty(N, dtype=np.int64)
        r[:half] = (c[:half] - lam[k]) % P
        r[half:] = (c[half:] + lam[k]) % P
        g.append(r.astype(np.uint32))
    return g, lam

"""Pins the end-to-end prove restatement (oracle/prove.py).  CPU only.

The oracle prover chains C-oracle pieces that have their own pins; what this
file checks is the glue and the verifier side, from outside them:

* the verifier's plain-integer field against the reference's QM31 KAT
  (STWO/core/fields/qm31.rs:150-171) and its inverse round trip;
* the OODS point is on the circle (x^2 + y^2 = 1);
* the vanishing polynomial of CanonicCoset(n) is zero on that coset and not
  off it (STWO/core/constraints.rs:12-36);
* the split-at-mid recombination (extract_composition_oods_eval) equals the
  full composition polynomial evaluated at the point;
* the verifier's fri_answers (restated from STWO/core/pcs/quotients.rs) equal
  the C oracle's compute_fri_quotients column at the query positions;
* the proof-of-work nonce is the smallest valid one;
* proofs verify, the byte layout round-trips, and a changed sampled value,
  queried value, hash witness, FRI witness, last-layer coefficient, nonce or
  commitment is rejected; a trace that breaks the AIR is refused by the prover
  (the reference's test_wide_fibonacci_constraints_fails, examples
  wide_fibonacci/mod.rs:159-177).

Proof bytes against the Rust SimdBackend are parity unpinned: no Rust
toolchain is present and the reference ships no serialized proof.
"""
import copy
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "oracle"))
import decommit as D  # noqa: E402
import oracle as O  # noqa: E402
import prove as PR  # noqa: E402

P = O.P


def ref_inputs(log_n):
    """generate_test_inputs of the reference's test and benchmark: a = 1, b = i."""
    return np.ones(1 << log_n, dtype=np.uint32), np.arange(1 << log_n, dtype=np.uint32)


def rand_inputs(log_n, seed):
    return O.fill_random(1 << log_n, seed, 0), O.fill_random(1 << log_n, seed, 1)


def test_field_kat():
    a, b = (1, 2, 3, 4), (4, 5, 6, 7)
    assert PR.q_mul(a, b) == (P - 71, 93, P - 16, 50)
    assert PR.q_mul(a, PR.q_inv(a)) == PR.ONE
    assert PR.q_mul(PR.q_mul(a, b), PR.q_inv(b)) == a
    rng = np.random.default_rng(1)
    for _ in range(20):
        x = tuple(int(v) for v in rng.integers(0, P, 4))
        y = tuple(int(v) for v in rng.integers(0, P, 4))
        assert list(PR.q_mul(x, y)) == O.qm31_mul(list(x), list(y))


def test_random_point_on_circle():
    ch = O.Channel()
    for _ in range(5):
        x, y = PR.random_point(ch)
        assert PR.q_add(PR.q_mul(x, x), PR.q_mul(y, y)) == PR.ONE


@pytest.mark.parametrize("n", [1, 3, 5])
def test_coset_vanishing(n):
    # CanonicCoset(n) = odds coset: initial 2^(30 - n), step 2^(31 - n)
    for i in range(1 << n):
        idx = (1 << (30 - n)) + i * (1 << (31 - n))
        assert PR.coset_vanishing(n, PR.pt_m31(idx)) == PR.ZERO
    for idx in (0, 1, 12345, 1 << (31 - n)):
        assert PR.coset_vanishing(n, PR.pt_m31(idx)) != PR.ZERO


def test_split_recombination_is_full_evaluation():
    """left(p) + pi^(L-2)(p.x) right(p) = full(p) for a polynomial of log size L."""
    L = 6
    rng = np.random.default_rng(3)
    coeffs = [rng.integers(0, P, 1 << L).astype(np.uint32) for _ in range(4)]
    ch = O.Channel()
    ch.mix_u64(99)
    pt = PR.random_point(ch)
    pnt = (list(pt[0]), list(pt[1]))
    n = 1 << (L - 1)
    full = [O.eval_at_point(c, pnt) for c in coeffs]
    halves = [O.eval_at_point(c[:n], pnt) for c in coeffs] + [O.eval_at_point(c[n:], pnt) for c in coeffs]
    got = PR.extract_composition_oods_eval([[], [[h] for h in halves]], pt, L - 1)
    assert got == PR.from_partial_evals(full)


@pytest.mark.parametrize("sizes", [[6, 6, 6], [4, 5, 6]])
def test_fri_answers_match_quotient_column(sizes):
    """Verifier-side answers at query positions = the prover's quotient column there."""
    L = max(sizes)
    rng = np.random.default_rng(7)
    cols = []
    for s in sizes:
        coef = np.zeros(1 << s, dtype=np.uint32)
        coef[: 1 << (s - 1)] = rng.integers(0, P, 1 << (s - 1))
        tw, _ = O.precompute_twiddles(O.half_odds_initial(s - 1), s - 1)
        cols.append(O.evaluate(coef, O.half_odds_initial(s - 1), s, tw))
    ch = O.Channel()
    p1, p2 = PR.random_point(ch), PR.random_point(ch)
    per_col = [[p1], [p1, p2], [p2]]
    samples, flat = [], []
    for c, pts in zip(cols, per_col):
        ss = []
        for pt in pts:
            v = tuple(int(x) for x in rng.integers(0, P, 4))
            ss.append((pt, v))
        samples.append(ss)
        flat.append([(list(pt[0]), list(pt[1]), list(v)) for pt, v in ss])
    alpha = (5, 6, 7, 8)
    quot = O.compute_fri_quotients(cols, flat, list(alpha), L)
    queries = [0, 3, 17, 40, (1 << L) - 1]
    qv = [[int(c[D.lift(p, L - s)]) for p in queries] for c, s in zip(cols, sizes)]
    ans = PR.fri_answers([sizes], [samples], alpha, queries, [qv], L)
    for a, p in zip(ans, queries):
        assert a == tuple(int(quot[k][p]) for k in range(4))


def test_pow_nonce_is_smallest():
    d = bytes(range(32))
    nonce = O.grind(d, 8)
    assert PR.pow_ok(d, 8, nonce)
    assert not any(PR.pow_ok(d, 8, k) for k in range(nonce))


CASES = [
    (4, 8, PR.PcsConfig()),
    (5, 100, PR.PcsConfig()),
    (6, 3, PR.PcsConfig(pow_bits=4, log_blowup_factor=2, log_last_layer_degree_bound=2, n_queries=10,
                        line_fold_step=2)),
]


@pytest.mark.parametrize("case", range(len(CASES)))
def test_prove_verifies_and_round_trips(case):
    log_n, n_cols, cfg = CASES[case]
    a, b = ref_inputs(log_n) if case != 2 else rand_inputs(log_n, 11)
    proof = PR.wide_fib_prove(a, b, n_cols, log_n, cfg)
    data = PR.proof_bytes(proof)
    back = PR.parse_proof(data)
    assert PR.proof_bytes(back) == data
    PR.verify(back, n_cols, log_n)
    # Fields the layout fixes: 3 trees, n_cols + 8 samples, queried values per query.
    assert len(back["commitments"]) == 3 and len(back["sampled_values"][1]) == n_cols
    assert all(len(c) == len(proof["query_positions"]) for c in back["queried_values"][1])


def _tampered(proof, path, fn):
    p = copy.deepcopy(proof)
    obj = p
    for k in path[:-1]:
        obj = obj[k]
    obj[path[-1]] = fn(obj[path[-1]])
    return p


def _bump(t):
    return tuple([(t[0] + 1) % P] + list(t[1:]))


def test_tampered_proofs_rejected():
    log_n, n_cols = 5, 12
    a, b = rand_inputs(log_n, 4)
    proof = PR.parse_proof(PR.proof_bytes(PR.wide_fib_prove(a, b, n_cols, log_n)))
    PR.verify(proof, n_cols, log_n)
    bad_nonce = next(k for k in range(1000) if k != proof["proof_of_work"])
    cases = [
        (["sampled_values", 1, 3, 0], _bump),
        (["sampled_values", 2, 5, 0], _bump),
        (["queried_values", 1, 0, 0], lambda v: (v + 1) % P),
        (["queried_values", 2, 7, 1], lambda v: (v + 1) % P),
        (["decommitments", 1, 0], _bump),
        (["fri", "first", "fri_witness", 0], _bump),
        (["fri", "first", "hash_witness", 0], _bump),
        (["fri", "inner", 0, "hash_witness", 0], _bump),
        (["fri", "last_layer_poly", 0], _bump),
        (["commitments", 1], _bump),
        (["commitments", 2], _bump),
        (["proof_of_work"], lambda v: bad_nonce),
    ]
    for path, fn in cases:
        bad = _tampered(proof, path, fn)
        with pytest.raises(PR.VerificationError):
            PR.verify(bad, n_cols, log_n)
    # a dropped hash witness
    bad = copy.deepcopy(proof)
    bad["decommitments"][2] = bad["decommitments"][2][1:]
    with pytest.raises(PR.VerificationError):
        PR.verify(bad, n_cols, log_n)
    # the wrong AIR width
    with pytest.raises(PR.VerificationError):
        PR.verify(proof, n_cols + 1, log_n)


def test_prover_refuses_broken_trace(monkeypatch):
    """A trace that breaks one constraint (the reference changes one cell,
    mod.rs:159-177) fails prove_ex's self-check."""
    log_n, n_cols = 4, 20
    a, b = ref_inputs(log_n)
    real = O.wide_fib_trace

    def broken(a_, b_, n):
        cols = real(a_, b_, n)
        cols[17][2] = 1
        return cols

    monkeypatch.setattr(O, "wide_fib_trace", broken)
    with pytest.raises(ValueError, match="ConstraintsNotSatisfied"):
        PR.wide_fib_prove(a, b, n_cols, log_n)

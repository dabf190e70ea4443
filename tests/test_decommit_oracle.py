"""Pins the decommitment restatement (oracle/decommit.py) before it checks the
library.  CPU only.

* The reference's own known answers for lifted decommitment
  (STWO/prover/vcs_lifted/prover.rs:232-303: queried values at positions 0,
  4 and 15, and the aux node values of a decommitment at position 1).
* The restated verifiers (STWO/core/vcs_lifted/verifier.rs and
  STWO/core/fri.rs) accept what the restated provers emit, and reject each
  planted mistake: a changed value, a changed, missing or extra witness
  hash, a changed FRI witness value, a changed last-layer coefficient.  The
  verifiers are restated from different reference code than the provers, so
  a mistake shared by a prover and its checker would have to be made twice.

Reference paths: STWO = vendor/stwo-upstream-dev-62b228e/crates/stwo/src.
"""
import os
import random
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "..", "oracle"))
import decommit as D  # noqa: E402
import oracle as O  # noqa: E402

U = np.uint32


def kat_tree():
    """prepare_merkle(): columns 0..4, 0..8, 0..16 committed at lifting log 4."""
    cols = [np.arange(1 << i, dtype=U) for i in range(2, 5)]
    return cols, O.lifted_commit(cols, 4)


@pytest.mark.parametrize("pos,want", [(0, [0, 0, 0]), (4, [0, 2, 4]), (15, [3, 7, 15])])
def test_lifted_decommitted_values_kat(pos, want):
    """test_lifted_decommitted_values (vcs_lifted/prover.rs:278-303)."""
    cols, tree = kat_tree()
    values, _, _ = D.merkle_lifted_decommit(tree, 4, cols, [pos])
    assert values == [[w] for w in want]


def test_decommitment_aux_kat():
    """test_decommitment_aux (vcs_lifted/prover.rs:338-357): decommitting
    position 1 visits nodes 0 and 1 of every layer below the root."""
    cols, tree = kat_tree()
    _, witness, aux = D.merkle_lifted_decommit(tree, 4, cols, [1])
    expected = [{0: D.tree_node(tree, l, 0), 1: D.tree_node(tree, l, 1)} for l in range(4, 0, -1)]
    assert aux == expected
    # Position 1's sibling 0 is known to the verifier only as a witness; above
    # the leaves each parent 0 needs its sibling 1.
    assert witness == [D.tree_node(tree, 4, 0)] + [D.tree_node(tree, l, 1) for l in range(3, 0, -1)]


def random_tree(seed, L, logs):
    rng = np.random.default_rng(seed)
    cols = [rng.integers(0, O.P, 1 << lg, dtype=np.uint64).astype(U) for lg in logs]
    return cols, O.lifted_commit(cols, L)


@pytest.mark.parametrize("seed", range(4))
def test_merkle_decommit_verifies(seed):
    rnd = random.Random(seed)
    L = 9
    logs = [rnd.randrange(1, L + 1) for _ in range(7)] + [L]
    cols, tree = random_tree(seed, L, logs)
    queries = sorted(rnd.randrange(0, 1 << L) for _ in range(20))
    queries += [queries[3], queries[7]]  # duplicates are allowed
    queries.sort()
    values, witness, _ = D.merkle_lifted_decommit(tree, L, cols, queries)
    root = D.tree_node(tree, 0, 0)
    D.merkle_lifted_verify(root, logs, queries, values, witness)

    bad = [list(v) for v in values]
    q = next(i for i in range(len(queries)) if queries.count(queries[i]) == 1)
    bad[2][q] = (bad[2][q] + 1) % O.P
    with pytest.raises(D.VerificationError, match="RootMismatch"):
        D.merkle_lifted_verify(root, logs, queries, bad, witness)
    w2 = list(witness)
    w2[len(w2) // 2] = tuple((x ^ 1) for x in w2[len(w2) // 2])
    with pytest.raises(D.VerificationError, match="RootMismatch"):
        D.merkle_lifted_verify(root, logs, queries, values, w2)
    with pytest.raises(D.VerificationError, match="WitnessTooShort"):
        D.merkle_lifted_verify(root, logs, queries, values, witness[:-1])
    with pytest.raises(D.VerificationError, match="WitnessTooLong"):
        D.merkle_lifted_verify(root, logs, queries, values, witness + [witness[0]])


def test_queries_fold_matches_domain_doubling():
    """test_folded_queries (core/queries.rs:104-139): a query of the domain of
    log 7 folded twice is the position, in the domain of log 5, of the point
    doubled twice (up to conjugation)."""
    n, m = 7, 5
    for q in range(1 << n):
        idx = D.circle_domain_index_at(O.half_odds_initial(n - 1), n - 1, D.bit_reverse_index(q, n))
        f = D.queries_fold([q], n - m)[0]
        fidx = D.circle_domain_index_at(O.half_odds_initial(m - 1), m - 1, D.bit_reverse_index(f, m))
        # the doubled point and the folded position's point share x (folding
        # may give the conjugate point)
        assert D.point((idx << (n - m)) & D.IDX_MASK)[0] == D.point(fidx)[0]


def low_degree_column(log_size, log_blowup, seed):
    """A QM31 evaluation on CanonicCoset(log_size) of a circle polynomial of
    degree < 2^(log_size - log_blowup), as 4 coordinate columns."""
    n = log_size - log_blowup
    rng = np.random.default_rng(seed)
    coeffs = [rng.integers(0, O.P, 1 << n, dtype=np.uint64).astype(U) for _ in range(4)]
    tw, _ = O.precompute_twiddles(O.half_odds_initial(log_size - 1), log_size - 1)
    return O.evaluate_polynomials(coeffs, log_blowup, tw)


def fri_case(log_size, cfg, seed):
    col = low_degree_column(log_size, cfg.log_blowup_factor, seed)
    ch = O.Channel()
    ch.mix_u64(0x1234 + seed)
    start = O.Channel.from_buffer_copy(ch)
    layers, roots, last = D.fri_layers(cfg, col, log_size, ch)
    queries, raw, proofs = D.fri_decommit(ch, cfg, layers)
    return col, start, layers, roots, last, queries, raw, proofs, ch


FRI_CASES = [(8, (1, 0, 5, 1)), (10, (2, 1, 8, 2)), (9, (1, 2, 12, 3)), (7, (1, 1, 3, 1)), (6, (1, 4, 4, 1))]


@pytest.mark.parametrize("log_size,cfgt", FRI_CASES)
def test_fri_decommit_verifies(log_size, cfgt):
    cfg = O.FriConfig(*cfgt)
    col, start, layers, roots, last, queries, raw, proofs, ch = fri_case(log_size, cfg, log_size)
    assert len(proofs) == len(D.fri_schedule(cfg, log_size))
    evals_of = lambda qs: [tuple(int(col[k][q]) for k in range(4)) for q in qs]  # noqa: E731
    vch = O.Channel.from_buffer_copy(start)
    D.fri_verify(vch, cfg, (proofs, [tuple(int(x) for x in c) for c in last]), log_size, evals_of)
    # the verifier's transcript ends where the prover's did
    assert bytes(vch.digest) == bytes(ch.digest) and vch.n_draws == ch.n_draws


@pytest.mark.parametrize("log_size,cfgt", FRI_CASES[:3])
def test_fri_verifier_rejects_planted_mistakes(log_size, cfgt):
    cfg = O.FriConfig(*cfgt)
    col, start, layers, roots, last, queries, raw, proofs, ch = fri_case(log_size, cfg, 100 + log_size)
    evals_of = lambda qs: [tuple(int(col[k][q]) for k in range(4)) for q in qs]  # noqa: E731
    poly = [tuple(int(x) for x in c) for c in last]

    def verify(pr, lp=poly):
        D.fri_verify(O.Channel.from_buffer_copy(start), cfg, (pr, lp), log_size, evals_of)

    verify(proofs)
    for k in range(len(proofs)):
        if proofs[k]["fri_witness"]:
            bad = [dict(p) for p in proofs]
            w = list(bad[k]["fri_witness"])
            w[0] = ((w[0][0] + 1) % O.P,) + tuple(w[0][1:])
            bad[k]["fri_witness"] = w
            with pytest.raises(D.VerificationError):
                verify(bad)
        if proofs[k]["hash_witness"]:
            bad = [dict(p) for p in proofs]
            bad[k]["hash_witness"] = bad[k]["hash_witness"][1:]
            with pytest.raises(D.VerificationError):
                verify(bad)
    bad_poly = list(poly)
    bad_poly[0] = ((bad_poly[0][0] + 1) % O.P,) + tuple(bad_poly[0][1:])
    with pytest.raises(D.VerificationError):
        verify(proofs, bad_poly)

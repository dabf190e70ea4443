"""GPU parity of the decommitment entry points against the oracle
(oracle/decommit.py), plus the restated reference verifiers run on the
library's output.

* stwo_hip_merkle_lifted_decommit vs MerkleProverLifted::decommit
  (STWO/prover/vcs_lifted/prover.rs:58-201): queried values, hash witness
  and aux node values, bit-exact; the reference's known answers
  (prover.rs:278-357); the lifted verifier accepts the library's witness.
* stwo_hip_fri_prover_{commit,decommit,layer_proof} vs FriProver::commit +
  FriProver::decommit (STWO/prover/fri.rs:133-489): roots, last layer,
  channel transcript, query positions and every layer's proof, bit-exact;
  FriVerifier (STWO/core/fri.rs) accepts the library's proof.
"""
import os
import random
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "oracle"))
import decommit as D  # noqa: E402
import oracle as O  # noqa: E402
import stwo_hip as sh  # noqa: E402

pytestmark = pytest.mark.gpu
U = np.uint32


@pytest.fixture(scope="module")
def ctx():
    c = sh.Context(0)
    yield c
    c.close()


def device_tree(ctx, cols, L):
    dcols = [ctx.to_device(c) for c in cols]
    layers = ctx.alloc(sh.tree_words(L))
    ctx.lifted_commit(dcols, [len(c).bit_length() - 1 for c in cols], L, layers)
    return dcols, layers


def as_tuples(a):
    return [tuple(int(x) for x in r) for r in a]


def test_merkle_decommit_reference_kats(ctx):
    cols = [np.arange(1 << i, dtype=U) for i in range(2, 5)]
    dcols, layers = device_tree(ctx, cols, 4)
    for pos, want in ((0, [0, 0, 0]), (4, [0, 2, 4]), (15, [3, 7, 15])):
        vals, _, _ = ctx.merkle_lifted_decommit(layers, 4, dcols, [2, 3, 4], [pos])
        assert vals[:, 0].tolist() == want, pos
    tree = layers.download()
    _, _, aux = ctx.merkle_lifted_decommit(layers, 4, dcols, [2, 3, 4], [1], with_aux=True)
    want = []
    for l in range(4, 0, -1):
        want += [(l, 0, D.tree_node(tree, l, 0)), (l, 1, D.tree_node(tree, l, 1))]
    assert aux == want


@pytest.mark.parametrize("seed,L", [(0, 6), (1, 10), (2, 16), (3, 1)])
def test_merkle_decommit_vs_oracle(ctx, seed, L):
    rnd = random.Random(seed)
    rng = np.random.default_rng(seed)
    logs = sorted(rnd.randrange(1, L + 1) for _ in range(9)) + [L]
    rnd.shuffle(logs)
    cols = [rng.integers(0, O.P, 1 << lg, dtype=np.uint64).astype(U) for lg in logs]
    dcols, layers = device_tree(ctx, cols, L)
    tree = layers.download()
    assert (tree == O.lifted_commit(cols, L)).all()
    queries = sorted(rnd.randrange(0, 1 << L) for _ in range(min(40, 1 << L)))
    queries = sorted(queries + queries[:3])  # with duplicates
    vals, wit, aux = ctx.merkle_lifted_decommit(layers, L, dcols, logs, queries, with_aux=True)
    ovals, owit, oaux = D.merkle_lifted_decommit(tree, L, cols, queries)
    assert vals.tolist() == ovals
    assert as_tuples(wit) == owit
    flat = [(l, i, h) for l, nodes in zip(range(L, 0, -1), oaux) for i, h in sorted(nodes.items())]
    assert aux == flat
    D.merkle_lifted_verify(D.tree_node(tree, 0, 0), logs, queries, vals.tolist(), as_tuples(wit))


def test_merkle_decommit_rejects_bad_queries(ctx):
    cols = [np.arange(16, dtype=U)]
    dcols, layers = device_tree(ctx, cols, 4)
    with pytest.raises(sh.StwoHipError):
        ctx.merkle_lifted_decommit(layers, 4, dcols, [4], [3, 2])  # not increasing
    with pytest.raises(sh.StwoHipError):
        ctx.merkle_lifted_decommit(layers, 4, dcols, [4], [16])  # outside 2^4
    with pytest.raises(sh.StwoHipError):
        ctx.merkle_lifted_decommit(layers, 4, dcols, [5], [1])  # column longer than the lifting size


def low_degree_column(log_size, log_blowup, seed):
    n = log_size - log_blowup
    rng = np.random.default_rng(seed)
    coeffs = [rng.integers(0, O.P, 1 << n, dtype=np.uint64).astype(U) for _ in range(4)]
    tw, _ = O.precompute_twiddles(O.half_odds_initial(log_size - 1), log_size - 1)
    return O.evaluate_polynomials(coeffs, log_blowup, tw)


FRI_CASES = [(8, (1, 0, 5, 1)), (10, (2, 1, 8, 2)), (9, (1, 2, 12, 3)), (7, (1, 1, 3, 1)), (6, (1, 4, 4, 1)),
             (14, (1, 3, 20, 1)), (16, (2, 2, 30, 4)), (13, (1, 1, 6, 5)), (20, (1, 0, 3, 1))]


@pytest.mark.parametrize("log_size,cfgt", FRI_CASES)
def test_fri_prover_decommit_vs_oracle(ctx, log_size, cfgt):
    cfg = O.FriConfig(*cfgt)
    scfg = sh.FriConfig(*cfgt)
    col = low_degree_column(log_size, cfg.log_blowup_factor, 50 + log_size)
    ch = O.Channel()
    ch.mix_u64(0xABCD + log_size)
    start = O.Channel.from_buffer_copy(ch)
    layers, roots, last = D.fri_layers(cfg, col, log_size, ch)
    after_commit = (bytes(ch.digest), ch.n_draws)
    queries, raw, proofs = D.fri_decommit(ch, cfg, layers)

    tw_root = sh.half_odds(log_size - 1)
    tw = ctx.precompute_twiddles(tw_root)
    dcol = [ctx.to_device(c) for c in col]
    pr, groots, glast, dig, nd = ctx.fri_prover_commit(bytes(start.digest), start.n_draws, scfg, dcol, log_size, tw)
    try:
        assert groots.tolist() == roots.tolist()
        assert glast.tolist() == last.tolist()
        assert (dig, nd) == after_commit
        gq, graw, dig2, nd2 = pr.decommit(dig, nd)
        assert gq.tolist() == queries and graw.tolist() == raw
        assert (dig2, nd2) == (bytes(ch.digest), ch.n_draws)
        gproofs = []
        for k, op in enumerate(proofs):
            lp = pr.layer_proof(k)
            assert tuple(int(x) for x in lp["commitment"]) == op["commitment"], k
            assert lp["positions"].tolist() == op["positions"], k
            assert as_tuples(lp["values"]) == op["values"], k
            assert as_tuples(lp["fri_witness"]) == op["fri_witness"], k
            assert as_tuples(lp["hash_witness"]) == op["hash_witness"], k
            gproofs.append(dict(commitment=tuple(int(x) for x in lp["commitment"]),
                                fri_witness=as_tuples(lp["fri_witness"]),
                                hash_witness=as_tuples(lp["hash_witness"])))
        with pytest.raises(sh.StwoHipError):
            pr.decommit(dig2, nd2)  # once per prover
    finally:
        pr.destroy()
        tw.destroy()
    evals_of = lambda qs: [tuple(int(col[k][q]) for k in range(4)) for q in qs]  # noqa: E731
    D.fri_verify(O.Channel.from_buffer_copy(start), cfg, (gproofs, as_tuples(glast)), log_size, evals_of)


@pytest.mark.parametrize("seed,pow_bits", [(0, 2), (1, 10), (0x1111222233334344, 16), (5, 21), (9, 23)])
def test_grind_vs_oracle(ctx, seed, pow_bits):
    """stwo_hip_grind vs SimdBackend's rule (smallest (hi, lo < 2^20) nonce,
    SIMD/grind.rs:22-62) computed by the oracle; pow_bits 2 after mix_u64(0)
    is test_grind_blake_is_deterministic's shape, 21-23 need hi > 0 often."""
    ch = O.Channel()
    ch.mix_u64(seed)
    d = ch.digest_bytes()
    assert ctx.grind(d, pow_bits) == O.grind(d, pow_bits)


def test_grind_reference_chain(ctx):
    """test_parallel_grind_with_high_pow_bits's loop (SIMD/grind.rs:201-211) at
    20 bits: grind, mix the nonce, mix the constant, repeat; each nonce equals
    the oracle's."""
    ch = O.Channel()
    ch.mix_u64(0x1111222233334344)
    for _ in range(4):
        d = ch.digest_bytes()
        n = ctx.grind(d, 20)
        assert n == O.grind(d, 20)
        ch.mix_u64(n)
        ch.mix_u64(0x1111222233334344)


def test_grind_rejects_more_than_32_bits(ctx):
    with pytest.raises(sh.StwoHipError):
        ctx.grind(bytes(32), 33)

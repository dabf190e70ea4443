"""GPU parity of the end-to-end wide-Fibonacci prove (stwo_hip_wide_fib_prove)
against the oracle's restatement of the reference's flow: the proof bytes
must be identical, and the proof must pass the restated verifier.  At the
prove benchmark's full size (256 columns x 2^18 rows) the oracle prover is
not run; there the library's proof is checked by the verifier, which
recomputes everything it checks from the proof and the channel."""
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "oracle"))
import oracle as O  # noqa: E402
import prove as PR  # noqa: E402
import stwo_hip as sh  # noqa: E402

pytestmark = pytest.mark.gpu
U = np.uint32


@pytest.fixture(scope="module")
def ctx():
    c = sh.Context(0)
    yield c
    c.close()


_tws = {}


def twiddles(ctx, E):
    if E not in _tws:
        _tws[E] = ctx.precompute_twiddles(sh.half_odds(E - 1))
    return _tws[E]


def lib_prove(ctx, a, b, n_cols, log_n, cfg, digest=bytes(32), n_draws=0, store_coeffs=True):
    da, db = ctx.to_device(a), ctx.to_device(b)
    try:
        return ctx.wide_fib_prove(da, db, n_cols, log_n, twiddles(ctx, log_n + cfg.fri.log_blowup_factor), cfg,
                                  digest, n_draws, store_coeffs=store_coeffs)
    finally:
        da.free()
        db.free()


def oracle_cfg(cfg):
    return PR.PcsConfig(cfg.pow_bits, cfg.fri.log_blowup_factor, cfg.fri.log_last_layer_degree_bound,
                        cfg.fri.n_queries, cfg.fri.line_fold_step)


def pcs(pow_bits=10, blowup=1, last=0, queries=3, step=1):
    return sh.PcsConfig(pow_bits, sh.FriConfig(blowup, last, queries, step))


CASES = [
    # (log_n, n_cols, config, inputs)
    (4, 8, pcs(), "ref"),
    (5, 100, pcs(), "ref"),
    (6, 3, pcs(4, 2, 2, 10, 2), "rand"),
    (8, 64, pcs(8, 1, 1, 7, 3), "rand"),
    (10, 256, pcs(), "ref"),
    (12, 16, pcs(12, 3, 3, 20, 4), "rand"),
    # the FRI commit's end and the grind's first range run on the device for
    # last layers of at most 2^6 values and pow_bits below 14; these two sit
    # on either side (2^6 values and 13 bits; 2^7 values and 14 bits)
    (7, 8, pcs(13, 1, 5, 4, 1), "rand"),
    (8, 8, pcs(14, 1, 6, 3, 1), "rand"),
]


def inputs(log_n, kind, seed=0):
    if kind == "ref":
        return np.ones(1 << log_n, dtype=U), np.arange(1 << log_n, dtype=U)
    return O.fill_random(1 << log_n, seed + log_n, 0), O.fill_random(1 << log_n, seed + log_n, 1)


@pytest.mark.parametrize("case", range(len(CASES)))
def test_prove_bytes_equal_oracle(ctx, case):
    log_n, n_cols, cfg, kind = CASES[case]
    a, b = inputs(log_n, kind)
    data, phases, qpos, digest, n_draws = lib_prove(ctx, a, b, n_cols, log_n, cfg)
    ch = O.Channel()
    want = PR.wide_fib_prove(a, b, n_cols, log_n, oracle_cfg(cfg), ch)
    assert qpos == want["query_positions"]
    assert data == PR.proof_bytes(want)
    assert digest == ch.digest_bytes() and n_draws == ch.n_draws
    PR.verify(PR.parse_proof(data), n_cols, log_n)
    assert set(phases) == set(sh.PROVE_PHASES) and all(v >= 0 for v in phases.values())


@pytest.mark.parametrize("case", [c for c in range(len(CASES)) if CASES[c][2].fri.log_blowup_factor == 1])
def test_prove_without_stored_coefficients_equals_oracle(ctx, case):
    """The PCS default (store_polynomials_coefficients = false): the trace LDE
    drops the coefficients and the OODS samples come from the committed
    evaluations through barycentric weights; the proof bytes and the channel
    are the oracle's (which samples the coefficients)."""
    log_n, n_cols, cfg, kind = CASES[case]
    a, b = inputs(log_n, kind, 3)
    data, _, qpos, digest, n_draws = lib_prove(ctx, a, b, n_cols, log_n, cfg, store_coeffs=False)
    ch = O.Channel()
    want = PR.wide_fib_prove(a, b, n_cols, log_n, oracle_cfg(cfg), ch)
    assert qpos == want["query_positions"]
    assert data == PR.proof_bytes(want)
    assert digest == ch.digest_bytes() and n_draws == ch.n_draws


def test_prove_plans_follow_the_shape(ctx):
    """The context keeps the last prove shape's launch plans: proves of
    alternating shapes and both sampling paths, each against the oracle."""
    seq = [(5, 100, True), (4, 8, False), (5, 100, False), (5, 100, True), (8, 64, True), (5, 100, True)]
    for i, (log_n, n_cols, store) in enumerate(seq):
        a, b = inputs(log_n, "rand", 20 + i)
        data, _, _, _, _ = lib_prove(ctx, a, b, n_cols, log_n, pcs(), store_coeffs=store)
        want = PR.wide_fib_prove(a, b, n_cols, log_n, PR.PcsConfig(), O.Channel())
        assert data == PR.proof_bytes(want), (i, log_n, n_cols, store)


def test_prove_without_stored_coefficients_needs_blowup_1(ctx):
    a, b = inputs(5, "ref")
    da, db = ctx.to_device(a), ctx.to_device(b)
    try:
        with pytest.raises(sh.StwoHipError) as e:
            ctx.wide_fib_prove(da, db, 8, 5, twiddles(ctx, 7), pcs(blowup=2), store_coeffs=False)
        assert e.value.code == -1
    finally:
        da.free()
        db.free()


def test_prove_continues_a_channel(ctx):
    """A non-default starting channel state (a prover that mixed other data first)."""
    log_n, n_cols = 5, 10
    a, b = inputs(log_n, "rand", 7)
    ch = O.Channel()
    ch.mix_u64(0x0123456789ABCDEF)
    ch.draw_u32s()
    data, _, _, digest, n_draws = lib_prove(ctx, a, b, n_cols, log_n, pcs(), ch.digest_bytes(), ch.n_draws)
    want = PR.wide_fib_prove(a, b, n_cols, log_n, PR.PcsConfig(), ch)
    assert data == PR.proof_bytes(want)
    assert digest == ch.digest_bytes() and n_draws == ch.n_draws


def test_prove_full_size_verifies(ctx):
    """The prove benchmark's workload: 256 columns x 2^18 rows, PcsConfig::default():
    the proof bytes equal the oracle's (about 1.5 s on 16 host cores) and the
    restated verifier accepts them and rejects a tampered copy."""
    log_n, n_cols = 18, 256
    a, b = inputs(log_n, "rand", 4)
    data, phases, qpos, digest, n_draws = lib_prove(ctx, a, b, n_cols, log_n, pcs())
    ch = O.Channel()
    want = PR.wide_fib_prove(a, b, n_cols, log_n, PR.PcsConfig(), ch)
    assert qpos == want["query_positions"]
    assert data == PR.proof_bytes(want)
    assert digest == ch.digest_bytes() and n_draws == ch.n_draws
    proof = PR.parse_proof(data)
    PR.verify(proof, n_cols, log_n)
    # the PCS default path (no stored coefficients) gives the same bytes
    data2, _, _, _, _ = lib_prove(ctx, a, b, n_cols, log_n, pcs(), store_coeffs=False)
    assert data2 == data
    bad = PR.parse_proof(data)
    bad["queried_values"][1][100][0] = (bad["queried_values"][1][100][0] + 1) % O.P
    with pytest.raises(PR.VerificationError):
        PR.verify(bad, n_cols, log_n)


def test_prove_rejects_bad_arguments(ctx):
    a, b = inputs(5, "ref")
    da, db = ctx.to_device(a), ctx.to_device(b)
    tw = twiddles(ctx, 6)
    try:
        # the last case's log_n + blowup wraps to 4 in 32 bits
        for n_cols, log_n, cfg in [(2, 5, pcs()), (8, 3, pcs()), (8, 5, pcs(blowup=0)), (8, 5, pcs(pow_bits=33)),
                                   (8, 5, pcs(blowup=(1 << 32) - 1))]:
            with pytest.raises(sh.StwoHipError) as e:
                ctx.wide_fib_prove(da, db, n_cols, log_n, tw, cfg)
            assert e.value.code == -1
    finally:
        da.free()
        db.free()

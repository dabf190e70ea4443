"""Pins the CPU oracle (oracle/stwo_oracle.c) before it is trusted as the checker.

Each check here comes from outside the oracle: the reference's own
known-answer tests, hashlib, brute-force evaluation fixtures written by
tests/golden/gen_golden.py, or an invariant the mathematics fixes.  A
plausible mistake in the oracle (a wrong sign, index, twiddle, byte order)
fails at least one of them.  CPU only.

Reference paths: STWO = vendor/stwo-upstream-dev-62b228e/crates/stwo/src.
"""
import hashlib
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "..", "oracle"))
import oracle  # noqa: E402

P = oracle.P
U = np.uint32


def arr(x):
    return np.array(x, dtype=U)


# ---------------------------------------------------------------- fields ---

def test_qm31_reference_kat():
    """STWO/core/fields/qm31.rs:157-171 (test_ops) and :150-155 (test_inverse)."""
    qm0, qm1 = [1, 2, 3, 4], [4, 5, 6, 7]
    assert oracle.qm31_mul(qm0, qm1) == [P - 71, 93, P - 16, 50]
    assert oracle.qm31_mul(qm0, oracle.qm31_inv(qm0)) == [1, 0, 0, 0]
    prod = oracle.qm31_mul(qm0, qm1)
    assert oracle.qm31_mul(prod, oracle.qm31_inv(qm1)) == qm0


def test_m31_reduce_edges():
    """M31::reduce on the extremes (STWO/core/fields/m31.rs:40-60 doc tests)."""
    assert oracle.m31_mul(P - 1, P - 1) == 1
    assert oracle.m31_mul(0, P - 1) == 0
    assert oracle.m31_mul(2, (P + 1) // 2) == 1
    for a in [1, 2, 3, 12345, P - 1, P - 2]:
        assert oracle.m31_mul(a, oracle.m31_inv(a)) == 1


def test_circle_generator_order():
    """M31_CIRCLE_GEN has order 2^31 (STWO/core/circle.rs:190-208): gen^(2^30) != 1, gen^(2^31) = 1."""
    assert oracle.index_to_point(1) == (2, 1268011823)
    assert oracle.index_to_point(1 << 30) != (1, 0)
    assert oracle.index_to_point(0) == (1, 0)
    x, y = oracle.index_to_point(987654321)
    assert (x * x + y * y) % P == 1


# --------------------------------------------------------------- blake2s ---

def test_blake2s_reference_kats():
    """STWO/core/vcs/blake2_hash.rs:134-152 (single_hash_test, hash_state_test)."""
    assert oracle.blake2s(b"a").hex() == "4a0d129873403037c2cd9b9048203687f6233fb6738956e0349bd4320fec3e90"
    assert oracle.blake2s(b"ab") == hashlib.blake2s(b"ab").digest()
    assert oracle.blake2s(b"") == hashlib.blake2s(b"").digest()


@pytest.mark.parametrize("n", [0, 1, 31, 32, 63, 64, 65, 127, 128, 129, 1000])
def test_blake2s_matches_hashlib(n):
    data = bytes((7 * i + 3) & 0xFF for i in range(n))
    assert oracle.blake2s(data) == hashlib.blake2s(data).digest()


# --------------------------------------------------------------- channel ---

def test_channel_reference_kats():
    """STWO/core/channel/blake2s.rs:214-247 (test_mix_u64, test_mix_u32s)."""
    ch = oracle.Channel()
    ch.mix_u64(0x1111222233334444)
    assert ch.digest_bytes().hex() == "bc9e3fc1d24e8897956d3359327397249d6bcacd224d927404e7ba4a77dc6ece"
    ch2 = oracle.Channel()
    ch2.mix_u32s([0x33334444, 0x11112222])
    assert ch2.digest_bytes() == ch.digest_bytes()
    ch3 = oracle.Channel()
    ch3.mix_u32s(list(range(1, 10)))
    assert ch3.digest_bytes().hex() == "7091768357bb1bb3346fdab6b357d7fa46b8fbe32c2e4324a0ffc294cbf9a1c7"


def test_channel_draws():
    """STWO/core/channel/blake2s.rs:164-200: n_draws counts, draws differ, and a draw
    equals the hashlib restatement H(digest || le32(n) || 0x00)."""
    ch = oracle.Channel()
    w0 = ch.draw_u32s()
    assert ch.n_draws == 1
    exp = hashlib.blake2s(bytes(32) + (0).to_bytes(4, "little") + b"\x00").digest()
    assert w0 == [int.from_bytes(exp[i:i + 4], "little") for i in range(0, 32, 4)]
    assert ch.draw_u32s() != w0
    f = ch.draw_secure_felt()
    assert all(0 <= v < P for v in f)


# -------------------------------------------------------------- twiddles ---

def test_twiddles_golden(golden):
    for case in golden("twiddles"):
        tw, itw = oracle.precompute_twiddles(case["initial"], case["log_size"])
        assert tw.tolist() == case["twiddles"], case["kind"]
        assert itw.tolist() == case["itwiddles"], case["kind"]


@pytest.mark.parametrize("log_size", [0, 1, 2, 3, 7, 12])
def test_bit_reverse_closed_form(log_size):
    """bit_reverse (STWO/core/utils.rs:91-101) is the permutation
    v'[i] = v[rev(i)] with rev reversing log_size bits (written here from
    the binary digits), and an involution."""
    n = 1 << log_size
    v = np.arange(1000, 1000 + n, dtype=np.uint32)
    got = oracle.bit_reverse(v.copy())
    rev = [int(format(i, "0%db" % log_size)[::-1], 2) if log_size else 0 for i in range(n)]
    assert got.tolist() == [int(v[r]) for r in rev]
    assert np.array_equal(oracle.bit_reverse(got.copy()), v)


def test_twiddle_inverse_invariant():
    tw, itw = oracle.precompute_twiddles(oracle.half_odds_initial(13), 13)
    assert np.all((tw.astype(np.uint64) * itw.astype(np.uint64)) % P == 1)


# ------------------------------------------------------------------- FFT ---

def _tree_for(case):
    # Root the tree at the evaluation domain's half coset (as the reference's
    # parity tests do, crates/stwo-metal/tests/metal_rfft_ifft_native.rs:29-30).
    return oracle.precompute_twiddles(case["domain_initial"], case["log_size"] - 1)


def test_evaluate_golden(golden):
    for case in golden("fft"):
        tw, itw = _tree_for(case)
        ev = oracle.evaluate(arr(case["coeffs"]), case["domain_initial"], case["log_size"], tw)
        assert ev.tolist() == case["evals"], (case["coeff_log"], case["log_size"])


def test_interpolate_golden(golden):
    for case in golden("fft"):
        if case["coeff_log"] != case["log_size"]:
            continue
        tw, itw = _tree_for(case)
        co = oracle.interpolate(arr(case["evals"]), case["domain_initial"], case["log_size"], itw)
        assert co.tolist() == case["coeffs"], case["log_size"]


def test_eval_at_point_kats():
    """CPUB/circle.rs tests test_evaluate_{2,4,8}_coeffs: evaluating 1..2^n on
    CanonicCoset(n) and bit-reversing gives eval_at_point at each domain point;
    checked here through the golden brute-force evaluator's definition for
    n = 1, 2, 3 with coefficients (1..2^n)."""
    sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden"))
    import gen_golden as g
    for n in [1, 2, 3]:
        coeffs = list(range(1, (1 << n) + 1))
        init = oracle.half_odds_initial(n - 1)
        tw, _ = oracle.precompute_twiddles(init, n - 1)
        ev = oracle.evaluate(arr(coeffs), init, n, tw)
        for i in range(1 << n):
            pt = g.domain_point(init, n - 1, g.bitrev(i, n))
            assert int(ev[i]) == g.circle_eval(coeffs, pt)
    # test_eval_at_point_with_4_coeffs asserts c0 + c1*y + c2*x + c3*x*y for coeffs [1, 3, 2, 4] at (5, 8)
    c = [1, 3, 2, 4]
    assert g.circle_eval(c, (5, 8)) == (c[0] + c[1] * 8 + c[2] * 5 + c[3] * 5 * 8) % P


@pytest.mark.parametrize("log_size", [11, 14, 17])
def test_fft_round_trip_and_extension(log_size):
    """interpolate(evaluate(c)) == c, and evaluating on a 2x domain then
    interpolating gives c zero-padded (extension invariant)."""
    coeffs = oracle.fill_random(1 << log_size, 99, log_size)
    init = oracle.half_odds_initial(log_size)
    tw, itw = oracle.precompute_twiddles(init, log_size)
    ev = oracle.evaluate(coeffs, init, log_size + 1, tw)
    back = oracle.interpolate(ev, init, log_size + 1, itw)
    assert np.array_equal(back[: 1 << log_size], coeffs)
    assert not back[1 << log_size:].any()
    # evaluation on the domain of the coefficients' own size, via a subtree of the same root
    ev_small = oracle.evaluate(coeffs, oracle.half_odds_initial(log_size - 1), log_size, tw)
    back2 = oracle.interpolate(ev_small, oracle.half_odds_initial(log_size - 1), log_size, itw)
    assert np.array_equal(back2, coeffs)


def test_canonical_outputs_on_edge_values():
    """All outputs in [0, P) for all-zero, all-(P-1) and alternating inputs."""
    n = 10
    init = oracle.half_odds_initial(n)
    tw, itw = oracle.precompute_twiddles(init, n)
    for col in [np.zeros(1 << n, U), np.full(1 << n, P - 1, U), np.tile(arr([0, P - 1]), 1 << (n - 1))]:
        ev = oracle.evaluate(col, init, n + 1, tw)
        assert int(ev.max()) < P
        co = oracle.interpolate(ev, init, n + 1, itw)
        assert int(co.max()) < P
    assert not oracle.evaluate(np.zeros(1 << n, U), init, n + 1, tw).any()


# ---------------------------------------------------------------- Merkle ---

def test_merkle_golden(golden):
    for case in golden("merkle"):
        cols = [arr(c) for c in case["cols"]]
        layers = oracle.merkle_commit(cols)
        assert layers[:8].tolist() == case["mixed_root"], case["shape"]
        L = max(case["shape"])
        leaf = layers[oracle.layer_offset(L): oracle.layer_offset(L + 1)].reshape(-1, 8)
        assert leaf.tolist() == case["mixed_leaf_layer"], case["shape"]
        if case["lifted_root"] is not None:
            lay = oracle.lifted_commit(cols, L)
            assert lay[:8].tolist() == case["lifted_root"], case["shape"]


def test_lifted_merkle_reference_kat():
    """STWO/prover/vcs_lifted/prover.rs:231-275 (prepare_merkle + test_lifted_merkle_leaves):
    columns 0..4, 0..8, 0..16 lifted to log 4."""
    cols = [np.arange(1 << i, dtype=U) for i in range(2, 5)]
    leaves = oracle.lifted_build_leaves(cols, 4).reshape(-1, 8)

    def hw(vals):
        d = hashlib.blake2s(b"".join(int(v).to_bytes(4, "little") for v in vals)).digest()
        return [int.from_bytes(d[i:i + 4], "little") for i in range(0, 32, 4)]

    assert leaves[0].tolist() == hw([0, 0, 0])
    assert leaves[4].tolist() == hw([0, 2, 4])
    assert leaves[15].tolist() == hw([3, 7, 15])


def test_empty_commitments():
    """Empty column sets commit to H("") (STWO/prover/vcs/prover.rs:40-44, CPUB/merkle_lifted.rs:46-48)."""
    exp = hashlib.blake2s(b"").digest()
    e = [int.from_bytes(exp[i:i + 4], "little") for i in range(0, 32, 4)]
    assert oracle.merkle_commit([])[:8].tolist() == e
    assert oracle.lifted_build_leaves([], 0).tolist() == e


def test_lifted_equals_mixed_for_equal_sizes():
    cols = [oracle.fill_random(1 << 9, 5, c) for c in range(20)]
    assert np.array_equal(oracle.merkle_commit(cols)[:8], oracle.lifted_commit(cols, 9)[:8])


def test_lift_poly_invariant():
    """STWO/prover/vcs_lifted/prover.rs:307-336 (test_bit_reverse_lifted_merkle_cpu):
    committing a small evaluation lifted to L equals committing the evaluation of
    f(pi^{L-n}(p)) on the big domain; the latter is computed by brute force."""
    sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden"))
    import gen_golden as g
    n, L = 3, 7
    coeffs = list(range(1 << n))
    init_small = oracle.half_odds_initial(n - 1)
    tw, _ = oracle.precompute_twiddles(init_small, n - 1)
    small = oracle.evaluate(arr(coeffs), init_small, n, tw)
    big_init = oracle.half_odds_initial(L - 1)
    lifted = []
    for i in range(1 << L):
        pt = g.domain_point(big_init, L - 1, g.bitrev(i, L))
        for _ in range(L - n):
            pt = g.pt_add(pt, pt)
        lifted.append(g.circle_eval(coeffs, pt))
    zero = np.zeros(1 << L, dtype=U)
    r1 = oracle.lifted_commit([arr(lifted), zero], L)[:8]
    r2 = oracle.lifted_commit([small, zero], L)[:8]
    r3 = oracle.merkle_commit([arr(lifted), zero])[:8]
    assert np.array_equal(r1, r2) and np.array_equal(r1, r3)


# ------------------------------------------------------------------ folds ---

def test_fold_circle_golden(golden):
    for case in golden("folds")["circle"]:
        N = case["log_size"]
        src = [arr([v[k] for v in case["src"]]) for k in range(4)]
        dst = [arr([v[k] for v in case["prev"]]) for k in range(4)]
        oracle.fold_circle_into_line(dst, src, oracle.half_odds_initial(N - 1), N, case["alpha"])
        got = np.stack(dst, axis=1).tolist()
        assert got == case["dst"], N


def test_fold_line_golden(golden):
    for case in golden("folds")["line"]:
        src = [arr([v[k] for v in case["src"]]) for k in range(4)]
        out = oracle.fold_line(src, case["line_initial"], case["log_size"], case["alpha"], case["fold_step"])
        got = np.stack(out, axis=1).tolist()
        assert got == case["dst"], (case["log_size"], case["fold_step"])


def test_eval_at_point_by_folding_kat():
    """CPUB/circle.rs test_cpu_eval_at_point_by_folding (:143-180 and the test
    module): folding an evaluation with the folding alphas of a secure point
    and dividing by 2^log_size gives the polynomial's value at the point."""
    sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden"))
    import gen_golden as g
    coeffs = [691, 805673, 5, 435684, 4832, 23876431, 197, 897346068]
    n = 10
    init = oracle.half_odds_initial(n - 1)
    tw, _ = oracle.precompute_twiddles(oracle.half_odds_initial(n), n)  # CanonicCoset(11) half coset
    ev = oracle.evaluate(arr(coeffs), init, n, tw)
    # SECURE_FIELD_CIRCLE_GEN (STWO/core/circle.rs:211-214), get_point(i) = gen.mul(i).
    gx, gy = (1, 0, 478637715, 513582971), (992285211, 649143431, 740191619, 1186584352)

    def sadd(p, q):
        return (g.qm_add(g.qm_mul(p[0], q[0]), g.qm_scale(g.qm_mul(p[1], q[1]), P - 1)),
                g.qm_add(g.qm_mul(p[0], q[1]), g.qm_mul(p[1], q[0])))

    def smul(k):
        res, cur = ((1, 0, 0, 0), (0, 0, 0, 0)), (gx, gy)
        while k:
            if k & 1:
                res = sadd(res, cur)
            cur = sadd(cur, cur)
            k >>= 1
        return res

    for idx in [348, 9736524, 13, 346752]:
        x, y = smul(idx)
        # get_folding_alphas (STWO/core/poly/utils.rs:44-67): [pi^{n-2}(x), ..., x, y]
        alphas = [None] * n
        alphas[n - 1] = y
        cur = x
        for i in range(n - 2, -1, -1):
            alphas[i] = cur
            cur = g.qm_add(g.qm_scale(g.qm_mul(cur, cur), 2), (P - 1, 0, 0, 0))
        src = [ev.copy()] + [np.zeros(1 << n, U) for _ in range(3)]
        dst = [np.zeros(1 << (n - 1), U) for _ in range(4)]
        oracle.fold_circle_into_line(dst, src, init, n, list(alphas.pop()))
        lg, linit = n - 1, oracle.half_odds_initial(n - 1)
        while lg > 0:
            dst = oracle.fold_line(dst, linit, lg, list(alphas.pop()), 1)
            linit = (linit * 2) % (1 << 31)
            lg -= 1
        folded = [int(d[0]) for d in dst]
        inv = pow(1 << n, P - 2, P)
        got = tuple(v * inv % P for v in folded)
        # direct evaluation with QM31 point coordinates
        maps = [y]
        cx = x
        for _ in range(1, 3):
            maps.append(cx)
            cx = g.qm_add(g.qm_scale(g.qm_mul(cx, cx), 2), (P - 1, 0, 0, 0))
        maps.reverse()
        want = g.fold([(c, 0, 0, 0) for c in coeffs], maps)
        assert got == want, idx


# -------------------------------------------------------------------- FRI ---

def test_fri_commit_golden(golden):
    for case in golden("fri"):
        lb, lld, nq, step = case["config"]
        cfg = oracle.FriConfig(lb, lld, nq, step)
        ch = oracle.Channel()
        ch.mix_root(arr([int.from_bytes(bytes(case["prior_root"][i:i + 4]), "little")
                         for i in range(0, 32, 4)]))
        col = [arr([v[k] for v in case["evals"]]) for k in range(4)]
        roots, last, _ = oracle.fri_commit(ch, cfg, col, case["log_size"])
        assert roots.tolist() == case["roots"], case["log_size"]
        assert last.tolist() == case["last_layer_poly"], case["log_size"]
        assert list(ch.digest_bytes()) == case["final_digest"], case["log_size"]


def test_fri_rejects_high_degree():
    """A random (not low-degree) evaluation fails the last-layer degree check
    (STWO/prover/fri.rs:281-285 panics with "invalid degree")."""
    n = 8
    col = [oracle.fill_random(1 << n, 7, k) for k in range(4)]
    with pytest.raises(ValueError):
        oracle.fri_commit(oracle.Channel(), oracle.FriConfig(1, 1, 3, 1), col, n)


# ------------------------------------------------------------ proof of work ---

def verify_pow_nonce(digest, n_bits, nonce):
    """Blake2sChannel::verify_pow_nonce (STWO/core/channel/blake2s.rs:122-139),
    restated with hashlib: trailing zeros of the first 16 bytes as a u128."""
    import struct
    pre = hashlib.blake2s(struct.pack("<I", 0x12345678) + bytes(12) + bytes(digest) +
                          struct.pack("<I", n_bits)).digest()
    res = hashlib.blake2s(pre + struct.pack("<Q", nonce)).digest()
    v = int.from_bytes(res[:16], "little")
    tz = (v & -v).bit_length() - 1 if v else 128
    return tz >= n_bits


@pytest.mark.parametrize("seed,pow_bits", [(0, 2), (1, 5), (2, 9), (3, 12), (0x1111222233334344, 8)])
def test_grind_is_smallest_valid_nonce(seed, pow_bits):
    """SimdBackend's grind (SIMD/grind.rs:22-62) returns the smallest valid
    nonce (hi << 32) | lo with lo < 2^20: checked by brute force with the
    reference's verify rule; pow_bits 2 with mix_u64(0) is the shape of
    test_grind_blake_is_deterministic."""
    ch = oracle.Channel()
    ch.mix_u64(seed)
    d = ch.digest_bytes()
    got = oracle.grind(d, pow_bits)
    assert verify_pow_nonce(d, pow_bits, got)
    assert got < (1 << 20)  # found at hi = 0 for these sizes
    assert not any(verify_pow_nonce(d, pow_bits, n) for n in range(got))
    assert oracle.pow_prefixed_digest(d, pow_bits) == hashlib.blake2s(
        (0x12345678).to_bytes(4, "little") + bytes(12) + d + pow_bits.to_bytes(4, "little")).digest()


def test_grind_hi_words_in_order():
    """With a window of 2^2 low values per high word the search must move to
    hi = 1, 2, ...: the answer is the smallest valid (hi, lo) in that lattice."""
    ch = oracle.Channel()
    ch.mix_u64(77)
    d = ch.digest_bytes()
    got = oracle.grind(d, 6, low_bits=2)
    hi, lo = got >> 32, got & 0xFFFFFFFF
    assert lo < 4 and verify_pow_nonce(d, 6, got)
    for h in range(hi + 1):
        for l in range(4):
            n = (h << 32) | l
            if n < got:
                assert not verify_pow_nonce(d, 6, n)

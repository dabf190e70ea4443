"""GPU parity: every hot-path entry point of libstwo_hip, called through the C
ABI, against the pinned CPU oracle and the golden fixtures.  Bit-exact
equality is the bar (all arithmetic is exact integer work).

Shapes follow the reference's parity tests (crates/stwo-metal/tests/*, cited
per test) plus the first benchmark workload (16 columns of 2^16 -> blowup 2),
checked in full.  Inputs are seeded (SplitMix64, u64 % P); the reference's
SmallRng vectors cannot be regenerated here.
"""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "..", "oracle"))
import oracle  # noqa: E402
import stwo_hip as sh  # noqa: E402

pytestmark = pytest.mark.gpu
P = sh.P
U = np.uint32


@pytest.fixture(scope="module")
def ctx():
    c = sh.Context(0)
    yield c
    c.close()


def arr(x):
    return np.array(x, dtype=U)


def rand(n, seed, stream):
    return oracle.fill_random(n, seed, stream)


# ---------------------------------------------------------------- twiddles --

def test_twiddles_golden(ctx, golden):
    """metal_twiddle_precompute_matches_cpu_on_{canonic,shifted}_cosets
    (crates/stwo-metal/tests/metal_twiddle_precompute_native.rs:17-49)."""
    for case in golden("twiddles"):
        tw = ctx.precompute_twiddles(sh.Coset(case["initial"], case["log_size"]))
        t, it = tw.download()
        assert t.tolist() == case["twiddles"], case
        assert it.tolist() == case["itwiddles"], case
        tw.destroy()


@pytest.mark.parametrize("log_size", [13, 21])
def test_twiddles_large_vs_oracle(ctx, log_size):
    root = sh.half_odds(log_size)
    tw = ctx.precompute_twiddles(root)
    t, it = tw.download()
    et, eit = oracle.precompute_twiddles(root.initial_index, log_size)
    assert np.array_equal(t, et) and np.array_equal(it, eit)
    tw.destroy()


def test_fill_random_matches_oracle(ctx):
    b = ctx.alloc(1 << 12)
    ctx.fill_random(b, 1 << 12, 7, 3)
    assert np.array_equal(b.download(), rand(1 << 12, 7, 3))
    b.free()


# ------------------------------------------------------------- bit reverse --

@pytest.mark.parametrize("log_size", [0, 1, 4, 10, 11, 12, 13, 16, 19])
def test_bit_reverse_column(ctx, log_size):
    """metal_base_field_bit_reverse_matches_cpu_backend (tests/metal_base_field_bit_reverse.rs:8): 0..N."""
    n = 1 << log_size
    v = np.arange(n, dtype=U)
    d = ctx.to_device(v)
    ctx.bit_reverse_column(d, log_size)
    assert np.array_equal(d.download(), oracle.bit_reverse(v))
    d.free()


@pytest.mark.parametrize("log_size,offset", [(12, 1), (17, 3), (20, 2)])
def test_bit_reverse_column_unaligned(ctx, log_size, offset):
    """A column that starts `offset` words into its buffer (not 16-byte
    aligned: the tiled kernel's one-word path); the words around it unchanged."""
    n = 1 << log_size
    v = rand(n, 7, log_size)
    buf = ctx.alloc(n + 8)
    guard = np.full(n + 8, 0x5A5A5A5A, dtype=U)
    buf.upload(guard)
    buf.upload(v, offset=offset)
    ctx.bit_reverse_column(buf.at(offset), log_size)
    got = buf.download()
    assert np.array_equal(got[offset:offset + n], oracle.bit_reverse(v))
    assert np.all(got[:offset] == 0x5A5A5A5A) and np.all(got[offset + n:] == 0x5A5A5A5A)
    buf.free()


def test_bit_reverse_secure_column(ctx):
    """metal_secure_field_bit_reverse_matches_cpu_backend (tests/metal_secure_field_bit_reverse.rs:8)."""
    log_size = 10
    cols = [rand(1 << log_size, 11, k) for k in range(4)]
    ds = [ctx.to_device(c) for c in cols]
    ctx.bit_reverse_secure_column(ds, log_size)
    for c, d in zip(cols, ds):
        assert np.array_equal(d.download(), oracle.bit_reverse(c))
        d.free()


# --------------------------------------------------------------------- FFT --

def test_fft_golden(ctx, golden):
    """Brute-force evaluations (log 1..10, blowups, non-canonic domains)."""
    for case in golden("fft"):
        N, n = case["log_size"], case["coeff_log"]
        dom = sh.Coset(case["domain_initial"], N - 1)
        tw = ctx.precompute_twiddles(dom)
        co = ctx.to_device(arr(case["coeffs"]))
        out = ctx.alloc(1 << N)
        ctx.evaluate(co, n, dom, tw, out)
        assert out.download().tolist() == case["evals"], (n, N)
        if n == N:
            ctx.interpolate(out, dom, tw)
            assert out.download().tolist() == case["coeffs"], N
        co.free()
        out.free()
        tw.destroy()


@pytest.mark.parametrize("log_size", [3, 8, 11, 12, 13, 15, 17, 18, 19, 20])
def test_interpolate_vs_oracle(ctx, log_size):
    """Includes metal_ifft_large_column_regression's log 17 (tests/metal_rfft_ifft_native.rs:60)."""
    root = sh.half_odds(log_size + 1)
    tw = ctx.precompute_twiddles(root)
    dom = sh.canonic_domain(log_size)
    ev = rand(1 << log_size, 42, log_size)
    d = ctx.to_device(ev)
    ctx.interpolate(d, dom, tw)
    _, itw = oracle.precompute_twiddles(root.initial_index, root.log_size)
    exp = oracle.interpolate(ev, dom.initial_index, log_size, itw)
    assert np.array_equal(d.download(), exp)
    d.free()
    tw.destroy()


@pytest.mark.parametrize("n,b", [(3, 0), (4, 1), (7, 1), (11, 1), (12, 1), (12, 2), (13, 1), (16, 2), (17, 1),
                                 (18, 2), (20, 2)])
def test_evaluate_vs_oracle(ctx, n, b):
    N = n + b
    root = sh.half_odds(N - 1)
    tw = ctx.precompute_twiddles(root)
    dom = sh.canonic_domain(N)
    coeffs = rand(1 << n, 5, n * 10 + b)
    co = ctx.to_device(coeffs)
    out = ctx.alloc(1 << N)
    ctx.evaluate(co, n, dom, tw, out)
    et, _ = oracle.precompute_twiddles(root.initial_index, root.log_size)
    exp = oracle.evaluate(coeffs, dom.initial_index, N, et)
    assert np.array_equal(out.download(), exp)
    co.free()
    out.free()
    tw.destroy()


def test_fft_all_pass_shapes(ctx):
    """Every pass/stage layout the planner produces: interpolate at log 4..22 and
    evaluate at log 4..21 with blowups 1..3, against the oracle."""
    top = 23
    root = sh.half_odds(top - 1)
    tw = ctx.precompute_twiddles(root)
    et, eit = oracle.precompute_twiddles(root.initial_index, root.log_size)
    for log_size in range(4, 23):
        dom = sh.canonic_domain(log_size)
        ev = rand(1 << log_size, 71, log_size)
        d = ctx.to_device(ev)
        ctx.interpolate(d, dom, tw)
        exp = oracle.interpolate(ev, dom.initial_index, log_size, eit)
        assert np.array_equal(d.download(), exp), ("interpolate", log_size)
        d.free()
    for n in range(4, 22):
        for b in (1, 2, 3):
            N = n + b
            if N > top:
                continue
            coeffs = rand(1 << n, 73, 100 * n + b)
            co = ctx.to_device(coeffs)
            out = ctx.alloc(1 << N)
            ctx.evaluate(co, n, sh.canonic_domain(N), tw, out)
            exp = oracle.evaluate(coeffs, sh.canonic_domain(N).initial_index, N, et)
            assert np.array_equal(out.download(), exp), ("evaluate", n, b)
            co.free()
            out.free()
    tw.destroy()


def test_rfft_ifft_shifted_domain_round_trip(ctx):
    """metal_rfft_and_ifft_match_cpu_on_shifted_domain (tests/metal_rfft_ifft_native.rs:21):
    CircleDomain(half_odds(7)), twiddles on its half coset."""
    dom = sh.half_odds(7)
    tw = ctx.precompute_twiddles(dom)
    coeffs = rand(1 << 8, 23, 0)
    co = ctx.to_device(coeffs)
    out = ctx.alloc(1 << 8)
    ctx.evaluate(co, 8, dom, tw, out)
    et, eit = oracle.precompute_twiddles(dom.initial_index, 7)
    ev = oracle.evaluate(coeffs, dom.initial_index, 8, et)
    assert np.array_equal(out.download(), ev)
    ctx.interpolate(out, dom, tw)
    assert np.array_equal(out.download(), coeffs)
    co.free()
    out.free()
    tw.destroy()


def test_edge_values_canonical(ctx):
    """All-zero, all-(P-1) and alternating columns stay canonical and exact."""
    n = 13
    root = sh.half_odds(n)
    tw = ctx.precompute_twiddles(root)
    et, _ = oracle.precompute_twiddles(root.initial_index, n)
    for col in [np.zeros(1 << n, U), np.full(1 << n, P - 1, U), np.tile(arr([0, P - 1]), 1 << (n - 1))]:
        co = ctx.to_device(col)
        out = ctx.alloc(1 << (n + 1))
        ctx.evaluate(co, n, sh.canonic_domain(n + 1), tw, out)
        got = out.download()
        assert int(got.max()) < P
        assert np.array_equal(got, oracle.evaluate(col, root.initial_index, n + 1, et))
        co.free()
        out.free()
    tw.destroy()


def test_batched_columns_mixed_sizes(ctx):
    """interpolate_columns + evaluate_polynomials over columns of several sizes."""
    logs = [3, 5, 9, 12, 12, 14, 16, 5, 13]
    b = 2
    root = sh.half_odds(max(logs) + b - 1)
    tw = ctx.precompute_twiddles(root)
    et, eit = oracle.precompute_twiddles(root.initial_index, root.log_size)
    evs = [rand(1 << l, 9, i) for i, l in enumerate(logs)]
    ds = [ctx.to_device(e) for e in evs]
    ctx.interpolate_columns(ds, logs, tw)
    coeffs = [e.copy() for e in evs]
    oracle.interpolate_columns(coeffs, eit)
    for d, c in zip(ds, coeffs):
        assert np.array_equal(d.download(), c)
    outs = [ctx.alloc(1 << (l + b)) for l in logs]
    ctx.evaluate_polynomials(ds, logs, b, tw, outs)
    exp = oracle.evaluate_polynomials(coeffs, b, et)
    for o, e in zip(outs, exp):
        assert np.array_equal(o.download(), e)
    for x in ds + outs:
        x.free()
    tw.destroy()


def test_domain_not_in_tree_is_rejected(ctx):
    """A domain that is not a doubling of the twiddle root is an EINVAL, not a launch."""
    tw = ctx.precompute_twiddles(sh.half_odds(10))
    d = ctx.alloc(1 << 8)
    with pytest.raises(sh.StwoHipError) as e:
        ctx.interpolate(d, sh.Coset(12345, 7), tw)
    assert e.value.code == -1
    with pytest.raises(sh.StwoHipError):
        ctx.interpolate(d, sh.canonic_domain(13), tw)  # larger than the tree
    d.free()
    tw.destroy()


# ------------------------------------------------------------------ Merkle --

def test_merkle_golden(ctx, golden):
    for case in golden("merkle"):
        cols = [ctx.to_device(arr(c)) for c in case["cols"]]
        shape = case["shape"]
        L = max(shape)
        layers = ctx.alloc(sh.tree_words(L))
        ctx.merkle_commit(cols, shape, layers)
        got = layers.download()
        assert got[:8].tolist() == case["mixed_root"], shape
        assert got[sh.layer_offset(L): sh.layer_offset(L + 1)].reshape(-1, 8).tolist() == case["mixed_leaf_layer"]
        if case["lifted_root"] is not None:
            ctx.lifted_commit(cols, shape, L, layers)
            assert layers.download(8).tolist() == case["lifted_root"], shape
        layers.free()
        for c in cols:
            c.free()


@pytest.mark.parametrize("mode", [sh.HASH_BLAKE2S, sh.HASH_BLAKE2S_M31])
def test_commit_on_layer_vs_oracle(ctx, mode):
    # > 16 columns up to 2^14 nodes: a quad per node, the columns staged
    # through LDS 240 at a time (one, two and three chunks, partial last
    # chunks, fewer than 64 nodes); at 2^15: a lane per node with the loads
    # one block ahead; otherwise a lane per node.
    for log_size, n_cols, with_prev in [(10, 0, True), (10, 1, False), (11, 16, False), (9, 17, True),
                                        (12, 33, True), (3, 5, True), (0, 2, True), (14, 40, False),
                                        (1, 18, False), (15, 20, True), (16, 17, False), (11, 241, True),
                                        (6, 480, False), (13, 250, False), (5, 300, True), (8, 721, True)]:
        cols = [rand(1 << log_size, 13, 100 * log_size + c) for c in range(n_cols)]
        prev = rand(16 << log_size, 17, log_size) if with_prev else None
        dcols = [ctx.to_device(c) for c in cols]
        dprev = ctx.to_device(prev) if with_prev else None
        out = ctx.alloc(8 << log_size)
        ctx.commit_on_layer(log_size, dprev, dcols, out, mode)
        exp = oracle.commit_on_layer(log_size, prev, cols, m31_output=bool(mode))
        assert np.array_equal(out.download(), exp), (log_size, n_cols, with_prev)
        for x in dcols + [out] + ([dprev] if dprev else []):
            x.free()


@pytest.mark.parametrize("mode", [sh.HASH_BLAKE2S, sh.HASH_BLAKE2S_M31])
def test_mixed_degree_commit_vs_oracle(ctx, mode):
    """MerkleProver::commit over columns injected at many layers (config-3 shape, small)."""
    logs = [14, 3, 9, 14, 11, 11, 2, 7, 13, 9, 1, 14]
    cols = [rand(1 << l, 19, i) for i, l in enumerate(logs)]
    dcols = [ctx.to_device(c) for c in cols]
    layers = ctx.alloc(sh.tree_words(max(logs)))
    ctx.merkle_commit(dcols, logs, layers, mode)
    exp = oracle.merkle_commit(cols, m31_output=bool(mode))
    assert np.array_equal(layers.download(), exp)
    layers.free()
    for d in dcols:
        d.free()


def test_empty_commitments(ctx):
    import hashlib
    e = np.frombuffer(hashlib.blake2s(b"").digest(), dtype="<u4")
    out = ctx.alloc(8)
    ctx.merkle_commit([], [], out)
    assert np.array_equal(out.download(), e)
    ctx.lifted_build_leaves([], [], 0, out)
    assert np.array_equal(out.download(), e)
    out.free()


@pytest.mark.parametrize("mode", [sh.HASH_BLAKE2S, sh.HASH_BLAKE2S_M31])
def test_lifted_vs_oracle(ctx, mode):
    """metal_blake2s_merkle_* (tests/metal_blake2s_channel_cpu_bridge.rs:56-110): lifted leaves
    of uneven columns, 32 equal columns, and a large next layer."""
    for logs, L in [([3, 4, 4], 4), ([8] * 32, 8), ([12, 5, 9, 12, 1, 7] * 3, 13)]:
        cols = [rand(1 << l, 29, 50 * i + l) for i, l in enumerate(logs)]
        dcols = [ctx.to_device(c) for c in cols]
        layers = ctx.alloc(sh.tree_words(L))
        ctx.lifted_commit(dcols, logs, L, layers, mode)
        exp = oracle.lifted_commit(cols, L, m31_output=bool(mode))
        assert np.array_equal(layers.download(), exp), logs
        layers.free()
        for d in dcols:
            d.free()
    prev = rand(8 << 12, 31, 0)
    dp = ctx.to_device(prev)
    out = ctx.alloc(8 << 11)
    ctx.lifted_build_next_layer(dp, 11, out, mode)
    exp = np.zeros(8 << 11, dtype=U)
    oracle.lib().or_lifted_build_next_layer(oracle._p(prev), 11, oracle._p(exp), int(mode))
    assert np.array_equal(out.download(), exp)
    dp.free()
    out.free()


def test_lifted_reference_kat(ctx):
    """test_lifted_merkle_leaves (STWO/prover/vcs_lifted/prover.rs:247-275)."""
    import hashlib
    cols = [np.arange(1 << i, dtype=U) for i in range(2, 5)]
    d = [ctx.to_device(c) for c in cols]
    out = ctx.alloc(8 << 4)
    ctx.lifted_build_leaves(d, [2, 3, 4], 4, out)
    leaves = out.download().reshape(-1, 8)

    def hw(vals):
        return np.frombuffer(hashlib.blake2s(b"".join(int(v).to_bytes(4, "little") for v in vals)).digest(), "<u4")

    assert np.array_equal(leaves[0], hw([0, 0, 0]))
    assert np.array_equal(leaves[4], hw([0, 2, 4]))
    assert np.array_equal(leaves[15], hw([3, 7, 15]))
    out.free()
    for x in d:
        x.free()


# -------------------------------------------------------------------- folds --

def test_folds_golden(ctx, golden):
    g = golden("folds")
    for case in g["circle"]:
        N = case["log_size"]
        if N < 3:
            continue
        tw = ctx.precompute_twiddles(sh.half_odds(N - 1))
        src = [ctx.to_device(arr([v[k] for v in case["src"]])) for k in range(4)]
        dst = [ctx.to_device(arr([v[k] for v in case["prev"]])) for k in range(4)]
        ctx.fold_circle_into_line(dst, src, sh.half_odds(N - 1), case["alpha"], tw)
        got = np.stack([d.download() for d in dst], axis=1).tolist()
        assert got == case["dst"], N
        for x in src + dst:
            x.free()
        tw.destroy()
    for case in g["line"]:
        n, step = case["log_size"], case["fold_step"]
        tw = ctx.precompute_twiddles(sh.Coset(case["line_initial"], n))
        src = [ctx.to_device(arr([v[k] for v in case["src"]])) for k in range(4)]
        dst = [ctx.alloc(1 << (n - step)) for _ in range(4)]
        ctx.fold_line(src, sh.Coset(case["line_initial"], n), case["alpha"], step, tw, dst)
        got = np.stack([d.download() for d in dst], axis=1).tolist()
        assert got == case["dst"], (n, step)
        for x in src + dst:
            x.free()
        tw.destroy()


def test_fri_first_layer_fold_reference_shape(ctx):
    """metal_fri_first_layer_fold_matches_cpu_backend (tests/metal_fri_first_layer_fold.rs:17):
    log 10 canonic, alpha = (1, 3, 5, 7), values 4i + k."""
    N = 10
    tw = ctx.precompute_twiddles(sh.half_odds(N - 1))
    src = [np.array([4 * i + k for i in range(1 << N)], dtype=U) for k in range(4)]
    dsrc = [ctx.to_device(s) for s in src]
    dst = [ctx.zeros(1 << (N - 1)) for _ in range(4)]
    ctx.fold_circle_into_line(dst, dsrc, sh.half_odds(N - 1), [1, 3, 5, 7], tw)
    exp = [np.zeros(1 << (N - 1), U) for _ in range(4)]
    oracle.fold_circle_into_line(exp, src, oracle.half_odds_initial(N - 1), N, [1, 3, 5, 7])
    for d, e in zip(dst, exp):
        assert np.array_equal(d.download(), e)
    for x in dsrc + dst:
        x.free()
    tw.destroy()


# (5, 5): the last of the two launches (4 steps, then 1) folds 2 values into 1.
@pytest.mark.parametrize("n,step", [(9, 2), (12, 1), (16, 3), (14, 4), (13, 6), (20, 1), (5, 5), (2, 1)])
def test_fold_line_vs_oracle(ctx, n, step):
    """Includes metal_line_fold_matches_cpu_backend's shape (tests/metal_line_fold.rs:12): log 9,
    fold_step 2, alpha (2, 4, 6, 8), values 8i + k."""
    dom = sh.half_odds(n)
    tw = ctx.precompute_twiddles(dom)
    if n == 9:
        src = [np.array([8 * i + k for i in range(1 << n)], dtype=U) for k in range(4)]
        alpha = [2, 4, 6, 8]
    else:
        src = [rand(1 << n, 37, 10 * n + k) for k in range(4)]
        alpha = [int(x) for x in rand(4, 41, n)]
    dsrc = [ctx.to_device(s) for s in src]
    dst = [ctx.alloc(1 << (n - step)) for _ in range(4)]
    ctx.fold_line(dsrc, dom, alpha, step, tw, dst)
    exp = oracle.fold_line(src, dom.initial_index, n, alpha, step)
    for d, e in zip(dst, exp):
        assert np.array_equal(d.download(), e)
    for x in dsrc + dst:
        x.free()
    tw.destroy()


@pytest.mark.parametrize("N", [3, 11, 18])
def test_fold_circle_accumulate_vs_oracle(ctx, N):
    tw = ctx.precompute_twiddles(sh.half_odds(N - 1))
    src = [rand(1 << N, 43, 10 * N + k) for k in range(4)]
    prev = [rand(1 << (N - 1), 47, 10 * N + k) for k in range(4)]
    alpha = [int(x) for x in rand(4, 53, N)]
    dsrc = [ctx.to_device(s) for s in src]
    dst = [ctx.to_device(p) for p in prev]
    ctx.fold_circle_into_line(dst, dsrc, sh.half_odds(N - 1), alpha, tw)
    exp = [p.copy() for p in prev]
    oracle.fold_circle_into_line(exp, src, oracle.half_odds_initial(N - 1), N, alpha)
    for d, e in zip(dst, exp):
        assert np.array_equal(d.download(), e)
    for x in dsrc + dst:
        x.free()
    tw.destroy()


# ---------------------------------------------------------------------- FRI --

def test_fri_commit_golden(ctx, golden):
    """FriProver::commit transcripts; includes metal_bounded_fri_prover's config
    (tests/metal_fri_prover.rs:24-26: log 10, FriConfig::new(3, 2, 3, 2))."""
    for case in golden("fri"):
        lb, lld, nq, step = case["config"]
        n = case["log_size"]
        tw = ctx.precompute_twiddles(sh.half_odds(n - 1))
        col = [ctx.to_device(arr([v[k] for v in case["evals"]])) for k in range(4)]
        ch = oracle.Channel()
        ch.mix_root(arr([int.from_bytes(bytes(case["prior_root"][i:i + 4]), "little") for i in range(0, 32, 4)]))
        roots, last, digest, nd = ctx.fri_commit(ch.digest_bytes(), ch.n_draws, sh.FriConfig(lb, lld, nq, step), col,
                                                  n, tw)
        assert roots.tolist() == case["roots"], n
        assert last.tolist() == case["last_layer_poly"], n
        assert list(digest) == case["final_digest"], n
        for x in col:
            x.free()
        tw.destroy()


def _low_degree_qm31_eval(ctx, log_size, log_blowup, seed, tw):
    """4 coordinate columns: random polys of degree < 2^(log_size - log_blowup) evaluated on CanonicCoset(log_size)."""
    n = log_size - log_blowup
    cols = []
    for k in range(4):
        co = ctx.alloc(1 << n)
        ctx.fill_random(co, 1 << n, seed, k)
        out = ctx.alloc(1 << log_size)
        ctx.evaluate(co, n, sh.canonic_domain(log_size), tw, out)
        co.free()
        cols.append(out)
    return cols


# 2^21: layers above 2^19 take node-level launches before the top-levels
# kernel; fold steps of 5 and 6: the unfused fold path (more than 4 steps),
# with layers below 2^8 committed outside the single-workgroup tail.  2^22
# with a last layer of 2^4: the benchmark's FRI chain at its full size.
@pytest.mark.parametrize("log_size,cfg", [(14, (1, 3, 3, 1)), (16, (2, 0, 3, 3)), (12, (1, 2, 3, 2)),
                                          (21, (1, 0, 3, 1)), (13, (1, 1, 3, 5)), (11, (2, 0, 3, 6)),
                                          (22, (1, 3, 3, 1))])
def test_fri_commit_vs_oracle(ctx, log_size, cfg):
    lb, lld, nq, step = cfg
    tw = ctx.precompute_twiddles(sh.half_odds(log_size - 1))
    col = _low_degree_qm31_eval(ctx, log_size, lb, 61, tw)
    host = [c.download() for c in col]
    ch = oracle.Channel()
    ch.mix_u64(0xDEADBEEF)
    roots, last, digest, nd = ctx.fri_commit(ch.digest_bytes(), ch.n_draws, sh.FriConfig(lb, lld, nq, step), col,
                                              log_size, tw)
    eroots, elast, _ = oracle.fri_commit(ch, oracle.FriConfig(lb, lld, nq, step), host, log_size)
    assert np.array_equal(roots, eroots)
    assert np.array_equal(last, elast)
    assert digest == ch.digest_bytes() and nd == ch.n_draws
    for c in col:
        c.free()
    tw.destroy()


def test_fri_rejects_high_degree(ctx):
    n = 10
    tw = ctx.precompute_twiddles(sh.half_odds(n - 1))
    col = [ctx.to_device(rand(1 << n, 67, k)) for k in range(4)]
    with pytest.raises(sh.StwoHipError) as e:
        ctx.fri_commit(bytes(32), 0, sh.FriConfig(1, 1, 3, 1), col, n, tw)
    assert e.value.code == -4
    for c in col:
        c.free()
    tw.destroy()


def test_fri_rejects_oversized_fold_step(ctx):
    """line_fold_step above 28 is rejected before any launch (a step near
    2^32 would wrap the schedule's layer arithmetic)."""
    n = 10
    tw = ctx.precompute_twiddles(sh.half_odds(n - 1))
    col = [ctx.to_device(rand(1 << n, 68, k)) for k in range(4)]
    for step in (29, 1 << 31, 0xFFFFFFFF):
        with pytest.raises(sh.StwoHipError) as e:
            ctx.fri_commit(bytes(32), 0, sh.FriConfig(1, 1, 3, step), col, n, tw)
        assert e.value.code == -1
    for c in col:
        c.free()
    tw.destroy()


def test_table_ring_wraps_and_large_tables(ctx):
    """Per-call tables through the pinned ring: eval_at_point calls whose
    column and work tables (16 B per column, 16 B per 2^11-coefficient block
    of each run of same-size columns) take about 3 MiB of the 8 MiB ring
    wrap it repeatedly, and a call whose tables exceed half of it (7.7 MB)
    takes the synchronous scratch path; every value must equal the value of
    the same column evaluated alone (tiny tables)."""
    a, b = ctx.alloc(1 << 18), ctx.alloc(1 << 17)
    ctx.fill_random(a, 1 << 18, 31, 0)
    ctx.fill_random(b, 1 << 17, 31, 1)
    pt = secure_point_for_tables()
    va = ctx.eval_at_point([a], [18], pt)[0]
    vb = ctx.eval_at_point([b], [17], pt)[0]
    for n_cols in (2000, 1600, 2000, 1200, 2000, 5000, 300):
        cols = [a if i % 2 == 0 else b for i in range(n_cols)]
        logs = [18 if i % 2 == 0 else 17 for i in range(n_cols)]
        got = ctx.eval_at_point(cols, logs, pt)
        assert np.array_equal(got[0::2], np.tile(va, (len(got[0::2]), 1)))
        assert np.array_equal(got[1::2], np.tile(vb, (len(got[1::2]), 1)))
    a.free()
    b.free()


def secure_point_for_tables():
    # any point works for the equality check; a fixed pair of QM31 coordinates
    return ([11, 22, 33, 44], [55, 66, 77, 88])


# ------------------------------------------------- first benchmark workload --

def test_config1_lde_merkle_full(ctx):
    """16 columns of 2^16 (seed 1): interpolate, evaluate at blowup 2 on
    CanonicCoset(17), mixed-degree Blake2s commit; every evaluation and every
    tree layer compared with the oracle."""
    n_cols, n, b = 16, 16, 1
    root = sh.half_odds(n + b - 1)
    tw = ctx.precompute_twiddles(root)
    et, eit = oracle.precompute_twiddles(root.initial_index, root.log_size)
    evs = [rand(1 << n, 1, c) for c in range(n_cols)]
    d = [ctx.to_device(e) for e in evs]
    ctx.interpolate_columns(d, [n] * n_cols, tw)
    outs = [ctx.alloc(1 << (n + b)) for _ in range(n_cols)]
    ctx.evaluate_polynomials(d, [n] * n_cols, b, tw, outs)
    layers = ctx.alloc(sh.tree_words(n + b))
    ctx.merkle_commit(outs, [n + b] * n_cols, layers)
    coeffs = [e.copy() for e in evs]
    oracle.interpolate_columns(coeffs, eit)
    exp_evals = oracle.evaluate_polynomials(coeffs, b, et)
    for dd, c in zip(d, coeffs):
        assert np.array_equal(dd.download(), c)
    for o, e in zip(outs, exp_evals):
        assert np.array_equal(o.download(), e)
    exp_layers = oracle.merkle_commit(exp_evals)
    assert np.array_equal(layers.download(), exp_layers)
    for x in d + outs + [layers]:
        x.free()
    tw.destroy()


# ---------------------------------------------------------- batched commitment --

def _geometric_logs(n_cols, lo, hi, seed):
    """log sizes lo + j with P(j) proportional to 0.7^j (the uneven-column workload's law)."""
    r = np.random.default_rng(seed)
    w = np.array([0.7 ** j for j in range(hi - lo + 1)])
    return [int(lo + j) for j in r.choice(len(w), size=n_cols, p=w / w.sum())]


@pytest.mark.parametrize("tree_kind,use_graph,mode", [(sh.TREE_MIXED, True, sh.HASH_BLAKE2S),
                                                      (sh.TREE_MIXED, False, sh.HASH_BLAKE2S_M31),
                                                      (sh.TREE_LIFTED, True, sh.HASH_BLAKE2S)])
def test_batch_lde_commit_vs_oracle(ctx, tree_kind, use_graph, mode):
    """Uneven columns (log 4..17) -> interpolate + evaluate (blowup 2) + one tree, three
    replays of the same plan (graph replays with eager work and syncs between them),
    each checked bit-exactly against the oracle."""
    logs = _geometric_logs(60, 4, 17, 5) + [17, 4, 13]
    b = 1
    max_log = max(logs) + b
    root = sh.half_odds(max_log - 1)
    tw = ctx.precompute_twiddles(root)
    et, eit = oracle.precompute_twiddles(root.initial_index, root.log_size)
    evs = [rand(1 << l, 83, i) for i, l in enumerate(logs)]
    cols = [ctx.to_device(e) for e in evs]
    outs = [ctx.alloc(1 << (l + b)) for l in logs]
    layers = ctx.alloc(sh.tree_words(max_log))
    plan = ctx.batch_create(cols, logs, b, tw, outs, layers, tree_kind, mode, use_graph)
    coeffs = [e.copy() for e in evs]
    oracle.interpolate_columns(coeffs, eit)
    exp_evals = oracle.evaluate_polynomials(coeffs, b, et)
    if tree_kind == sh.TREE_MIXED:
        exp_tree = oracle.merkle_commit(exp_evals, m31_output=bool(mode))
    else:
        exp_tree = oracle.lifted_commit(exp_evals, max_log, m31_output=bool(mode))
    for rep in range(3):
        for c, e in zip(cols, evs):   # eager work between replays: restore the inputs
            c.upload(e)
        layers.upload(np.zeros(sh.tree_words(max_log), dtype=U))
        plan.run()
        ctx.sync()
        for c, e in zip(cols, coeffs):
            assert np.array_equal(c.download(), e), ("coeffs", rep)
        for o, e in zip(outs, exp_evals):
            assert np.array_equal(o.download(), e), ("evals", rep)
        assert np.array_equal(layers.download(), exp_tree), ("tree", rep)
    plan.destroy()
    for x in cols + outs + [layers]:
        x.free()
    tw.destroy()


def _lde_expect(evs, logs, b, tw_root, tree_kind, mode):
    et, eit = oracle.precompute_twiddles(tw_root.initial_index, tw_root.log_size)
    coeffs = [e.copy() for e in evs]
    oracle.interpolate_columns(coeffs, eit)
    exp_evals = oracle.evaluate_polynomials(coeffs, b, et)
    if tree_kind == sh.TREE_MIXED:
        tree = oracle.merkle_commit(exp_evals, m31_output=bool(mode))
    else:
        tree = oracle.lifted_commit(exp_evals, max(logs) + b, m31_output=bool(mode))
    return coeffs, exp_evals, tree


@pytest.mark.parametrize("tree_kind,mode", [(sh.TREE_MIXED, sh.HASH_BLAKE2S), (sh.TREE_LIFTED, sh.HASH_BLAKE2S_M31)])
def test_lde_commit_per_column_vs_oracle(ctx, tree_kind, mode):
    """The plain per-column entry point: coefficients, evaluations and tree
    bit-exact against the oracle (uneven columns log 4..16, blowup 2)."""
    logs = _geometric_logs(40, 4, 16, 9) + [16, 4]
    b = 1
    root = sh.half_odds(max(logs) + b - 1)
    tw = ctx.precompute_twiddles(root)
    evs = [rand(1 << l, 91, i) for i, l in enumerate(logs)]
    cols = [ctx.to_device(e) for e in evs]
    outs = [ctx.alloc(1 << (l + b)) for l in logs]
    layers = ctx.alloc(sh.tree_words(max(logs) + b))
    ctx.lde_commit_per_column(cols, logs, b, tw, outs, layers, tree_kind, mode)
    ctx.sync()
    coeffs, exp_evals, tree = _lde_expect(evs, logs, b, root, tree_kind, mode)
    for c, e in zip(cols, coeffs):
        assert np.array_equal(c.download(), e)
    for o, e in zip(outs, exp_evals):
        assert np.array_equal(o.download(), e)
    assert np.array_equal(layers.download(), tree)
    for x in cols + outs + [layers]:
        x.free()
    tw.destroy()


@pytest.mark.parametrize("use_graph", [True, False])
def test_batch_multi_sets_vs_oracle(ctx, use_graph):
    """Three column sets in one plan -- blowups 2, 4 and 2, mixed and lifted
    trees, shared log sizes across sets (their FFT launches are shared) --
    coefficients dropped; every set's evaluations and tree bit-exact against
    the oracle on two replays."""
    specs = [(_geometric_logs(20, 5, 14, 21) + [14], 1, sh.TREE_MIXED, sh.HASH_BLAKE2S),
             (_geometric_logs(12, 5, 12, 22) + [12, 5], 2, sh.TREE_LIFTED, sh.HASH_BLAKE2S),
             ([14, 9, 9, 6], 1, sh.TREE_MIXED, sh.HASH_BLAKE2S_M31)]
    max_log = max(max(l) + b for l, b, _, _ in specs)
    root = sh.half_odds(max_log - 1)
    tw = ctx.precompute_twiddles(root)
    sets, bufs, expect = [], [], []
    for si, (logs, b, kind, mode) in enumerate(specs):
        evs = [rand(1 << l, 300 + si, i) for i, l in enumerate(logs)]
        cols = [ctx.to_device(e) for e in evs]
        outs = [ctx.alloc(1 << (l + b)) for l in logs]
        layers = ctx.alloc(sh.tree_words(max(logs) + b))
        sets.append((cols, logs, b, outs, layers, kind, mode))
        bufs += cols + outs + [layers]
        expect.append((evs, _lde_expect(evs, logs, b, root, kind, mode)))
    plan = ctx.batch_create_multi(sets, tw, use_graph=use_graph, drop_coeffs=True)
    for rep in range(2):
        for (cols, logs, b, outs, layers, kind, mode), (evs, _) in zip(sets, expect):
            for c, e in zip(cols, evs):
                c.upload(e)
            layers.upload(np.zeros(sh.tree_words(max(logs) + b), dtype=U))
        plan.run()
        ctx.sync()
        for si, ((cols, logs, b, outs, layers, kind, mode), (_, (_, exp_evals, tree))) in enumerate(zip(sets, expect)):
            for o, e in zip(outs, exp_evals):
                assert np.array_equal(o.download(), e), ("evals", si, rep)
            assert np.array_equal(layers.download(), tree), ("tree", si, rep)
    plan.destroy()
    for x in bufs:
        x.free()
    tw.destroy()


@pytest.mark.parametrize("b", [1, 2])
def test_batch_ifft_then_rfft_sets(ctx, b):
    """A composition-polynomial plan: an IFFT-only set (4 columns of log 13,
    interpolated in place) and an RFFT-only set evaluating their halves
    (coefficients [0, 2^12) and [2^12, 2^13) of each, 8 columns of log 12) at
    the blowup into a lifted tree; plus an LDE set of two other columns in the
    same plan.  Coefficients, evaluations and both trees bit-exact against the
    oracle, on two graph replays."""
    C = 13
    n = C - 1
    root = sh.half_odds(C + b - 1)
    tw = ctx.precompute_twiddles(root)
    et, eit = oracle.precompute_twiddles(root.initial_index, root.log_size)
    comp_ev = [rand(1 << C, 400, k) for k in range(4)]
    comp = [ctx.to_device(e) for e in comp_ev]
    halves = [c.at(0) for c in comp] + [c.at(1 << n) for c in comp]
    hout = [ctx.alloc(1 << (n + b)) for _ in range(8)]
    htree = ctx.alloc(sh.tree_words(n + b))
    lde_ev = [rand(1 << 11, 401, k) for k in range(2)]
    lde = [ctx.to_device(e) for e in lde_ev]
    lout = [ctx.alloc(1 << (11 + b)) for _ in range(2)]
    ltree = ctx.alloc(sh.tree_words(11 + b))
    sets = [(comp, [C] * 4, 0, None, None, 0, 0, sh.BATCH_IFFT),
            (halves, [n] * 8, b, hout, htree, sh.TREE_LIFTED, sh.HASH_BLAKE2S, sh.BATCH_RFFT),
            (lde, [11, 11], b, lout, ltree, sh.TREE_MIXED, sh.HASH_BLAKE2S, sh.BATCH_LDE)]
    plan = ctx.batch_create_multi(sets, tw, use_graph=True)
    coeffs = [e.copy() for e in comp_ev]
    oracle.interpolate_columns(coeffs, eit)
    hc = [c[: 1 << n].copy() for c in coeffs] + [c[1 << n:].copy() for c in coeffs]
    exp_h = oracle.evaluate_polynomials(hc, b, et)
    exp_htree = oracle.lifted_commit(exp_h, n + b)
    _, exp_l, exp_ltree = _lde_expect(lde_ev, [11, 11], b, root, sh.TREE_MIXED, sh.HASH_BLAKE2S)
    for rep in range(2):
        for d, e in zip(comp + lde, comp_ev + lde_ev):
            d.upload(e)
        plan.run()
        ctx.sync()
        for d, e in zip(comp, coeffs):
            assert np.array_equal(d.download(), e), ("coeffs", rep)
        for o, e in zip(hout + lout, exp_h + exp_l):
            assert np.array_equal(o.download(), e), ("evals", rep)
        assert np.array_equal(htree.download(), exp_htree), ("halves tree", rep)
        assert np.array_equal(ltree.download(), exp_ltree), ("lde tree", rep)
    plan.destroy()
    for x in comp + hout + [htree] + lde + lout + [ltree]:
        x.free()
    tw.destroy()


def test_unknown_hash_mode_and_wrapping_blowup_rejected(ctx):
    """Hash modes other than Blake2s / Blake2s-M31 and a log blowup whose sum
    with the column's log size wraps in 32 bits fail with EINVAL before any
    launch."""
    tw = ctx.precompute_twiddles(sh.half_odds(8))
    c = ctx.to_device(np.arange(16, dtype=np.uint32))
    o = ctx.alloc(32)
    layers = ctx.alloc(sh.tree_words(5))
    try:
        for call in (lambda: ctx.commit_on_layer(4, None, [c], o, mode=2),
                     lambda: ctx.merkle_commit([c], [4], layers, mode=-1),
                     lambda: ctx.evaluate_polynomials([c], [4], (1 << 32) - 1, tw, [o]),
                     lambda: ctx.batch_create([c], [4], 1, tw, [o], layers, mode=3),
                     lambda: ctx.batch_create([c], [4], (1 << 32) - 2, tw, [o], layers)):
            with pytest.raises(sh.StwoHipError) as e:
                call()
            assert e.value.code == -1
    finally:
        for x in (c, o, layers):
            x.free()
        tw.destroy()


def test_batch_rejects_small_columns(ctx):
    tw = ctx.precompute_twiddles(sh.half_odds(8))
    c = ctx.alloc(8)
    o = ctx.alloc(16)
    layers = ctx.alloc(sh.tree_words(4))
    with pytest.raises(sh.StwoHipError):
        ctx.batch_create([c], [3], 1, tw, [o], layers)
    for x in (c, o, layers):
        x.free()
    tw.destroy()

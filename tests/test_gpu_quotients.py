"""GPU parity of stwo_hip_compute_fri_quotients against the oracle's
restatement of compute_fri_quotients (STWO/prover/pcs/quotient_ops.rs:131-232),
bit-exact, plus the low-degree property the reference's
test_quotients_are_low_degree checks, on the library's own output."""
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "oracle"))
sys.path.insert(0, HERE)
import oracle as O  # noqa: E402
import stwo_hip as sh  # noqa: E402
from test_quotients_oracle import ALPHA, domain_eval, eval_at, interp, lifted_poly, secure_point  # noqa: E402

pytestmark = pytest.mark.gpu
U = np.uint32
P = O.P


@pytest.fixture(scope="module")
def ctx():
    c = sh.Context(0)
    yield c
    c.close()


def run(ctx, cols, samples, L):
    d = [ctx.to_device(c) for c in cols]
    out = [ctx.alloc(1 << L) for _ in range(4)]
    ctx.compute_fri_quotients(d, [len(c).bit_length() - 1 for c in cols], samples, ALPHA, L, out)
    res = [o.download() for o in out]
    for x in d + out:
        x.free()
    return res


def test_quotients_reference_shape(ctx):
    L = 7
    coeffs = np.arange(8, dtype=U)
    ev = domain_eval(coeffs, L)
    pts = [secure_point(k) for k in (0x1234567, 0xABCDEF01)]
    samples = [[(p[0], p[1], eval_at(coeffs, p)) for p in pts]]
    got = run(ctx, [ev], samples, L)
    want = O.compute_fri_quotients([ev], samples, ALPHA, L)
    for g, w in zip(got, want):
        assert np.array_equal(g, w)
    for g in got:
        assert not interp(g, L)[(1 << 3) - 1:].any()


@pytest.mark.parametrize("seed", range(2))
def test_quotients_lifted_mixed(ctx, seed):
    rng = np.random.default_rng(seed)
    L, b = 8, 2
    logs = [8, 5, 6, 8, 4, 7, 1]
    cols, samples = [], []
    z = [secure_point(int(k)) for k in rng.integers(1, 1 << 62, 3)]
    for i, lg in enumerate(logs):
        if lg > b:
            c = rng.integers(0, P, 1 << (lg - b), dtype=np.uint64).astype(U)
            ev = domain_eval(c, lg)
        else:
            ev = rng.integers(0, P, 1 << lg, dtype=np.uint64).astype(U)
        cols.append(ev)
        lp = lifted_poly(ev, L)
        n = i % 4  # 0..3 samples, 2 gets the periodicity sample
        samples.append([(p[0], p[1], eval_at(lp, p)) for p in z[:n]])
    got = run(ctx, cols, samples, L)
    want = O.compute_fri_quotients(cols, samples, ALPHA, L)
    for g, w in zip(got, want):
        assert np.array_equal(g, w)


def test_quotients_wide_vs_oracle(ctx):
    """96 columns of log 14..16 lifted to 16, two samples each (an OODS point
    and its predecessor, as a trace with a mask of two rows has)."""
    rng = np.random.default_rng(11)
    L = 16
    logs = [14 + (i % 3) for i in range(96)]
    cols = [rng.integers(0, P, 1 << lg, dtype=np.uint64).astype(U) for lg in logs]
    z1, z2 = secure_point(0x55AA55AA), secure_point(0x1234)
    samples = []
    for i in range(96):
        v1 = tuple(int(x) for x in rng.integers(0, P, 4))
        v2 = tuple(int(x) for x in rng.integers(0, P, 4))
        samples.append([(z1[0], z1[1], v1), (z2[0], z2[1], v2)] if i % 5 else [(z1[0], z1[1], v1)])
    got = run(ctx, cols, samples, L)
    want = O.compute_fri_quotients(cols, samples, ALPHA, L)
    for g, w in zip(got, want):
        assert np.array_equal(g, w)


def test_quotients_reject_bad_columns(ctx):
    d = ctx.to_device(np.zeros(16, U))
    out = [ctx.alloc(1 << 3) for _ in range(4)]
    with pytest.raises(sh.StwoHipError):
        ctx.compute_fri_quotients([d], [4], [[]], ALPHA, 3, out)  # column longer than the lifting size
    for x in [d] + out:
        x.free()


def test_eval_at_point_vs_oracle(ctx):
    """stwo_hip_eval_at_point vs the oracle (pinned by the FFT on domain
    points and the reference's 4-coefficient KAT), logs 0..22, one call."""
    rng = np.random.default_rng(5)
    logs = [0, 1, 2, 3, 10, 11, 12, 13, 17, 22, 11, 4]
    cols = [rng.integers(0, P, 1 << lg, dtype=np.uint64).astype(U) for lg in logs]
    d = [ctx.to_device(c) for c in cols]
    for k in (0xDEADBEEF, 42):
        pt = secure_point(k)
        got = ctx.eval_at_point(d, logs, pt)
        for c, g in zip(cols, got):
            assert tuple(int(v) for v in g) == O.eval_at_point(c, pt)
    for x in d:
        x.free()


def test_eval_at_point_grouped_and_extreme_values(ctx):
    """Runs of same-size columns (grouped 8 to a workgroup, with a remainder)
    and coefficients all P - 1, which drive the 64-bit accumulation to its
    largest sums; bit-exact against the oracle."""
    rng = np.random.default_rng(9)
    logs = [12] * 19 + [3] * 9 + [11] * 2 + [14]
    cols = [rng.integers(0, P, 1 << lg, dtype=np.uint64).astype(U) for lg in logs]
    cols[3][:] = P - 1
    cols[20][:] = P - 1
    cols[-1][:] = P - 1
    d = [ctx.to_device(c) for c in cols]
    pt = secure_point(0x5EED)
    got = ctx.eval_at_point(d, logs, pt)
    for c, g in zip(cols, got):
        assert tuple(int(v) for v in g) == O.eval_at_point(c, pt)
    for x in d:
        x.free()


def test_eval_at_point_unaligned_columns(ctx):
    """Column pointers at word offsets 1, 3, 4 and 8 of a buffer (not 32-byte
    aligned: whole 2^11 blocks then take the per-word loads) beside aligned
    ones in the same workgroups; bit-exact against the oracle."""
    rng = np.random.default_rng(13)
    logs = [13, 13, 13, 13, 13, 12, 11]
    offs = [1, 0, 3, 4, 8, 0, 5]
    cols = [rng.integers(0, P, 1 << lg, dtype=np.uint64).astype(U) for lg in logs]
    bufs = []
    ptrs = []
    for c, o in zip(cols, offs):
        b = ctx.alloc(len(c) + 8)
        b.upload(c, offset=o)
        bufs.append(b)
        ptrs.append(b.at(o))
    pt = secure_point(0xA11)
    got = ctx.eval_at_point(ptrs, logs, pt)
    for c, g in zip(cols, got):
        assert tuple(int(v) for v in g) == O.eval_at_point(c, pt)
    for b in bufs:
        b.free()


@pytest.mark.parametrize("log_n,n_cols,offset", [(5, 8, 0), (12, 100, 0), (16, 256, 0), (9, 13, 1)])
def test_wide_fib_trace_and_constraints_vs_oracle(ctx, log_n, n_cols, offset):
    """Trace generation and the constraint quotients, bit-exact against the
    oracle (two rows per lane; with offset 1 the extended columns start one
    word into their buffers, one row per lane); the GPU's quotients are low
    degree (checked at log 5)."""
    rng = np.random.default_rng(log_n)
    a = rng.integers(0, P, 1 << log_n, dtype=np.uint64).astype(U)
    b = rng.integers(0, P, 1 << log_n, dtype=np.uint64).astype(U)
    da, db = ctx.to_device(a), ctx.to_device(b)
    cols = [ctx.alloc(1 << log_n) for _ in range(n_cols)]
    ctx.wide_fib_trace(cols, log_n, da, db)
    tr = O.wide_fib_trace(a, b, n_cols)
    for c, t in zip(cols, tr):
        assert np.array_equal(c.download(), t)
    # extend on the host oracle (the GPU FFT is checked elsewhere), then constrain on both
    tw, itw = O.precompute_twiddles(O.half_odds_initial(log_n), log_n)
    coeffs = [t.copy() for t in tr]
    O.interpolate_columns(coeffs, itw)
    ext = O.evaluate_polynomials(coeffs, 1, tw)
    ebufs = [ctx.alloc(len(e) + offset) for e in ext]
    for bf, e in zip(ebufs, ext):
        bf.upload(e, offset=offset)
    dext = [bf.at(offset) for bf in ebufs]
    base = [rng.integers(0, P, 2 << log_n, dtype=np.uint64).astype(U) for _ in range(4)]
    acc = [ctx.to_device(x) for x in base]
    ctx.wide_fib_constraint_quotients(dext, log_n, ALPHA, acc)
    want = O.wide_fib_constraint_quotients(ext, log_n, ALPHA, [x.copy() for x in base])
    got = [x.download() for x in acc]
    for g, w in zip(got, want):
        assert np.array_equal(g, w)
    if log_n == 5:
        zero = [ctx.to_device(np.zeros(2 << log_n, U)) for _ in range(4)]
        ctx.wide_fib_constraint_quotients(dext, log_n, ALPHA, zero)
        for z in zero:
            assert not interp(z.download(), log_n + 1)[(1 << log_n) + 1:].any()
            z.free()
    for x in [da, db] + cols + ebufs + acc:
        x.free()

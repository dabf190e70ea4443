"""stwo_hip_extend_evals (interpolate_columns + evaluate_polynomials with
the IFFT's last pass and the RFFT's first pass fused into one kernel)
against the oracle, bit-exact, over every pass geometry the fusion meets:
no fusion (n <= 12: the low pass alone), the fused pass on each
compile-time strided schedule (n = 13..16: one strided pass of 1..4 layers
on a 2^10 tile; 17, 18, 19: 5, 6, 7 layers; 20: the 8-layer pair; n >= 21,
where the fused pair is the top 8-layer pass), blowups 0..3, with and
without keeping the coefficients; and against the library's own two-call
path on the config-2 shape."""
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "oracle"))
import oracle  # noqa: E402
import stwo_hip as sh  # noqa: E402

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = sh.Context(0)
    yield c
    c.close()


SHAPES = [(4, 1, 3), (10, 2, 5), (12, 1, 4), (13, 1, 3), (14, 2, 2), (15, 1, 3), (16, 1, 4), (16, 0, 2), (17, 3, 2),
          (18, 1, 2), (19, 2, 2), (20, 2, 3), (21, 1, 2), (22, 1, 1), (23, 2, 1)]


@pytest.mark.parametrize("n,b,C", SHAPES)
@pytest.mark.parametrize("store", [False, True])
def test_extend_evals_vs_oracle(ctx, n, b, C, store):
    N = n + b
    print("extend_evals n=%d b=%d cols=%d store=%s" % (n, b, C, store), flush=True)
    root = max(N, 2) - 1
    tw = ctx.precompute_twiddles(sh.half_odds(root))
    etw, eitw = oracle.precompute_twiddles(oracle.half_odds_initial(root), root)
    evs = [oracle.fill_random(1 << n, 40 + n, c) for c in range(C)]
    if C > 1:
        evs[1][:] = oracle.P - 1
    cols = [ctx.to_device(e) for e in evs]
    outs = [ctx.alloc(1 << N) for _ in range(C)]
    ctx.extend_evals(cols, [n] * C, b, tw, outs, store_coeffs=store)
    coeffs = [e.copy() for e in evs]
    oracle.interpolate_columns(coeffs, eitw)
    want = oracle.evaluate_polynomials(coeffs, b, etw)
    for o, w in zip(outs, want):
        assert np.array_equal(o.download(), w)
    if store:
        for c, w in zip(cols, coeffs):
            assert np.array_equal(c.download(), w)
    for x in cols + outs:
        x.free()
    tw.destroy()


def test_extend_evals_mixed_sizes_equals_two_calls(ctx):
    """Columns of several sizes in one call (grouped by size) against
    interpolate_columns + evaluate_polynomials on copies."""
    logs = [20, 5, 13, 20, 16, 21, 12, 20]
    b = 2
    tw = ctx.precompute_twiddles(sh.half_odds(max(logs) + b - 1))
    src = []
    for i, lg in enumerate(logs):
        d = ctx.alloc(1 << lg)
        ctx.fill_random(d, 1 << lg, 77, i)
        src.append(d)
    a = [ctx.alloc(1 << lg) for lg in logs]
    c = [ctx.alloc(1 << lg) for lg in logs]
    for x, y, s in zip(a, c, src):
        ctx._check(sh.lib().stwo_hip_copy(ctx.h, x.ptr, s.ptr, s.n))
        ctx._check(sh.lib().stwo_hip_copy(ctx.h, y.ptr, s.ptr, s.n))
    oa = [ctx.alloc(1 << (lg + b)) for lg in logs]
    oc = [ctx.alloc(1 << (lg + b)) for lg in logs]
    ctx.extend_evals(a, logs, b, tw, oa, store_coeffs=True)
    ctx.interpolate_columns(c, logs, tw)
    ctx.evaluate_polynomials(c, logs, b, tw, oc)
    for x, y in zip(oa, oc):
        assert np.array_equal(x.download(), y.download())
    for x, y in zip(a, c):
        assert np.array_equal(x.download(), y.download())
    for x in src + a + c + oa + oc:
        x.free()
    tw.destroy()

"""The FFT autotuner: stwo_hip_fft_autotune times the planner's candidate
plans on the device and keeps the fastest per shape; every candidate plan
must compute the same transform.  Each candidate (low-pass tile 2^10..2^13,
strided passes of at most 5..8 layers, workgroup size) is forced through
stwo_hip_fft_tune_set on shapes that give it distinct plans and checked
bit-exactly against the oracle, through interpolate, evaluate, extend_evals
and a batched plan."""
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


CANDIDATES = [(t0, sm, wg) for t0 in (10, 11, 12, 13) for sm in (5, 6, 7, 8) for wg in (0, 9)]


@pytest.mark.parametrize("n,b", [(11, 1), (14, 2), (17, 1), (21, 1)])
def test_every_candidate_plan_is_exact(ctx, n, b):
    N = n + b
    C = 2
    tw = ctx.precompute_twiddles(sh.half_odds(N - 1))
    etw, eitw = oracle.precompute_twiddles(oracle.half_odds_initial(N - 1), N - 1)
    evs = [oracle.fill_random(1 << n, 90 + n, c) for c in range(C)]
    coeffs = [e.copy() for e in evs]
    oracle.interpolate_columns(coeffs, eitw)
    want = oracle.evaluate_polynomials(coeffs, b, etw)
    seen = set()
    for (t0, sm, wg) in CANDIDATES:
        key = (str(sh.fft_plan(n, n, True, t0, sm, wg)), str(sh.fft_plan(n, N, False, t0, sm, wg)))
        if key in seen:
            continue
        seen.add(key)
        print("candidate n=%d b=%d t0=%d smax=%d wg=%d" % (n, b, t0, sm, wg), flush=True)
        ctx.fft_tune_set([(n, n, 1, t0, sm, wg), (n, N, 0, t0, sm, wg)])
        cols = [ctx.to_device(e) for e in evs]
        outs = [ctx.alloc(1 << N) for _ in range(C)]
        ctx.interpolate_columns(cols, [n] * C, tw)
        for c, w in zip(cols, coeffs):
            assert np.array_equal(c.download(), w), (t0, sm, wg)
        ctx.evaluate_polynomials(cols, [n] * C, b, tw, outs)
        for o, w in zip(outs, want):
            assert np.array_equal(o.download(), w), (t0, sm, wg)
        for c, e in zip(cols, evs):
            c.upload(e)
        ctx.extend_evals(cols, [n] * C, b, tw, outs, store_coeffs=True)
        for o, w in zip(outs, want):
            assert np.array_equal(o.download(), w), (t0, sm, wg)
        for x in cols + outs:
            x.free()
    ctx.fft_tune_clear()
    assert len(seen) > 1
    tw.destroy()


def test_autotune_then_batched_plan(ctx):
    """Autotune a set of uneven columns, then the batched plan (which takes
    the tuned choices) gives the oracle's tree root and evaluations."""
    logs = [10, 10, 11, 12, 13, 13, 15, 17, 18, 20]
    b = 1
    max_log = max(logs) + b
    tw = ctx.precompute_twiddles(sh.half_odds(max_log - 1))
    rep = ctx.fft_autotune(logs, b, tw)
    assert len(rep) == 2 * len(set(logs))
    for (n, N, inv, t0, sm, wg, best_us, def_us) in rep:
        assert 10 <= t0 <= 13 and 5 <= sm <= 8 and wg in (0, 9)
        assert best_us <= def_us + 1
    evs = [oracle.fill_random(1 << lg, 123, i) for i, lg in enumerate(logs)]
    cols = [ctx.to_device(e) for e in evs]
    outs = [ctx.alloc(1 << (lg + b)) for lg in logs]
    layers = ctx.alloc(sh.tree_words(max_log))
    plan = ctx.batch_create(cols, logs, b, tw, outs, layers, sh.TREE_MIXED, sh.HASH_BLAKE2S, True)
    plan.run()
    ctx.sync()
    etw, eitw = oracle.precompute_twiddles(oracle.half_odds_initial(max_log - 1), max_log - 1)
    coeffs = [e.copy() for e in evs]
    oracle.interpolate_columns(coeffs, eitw)
    want = oracle.evaluate_polynomials(coeffs, b, etw)
    for o, w in zip(outs, want):
        assert np.array_equal(o.download(), w)
    assert np.array_equal(layers.download(8), oracle.merkle_commit(want)[:8])
    plan.destroy()
    ctx.fft_tune_clear()
    for x in cols + outs + [layers]:
        x.free()
    tw.destroy()

"""The benchmark workloads at their full sizes, bit-exact against the oracle
(C, OpenMP): the plain per-column path on 256 columns of 2^20 at blowup 4
(the headline workload) and the batched plan on 1,024 uneven columns
(log 10..22, blowup 2, one mixed-degree tree).  Every tree layer is
compared, plus the coefficients and evaluations of a few columns.  Inputs
come from the device generator (seeded SplitMix64, pinned against the
oracle's by test_fill_random_matches_oracle)."""
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "oracle"))
sys.path.insert(0, HERE)
import oracle  # noqa: E402
import stwo_hip as sh  # noqa: E402

pytestmark = pytest.mark.gpu
U = np.uint32


@pytest.fixture(scope="module")
def ctx():
    c = sh.Context(0)
    yield c
    c.close()


def _device_inputs(ctx, logs, seed):
    cols = []
    for c, lg in enumerate(logs):
        d = ctx.alloc(1 << lg)
        ctx.fill_random(d, 1 << lg, seed, c)
        cols.append(d)
    return cols


def _uneven_logs(n_cols, lo, hi, seed):
    """log_size = lo + j with P(j) proportional to 0.7^j (seeded)."""
    rng = np.random.default_rng(seed)
    w = np.array([0.7 ** j for j in range(hi - lo + 1)])
    return [int(lo + j) for j in rng.choice(len(w), size=n_cols, p=w / w.sum())]


def test_plain_path_256_columns_2e20_blowup4(ctx):
    n_cols, n, b, seed = 256, 20, 2, 2
    root = sh.half_odds(n + b - 1)
    tw = ctx.precompute_twiddles(root)
    cols = _device_inputs(ctx, [n] * n_cols, seed)
    outs = [ctx.alloc(1 << (n + b)) for _ in range(n_cols)]
    layers = ctx.alloc(sh.tree_words(n + b))
    try:
        ctx.interpolate_columns(cols, [n] * n_cols, tw)
        ctx.evaluate_polynomials(cols, [n] * n_cols, b, tw, outs)
        ctx.merkle_commit(outs, [n + b] * n_cols, layers)
        got_tree = layers.download()
        spot = (0, 101, 255)
        got_coeffs = {c: cols[c].download() for c in spot}
        got_evals = {c: outs[c].download() for c in spot}
    finally:
        for x in cols + outs + [layers]:
            x.free()
        tw.destroy()
    et, eit = oracle.precompute_twiddles(root.initial_index, root.log_size)
    evs = [oracle.fill_random(1 << n, seed, c) for c in range(n_cols)]
    oracle.interpolate_columns(evs, eit)
    for c in spot:
        assert np.array_equal(got_coeffs[c], evs[c]), ("coefficients", c)
    exp = oracle.evaluate_polynomials(evs, b, et)
    del evs
    for c in spot:
        assert np.array_equal(got_evals[c], exp[c]), ("evaluations", c)
    exp_tree = oracle.merkle_commit(exp)
    assert np.array_equal(got_tree[:8], exp_tree[:8]), "root"
    assert np.array_equal(got_tree, exp_tree)


def test_batched_plan_1024_uneven_columns(ctx):
    logs = _uneven_logs(1024, 10, 22, 3)
    b, seed = 1, 3
    max_log = max(logs) + b
    root = sh.half_odds(max_log - 1)
    tw = ctx.precompute_twiddles(root)
    cols = _device_inputs(ctx, logs, seed)
    outs = [ctx.alloc(1 << (lg + b)) for lg in logs]
    layers = ctx.alloc(sh.tree_words(max_log))
    spot = [int(np.argmax(logs)), int(np.argmin(logs)), 512, 1023]
    try:
        plan = ctx.batch_create(cols, logs, b, tw, outs, layers, sh.TREE_MIXED, sh.HASH_BLAKE2S, True)
        plan.run()
        ctx.sync()
        got_tree = layers.download()
        got_coeffs = {c: cols[c].download() for c in spot}
        got_evals = {c: outs[c].download() for c in spot}
        plan.destroy()
    finally:
        for x in cols + outs + [layers]:
            x.free()
        tw.destroy()
    et, eit = oracle.precompute_twiddles(root.initial_index, root.log_size)
    evs = [oracle.fill_random(1 << lg, seed, c) for c, lg in enumerate(logs)]
    oracle.interpolate_columns(evs, eit)
    for c in spot:
        assert np.array_equal(got_coeffs[c], evs[c]), ("coefficients", c, logs[c])
    exp = oracle.evaluate_polynomials(evs, b, et)
    for c in spot:
        assert np.array_equal(got_evals[c], exp[c]), ("evaluations", c, logs[c])
    exp_tree = oracle.merkle_commit(exp)
    assert np.array_equal(got_tree[:8], exp_tree[:8]), "root"
    assert np.array_equal(got_tree, exp_tree)

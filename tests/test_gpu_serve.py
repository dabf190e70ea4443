"""GPU tests of the continuous-batching server: every job's result (trace
root, FRI roots, last layer, final channel digest) is the same alone and in a
mix of concurrent jobs, and matches the oracle."""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "..", "oracle"))
import oracle  # noqa: E402
import stwo_hip as sh  # noqa: E402

pytestmark = pytest.mark.gpu

JOBS = [(8, 12, 1, 101, (1, 1, 3, 1)), (16, 13, 2, 102, (2, 2, 3, 2)), (4, 10, 1, 103, (1, 0, 3, 1)),
        (12, 12, 2, 104, (2, 1, 3, 3)), (8, 14, 1, 105, (1, 3, 3, 1)), (6, 11, 2, 106, (2, 0, 3, 2)),
        (20, 9, 1, 107, (1, 2, 3, 1)), (5, 15, 1, 108, (1, 1, 3, 2))]


def _cfg(t):
    lb, lld, nq, step = t
    return sh.FriConfig(lb, lld, nq, step)


def test_jobs_alone_equal_mix():
    with sh.Server(0, 2 << 30, 17, 4) as srv:
        ids = [srv.submit(n, l, b, seed, _cfg(f)) for (n, l, b, seed, f) in JOBS]
        mix = [srv.wait(i) for i in ids]
    for r in mix:
        assert r.status == 0
    with sh.Server(0, 2 << 30, 17, 1) as srv:
        alone = [srv.wait(srv.submit(n, l, b, seed, _cfg(f))) for (n, l, b, seed, f) in JOBS]
    for a, m, job in zip(alone, mix, JOBS):
        assert a.key() == m.key(), job


def _oracle_job(n_cols, log_size, log_blowup, seed, fri):
    L = log_size + log_blowup
    init = oracle.half_odds_initial(L - 1)
    tw, itw = oracle.precompute_twiddles(init, L - 1)
    evs = [oracle.fill_random(1 << log_size, seed, c) for c in range(n_cols)]
    oracle.interpolate_columns(evs, itw)
    outs = oracle.evaluate_polynomials(evs, log_blowup, tw)
    root = oracle.lifted_commit(outs, L)[:8]
    ch = oracle.Channel()
    ch.mix_root(root)
    lb, lld, nq, step = fri
    roots, last, _ = oracle.fri_commit(ch, oracle.FriConfig(lb, lld, nq, step), outs[:4], L)
    return root, roots, last, ch.digest_bytes()


def test_jobs_match_oracle():
    with sh.Server(0, 1 << 30, 17, 2) as srv:
        picks = [JOBS[0], JOBS[2], JOBS[3], JOBS[6]]
        ids = [srv.submit(n, l, b, seed, _cfg(f)) for (n, l, b, seed, f) in picks]
        res = [srv.wait(i) for i in ids]
    for r, (n, l, b, seed, f) in zip(res, picks):
        assert r.status == 0
        root, roots, last, digest = _oracle_job(n, l, b, seed, f)
        assert list(r.trace_root) == root.tolist()
        assert r.n_fri_roots == len(roots)
        assert list(r.fri_roots[: 8 * len(roots)]) == roots.reshape(-1).tolist()
        k = 1 << f[1]
        assert list(r.last_layer_poly[: 4 * k]) == last.reshape(-1).tolist()
        assert bytes(r.final_digest) == digest


def test_admission_respects_budget_and_rejects_oversize():
    # a budget that holds one job of this size at a time: jobs still all complete
    one = 4 * ((8 << 13) + (8 << 14) + ((2 << 14) - 1) * 8 + 256)
    with sh.Server(0, int(one * 1.5), 17, 4) as srv:
        ids = [srv.submit(8, 13, 1, 200 + i, _cfg((1, 1, 3, 1))) for i in range(5)]
        res = [srv.wait(i) for i in ids]
        assert all(r.status == 0 for r in res)
        # admission is serialised by memory: no two jobs overlap in time
        spans = sorted((r.t_admit_ms, r.t_done_ms) for r in res)
        for (a0, d0), (a1, d1) in zip(spans, spans[1:]):
            assert a1 >= d0 - 1e-6
        with pytest.raises(sh.StwoHipError) as e:
            srv.submit(64, 16, 1, 1, _cfg((1, 1, 3, 1)))
        assert e.value.code == -2

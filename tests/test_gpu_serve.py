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
        srv.hold(True)  # the burst queues whole, then is admitted as groups
        ids = [srv.submit(n, l, b, seed, _cfg(f)) for (n, l, b, seed, f) in JOBS]
        srv.hold(False)
        mix = [srv.wait(i) for i in ids]
    for r in mix:
        assert r.status == 0
    assert min(r.group_size for r in mix) > 1  # every job was co-batched
    with sh.Server(0, 2 << 30, 17, 1, max_group=1) as srv:
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
    with sh.Server(0, 1 << 30, 17, 1) as probe:
        one = probe.job_bytes(8, 13, 1, fri=_cfg((1, 1, 3, 1)))
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


def test_caller_device_columns_and_host_data():
    """Jobs that bring their columns (device pointers, or host data) give the
    results of the same columns generated from the seed; the caller's device
    columns are left unchanged."""
    n, l, b, seed, f = JOBS[1]
    host = np.concatenate([oracle.fill_random(1 << l, seed, c) for c in range(n)])
    with sh.Server(0, 1 << 30, 17, 2) as srv:
        ctx = sh.Context(0)
        dev = []
        for c in range(n):
            d = ctx.alloc(1 << l)
            ctx.fill_random(d, 1 << l, seed, c)
            dev.append(d)
        ctx.sync()
        ids = [srv.submit(n, l, b, seed, _cfg(f)), srv.submit(n, l, b, 0, _cfg(f), cols=dev),
               srv.submit(n, l, b, 0, _cfg(f), host=host)]
        res = [srv.wait(i) for i in ids]
        for c, d in enumerate(dev):
            assert np.array_equal(d.download(), host[c << l:(c + 1) << l])
        for d in dev:
            d.free()
        ctx.close()
    assert res[0].status == 0 and res[1].key() == res[0].key() and res[2].key() == res[0].key()


def _oracle_proof(n_cols, log_n, seed):
    import prove as oracle_prove
    a, b = oracle.fill_random(1 << log_n, seed, 0), oracle.fill_random(1 << log_n, seed, 1)
    return oracle_prove.proof_bytes(oracle_prove.wide_fib_prove(a, b, n_cols, log_n))


def test_prove_jobs_alone_and_in_a_mix():
    """Wide-Fibonacci prove jobs through the server: the proof bytes equal
    the oracle's, alone and submitted among commit jobs; the commit jobs'
    results do not change."""
    proves = [(16, 10, 41), (24, 12, 42)]
    want = [_oracle_proof(n, l, s) for (n, l, s) in proves]
    with sh.Server(0, 2 << 30, 17, 1) as srv:
        alone = [srv.wait(srv.submit_prove(n, l, s)) for (n, l, s) in proves]
    for (r, data), w in zip(alone, want):
        assert r.status == 0 and data == w
    with sh.Server(0, 2 << 30, 17, 3) as srv:
        ids = []
        for k, job in enumerate(JOBS[:4]):
            n, l, b, seed, f = job
            ids.append(("c", srv.submit(n, l, b, seed, _cfg(f))))
            if k < len(proves):
                ids.append(("p", srv.submit_prove(*proves[k])))
        out = [(t, srv.wait(i)) for t, i in ids]
    proofs = [x for t, x in out if t == "p"]
    for (r, data), w in zip(proofs, want):
        assert r.status == 0 and data == w
    with sh.Server(0, 2 << 30, 17, 1, max_group=1) as srv:
        alone_c = [srv.wait(srv.submit(n, l, b, seed, _cfg(f))) for (n, l, b, seed, f) in JOBS[:4]]
    mixed_c = [x for t, x in out if t == "c"]
    for a, m in zip(alone_c, mixed_c):
        assert a.key() == m.key()


def test_large_job_alone_and_in_a_mix():
    """A job of 64 columns of 2^18 at blowup 4 (2^20 evaluations) against the
    oracle, run alone and co-batched with smaller jobs."""
    big = (64, 18, 2, 777, (2, 2, 3, 1))
    root, roots, last, digest = _oracle_job(*big)
    mix_jobs = [JOBS[0], big, JOBS[3], JOBS[4]]
    with sh.Server(0, 8 << 30, 20, 2) as srv:
        alone = srv.wait(srv.submit(*big[:4], _cfg(big[4])))
        srv.hold(True)
        ids = [srv.submit(n, l, b, seed, _cfg(f)) for (n, l, b, seed, f) in mix_jobs]
        srv.hold(False)
        res = [srv.wait(i) for i in ids]
    for r in (alone, res[1]):
        assert r.status == 0
        assert list(r.trace_root) == root.tolist()
        assert list(r.fri_roots[: 8 * len(roots)]) == roots.reshape(-1).tolist()
        assert list(r.last_layer_poly[: 4 * 4]) == last.reshape(-1).tolist()
        assert bytes(r.final_digest) == digest
    assert res[1].group_size > 1


def test_device_memory_stays_within_the_budget():
    """Running jobs (commit and prove) allocates no device memory beyond the
    server's budget block: the device's free memory after creation, after a
    warm-up job and after jobs whose FRI and prove scratch alone would take
    more than 100 MiB, stays within 32 MiB."""
    import ctypes
    sh.lib()  # the HIP runtime the library links (a second runtime in the process would not see the GPU)
    hip = ctypes.CDLL("libamdhip64.so.7")

    def free_bytes():
        f, t = ctypes.c_size_t(), ctypes.c_size_t()
        assert hip.hipDeviceSynchronize() == 0
        assert hip.hipMemGetInfo(ctypes.byref(f), ctypes.byref(t)) == 0
        return f.value

    with sh.Server(0, 4 << 30, 21, 2) as srv:
        srv.wait(srv.submit(8, 10, 1, 5, _cfg((1, 1, 3, 1))))
        srv.wait(srv.submit_prove(8, 10, 5))
        free0 = free_bytes()
        ids = [srv.submit(16, 20, 1, 9, _cfg((1, 1, 3, 1))), srv.submit(32, 17, 2, 10, _cfg((2, 1, 3, 2))),
               srv.submit_prove(64, 14, 11)]
        res = [srv.wait(i) for i in ids]
        assert res[0].status == 0 and res[1].status == 0 and res[2][0].status == 0
        free1 = free_bytes()
    assert free0 - free1 < (32 << 20), (free0, free1)

"""Error paths under device-allocation failure.

`stwo_hip_ctx_fail_allocs_after(ctx, k)` lets a context's next k device
allocations succeed and fails every later one, as an exhausted device would.
For each call below, a fresh context (its inputs, outputs and twiddles
allocated before the hook) fails its k-th allocation for k = 0, 1, ... until
the call needs no more than k: every failing call must return
STWO_HIP_ENOMEM with a message, and the same context, with the hook off, must
then give the same result as a context that never saw a failure, so no
failure leaves half-built plans, scratch or channel state behind.  Once the
contexts are destroyed the device's free memory is not below what it was
after the first clean call (which also holds what the HIP runtime keeps
for itself once a kernel has run, such as code objects and wave scratch)."""
import ctypes
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "..", "oracle"))
import oracle as O  # noqa: E402
import stwo_hip as sh  # noqa: E402

pytestmark = pytest.mark.gpu

ENOMEM = -2
U = np.uint32


def free_bytes():
    sh.lib()  # the HIP runtime the library links
    hip = ctypes.CDLL("libamdhip64.so.7")
    f, t = ctypes.c_size_t(), ctypes.c_size_t()
    assert hip.hipDeviceSynchronize() == 0
    assert hip.hipMemGetInfo(ctypes.byref(f), ctypes.byref(t)) == 0
    return f.value


def release(state):
    """Frees the device buffers and twiddle trees a prep function made."""
    if isinstance(state, (list, tuple)):
        for s in state:
            release(s)
    elif isinstance(state, sh.DeviceBuf):
        state.free()
    elif isinstance(state, sh.Twiddles):
        state.destroy()


def fail_each_allocation(prep, call, max_k=64):
    """Runs call(ctx, state) on fresh contexts with the k-th allocation
    failing, k = 0, 1, ...; returns (number of failing k, clean result)."""
    with sh.Context(0) as ctx:
        state = prep(ctx)
        expect = call(ctx, state)
        release(state)
    free0 = free_bytes()
    for k in range(max_k):
        with sh.Context(0) as ctx:
            state = prep(ctx)
            try:
                ctx.fail_allocs_after(k)
                try:
                    got = call(ctx, state)
                    failed = False
                except sh.StwoHipError as e:
                    assert e.code == ENOMEM, str(e)
                    assert "memory" in str(e) or "allocation" in str(e), str(e)
                    failed = True
                ctx.fail_allocs_after(-1)
                again = call(ctx, state)
                assert again == expect, "context unusable after an allocation failure at k = %d" % k
            finally:
                release(state)
        # The HIP runtime's own holdings (memory of the hardware queues behind
        # its streams, reused across streams) move free memory by tens of MiB
        # either way; a leaked context holds at least its 8 MiB upload ring,
        # and leaks add up over the loop against the fixed free0.
        assert free_bytes() > free0 - (64 << 20), "device memory not released after k = %d" % k
        if not failed:
            assert got == expect
            return k, expect
    raise AssertionError("the call still failed with %d allocations allowed" % max_k)


def test_prove_under_allocation_failures():
    """wide_fib_prove (coefficients stored and the PCS default path): the
    batch plans, the prove and FRI scratch and the kept tables are all built
    after the hook, so each of their allocations fails in turn."""
    log_n, n_cols = 6, 16
    a = np.arange(1, (1 << log_n) + 1, dtype=U)
    b = np.arange(7, (1 << log_n) + 7, dtype=U)
    cfg = sh.PcsConfig(10, sh.FriConfig(1, 0, 3, 1))

    def prep(ctx):
        return ctx.to_device(a), ctx.to_device(b), ctx.precompute_twiddles(sh.half_odds(log_n))

    for store in (True, False):
        def call(ctx, st):
            da, db, tw = st
            data, _, qpos, digest, n_draws = ctx.wide_fib_prove(da, db, n_cols, log_n, tw, cfg, store_coeffs=store)
            return data, list(qpos), digest, n_draws

        n_fail, _ = fail_each_allocation(prep, call)
        assert n_fail >= 3  # the prove allocates its buffer, plans and FRI memory itself


def test_commit_and_batch_under_allocation_failures():
    """extend_evals + lifted commit, and a batch plan created and run, with
    each allocation failing in turn."""
    logs = [9, 8, 8, 6]
    rng = np.random.default_rng(5)
    host = [rng.integers(0, 2**31 - 1, 1 << n, dtype=np.uint64).astype(U) for n in logs]
    L = max(logs) + 1

    def prep(ctx):
        cols = [ctx.to_device(h) for h in host]
        outs = [ctx.zeros(1 << (n + 1)) for n in logs]
        layers = ctx.zeros(sh.tree_words(L))
        tw = ctx.precompute_twiddles(sh.half_odds(L - 1))
        return cols, outs, layers, tw

    def commit(ctx, st):
        cols, outs, layers, tw = st
        for c, h in zip(cols, host):  # the coefficients are stored over the inputs: fresh inputs per run
            c.upload(h)
        ctx.extend_evals(cols, logs, 1, tw, outs, store_coeffs=True)
        ctx.lifted_commit(outs, [n + 1 for n in logs], L, layers)
        ctx.sync()
        return layers.download(8).tolist()

    def batch(ctx, st):
        cols, outs, layers, tw = st
        for c, h in zip(cols, host):  # the plan stores the coefficients over the inputs
            c.upload(h)
        bt = ctx.batch_create(cols, logs, 1, tw, outs, layers, tree_kind=sh.TREE_MIXED)
        try:
            bt.run()
            ctx.sync()
            return layers.download(8).tolist()
        finally:
            bt.destroy()

    # extend_evals and the lifted commit take their tables from the context's
    # upload ring: at these sizes they allocate no device memory at all
    assert fail_each_allocation(prep, commit)[0] == 0
    assert fail_each_allocation(prep, batch)[0] >= 1


def test_twiddles_and_fri_prover_under_allocation_failures():
    """precompute_twiddles and a FRI prover (commit, decommit, layer proofs)
    with each allocation failing in turn."""
    log_size = 10
    cfg = sh.FriConfig(1, 1, 5, 2)
    # FRI takes a low-degree column: evaluations of 2^(log_size - 1) random coefficients
    rng = np.random.default_rng(9)
    coeffs = [rng.integers(0, O.P, 1 << (log_size - 1), dtype=np.uint64).astype(U) for _ in range(4)]
    otw, _ = O.precompute_twiddles(O.half_odds_initial(log_size - 1), log_size - 1)
    host = O.evaluate_polynomials(coeffs, 1, otw)

    def prep_tw(ctx):
        return None

    def twiddles(ctx, _):
        tw = ctx.precompute_twiddles(sh.half_odds(log_size - 1))
        try:
            return tw.download()[0][:16].tolist()
        finally:
            tw.destroy()

    def prep(ctx):
        return [ctx.to_device(h) for h in host], ctx.precompute_twiddles(sh.half_odds(log_size - 1))

    def fri(ctx, st):
        col4, tw = st
        pr, roots, last, digest, n_draws = ctx.fri_prover_commit(bytes(32), 0, cfg, col4, log_size, tw)
        try:
            pos, raw, digest2, n_draws2 = pr.decommit(digest, n_draws)
            lp = pr.layer_proof(0)
            return (roots.tolist(), last.tolist(), digest2, n_draws2, pos.tolist(), lp["hash_witness"].tolist(),
                    lp["fri_witness"].tolist())
        finally:
            pr.destroy()

    assert fail_each_allocation(prep_tw, twiddles)[0] >= 1
    assert fail_each_allocation(prep, fri)[0] >= 1

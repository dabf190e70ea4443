"""Pins the wide-Fibonacci restatement (or_wide_fib_trace,
or_wide_fib_constraint_quotients in oracle/stwo_oracle.c).  CPU only.

* Trace: the reference test's inputs a = 1, b = row
  (examples/src/wide_fibonacci/mod.rs:132-139) give the sequences written out
  by hand here (1, 0, 1, 1, 2, 5, 29, 866 and 1, 3, 10, 109, 11981, ...).
* Constraint quotients: for a trace that satisfies the AIR, the
  random-linear combination of the constraints divided by the trace domain's
  vanishing polynomial is a polynomial -- its coefficients above index
  2^log_n on CanonicCoset(log_n + 1) are zero; changing one trace value makes
  it fill the whole domain.  The vanishing polynomial here takes the general
  coset shift (-initial + step / 2); the library uses that it is zero for
  canonic cosets.
"""
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "oracle"))
sys.path.insert(0, HERE)
import oracle as O  # noqa: E402
from test_quotients_oracle import interp  # noqa: E402

U = np.uint32
ALPHA = (12345, 678, 91011, 1213)


def test_trace_reference_inputs():
    a = np.ones(32, U)
    b = np.arange(32, dtype=U)
    tr = O.wide_fib_trace(a, b, 8)
    assert [int(c[0]) for c in tr] == [1, 0, 1, 1, 2, 5, 29, 866]
    assert [int(c[3]) for c in tr][:5] == [1, 3, 10, 109, 11981]


def extend(trace, log_n):
    tw, itw = O.precompute_twiddles(O.half_odds_initial(log_n), log_n)
    cols = [c.copy() for c in trace]
    O.interpolate_columns(cols, itw)
    return O.evaluate_polynomials(cols, 1, tw)


@pytest.mark.parametrize("log_n,n_cols", [(3, 3), (5, 8), (7, 17)])
def test_constraint_quotients_low_degree(log_n, n_cols):
    rng = np.random.default_rng(log_n)
    a = rng.integers(0, O.P, 1 << log_n, dtype=np.uint64).astype(U)
    b = rng.integers(0, O.P, 1 << log_n, dtype=np.uint64).astype(U)
    tr = O.wide_fib_trace(a, b, n_cols)
    acc = O.wide_fib_constraint_quotients(extend(tr, log_n), log_n, ALPHA)
    for q in acc:
        assert not interp(q, log_n + 1)[(1 << log_n) + 1:].any()
    bad = [c.copy() for c in tr]
    bad[n_cols - 1][1] = (int(bad[n_cols - 1][1]) + 1) % O.P
    acc = O.wide_fib_constraint_quotients(extend(bad, log_n), log_n, ALPHA)
    assert any(interp(q, log_n + 1)[(1 << log_n) + 1:].any() for q in acc)


def test_constraint_quotients_accumulate():
    """The result is added to the accumulator it is given."""
    log_n = 4
    tr = O.wide_fib_trace(np.ones(16, U), np.arange(16, dtype=U), 5)
    ext = extend(tr, log_n)
    base = [np.full(32, 7 + k, U) for k in range(4)]
    got = O.wide_fib_constraint_quotients(ext, log_n, ALPHA, [c.copy() for c in base])
    alone = O.wide_fib_constraint_quotients(ext, log_n, ALPHA)
    for g, a0, b0 in zip(got, alone, base):
        assert np.array_equal(g, (a0.astype(np.uint64) + b0) % O.P)

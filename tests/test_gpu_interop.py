"""GPU parity of the SimdBackend layout entry points against
oracle/simd_layout.py: coefficient transposition (after a real IFFT of this
library, whose natural order the oracle pins), doubled twiddle trees, and
mapping the redundant P to 0."""
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "oracle"))
import oracle as O  # noqa: E402
import simd_layout as S  # noqa: E402
import stwo_hip as sh  # noqa: E402

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = sh.Context(0)
    yield c
    c.close()


@pytest.mark.parametrize("log_size", [10, 16, 17, 18, 21])
def test_coeffs_simd_transpose(ctx, log_size):
    tw = ctx.precompute_twiddles(sh.half_odds(log_size - 1))
    ev = O.fill_random(1 << log_size, 31, log_size)
    d = ctx.to_device(ev)
    ctx.interpolate(d, sh.canonic_domain(log_size), tw)
    natural = d.download()
    ctx.coeffs_simd_transpose(d, log_size)
    assert np.array_equal(d.download(), S.simd_coeff_order(natural, log_size))
    ctx.coeffs_simd_transpose(d, log_size)
    assert np.array_equal(d.download(), natural)
    d.free()
    tw.destroy()


@pytest.mark.parametrize("log_size", [3, 4, 5, 12])
def test_twiddles_simd_doubled(ctx, log_size):
    root = sh.half_odds(log_size)
    tw = ctx.precompute_twiddles(root)
    n = 1 << log_size
    out_tw, out_itw = ctx.alloc(n), ctx.alloc(n)
    ctx.twiddles_simd_doubled(tw, out_tw, out_itw)
    t, it = O.precompute_twiddles(root.initial_index, root.log_size)
    assert np.array_equal(out_tw.download(), S.doubled(t))
    assert np.array_equal(out_itw.download(), S.doubled(it))
    out_tw.free()
    out_itw.free()
    tw.destroy()


def test_canonicalize(ctx):
    v = O.fill_random(5000, 3, 3)
    v[::7] = O.P
    d = ctx.to_device(v)
    ctx.canonicalize(d, len(v))
    want = v.copy()
    want[want == O.P] = 0
    assert np.array_equal(d.download(), want)
    d.free()

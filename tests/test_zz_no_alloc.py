"""stwo_hip_eval_at_point_noalloc (the entry point that allocates nothing:
its scratch is a fixed per-device array) against stwo_hip_eval_at_point,
bit-exact, through the C ABI: two calls in a row, the second smaller, so
the second call runs on what the first left in the shared array; the small
call is also checked against the oracle.  Sizes beyond the array are
rejected before any launch."""
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "oracle"))
sys.path.insert(0, HERE)
import oracle as O  # noqa: E402
import stwo_hip as sh  # noqa: E402
from test_quotients_oracle import secure_point  # noqa: E402

pytestmark = pytest.mark.gpu
U = np.uint32


@pytest.fixture(scope="module")
def ctx():
    c = sh.Context(0)
    yield c
    c.close()


def device_cols(ctx, logs, seed):
    cols = []
    for i, lg in enumerate(logs):
        b = ctx.alloc(1 << lg)
        ctx.fill_random(b, 1 << lg, seed, i)
        cols.append(b)
    return cols


def test_noalloc_matches_allocating_entry_twice(ctx):
    # First call: 128 columns of 2^18 plus odd sizes (32,780 blocks of 2^11,
    # 139,900 of the array's 262,144 words); second call: far smaller.
    big_logs = [18] * 128 + [20, 11, 5, 0, 13, 13, 2]
    small_logs = [3, 11, 12, 0, 7]
    big = device_cols(ctx, big_logs, 21)
    small = device_cols(ctx, small_logs, 22)
    pt1, pt2 = secure_point(0xC0FFEE), secure_point(77)
    print("noalloc: call 1, %d columns" % len(big_logs), flush=True)
    want1 = ctx.eval_at_point(big, big_logs, pt1)
    got1 = ctx.eval_at_point(big, big_logs, pt1, noalloc=True)
    assert np.array_equal(got1, want1)
    print("noalloc: call 2, %d columns" % len(small_logs), flush=True)
    want2 = ctx.eval_at_point(small, small_logs, pt2)
    got2 = ctx.eval_at_point(small, small_logs, pt2, noalloc=True)
    assert np.array_equal(got2, want2)
    for c, lg, g in zip(small, small_logs, got2):
        assert tuple(int(v) for v in g) == O.eval_at_point(c.download(), pt2)
    # the first call again, after the smaller one
    got1b = ctx.eval_at_point(big, big_logs, pt1, noalloc=True)
    assert np.array_equal(got1b, want1)
    for x in big + small:
        x.free()


def test_noalloc_rejects_beyond_scratch(ctx):
    # 256 columns of 2^20 coefficients: 131,072 blocks, 525,312 words > 2^18.
    one = ctx.alloc(1 << 20)
    with pytest.raises(sh.StwoHipError):
        ctx.eval_at_point([one] * 256, [20] * 256, secure_point(5), noalloc=True)
    # the same set through the allocating entry point still works
    ctx.fill_random(one, 1 << 20, 3, 0)
    want = ctx.eval_at_point([one] * 2, [20] * 2, secure_point(5))
    got = ctx.eval_at_point([one] * 2, [20] * 2, secure_point(5), noalloc=True)
    assert np.array_equal(got, want)
    one.free()

"""A CPU model of the radix-16 FFT kernels' addressing (fft_pass_body and
lde_pass_kernel in fft.hip), run over every pass the planner makes
(stwo_hip_debug_fft_plan, which launches nothing): for every lane of every
workgroup -- including the idle lanes past the last tile, which still read
twiddles -- each data load stays inside the pass's input (2^in_log words of
the column, or of the coefficient domain), each store inside the output
(2^N words), each twiddle read inside its own tree level (level of line
layer s: 2^(N-1-s) words; the circle layer's (x, y) pairs: the 2^(N-2) words
of line layer 1's level), for one column and for three (the idle-lane case).
This is the check that found the twiddle over-read of idle lanes after
tiles were renumbered (a copy of the model with that numbering flags 35
shapes); it runs on every planner shape, N = 4..28, every blowup (for
passes of more than 2,048 workgroups, the lanes of the first and last 256
workgroups: the bounds grow with the tile index)."""
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "stwo-metal_amd"))
import stwo_hip as sh  # noqa: E402

MAX_WGS_FULL = 1 << 11


def lanes_of(p, n_cols, N):
    """(local, active, tile, col) per lane of the workgroups modelled."""
    T, wg_log = p["T"], p["wg"]
    tpt_log = T - 4
    dom_log = N - p["rep"]
    tpc_log = dom_log - T
    total_tiles = n_cols << tpc_log
    tiles_per_wg = 1 << (wg_log + 4 - T)
    wgs = -(-total_tiles // tiles_per_wg)
    assert wgs * tiles_per_wg <= 1 << 32
    if wgs <= MAX_WGS_FULL:
        wg = np.arange(wgs, dtype=np.int64)
    else:
        wg = np.concatenate([np.arange(256), np.arange(wgs - 256, wgs)]).astype(np.int64)
    t = np.arange(1 << wg_log, dtype=np.int64)
    wg, t = np.meshgrid(wg, t, indexing="ij")
    wg, t = wg.ravel(), t.ravel()
    local = t & ((1 << tpt_log) - 1)
    slot = t >> tpt_log
    gtile = (wg << (wg_log - tpt_log)) | slot
    active = gtile < total_tiles
    tile = np.where(active, gtile & ((1 << tpc_log) - 1), 0)
    col = np.where(active, gtile >> tpc_log, 0)
    assert col.max() < n_cols
    return local, active, tile, wgs


def check_pass(p, N, n_cols):
    T, lo, w, m = p["T"], p["lo"], p["w"], p["m"]
    local, active, tile, _ = lanes_of(p, n_cols, N)
    ih_bits = lo - w
    ih = tile & ((1 << ih_bits) - 1)
    o = tile >> ih_bits
    wmask = (1 << w) - 1
    in_mask = (1 << p["in_log"]) - 1
    dom_log = N - p["rep"]
    top_b0 = T - 4

    def group_base(b0):
        return ((local >> b0) << (b0 + 4)) | (local & ((1 << b0) - 1))

    for rep in range(1 << p["rep"]):
        tile_gb = (rep << dom_log) | (o << (lo + m)) | (ih << w)

        def gidx(b):
            return tile_gb | ((b >> w) << lo) | (b & wmask)

        # loads (active lanes): the top layout, or 16 consecutive words (vec_io bit 0)
        if p["vec_io"] & 1:
            a = (gidx(local << 4) & in_mask)[active]
            assert a.size == 0 or (a.max() + 15 < 1 << p["in_log"] and (a % 16 == 0).all())
        else:
            gb = gidx(local)
            sh_ = top_b0 - w + lo
            jmask = in_mask >> sh_
            for j in range(16):
                a = ((gb & in_mask) + ((j & jmask) << sh_))[active]
                assert a.size == 0 or a.max() < 1 << p["in_log"], (p, j)
        # twiddle reads of every stage (every lane, idle ones included)
        for (b0, r_lo, r_hi) in p["stages"]:
            gb = gidx(group_base(b0))
            sh_ = b0 - w + lo
            for r in range(r_lo, r_hi):
                s = sh_ + r
                hs = gb >> (s + 1)
                if s == 0:
                    last = (hs >> 1) + 3
                    assert last.max() < 1 << (N - 2), (p, "circle")
                else:
                    cnt = 8 >> r
                    last = hs + cnt - 1
                    assert last.max() < 1 << (N - 1 - s), (p, s)
        # stores (active lanes) from the final layout
        final = p["stages"][-1][0] if p["stages"] else top_b0
        if (p["vec_io"] & 2) and final == 0:
            a = gidx(local << 4)[active]
            assert a.size == 0 or a.max() + 15 < 1 << N
        else:
            cur = top_b0 if (final != top_b0 and final < 5) else final
            gb = gidx(group_base(cur))
            sh_ = cur - w + lo
            for j in range(16):
                a = (gb + (j << sh_))[active]
                assert a.size == 0 or a.max() < 1 << N, (p, j)


@pytest.mark.parametrize("N", list(range(4, 29)))
def test_pass_addressing_in_bounds(N):
    for inverse in (True, False):
        for n in ([N] if inverse else range(max(1, N - 4), N + 1)):
            for p in sh.fft_plan(n, N, inverse):
                for n_cols in (1, 3):
                    check_pass(p, N, n_cols)


@pytest.mark.parametrize("n", list(range(4, 27)))
def test_fused_lde_pair_shares_the_tile_mapping(n):
    """The fused pass runs the IFFT's last pass and the RFFT's first pass on
    the IFFT pass's tile mapping: where the host fuses them (same T, lo, w,
    m, strided, T >= 10, the forward pass tiling the coefficient domain) the
    forward pass's own mapping is the same one, so the checks above cover
    every address the fused kernel forms."""
    for b in range(0, 3):
        if n + b > 28:
            continue
        li = sh.fft_plan(n, n, True)[-1]
        ff = sh.fft_plan(n, n + b, False)[0]
        fused = (li["T"] == ff["T"] and li["lo"] == ff["lo"] and li["w"] == ff["w"] and li["m"] == ff["m"]
                 and li["lo"] > 0 and li["T"] >= 10 and ff["rep"] == b and li["wg"] == ff["wg"]
                 and li["wgs"] == ff["wgs"])
        if n > 12:
            assert fused, (n, b, li, ff)
            # same tiles of the coefficient domain: N_i = n, N_f - rep = n
            assert n == (n + b) - ff["rep"]
        else:
            assert not fused


def tuned_plans(n, N, inverse):
    """The distinct plans stwo_hip_fft_autotune can pick for a shape."""
    seen = {}
    for t0 in range(10, 14):
        for smax in range(5, 9):
            for wg in (0, 9):
                ps = sh.fft_plan(n, N, inverse, t0, smax, wg)
                key = tuple(tuple(sorted((k, str(v)) for k, v in p.items())) for p in ps)
                seen.setdefault(key, ps)
    return list(seen.values())


@pytest.mark.parametrize("N", list(range(4, 29)))
def test_tuned_plans_addressing_in_bounds(N):
    """Every pass of every plan the autotuner may choose, under the same
    model (each distinct pass once)."""
    done = set()
    for inverse in (True, False):
        for n in ([N] if inverse else range(max(1, N - 3), N + 1)):
            for ps in tuned_plans(n, N, inverse):
                for p in ps:
                    key = tuple(sorted((k, str(v)) for k, v in p.items() if k != "scaled"))
                    if key not in done:
                        done.add(key)
                        check_pass(p, N, 3)

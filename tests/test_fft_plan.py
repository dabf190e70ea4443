"""CPU checks of the FFT pass planner through stwo_hip_debug_fft_plan (which
launches nothing): for every transform shape the library accepts, each
butterfly layer is applied exactly once, in the order the circle FFT needs
(RFFT top layer first, IFFT circle layer first), and every pass's tile and
register groups are consistent with the kernel's assumptions."""
import ctypes
import os

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(REPO, "stwo-metal_amd", "libstwo_hip.so")
SUPPORTED = {(0, 4), (0, 3), (0, 2), (0, 1), (1, 2), (1, 3), (1, 4), (2, 3), (2, 4), (3, 4)}


@pytest.fixture(scope="module")
def lib():
    L = ctypes.CDLL(LIB)
    L.stwo_hip_debug_fft_plan.argtypes = [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int,
                                          ctypes.POINTER(ctypes.c_uint32), ctypes.c_uint32,
                                          ctypes.POINTER(ctypes.c_uint32)]
    return L


def plan(lib, n, N, inverse):
    buf = (ctypes.c_uint32 * (16 * 8))()
    k = ctypes.c_uint32()
    rc = lib.stwo_hip_debug_fft_plan(n, N, int(inverse), buf, 8, ctypes.byref(k))
    assert rc == 0, (n, N, inverse)
    out = []
    for i in range(k.value):
        w = list(buf[16 * i: 16 * i + 16])
        stages = []
        for j in range(w[8]):
            v = w[9 + j]
            stages.append((v >> 16, (v >> 8) & 0xFF, v & 0xFF))
        out.append(dict(T=w[0], lo=w[1], w=w[2], m=w[3], rep=w[4], wg=w[5], in_log=w[6], scaled=w[7],
                        stages=stages, wgs=w[13], vec_io=w[14]))
    return out



@pytest.mark.parametrize("N", list(range(4, 29)))
def test_every_layer_once_in_order(lib, N):
    for inverse in (False, True):
        for n in ([N] if inverse else range(0, N + 1)):
            L = N if inverse else n
            passes = plan(lib, n, N, inverse)
            seen = []
            for k, p in enumerate(passes):
                assert p["T"] == p["m"] + p["w"] and 4 <= p["T"] <= 13
                assert p["w"] <= p["lo"] or p["w"] == 0
                assert p["lo"] + p["m"] <= N
                assert p["wg"] == max(8, p["T"] - 4)
                assert p["in_log"] == (n if (not inverse and k == 0) else N)
                assert p["scaled"] == (1 if (inverse and k == len(passes) - 1) else 0)
                if p["rep"]:
                    assert not inverse and k == 0 and p["rep"] == N - n and p["lo"] + p["m"] <= n
                for (b0, rl, rh) in p["stages"]:
                    if (rl, rh) == (0, 0):
                        continue
                    assert (rl, rh) in SUPPORTED
                    assert b0 >= p["w"] and b0 + 4 <= p["T"], (n, N, inverse, p)
                    rs = range(rl, rh) if inverse else range(rh - 1, rl - 1, -1)
                    for r in rs:
                        seen.append(b0 + r - p["w"] + p["lo"])
            assert sorted(seen) == list(range(L)), (n, N, inverse, passes)
            expect = list(range(L)) if inverse else list(range(L - 1, -1, -1))
            assert seen == expect, (n, N, inverse, seen)


@pytest.mark.parametrize("N", list(range(4, 29)))
def test_vector_io_only_on_contiguous_layout_zero(lib, N):
    """128-bit loads/stores are used only where a lane's 16 elements are 16
    consecutive words: layout b0 = 0 of a contiguous tile (w = 0), and for
    loads an input index mask that keeps the low 4 bits (in_log >= 4)."""
    for inverse in (False, True):
        for n in ([N] if inverse else range(0, N + 1)):
            for p in plan(lib, n, N, inverse):
                v = p["vec_io"]
                if v & 1:
                    assert p["w"] == 0 and p["lo"] == 0 and p["stages"][0][0] == 0 and p["in_log"] >= 4, (n, N, p)
                if v & 2:
                    assert p["w"] == 0 and p["lo"] == 0 and p["stages"][-1][0] == 0, (n, N, p)
                if p["w"] == 0 and p["stages"][0][2] > 0:
                    assert bool(v & 1) == (p["stages"][0][0] == 0 and p["in_log"] >= 4)
                    assert bool(v & 2) == (p["stages"][-1][0] == 0)

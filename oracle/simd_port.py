"""ctypes binding of the SimdBackend-equivalent CPU port (simd_port.cpp).

Test and CPU-baseline infrastructure only: tests/ and bench.py's
cpu_baseline leg load it; the library never does.  Picks the AVX-512 build
(16 lanes, SimdBackend's PackedM31 width) when the host CPU has avx512f,
else the AVX2 build (8 lanes)."""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_lib = None


def host_has_avx512():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("flags"):
                    return " avx512f " in line + " "
    except OSError:
        pass
    return False


def lib(variant=None):
    global _lib
    if variant is None and _lib is not None:
        return _lib
    v = variant or ("v4" if host_has_avx512() else "v3")
    # STWO_SIMDPORT_LIB: another build of the same source (the sanitizer build)
    path = os.environ.get("STWO_SIMDPORT_LIB") or os.path.join(HERE, "libsimdport_%s.so" % v)
    if not os.path.exists(path):
        raise RuntimeError("%s is not built (make -C oracle)" % path)
    L = ctypes.CDLL(path)
    pp = ctypes.POINTER(ctypes.c_void_p)
    u32p = ctypes.POINTER(ctypes.c_uint32)
    L.sp_interpolate_columns.argtypes = [pp, ctypes.c_size_t, ctypes.c_uint32, u32p, ctypes.c_size_t]
    L.sp_evaluate_columns.argtypes = [pp, ctypes.c_size_t, ctypes.c_uint32, ctypes.c_uint32, u32p, ctypes.c_size_t,
                                      pp]
    L.sp_merkle_commit.argtypes = [pp, ctypes.c_size_t, ctypes.c_uint32, u32p]
    if variant is None:
        _lib = L
    return L


def _ptrs(arrs):
    return (ctypes.c_void_p * max(1, len(arrs)))(*[a.ctypes.data for a in arrs])


def doubled(tree):
    """The twiddle tree as SimdBackend stores it: 2t as u32."""
    return (tree.astype(np.uint64) * 2).astype(np.uint32)


class Port:
    """The port on one set of doubled trees (from oracle.precompute_twiddles)."""

    def __init__(self, tw, itw, variant=None):
        self.L = lib(variant)
        self.tw2 = np.ascontiguousarray(doubled(tw))
        self.itw2 = np.ascontiguousarray(doubled(itw))

    @property
    def width(self):
        return self.L.sp_simd_width()

    @property
    def threads(self):
        return self.L.sp_num_threads()

    def interpolate_columns(self, cols, log_n):
        rc = self.L.sp_interpolate_columns(_ptrs(cols), len(cols), log_n,
                                           self.itw2.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), len(self.itw2))
        assert rc == 0, "sp_interpolate_columns rejected log %d" % log_n

    def evaluate_columns(self, coeffs, log_n, log_blowup, outs=None):
        if outs is None:
            outs = [np.empty(1 << (log_n + log_blowup), dtype=np.uint32) for _ in coeffs]
        rc = self.L.sp_evaluate_columns(_ptrs(coeffs), len(coeffs), log_n, log_blowup,
                                        self.tw2.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), len(self.tw2),
                                        _ptrs(outs))
        assert rc == 0, "sp_evaluate_columns rejected log %d + %d" % (log_n, log_blowup)
        return outs

    def merkle_commit(self, cols, log_size, layers=None):
        if layers is None:
            layers = np.empty(((2 << log_size) - 1) * 8, dtype=np.uint32)
        rc = self.L.sp_merkle_commit(_ptrs(cols), len(cols), log_size,
                                     layers.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)))
        assert rc == 0, "sp_merkle_commit rejected"
        return layers

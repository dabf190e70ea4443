"""Python binding of libstwo_hip (the C ABI in include/stwo_hip.h) via ctypes.

The method names follow stwo's Backend traits (PolyOps::interpolate_columns,
evaluate_polynomials, ColumnOps::bit_reverse_column,
MerkleOps::commit_on_layer, MerkleOpsLifted::build_leaves,
FriOps::fold_circle_into_line / fold_line, FriProver::commit) so tests read
like the reference's own trait-level tests.  Device columns are `DeviceBuf`
objects owning a device allocation; numpy uint32 arrays cross the host
boundary.  There is no CPU fallback: if the library is missing or no GPU is
visible, loading fails with an exception.
"""
import ctypes
import time
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libstwo_hip.so")
P = (1 << 31) - 1

HASH_BLAKE2S = 0
HASH_BLAKE2S_M31 = 1
TREE_MIXED = 0
BATCH_GRAPH, BATCH_DROP_COEFFS = 1, 2
TREE_LIFTED = 1

_u32p = ctypes.POINTER(ctypes.c_uint32)
_vp = ctypes.c_void_p


class StwoHipError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("stwo_hip error %d: %s" % (code, msg))
        self.code = code


class Coset(ctypes.Structure):
    _fields_ = [("initial_index", ctypes.c_uint32), ("log_size", ctypes.c_uint32)]


class FriConfig(ctypes.Structure):
    _fields_ = [("log_blowup_factor", ctypes.c_uint32), ("log_last_layer_degree_bound", ctypes.c_uint32),
                ("n_queries", ctypes.c_uint32), ("line_fold_step", ctypes.c_uint32)]


class PcsConfig(ctypes.Structure):
    """PcsConfig (lifting_log_size None); the default is PcsConfig::default()."""
    _fields_ = [("pow_bits", ctypes.c_uint32), ("fri", FriConfig)]

    @staticmethod
    def default():
        return PcsConfig(10, FriConfig(1, 0, 3, 1))


PROVE_PHASES = ("trace_commit", "composition_commit", "oods_samples", "fri_quotients", "fri_commit",
                "proof_of_work", "decommit")


JOB_COMMIT_FRI, JOB_PROVE = 0, 1


class BatchSet(ctypes.Structure):
    _fields_ = [("cols", ctypes.c_void_p), ("log_sizes", ctypes.c_void_p), ("n_cols", ctypes.c_size_t),
                ("log_blowup", ctypes.c_uint32), ("outs", ctypes.c_void_p), ("layers", ctypes.c_void_p),
                ("tree_kind", ctypes.c_int), ("hash_mode", ctypes.c_int), ("stage", ctypes.c_int)]


BATCH_LDE, BATCH_IFFT, BATCH_RFFT = 0, 1, 2


class JobDesc(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_uint32), ("n_cols", ctypes.c_uint32), ("log_size", ctypes.c_uint32),
                ("log_blowup", ctypes.c_uint32), ("seed", ctypes.c_uint64), ("cols", ctypes.c_void_p),
                ("host", ctypes.c_void_p), ("fri", FriConfig), ("pcs", PcsConfig), ("proof_out", ctypes.c_void_p),
                ("proof_cap", ctypes.c_size_t)]


class JobResult(ctypes.Structure):
    _fields_ = [("status", ctypes.c_int), ("trace_root", ctypes.c_uint32 * 8), ("n_fri_roots", ctypes.c_uint32),
                ("fri_roots", ctypes.c_uint32 * 256), ("last_layer_poly", ctypes.c_uint32 * 64),
                ("final_digest", ctypes.c_uint8 * 32), ("proof_len", ctypes.c_size_t), ("group_size", ctypes.c_uint32),
                ("t_submit_ms", ctypes.c_double), ("t_admit_ms", ctypes.c_double), ("t_start_ms", ctypes.c_double),
                ("t_done_ms", ctypes.c_double)]

    def key(self):
        """Everything the job computed (for alone-vs-mix comparisons)."""
        n = self.n_fri_roots
        return (self.status, tuple(self.trace_root), tuple(self.fri_roots[: 8 * n]), tuple(self.last_layer_poly),
                bytes(self.final_digest), self.proof_len)


def half_odds(n):
    """Coset::half_odds(n) (STWO/core/circle.rs:328-330)."""
    return Coset((1 << (29 - n)) & ((1 << 31) - 1), n)


def canonic_domain(log_size):
    """Half coset of CanonicCoset(log_size).circle_domain()."""
    return half_odds(log_size - 1)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FileNotFoundError("libstwo_hip.so is not built (run __graft_entry__.build())")
        L = ctypes.CDLL(LIB_PATH)
        sig = {
            "stwo_hip_ctx_create": [ctypes.c_int, ctypes.POINTER(_vp)],
            "stwo_hip_ctx_destroy": [_vp],
            "stwo_hip_last_error": [_vp, ctypes.c_char_p, ctypes.c_size_t],
            "stwo_hip_sync": [_vp],
            "stwo_hip_alloc": [_vp, ctypes.c_size_t, ctypes.POINTER(_vp)],
            "stwo_hip_free": [_vp, _vp],
            "stwo_hip_upload": [_vp, _vp, _vp, ctypes.c_size_t],
            "stwo_hip_download": [_vp, _vp, _vp, ctypes.c_size_t],
            "stwo_hip_copy": [_vp, _vp, _vp, ctypes.c_size_t],
            "stwo_hip_memset_zero": [_vp, _vp, ctypes.c_size_t],
            "stwo_hip_fill_random": [_vp, _vp, ctypes.c_size_t, ctypes.c_uint64, ctypes.c_uint64],
            "stwo_hip_twiddles_create": [_vp, Coset, ctypes.POINTER(_vp)],
            "stwo_hip_twiddles_destroy": [_vp, _vp],
            "stwo_hip_twiddles_buffers": [_vp, ctypes.POINTER(_vp), ctypes.POINTER(_vp), ctypes.POINTER(ctypes.c_size_t)],
            "stwo_hip_bit_reverse_column": [_vp, _vp, ctypes.c_uint32],
            "stwo_hip_bit_reverse_secure_column": [_vp, ctypes.POINTER(_vp), ctypes.c_uint32],
            "stwo_hip_interpolate": [_vp, _vp, Coset, _vp],
            "stwo_hip_interpolate_columns": [_vp, ctypes.POINTER(_vp), _u32p, ctypes.c_size_t, _vp],
            "stwo_hip_evaluate": [_vp, _vp, ctypes.c_uint32, Coset, _vp, _vp],
            "stwo_hip_evaluate_polynomials": [_vp, ctypes.POINTER(_vp), _u32p, ctypes.c_size_t, ctypes.c_uint32, _vp,
                                              ctypes.POINTER(_vp)],
            "stwo_hip_barycentric_weights": [_vp, ctypes.c_uint32, _u32p, ctypes.POINTER(_vp)],
            "stwo_hip_barycentric_eval_at_point": [_vp, ctypes.POINTER(_vp), ctypes.c_size_t, ctypes.c_uint32,
                                                   ctypes.POINTER(_vp), _u32p],
            "stwo_hip_eval_at_point_by_folding": [_vp, _vp, ctypes.c_uint32, _u32p, _vp, _u32p],
            "stwo_hip_split_at_mid": [_vp, _vp, ctypes.c_uint32, _vp, _vp],
            "stwo_hip_fri_decompose": [_vp, ctypes.POINTER(_vp), ctypes.c_uint32, ctypes.POINTER(_vp), _u32p],
            "stwo_hip_lde_commit_per_column": [_vp, ctypes.POINTER(_vp), _u32p, ctypes.c_size_t, ctypes.c_uint32,
                                               _vp, ctypes.POINTER(_vp), _vp, ctypes.c_int, ctypes.c_int],
            "stwo_hip_extend_evals": [_vp, ctypes.POINTER(_vp), _u32p, ctypes.c_size_t, ctypes.c_uint32, _vp,
                                      ctypes.POINTER(_vp), ctypes.c_int],
            "stwo_hip_merkle_commit_on_layer": [_vp, ctypes.c_uint32, _vp, ctypes.POINTER(_vp), ctypes.c_size_t, _vp,
                                                ctypes.c_int],
            "stwo_hip_merkle_commit": [_vp, ctypes.POINTER(_vp), _u32p, ctypes.c_size_t, _vp, ctypes.c_int],
            "stwo_hip_merkle_lifted_build_leaves": [_vp, ctypes.POINTER(_vp), _u32p, ctypes.c_size_t, ctypes.c_uint32,
                                                    _vp, ctypes.c_int],
            "stwo_hip_merkle_lifted_build_next_layer": [_vp, _vp, ctypes.c_uint32, _vp, ctypes.c_int],
            "stwo_hip_merkle_lifted_commit": [_vp, ctypes.POINTER(_vp), _u32p, ctypes.c_size_t, ctypes.c_uint32, _vp,
                                              ctypes.c_int],
            "stwo_hip_fold_circle_into_line": [_vp, ctypes.POINTER(_vp), ctypes.POINTER(_vp), Coset, _u32p, _vp],
            "stwo_hip_fold_line": [_vp, ctypes.POINTER(_vp), Coset, _u32p, ctypes.c_uint32, _vp, ctypes.POINTER(_vp)],
            "stwo_hip_fri_commit": [_vp, ctypes.c_char_p, _u32p, FriConfig, ctypes.POINTER(_vp), ctypes.c_uint32, _vp,
                                    _u32p, _u32p, _u32p],
            "stwo_hip_batch_create": [_vp, ctypes.POINTER(_vp), _u32p, ctypes.c_size_t, ctypes.c_uint32, _vp,
                                      ctypes.POINTER(_vp), _vp, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                      ctypes.POINTER(_vp)],
            "stwo_hip_batch_create_multi": [_vp, _vp, ctypes.c_size_t, _vp, ctypes.c_int, ctypes.POINTER(_vp)],
            "stwo_hip_batch_run": [_vp, _vp],
            "stwo_hip_batch_destroy": [_vp, _vp],
            "stwo_hip_server_create": [ctypes.c_int, ctypes.c_size_t, ctypes.c_uint32, ctypes.c_uint32,
                                       ctypes.POINTER(_vp)],
            "stwo_hip_server_set_max_group": [_vp, ctypes.c_uint32],
            "stwo_hip_server_hold": [_vp, ctypes.c_int],
            "stwo_hip_server_job_bytes": [_vp, ctypes.POINTER(JobDesc), ctypes.POINTER(ctypes.c_size_t)],
            "stwo_hip_server_submit": [_vp, ctypes.POINTER(JobDesc), ctypes.POINTER(ctypes.c_uint64)],
            "stwo_hip_server_poll": [_vp, ctypes.c_uint64, ctypes.POINTER(JobResult)],
            "stwo_hip_server_wait": [_vp, ctypes.c_uint64, ctypes.POINTER(JobResult)],
            "stwo_hip_server_stats": [_vp, ctypes.POINTER(ctypes.c_size_t), ctypes.POINTER(ctypes.c_size_t),
                                      ctypes.POINTER(ctypes.c_uint32), ctypes.POINTER(ctypes.c_uint32)],
            "stwo_hip_server_destroy": [_vp],
            "stwo_hip_prof_enable": [_vp, ctypes.c_int],
            "stwo_hip_prof_reset": [_vp],
            "stwo_hip_prof_read": [_vp, ctypes.c_char_p, ctypes.c_size_t, ctypes.POINTER(ctypes.c_uint64),
                                   ctypes.POINTER(ctypes.c_double), ctypes.c_size_t, ctypes.POINTER(ctypes.c_size_t)],
            "stwo_hip_merkle_lifted_decommit": [_vp, _vp, ctypes.c_uint32, ctypes.POINTER(_vp), _u32p, ctypes.c_size_t,
                                                _u32p, ctypes.c_size_t, _u32p, _u32p, ctypes.c_size_t,
                                                ctypes.POINTER(ctypes.c_size_t), _u32p, ctypes.c_size_t,
                                                ctypes.POINTER(ctypes.c_size_t)],
            "stwo_hip_fri_prover_commit": [_vp, ctypes.c_char_p, _u32p, FriConfig, ctypes.POINTER(_vp),
                                           ctypes.c_uint32, _vp, _u32p, _u32p, _u32p, ctypes.POINTER(_vp)],
            "stwo_hip_fri_prover_decommit": [_vp, _vp, ctypes.c_char_p, _u32p, _u32p, _u32p, _u32p],
            "stwo_hip_fri_prover_layer_proof": [_vp, _vp, ctypes.c_uint32, _u32p, _u32p, ctypes.c_size_t,
                                                ctypes.POINTER(ctypes.c_size_t), _u32p, ctypes.c_size_t,
                                                ctypes.POINTER(ctypes.c_size_t), _u32p, _u32p, ctypes.c_size_t,
                                                ctypes.POINTER(ctypes.c_size_t)],
            "stwo_hip_fri_prover_destroy": [_vp, _vp],
            "stwo_hip_coeffs_simd_transpose": [_vp, _vp, ctypes.c_uint32],
            "stwo_hip_twiddles_simd_doubled": [_vp, _vp, _vp, _vp],
            "stwo_hip_canonicalize": [_vp, _vp, ctypes.c_size_t],
            "stwo_hip_compute_fri_quotients": [_vp, ctypes.POINTER(_vp), _u32p, ctypes.c_size_t, ctypes.c_void_p,
                                               _u32p, _u32p, ctypes.c_uint32, ctypes.POINTER(_vp)],
            "stwo_hip_eval_at_point": [_vp, ctypes.POINTER(_vp), _u32p, ctypes.c_size_t, _u32p, _u32p],
            "stwo_hip_eval_at_point_noalloc": [_vp, ctypes.POINTER(_vp), _u32p, ctypes.c_size_t, _u32p, _u32p],
            "stwo_hip_wide_fib_trace": [_vp, ctypes.POINTER(_vp), ctypes.c_uint32, ctypes.c_uint32, _vp, _vp],
            "stwo_hip_wide_fib_constraint_quotients": [_vp, ctypes.POINTER(_vp), ctypes.c_uint32, ctypes.c_uint32,
                                                       _u32p, ctypes.POINTER(_vp)],
            "stwo_hip_grind": [_vp, ctypes.c_char_p, ctypes.c_uint32, ctypes.POINTER(ctypes.c_uint64)],
            "stwo_hip_wide_fib_prove": [_vp, _vp, _vp, ctypes.c_uint32, ctypes.c_uint32, PcsConfig, _vp,
                                        ctypes.c_char_p, _u32p, ctypes.POINTER(_vp)],
            "stwo_hip_wide_fib_prove_ex": [_vp, _vp, _vp, ctypes.c_uint32, ctypes.c_uint32, PcsConfig, _vp, ctypes.c_int,
                                           ctypes.c_char_p, _u32p, ctypes.POINTER(_vp)],
            "stwo_hip_stark_proof_bytes": [_vp, ctypes.c_char_p, ctypes.c_size_t, ctypes.POINTER(ctypes.c_size_t)],
            "stwo_hip_stark_proof_info": [_vp, ctypes.POINTER(ctypes.c_double), _u32p, ctypes.c_size_t,
                                          ctypes.POINTER(ctypes.c_size_t)],
            "stwo_hip_stark_proof_destroy": [_vp],
            "stwo_hip_debug_fft_plan": [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int, _u32p, ctypes.c_uint32,
                                        _u32p],
            "stwo_hip_debug_fft_plan_tuned": [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int, ctypes.c_uint32,
                                              ctypes.c_uint32, ctypes.c_uint32, _u32p, ctypes.c_uint32, _u32p],
            "stwo_hip_fft_autotune": [_vp, _u32p, ctypes.c_size_t, ctypes.c_uint32, _vp, _u32p, ctypes.c_size_t,
                                      ctypes.POINTER(ctypes.c_size_t)],
            "stwo_hip_fft_tune_set": [_vp, _u32p, ctypes.c_size_t],
            "stwo_hip_fft_tune_clear": [_vp],
        }
        for name, args in sig.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = ctypes.c_int
        L.stwo_hip_stream.argtypes = [_vp]
        L.stwo_hip_stream.restype = _vp
        L.stwo_hip_ctx_fail_allocs_after.argtypes = [_vp, ctypes.c_longlong]
        _lib = L
    return _lib


def _ptrs(bufs):
    arr = (_vp * max(1, len(bufs)))()
    for i, b in enumerate(bufs):
        arr[i] = b.ptr if isinstance(b, DeviceBuf) else b
    return arr


class DeviceBuf:
    """A device allocation of `n` u32 words owned by a Context."""

    def __init__(self, ctx, n, ptr):
        self.ctx, self.n, self.ptr = ctx, n, ptr

    def download(self, count=None, offset=0):
        count = self.n - offset if count is None else count
        out = np.empty(count, dtype=np.uint32)
        self.ctx._check(lib().stwo_hip_download(self.ctx.h, out.ctypes.data, self.ptr + 4 * offset, count))
        return out

    def upload(self, arr, offset=0):
        a = np.ascontiguousarray(arr, dtype=np.uint32)
        assert offset + len(a) <= self.n
        self.ctx._check(lib().stwo_hip_upload(self.ctx.h, self.ptr + 4 * offset, a.ctypes.data, len(a)))

    def at(self, offset_words):
        """Raw device address of word `offset_words` (for sub-buffers)."""
        return self.ptr + 4 * offset_words

    def free(self):
        if self.ptr:
            self.ctx._check(lib().stwo_hip_free(self.ctx.h, self.ptr))
            self.ptr = None


class Twiddles:
    def __init__(self, ctx, handle, root):
        self.ctx, self.h, self.root = ctx, handle, root

    def buffers(self):
        t, it, n = _vp(), _vp(), ctypes.c_size_t()
        lib().stwo_hip_twiddles_buffers(self.h, ctypes.byref(t), ctypes.byref(it), ctypes.byref(n))
        return t.value, it.value, n.value

    def download(self):
        t, it, n = self.buffers()
        tw = np.empty(n, dtype=np.uint32)
        itw = np.empty(n, dtype=np.uint32)
        self.ctx._check(lib().stwo_hip_download(self.ctx.h, tw.ctypes.data, t, n))
        self.ctx._check(lib().stwo_hip_download(self.ctx.h, itw.ctypes.data, it, n))
        return tw, itw

    def destroy(self):
        if self.h:
            self.ctx._check(lib().stwo_hip_twiddles_destroy(self.ctx.h, self.h))
            self.h = None


class Batch:
    """A batched LDE + Merkle plan (stwo_hip_batch)."""

    def __init__(self, ctx, handle):
        self.ctx, self.h = ctx, handle

    def run(self):
        self.ctx._check(lib().stwo_hip_batch_run(self.ctx.h, self.h))

    def destroy(self):
        if self.h:
            self.ctx._check(lib().stwo_hip_batch_destroy(self.ctx.h, self.h))
            self.h = None


class Context:
    """One device + one stream (stwo_hip_ctx)."""

    def __init__(self, device=0):
        h = _vp()
        rc = lib().stwo_hip_ctx_create(device, ctypes.byref(h))
        if rc != 0:
            raise StwoHipError(rc, "stwo_hip_ctx_create(%d) failed (no GPU?)" % device)
        self.h = h.value

    def close(self):
        if self.h:
            lib().stwo_hip_ctx_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def last_error(self):
        buf = ctypes.create_string_buffer(1024)
        lib().stwo_hip_last_error(self.h, buf, len(buf))
        return buf.value.decode()

    def _check(self, rc):
        if rc != 0:
            raise StwoHipError(rc, self.last_error())

    def fail_allocs_after(self, n):
        """Test hook: the next n device allocations succeed, later ones fail
        (STWO_HIP_ENOMEM); n < 0 turns it off."""
        self._check(lib().stwo_hip_ctx_fail_allocs_after(self.h, n))

    # -- memory ---------------------------------------------------------------
    def alloc(self, n):
        p = _vp()
        self._check(lib().stwo_hip_alloc(self.h, n, ctypes.byref(p)))
        return DeviceBuf(self, n, p.value)

    def to_device(self, arr):
        a = np.ascontiguousarray(arr, dtype=np.uint32)
        b = self.alloc(len(a))
        b.upload(a)
        return b

    def zeros(self, n):
        b = self.alloc(n)
        self._check(lib().stwo_hip_memset_zero(self.h, b.ptr, n))
        return b

    def fill_random(self, buf, n, seed, stream, offset=0):
        self._check(lib().stwo_hip_fill_random(self.h, buf.ptr + 4 * offset, n, seed, stream))

    def sync(self):
        self._check(lib().stwo_hip_sync(self.h))

    # -- twiddles -------------------------------------------------------------
    def precompute_twiddles(self, root):
        h = _vp()
        self._check(lib().stwo_hip_twiddles_create(self.h, root, ctypes.byref(h)))
        return Twiddles(self, h.value, root)

    # -- ColumnOps --------------------------------------------------------------
    def bit_reverse_column(self, col, log_size):
        self._check(lib().stwo_hip_bit_reverse_column(self.h, col.ptr if isinstance(col, DeviceBuf) else col, log_size))

    def bit_reverse_secure_column(self, coords, log_size):
        self._check(lib().stwo_hip_bit_reverse_secure_column(self.h, _ptrs(coords), log_size))

    # -- PolyOps ----------------------------------------------------------------
    def interpolate(self, col, domain, tw):
        self._check(lib().stwo_hip_interpolate(self.h, col.ptr if isinstance(col, DeviceBuf) else col, domain, tw.h))

    def interpolate_columns(self, cols, log_sizes, tw):
        ls = np.ascontiguousarray(log_sizes, dtype=np.uint32)
        self._check(lib().stwo_hip_interpolate_columns(self.h, _ptrs(cols), ls.ctypes.data_as(_u32p), len(cols), tw.h))

    def evaluate(self, coeffs, coeffs_log_size, domain, tw, out):
        self._check(lib().stwo_hip_evaluate(self.h, coeffs.ptr if isinstance(coeffs, DeviceBuf) else coeffs,
                                            coeffs_log_size, domain, tw.h,
                                            out.ptr if isinstance(out, DeviceBuf) else out))

    def evaluate_polynomials(self, coeffs, log_sizes, log_blowup, tw, outs):
        ls = np.ascontiguousarray(log_sizes, dtype=np.uint32)
        self._check(lib().stwo_hip_evaluate_polynomials(self.h, _ptrs(coeffs), ls.ctypes.data_as(_u32p), len(coeffs),
                                                        log_blowup, tw.h, _ptrs(outs)))

    # -- the other PolyOps / FriOps methods ---------------------------------------
    @staticmethod
    def _pt(point):
        return np.ascontiguousarray(list(point[0]) + list(point[1]), dtype=np.uint32)

    def barycentric_weights(self, log_size, point, out4):
        pt = self._pt(point)
        self._check(lib().stwo_hip_barycentric_weights(self.h, log_size, pt.ctypes.data_as(_u32p), _ptrs(out4)))

    def barycentric_eval_at_point(self, evals, log_size, weights4):
        """[n][4] values (PolyOps::barycentric_eval_at_point of each column)."""
        out = np.zeros(4 * max(1, len(evals)), dtype=np.uint32)
        self._check(lib().stwo_hip_barycentric_eval_at_point(self.h, _ptrs(evals), len(evals), log_size,
                                                             _ptrs(weights4), out.ctypes.data_as(_u32p)))
        return out[: 4 * len(evals)].reshape(-1, 4)

    def eval_at_point_by_folding(self, evals, log_size, point, tw):
        pt = self._pt(point)
        out = np.zeros(4, dtype=np.uint32)
        self._check(lib().stwo_hip_eval_at_point_by_folding(self.h, evals.ptr if isinstance(evals, DeviceBuf)
                                                            else evals, log_size, pt.ctypes.data_as(_u32p), tw.h,
                                                            out.ctypes.data_as(_u32p)))
        return tuple(int(v) for v in out)

    def split_at_mid(self, coeffs, log_size, left, right):
        self._check(lib().stwo_hip_split_at_mid(self.h, *[b.ptr if isinstance(b, DeviceBuf) else b
                                                          for b in (coeffs,)], log_size,
                                                *[b.ptr if isinstance(b, DeviceBuf) else b for b in (left, right)]))

    def fri_decompose(self, eval4, log_size, g4):
        lam = np.zeros(4, dtype=np.uint32)
        self._check(lib().stwo_hip_fri_decompose(self.h, _ptrs(eval4), log_size, _ptrs(g4),
                                                 lam.ctypes.data_as(_u32p)))
        return tuple(int(v) for v in lam)

    def fft_autotune(self, log_sizes, log_blowup, tw):
        """stwo_hip_fft_autotune: returns [(n, N, inverse, t0, smax, wg, best_us, default_us)]."""
        ls = np.ascontiguousarray(list(log_sizes) or [0], dtype=np.uint32)
        rep = np.zeros(8 * 64, dtype=np.uint32)
        nr = ctypes.c_size_t()
        self._check(lib().stwo_hip_fft_autotune(self.h, ls.ctypes.data_as(_u32p), len(log_sizes), log_blowup, tw.h,
                                                rep.ctypes.data_as(_u32p), 64, ctypes.byref(nr)))
        return [tuple(int(v) for v in rep[8 * i: 8 * i + 8]) for i in range(min(nr.value, 64))]

    def fft_tune_set(self, entries):
        """entries: [(n, N, inverse, t0, smax, wg)]."""
        e = np.ascontiguousarray([v for t in entries for v in t] or [0], dtype=np.uint32)
        self._check(lib().stwo_hip_fft_tune_set(self.h, e.ctypes.data_as(_u32p), len(entries)))

    def fft_tune_clear(self):
        self._check(lib().stwo_hip_fft_tune_clear(self.h))

    def extend_evals(self, cols, log_sizes, log_blowup, tw, outs, store_coeffs=False):
        """interpolate_columns + evaluate_polynomials in one call (TreeBuilder::extend_evals);
        cols are consumed (coefficients afterwards only with store_coeffs)."""
        ls = np.ascontiguousarray(log_sizes, dtype=np.uint32)
        self._check(lib().stwo_hip_extend_evals(self.h, _ptrs(cols), ls.ctypes.data_as(_u32p), len(cols), log_blowup,
                                                tw.h, _ptrs(outs), int(bool(store_coeffs))))

    # -- Merkle -----------------------------------------------------------------
    def commit_on_layer(self, log_size, prev, cols, out, mode=HASH_BLAKE2S):
        pp = (prev.ptr if isinstance(prev, DeviceBuf) else prev) if prev is not None else None
        self._check(lib().stwo_hip_merkle_commit_on_layer(self.h, log_size, pp, _ptrs(cols), len(cols),
                                                          out.ptr if isinstance(out, DeviceBuf) else out, mode))

    def merkle_commit(self, cols, log_sizes, layers, mode=HASH_BLAKE2S):
        ls = np.ascontiguousarray(list(log_sizes) or [0], dtype=np.uint32)
        self._check(lib().stwo_hip_merkle_commit(self.h, _ptrs(cols), ls.ctypes.data_as(_u32p), len(cols),
                                                 layers.ptr if isinstance(layers, DeviceBuf) else layers, mode))

    def lifted_build_leaves(self, cols, log_sizes, lifting_log, out, mode=HASH_BLAKE2S):
        ls = np.ascontiguousarray(list(log_sizes) or [0], dtype=np.uint32)
        self._check(lib().stwo_hip_merkle_lifted_build_leaves(self.h, _ptrs(cols), ls.ctypes.data_as(_u32p),
                                                              len(cols), lifting_log,
                                                              out.ptr if isinstance(out, DeviceBuf) else out, mode))

    def lifted_build_next_layer(self, prev, out_log, out, mode=HASH_BLAKE2S):
        self._check(lib().stwo_hip_merkle_lifted_build_next_layer(
            self.h, prev.ptr if isinstance(prev, DeviceBuf) else prev, out_log,
            out.ptr if isinstance(out, DeviceBuf) else out, mode))

    def lifted_commit(self, cols, log_sizes, lifting_log, layers, mode=HASH_BLAKE2S):
        ls = np.ascontiguousarray(list(log_sizes) or [0], dtype=np.uint32)
        self._check(lib().stwo_hip_merkle_lifted_commit(self.h, _ptrs(cols), ls.ctypes.data_as(_u32p), len(cols),
                                                        lifting_log,
                                                        layers.ptr if isinstance(layers, DeviceBuf) else layers, mode))

    # -- FRI --------------------------------------------------------------------
    def fold_circle_into_line(self, dst4, src4, src_domain, alpha, tw):
        a = np.ascontiguousarray(alpha, dtype=np.uint32)
        self._check(lib().stwo_hip_fold_circle_into_line(self.h, _ptrs(dst4), _ptrs(src4), src_domain,
                                                         a.ctypes.data_as(_u32p), tw.h))

    def fold_line(self, src4, line_domain, alpha, fold_step, tw, dst4):
        a = np.ascontiguousarray(alpha, dtype=np.uint32)
        self._check(lib().stwo_hip_fold_line(self.h, _ptrs(src4), line_domain, a.ctypes.data_as(_u32p), fold_step,
                                             tw.h, _ptrs(dst4)))

    def fri_commit(self, digest: bytes, n_draws: int, config: FriConfig, column4, log_size, tw):
        """FriProver::commit; returns (roots [n][8], last_layer_poly [k][4], digest, n_draws)."""
        d = ctypes.create_string_buffer(bytes(digest), 32)
        nd = np.array([n_draws], dtype=np.uint32)
        roots = np.zeros(8 * (log_size + 1), dtype=np.uint32)
        nr = np.zeros(1, dtype=np.uint32)
        bound = 1 << config.log_last_layer_degree_bound
        last = np.zeros(4 * bound, dtype=np.uint32)
        self._check(lib().stwo_hip_fri_commit(self.h, d, nd.ctypes.data_as(_u32p), config, _ptrs(column4), log_size,
                                              tw.h, roots.ctypes.data_as(_u32p), nr.ctypes.data_as(_u32p),
                                              last.ctypes.data_as(_u32p)))
        n = int(nr[0])
        return roots[: 8 * n].reshape(n, 8), last.reshape(bound, 4), d.raw[:32], int(nd[0])

    def fri_prover_commit(self, digest: bytes, n_draws: int, config: FriConfig, column4, log_size, tw):
        """FriProver::commit keeping the layers; returns (FriProver, roots,
        last_layer_poly, digest, n_draws).  column4 must outlive the prover."""
        d = ctypes.create_string_buffer(bytes(digest), 32)
        nd = np.array([n_draws], dtype=np.uint32)
        roots = np.zeros(8 * (log_size + 1), dtype=np.uint32)
        nr = np.zeros(1, dtype=np.uint32)
        bound = 1 << config.log_last_layer_degree_bound
        last = np.zeros(4 * bound, dtype=np.uint32)
        h = _vp()
        self._check(lib().stwo_hip_fri_prover_commit(self.h, d, nd.ctypes.data_as(_u32p), config, _ptrs(column4),
                                                     log_size, tw.h, roots.ctypes.data_as(_u32p),
                                                     nr.ctypes.data_as(_u32p), last.ctypes.data_as(_u32p),
                                                     ctypes.byref(h)))
        n = int(nr[0])
        pr = FriProver(self, h.value, config, n, column4)
        return pr, roots[: 8 * n].reshape(n, 8), last.reshape(bound, 4), d.raw[:32], int(nd[0])

    # -- SimdBackend layout interop ---------------------------------------------
    def coeffs_simd_transpose(self, col, log_size):
        self._check(lib().stwo_hip_coeffs_simd_transpose(self.h, col.ptr if isinstance(col, DeviceBuf) else col,
                                                         log_size))

    def twiddles_simd_doubled(self, tw, out_tw, out_itw):
        self._check(lib().stwo_hip_twiddles_simd_doubled(self.h, tw.h, out_tw.ptr, out_itw.ptr))

    def canonicalize(self, col, n):
        self._check(lib().stwo_hip_canonicalize(self.h, col.ptr if isinstance(col, DeviceBuf) else col, n))

    def compute_fri_quotients(self, cols, log_sizes, samples_per_col, random_coeff, lifting_log, out4):
        """compute_fri_quotients; samples_per_col: per column a list of
        (x[4], y[4], value[4]) PointSamples; out4: 4 device columns of 2^lifting_log."""
        flat, counts = [], []
        for ss in samples_per_col:
            counts.append(len(ss))
            for x, y, v in ss:
                flat.extend(list(x) + list(y) + list(v))
        fl = np.array(flat or [0], dtype=np.uint32)
        cn = np.array(counts or [0], dtype=np.uint32)
        ls = np.ascontiguousarray(list(log_sizes) or [0], dtype=np.uint32)
        rc = np.ascontiguousarray(random_coeff, dtype=np.uint32)
        self._check(lib().stwo_hip_compute_fri_quotients(self.h, _ptrs(cols), ls.ctypes.data_as(_u32p), len(cols),
                                                         fl.ctypes.data, cn.ctypes.data_as(_u32p),
                                                         rc.ctypes.data_as(_u32p), lifting_log, _ptrs(out4)))

    def eval_at_point(self, coeff_cols, log_sizes, point, noalloc=False):
        """PolyOps::eval_at_point of each polynomial at the secure point (x[4], y[4]); returns [n][4].
        noalloc: through stwo_hip_eval_at_point_noalloc (fixed per-device scratch)."""
        pt = np.ascontiguousarray(list(point[0]) + list(point[1]), dtype=np.uint32)
        ls = np.ascontiguousarray(list(log_sizes) or [0], dtype=np.uint32)
        out = np.zeros(4 * max(1, len(coeff_cols)), dtype=np.uint32)
        fn = lib().stwo_hip_eval_at_point_noalloc if noalloc else lib().stwo_hip_eval_at_point
        self._check(fn(self.h, _ptrs(coeff_cols), ls.ctypes.data_as(_u32p), len(coeff_cols),
                       pt.ctypes.data_as(_u32p), out.ctypes.data_as(_u32p)))
        return out[: 4 * len(coeff_cols)].reshape(-1, 4)

    def wide_fib_trace(self, cols, log_n, a, b):
        self._check(lib().stwo_hip_wide_fib_trace(self.h, _ptrs(cols), len(cols), log_n,
                                                  a.ptr if isinstance(a, DeviceBuf) else a,
                                                  b.ptr if isinstance(b, DeviceBuf) else b))

    def wide_fib_constraint_quotients(self, cols, log_n, random_coeff, accum4):
        rc = np.ascontiguousarray(random_coeff, dtype=np.uint32)
        self._check(lib().stwo_hip_wide_fib_constraint_quotients(self.h, _ptrs(cols), len(cols), log_n,
                                                                 rc.ctypes.data_as(_u32p), _ptrs(accum4)))

    def grind(self, digest: bytes, pow_bits):
        """GrindOps::grind (SimdBackend rule) for the Blake2s channel state `digest`."""
        n = ctypes.c_uint64()
        self._check(lib().stwo_hip_grind(self.h, bytes(digest), pow_bits, ctypes.byref(n)))
        return n.value

    def wide_fib_prove(self, a, b, n_cols, log_n, tw, config=None, digest=bytes(32), n_draws=0, store_coeffs=True):
        """The wide-Fibonacci prove (empty preprocessed tree, trace tree, prove)
        from the device inputs a, b (2^log_n words each).  store_coeffs=True:
        stwo_hip_wide_fib_prove (coefficients kept, as the reference's
        benchmark sets); False: stwo_hip_wide_fib_prove_ex without stored
        coefficients (the PCS default).  Returns (proof bytes, phase times in
        ms by name, sorted query positions, channel digest, channel n_draws)."""
        cfg = config if config is not None else PcsConfig.default()
        d = ctypes.create_string_buffer(bytes(digest), 32)
        nd = np.array([n_draws], dtype=np.uint32)
        h = _vp()
        pa = a.ptr if isinstance(a, DeviceBuf) else a
        pb = b.ptr if isinstance(b, DeviceBuf) else b
        t0 = time.perf_counter()
        if store_coeffs:
            self._check(lib().stwo_hip_wide_fib_prove(self.h, pa, pb, n_cols, log_n, cfg, tw.h, d,
                                                      nd.ctypes.data_as(_u32p), ctypes.byref(h)))
        else:
            self._check(lib().stwo_hip_wide_fib_prove_ex(self.h, pa, pb, n_cols, log_n, cfg, tw.h, 0, d,
                                                         nd.ctypes.data_as(_u32p), ctypes.byref(h)))
        # wall time of the prove call itself (the proof object on the host,
        # before serialization)
        self.last_prove_ms = (time.perf_counter() - t0) * 1e3
        try:
            n = ctypes.c_size_t()
            self._check(lib().stwo_hip_stark_proof_bytes(h, None, 0, ctypes.byref(n)))
            buf = ctypes.create_string_buffer(n.value)
            self._check(lib().stwo_hip_stark_proof_bytes(h, buf, n.value, ctypes.byref(n)))
            ms = (ctypes.c_double * len(PROVE_PHASES))()
            npos = ctypes.c_size_t()
            self._check(lib().stwo_hip_stark_proof_info(h, ms, None, 0, ctypes.byref(npos)))
            pos = np.zeros(max(1, npos.value), dtype=np.uint32)
            self._check(lib().stwo_hip_stark_proof_info(h, ms, pos.ctypes.data_as(_u32p), npos.value,
                                                        ctypes.byref(npos)))
        finally:
            lib().stwo_hip_stark_proof_destroy(h)
        return (buf.raw[: n.value], dict(zip(PROVE_PHASES, list(ms))), [int(x) for x in pos[: npos.value]],
                d.raw[:32], int(nd[0]))

    def merkle_lifted_decommit(self, layers, lifting_log, cols, log_sizes, queries, with_aux=False):
        """MerkleProverLifted::decommit on a packed lifted tree; returns
        (values [n_cols][n_queries], hash_witness [k][8], aux [(layer, index, hash[8])] or None)."""
        q = np.ascontiguousarray(queries, dtype=np.uint32)
        ls = np.ascontiguousarray(list(log_sizes) or [0], dtype=np.uint32)
        nq, nc = len(q), len(cols)
        values = np.zeros(max(1, nc * nq), dtype=np.uint32)
        cap = max(1, nq * (lifting_log + 1))
        wit = np.zeros(8 * cap, dtype=np.uint32)
        nw = ctypes.c_size_t()
        aux = np.zeros(10 * 2 * cap, dtype=np.uint32) if with_aux else None
        na = ctypes.c_size_t()
        self._check(lib().stwo_hip_merkle_lifted_decommit(
            self.h, layers.ptr if isinstance(layers, DeviceBuf) else layers, lifting_log, _ptrs(cols),
            ls.ctypes.data_as(_u32p), nc, q.ctypes.data_as(_u32p), nq, values.ctypes.data_as(_u32p),
            wit.ctypes.data_as(_u32p), cap, ctypes.byref(nw),
            aux.ctypes.data_as(_u32p) if with_aux else None, 2 * cap, ctypes.byref(na)))
        vals = values[: nc * nq].reshape(nc, nq) if nc else np.zeros((0, nq), dtype=np.uint32)
        w = wit[: 8 * nw.value].reshape(-1, 8)
        a = None
        if with_aux:
            r = aux[: 10 * na.value].reshape(-1, 10)
            a = [(int(x[0]), int(x[1]), tuple(int(v) for v in x[2:])) for x in r]
        return vals, w, a

    # -- batched commitment -------------------------------------------------------
    def batch_create(self, cols, log_sizes, log_blowup, tw, outs, layers, tree_kind=TREE_MIXED,
                     mode=HASH_BLAKE2S, use_graph=True, drop_coeffs=False):
        """stwo_hip_batch_create; drop_coeffs: the fused pass does not store the
        coefficients (cols are left unspecified)."""
        ls = np.ascontiguousarray(log_sizes, dtype=np.uint32)
        h = _vp()
        flags = (BATCH_GRAPH if use_graph else 0) | (BATCH_DROP_COEFFS if drop_coeffs else 0)
        self._check(lib().stwo_hip_batch_create(self.h, _ptrs(cols), ls.ctypes.data_as(_u32p), len(cols), log_blowup,
                                                tw.h, _ptrs(outs),
                                                layers.ptr if isinstance(layers, DeviceBuf) else layers,
                                                tree_kind, mode, flags, ctypes.byref(h)))
        return Batch(self, h.value)

    def batch_create_multi(self, sets, tw, use_graph=True, drop_coeffs=False):
        """stwo_hip_batch_create_multi; sets: [(cols, log_sizes, log_blowup, outs, layers, tree_kind, mode[,
        stage])] (outs / layers None where the stage has none)."""
        keep, arr = [], (BatchSet * max(1, len(sets)))()
        for i, st in enumerate(sets):
            cols, logs, b, outs, layers, kind, mode = st[:7]
            stage = st[7] if len(st) > 7 else BATCH_LDE
            cp = _ptrs(cols)
            op = _ptrs(outs) if outs is not None else None
            ls = np.ascontiguousarray(logs, dtype=np.uint32)
            keep += [cp, op, ls]
            arr[i] = BatchSet(ctypes.cast(cp, ctypes.c_void_p), ls.ctypes.data, len(cols), b,
                              ctypes.cast(op, ctypes.c_void_p) if op is not None else None,
                              layers.ptr if isinstance(layers, DeviceBuf) else layers, kind, mode, stage)
        h = _vp()
        flags = (BATCH_GRAPH if use_graph else 0) | (BATCH_DROP_COEFFS if drop_coeffs else 0)
        self._check(lib().stwo_hip_batch_create_multi(self.h, ctypes.addressof(arr), len(sets), tw.h, flags,
                                                      ctypes.byref(h)))
        del keep
        return Batch(self, h.value)

    def lde_commit_per_column(self, cols, log_sizes, log_blowup, tw, outs, layers, tree_kind=TREE_MIXED,
                              mode=HASH_BLAKE2S):
        """stwo_hip_lde_commit_per_column: the plain per-column form of a batch plan's work."""
        ls = np.ascontiguousarray(log_sizes, dtype=np.uint32)
        self._check(lib().stwo_hip_lde_commit_per_column(self.h, _ptrs(cols), ls.ctypes.data_as(_u32p), len(cols),
                                                         log_blowup, tw.h, _ptrs(outs),
                                                         layers.ptr if isinstance(layers, DeviceBuf) else layers,
                                                         tree_kind, mode))

    # -- profiling --------------------------------------------------------------
    def prof_enable(self, on=True):
        self._check(lib().stwo_hip_prof_enable(self.h, int(on)))

    def prof_reset(self):
        self._check(lib().stwo_hip_prof_reset(self.h))

    def prof_read(self):
        """{kernel name: (launches, total ms)} since the last reset."""
        names = ctypes.create_string_buffer(8192)
        counts = (ctypes.c_uint64 * 64)()
        ms = (ctypes.c_double * 64)()
        n = ctypes.c_size_t()
        self._check(lib().stwo_hip_prof_read(self.h, names, len(names), counts, ms, 64, ctypes.byref(n)))
        keys = names.raw.split(b"\0")[: n.value]
        return {k.decode(): (int(counts[i]), float(ms[i])) for i, k in enumerate(keys)}


class FriProver:
    """A committed FRI prover (stwo_hip_fri_prover) kept for decommitment."""

    def __init__(self, ctx, h, config, n_layers, column4):
        self.ctx, self.h, self.config, self.n_layers = ctx, h, config, n_layers
        self._column = column4  # borrowed by the prover

    def decommit(self, digest: bytes, n_draws: int):
        """FriProver::decommit; returns (sorted query positions, raw queries, digest, n_draws)."""
        d = ctypes.create_string_buffer(bytes(digest), 32)
        nd = np.array([n_draws], dtype=np.uint32)
        nq = self.config.n_queries
        pos = np.zeros(nq, dtype=np.uint32)
        npos = np.zeros(1, dtype=np.uint32)
        raw = np.zeros(nq, dtype=np.uint32)
        self.ctx._check(lib().stwo_hip_fri_prover_decommit(self.ctx.h, self.h, d, nd.ctypes.data_as(_u32p),
                                                           pos.ctypes.data_as(_u32p), npos.ctypes.data_as(_u32p),
                                                           raw.ctypes.data_as(_u32p)))
        return pos[: int(npos[0])], raw, d.raw[:32], int(nd[0])

    def layer_proof(self, layer):
        """FriLayerProof of committed layer `layer` as a dict: commitment [8],
        fri_witness [k][4], hash_witness [k][8], positions, values [k][4]."""
        com = np.zeros(8, dtype=np.uint32)
        nf, nh, npos = ctypes.c_size_t(), ctypes.c_size_t(), ctypes.c_size_t()
        # A first call with no capacity reports the counts.
        lib().stwo_hip_fri_prover_layer_proof(self.ctx.h, self.h, layer, com.ctypes.data_as(_u32p), None, 0,
                                              ctypes.byref(nf), None, 0, ctypes.byref(nh), None, None, 0,
                                              ctypes.byref(npos))
        fw = np.zeros(4 * nf.value + 1, dtype=np.uint32)
        hw = np.zeros(8 * nh.value + 1, dtype=np.uint32)
        ps = np.zeros(npos.value + 1, dtype=np.uint32)
        vs = np.zeros(4 * npos.value + 1, dtype=np.uint32)
        self.ctx._check(lib().stwo_hip_fri_prover_layer_proof(
            self.ctx.h, self.h, layer, com.ctypes.data_as(_u32p), fw.ctypes.data_as(_u32p), nf.value,
            ctypes.byref(nf), hw.ctypes.data_as(_u32p), nh.value, ctypes.byref(nh), ps.ctypes.data_as(_u32p),
            vs.ctypes.data_as(_u32p), npos.value, ctypes.byref(npos)))
        return dict(commitment=com, fri_witness=fw[: 4 * nf.value].reshape(-1, 4),
                    hash_witness=hw[: 8 * nh.value].reshape(-1, 8), positions=ps[: npos.value],
                    values=vs[: 4 * npos.value].reshape(-1, 4))

    def destroy(self):
        if self.h:
            self.ctx._check(lib().stwo_hip_fri_prover_destroy(self.ctx.h, self.h))
            self.h = None


def fft_plan(n, N, inverse, t0=0, smax=0, wg=0):
    """The FFT passes the library runs for a transform of 2^n values (RFFT:
    coefficients) on a domain of log size N: one dict per pass, in launch
    order (stwo_hip_debug_fft_plan_tuned; launches nothing).  t0, smax, wg:
    planner choices as for Context.fft_tune_set (0 = default)."""
    buf = (ctypes.c_uint32 * (16 * 8))()
    k = ctypes.c_uint32()
    rc = lib().stwo_hip_debug_fft_plan_tuned(n, N, int(inverse), t0, smax, wg, buf, 8, ctypes.byref(k))
    if rc != 0:
        raise StwoHipError(rc, "fft plan (n=%d, N=%d) rejected" % (n, N))
    out = []
    for i in range(k.value):
        w = list(buf[16 * i: 16 * i + 16])
        stages = [(v >> 16, (v >> 8) & 0xFF, v & 0xFF) for v in w[9: 9 + w[8]]]
        out.append(dict(T=w[0], lo=w[1], w=w[2], m=w[3], rep=w[4], wg=w[5], in_log=w[6], scaled=w[7],
                        stages=stages, wgs=w[13], vec_io=w[14], fixed=w[15]))
    return out


def layer_offset(l):
    """Word offset of layer l in the packed tree layout (root first)."""
    return ((1 << l) - 1) * 8


def tree_words(max_log):
    return ((2 << max_log) - 1) * 8


class Server:
    """Continuous-batching server of commit + FRI and prove jobs (stwo_hip_server)."""

    def __init__(self, device=0, hbm_budget_bytes=8 << 30, max_log_size=24, n_slots=4, max_group=None):
        h = _vp()
        rc = lib().stwo_hip_server_create(device, hbm_budget_bytes, max_log_size, n_slots, ctypes.byref(h))
        if rc != 0:
            raise StwoHipError(rc, "stwo_hip_server_create failed")
        self.h = h.value
        self._keep = {}  # caller data of pending jobs (kept alive until wait)
        if max_group is not None:
            rc = lib().stwo_hip_server_set_max_group(self.h, max_group)
            if rc != 0:
                raise StwoHipError(rc, "set_max_group rejected")

    def _desc(self, kind, n_cols, log_size, log_blowup, seed, fri=None, pcs=None, cols=None, host=None,
              proof_cap=0):
        d = JobDesc()
        d.kind, d.n_cols, d.log_size, d.log_blowup, d.seed = kind, n_cols, log_size, log_blowup, seed
        keep = []
        if cols is not None:
            arr = (ctypes.c_void_p * len(cols))(*[c.ptr if isinstance(c, DeviceBuf) else c for c in cols])
            keep.append(arr)
            d.cols = ctypes.cast(arr, ctypes.c_void_p)
        if host is not None:
            h = np.ascontiguousarray(host, dtype=np.uint32)
            keep.append(h)
            d.host = h.ctypes.data
        if fri is not None:
            d.fri = fri
        if pcs is not None:
            d.pcs = pcs
        if proof_cap:
            buf = (ctypes.c_uint8 * proof_cap)()
            keep.append(buf)
            d.proof_out = ctypes.cast(buf, ctypes.c_void_p)
            d.proof_cap = proof_cap
        return d, keep

    def _submit(self, d, keep):
        jid = ctypes.c_uint64()
        rc = lib().stwo_hip_server_submit(self.h, ctypes.byref(d), ctypes.byref(jid))
        if rc != 0:
            raise StwoHipError(rc, "submit rejected")
        self._keep[jid.value] = keep
        return jid.value

    def hold(self, on=True):
        """stwo_hip_server_hold: while held, submitted jobs queue without being admitted."""
        rc = lib().stwo_hip_server_hold(self.h, int(bool(on)))
        if rc != 0:
            raise StwoHipError(rc, "hold rejected")

    def submit(self, n_cols, log_size, log_blowup, seed, fri, cols=None, host=None):
        """A commit + FRI job: seeded columns, or the caller's device columns
        `cols` / host data `host` (n_cols x 2^log_size words)."""
        d, keep = self._desc(JOB_COMMIT_FRI, n_cols, log_size, log_blowup, seed, fri=fri, cols=cols, host=host)
        return self._submit(d, keep)

    def submit_prove(self, n_cols, log_size, seed, pcs=None, cols=None, host=None, proof_cap=1 << 20):
        """A wide-Fibonacci prove job; wait() then returns (result, proof bytes)."""
        pcs = pcs or PcsConfig.default()
        d, keep = self._desc(JOB_PROVE, n_cols, log_size, pcs.fri.log_blowup_factor, seed, pcs=pcs, cols=cols,
                             host=host, proof_cap=proof_cap)
        return self._submit(d, keep)

    def job_bytes(self, n_cols, log_size, log_blowup, fri=None, pcs=None, kind=JOB_COMMIT_FRI):
        d, keep = self._desc(kind, n_cols, log_size, log_blowup, 1, fri=fri, pcs=pcs,
                             proof_cap=1 if kind == JOB_PROVE else 0)
        b = ctypes.c_size_t()
        rc = lib().stwo_hip_server_job_bytes(self.h, ctypes.byref(d), ctypes.byref(b))
        if rc != 0:
            raise StwoHipError(rc, "job_bytes rejected")
        return b.value

    def poll(self, jid):
        r = JobResult()
        rc = lib().stwo_hip_server_poll(self.h, jid, ctypes.byref(r))
        return r if rc == 0 else None

    def wait(self, jid):
        r = JobResult()
        rc = lib().stwo_hip_server_wait(self.h, jid, ctypes.byref(r))
        keep = self._keep.pop(jid, [])
        if rc != 0:
            raise StwoHipError(rc, "wait failed")
        for k in keep:
            if isinstance(k, ctypes.Array) and k._type_ is ctypes.c_uint8:
                return r, bytes(k[: r.proof_len])
        return r

    def stats(self):
        u, c = ctypes.c_size_t(), ctypes.c_size_t()
        p, r = ctypes.c_uint32(), ctypes.c_uint32()
        lib().stwo_hip_server_stats(self.h, ctypes.byref(u), ctypes.byref(c), ctypes.byref(p), ctypes.byref(r))
        return {"pool_used": u.value, "pool_capacity": c.value, "pending": p.value, "running": r.value}

    def close(self):
        if self.h:
            lib().stwo_hip_server_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

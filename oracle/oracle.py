"""ctypes wrapper over the CPU oracle (liboracle.so built from stwo_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker.  The product (the HIP library)
never imports or links this module.

All columns are numpy uint32 arrays; QM31 columns are lists of 4 coordinate
arrays (SecureColumnByCoords, STWO/prover/secure_column.rs); hashes are
uint32 arrays of 8 words (32 little-endian bytes) per node.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# STWO_ORACLE_LIB: another build of the same source (the sanitizer build).
_LIB_PATH = os.environ.get("STWO_ORACLE_LIB") or os.path.join(_HERE, "liboracle.so")
P = (1 << 31) - 1

_u32p = ctypes.POINTER(ctypes.c_uint32)
_lib = None


def build():
    subprocess.run(["make", "-s", "-C", _HERE, "-B", "liboracle.so"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.or_m31_mul.restype = ctypes.c_uint32
        L.or_m31_add.restype = ctypes.c_uint32
        L.or_m31_sub.restype = ctypes.c_uint32
        L.or_m31_inv.restype = ctypes.c_uint32
        L.or_half_odds_initial.restype = ctypes.c_uint32
        for name in ["or_m31_mul", "or_m31_add", "or_m31_sub"]:
            getattr(L, name).argtypes = [ctypes.c_uint32, ctypes.c_uint32]
        L.or_m31_inv.argtypes = [ctypes.c_uint32]
        L.or_half_odds_initial.argtypes = [ctypes.c_uint32]
        L.or_interpolate.argtypes = [_u32p, ctypes.c_uint32, ctypes.c_uint32, _u32p, ctypes.c_size_t]
        L.or_evaluate.argtypes = [_u32p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, _u32p,
                                  ctypes.c_size_t, _u32p]
        L.or_coset_at.argtypes = [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint64, _u32p]
        L.or_index_to_point.argtypes = [ctypes.c_uint32, _u32p]
        L.or_fill_random.argtypes = [_u32p, ctypes.c_size_t, ctypes.c_uint64, ctypes.c_uint64]
        L.or_commit_on_layer.argtypes = [ctypes.c_uint32, _u32p, ctypes.POINTER(_u32p), ctypes.c_size_t,
                                         _u32p, ctypes.c_int]
        L.or_merkle_commit.argtypes = [ctypes.POINTER(_u32p), _u32p, ctypes.c_size_t, _u32p, ctypes.c_int]
        L.or_lifted_build_leaves.argtypes = [ctypes.POINTER(_u32p), _u32p, ctypes.c_size_t, ctypes.c_uint32,
                                             _u32p, ctypes.c_int]
        L.or_lifted_build_next_layer.argtypes = [_u32p, ctypes.c_uint32, _u32p, ctypes.c_int]
        L.or_lifted_commit.argtypes = [ctypes.POINTER(_u32p), _u32p, ctypes.c_size_t, ctypes.c_uint32, _u32p,
                                       ctypes.c_int]
        L.or_fold_line.argtypes = [ctypes.POINTER(_u32p), ctypes.c_uint32, ctypes.c_uint32, _u32p,
                                   ctypes.c_uint32, ctypes.POINTER(_u32p)]
        L.or_fold_circle_into_line.argtypes = [ctypes.POINTER(_u32p), ctypes.POINTER(_u32p), ctypes.c_uint32,
                                               ctypes.c_uint32, _u32p]
        L.or_interpolate_columns.argtypes = [ctypes.POINTER(_u32p), _u32p, ctypes.c_size_t, _u32p,
                                             ctypes.c_size_t]
        L.or_evaluate_polynomials.argtypes = [ctypes.POINTER(_u32p), _u32p, ctypes.c_size_t, ctypes.c_uint32,
                                              _u32p, ctypes.c_size_t, ctypes.POINTER(_u32p)]
        L.or_blake2s.argtypes = [ctypes.c_char_p, ctypes.c_size_t, ctypes.c_char_p]
        L.or_num_threads.restype = ctypes.c_int
        L.or_compute_fri_quotients.argtypes = [ctypes.POINTER(_u32p), _u32p, ctypes.c_size_t, ctypes.c_void_p, _u32p,
                                               _u32p, ctypes.c_uint32, ctypes.POINTER(_u32p)]
        L.or_eval_at_point.argtypes = [_u32p, ctypes.c_uint32, _u32p, _u32p, _u32p]
        L.or_wide_fib_trace.argtypes = [ctypes.POINTER(_u32p), ctypes.c_uint32, ctypes.c_uint32, _u32p, _u32p]
        L.or_wide_fib_constraint_quotients.argtypes = [ctypes.POINTER(_u32p), ctypes.c_uint32, ctypes.c_uint32,
                                                       _u32p, ctypes.POINTER(_u32p)]
        L.or_pow_prefixed_digest.argtypes = [ctypes.c_char_p, ctypes.c_uint32, ctypes.c_char_p]
        L.or_grind.argtypes = [ctypes.c_char_p, ctypes.c_uint32, ctypes.c_uint32, ctypes.POINTER(ctypes.c_uint64)]
        _lib = L
    return _lib


def _p(a):
    assert a.dtype == np.uint32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_u32p)


def _ptr_array(arrays):
    arr = (_u32p * max(1, len(arrays)))()
    for i, a in enumerate(arrays):
        arr[i] = _p(a)
    return arr


def half_odds_initial(n):
    """Coset::half_odds(n).initial_index = 2^(29-n) (STWO/core/circle.rs:328-330)."""
    return 1 << (29 - n)


# --- fields / circle -------------------------------------------------------

def m31_mul(a, b):
    return lib().or_m31_mul(a, b)


def m31_inv(a):
    return lib().or_m31_inv(a)


def qm31_mul(a, b):
    A = np.array(a, dtype=np.uint32)
    B = np.array(b, dtype=np.uint32)
    out = np.zeros(4, dtype=np.uint32)
    lib().or_qm31_mul(_p(A), _p(B), _p(out))
    return [int(v) for v in out]


def qm31_inv(a):
    A = np.array(a, dtype=np.uint32)
    out = np.zeros(4, dtype=np.uint32)
    lib().or_qm31_inv(_p(A), _p(out))
    return [int(v) for v in out]


def index_to_point(idx):
    out = np.zeros(2, dtype=np.uint32)
    lib().or_index_to_point(idx, _p(out))
    return int(out[0]), int(out[1])


def coset_at(initial, log_size, i):
    out = np.zeros(2, dtype=np.uint32)
    lib().or_coset_at(initial, log_size, i, _p(out))
    return int(out[0]), int(out[1])


def bit_reverse(v):
    v = np.ascontiguousarray(v, dtype=np.uint32).copy()
    lib().or_bit_reverse(_p(v), int(len(v)).bit_length() - 1)
    return v


def fill_random(n, seed, stream):
    out = np.empty(n, dtype=np.uint32)
    lib().or_fill_random(_p(out), n, seed, stream)
    return out


# --- twiddles / FFT --------------------------------------------------------

def precompute_twiddles(initial, log_size):
    """slow_precompute_twiddles + inverses (CPUB/circle.rs:252-311)."""
    n = 1 << log_size
    tw = np.zeros(n, dtype=np.uint32)
    itw = np.zeros(n, dtype=np.uint32)
    lib().or_precompute_twiddles(ctypes.c_uint32(initial), ctypes.c_uint32(log_size), _p(tw), _p(itw))
    return tw, itw


def interpolate(values, domain_initial, log_size, itw):
    v = np.ascontiguousarray(values, dtype=np.uint32).copy()
    rc = lib().or_interpolate(_p(v), domain_initial, log_size, _p(itw), len(itw))
    assert rc == 0, rc
    return v


def evaluate(coeffs, domain_initial, log_size, tw):
    c = np.ascontiguousarray(coeffs, dtype=np.uint32)
    out = np.zeros(1 << log_size, dtype=np.uint32)
    rc = lib().or_evaluate(_p(c), int(len(c)).bit_length() - 1, domain_initial, log_size, _p(tw), len(tw), _p(out))
    assert rc == 0, rc
    return out


def interpolate_columns(cols, itw):
    """In place over canonic domains (TreeBuilder::extend_evals's call)."""
    logs = np.array([int(len(c)).bit_length() - 1 for c in cols], dtype=np.uint32)
    rc = lib().or_interpolate_columns(_ptr_array(cols), _p(logs), len(cols), _p(itw), len(itw))
    assert rc == 0


def evaluate_polynomials(coeffs, log_blowup, tw):
    logs = np.array([int(len(c)).bit_length() - 1 for c in coeffs], dtype=np.uint32)
    outs = [np.empty(len(c) << log_blowup, dtype=np.uint32) for c in coeffs]
    rc = lib().or_evaluate_polynomials(_ptr_array(coeffs), _p(logs), len(coeffs), log_blowup, _p(tw), len(tw),
                                       _ptr_array(outs))
    assert rc == 0
    return outs


# --- Blake2s / Merkle -------------------------------------------------------

def blake2s(data: bytes) -> bytes:
    out = ctypes.create_string_buffer(32)
    lib().or_blake2s(data, len(data), out)
    return out.raw


def commit_on_layer(log_size, prev, cols, m31_output=False):
    out = np.zeros(8 << log_size, dtype=np.uint32)
    prevp = _p(prev) if prev is not None else None
    lib().or_commit_on_layer(log_size, prevp, _ptr_array(cols), len(cols), _p(out), int(m31_output))
    return out


def layer_offset(l):
    """Word offset of layer l in the packed layout (root first)."""
    return ((1 << l) - 1) * 8


def merkle_commit(cols, m31_output=False):
    """MerkleProver::commit; returns packed layers (root first)."""
    logs = np.array([int(len(c)).bit_length() - 1 for c in cols], dtype=np.uint32)
    max_log = int(logs.max()) if len(cols) else 0
    out = np.zeros(((2 << max_log) - 1) * 8, dtype=np.uint32)
    lib().or_merkle_commit(_ptr_array(cols), _p(logs) if len(cols) else None, len(cols), _p(out), int(m31_output))
    return out


def lifted_build_leaves(cols, lifting_log, m31_output=False):
    logs = np.array([int(len(c)).bit_length() - 1 for c in cols] or [0], dtype=np.uint32)
    out = np.zeros(8 << (lifting_log if cols else 0), dtype=np.uint32)
    rc = lib().or_lifted_build_leaves(_ptr_array(cols), _p(logs), len(cols), lifting_log, _p(out), int(m31_output))
    assert rc == 0
    return out


def lifted_commit(cols, lifting_log, m31_output=False):
    logs = np.array([int(len(c)).bit_length() - 1 for c in cols] or [0], dtype=np.uint32)
    out = np.zeros(((2 << lifting_log) - 1) * 8, dtype=np.uint32)
    rc = lib().or_lifted_commit(_ptr_array(cols), _p(logs), len(cols), lifting_log, _p(out), int(m31_output))
    assert rc == 0
    return out


# --- FRI --------------------------------------------------------------------

def fold_line(src4, line_initial, log_n, alpha, fold_step=1):
    a = np.array(alpha, dtype=np.uint32)
    dst = [np.zeros(1 << (log_n - fold_step), dtype=np.uint32) for _ in range(4)]
    rc = lib().or_fold_line(_ptr_array([np.ascontiguousarray(s, dtype=np.uint32) for s in src4]), line_initial,
                            log_n, _p(a), fold_step, _ptr_array(dst))
    assert rc == 0
    return dst


def fold_circle_into_line(dst4, src4, domain_initial, log_n, alpha):
    """Accumulates into dst4 (modified in place) and returns it."""
    a = np.array(alpha, dtype=np.uint32)
    rc = lib().or_fold_circle_into_line(_ptr_array(dst4), _ptr_array(
        [np.ascontiguousarray(s, dtype=np.uint32) for s in src4]), domain_initial, log_n, _p(a))
    assert rc == 0
    return dst4


class Channel(ctypes.Structure):
    """Blake2sChannel state (STWO/core/channel/blake2s.rs:20-24)."""
    _fields_ = [("digest", ctypes.c_uint8 * 32), ("n_draws", ctypes.c_uint32)]

    def mix_root(self, root_words):
        r = np.ascontiguousarray(root_words, dtype=np.uint32)
        lib().or_channel_mix_root(ctypes.byref(self), r.ctypes.data_as(ctypes.c_char_p))

    def mix_u32s(self, words):
        w = np.ascontiguousarray(words, dtype=np.uint32)
        lib().or_channel_mix_u32s(ctypes.byref(self), _p(w), ctypes.c_size_t(len(w)))

    def mix_u64(self, v):
        self.mix_u32s([v & 0xFFFFFFFF, v >> 32])

    def mix_felts(self, felts):
        f = np.ascontiguousarray(np.array(felts, dtype=np.uint32).reshape(-1))
        lib().or_channel_mix_felts(ctypes.byref(self), _p(f), ctypes.c_size_t(len(f) // 4))

    def draw_u32s(self):
        out = np.zeros(8, dtype=np.uint32)
        lib().or_channel_draw_u32s(ctypes.byref(self), _p(out))
        return [int(x) for x in out]

    def draw_secure_felt(self):
        out = np.zeros(4, dtype=np.uint32)
        lib().or_channel_draw_secure_felt(ctypes.byref(self), _p(out))
        return [int(x) for x in out]

    def digest_bytes(self):
        return bytes(self.digest)


class FriConfig(ctypes.Structure):
    """FriConfig (STWO/core/fri.rs:31-79)."""
    _fields_ = [("log_blowup_factor", ctypes.c_uint32), ("log_last_layer_degree_bound", ctypes.c_uint32),
                ("n_queries", ctypes.c_uint32), ("line_fold_step", ctypes.c_uint32)]


def fri_commit(channel, cfg, col4, log_size, keep_layers=False):
    """FriProver::commit; returns (roots [n][8], last_layer_poly [k][4], layers)."""
    max_layers = log_size + 1
    roots = np.zeros(8 * max_layers, dtype=np.uint32)
    bound = 1 << cfg.log_last_layer_degree_bound
    last = np.zeros(4 * bound, dtype=np.uint32)
    layer_bufs = None
    lp = None
    if keep_layers:
        layer_bufs = [np.zeros(4 << (log_size - 1 - i), dtype=np.uint32) for i in range(log_size)]
        lp = _ptr_array(layer_bufs)
    n = lib().or_fri_commit(ctypes.byref(channel), ctypes.byref(cfg),
                            _ptr_array([np.ascontiguousarray(c, dtype=np.uint32) for c in col4]),
                            log_size, _p(roots), _p(last), lp)
    if n < 0:
        raise ValueError("fri_commit failed: %d" % n)
    return roots[: 8 * n].reshape(n, 8), last.reshape(bound, 4), layer_bufs


def eval_at_point(coeffs, point):
    """CpuBackend::eval_at_point at a secure point ((x[4], y[4]))."""
    c = np.ascontiguousarray(coeffs, dtype=np.uint32)
    x = np.array(point[0], dtype=np.uint32)
    y = np.array(point[1], dtype=np.uint32)
    out = np.zeros(4, dtype=np.uint32)
    lib().or_eval_at_point(_p(c), int(len(c)).bit_length() - 1, _p(x), _p(y), _p(out))
    return tuple(int(v) for v in out)


# --- wide-Fibonacci AIR ----------------------------------------------------------

def wide_fib_trace(a, b, n_cols):
    a = np.ascontiguousarray(a, dtype=np.uint32)
    b = np.ascontiguousarray(b, dtype=np.uint32)
    log_n = int(len(a)).bit_length() - 1
    cols = [np.zeros(len(a), dtype=np.uint32) for _ in range(n_cols)]
    lib().or_wide_fib_trace(_ptr_array(cols), n_cols, log_n, _p(a), _p(b))
    return cols


def wide_fib_constraint_quotients(ext_cols, log_n, alpha, accum=None):
    """Adds the component's constraint quotients to accum (4 columns of 2^(log_n+1)); returns accum."""
    if accum is None:
        accum = [np.zeros(2 << log_n, dtype=np.uint32) for _ in range(4)]
    al = np.array(alpha, dtype=np.uint32)
    rc = lib().or_wide_fib_constraint_quotients(_ptr_array(ext_cols), len(ext_cols), log_n, _p(al), _ptr_array(accum))
    assert rc == 0
    return accum


# --- FRI quotients -------------------------------------------------------------

def pack_samples(samples_per_col):
    """[(x[4], y[4], value[4]), ...] per column -> (flat u32 array of 12 words per
    sample, counts) in the C layout {point.x, point.y, value}."""
    flat, counts = [], []
    for ss in samples_per_col:
        counts.append(len(ss))
        for x, y, v in ss:
            flat.extend(list(x) + list(y) + list(v))
    return np.array(flat or [0], dtype=np.uint32), np.array(counts or [0], dtype=np.uint32)


def compute_fri_quotients(cols, samples_per_col, random_coeff, lifting_log):
    """compute_fri_quotients over the flattened columns; returns 4 coordinate columns of 2^lifting_log."""
    logs = np.array([int(len(c)).bit_length() - 1 for c in cols], dtype=np.uint32)
    flat, counts = pack_samples(samples_per_col)
    rc = np.array(random_coeff, dtype=np.uint32)
    out = [np.zeros(1 << lifting_log, dtype=np.uint32) for _ in range(4)]
    r = lib().or_compute_fri_quotients(_ptr_array(cols), _p(logs), len(cols), flat.ctypes.data, _p(counts), _p(rc),
                                       lifting_log, _ptr_array(out))
    if r != 0:
        raise ValueError("compute_fri_quotients rejected its input")
    return out


# --- proof of work -------------------------------------------------------------

def pow_prefixed_digest(digest: bytes, pow_bits):
    out = ctypes.create_string_buffer(32)
    lib().or_pow_prefixed_digest(bytes(digest), pow_bits, out)
    return out.raw


def grind(digest: bytes, pow_bits, low_bits=20):
    """SimdBackend's grind for the Blake2s channel: smallest (hi << 32) | lo, lo < 2^low_bits."""
    n = ctypes.c_uint64()
    rc = lib().or_grind(bytes(digest), pow_bits, low_bits, ctypes.byref(n))
    if rc != 0:
        raise ValueError("grind failed")
    return n.value


def num_threads():
    return lib().or_num_threads()

"""CPU restatement of SimdBackend's data layouts (test infrastructure only:
imported by tests/ as the checker).

transpose_vecs follows STWO/prover/backend/simd/fft/mod.rs:22-56 (numpy over
16-word vectors); doubled twiddles follow STWO/prover/backend/simd/circle.rs:449-486.
"""
import numpy as np

LOG_N_LANES = 4
CACHED_FFT_LOG_SIZE = 16


def transpose_vecs(values, log_n_vecs):
    v = np.array(values, dtype=np.uint32).reshape(1 << log_n_vecs, 1 << LOG_N_LANES).copy()
    half = log_n_vecs // 2
    out = v.copy()
    for a in range(1 << half):
        for b in range(1 << (log_n_vecs & 1)):
            for c in range(1 << half):
                i = (a << (log_n_vecs - half)) | (b << half) | c
                j = (c << (log_n_vecs - half)) | (b << half) | a
                out[i], out[j] = v[j], v[i]
    return out.reshape(-1)


def simd_coeff_order(natural, log_size):
    """SimdBackend's storage order of natural-order coefficients."""
    if log_size <= CACHED_FFT_LOG_SIZE:
        return np.array(natural, dtype=np.uint32)
    return transpose_vecs(natural, log_size - LOG_N_LANES)


def doubled(tw):
    return (np.array(tw, dtype=np.uint64) * 2).astype(np.uint32)

// Data-layout interop with SimdBackend, so columns can move between this
// library and a SimdBackend-driven prover.
//
// * Coefficient order.  Above log 16 (CACHED_FFT_LOG_SIZE) SimdBackend's IFFT
//   stops with its 16-element vectors transposed: vector index bits abc are
//   stored at cba, |a| = |c| = floor(v / 2), |b| = v mod 2, v = log2 of the
//   number of vectors (transpose_vecs, SIMD/fft/mod.rs:22-56 called from
//   SIMD/fft/ifft.rs:42-56; SIMD/circle.rs:74-82 describes the order and
//   split_at_mid transposes back before splitting, SIMD/circle.rs:503-512).
//   This library keeps natural order.  The map is an involution, so one entry
//   point converts both ways.
// * Twiddles.  SimdBackend stores each twiddle doubled, 2t as u32, in the same
//   tree layout (SIMD/circle.rs:449-486).
// * Redundant values.  SimdBackend's M31 values may be P, the redundant zero
//   (SIMD/m31.rs:58); this library takes canonical values.
#include "common.hpp"
#include "field.hpp"

namespace stwo_hip {

constexpr uint32_t kSimdLogLanes = 4;     // 16 M31 per SimdBackend vector
constexpr uint32_t kSimdCachedLog = 16;   // CACHED_FFT_LOG_SIZE

// One lane per vector index i: swaps vector i with its transpose j when i < j.
__global__ void __launch_bounds__(256) transpose_vecs_kernel(uint32_t *col, uint32_t log_n_vecs) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >> log_n_vecs) return;
  const uint32_t half = log_n_vecs / 2;
  const uint32_t hi_shift = log_n_vecs - half;
  const uint32_t a = i >> hi_shift;
  const uint32_t b = (i >> half) & ((1u << (log_n_vecs & 1)) - 1);
  const uint32_t c = i & ((1u << half) - 1);
  const uint32_t j = (c << hi_shift) | (b << half) | a;
  if (i >= j) return;
  uint4 *vi = reinterpret_cast<uint4 *>(col + ((size_t)i << kSimdLogLanes));
  uint4 *vj = reinterpret_cast<uint4 *>(col + ((size_t)j << kSimdLogLanes));
  uint4 ti[4], tj[4];
#pragma unroll
  for (int k = 0; k < 4; k++) {
    ti[k] = vi[k];
    tj[k] = vj[k];
  }
#pragma unroll
  for (int k = 0; k < 4; k++) {
    vi[k] = tj[k];
    vj[k] = ti[k];
  }
}

__global__ void __launch_bounds__(256) double_words_kernel(const uint32_t *in, uint32_t *out, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = in[i] << 1;
}

__global__ void __launch_bounds__(256) canonicalize_kernel(uint32_t *col, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && col[i] == P) col[i] = 0;
}

}  // namespace stwo_hip

using namespace stwo_hip;

extern "C" {

int stwo_hip_coeffs_simd_transpose(stwo_hip_ctx *ctx, uint32_t *col, uint32_t log_size) {
  if (!ctx) return STWO_HIP_EINVAL;
  STWO_HIP_CHECK_ARG(ctx, col != nullptr && ((uintptr_t)col & 15) == 0,
                     "coeffs_simd_transpose: column NULL or not 16-byte aligned");
  STWO_HIP_CHECK_ARG(ctx, log_size <= 31, "coeffs_simd_transpose: log size %u too large", log_size);
  if (log_size <= kSimdCachedLog) return STWO_HIP_OK;  // SimdBackend keeps natural order there
  const uint32_t log_n_vecs = log_size - kSimdLogLanes;
  const uint64_t n = 1ull << log_n_vecs;
  return launch(ctx, "coeffs_simd_transpose", transpose_vecs_kernel, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0,
                col, log_n_vecs);
}

int stwo_hip_twiddles_simd_doubled(stwo_hip_ctx *ctx, const stwo_hip_twiddles *tw, uint32_t *out_twiddles,
                                   uint32_t *out_itwiddles) {
  if (!ctx) return STWO_HIP_EINVAL;
  STWO_HIP_CHECK_ARG(ctx, tw && tw->device == ctx->device, "twiddles missing or from another device");
  STWO_HIP_CHECK_ARG(ctx, out_twiddles && out_itwiddles, "twiddles_simd_doubled: NULL output");
  const size_t n = tw->len;
  const uint32_t blocks = (uint32_t)((n + 255) / 256);
  int rc = launch(ctx, "twiddles_simd_doubled", double_words_kernel, dim3(blocks), dim3(256), 0,
                  (const uint32_t *)tw->d_tw, out_twiddles, n);
  if (rc) return rc;
  return launch(ctx, "twiddles_simd_doubled", double_words_kernel, dim3(blocks), dim3(256), 0,
                (const uint32_t *)tw->d_itw, out_itwiddles, n);
}

int stwo_hip_canonicalize(stwo_hip_ctx *ctx, uint32_t *col, size_t n) {
  if (!ctx) return STWO_HIP_EINVAL;
  STWO_HIP_CHECK_ARG(ctx, col || !n, "canonicalize: NULL column");
  STWO_HIP_CHECK_ARG(ctx, n <= (1ull << 36), "canonicalize: %zu words too many", n);
  return launch(ctx, "canonicalize", canonicalize_kernel, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, col, n);
}

}  // extern "C"

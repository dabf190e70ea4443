// The PolyOps / FriOps methods beside the FFT: barycentric evaluation at an
// out-of-domain point (the PCS's sampling path when coefficients are not
// stored), evaluation by folding, split_at_mid and FriOps::decompose.
//
// barycentric_weights (CpuBackend, CPUB/circle.rs:101-132): for the
// bit-reversed domain point q_i of CanonicCoset(n),
//   w_i = vn(p) / (s_i * v_i(p)),
//   s_i = -2 q_i.y * 4^(n-1) * prod_{k=1}^{n-1} x_k,  x_1 = q_i.x, x_{k+1} = 2 x_k^2 - 1
//         (coset_vanishing_derivative over CanonicCoset(k).coset, whose shift is zero),
//   v_i(p) = h.y / (1 + h.x),  h = p - q_i  (point_vanishing, core/constraints.rs:87-93),
//   vn(p) = x of p doubled n - 1 times (coset_vanishing of CanonicCoset(n).coset).
// One lane per domain point: the point from its index (the half coset's
// initial point plus the step multiples of the index's bits), the M31
// product, and one QM31 inverse (w_i = vn (1 + h.x) / (s_i h.y)).  Field
// values are unique, so the result equals any other correct evaluation of
// the formula (the reference's batch inversion included).
//
// barycentric_eval_at_point (CPUB/circle.rs:134-141): sum_i evals[i] * w_i,
// batched over columns of one size: a workgroup takes a 2^11-row block of
// up to 32 columns, each lane 8 consecutive rows (16-byte loads when the
// columns allow); it loads the block's weights once and runs the columns 8 at
// a time, with 64-bit accumulation per limb and an LDS reduction per column;
// a second kernel sums each column's block partials.
#include <string.h>

#include <algorithm>
#include <vector>

#include "common.hpp"
#include "field.hpp"
#include "internal.hpp"

namespace stwo_hip {

constexpr uint32_t kBaryBlockLog = 11;
constexpr uint32_t kBaryCols = 4;     // columns per inner step
constexpr uint32_t kBaryWgCols = 32;  // columns per workgroup (the weights are loaded once for them)

struct BaryPointArgs {
  uint32_t n;              // log size of the domain
  cpoint init;             // half coset's initial point
  cpoint step[30];         // step * 2^j, j < n - 1
  uint32_t four_pow;       // 4^(n-1)
  uint32_t px[4], py[4];   // the sample point
  uint32_t vn[4];          // vn(p)
};

// d_pt: the sample point x, y and vn(p) on the device (12 words), in place
// of a.px, a.py and a.vn, when not NULL.
__global__ void __launch_bounds__(256) barycentric_weights_kernel(BaryPointArgs a, const uint32_t *d_pt,
                                                                  uint32_t *w0, uint32_t *w1, uint32_t *w2,
                                                                  uint32_t *w3) {
  const uint32_t n = a.n;
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (1u << n)) return;
  const uint32_t k = bitrev32(i, n);
  const uint32_t half = 1u << (n - 1);
  const uint32_t kk = k & (half - 1);
  cpoint q = a.init;
  for (uint32_t j = 0; j + 1 < n; j++)
    if ((kk >> j) & 1) q = cp_add(q, a.step[j]);
  if (k >= half) q.y = m31_neg(q.y);  // the second half: the conjugate points
  uint32_t prod = 1, x = q.x;
  for (uint32_t j = 1; j < n; j++) {
    prod = m31_mul(prod, x);
    const uint32_t sq = m31_mul(x, x);
    x = m31_sub(m31_add(sq, sq), 1);
  }
  const uint32_t s = m31_mul(m31_mul(m31_neg(m31_add(q.y, q.y)), a.four_pow), prod);
  uint32_t pp[12];  // x, y, vn
  if (d_pt) {
#pragma unroll
    for (int k = 0; k < 12; k++) pp[k] = as_global(d_pt)[k];
  } else {
#pragma unroll
    for (int k = 0; k < 4; k++) {
      pp[k] = a.px[k];
      pp[4 + k] = a.py[k];
      pp[8 + k] = a.vn[k];
    }
  }
  const qm31 px = qm31_make(pp[0], pp[1], pp[2], pp[3]);
  const qm31 py = qm31_make(pp[4], pp[5], pp[6], pp[7]);
  const qm31 vn = qm31_make(pp[8], pp[9], pp[10], pp[11]);
  // h = p - q = p + (q.x, -q.y)
  const qm31 hx = qm31_add(qm31_mul_m31(px, q.x), qm31_mul_m31(py, q.y));
  const qm31 hy = qm31_sub(qm31_mul_m31(py, q.x), qm31_mul_m31(px, q.y));
  const qm31 num = qm31_mul(vn, qm31_add(qm31_make(1, 0, 0, 0), hx));
  const qm31 w = qm31_mul(num, qm31_inv(qm31_mul_m31(hy, s)));
  as_global(w0)[i] = w.v[0];
  as_global(w1)[i] = w.v[1];
  as_global(w2)[i] = w.v[2];
  as_global(w3)[i] = w.v[3];
}

struct BaryWork {
  uint32_t col0, n, block;
};

// Block `w.block` (2^B rows, B = min(log, 11)) of columns col0 .. col0 + n - 1
// (n <= kBaryWgCols).  VEC: every column and weight array is 16-byte aligned
// and B = 11, so a lane's 8 rows are two 16-byte loads, all in the block.
template <bool VEC>
__global__ void __launch_bounds__(256) barycentric_dot_kernel(const uint32_t *const *cols, const BaryWork *work,
                                                              uint32_t log_size, const uint32_t *w0,
                                                              const uint32_t *w1, const uint32_t *w2,
                                                              const uint32_t *w3, uint32_t n_blocks,
                                                              uint32_t *partials) {
  __shared__ uint32_t red[kBaryCols * 4][257];
  __shared__ uint32_t sums[kBaryCols * 4];
  const BaryWork wk = work[blockIdx.x];
  const uint32_t B = log_size < kBaryBlockLog ? log_size : kBaryBlockLog;
  const uint32_t nb = 1u << B;
  const uint32_t r0 = (wk.block << B) + threadIdx.x * 8;
  const uint32_t valid = threadIdx.x * 8 < nb ? min(8u, nb - threadIdx.x * 8) : 0u;
  auto load8 = [&](const uint32_t *src, uint32_t (&v)[8]) {
    if (VEC) {
      // whole blocks only: every lane's 8 rows are in the block
      const uint4 a = ld_global4(src + r0), b = ld_global4(src + r0 + 4);
      v[0] = a.x, v[1] = a.y, v[2] = a.z, v[3] = a.w, v[4] = b.x, v[5] = b.y, v[6] = b.z, v[7] = b.w;
    } else {
#pragma unroll
      for (int e = 0; e < 8; e++) v[e] = (uint32_t)e < valid ? as_global(src)[r0 + e] : 0u;
    }
  };
  uint32_t w[4][8];
  load8(w0, w[0]);
  load8(w1, w[1]);
  load8(w2, w[2]);
  load8(w3, w[3]);
  for (uint32_t g0 = 0; g0 < wk.n; g0 += kBaryCols) {
    const uint32_t gn = min(kBaryCols, wk.n - g0);
    uint32_t cv[kBaryCols][8];
#pragma unroll
    for (uint32_t j = 0; j < kBaryCols; j++) {
      if (j < gn) {
        load8(cols[wk.col0 + g0 + j], cv[j]);
      } else {
#pragma unroll
        for (int e = 0; e < 8; e++) cv[j][e] = 0;
      }
    }
#pragma unroll
    for (uint32_t j = 0; j < kBaryCols; j++) {
      if (j >= gn) break;
      // The weights and the column's words as fresh 32-bit values for each
      // column: otherwise the compiler keeps their 64-bit zero extensions
      // live across the columns, two registers per word.
#pragma unroll
      for (int k = 0; k < 4; k++)
#pragma unroll
        for (int e = 0; e < 8; e++) asm volatile("" : "+v"(w[k][e]));
#pragma unroll
      for (int e = 0; e < 8; e++) asm volatile("" : "+v"(cv[j][e]));
      uint64_t acc[4] = {0, 0, 0, 0};
#pragma unroll
      for (int e = 0; e < 8; e++) {
#pragma unroll
        for (int k = 0; k < 4; k++) acc[k] += (uint64_t)cv[j][e] * w[k][e];
        if (e == 2 || e == 5)
#pragma unroll
          for (int k = 0; k < 4; k++) acc[k] = m31_fold64(acc[k]);
      }
#pragma unroll
      for (int k = 0; k < 4; k++) red[4 * j + k][threadIdx.x] = m31_reduce64(acc[k]);
      // one column's products at a time: interleaving the columns would hold
      // every column's accumulators at once and cut the waves per SIMD
      __builtin_amdgcn_sched_barrier(0);
    }
    __syncthreads();
    const uint32_t r = threadIdx.x >> 3, seg = threadIdx.x & 7;
    if (r < 4 * gn) {
      uint64_t sacc = 0;
#pragma unroll
      for (int i = 0; i < 32; i++) sacc += red[r][seg * 32 + i];
      uint32_t v = m31_reduce64(sacc);
#pragma unroll
      for (int o = 1; o < 8; o <<= 1) v = m31_add(v, __shfl_xor(v, o, 8));
      if (seg == 0) sums[r] = v;
    }
    __syncthreads();
    if (threadIdx.x < 4 * gn) {
      const uint32_t j = threadIdx.x >> 2, k = threadIdx.x & 3;
      partials[4 * ((size_t)(wk.col0 + g0 + j) * n_blocks + wk.block) + k] = sums[threadIdx.x];
    }
    // the next group rewrites red and sums only after every lane passed here
    __syncthreads();
  }
}

// One workgroup per column: the sum of its n_blocks partials.
__global__ void __launch_bounds__(256) barycentric_sum_kernel(const uint32_t *partials, uint32_t n_blocks,
                                                              uint32_t *out) {
  __shared__ uint32_t red[4][256];
  const uint32_t c = blockIdx.x;
  uint64_t acc[4] = {0, 0, 0, 0};
  for (uint32_t b = threadIdx.x; b < n_blocks; b += blockDim.x)
#pragma unroll
    for (int k = 0; k < 4; k++) acc[k] += partials[4 * ((size_t)c * n_blocks + b) + k];
#pragma unroll
  for (int k = 0; k < 4; k++) red[k][threadIdx.x] = m31_reduce64(acc[k]);
  __syncthreads();
  for (uint32_t s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s)
#pragma unroll
      for (int k = 0; k < 4; k++) red[k][threadIdx.x] = m31_add(red[k][threadIdx.x], red[k][threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x < 4) out[4 * (size_t)c + threadIdx.x] = red[threadIdx.x][0];
}

// ---- FriOps::decompose (CPUB/fri.rs:41-63,141-156) ---------------------------
// lambda = (sum of the first half - sum of the second half) / N over each
// coordinate of the bit-reversed secure evaluation; g = f - lambda on the
// first half, f + lambda on the second.  Partial sums per workgroup (64-bit,
// folded), one workgroup combines them into lambda (device memory), then the
// elementwise pass reads it.
constexpr uint32_t kDecompWgs = 256;

__global__ void __launch_bounds__(256) decompose_partial_kernel(const uint32_t *c0, const uint32_t *c1,
                                                                const uint32_t *c2, const uint32_t *c3,
                                                                uint32_t log_size, uint32_t *partials) {
  __shared__ uint32_t red[8][256];
  const uint32_t N = 1u << log_size, half = N >> 1;
  const uint32_t *cs[4] = {c0, c1, c2, c3};
  uint64_t a[4] = {0, 0, 0, 0}, b[4] = {0, 0, 0, 0};
  uint32_t cnt = 0;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < half; i += gridDim.x * blockDim.x) {
#pragma unroll
    for (int k = 0; k < 4; k++) {
      a[k] += cs[k][i];
      b[k] += cs[k][i + half];
    }
    if (++cnt == 1u << 30) {  // keep the 64-bit sums far from overflow
#pragma unroll
      for (int k = 0; k < 4; k++) {
        a[k] = m31_fold64(a[k]);
        b[k] = m31_fold64(b[k]);
      }
      cnt = 0;
    }
  }
#pragma unroll
  for (int k = 0; k < 4; k++) {
    red[k][threadIdx.x] = m31_reduce64(a[k]);
    red[4 + k][threadIdx.x] = m31_reduce64(b[k]);
  }
  __syncthreads();
  for (uint32_t s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s)
#pragma unroll
      for (int k = 0; k < 8; k++) red[k][threadIdx.x] = m31_add(red[k][threadIdx.x], red[k][threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x < 8) partials[8 * blockIdx.x + threadIdx.x] = red[threadIdx.x][0];
}

__global__ void __launch_bounds__(64) decompose_lambda_kernel(const uint32_t *partials, uint32_t n_parts,
                                                              uint32_t n_inv, uint32_t *lambda) {
  if (threadIdx.x >= 4) return;
  uint64_t a = 0, b = 0;
  for (uint32_t p = 0; p < n_parts; p++) {
    a += partials[8 * p + threadIdx.x];
    b += partials[8 * p + 4 + threadIdx.x];
  }
  lambda[threadIdx.x] = m31_mul(m31_sub(m31_reduce64(a), m31_reduce64(b)), n_inv);
}

__global__ void __launch_bounds__(256) decompose_apply_kernel(const uint32_t *c0, const uint32_t *c1,
                                                              const uint32_t *c2, const uint32_t *c3,
                                                              uint32_t *g0, uint32_t *g1, uint32_t *g2, uint32_t *g3,
                                                              uint32_t log_size, const uint32_t *lambda) {
  const uint32_t N = 1u << log_size, half = N >> 1;
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const uint32_t *cs[4] = {c0, c1, c2, c3};
  uint32_t *gs[4] = {g0, g1, g2, g3};
#pragma unroll
  for (int k = 0; k < 4; k++) {
    const uint32_t v = cs[k][i], l = lambda[k];
    gs[k][i] = i < half ? m31_sub(v, l) : m31_add(v, l);
  }
}

// The point from the host (p) or the device (d_pt, 12 words: x, y, vn).
static int barycentric_weights_impl(stwo_hip_ctx *ctx, uint32_t log_size, const stwo_hip_secure_point *p,
                                    const uint32_t *d_pt, uint32_t *const out[4]) {
  if (!ctx) return STWO_HIP_EINVAL;
  STWO_HIP_CHECK_ARG(ctx, (p || d_pt) && out && out[0] && out[1] && out[2] && out[3],
                     "barycentric_weights: NULL argument");
  STWO_HIP_CHECK_ARG(ctx, log_size >= 1 && log_size <= 30, "barycentric_weights: log size %u outside [1, 30]",
                     log_size);
  BaryPointArgs a{};
  a.n = log_size;
  const uint32_t init = half_odds_initial(log_size - 1);
  const uint32_t step = coset_step(log_size - 1);
  a.init = host_index_to_point(init);
  for (uint32_t j = 0; j + 1 < log_size; j++) a.step[j] = host_index_to_point((uint32_t)(((uint64_t)step << j) & IDX_MASK));
  a.four_pow = m31_pow(4, log_size - 1);
  if (p) {
    memcpy(a.px, p->x, 16);
    memcpy(a.py, p->y, 16);
    qm31 x = qm31_make(p->x[0], p->x[1], p->x[2], p->x[3]);
    for (uint32_t j = 1; j < log_size; j++) {
      const qm31 sq = qm31_mul(x, x);
      x = qm31_sub(qm31_add(sq, sq), qm31_make(1, 0, 0, 0));
    }
    memcpy(a.vn, x.v, 16);
  }
  const uint64_t n = 1ull << log_size;
  return launch(ctx, "barycentric_weights", barycentric_weights_kernel, dim3((uint32_t)((n + 255) / 256)), dim3(256),
                0, a, d_pt, out[0], out[1], out[2], out[3]);
}

int barycentric_weights_dev(stwo_hip_ctx *ctx, uint32_t log_size, const uint32_t *d_point, uint32_t *const out[4]) {
  if (!ctx) return STWO_HIP_EINVAL;
  STWO_HIP_CHECK_ARG(ctx, d_point, "barycentric_weights: NULL device point");
  return barycentric_weights_impl(ctx, log_size, nullptr, d_point, out);
}

// The values to the host (values) or left on the device (d_values).
static int barycentric_eval_impl(stwo_hip_ctx *ctx, const uint32_t *const *evals, size_t n_cols, uint32_t log_size,
                                 const uint32_t *const weights[4], uint32_t *values, uint32_t *d_values) {
  if (!ctx) return STWO_HIP_EINVAL;
  STWO_HIP_CHECK_ARG(ctx, (evals || !n_cols) && weights && (values || d_values),
                     "barycentric_eval_at_point: NULL argument");
  if (n_cols == 0) return STWO_HIP_OK;
  STWO_HIP_CHECK_ARG(ctx, log_size >= 1 && log_size <= 30 && n_cols <= 65536,
                     "barycentric_eval_at_point: log size %u or %zu columns out of range", log_size, n_cols);
  for (int k = 0; k < 4; k++) STWO_HIP_CHECK_ARG(ctx, weights[k], "barycentric_eval_at_point: NULL weights");
  for (size_t c = 0; c < n_cols; c++) STWO_HIP_CHECK_ARG(ctx, evals[c], "barycentric_eval_at_point: NULL column");
  const uint32_t n_blocks = 1u << (log_size > kBaryBlockLog ? log_size - kBaryBlockLog : 0);
  std::vector<BaryWork> work;
  for (uint32_t c = 0; c < n_cols; c += kBaryWgCols)
    for (uint32_t b = 0; b < n_blocks; b++)
      work.push_back({c, (uint32_t)std::min<size_t>(kBaryWgCols, n_cols - c), b});
  bool vec = log_size >= kBaryBlockLog;  // whole 2^11-row blocks
  for (int k = 0; k < 4; k++) vec = vec && ((uintptr_t)weights[k] & 15) == 0;
  for (size_t c = 0; c < n_cols; c++) vec = vec && ((uintptr_t)evals[c] & 15) == 0;
  void *d[2];
  const void *srcs[2] = {evals, work.data()};
  const size_t bytes[2] = {n_cols * sizeof(uint32_t *), work.size() * sizeof(BaryWork)};
  // with the values left on the device (the prove) the tables repeat from call to call
  int rc = d_values ? upload_tables_cached(ctx, "barycentric_eval", 2, srcs, bytes, d)
                    : upload_tables(ctx, 2, srcs, bytes, d);
  if (rc) return rc;
  void *part = nullptr, *out = d_values;
  rc = get_scratch(ctx, "bary_partials", n_cols * (size_t)n_blocks * 16, &part);
  if (rc) return rc;
  if (!out) {
    rc = get_scratch(ctx, "bary_out", n_cols * 16, &out);
    if (rc) return rc;
  }
  rc = launch(ctx, "barycentric_dot", vec ? barycentric_dot_kernel<true> : barycentric_dot_kernel<false>,
              dim3((uint32_t)work.size()), dim3(256), 0,
              (const uint32_t *const *)d[0], (const BaryWork *)d[1], log_size, weights[0], weights[1], weights[2],
              weights[3], n_blocks, (uint32_t *)part);
  if (rc) return rc;
  rc = launch(ctx, "barycentric_sum", barycentric_sum_kernel, dim3((uint32_t)n_cols), dim3(256), 0,
              (const uint32_t *)part, n_blocks, (uint32_t *)out);
  if (rc) return rc;
  if (d_values) return STWO_HIP_OK;
  return stwo_hip_download(ctx, values, (const uint32_t *)out, 4 * n_cols);
}

int barycentric_eval_dev(stwo_hip_ctx *ctx, const uint32_t *const *evals, size_t n_cols, uint32_t log_size,
                         const uint32_t *const weights[4], uint32_t *d_values) {
  if (!ctx) return STWO_HIP_EINVAL;
  STWO_HIP_CHECK_ARG(ctx, d_values, "barycentric_eval_at_point: NULL device values");
  return barycentric_eval_impl(ctx, evals, n_cols, log_size, weights, nullptr, d_values);
}

}  // namespace stwo_hip

using namespace stwo_hip;

extern "C" {

int stwo_hip_barycentric_weights(stwo_hip_ctx *ctx, uint32_t log_size, const stwo_hip_secure_point *p,
                                 uint32_t *const out[4]) {
  if (!ctx) return STWO_HIP_EINVAL;
  STWO_HIP_CHECK_ARG(ctx, p, "barycentric_weights: NULL argument");
  return barycentric_weights_impl(ctx, log_size, p, nullptr, out);
}

int stwo_hip_barycentric_eval_at_point(stwo_hip_ctx *ctx, const uint32_t *const *evals, size_t n_cols,
                                       uint32_t log_size, const uint32_t *const weights[4], uint32_t *values) {
  if (!ctx) return STWO_HIP_EINVAL;
  STWO_HIP_CHECK_ARG(ctx, values, "barycentric_eval_at_point: NULL argument");
  return barycentric_eval_impl(ctx, evals, n_cols, log_size, weights, values, nullptr);
}

int stwo_hip_eval_at_point_by_folding(stwo_hip_ctx *ctx, const uint32_t *evals, uint32_t log_size,
                                      const stwo_hip_secure_point *p, const stwo_hip_twiddles *tw,
                                      uint32_t value[4]) {
  if (!ctx) return STWO_HIP_EINVAL;
  STWO_HIP_CHECK_ARG(ctx, evals && p && value, "eval_at_point_by_folding: NULL argument");
  STWO_HIP_CHECK_ARG(ctx, log_size >= 3 && log_size <= 28, "eval_at_point_by_folding: log size %u outside [3, 28]",
                     log_size);
  // get_folding_alphas (core/poly/utils.rs:45-70): [pi^(n-2)(x), ..., x, y],
  // used from the back: y for the circle fold, then x, pi(x), ...
  std::vector<qm31> alphas(log_size);
  alphas[log_size - 1] = qm31_make(p->y[0], p->y[1], p->y[2], p->y[3]);
  qm31 x = qm31_make(p->x[0], p->x[1], p->x[2], p->x[3]);
  for (int i = (int)log_size - 2; i >= 0; i--) {
    alphas[i] = x;
    const qm31 sq = qm31_mul(x, x);
    x = qm31_sub(qm31_add(sq, sq), qm31_make(1, 0, 0, 0));
  }
  const size_t half = (size_t)1 << (log_size - 1);
  // scratch: three zero coordinates of the source (the evaluation as secure
  // values), the line evaluation and its fold (4 coordinates each)
  void *mem = nullptr;
  int rc = get_scratch(ctx, "fold_eval", (3 * 2 * half + 8 * half) * 4, &mem);
  if (rc) return rc;
  uint32_t *zero = (uint32_t *)mem, *la = zero + 3 * 2 * half, *lb = la + 4 * half;
  STWO_HIP_TRY(ctx, hipMemsetAsync(zero, 0, 3 * 2 * half * 4, ctx->stream));
  STWO_HIP_TRY(ctx, hipMemsetAsync(la, 0, 4 * half * 4, ctx->stream));
  const uint32_t *src[4] = {evals, zero, zero + 2 * half, zero + 4 * half};
  uint32_t *dst[4] = {la, la + half, la + 2 * half, la + 3 * half};
  uint32_t alpha[4];
  memcpy(alpha, alphas.back().v, 16);
  alphas.pop_back();
  rc = stwo_hip_fold_circle_into_line(ctx, dst, src, {half_odds_initial(log_size - 1), log_size - 1}, alpha, tw);
  if (rc) return rc;
  uint32_t lg = log_size - 1;
  uint32_t line_init = half_odds_initial(log_size - 1);
  uint32_t *cur = la, *nxt = lb;
  while (lg > 0) {
    const size_t n = (size_t)1 << lg;
    const uint32_t *s4[4] = {cur, cur + n, cur + 2 * n, cur + 3 * n};
    uint32_t *d4[4] = {nxt, nxt + n / 2, nxt + n, nxt + 3 * n / 2};
    memcpy(alpha, alphas.back().v, 16);
    alphas.pop_back();
    rc = stwo_hip_fold_line(ctx, s4, {line_init, lg}, alpha, 1, tw, d4);
    if (rc) return rc;
    line_init = (uint32_t)(((uint64_t)line_init << 1) & IDX_MASK);
    lg--;
    std::swap(cur, nxt);
  }
  uint32_t v[4];
  const DownloadReq reqs[4] = {{v, cur, 1}, {v + 1, cur + 1, 1}, {v + 2, cur + 2, 1}, {v + 3, cur + 3, 1}};
  rc = download_many(ctx, reqs, 4);
  if (rc) return rc;
  const uint32_t inv = m31_inv(m31_pow(2, log_size));
  for (int k = 0; k < 4; k++) value[k] = m31_mul(v[k], inv);
  return STWO_HIP_OK;
}

int stwo_hip_split_at_mid(stwo_hip_ctx *ctx, const uint32_t *coeffs, uint32_t log_size, uint32_t *left,
                          uint32_t *right) {
  if (!ctx) return STWO_HIP_EINVAL;
  STWO_HIP_CHECK_ARG(ctx, coeffs && left && right && log_size >= 1 && log_size <= 30,
                     "split_at_mid: bad arguments");
  const size_t half = (size_t)1 << (log_size - 1);
  if (left != coeffs) STWO_HIP_TRY(ctx, hipMemcpyAsync(left, coeffs, half * 4, hipMemcpyDeviceToDevice, ctx->stream));
  STWO_HIP_TRY(ctx, hipMemcpyAsync(right, coeffs + half, half * 4, hipMemcpyDeviceToDevice, ctx->stream));
  return STWO_HIP_OK;
}

int stwo_hip_fri_decompose(stwo_hip_ctx *ctx, const uint32_t *const eval[4], uint32_t log_size, uint32_t *const g[4],
                           uint32_t lambda[4]) {
  if (!ctx) return STWO_HIP_EINVAL;
  STWO_HIP_CHECK_ARG(ctx, eval && g && lambda && log_size >= 1 && log_size <= 30, "fri_decompose: bad arguments");
  for (int k = 0; k < 4; k++) STWO_HIP_CHECK_ARG(ctx, eval[k] && g[k], "fri_decompose: NULL coordinate");
  const uint32_t half = 1u << (log_size - 1);
  const uint32_t wgs = std::max<uint32_t>(1, std::min<uint32_t>(kDecompWgs, (half + 255) / 256));
  void *mem = nullptr;
  int rc = get_scratch(ctx, "decompose", (8 * (size_t)kDecompWgs + 4) * 4, &mem);
  if (rc) return rc;
  uint32_t *part = (uint32_t *)mem, *lam = part + 8 * kDecompWgs;
  rc = launch(ctx, "decompose_partial", decompose_partial_kernel, dim3(wgs), dim3(256), 0, eval[0], eval[1], eval[2],
              eval[3], log_size, part);
  if (rc) return rc;
  const uint32_t n_inv = m31_inv((uint32_t)(((uint64_t)1 << log_size) % P));
  rc = launch(ctx, "decompose_lambda", decompose_lambda_kernel, dim3(1), dim3(64), 0, (const uint32_t *)part, wgs,
              n_inv, lam);
  if (rc) return rc;
  const uint64_t N = 1ull << log_size;
  rc = launch(ctx, "decompose_apply", decompose_apply_kernel, dim3((uint32_t)((N + 255) / 256)), dim3(256), 0,
              eval[0], eval[1], eval[2], eval[3], g[0], g[1], g[2], g[3], log_size, (const uint32_t *)lam);
  if (rc) return rc;
  return stwo_hip_download(ctx, lambda, lam, 4);
}

}  // extern "C"

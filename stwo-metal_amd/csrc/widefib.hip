// The wide-Fibonacci AIR of the reference's prove benchmark
// (vendor/stwo-upstream-dev-62b228e/crates/examples/src/wide_fibonacci/mod.rs):
// trace generation and the constraint-quotient evaluation that the
// composition polynomial is built from.
//
// Trace (generate_trace, mod.rs:27-50): row i holds one sequence of N
// values, a, b, then x_k = x_{k-2}^2 + x_{k-1}^2; column k is stored with
// row i at index i.
//
// Constraint quotients (FrameworkComponent::evaluate_constraint_quotients_on_domain,
// constraint-framework/src/prover/component_prover.rs:170-274): on
// CanonicCoset(log_n + 1) (the trace committed at blowup 1), row r gets
//   accum[r] + (sum_k alpha^(N-3-k) (x_{k+2} - x_k^2 - x_{k+1}^2)) * denom_inv[r >> log_n]
// where the powers come from DomainEvaluationAccumulator (generated
// alpha^0.., reversed for the component, STWO/prover/air/accumulation.rs:30-68,
// CpuDomainEvaluator::add_constraint) and denom_inv holds the inverses of
// the trace coset's vanishing polynomial at the first 2^(eval - trace)
// points of the evaluation domain, bit-reversed (coset_vanishing,
// STWO/core/constraints.rs:12-35).
#include <string.h>

#include <vector>

#include "common.hpp"
#include "field.hpp"
#include "internal.hpp"

namespace stwo_hip {

__global__ void __launch_bounds__(256) wide_fib_trace_kernel(uint32_t *const *cols, uint32_t n_cols, uint32_t log_n,
                                                             const uint32_t *a_in, const uint32_t *b_in) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >> log_n) return;
  uint32_t a = a_in[i], b = b_in[i];
  as_global(cols[0])[i] = a;
  as_global(cols[1])[i] = b;
  for (uint32_t k = 2; k < n_cols; k++) {
    const uint32_t c = m31_add(m31_mul(a, a), m31_mul(b, b));
    as_global(cols[k])[i] = c;
    a = b;
    b = c;
  }
}

// The row's combination regrouped by column: with weights w_j = alpha^(N-3-j),
//   sum_j w_j (x_{j+2} - x_j^2 - x_{j+1}^2) = sum_i (x_i W1_i + x_i^2 W2_i),
// W1_i = w_{i-2} (i >= 2) and W2_i = -(w_i [i <= N-3] + w_{i-1} [1 <= i <= N-2]),
// so each column costs one square and 8 multiply-adds into 64-bit
// accumulators: sum x W1 and sum x^2 W2 apart, one product per column each,
// folded after every third column (three products below 2^62 fit above a
// folded value below 2^34).  Columns go 8 at a time, their loads issued
// together.
__device__ __forceinline__ void wf_column(uint64_t (&a1)[4], uint64_t (&a2)[4], uint32_t x, const uint32_t *w) {
  const uint32_t sq = m31_mul(x, x);
#pragma unroll
  for (int k = 0; k < 4; k++) {
    a1[k] += (uint64_t)x * w[k];  // W1 (words 0..3)
    a2[k] += (uint64_t)sq * w[4 + k];  // W2 (words 4..7)
  }
}
__device__ __forceinline__ void wf_fold(uint64_t (&a1)[4], uint64_t (&a2)[4]) {
#pragma unroll
  for (int k = 0; k < 4; k++) {
    a1[k] = m31_fold64(a1[k]);
    a2[k] = m31_fold64(a2[k]);
  }
}
// R consecutive rows per lane (R = 2: one 8-byte load per column, which
// doubles the bytes each load instruction carries).
template <int R>
__global__ void __launch_bounds__(256) wide_fib_quotients_kernel(const uint32_t *const *cols, uint32_t n_cols,
                                                                 uint32_t eval_log, uint32_t log_n,
                                                                 const uint32_t *weights, const uint32_t *denom_inv,
                                                                 uint32_t *acc0, uint32_t *acc1, uint32_t *acc2,
                                                                 uint32_t *acc3, uint32_t accumulate) {
  const uint32_t r0 = (blockIdx.x * blockDim.x + threadIdx.x) * R;
  if (r0 >> eval_log) return;
  uint64_t a1[R][4], a2[R][4];
#pragma unroll
  for (int rr = 0; rr < R; rr++)
#pragma unroll
    for (int k = 0; k < 4; k++) a1[rr][k] = a2[rr][k] = 0;
  auto load = [&](const uint32_t *c, uint32_t (&x)[R]) {
    if constexpr (R == 2) {
      const uint2 v = ld_global2(c + r0);
      x[0] = v.x;
      x[1] = v.y;
    } else {
      x[0] = as_global(c)[r0];
    }
  };
  uint32_t i = 0;
  for (; i + 8 <= n_cols; i += 8) {
    uint32_t x[8][R];
#pragma unroll
    for (int u = 0; u < 8; u++) load(cols[i + u], x[u]);
#pragma unroll
    for (int u = 0; u < 8; u++) {
#pragma unroll
      for (int rr = 0; rr < R; rr++) {
        wf_column(a1[rr], a2[rr], x[u][rr], weights + 8 * (i + u));
        if (u == 2 || u == 5 || u == 7) wf_fold(a1[rr], a2[rr]);
      }
    }
  }
  for (; i < n_cols; i++) {
    uint32_t x[R];
    load(cols[i], x);
#pragma unroll
    for (int rr = 0; rr < R; rr++) {
      wf_column(a1[rr], a2[rr], x[rr], weights + 8 * i);
      wf_fold(a1[rr], a2[rr]);
    }
  }
#pragma unroll
  for (int rr = 0; rr < R; rr++) {
    const uint32_t r = r0 + rr;
    qm31 res;
#pragma unroll
    for (int k = 0; k < 4; k++) res.v[k] = m31_add(m31_reduce64(a1[rr][k]), m31_reduce64(a2[rr][k]));
    res = qm31_mul_m31(res, denom_inv[r >> log_n]);
    g_u32 *o[4] = {as_global(acc0), as_global(acc1), as_global(acc2), as_global(acc3)};
#pragma unroll
    for (int k = 0; k < 4; k++) o[k][r] = accumulate ? m31_add(o[k][r], res.v[k]) : res.v[k];
  }
}

// The per-column weights from alpha on the device (one lane per column):
// with w_j = alpha^(n_cons - 1 - j), W1_i = w_{i-2} (i >= 2) and
// W2_i = -(w_i [i < n_cons] + w_{i-1} [1 <= i <= n_cons]); 8 words per
// column (W1 then W2).
__device__ __forceinline__ qm31 qm31_pow(qm31 a, uint32_t e) {
  qm31 r = qm31_make(1, 0, 0, 0);
  while (e) {
    if (e & 1) r = qm31_mul(r, a);
    a = qm31_mul(a, a);
    e >>= 1;
  }
  return r;
}
__global__ void __launch_bounds__(256) wide_fib_weights_kernel(const uint32_t *alpha_dev, uint32_t n_cols,
                                                               uint32_t *weights) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_cols) return;
  const qm31 alpha = qm31_make(alpha_dev[0], alpha_dev[1], alpha_dev[2], alpha_dev[3]);
  const uint32_t n_cons = n_cols - 2;
  const qm31 zero = qm31_make(0, 0, 0, 0);
  // w_j for j = i - 2, i - 1, i: alpha^(n_cons + 1 - i), ^(n_cons - i), ^(n_cons - 1 - i)
  const qm31 w1 = i >= 2 ? qm31_pow(alpha, n_cons + 1 - i) : zero;
  qm31 w2 = zero;
  if (i < n_cons) w2 = qm31_add(w2, qm31_pow(alpha, n_cons - 1 - i));
  if (i >= 1 && i - 1 < n_cons) w2 = qm31_add(w2, qm31_pow(alpha, n_cons - i));
  w2 = qm31_sub(zero, w2);
#pragma unroll
  for (int k = 0; k < 4; k++) {
    weights[8 * (size_t)i + k] = w1.v[k];
    weights[8 * (size_t)i + 4 + k] = w2.v[k];
  }
}

int wide_fib_constraint_quotients_impl(stwo_hip_ctx *ctx, const uint32_t *const *cols, uint32_t n_cols,
                                       uint32_t log_n, const uint32_t random_coeff[4], uint32_t *const accum[4],
                                       bool accumulate, const uint32_t *alpha_dev) {
  if (!ctx) return STWO_HIP_EINVAL;
  STWO_HIP_CHECK_ARG(ctx, cols && (random_coeff || alpha_dev) && accum && n_cols >= 3 && n_cols <= 4096 &&
                              log_n >= 1 && log_n <= 27,
                     "wide_fib_constraint_quotients: bad arguments");
  for (uint32_t k = 0; k < n_cols; k++)
    STWO_HIP_CHECK_ARG(ctx, cols[k] != nullptr, "wide_fib_constraint_quotients: NULL column");
  for (int k = 0; k < 4; k++) STWO_HIP_CHECK_ARG(ctx, accum[k] != nullptr, "wide_fib_constraint_quotients: NULL accumulator");
  const uint32_t eval_log = log_n + 1;
  // Vanishing-polynomial inverses at the first 2^(eval_log - log_n) = 2 points of
  // CanonicCoset(eval_log).circle_domain(), bit-reversed.  The trace coset is
  // canonic, so its shift (-initial + step / 2) is zero and v(p) = pi^(log_n - 1)(p.x).
  uint32_t dinv[2];
  for (uint32_t i = 0; i < 2; i++) {
    // domain.at(i) for i < half: half coset half_odds(eval_log - 1) at i
    const uint32_t idx =
        (uint32_t)((half_odds_initial(eval_log - 1) + (uint64_t)coset_step(eval_log - 1) * i) & IDX_MASK);
    uint32_t x = host_index_to_point(idx).x;
    for (uint32_t d = 1; d < log_n; d++) x = m31_sub(m31_add(m31_mul(x, x), m31_mul(x, x)), 1);
    STWO_HIP_CHECK_ARG(ctx, x != 0, "wide_fib_constraint_quotients: vanishing polynomial is zero");
    dinv[bitrev32(i, 1)] = m31_inv(x);
  }
  void *d_cols = nullptr, *d_alpha = nullptr, *d_dinv = nullptr;
  {
    const uint32_t unused[4] = {0, 0, 0, 0};
    const void *srcs[3] = {cols, random_coeff ? (const void *)random_coeff : (const void *)unused, dinv};
    const size_t bytes[3] = {n_cols * sizeof(uint32_t *), 16, sizeof dinv};
    void *d[3];
    // with alpha on the device (the prove) the tables repeat from call to call
    int rc = alpha_dev ? upload_tables_cached(ctx, "wide_fib_constraint_quotients", 3, srcs, bytes, d)
                       : upload_tables(ctx, 3, srcs, bytes, d);
    if (rc) return rc;
    d_cols = d[0];
    d_alpha = d[1];
    d_dinv = d[2];
  }
  // alpha: the caller's (uploaded) or one already on the device
  void *d_pw = nullptr;
  int rc = get_scratch(ctx, "wide_fib_weights", 32 * (size_t)n_cols, &d_pw);
  if (rc) return rc;
  rc = launch(ctx, "wide_fib_weights", wide_fib_weights_kernel, dim3((n_cols + 255) / 256), dim3(256), 0,
              alpha_dev ? alpha_dev : (const uint32_t *)d_alpha, n_cols, (uint32_t *)d_pw);
  if (rc) return rc;
  // Two rows per lane when every column is 8-byte aligned (eval_log >= 2).
  bool pairs = true;
  for (uint32_t k = 0; k < n_cols; k++) pairs = pairs && ((uintptr_t)cols[k] & 7) == 0;
  const uint64_t n = (1ull << eval_log) >> (pairs ? 1 : 0);
  return launch(ctx, "wide_fib_constraint_quotients",
                pairs ? wide_fib_quotients_kernel<2> : wide_fib_quotients_kernel<1>,
                dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, (const uint32_t *const *)d_cols, n_cols, eval_log,
                log_n, (const uint32_t *)d_pw, (const uint32_t *)d_dinv, accum[0], accum[1], accum[2], accum[3],
                (uint32_t)accumulate);
}

}  // namespace stwo_hip

using namespace stwo_hip;

extern "C" {

int stwo_hip_wide_fib_trace(stwo_hip_ctx *ctx, uint32_t *const *cols, uint32_t n_cols, uint32_t log_n,
                            const uint32_t *a, const uint32_t *b) {
  if (!ctx) return STWO_HIP_EINVAL;
  STWO_HIP_CHECK_ARG(ctx, cols && a && b && n_cols >= 2 && n_cols <= 4096 && log_n >= 1 && log_n <= 28,
                     "wide_fib_trace: bad arguments");
  for (uint32_t k = 0; k < n_cols; k++) STWO_HIP_CHECK_ARG(ctx, cols[k] != nullptr, "wide_fib_trace: NULL column");
  void *d_tab = nullptr;
  int rc = upload_table(ctx, cols, n_cols * sizeof(uint32_t *), &d_tab);
  if (rc) return rc;
  const uint64_t n = 1ull << log_n;
  return launch(ctx, "wide_fib_trace", wide_fib_trace_kernel, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0,
                (uint32_t *const *)d_tab, n_cols, log_n, a, b);
}

int stwo_hip_wide_fib_constraint_quotients(stwo_hip_ctx *ctx, const uint32_t *const *cols, uint32_t n_cols,
                                           uint32_t log_n, const uint32_t random_coeff[4], uint32_t *const accum[4]) {
  if (!ctx) return STWO_HIP_EINVAL;
  STWO_HIP_CHECK_ARG(ctx, random_coeff != nullptr, "wide_fib_constraint_quotients: NULL random coefficient");
  return wide_fib_constraint_quotients_impl(ctx, cols, n_cols, log_n, random_coeff, accum, true, nullptr);
}

}  // extern "C"

// Evaluation of circle polynomials at a point of the secure-field circle
// (PolyOps::eval_at_point, STWO/prover/poly/circle/ops.rs; CpuBackend
// CPUB/circle.rs:82-99; Metal batch_eval_at_point, SH/eval_at_point.metal),
// the out-of-domain sampling step of the PCS.
//
// For coefficients c_0..c_{2^n - 1} (natural FFT-basis order) and the point
// (x, y), eval_at_point = fold(c, [pi^(n-2)(x), ..., pi(x), x, y])
// (STWO/core/poly/utils.rs:27-38), i.e. sum_j c_j * prod_{bit k of j} g_k with
// g_0 = y, g_1 = x, g_k = pi^(k-1)(x), pi(x) = 2x^2 - 1.  A workgroup folds a
// block of 2^B consecutive coefficients through LDS with g_0..g_{B-1} (the
// fold's bottom levels) and multiplies the result by the product of g_k over
// the set bits of the block index (the top levels); a second kernel sums the
// blocks of each column.  Sums of field elements are exact, so the grouping
// does not change the result.
#include <string.h>

#include <algorithm>
#include <vector>

#include "common.hpp"
#include "field.hpp"

namespace stwo_hip {

constexpr uint32_t kEvalBlockLog = 11;  // 2048 coefficients per workgroup (8 KiB of LDS + 16 KiB of QM31)

struct EvalCol {
  const uint32_t *coeffs;
  uint32_t log_size;
  uint32_t block_begin;  // first block (partial) index of this column
};
struct EvalFactors {
  uint32_t g[32][4];  // g_0 = y, g_1 = x, g_k = pi^(k-1)(x)
};

__device__ __forceinline__ qm31 ld_factor(const EvalFactors &f, uint32_t k) {
  return qm31_make(f.g[k][0], f.g[k][1], f.g[k][2], f.g[k][3]);
}

// grid.x = total blocks over all columns; block `b` belongs to the column
// whose [block_begin, next block_begin) holds it.
__global__ void __launch_bounds__(256) eval_point_blocks_kernel(const EvalCol *cols, uint32_t n_cols,
                                                                EvalFactors f, uint32_t *partials) {
  __shared__ uint32_t sc[1u << kEvalBlockLog];
  __shared__ qm31 lv[1u << (kEvalBlockLog - 1)];
  // the last column whose first block is <= blockIdx.x (every column has >= 1 block)
  uint32_t lo = 0, hi = n_cols - 1;
  while (lo < hi) {
    const uint32_t mid = (lo + hi + 1) >> 1;
    if (cols[mid].block_begin <= blockIdx.x) lo = mid;
    else hi = mid - 1;
  }
  const uint32_t ci = lo;
  const EvalCol col = cols[ci];
  const uint32_t blk = blockIdx.x - col.block_begin;
  const uint32_t B = col.log_size < kEvalBlockLog ? col.log_size : kEvalBlockLog;
  const uint32_t nb = 1u << B;
  for (uint32_t t = threadIdx.x; t < nb; t += blockDim.x) sc[t] = as_global(col.coeffs)[((size_t)blk << B) + t];
  __syncthreads();
  qm31 res;
  if (B == 0) {
    res = qm31_make(sc[0], 0, 0, 0);
  } else {
    // bottom level: c[2t] + c[2t + 1] * g_0
    const qm31 g0 = ld_factor(f, 0);
    for (uint32_t t = threadIdx.x; t < (nb >> 1); t += blockDim.x) {
      qm31 v = qm31_mul_m31(g0, sc[2 * t + 1]);
      v.v[0] = m31_add(v.v[0], sc[2 * t]);
      lv[t] = v;
    }
    for (uint32_t k = 1; k < B; k++) {
      __syncthreads();
      const uint32_t n = nb >> (k + 1);
      const qm31 g = ld_factor(f, k);
      qm31 tmp[4];
      uint32_t cnt = 0;
      for (uint32_t t = threadIdx.x; t < n; t += blockDim.x) tmp[cnt++] = qm31_add(lv[2 * t], qm31_mul(lv[2 * t + 1], g));
      __syncthreads();
      cnt = 0;
      for (uint32_t t = threadIdx.x; t < n; t += blockDim.x) lv[t] = tmp[cnt++];
    }
    __syncthreads();
    res = lv[0];
  }
  if (threadIdx.x == 0) {
    // top levels: the block index's set bits select g_B, g_{B+1}, ...
    for (uint32_t k = B; k < col.log_size; k++)
      if (blk >> (k - B) & 1) res = qm31_mul(res, ld_factor(f, k));
    for (int w = 0; w < 4; w++) partials[4 * (size_t)blockIdx.x + w] = res.v[w];
  }
}

// One workgroup per column: the sum of its block partials.
__global__ void __launch_bounds__(256) eval_point_sum_kernel(const EvalCol *cols, uint32_t n_cols,
                                                             uint32_t total_blocks, const uint32_t *partials,
                                                             uint32_t *out) {
  __shared__ qm31 red[256];
  const uint32_t ci = blockIdx.x;
  const uint32_t b0 = cols[ci].block_begin;
  const uint32_t b1 = ci + 1 < n_cols ? cols[ci + 1].block_begin : total_blocks;
  qm31 acc = qm31_make(0, 0, 0, 0);
  for (uint32_t b = b0 + threadIdx.x; b < b1; b += blockDim.x)
    acc = qm31_add(acc, qm31_make(partials[4 * (size_t)b], partials[4 * (size_t)b + 1], partials[4 * (size_t)b + 2],
                                  partials[4 * (size_t)b + 3]));
  red[threadIdx.x] = acc;
  __syncthreads();
  for (uint32_t s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] = qm31_add(red[threadIdx.x], red[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int w = 0; w < 4; w++) out[4 * (size_t)ci + w] = red[0].v[w];
}

}  // namespace stwo_hip

using namespace stwo_hip;

extern "C" {

int stwo_hip_eval_at_point(stwo_hip_ctx *ctx, const uint32_t *const *coeffs, const uint32_t *log_sizes, size_t n_cols,
                           const stwo_hip_secure_point *point, uint32_t *values) {
  if (!ctx) return STWO_HIP_EINVAL;
  STWO_HIP_CHECK_ARG(ctx, point && values && (coeffs || !n_cols) && (log_sizes || !n_cols),
                     "eval_at_point: NULL argument");
  if (n_cols == 0) return STWO_HIP_OK;
  STWO_HIP_CHECK_ARG(ctx, n_cols <= 65536, "eval_at_point: %zu columns in one call (at most 65536)", n_cols);
  std::vector<EvalCol> tab(n_cols);
  uint64_t total = 0;
  uint32_t max_log = 0;
  for (size_t c = 0; c < n_cols; c++) {
    STWO_HIP_CHECK_ARG(ctx, coeffs[c] && log_sizes[c] <= 30, "eval_at_point: column %zu invalid", c);
    tab[c].coeffs = coeffs[c];
    tab[c].log_size = log_sizes[c];
    tab[c].block_begin = (uint32_t)total;
    total += 1ull << (log_sizes[c] > kEvalBlockLog ? log_sizes[c] - kEvalBlockLog : 0);
    max_log = std::max(max_log, log_sizes[c]);
  }
  STWO_HIP_CHECK_ARG(ctx, total < (1ull << 31), "eval_at_point: too many blocks");
  // Factors g_0 = y, g_1 = x, g_k = pi^(k-1)(x) (CPUB/circle.rs:82-99).
  EvalFactors f{};
  qm31 x = qm31_make(point->x[0], point->x[1], point->x[2], point->x[3]);
  memcpy(f.g[0], point->y, 16);
  for (uint32_t k = 1; k < 32 && k < max_log + 1; k++) {
    memcpy(f.g[k], x.v, 16);
    qm31 sq = qm31_mul(x, x);
    x = qm31_sub(qm31_add(sq, sq), qm31_make(1, 0, 0, 0));
  }
  void *d_tab = nullptr, *d_part = nullptr, *d_out = nullptr;
  int rc = upload_table(ctx, tab.data(), tab.size() * sizeof(EvalCol), &d_tab);
  if (rc) return rc;
  rc = get_scratch(ctx, "eval_point_partials", total * 16, &d_part);
  if (rc) return rc;
  rc = get_scratch(ctx, "eval_point_out", n_cols * 16, &d_out);
  if (rc) return rc;
  rc = launch(ctx, "eval_at_point_blocks", eval_point_blocks_kernel, dim3((uint32_t)total), dim3(256), 0,
              (const EvalCol *)d_tab, (uint32_t)n_cols, f, (uint32_t *)d_part);
  if (rc) return rc;
  rc = launch(ctx, "eval_at_point_sum", eval_point_sum_kernel, dim3((uint32_t)n_cols), dim3(256), 0,
              (const EvalCol *)d_tab, (uint32_t)n_cols, (uint32_t)total, (const uint32_t *)d_part, (uint32_t *)d_out);
  if (rc) return rc;
  return stwo_hip_download(ctx, values, (const uint32_t *)d_out, 4 * n_cols);
}

}  // extern "C"

// Evaluation of circle polynomials at a point of the secure-field circle
// (PolyOps::eval_at_point, STWO/prover/poly/circle/ops.rs; CpuBackend
// CPUB/circle.rs:82-99; Metal batch_eval_at_point, SH/eval_at_point.metal),
// the out-of-domain sampling step of the PCS.
//
// For coefficients c_0..c_{2^n - 1} (natural FFT-basis order) and the point
// (x, y), eval_at_point = fold(c, [pi^(n-2)(x), ..., pi(x), x, y])
// (STWO/core/poly/utils.rs:27-38), i.e. sum_j c_j * basis_j with
// basis_j = prod of g_k over the set bits k of j, g_0 = y, g_1 = x,
// g_k = pi^(k-1)(x), pi(x) = 2x^2 - 1.  The basis does not depend on the
// column, so one table of the low 2^11 basis values serves every column: a
// workgroup takes one 2^11-coefficient block of up to 8 columns of the same
// size, forms the dot products with the table (64-bit accumulation, one
// multiply-add per limb and coefficient), and multiplies by the product of
// g_k over the block index's bits; a second kernel sums each column's
// blocks.  Sums of field elements are exact, so the grouping does not change
// the result.
#include <string.h>

#include <algorithm>
#include <mutex>
#include <vector>

#include "common.hpp"
#include "field.hpp"
#include "internal.hpp"

namespace stwo_hip {

constexpr uint32_t kEvalBlockLog = 11;  // 2048 coefficients per workgroup and block

struct EvalCol {
  const uint32_t *coeffs;
  uint32_t log_size;
  uint32_t block_begin;  // first block (partial) index of this column
};
struct EvalFactors {
  uint32_t g[32][4];  // g_0 = y, g_1 = x, g_k = pi^(k-1)(x)
};
static_assert(sizeof(EvalFactors) == 4 * kOodsFactorWords, "the device OODS point's factor layout");

__device__ __forceinline__ qm31 ld_factor(const EvalFactors *f, uint32_t k) {
  return qm31_make(f->g[k][0], f->g[k][1], f->g[k][2], f->g[k][3]);
}

// Basis values of the low kEvalBlockLog bits: basis[t] = prod of g_k over the
// set bits k of t (k < kEvalBlockLog), 4 words each.  The basis of index
// j = (b << B) + t is basis[t] times the product of g_k over the set bits of
// b at positions k >= B, so one table serves every column and block.
__global__ void __launch_bounds__(256) eval_basis_kernel(const EvalFactors *f, uint32_t *basis, uint32_t n) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  qm31 r = qm31_make(1, 0, 0, 0);
  for (uint32_t k = 0; (t >> k) != 0; k++)
    if ((t >> k) & 1) r = qm31_mul(r, ld_factor(f, k));
  for (int w = 0; w < 4; w++) basis[4 * (size_t)t + w] = r.v[w];
}

// One workgroup: block `block` (2^B coefficients, B = min(log, kEvalBlockLog))
// of `n` <= kDotCols consecutive columns of the same log size.  Lane t holds
// the basis values of coefficients 8t .. 8t + 7 in registers and, per
// column, accumulates sum c_j * basis_j limb by limb in 64 bits
// (v_mad_u64_u32; each product is below 2^62, so three fit before a fold
// (x & P) + (x >> 31), which keeps the sum below 2^34).  The 256 lanes' sums
// are reduced through LDS, multiplied by the block's factor and written as
// the column's partial for the block.
constexpr uint32_t kDotCols = 8;
struct EvalWork {
  uint32_t col0, n, block, log_size;
};


__global__ void __launch_bounds__(256) eval_point_dot_kernel(const EvalCol *cols, const EvalWork *work,
                                                             const EvalFactors *f,
                                                             const uint32_t *basis, uint32_t *partials) {
  __shared__ uint32_t red[kDotCols * 4][257];
  __shared__ uint32_t sums[kDotCols * 4];
  const EvalWork w = work[blockIdx.x];
  const uint32_t B = w.log_size < kEvalBlockLog ? w.log_size : kEvalBlockLog;
  const uint32_t nb = 1u << B;
  const uint32_t t0 = threadIdx.x * 8;
  qm31 b[8];
#pragma unroll
  for (int i = 0; i < 8; i++) {
    const uint32_t t = t0 + i;
    b[i] = t < nb ? qm31_make(basis[4 * t], basis[4 * t + 1], basis[4 * t + 2], basis[4 * t + 3])
                  : qm31_make(0, 0, 0, 0);
  }
  // Every coefficient load of the workgroup is issued before the first
  // product (one memory latency per workgroup rather than one per column):
  // 8 words per lane and column, as two 16-byte loads when the block is
  // whole and the column pointer 32-byte aligned (a wave-uniform test).
  uint32_t cv[kDotCols][8];
#pragma unroll
  for (uint32_t j = 0; j < kDotCols; j++) {
    if (j < w.n) {
      const uint32_t *c = cols[w.col0 + j].coeffs + ((size_t)w.block << B);
      if (nb == (1u << kEvalBlockLog) && ((uintptr_t)c & 31) == 0) {
        const u32x4 lo = *(gconst_u4 *)(as_global(c) + t0), hi = *(gconst_u4 *)(as_global(c) + t0 + 4);
        cv[j][0] = lo.x;
        cv[j][1] = lo.y;
        cv[j][2] = lo.z;
        cv[j][3] = lo.w;
        cv[j][4] = hi.x;
        cv[j][5] = hi.y;
        cv[j][6] = hi.z;
        cv[j][7] = hi.w;
      } else {
#pragma unroll
        for (int i = 0; i < 8; i++) cv[j][i] = (t0 + i < nb) ? as_global(c)[t0 + i] : 0;
      }
    }
  }
#pragma unroll
  for (uint32_t j = 0; j < kDotCols; j++) {
    if (j >= w.n) break;
    uint64_t acc[4] = {0, 0, 0, 0};
#pragma unroll
    for (int i = 0; i < 8; i++) {
#pragma unroll
      for (int k = 0; k < 4; k++) acc[k] += (uint64_t)cv[j][i] * b[i].v[k];
      if (i == 2 || i == 5)
#pragma unroll
        for (int k = 0; k < 4; k++) acc[k] = m31_fold64(acc[k]);
    }
#pragma unroll
    for (int k = 0; k < 4; k++) red[4 * j + k][threadIdx.x] = m31_reduce64(acc[k]);
  }
  __syncthreads();
  // Row r (column r / 4, limb r % 4): 8 lanes sum 32 entries each, then
  // combine across the 8 lanes (consecutive lanes of one wave).
  const uint32_t r = threadIdx.x >> 3, seg = threadIdx.x & 7;
  if (r < 4 * w.n) {
    uint64_t s = 0;
#pragma unroll
    for (int i = 0; i < 32; i++) s += red[r][seg * 32 + i];
    uint32_t v = m31_reduce64(s);
#pragma unroll
    for (int o = 1; o < 8; o <<= 1) v = m31_add(v, __shfl_xor(v, o, 8));
    if (seg == 0) sums[r] = v;
  }
  __syncthreads();
  if (threadIdx.x < w.n) {
    const uint32_t j = threadIdx.x;
    qm31 res = qm31_make(sums[4 * j], sums[4 * j + 1], sums[4 * j + 2], sums[4 * j + 3]);
    // the block index's set bits select g_B, g_{B+1}, ...
    for (uint32_t k = B; k < w.log_size; k++)
      if ((w.block >> (k - B)) & 1) res = qm31_mul(res, ld_factor(f, k));
    const size_t o = 4 * ((size_t)cols[w.col0 + j].block_begin + w.block);
    for (int q = 0; q < 4; q++) partials[o + q] = res.v[q];
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

// ---- fixed scratch for the entry point that allocates nothing -------------
//
// stwo_hip_eval_at_point_noalloc keeps its basis table, block partials and
// per-column sums in this file-scope array (one copy per device, held until
// the process ends) instead of buffers it would allocate.  Layout in words:
// basis [0, kNaBasisWords), sums [kNaBasisWords, + 4 n_cols), partials after
// them.  Host code never passes the array's name to a kernel or a copy (on
// the host it is only a stand-in for the symbol): it uses the per-device
// address that hipGetSymbolAddress returned at context creation.
constexpr size_t kNaScratchWords = 1u << 18;  // 1 MiB
constexpr size_t kNaBasisWords = (size_t)4 << kEvalBlockLog;
static __device__ __align__(256) uint32_t scratch[kNaScratchWords];

constexpr int kMaxDevices = 64;
static std::mutex g_na_mu;
static uint32_t *g_na_ptr[kMaxDevices];

}  // namespace stwo_hip

namespace stwo_hip {

// Resolves this device's address of `scratch` (the device is current).
// Called once per device, from context creation.
int eval_point_scratch_setup(stwo_hip_ctx *ctx) {
  const int dev = ctx->device;
  if (dev < 0 || dev >= kMaxDevices)
    return set_error(ctx, STWO_HIP_EDEVICE, "device %d beyond the %d the library supports", dev, kMaxDevices);
  std::lock_guard<std::mutex> lk(g_na_mu);
  if (g_na_ptr[dev]) return STWO_HIP_OK;
  void *ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !ptr)
    return set_error(ctx, STWO_HIP_EDEVICE, "hipGetSymbolAddress(scratch) on device %d: %s", dev,
                     hipGetErrorString(e));
  g_na_ptr[dev] = (uint32_t *)ptr;
  return STWO_HIP_OK;
}

// `na_scratch`: NULL = buffers from the context's scratch pool (allocated on
// first use or growth); otherwise the fixed per-device array, with every
// size checked against it and no allocating call made.  `d_factors`: the
// point's factors already on the device (EvalFactors layout), `point` unused;
// `d_values`: the values stay on the device there (4 words per column),
// `values` unused.
static int eval_at_point_impl(stwo_hip_ctx *ctx, const uint32_t *const *coeffs, const uint32_t *log_sizes,
                              size_t n_cols, const stwo_hip_secure_point *point, uint32_t *values,
                              uint32_t *na_scratch, const uint32_t *d_factors = nullptr,
                              uint32_t *d_values = nullptr) {
  if (!ctx) return STWO_HIP_EINVAL;
  STWO_HIP_CHECK_ARG(ctx, (point || d_factors) && (values || d_values) && (coeffs || !n_cols) &&
                              (log_sizes || !n_cols),
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
  if (!d_factors) {
    qm31 x = qm31_make(point->x[0], point->x[1], point->x[2], point->x[3]);
    memcpy(f.g[0], point->y, 16);
    for (uint32_t k = 1; k < 32 && k < max_log + 1; k++) {
      memcpy(f.g[k], x.v, 16);
      qm31 sq = qm31_mul(x, x);
      x = qm31_sub(qm31_add(sq, sq), qm31_make(1, 0, 0, 0));
    }
  }
  // Work items: runs of up to kDotCols consecutive columns of one log size, per block.
  std::vector<EvalWork> work;
  for (size_t c = 0; c < n_cols;) {
    size_t e = c + 1;
    while (e < n_cols && e - c < kDotCols && log_sizes[e] == log_sizes[c]) e++;
    const uint32_t lg = log_sizes[c];
    const uint32_t nblk = 1u << (lg > kEvalBlockLog ? lg - kEvalBlockLog : 0);
    for (uint32_t k = 0; k < nblk; k++) work.push_back({(uint32_t)c, (uint32_t)(e - c), k, lg});
    c = e;
  }
  const uint32_t n_basis = 1u << std::min(max_log, kEvalBlockLog);
  if (na_scratch) {
    // Everything must fit the fixed array, and the tables the ring arena
    // (larger ones would take an allocated buffer).
    STWO_HIP_CHECK_ARG(ctx, kNaBasisWords + 4 * ((uint64_t)n_cols + total) <= kNaScratchWords,
                       "eval_at_point_noalloc: %zu columns and %llu blocks of 2^11 coefficients need %llu words of "
                       "scratch, more than the %zu available",
                       n_cols, (unsigned long long)total,
                       (unsigned long long)(kNaBasisWords + 4 * ((uint64_t)n_cols + total)), kNaScratchWords);
    STWO_HIP_CHECK_ARG(ctx,
                       tab.size() * sizeof(EvalCol) + work.size() * sizeof(EvalWork) + sizeof(EvalFactors) + 512 <=
                           ctx->arena_cap / 2,
                       "eval_at_point_noalloc: %zu work items exceed the table arena", work.size());
  }
  void *d_tab = nullptr, *d_work = nullptr, *d_basis = nullptr, *d_part = nullptr, *d_out = nullptr;
  const EvalFactors *d_f = (const EvalFactors *)d_factors;
  int rc;
  {
    const void *srcs[3] = {tab.data(), work.data(), &f};
    const size_t bytes[3] = {tab.size() * sizeof(EvalCol), work.size() * sizeof(EvalWork), sizeof f};
    void *d[3];
    // with the point on the device (the prove) the tables repeat from call to call
    rc = (d_factors && !na_scratch) ? upload_tables_cached(ctx, "eval_at_point", 2, srcs, bytes, d)
                                    : upload_tables(ctx, d_factors ? 2 : 3, srcs, bytes, d);
    if (rc) return rc;
    d_tab = d[0];
    d_work = d[1];
    if (!d_factors) d_f = (const EvalFactors *)d[2];
  }
  if (na_scratch) {
    d_basis = na_scratch;
    d_out = na_scratch + kNaBasisWords;
    d_part = na_scratch + kNaBasisWords + 4 * n_cols;
  } else {
    rc = get_scratch(ctx, "eval_point_basis", (size_t)n_basis * 16, &d_basis);
    if (rc) return rc;
    rc = get_scratch(ctx, "eval_point_partials", total * 16, &d_part);
    if (rc) return rc;
    rc = get_scratch(ctx, "eval_point_out", n_cols * 16, &d_out);
    if (rc) return rc;
  }
  if (d_values) d_out = d_values;
  rc = launch(ctx, "eval_at_point_basis", eval_basis_kernel, dim3((n_basis + 255) / 256), dim3(256), 0, d_f,
              (uint32_t *)d_basis, n_basis);
  if (rc) return rc;
  rc = launch(ctx, "eval_at_point_dot", eval_point_dot_kernel, dim3((uint32_t)work.size()), dim3(256), 0,
              (const EvalCol *)d_tab, (const EvalWork *)d_work, d_f, (const uint32_t *)d_basis, (uint32_t *)d_part);
  if (rc) return rc;
  rc = launch(ctx, "eval_at_point_sum", eval_point_sum_kernel, dim3((uint32_t)n_cols), dim3(256), 0,
              (const EvalCol *)d_tab, (uint32_t)n_cols, (uint32_t)total, (const uint32_t *)d_part, (uint32_t *)d_out);
  if (rc) return rc;
  if (d_values) return STWO_HIP_OK;
  return stwo_hip_download(ctx, values, (const uint32_t *)d_out, 4 * n_cols);
}

int eval_at_point_dev(stwo_hip_ctx *ctx, const uint32_t *const *coeffs, const uint32_t *log_sizes, size_t n_cols,
                      const uint32_t *d_factors, uint32_t *d_values) {
  if (!ctx) return STWO_HIP_EINVAL;
  STWO_HIP_CHECK_ARG(ctx, d_factors && d_values, "eval_at_point: NULL device factors or values");
  return eval_at_point_impl(ctx, coeffs, log_sizes, n_cols, nullptr, nullptr, nullptr, d_factors, d_values);
}

}  // namespace stwo_hip

using namespace stwo_hip;

extern "C" {

int stwo_hip_eval_at_point(stwo_hip_ctx *ctx, const uint32_t *const *coeffs, const uint32_t *log_sizes, size_t n_cols,
                           const stwo_hip_secure_point *point, uint32_t *values) {
  return eval_at_point_impl(ctx, coeffs, log_sizes, n_cols, point, values, nullptr);
}

int stwo_hip_eval_at_point_noalloc(stwo_hip_ctx *ctx, const uint32_t *const *coeffs, const uint32_t *log_sizes,
                                   size_t n_cols, const stwo_hip_secure_point *point, uint32_t *values) {
  if (!ctx) return STWO_HIP_EINVAL;
  uint32_t *ptr = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_na_mu);
    if (ctx->device >= 0 && ctx->device < kMaxDevices) ptr = g_na_ptr[ctx->device];
  }
  STWO_HIP_CHECK_ARG(ctx, ptr != nullptr, "eval_at_point_noalloc: scratch of device %d not set up", ctx->device);
  int rc = arena_quiesce(ctx);
  if (rc) return rc;
  return eval_at_point_impl(ctx, coeffs, log_sizes, n_cols, point, values, ptr);
}

}  // extern "C"

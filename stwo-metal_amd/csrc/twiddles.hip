// Twiddle trees, built on the device and kept resident.
//
// Layout = CpuBackend's slow_precompute_twiddles (CPUB/circle.rs:294-311):
// for each doubling C_l of the root coset (l = 0, 1, ...), the x-coordinates
// of its first half in bit-reversed order, then a final pad of 1; the tree
// length equals the root size.  The inverse tree holds elementwise inverses
// (PolyOps::precompute_twiddles, CPUB/circle.rs:252-292).
//
// The reference builds one level per dispatch by square-and-multiply, then
// inverts on the GPU and reads back (MB/poly.rs:185-224, SH/twiddles.metal:12,
// SH/fields.metal:3).  Here one launch computes every entry directly: entry k
// finds its level from the distance to the end of the tree, forms its point
// index and sums the set bits' powers of the generator from a 31-entry
// constant table.
#include "common.hpp"
#include "field.hpp"

namespace stwo_hip {

__constant__ cpoint c_gen_pow2[31];  // M31_CIRCLE_GEN * 2^k

__device__ __forceinline__ cpoint dev_index_to_point(uint32_t idx) {
  cpoint r = {1, 0};
#pragma unroll 1
  for (int k = 0; k < 31; k++)
    if (idx >> k & 1) r = cp_add(r, c_gen_pow2[k]);
  return r;
}

// Also writes both trees doubled (tw2 = 2 tw, itw2 = 2 itw as u32), the
// operands of the radix-16 FFT's multiplications.
__global__ void __launch_bounds__(256) twiddle_tree_kernel(uint32_t *tw, uint32_t *itw, uint32_t *tw2, uint32_t *itw2,
                                                           uint32_t root_initial, uint32_t r) {
  uint32_t len = 1u << r;
  uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= len) return;
  if (k == len - 1) {
    tw[k] = 1;
    itw[k] = 1;
    tw2[k] = 2;
    itw2[k] = 2;
    return;
  }
  uint32_t d = len - k;                  // >= 2
  uint32_t cl = 32 - __clz(d - 1);       // ceil(log2 d) = log size of this level's coset
  uint32_t level = r - cl;
  uint32_t j = k - (len - (1u << cl));   // index within the level, < 2^(cl-1)
  uint32_t initial = (uint32_t)(((uint64_t)root_initial << level) & IDX_MASK);
  uint32_t idx = (initial + bitrev32(j, cl - 1) * coset_step(cl)) & IDX_MASK;
  uint32_t x = dev_index_to_point(idx).x;
  uint32_t xi = m31_inv(x);
  tw[k] = x;
  itw[k] = xi;
  tw2[k] = x << 1;
  itw2[k] = xi << 1;
}

static int upload_gen_table(stwo_hip_ctx *ctx) {
  cpoint pw[31];
  cpoint cur = {2, 1268011823u};
  for (int k = 0; k < 31; k++) {
    pw[k] = cur;
    cur = cp_add(cur, cur);
  }
  STWO_HIP_TRY(ctx, hipMemcpyToSymbolAsync(HIP_SYMBOL(c_gen_pow2), pw, sizeof pw, 0, hipMemcpyHostToDevice,
                                           ctx->stream));
  return STWO_HIP_OK;
}

}  // namespace stwo_hip

using namespace stwo_hip;

extern "C" {

int stwo_hip_twiddles_create(stwo_hip_ctx *ctx, stwo_hip_coset root, stwo_hip_twiddles **out) {
  if (!ctx || !out) return STWO_HIP_EINVAL;
  *out = nullptr;
  STWO_HIP_CHECK_ARG(ctx, root.log_size <= 30, "twiddle root log size %u exceeds 30", root.log_size);
  int rc = upload_gen_table(ctx);
  if (rc) return rc;
  stwo_hip_twiddles *tw = new (std::nothrow) stwo_hip_twiddles();
  if (!tw) return set_error(ctx, STWO_HIP_ENOMEM, "host allocation failed");
  tw->device = ctx->device;
  tw->root = {root.initial_index & IDX_MASK, root.log_size};
  tw->len = (size_t)1 << root.log_size;
  // One allocation: tree | inverse tree | both doubled.
  if (ctx_malloc(ctx, (void **)&tw->d_tw, tw->len * 16) != hipSuccess) {
    (void)hipGetLastError();  // clear the sticky error of the failed allocation
    delete tw;
    return set_error(ctx, STWO_HIP_ENOMEM, "twiddle tree of %zu entries: out of device memory", (size_t)1 << root.log_size);
  }
  tw->d_itw = tw->d_tw + tw->len;
  tw->d_tw2 = tw->d_tw + 2 * tw->len;
  tw->d_itw2 = tw->d_tw + 3 * tw->len;
  rc = launch(ctx, "twiddle_tree", twiddle_tree_kernel, dim3((uint32_t)((tw->len + 255) / 256)), dim3(256), 0,
              tw->d_tw, tw->d_itw, tw->d_tw2, tw->d_itw2, tw->root.initial_index, root.log_size);
  if (rc == STWO_HIP_OK && hipStreamSynchronize(ctx->stream) != hipSuccess)
    rc = set_error(ctx, STWO_HIP_EDEVICE, "twiddle tree build failed");
  if (rc) {
    (void)hipFree(tw->d_tw);
    delete tw;
    return rc;
  }
  *out = tw;
  return STWO_HIP_OK;
}

int stwo_hip_twiddles_destroy(stwo_hip_ctx *ctx, stwo_hip_twiddles *tw) {
  if (!ctx || !tw) return STWO_HIP_EINVAL;
  STWO_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  (void)hipFree(tw->d_tw);
  delete tw;
  return STWO_HIP_OK;
}

int stwo_hip_twiddles_buffers(const stwo_hip_twiddles *tw, const uint32_t **twiddles, const uint32_t **itwiddles,
                              size_t *len) {
  if (!tw) return STWO_HIP_EINVAL;
  if (twiddles) *twiddles = tw->d_tw;
  if (itwiddles) *itwiddles = tw->d_itw;
  if (len) *len = tw->len;
  return STWO_HIP_OK;
}

}  // extern "C"

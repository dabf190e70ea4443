// Blake2s Merkle commitments: mixed-degree (MerkleOps::commit_on_layer,
// MerkleProver::commit) and lifted (MerkleOpsLifted).
//
// One hash per lane.  A node's message is streamed in 64-byte blocks exactly
// as the `blake2` crate buffers it (the last block, full or partial, is the
// finalized one, `t` counts the bytes so far):
//   mixed node i  = [prev[2i] || prev[2i+1]] || le32(c0[i]) || le32(c1[i]) || ...
//                   (STWO/core/vcs/blake2_merkle.rs:17-38, SIMD/blake2s.rs:60-144)
//   lifted leaf i = le32(c[lift(i, L - log c)]) for columns in ascending
//                   length order (CPUB/merkle_lifted.rs:44-78)
//   lifted node i = prev[2i] || prev[2i+1] (vcs_lifted/blake2_merkle.rs:18-27)
// Column words of one block are 16 columns at the same row: lane i reads
// column c at row i, so each load instruction is a coalesced run over the
// wave.  Column pointers come from a per-call device table (uniform across
// the wave -> scalar loads).
//
// Node layers without injected columns run in `merkle_nodes_kernel`, which
// hashes up to 8 consecutive layers per launch: each 256-lane workgroup
// hashes 256 nodes, keeps the digests in LDS and halves the active lanes per
// layer, writing every layer to global memory.
#include <algorithm>
#include <vector>

#include "blake2s.hpp"
#include "common.hpp"
#include "field.hpp"
#include "internal.hpp"
#include "merkle_dev.hpp"

namespace stwo_hip {


__device__ __forceinline__ void load_block64(uint32_t m[16], const uint32_t *src) {
  const uint4 *s = reinterpret_cast<const uint4 *>(src);
#pragma unroll
  for (int k = 0; k < 4; k++) {
    uint4 v = s[k];
    m[4 * k] = v.x;
    m[4 * k + 1] = v.y;
    m[4 * k + 2] = v.z;
    m[4 * k + 3] = v.w;
  }
}

// Word idx of a column.  Layers are at most 2^30 rows, so the byte offset
// fits 32 bits: the load takes the column pointer as an SGPR base and a
// 32-bit lane offset (no per-column 64-bit address arithmetic).
__device__ __forceinline__ uint32_t ld_col_word(const uint32_t *col, uint32_t idx) {
  return *(gconst_u32 *)((gconst_u8 *)as_global(col) + (idx << 2));
}

// Mixed-degree layer: optional previous layer, then the layer's columns.
// `lifted` selects the lifted leaf index for column words.
// AHEAD: each block's column words are loaded one compression ahead, so
// their latency overlaps the previous compression.  Used for layers of few
// nodes, where a node's chain of blocks is the launch's critical path; on
// large layers the other waves hide the latency and the extra registers and
// copies cost more than they save.
template <bool AHEAD>
__global__ void __launch_bounds__(256) merkle_layer_kernel(const uint32_t *prev, const ColRef *cols, uint32_t n_cols,
                                                           uint32_t log_size, uint32_t *out, uint32_t m31,
                                                           uint32_t lifted) {
  const uint32_t n = 1u << log_size;
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  // The words of columns c0 .. c0 + 15 at this row (zeros past n_cols).
  auto load_cols = [&](uint32_t c0, uint32_t (&dst)[16]) {
    const uint32_t cnt = min(16u, n_cols - c0);
#pragma unroll
    for (uint32_t k = 0; k < 16; k++) {
      uint32_t v = 0;
      if (k < cnt) {
        const ColRef cr = cols[c0 + k];
        uint32_t idx = lifted ? (((i >> (cr.shift + 1)) << 1) + (i & 1)) : i;
        v = ld_col_word(cr.p, idx);
      }
      dst[k] = v;
    }
  };
  b2_state s;
  b2_init(s);
  uint32_t m[16], nx[16];
  uint32_t t = 0;
  if (AHEAD && n_cols) load_cols(0, nx);
  if (prev) {
    load_block64(m, prev + 16 * (size_t)i);
    t = 64;
    b2_compress(s, m, t, n_cols == 0);
  } else if (n_cols == 0) {
#pragma unroll
    for (int k = 0; k < 16; k++) m[k] = 0;
    b2_compress(s, m, 0, true);
  }
  for (uint32_t c0 = 0; c0 < n_cols; c0 += 16) {
    const uint32_t cnt = min(16u, n_cols - c0);
    if (AHEAD) {
#pragma unroll
      for (int k = 0; k < 16; k++) m[k] = nx[k];
      if (c0 + 16 < n_cols) load_cols(c0 + 16, nx);
    } else {
      load_cols(c0, m);
    }
    t += 4 * cnt;
    b2_compress(s, m, t, c0 + 16 >= n_cols);
  }
  store_hash(out + 8 * (size_t)i, s, m31);
}

// Layers of few nodes, where a node's chain of blocks is the launch's
// critical path: one node per quad (b2_compress_quad, about 410
// instructions per lane instead of 960 on one lane), each block's column
// words loaded one compression ahead.  Same inputs and outputs as
// merkle_layer_kernel; the grid has 4 lanes per node.
__global__ void __launch_bounds__(256) merkle_layer_quad_kernel(const uint32_t *prev, const ColRef *cols,
                                                                uint32_t n_cols, uint32_t log_size, uint32_t *out,
                                                                uint32_t m31, uint32_t lifted) {
  const uint32_t n = 1u << log_size;
  const uint32_t i = (blockIdx.x * blockDim.x + threadIdx.x) >> 2, j = threadIdx.x & 3;
  if (i >= n) return;  // a quad's four lanes share i and leave together
  auto load_cols = [&](uint32_t c0, uint32_t (&dst)[16]) {
    const uint32_t cnt = min(16u, n_cols - c0);
#pragma unroll
    for (uint32_t k = 0; k < 16; k++) {
      uint32_t v = 0;
      if (k < cnt) {
        const ColRef cr = cols[c0 + k];
        uint32_t idx = lifted ? (((i >> (cr.shift + 1)) << 1) + (i & 1)) : i;
        v = ld_col_word(cr.p, idx);
      }
      dst[k] = v;
    }
  };
  uint32_t lo = b2_iv_lo(j), hi = b2_iv_hi(j);
  uint32_t m[16], nx[16];
  uint32_t t = 0;
  if (n_cols) load_cols(0, nx);
  if (prev) {
    load_block64(m, prev + 16 * (size_t)i);
    t = 64;
    b2_compress_quad(lo, hi, m, t, n_cols == 0);
  } else if (n_cols == 0) {
#pragma unroll
    for (int k = 0; k < 16; k++) m[k] = 0;
    b2_compress_quad(lo, hi, m, 0, true);
  }
  for (uint32_t c0 = 0; c0 < n_cols; c0 += 16) {
    const uint32_t cnt = min(16u, n_cols - c0);
#pragma unroll
    for (int k = 0; k < 16; k++) m[k] = nx[k];
    if (c0 + 16 < n_cols) load_cols(c0 + 16, nx);
    t += 4 * cnt;
    b2_compress_quad(lo, hi, m, t, c0 + 16 >= n_cols);
  }
  if (m31) {
    lo = b2_reduce_word(lo);
    hi = b2_reduce_word(hi);
  }
  g_u32 *o = as_global(out + 8 * (size_t)i);
  o[j] = lo;
  o[4 + j] = hi;
}

// Small mixed-degree layers with many columns (up to 2^14 nodes): the chain
// of blocks per node is the critical path and, with one block loaded ahead,
// each block waits on a table load followed by a column load.  Here a
// workgroup owns 64 consecutive nodes (a quad each) and stages its columns
// through LDS in chunks of MQ_CHUNK columns: the chunk's column pointers
// first, then every column's 64 words (one coalesced 256-byte run per wave
// and load, 32 loads in flight per lane), so a chunk costs about two memory
// round trips and its blocks then read their words from LDS.  Same inputs
// and outputs as merkle_layer_kernel (mixed layers only, not lifted).
constexpr uint32_t MQ_CHUNK = 240;  // 15 blocks; 61,440 B of words + 1,920 B of pointers
__global__ void __launch_bounds__(256) merkle_layer_quad_lds_kernel(const uint32_t *prev, const ColRef *cols,
                                                                    uint32_t n_cols, uint32_t log_size,
                                                                    uint32_t *out, uint32_t m31) {
  __shared__ const uint32_t *ptrs[MQ_CHUNK];
  __shared__ uint32_t tile[MQ_CHUNK * 64];
  const uint32_t n = 1u << log_size;
  const uint32_t tid = threadIdx.x, node = tid >> 2, j = tid & 3;
  const uint32_t i0 = blockIdx.x * 64, i = i0 + node;
  const bool active = i < n;  // every lane stays for the barriers
  uint32_t lo = b2_iv_lo(j), hi = b2_iv_hi(j);
  uint32_t m[16];
  uint32_t t = 0;
  if (prev) {
    if (active) {
      load_block64(m, prev + 16 * (size_t)i);
      b2_compress_quad(lo, hi, m, 64, n_cols == 0);
    }
    t = 64;
  } else if (n_cols == 0) {
#pragma unroll
    for (int k = 0; k < 16; k++) m[k] = 0;
    if (active) b2_compress_quad(lo, hi, m, 0, true);
  }
  const uint32_t w = tid & 63, row = i0 + w;
  for (uint32_t cb = 0; cb < n_cols; cb += MQ_CHUNK) {
    const uint32_t cnt = min(MQ_CHUNK, n_cols - cb), words = cnt * 64;
    __syncthreads();  // the previous chunk's words are consumed
    for (uint32_t c = tid; c < cnt; c += 256) ptrs[c] = cols[cb + c].p;
    __syncthreads();
    for (uint32_t e0 = 0; e0 < words; e0 += 256 * 32) {
      uint32_t v[32];
#pragma unroll
      for (uint32_t k = 0; k < 32; k++) {
        const uint32_t e = e0 + tid + 256 * k;  // column e >> 6 is uniform over the wave
        v[k] = (e < words && row < n) ? ld_col_word(ptrs[e >> 6], row) : 0u;
      }
#pragma unroll
      for (uint32_t k = 0; k < 32; k++) {
        const uint32_t e = e0 + tid + 256 * k;
        if (e < words) tile[e] = v[k];
      }
    }
    __syncthreads();
    for (uint32_t b = 0; b < cnt; b += 16) {
      const uint32_t bc = min(16u, cnt - b);
#pragma unroll
      for (uint32_t k = 0; k < 16; k++) m[k] = k < bc ? tile[(b + k) * 64 + node] : 0u;
      t += 4 * bc;
      if (active) b2_compress_quad(lo, hi, m, t, cb + b + 16 >= n_cols);
    }
  }
  if (!active) return;
  if (m31) {
    lo = b2_reduce_word(lo);
    hi = b2_reduce_word(hi);
  }
  g_u32 *o = as_global(out + 8 * (size_t)i);
  o[j] = lo;
  o[4 + j] = hi;
}

// Hashes `n_layers` (<= 8) consecutive node layers below `prev` (of log size
// out_log + 1).  Layer k (k = 0 first) has log size out_log - k and is
// written at out_base + offset(out_log - k) in the packed layout.
__global__ void __launch_bounds__(256) merkle_nodes_kernel(const uint32_t *prev, uint32_t *packed, uint32_t out_log,
                                                           uint32_t n_layers, uint32_t m31) {
  __shared__ uint32_t hs[256][9];
  const uint32_t n = 1u << out_log;
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  uint32_t m[16];
  b2_state s;
  if (i < n) {
    load_block64(m, prev + 16 * (size_t)i);
    b2_init(s);
    b2_compress(s, m, 64, true);
    if (m31)
#pragma unroll
      for (int k = 0; k < 8; k++) s.h[k] = b2_reduce_word(s.h[k]);
    store_hash(packed + (((size_t)1 << out_log) - 1) * 8 + 8 * (size_t)i, s, 0);
#pragma unroll
    for (int k = 0; k < 8; k++) hs[threadIdx.x][k] = s.h[k];
  }
  subtree_levels(hs, packed, out_log, n_layers - 1, m31);
}

static inline size_t layer_off(uint32_t l) { return tree_layer_off(l); }

int launch_merkle_layer(stwo_hip_ctx *ctx, const uint32_t *prev, const ColRef *d_cols, uint32_t n_cols,
                        uint32_t log_size, uint32_t *out, int mode, bool lifted) {
  STWO_HIP_CHECK_ARG(ctx, mode == STWO_HIP_HASH_BLAKE2S || mode == STWO_HIP_HASH_BLAKE2S_M31,
                     "merkle: unknown hash mode %d", mode);
  uint32_t n = 1u << log_size;
  const char *name = lifted ? "merkle_lifted_leaves" : "merkle_commit_on_layer";
  const uint32_t m31 = (uint32_t)(mode == STWO_HIP_HASH_BLAKE2S_M31);
  // Up to 2^14 nodes with more than one block per node, the layer is
  // latency-bound even at four lanes per node: a quad per node (mixed layers
  // with the columns staged through LDS, lifted ones loading a block ahead).  Below 2^16
  // nodes, a lane per node with the loads one block ahead.
  if (log_size <= 14 && n_cols > 16 && !lifted)
    return launch(ctx, name, merkle_layer_quad_lds_kernel, dim3((n + 63) / 64), dim3(256), 0, prev, d_cols, n_cols,
                  log_size, out, m31);
  if (log_size <= 14 && n_cols > 16)
    return launch(ctx, name, merkle_layer_quad_kernel, dim3((4 * n + 255) / 256), dim3(256), 0, prev, d_cols, n_cols,
                  log_size, out, m31, (uint32_t)lifted);
  const bool ahead = log_size < 16 && n_cols > 16;
  return launch(ctx, name, ahead ? merkle_layer_kernel<true> : merkle_layer_kernel<false>, dim3((n + 255) / 256),
                dim3(256), 0, prev, d_cols, n_cols, log_size, out, m31, (uint32_t)lifted);
}

static int launch_layer(stwo_hip_ctx *ctx, const uint32_t *prev, const std::vector<ColRef> &cols, uint32_t log_size,
                        uint32_t *out, int mode, bool lifted) {
  void *dtab = nullptr;
  if (!cols.empty()) {
    int rc = upload_table(ctx, cols.data(), cols.size() * sizeof(ColRef), &dtab);
    if (rc) return rc;
  }
  return launch_merkle_layer(ctx, prev, (const ColRef *)dtab, (uint32_t)cols.size(), log_size, out, mode, lifted);
}

// Hashes node layers top_log-1 .. stop_log of a packed tree whose layer top_log is filled.
int hash_node_layers(stwo_hip_ctx *ctx, uint32_t *layers, uint32_t top_log, uint32_t stop_log, int mode) {
  // layers [stop_log, top_log) are produced, from top_log - 1 down.
  STWO_HIP_CHECK_ARG(ctx, mode == STWO_HIP_HASH_BLAKE2S || mode == STWO_HIP_HASH_BLAKE2S_M31,
                     "merkle: unknown hash mode %d", mode);
  uint32_t l = top_log;
  while (l > stop_log) {
    uint32_t out_log = l - 1;
    uint32_t n_layers = std::min<uint32_t>(8, out_log - stop_log + 1);
    // within one workgroup of 256 outputs we can go down at most 8 layers
    uint32_t n_out = 1u << out_log;
    uint32_t blocks = (n_out + 255) / 256;
    if (n_out < 256) n_layers = std::min<uint32_t>(n_layers, out_log + 1);
    int rc = launch(ctx, "merkle_nodes", merkle_nodes_kernel, dim3(blocks), dim3(256), 0,
                    (const uint32_t *)(layers + layer_off(l)), layers, out_log, n_layers,
                    (uint32_t)(mode == STWO_HIP_HASH_BLAKE2S_M31));
    if (rc) return rc;
    l = out_log + 1 - n_layers;
  }
  return STWO_HIP_OK;
}

}  // namespace stwo_hip

using namespace stwo_hip;

extern "C" {

int stwo_hip_merkle_commit_on_layer(stwo_hip_ctx *ctx, uint32_t log_size, const uint32_t *prev,
                                    const uint32_t *const *cols, size_t n_cols, uint32_t *out, int hash_mode) {
  if (!ctx) return STWO_HIP_EINVAL;
  STWO_HIP_CHECK_ARG(ctx, out != nullptr && log_size <= 30, "commit_on_layer: bad output");
  STWO_HIP_CHECK_ARG(ctx, n_cols == 0 || cols != nullptr, "commit_on_layer: NULL column array");
  std::vector<ColRef> refs(n_cols);
  for (size_t c = 0; c < n_cols; c++) {
    STWO_HIP_CHECK_ARG(ctx, cols[c] != nullptr, "commit_on_layer: NULL column %zu", c);
    refs[c] = {cols[c], 0, 0};
  }
  return launch_layer(ctx, prev, refs, log_size, out, hash_mode, false);
}

int stwo_hip_merkle_commit(stwo_hip_ctx *ctx, const uint32_t *const *cols, const uint32_t *log_sizes, size_t n_cols,
                           uint32_t *layers, int hash_mode) {
  if (!ctx) return STWO_HIP_EINVAL;
  STWO_HIP_CHECK_ARG(ctx, layers != nullptr, "merkle_commit: NULL output");
  if (n_cols == 0) return launch_layer(ctx, nullptr, {}, 0, layers, hash_mode, false);
  STWO_HIP_CHECK_ARG(ctx, cols && log_sizes, "merkle_commit: NULL arrays");
  uint32_t max_log = 0;
  for (size_t c = 0; c < n_cols; c++) {
    STWO_HIP_CHECK_ARG(ctx, cols[c] && log_sizes[c] <= 30, "merkle_commit: bad column %zu", c);
    max_log = std::max(max_log, log_sizes[c]);
  }
  int l = (int)max_log;
  while (l >= 0) {
    std::vector<ColRef> refs;
    for (size_t c = 0; c < n_cols; c++)  // input order == stable order among equal lengths
      if (log_sizes[c] == (uint32_t)l) refs.push_back({cols[c], 0, 0});
    const uint32_t *prev = (l == (int)max_log) ? nullptr : layers + layer_off(l + 1);
    if (refs.empty() && prev) {
      // Run of pure node layers down to the next injected layer (or the root).
      int next = -1;
      for (size_t c = 0; c < n_cols; c++)
        if ((int)log_sizes[c] < l) next = std::max(next, (int)log_sizes[c]);
      uint32_t stop = (uint32_t)(next + 1);
      int rc = hash_node_layers(ctx, layers, (uint32_t)l + 1, stop, hash_mode);
      if (rc) return rc;
      l = (int)stop - 1;
      continue;
    }
    int rc = launch_layer(ctx, prev, refs, (uint32_t)l, layers + layer_off((uint32_t)l), hash_mode, false);
    if (rc) return rc;
    l--;
  }
  return STWO_HIP_OK;
}

int stwo_hip_merkle_lifted_build_leaves(stwo_hip_ctx *ctx, const uint32_t *const *cols, const uint32_t *log_sizes,
                                        size_t n_cols, uint32_t lifting_log_size, uint32_t *out, int hash_mode) {
  if (!ctx) return STWO_HIP_EINVAL;
  STWO_HIP_CHECK_ARG(ctx, out != nullptr, "lifted_build_leaves: NULL output");
  if (n_cols == 0) return launch_layer(ctx, nullptr, {}, 0, out, hash_mode, true);
  STWO_HIP_CHECK_ARG(ctx, cols && log_sizes && lifting_log_size <= 30, "lifted_build_leaves: bad arguments");
  std::vector<size_t> order(n_cols);
  for (size_t c = 0; c < n_cols; c++) {
    order[c] = c;
    STWO_HIP_CHECK_ARG(ctx, cols[c] != nullptr, "lifted_build_leaves: NULL column %zu", c);
    STWO_HIP_CHECK_ARG(ctx, log_sizes[c] >= 1, "A column must be of length >= 2.");
    STWO_HIP_CHECK_ARG(ctx, log_sizes[c] <= lifting_log_size, "column %zu longer than the lifting size", c);
  }
  std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) { return log_sizes[a] < log_sizes[b]; });
  std::vector<ColRef> refs;
  for (size_t c : order) refs.push_back({cols[c], lifting_log_size - log_sizes[c], 0});
  return launch_layer(ctx, nullptr, refs, lifting_log_size, out, hash_mode, true);
}

int stwo_hip_merkle_lifted_build_next_layer(stwo_hip_ctx *ctx, const uint32_t *prev, uint32_t out_log_size,
                                            uint32_t *out, int hash_mode) {
  if (!ctx) return STWO_HIP_EINVAL;
  STWO_HIP_CHECK_ARG(ctx, prev && out && out_log_size < 30, "lifted_build_next_layer: bad arguments");
  return launch_layer(ctx, prev, {}, out_log_size, out, hash_mode, false);
}

int stwo_hip_merkle_lifted_commit(stwo_hip_ctx *ctx, const uint32_t *const *cols, const uint32_t *log_sizes,
                                  size_t n_cols, uint32_t lifting_log_size, uint32_t *layers, int hash_mode) {
  if (!ctx) return STWO_HIP_EINVAL;
  STWO_HIP_CHECK_ARG(ctx, layers != nullptr, "lifted_commit: NULL output");
  if (n_cols == 0) return stwo_hip_merkle_lifted_build_leaves(ctx, cols, log_sizes, 0, 0, layers, hash_mode);
  int rc = stwo_hip_merkle_lifted_build_leaves(ctx, cols, log_sizes, n_cols, lifting_log_size,
                                               layers + layer_off(lifting_log_size), hash_mode);
  if (rc) return rc;
  return hash_node_layers(ctx, layers, lifting_log_size, 0, hash_mode);
}

}  // extern "C"

// Device pieces shared by the Merkle node kernels and the fused FRI layer
// kernel: digest stores and the workgroup subtree over 256 held digests.
#pragma once
#include <stdint.h>

#include "blake2s.hpp"
#include "common.hpp"
#include "internal.hpp"

namespace stwo_hip {

__device__ __forceinline__ void store_hash(uint32_t *dst, const b2_state &s, uint32_t m31) {
  uint4 a, b;
  if (m31) {
    a = make_uint4(b2_reduce_word(s.h[0]), b2_reduce_word(s.h[1]), b2_reduce_word(s.h[2]), b2_reduce_word(s.h[3]));
    b = make_uint4(b2_reduce_word(s.h[4]), b2_reduce_word(s.h[5]), b2_reduce_word(s.h[6]), b2_reduce_word(s.h[7]));
  } else {
    a = make_uint4(s.h[0], s.h[1], s.h[2], s.h[3]);
    b = make_uint4(s.h[4], s.h[5], s.h[6], s.h[7]);
  }
  reinterpret_cast<uint4 *>(dst)[0] = a;
  reinterpret_cast<uint4 *>(dst)[1] = b;
}

// Node layers below a layer of log size `top_log` whose digests this
// 256-lane workgroup holds in hs[0 .. min(256, 2^top_log)) (node
// blockIdx.x * 256 + t of that layer in hs[t]).  Hashes `n_levels` layers,
// top_log - 1 down to top_log - n_levels, H(left || right) each, writing
// layer l at packed + tree_layer_off(l) and keeping its digests in hs.  The
// first hs writes must be done before the call (it starts with a barrier).
// Levels of at most 64 hashes use a quad per hash (b2_compress_quad): their
// latency is the critical path of the remaining levels.
__device__ __forceinline__ void subtree_levels(uint32_t (&hs)[256][9], uint32_t *packed, uint32_t top_log,
                                               uint32_t n_levels, uint32_t m31) {
  uint32_t m[16];
  uint32_t active = min(256u, 1u << top_log);
  for (uint32_t layer = 1; layer <= n_levels; layer++) {
    __syncthreads();
    active >>= 1;
    const uint32_t lg = top_log - layer;
    if (active <= 64) {
      const uint32_t q = threadIdx.x >> 2, j = threadIdx.x & 3;
      const bool work = q < active;
      uint32_t lo = b2_iv_lo(j), hi = b2_iv_hi(j);
      if (work) {
#pragma unroll
        for (int k = 0; k < 8; k++) {
          m[k] = hs[2 * q][k];
          m[8 + k] = hs[2 * q + 1][k];
        }
        b2_compress_quad(lo, hi, m, 64, true);
        if (m31) {
          lo = b2_reduce_word(lo);
          hi = b2_reduce_word(hi);
        }
        g_u32 *o = as_global(packed + tree_layer_off(lg) + 8 * ((size_t)blockIdx.x * (256u >> layer) + q));
        o[j] = lo;
        o[4 + j] = hi;
      }
      __syncthreads();
      if (work) {
        hs[q][j] = lo;
        hs[q][4 + j] = hi;
      }
      continue;
    }
    const uint32_t gi = blockIdx.x * (256u >> layer) + threadIdx.x;
    b2_state s2;
    const bool work = threadIdx.x < active;
    if (work) {
#pragma unroll
      for (int k = 0; k < 8; k++) {
        m[k] = hs[2 * threadIdx.x][k];
        m[8 + k] = hs[2 * threadIdx.x + 1][k];
      }
      b2_init(s2);
      b2_compress(s2, m, 64, true);
      if (m31)
#pragma unroll
        for (int k = 0; k < 8; k++) s2.h[k] = b2_reduce_word(s2.h[k]);
      store_hash(packed + tree_layer_off(lg) + 8 * (size_t)gi, s2, 0);
    }
    __syncthreads();
    if (work)
#pragma unroll
      for (int k = 0; k < 8; k++) hs[threadIdx.x][k] = s2.h[k];
  }
}

}  // namespace stwo_hip

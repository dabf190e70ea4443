// Proof of work for the Blake2s channel (GrindOps::grind).
//
// SimdBackend's grind (SIMD/grind.rs:22-62) returns the smallest nonce
// (hi << 32) | lo with lo < 2^20 (GRIND_LOW_BITS), scanning hi = 0, 1, ...,
// such that H(prefixed || le64(nonce)) has at least pow_bits trailing zero
// bits in its first word, where prefixed = H(POW_PREFIX || 0^12 || digest ||
// le32(pow_bits)) (STWO/core/channel/blake2s.rs:122-139).  Here one lane
// hashes one nonce (a single 40-byte Blake2s block) and a launch covers
// 2^hb consecutive hi words; valid lanes take the minimum with a 64-bit
// atomic.  Launches go in nonce order, so the first launch that finds a
// nonce holds the smallest one.  The first launch covers only the lowest
// 2^(pow_bits + 6) lo words of hi = 0 (it misses with probability about
// e^-64), so the usual search hashes a few thousand nonces, not 2^20.
#include <string.h>

#include <algorithm>

#include "blake2s.hpp"
#include "channel.hpp"
#include "common.hpp"
#include "field.hpp"
#include "internal.hpp"

namespace stwo_hip {

constexpr uint32_t kGrindLowBits = 20;

// Lane gid hashes nonce (hi0 + (gid >> 20), lo0 + (gid & (2^20 - 1))).
struct GrindArgs {
  uint32_t prefixed[8];
  uint32_t hi0, lo0;
  uint32_t pow_bits;
};

// d_prefixed: the prefix in device memory, in place of a.prefixed, when not NULL.
__global__ void __launch_bounds__(256) grind_kernel(GrindArgs a, const uint32_t *d_prefixed, unsigned long long *best) {
  const uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t lo = a.lo0 + (gid & ((1u << kGrindLowBits) - 1));
  const uint32_t hi = a.hi0 + (gid >> kGrindLowBits);
  const uint32_t *pre = d_prefixed ? d_prefixed : a.prefixed;
  uint32_t m[16];
#pragma unroll
  for (int k = 0; k < 8; k++) m[k] = pre[k];
  m[8] = lo;
  m[9] = hi;
#pragma unroll
  for (int k = 10; k < 16; k++) m[k] = 0;
  b2_state s;
  b2_init(s);
  b2_compress(s, m, 40, true);
  const uint32_t w = s.h[0];
  const uint32_t tz = w ? (uint32_t)__builtin_ctz(w) : 32u;
  if (tz >= a.pow_bits) atomicMin(best, ((unsigned long long)hi << 32) | lo);
}

__global__ void grind_reset_kernel(unsigned long long *best) {
  if (threadIdx.x == 0) *best = ~0ull;
}

// The first search range of stwo_hip_grind: lo < 2^(pow_bits + 6), hi = 0.
static uint32_t first_range_log(uint32_t pow_bits) { return std::max<uint32_t>(8, pow_bits + 6); }

int grind_first_range(stwo_hip_ctx *ctx, const uint32_t *d_prefixed, uint32_t pow_bits, unsigned long long *best) {
  STWO_HIP_CHECK_ARG(ctx, first_range_log(pow_bits) < kGrindLowBits, "internal: grind range for %u bits", pow_bits);
  GrindArgs a{};
  a.pow_bits = pow_bits;
  return launch(ctx, "grind", grind_kernel, dim3((1u << first_range_log(pow_bits)) / 256), dim3(256), 0, a,
                d_prefixed, best);
}

}  // namespace stwo_hip

using namespace stwo_hip;

extern "C" {

int stwo_hip_grind(stwo_hip_ctx *ctx, const uint8_t channel_digest[32], uint32_t pow_bits, uint64_t *nonce) {
  if (!ctx) return STWO_HIP_EINVAL;
  STWO_HIP_CHECK_ARG(ctx, channel_digest && nonce, "grind: NULL argument");
  STWO_HIP_CHECK_ARG(ctx, pow_bits <= 32, "pow_bits > 32 is not supported");
  GrindArgs a{};
  {
    uint8_t in[52];
    memcpy(in, &kPowPrefix, 4);
    memset(in + 4, 0, 12);
    memcpy(in + 16, channel_digest, 32);
    memcpy(in + 48, &pow_bits, 4);
    b2_host h;
    h.update(in, 52);
    uint8_t d[32];
    h.finalize(d);
    memcpy(a.prefixed, d, 32);
  }
  a.pow_bits = pow_bits;
  void *d_best = nullptr;
  int rc = get_scratch(ctx, "grind_best", 8, &d_best);
  if (rc) return rc;
  // Searches `lanes` nonces from (hi0, lo0) on; 1 when one was found.
  auto search = [&](uint32_t hi0, uint32_t lo0, uint32_t lanes) -> int {
    int r = launch(ctx, "grind_reset", grind_reset_kernel, dim3(1), dim3(64), 0, (unsigned long long *)d_best);
    if (r) return r;
    a.hi0 = hi0;
    a.lo0 = lo0;
    r = launch(ctx, "grind", grind_kernel, dim3(lanes / 256), dim3(256), 0, a, (const uint32_t *)nullptr,
               (unsigned long long *)d_best);
    if (r) return r;
    uint64_t best = 0;
    r = stwo_hip_download(ctx, (uint32_t *)&best, (const uint32_t *)d_best, 2);
    if (r) return r;
    if (best == ~0ull) return 0;
    // The reference asserts both halves are below P (SIMD/grind.rs:50-57).
    STWO_HIP_CHECK_ARG(ctx, (best >> 32) < P && (uint32_t)best < P, "grind: nonce halves not below P");
    *nonce = best;
    return 1;
  };
  uint64_t h_start = 0;
  // lo < 2^k0, then the rest of hi = 0 (both multiples of 256 lanes).
  const uint32_t k0 = first_range_log(pow_bits);
  if (k0 < kGrindLowBits) {
    rc = search(0, 0, 1u << k0);
    if (rc) return rc == 1 ? STWO_HIP_OK : rc;
    rc = search(0, 1u << k0, (1u << kGrindLowBits) - (1u << k0));
    if (rc) return rc == 1 ? STWO_HIP_OK : rc;
    h_start = 1;
  }
  // About 2^(pow_bits + 1) attempts per launch (at least one hi word, at most 64).
  const uint32_t hb = pow_bits + 1 > kGrindLowBits ? std::min<uint32_t>(pow_bits + 1 - kGrindLowBits, 6) : 0;
  const uint32_t lanes = 1u << (kGrindLowBits + hb);
  for (uint64_t h0 = h_start; h0 < P; h0 += (1u << hb)) {
    rc = search((uint32_t)h0, 0, lanes);
    if (rc) return rc == 1 ? STWO_HIP_OK : rc;
  }
  return set_error(ctx, STWO_HIP_EDEVICE, "Grind failed to find a solution.");
}

}  // extern "C"

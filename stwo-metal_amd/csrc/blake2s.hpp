// Blake2s-256 compression (RFC 7693) for gfx950 lanes and the host.
//
// Unkeyed, 32-byte digest: h0 = IV0 ^ 0x01010020 (SIMD/blake2s.rs:23-33);
// SIGMA schedule SIMD/blake2s.rs:37-48; `t` is the byte count including the
// current block, `last` sets f0 = 0xFFFFFFFF (SIMD/blake2s.rs:163-186).
// Messages below 4 GiB only (t >> 32 = 0), which covers every node here.
//
// Device: one hash per lane, the 16 message words and 16 state words live in
// VGPRs; rotations compile to v_alignbit_b32 and the 3-input adds to
// v_add3_u32.  No LDS, no shuffles.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>

namespace stwo_hip {

__host__ __device__ __forceinline__ uint32_t b2_rotr(uint32_t x, uint32_t n) { return (x >> n) | (x << (32 - n)); }

#define STWO_B2_IV0 0x6A09E667u
#define STWO_B2_IV1 0xBB67AE85u
#define STWO_B2_IV2 0x3C6EF372u
#define STWO_B2_IV3 0xA54FF53Au
#define STWO_B2_IV4 0x510E527Fu
#define STWO_B2_IV5 0x9B05688Cu
#define STWO_B2_IV6 0x1F83D9ABu
#define STWO_B2_IV7 0x5BE0CD19u

struct b2_state {
  uint32_t h[8];
};

__host__ __device__ __forceinline__ void b2_init(b2_state &s) {
  s.h[0] = STWO_B2_IV0 ^ 0x01010020u;
  s.h[1] = STWO_B2_IV1;
  s.h[2] = STWO_B2_IV2;
  s.h[3] = STWO_B2_IV3;
  s.h[4] = STWO_B2_IV4;
  s.h[5] = STWO_B2_IV5;
  s.h[6] = STWO_B2_IV6;
  s.h[7] = STWO_B2_IV7;
}

#define STWO_B2_G(a, b, c, d, x, y)  \
  v[a] = v[a] + v[b] + (x);          \
  v[d] = b2_rotr(v[d] ^ v[a], 16);   \
  v[c] = v[c] + v[d];                \
  v[b] = b2_rotr(v[b] ^ v[c], 12);   \
  v[a] = v[a] + v[b] + (y);          \
  v[d] = b2_rotr(v[d] ^ v[a], 8);    \
  v[c] = v[c] + v[d];                \
  v[b] = b2_rotr(v[b] ^ v[c], 7);

#define STWO_B2_ROUND(s0, s1, s2, s3, s4, s5, s6, s7, s8, s9, s10, s11, s12, s13, s14, s15) \
  STWO_B2_G(0, 4, 8, 12, m[s0], m[s1])                                                      \
  STWO_B2_G(1, 5, 9, 13, m[s2], m[s3])                                                      \
  STWO_B2_G(2, 6, 10, 14, m[s4], m[s5])                                                     \
  STWO_B2_G(3, 7, 11, 15, m[s6], m[s7])                                                     \
  STWO_B2_G(0, 5, 10, 15, m[s8], m[s9])                                                     \
  STWO_B2_G(1, 6, 11, 12, m[s10], m[s11])                                                   \
  STWO_B2_G(2, 7, 8, 13, m[s12], m[s13])                                                    \
  STWO_B2_G(3, 4, 9, 14, m[s14], m[s15])

// Compresses one 64-byte block `m` into `s`.
__host__ __device__ __forceinline__ void b2_compress(b2_state &s, const uint32_t m[16], uint32_t t, bool last) {
  uint32_t v[16];
#pragma unroll
  for (int i = 0; i < 8; i++) v[i] = s.h[i];
  v[8] = STWO_B2_IV0;
  v[9] = STWO_B2_IV1;
  v[10] = STWO_B2_IV2;
  v[11] = STWO_B2_IV3;
  v[12] = STWO_B2_IV4 ^ t;
  v[13] = STWO_B2_IV5;
  v[14] = last ? ~STWO_B2_IV6 : STWO_B2_IV6;
  v[15] = STWO_B2_IV7;
  STWO_B2_ROUND(0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15)
  STWO_B2_ROUND(14, 10, 4, 8, 9, 15, 13, 6, 1, 12, 0, 2, 11, 7, 5, 3)
  STWO_B2_ROUND(11, 8, 12, 0, 5, 2, 15, 13, 10, 14, 3, 6, 7, 1, 9, 4)
  STWO_B2_ROUND(7, 9, 3, 1, 13, 12, 11, 14, 2, 6, 5, 10, 4, 0, 15, 8)
  STWO_B2_ROUND(9, 0, 5, 7, 2, 4, 10, 15, 14, 1, 11, 12, 6, 8, 3, 13)
  STWO_B2_ROUND(2, 12, 6, 10, 0, 11, 8, 3, 4, 13, 7, 5, 15, 14, 1, 9)
  STWO_B2_ROUND(12, 5, 1, 15, 14, 13, 4, 10, 0, 7, 6, 3, 9, 2, 8, 11)
  STWO_B2_ROUND(13, 11, 7, 14, 12, 1, 3, 9, 5, 0, 15, 4, 8, 6, 2, 10)
  STWO_B2_ROUND(6, 15, 14, 9, 11, 3, 0, 8, 12, 2, 13, 7, 1, 4, 10, 5)
  STWO_B2_ROUND(10, 2, 8, 4, 7, 6, 1, 5, 15, 11, 9, 14, 3, 12, 13, 0)
#pragma unroll
  for (int i = 0; i < 8; i++) s.h[i] ^= v[i] ^ v[i + 8];
}

// ---- one compression on the 4 lanes of an aligned quad ----------------------
//
// For latency-bound hashing (the top levels of a Merkle tree, the channel):
// lane j = lane & 3 owns state column j (a = v[j], b = v[4+j], c = v[8+j],
// d = v[12+j]) and runs that column's G; for the diagonal half-round b, c, d
// are rotated by 1, 2, 3 lanes within the quad (DPP quad_perm, no LDS) so
// lane j holds G_j's operands v[j], v[4+(j+1)%4], v[8+(j+2)%4], v[12+(j+3)%4],
// then rotated back.  A lane's message words are selected from the quad's
// common message by its two lane bits.  About 420 instructions per lane
// instead of 960 on one lane.  Every lane of the quad must be active and pass
// the same m, t and last.
template <int CTRL>
__device__ __forceinline__ uint32_t b2_quad_dpp(uint32_t v) {
  return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, CTRL, 0xF, 0xF, true);
}
__device__ __forceinline__ uint32_t b2_sel4(bool j0, bool j1, uint32_t w0, uint32_t w1, uint32_t w2, uint32_t w3) {
  const uint32_t lo = j0 ? w1 : w0, hi = j0 ? w3 : w2;
  return j1 ? hi : lo;
}

#define STWO_B2_QG(x, y)          \
  a = a + b + (x);                \
  d = b2_rotr(d ^ a, 16);         \
  c = c + d;                      \
  b = b2_rotr(b ^ c, 12);         \
  a = a + b + (y);                \
  d = b2_rotr(d ^ a, 8);          \
  c = c + d;                      \
  b = b2_rotr(b ^ c, 7);

#define STWO_B2_QROUND(s0, s1, s2, s3, s4, s5, s6, s7, s8, s9, s10, s11, s12, s13, s14, s15)                \
  STWO_B2_QG(b2_sel4(j0, j1, m[s0], m[s2], m[s4], m[s6]), b2_sel4(j0, j1, m[s1], m[s3], m[s5], m[s7]))      \
  b = b2_quad_dpp<0x39>(b);                                                                                  \
  c = b2_quad_dpp<0x4E>(c);                                                                                  \
  d = b2_quad_dpp<0x93>(d);                                                                                  \
  STWO_B2_QG(b2_sel4(j0, j1, m[s8], m[s10], m[s12], m[s14]), b2_sel4(j0, j1, m[s9], m[s11], m[s13], m[s15])) \
  b = b2_quad_dpp<0x93>(b);                                                                                  \
  c = b2_quad_dpp<0x4E>(c);                                                                                  \
  d = b2_quad_dpp<0x39>(d);

// h_lo / h_hi: chain words j and 4 + j of this lane (j = lane & 3), updated
// in place.
__device__ __forceinline__ void b2_compress_quad(uint32_t &h_lo, uint32_t &h_hi, const uint32_t m[16], uint32_t t,
                                                 bool last) {
  const uint32_t j = __lane_id() & 3;
  const bool j0 = j & 1, j1 = j & 2;
  const uint32_t iv_lo = b2_sel4(j0, j1, STWO_B2_IV0, STWO_B2_IV1, STWO_B2_IV2, STWO_B2_IV3);
  const uint32_t iv_hi = b2_sel4(j0, j1, STWO_B2_IV4, STWO_B2_IV5, STWO_B2_IV6, STWO_B2_IV7);
  uint32_t a = h_lo, b = h_hi, c = iv_lo;
  uint32_t d = iv_hi ^ (j == 0 ? t : 0u) ^ ((j == 2 && last) ? 0xFFFFFFFFu : 0u);
  STWO_B2_QROUND(0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15)
  STWO_B2_QROUND(14, 10, 4, 8, 9, 15, 13, 6, 1, 12, 0, 2, 11, 7, 5, 3)
  STWO_B2_QROUND(11, 8, 12, 0, 5, 2, 15, 13, 10, 14, 3, 6, 7, 1, 9, 4)
  STWO_B2_QROUND(7, 9, 3, 1, 13, 12, 11, 14, 2, 6, 5, 10, 4, 0, 15, 8)
  STWO_B2_QROUND(9, 0, 5, 7, 2, 4, 10, 15, 14, 1, 11, 12, 6, 8, 3, 13)
  STWO_B2_QROUND(2, 12, 6, 10, 0, 11, 8, 3, 4, 13, 7, 5, 15, 14, 1, 9)
  STWO_B2_QROUND(12, 5, 1, 15, 14, 13, 4, 10, 0, 7, 6, 3, 9, 2, 8, 11)
  STWO_B2_QROUND(13, 11, 7, 14, 12, 1, 3, 9, 5, 0, 15, 4, 8, 6, 2, 10)
  STWO_B2_QROUND(6, 15, 14, 9, 11, 3, 0, 8, 12, 2, 13, 7, 1, 4, 10, 5)
  STWO_B2_QROUND(10, 2, 8, 4, 7, 6, 1, 5, 15, 11, 9, 14, 3, 12, 13, 0)
  h_lo ^= a ^ c;
  h_hi ^= b ^ d;
}

// Initial chain words j and 4 + j of an unkeyed 32-byte Blake2s (b2_init).
__device__ __forceinline__ uint32_t b2_iv_lo(uint32_t j) {
  return j == 0 ? (STWO_B2_IV0 ^ 0x01010020u) : j == 1 ? STWO_B2_IV1 : j == 2 ? STWO_B2_IV2 : STWO_B2_IV3;
}
__device__ __forceinline__ uint32_t b2_iv_hi(uint32_t j) {
  return j == 0 ? STWO_B2_IV4 : j == 1 ? STWO_B2_IV5 : j == 2 ? STWO_B2_IV6 : STWO_B2_IV7;
}

// Word k (0..7) of the quad's chain value, in every lane of the quad.
__device__ __forceinline__ void b2_quad_gather(uint32_t h_lo, uint32_t h_hi, uint32_t out[8]) {
  out[0] = b2_quad_dpp<0x00>(h_lo);
  out[1] = b2_quad_dpp<0x55>(h_lo);
  out[2] = b2_quad_dpp<0xAA>(h_lo);
  out[3] = b2_quad_dpp<0xFF>(h_lo);
  out[4] = b2_quad_dpp<0x00>(h_hi);
  out[5] = b2_quad_dpp<0x55>(h_hi);
  out[6] = b2_quad_dpp<0xAA>(h_hi);
  out[7] = b2_quad_dpp<0xFF>(h_hi);
}

// reduce_to_m31 on each output word (STWO/core/vcs/blake2_hash.rs:111-119):
// M31::reduce of a u32 = v mod P computed as ((v >> 31) + v + 1 >> 31) + v & P.
__host__ __device__ __forceinline__ uint32_t b2_reduce_word(uint32_t w) {
  uint64_t v = w;
  return (uint32_t)(((((v >> 31) + v + 1) >> 31) + v) & 0x7fffffffull);
}

// Host-side streaming hasher (same buffering as the `blake2` crate: the last
// block, full or partial, is the finalized one).
struct b2_host {
  b2_state s;
  uint8_t buf[64];
  uint32_t buflen = 0;
  uint32_t t = 0;
  b2_host() { b2_init(s); }
  void update(const void *data, size_t len) {
    const uint8_t *d = (const uint8_t *)data;
    while (len) {
      if (buflen == 64) {
        uint32_t m[16];
        memcpy(m, buf, 64);
        t += 64;
        b2_compress(s, m, t, false);
        buflen = 0;
      }
      size_t take = 64 - buflen;
      if (take > len) take = len;
      memcpy(buf + buflen, d, take);
      buflen += (uint32_t)take;
      d += take;
      len -= take;
    }
  }
  void finalize(uint8_t out[32]) {
    t += buflen;
    memset(buf + buflen, 0, 64 - buflen);
    uint32_t m[16];
    memcpy(m, buf, 64);
    b2_compress(s, m, t, true);
    memcpy(out, s.h, 32);
  }
};

}  // namespace stwo_hip

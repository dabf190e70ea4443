// Host-side Blake2s Fiat-Shamir channel (Blake2sChannel,
// STWO/core/channel/blake2s.rs), shared by the FRI layer loop and the prover.
#pragma once
#include <stdint.h>
#include <string.h>

#include "blake2s.hpp"
#include "field.hpp"

namespace stwo_hip {

// POW_PREFIX of the proof-of-work hash (STWO/core/channel/blake2s.rs:122-139).
constexpr uint32_t kPowPrefix = 0x12345678u;

struct HostChannel {
  uint8_t digest[32];
  uint32_t n_draws;

  void update(const uint8_t d[32]) {
    memcpy(digest, d, 32);
    n_draws = 0;
  }
  // MerkleChannel::mix_root: digest = H(digest || root), vcs_lifted/blake2_merkle.rs:49-55.
  void mix_root(const uint32_t root[8]) {
    b2_host h;
    h.update(digest, 32);
    h.update(root, 32);
    uint8_t o[32];
    h.finalize(o);
    update(o);
  }
  // mix_felts, blake2s.rs:58-68.
  void mix_felts(const uint32_t *felts4, size_t n) {
    b2_host h;
    h.update(digest, 32);
    h.update(felts4, 16 * n);
    uint8_t o[32];
    h.finalize(o);
    update(o);
  }
  // mix_u32s: digest = H(digest || le32(w0) || le32(w1) ...), blake2s.rs:71-79.
  void mix_u32s(const uint32_t *words, size_t n) {
    b2_host h;
    h.update(digest, 32);
    h.update(words, 4 * n);
    uint8_t o[32];
    h.finalize(o);
    update(o);
  }
  // mix_u64 = mix_u32s([lo, hi]), blake2s.rs:81-83.
  void mix_u64(uint64_t v) {
    const uint32_t w[2] = {(uint32_t)v, (uint32_t)(v >> 32)};
    mix_u32s(w, 2);
  }
  // draw_u32s: H(digest || le32(n_draws) || 0), blake2s.rs:103-117.
  void draw_u32s(uint32_t out[8]) {
    uint8_t in[37];
    memcpy(in, digest, 32);
    memcpy(in + 32, &n_draws, 4);
    in[36] = 0;
    n_draws++;
    b2_host h;
    h.update(in, 37);
    h.finalize((uint8_t *)out);
  }
  // draw_secure_felt via draw_base_felts (retry unless all words < 2P), blake2s.rs:36-51,85-88.
  qm31 draw_secure_felt() {
    uint32_t w[8];
    for (;;) {
      draw_u32s(w);
      bool ok = true;
      for (int i = 0; i < 8; i++) ok = ok && w[i] < 2 * P;
      if (ok) break;
    }
    return qm31_make(b2_reduce_word(w[0]), b2_reduce_word(w[1]), b2_reduce_word(w[2]), b2_reduce_word(w[3]));
  }
};

}  // namespace stwo_hip

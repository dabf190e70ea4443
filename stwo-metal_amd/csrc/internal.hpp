// Internal interfaces shared between translation units of libstwo_hip (not
// part of the C ABI): FFT pass descriptors and the launchers the batched
// path reuses with tables it uploads itself.
#pragma once
#include <stdint.h>

#include <vector>

#include "common.hpp"

namespace stwo_hip {

struct FftCol {
  const uint32_t *in;
  uint32_t *out;
};

constexpr uint32_t kMaxStages = 4;

struct FftStage {
  uint32_t b0, r_lo, r_hi;
};

// One pass of the radix-16 FFT over a group of same-shape columns.
struct FftPass2 {
  const FftCol *cols;
  uint32_t n_cols;
  const uint32_t *tw_end;
  uint32_t N, T, lo, w, m;
  uint32_t n_stages;
  FftStage st[kMaxStages];
  uint32_t in_log;
  uint32_t scale;
  uint32_t rep_log;  // output blocks per tile (first RFFT pass of a blowup)
};

struct PassLaunch {
  FftPass2 p;
  uint32_t wgs;
  uint32_t wg_log;  // log2 lanes per workgroup (8 or 9)
};

int make_group_passes(stwo_hip_ctx *ctx, uint32_t n, uint32_t N, bool inverse, const FftCol *d_tab,
                      const FftCol *d_tab_inplace, uint32_t n_cols, const uint32_t *tw_end,
                      std::vector<PassLaunch> &out);
int launch_fft_pass(stwo_hip_ctx *ctx, bool inverse, const FftPass2 &p, uint32_t wgs, uint32_t wg_log);
int launch_fft_multi(stwo_hip_ctx *ctx, bool inverse, const FftPass2 *d_groups, const uint32_t *d_wg_start,
                     uint32_t n_groups, uint32_t total_wgs, uint32_t wg_log);

// Merkle column reference (pointer + lifting shift).
struct ColRef {
  const uint32_t *p;
  uint32_t shift;  // lifted: L - log_size(col); mixed: unused
  uint32_t pad;
};

// One Merkle layer over a device table of columns (mixed layer or lifted leaves).
int launch_merkle_layer(stwo_hip_ctx *ctx, const uint32_t *prev, const ColRef *d_cols, uint32_t n_cols,
                        uint32_t log_size, uint32_t *out, int mode, bool lifted);
// Node layers [stop_log, top_log) of a packed tree whose layer top_log is filled.
int hash_node_layers(stwo_hip_ctx *ctx, uint32_t *layers, uint32_t top_log, uint32_t stop_log, int mode);

inline size_t tree_layer_off(uint32_t l) { return (((size_t)1 << l) - 1) * 8; }

}  // namespace stwo_hip

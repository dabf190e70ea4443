// Internal interfaces shared between translation units of libstwo_hip (not
// part of the C ABI): FFT pass descriptors and the launchers the batched
// path reuses with tables it uploads itself.
#pragma once
#include <hip/hip_runtime.h>
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
  // Bit 0: the first stage's layout (b0 = 0, 16 consecutive words per lane)
  // is loaded straight from HBM with 128-bit loads; bit 1: the last stage's
  // layout (b0 = 0) is stored the same way.  Set only for contiguous tiles
  // (w = 0) of 16-byte aligned columns.
  uint32_t vec_io;
  // The compile-time schedule the pass matches (set by make_group_passes):
  // 1 = the 12-layer pass of 3 full stages (wg8), 2 = the 8-layer strided
  // pass of 2 full stages on a 2^13 tile (wg9), 3 / 4 / 5 = strided passes
  // of 5 / 6 / 7 layers (a full stage and a partial one; wg8, wg8, wg9),
  // 6..12 = single-stage passes of 1 to 4 strided layers on a 2^10 tile
  // (wg8), 0 = none (run-time schedule).
  uint32_t fixed;
};

struct PassLaunch {
  FftPass2 p;
  uint32_t wgs;
  uint32_t wg_log;  // log2 lanes per workgroup (8 or 9)
};

// `aligned16`: every column pointer of the group is 16-byte aligned.
// `tune`: the planner choices to use; NULL = the context's choice for the
// shape (stwo_hip_fft_autotune), else the defaults.
int make_group_passes(stwo_hip_ctx *ctx, uint32_t n, uint32_t N, bool inverse, const FftCol *d_tab,
                      const FftCol *d_tab_inplace, uint32_t n_cols, const uint32_t *tw_end, bool aligned16,
                      std::vector<PassLaunch> &out, const FftTune *tune = nullptr);
int launch_fft_pass(stwo_hip_ctx *ctx, bool inverse, const FftPass2 &p, uint32_t wgs, uint32_t wg_log);
// `kind`: the schedule every group of the launch matches: 1 to 5 (FftPass2
// fixed, with its workgroup size), -1 (tiles of 2^10 or more: wave-uniform
// columns) or 0 (any).
int launch_fft_multi(stwo_hip_ctx *ctx, bool inverse, const FftPass2 *d_groups, const uint32_t *d_wg_start,
                     uint32_t n_groups, uint32_t total_wgs, uint32_t wg_log, int kind);

// The IFFT's last pass and the RFFT's first pass of one column group run as
// one fused kernel when their tiles match (lde_fusable); kind 2 = the
// compile-time 8-layer pair, 0 = run-time stage lists.
bool lde_fusable(const PassLaunch &last_inv, const PassLaunch &first_fwd, uint32_t log_blowup);
int lde_kind(const PassLaunch &last_inv, const PassLaunch &first_fwd);
int launch_lde_multi(stwo_hip_ctx *ctx, const FftPass2 *d_pis, const FftPass2 *d_pfs, const uint32_t *d_wg_start,
                     uint32_t n_groups, uint32_t total_wgs, uint32_t wg_log, int kind, uint32_t store_coeffs);

// Upper bound on the device bytes stwo_hip_wide_fib_prove takes for itself
// (its buffers, the FRI prover's memory, OODS scratch, a margin for tables).
size_t wide_fib_prove_device_bytes(uint32_t n_cols, uint32_t log_n, const stwo_hip_pcs_config &config);

// One column set of a batched plan (stwo_hip_batch_create's arguments).
// A column set of a batch plan.  stage: what the plan does to its columns --
// BATCH_LDE interpolates them in place and evaluates them at the blowup into
// outs; BATCH_IFFT only interpolates them in place (no outs, no tree);
// BATCH_RFFT takes them as coefficients and evaluates them into outs.  Every
// set's IFFT passes run before any set's RFFT pass, so a BATCH_RFFT set may
// read what a BATCH_IFFT set of the same plan writes.  layers: the set's
// tree (required for BATCH_LDE, optional for BATCH_RFFT).
enum BatchStage { BATCH_LDE = 0, BATCH_IFFT = 1, BATCH_RFFT = 2 };
struct BatchTree {
  uint32_t *const *cols;
  const uint32_t *log_sizes;
  size_t n_cols;
  uint32_t log_blowup;
  uint32_t *const *outs;
  uint32_t *layers;
  int tree_kind, hash_mode;
  int stage = BATCH_LDE;
};
int batch_create_trees(stwo_hip_ctx *ctx, const BatchTree *trees, size_t n_trees, const stwo_hip_twiddles *tw,
                       int flags, stwo_hip_batch **out);

// stwo_hip_wide_fib_constraint_quotients, writing the quotients over accum
// instead of adding them when !accumulate (a single component needs no zeroed
// accumulator); the random coefficient from the host (random_coeff) or
// already on the device (alpha_dev, 4 words, random_coeff NULL).
int wide_fib_constraint_quotients_impl(stwo_hip_ctx *ctx, const uint32_t *const *cols, uint32_t n_cols,
                                       uint32_t log_n, const uint32_t random_coeff[4], uint32_t *const accum[4],
                                       bool accumulate, const uint32_t *alpha_dev);

// ---- the prove's Fiat-Shamir steps on the device ---------------------------

// One wave: chan (10 words: digest, n_draws, error flag; first set to
// [init_digest, 0, 0] when init_digest is not NULL) mixes the Merkle root at
// `root` (copied to root_out) and draws a secure felt into draw_out.  With
// point_out, the draw is turned into the OODS point: point_out gets
// kOodsFactorWords of eval_at_point factors (g_0 = y, g_1 = x,
// g_k = pi^(k-1)(x) for k <= point_log, zeros up to k = 31), then x, y and
// pi^(point_log - 1)(x).  Columns sampled with these factors must have log
// sizes of at most point_log.
constexpr uint32_t kOodsFactorWords = 128;
constexpr uint32_t kOodsPointWords = kOodsFactorWords + 12;
int prove_channel_step(stwo_hip_ctx *ctx, uint32_t *chan, const uint8_t *init_digest, const uint32_t *root,
                       uint32_t *root_out, uint32_t *draw_out, uint32_t *point_out, uint32_t point_log);

// stwo_hip_eval_at_point with the point's factors on the device (the first
// kOodsFactorWords words of prove_channel_step's point_out) and the values
// left on the device (4 words per column).
int eval_at_point_dev(stwo_hip_ctx *ctx, const uint32_t *const *coeffs, const uint32_t *log_sizes, size_t n_cols,
                      const uint32_t *d_factors, uint32_t *d_values);

// stwo_hip_barycentric_weights with the point on the device (x, y and
// pi^(log_size - 1)(x), 12 words), and stwo_hip_barycentric_eval_at_point
// with the values left on the device.
int barycentric_weights_dev(stwo_hip_ctx *ctx, uint32_t log_size, const uint32_t *d_point, uint32_t *const out[4]);
int barycentric_eval_dev(stwo_hip_ctx *ctx, const uint32_t *const *evals, size_t n_cols, uint32_t log_size,
                         const uint32_t *const weights[4], uint32_t *d_values);

// ---- decommitment: batched device gather -----------------------------------

// Collects (device source, word count) reads; run() gathers them all with
// one kernel launch into `out` (host), in the order they were added.
struct Gatherer {
  struct Op {
    uintptr_t src;
    uint32_t n, dst;
  };
  std::vector<Op> ops_;
  size_t words = 0;
  std::vector<uint32_t> out;
  // Returns the word offset of this read in `out`.
  size_t add(const uint32_t *src, uint32_t n);
  void reserve(size_t more_ops);
  int run(stwo_hip_ctx *ctx);
};

// Hash-witness traversal of a lifted Merkle decommitment: one entry per
// visited parent, from the leaf layer up.
struct MerkleDecommitPlan {
  struct Parent {
    uint32_t child_layer;  // log size of the layer holding the two children
    uint32_t parent;       // parent index (children 2 * parent, 2 * parent + 1)
    int witness_child;     // child (0 or 1) that is a witness, or -1
    size_t off;            // gather offset of the two children (16 words)
  };
  std::vector<Parent> parents;
};
void plan_merkle_decommit(Gatherer &g, const uint32_t *tree, uint32_t L, const std::vector<uint32_t> &queries,
                          MerkleDecommitPlan &plan);
// witness: 8 words per hash; aux (optional): 10 words per node
// [layer log size, index, hash].
void finish_merkle_decommit(const Gatherer &g, const MerkleDecommitPlan &plan, std::vector<uint32_t> &witness,
                            std::vector<uint32_t> *aux);

// ---- FRI prover state kept for decommitment -------------------------------

struct FriLayerShape {
  uint32_t log_size;   // log size of the layer's evaluation
  uint32_t fold_step;  // folds from this layer to the next
};
std::vector<FriLayerShape> fri_schedule(const stwo_hip_fri_config &cfg, uint32_t log_size);

struct FriLayerKeep {
  FriLayerShape shape;
  const uint32_t *cols[4];  // coordinate columns of the committed evaluation
  uint32_t *tree;           // packed lifted Merkle tree of the 4 columns
  uint32_t root[8];
};

// One committed layer's decommitment (FriLayerProof + FriLayerProofAux).
struct FriLayerDecommit {
  std::vector<uint32_t> positions;     // decommitment positions (sorted)
  std::vector<uint32_t> values;        // 4 words per position
  std::vector<uint32_t> fri_witness;   // 4 words per witness value
  std::vector<uint32_t> hash_witness;  // 8 words per hash
};

}  // namespace stwo_hip

struct stwo_hip_fri_prover {
  int device = 0;
  stwo_hip_fri_config cfg{};
  uint32_t log_size = 0;
  uint32_t *mem = nullptr;  // trees and folded evaluations, one allocation
  size_t mem_words = 0;
  bool owns_mem = true;     // false: context scratch memory
  std::vector<stwo_hip::FriLayerKeep> layers;  // committed layers, first (circle) layer first
  std::vector<uint32_t *> evals;               // output of fold f (4 columns of 2^log each)
  std::vector<uint32_t> last_layer_poly;       // 4 words per coefficient
  uint32_t last_log = 0;
  uint32_t *last_layer_eval[4] = {nullptr, nullptr, nullptr, nullptr};
  bool decommitted = false;
  std::vector<uint32_t> query_positions;  // sorted, deduplicated
  std::vector<uint32_t> unsorted_queries;
  std::vector<stwo_hip::FriLayerDecommit> proofs;
};

namespace stwo_hip {

// stwo_hip_fri_prover_commit whose memory is the context's scratch buffer
// `scratch_key` (kept by the context, reused by the next call with the same
// key) when scratch_key is not NULL.  One such prover per key at a time.
// With pow_bits >= 0 (the prove), the commit's end (last layer, its mix
// into the channel) and the proof of work's first search range may run on
// the device before the one synchronization: *nonce_found then says whether
// *nonce holds the grind's result (otherwise stwo_hip_grind finds it).
int fri_prover_commit_mem(stwo_hip_ctx *ctx, uint8_t channel_digest[32], uint32_t *channel_n_draws,
                          stwo_hip_fri_config cfg, const uint32_t *const column[4], uint32_t log_size,
                          const stwo_hip_twiddles *tw, uint32_t *roots, uint32_t *n_roots, uint32_t *last_layer_poly,
                          const char *scratch_key, stwo_hip_fri_prover **out, int pow_bits = -1,
                          uint64_t *nonce = nullptr, bool *nonce_found = nullptr);

// The proof of work's first search range (the lowest 2^(pow_bits + 6) lo
// words of hi = 0) with the prefix H(POW_PREFIX || 0^12 || digest ||
// pow_bits) already on the device and *best reset to all ones: the smallest
// valid nonce of the range is left in *best (all ones if none).  Requires
// pow_bits + 6 < 20.
int grind_first_range(stwo_hip_ctx *ctx, const uint32_t *d_prefixed, uint32_t pow_bits, unsigned long long *best);

// stwo_hip_fri_prover_decommit in two halves around a caller's gather, so
// that other reads (the trace trees' decommitments) share its launch and
// synchronisation: fri_decommit_plan draws the queries from the channel (the
// digest and draw count are updated) and adds every FRI read to g;
// fri_decommit_finish, after g.run, fills the prover's layer proofs.
struct FriDecommitPlan {
  struct Layer {
    std::vector<uint32_t> positions;
    std::vector<size_t> val_off;   // gather offset of each coordinate read (4 per position)
    std::vector<uint8_t> witness;  // 1 if the position is not a query
    MerkleDecommitPlan merkle;
  };
  std::vector<Layer> layers;
  std::vector<uint32_t> queries, raw;  // sorted and deduplicated; as drawn
};
int fri_decommit_plan(stwo_hip_ctx *ctx, stwo_hip_fri_prover *pr, uint8_t channel_digest[32],
                      uint32_t *channel_n_draws, Gatherer &g, FriDecommitPlan &plan);
void fri_decommit_finish(stwo_hip_fri_prover *pr, const Gatherer &g, FriDecommitPlan &plan);

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

__host__ __device__ inline size_t tree_layer_off(uint32_t l) { return (((size_t)1 << l) - 1) * 8; }

}  // namespace stwo_hip

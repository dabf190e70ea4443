// FRI folds (FriOps) and the FRI commitment driver (FriProver::commit).
//
// Fold factors come from the resident inverse twiddle tree, as SimdBackend
// does (SIMD/fri.rs:32-172): for a circle domain of log size N the factor of
// pair i is 1/p.y with p = domain.at(bitrev(2i, N)) -- the inverse circle
// twiddle of pair i -- and for a line domain of log size n the factor of
// pair i is 1/x with x = domain.at(bitrev(2i, n)) -- entry i of the inverse
// tree level of that domain, itw[len - 2^n + i].  QM31 columns are 4
// coordinate columns (SoA), so each lane's loads are coalesced.
//
// fold_line folds up to 4 steps per launch with the 2^step inputs of an
// output held in VGPRs (alpha, alpha^2, ... per step, SIMD/fri.rs:52-111).
#include <string.h>

#include <algorithm>

#include <vector>

#include "blake2s.hpp"
#include "channel.hpp"
#include "common.hpp"
#include "field.hpp"
#include "internal.hpp"
#include "merkle_dev.hpp"

namespace stwo_hip {

struct Qm31Cols {
  uint32_t *c[4];
};
struct Qm31ColsConst {
  const uint32_t *c[4];
};

__device__ __forceinline__ uint32_t circle_itwiddle(const uint32_t *itw_end, uint32_t N, uint32_t h) {
  const uint32_t *l0 = itw_end - ((int64_t)1 << (N - 1));
  uint32_t x = l0[2 * (h >> 2)], y = l0[2 * (h >> 2) + 1];
  uint32_t v = (h & 2) ? x : y;
  bool neg = ((h & 3) == 1) || ((h & 3) == 2);
  return neg ? m31_neg(v) : v;
}

__device__ __forceinline__ qm31 load_qm31(const uint32_t *p) { return qm31_make(p[0], p[1], p[2], p[3]); }

// Per-output folds (shared by the standalone fold kernels, the fused FRI
// layer kernel and the small-layer tail).  Circle: output i of the fold of a
// circle evaluation of log size N (f0 = f[2i] + f[2i+1], f1 = (f[2i] -
// f[2i+1]) / p.y, f0 + alpha f1).
__device__ __forceinline__ qm31 fold_circle_at(const Qm31ColsConst &src, const uint32_t *itw_end, uint32_t N,
                                               const qm31 &alpha, uint32_t i) {
  const uint32_t iy = circle_itwiddle(itw_end, N, i);
  qm31 f0, f1;
#pragma unroll
  for (int k = 0; k < 4; k++) {
    uint2 v = reinterpret_cast<const uint2 *>(src.c[k])[i];
    f0.v[k] = m31_add(v.x, v.y);
    f1.v[k] = m31_mul(m31_sub(v.x, v.y), iy);
  }
  return qm31_add(f0, qm31_mul(alpha, f1));
}

// Line: output i of STEP successive folds of a line evaluation of log size n
// (the 2^STEP inputs of the output in VGPRs; alpha, alpha^2, ... per fold).
template <int STEP>
__device__ __forceinline__ qm31 fold_line_at(const Qm31ColsConst &src, const uint32_t *itw_end, uint32_t n,
                                             qm31 a, uint32_t i) {
  constexpr int W = 1 << STEP;
  qm31 v[W];
#pragma unroll
  for (int k = 0; k < 4; k++) {
    const uint32_t *s = src.c[k] + (size_t)i * W;
#pragma unroll
    for (int j = 0; j < W; j++) v[j].v[k] = s[j];
  }
#pragma unroll
  for (int layer = 0; layer < STEP; layer++) {
    const int half = W >> (layer + 1);
    const uint32_t *lvl = itw_end - ((int64_t)1 << (n - layer));
#pragma unroll
    for (int j = 0; j < half; j++) {
      uint32_t xi = lvl[(size_t)i * half + j];
      qm31 f0, f1;
#pragma unroll
      for (int k = 0; k < 4; k++) {
        f0.v[k] = m31_add(v[2 * j].v[k], v[2 * j + 1].v[k]);
        f1.v[k] = m31_mul(m31_sub(v[2 * j].v[k], v[2 * j + 1].v[k]), xi);
      }
      v[j] = qm31_add(f0, qm31_mul(a, f1));
    }
    a = qm31_mul(a, a);
  }
  return v[0];
}

// `alpha_dev`, when not NULL, holds alpha in device memory (written by the
// device channel) and replaces the by-value alpha.
__global__ void __launch_bounds__(256) fold_circle_kernel(Qm31Cols dst, Qm31ColsConst src, const uint32_t *itw_end,
                                                          uint32_t N, qm31 alpha, qm31 alpha_sq, uint32_t accumulate,
                                                          const uint32_t *alpha_dev) {
  const uint32_t n_out = 1u << (N - 1);
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_out) return;
  if (alpha_dev) {
    alpha = load_qm31(alpha_dev);
    alpha_sq = qm31_mul(alpha, alpha);
  }
  qm31 r = fold_circle_at(src, itw_end, N, alpha, i);
  if (accumulate) {
    qm31 d = qm31_make(dst.c[0][i], dst.c[1][i], dst.c[2][i], dst.c[3][i]);
    r = qm31_add(qm31_mul(d, alpha_sq), r);
  }
#pragma unroll
  for (int k = 0; k < 4; k++) dst.c[k][i] = r.v[k];
}

template <int STEP>
__global__ void __launch_bounds__(256) fold_line_kernel(Qm31Cols dst, Qm31ColsConst src, const uint32_t *itw_end,
                                                        uint32_t n, qm31 alpha, const uint32_t *alpha_dev) {
  const uint32_t n_out = 1u << (n - STEP);
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_out) return;
  if (alpha_dev) alpha = load_qm31(alpha_dev);
  const qm31 r = fold_line_at<STEP>(src, itw_end, n, alpha, i);
#pragma unroll
  for (int k = 0; k < 4; k++) dst.c[k][i] = r.v[k];
}

// out = in^(2^k) (alpha for the folds after the first 4 of a longer fold step).
__global__ void alpha_pow2k_kernel(const uint32_t *in, uint32_t *out, uint32_t k) {
  if (threadIdx.x != 0) return;
  qm31 a = load_qm31(in);
  for (uint32_t i = 0; i < k; i++) a = qm31_mul(a, a);
  for (int j = 0; j < 4; j++) out[j] = a.v[j];
}

// One Fiat-Shamir step of the FRI layer loop on the device, for the
// Blake2sChannel (STWO/core/channel/blake2s.rs): digest = H(digest || root)
// (MerkleChannel::mix_root, STWO/core/vcs_lifted/blake2_merkle.rs:49-55),
// n_draws = 0, then alpha = draw_secure_felt(): draw H(digest || le32(n_draws)
// || 0x00) until all 8 words are < 2P (blake2s.rs:36-51,103-117) and reduce
// the first 4.  chan = [digest 8 words, n_draws, error flag].  The root is
// also copied to root_out.  The retry loop is bounded (each try fails with
// probability about 2^-28); running out sets the error flag.
// Run by the 4 lanes of one quad (lanes 0-3 of a wave, all active), each
// compression shared by the quad (b2_compress_quad).
__device__ void channel_step(uint32_t *chan, const uint32_t *root, uint32_t *root_out, uint32_t *alpha_out) {
  const uint32_t j = __lane_id() & 3;
  const uint32_t init_lo = b2_iv_lo(j), init_hi = b2_iv_hi(j);
  uint32_t m[16];
  for (int k = 0; k < 8; k++) {
    m[k] = chan[k];
    m[8 + k] = root[k];
  }
  root_out[j] = root[j];
  root_out[4 + j] = root[4 + j];
  uint32_t lo = init_lo, hi = init_hi;
  b2_compress_quad(lo, hi, m, 64, true);
  uint32_t digest[8];
  b2_quad_gather(lo, hi, digest);
  uint32_t n_draws = 0;
  bool ok = false;
  for (int tries = 0; tries < 64 && !ok; tries++) {
    for (int k = 0; k < 8; k++) m[k] = digest[k];
    m[8] = n_draws;
    for (int k = 9; k < 16; k++) m[k] = 0;  // the 0x00 separator byte and zero padding
    lo = init_lo;
    hi = init_hi;
    b2_compress_quad(lo, hi, m, 37, true);
    n_draws++;
    const uint64_t bal = __ballot(lo < 2u * P && hi < 2u * P);
    ok = ((bal >> (__lane_id() & ~3u)) & 0xF) == 0xF;
  }
  chan[j] = digest[j];
  chan[4 + j] = digest[4 + j];
  if (j == 0) {
    chan[8] = n_draws;
    if (!ok) chan[9] = 1;
  }
  alpha_out[j] = b2_reduce_word(lo);
}

// ---- the small layers in one launch ---------------------------------------
//
// Below kTailLog the per-layer work (a few thousand hashes) is far shorter
// than the launches it takes (leaves, node layers, channel step, fold), so
// one workgroup runs every remaining inner layer: lifted leaves of the 4
// coordinate columns, the node layers, the channel step, then the line fold
// into the next layer's evaluation, with a workgroup barrier between phases
// (all of a workgroup's waves share one CU's L1, so its global writes are
// visible to the workgroup after __syncthreads).  It writes the same trees,
// roots, channel state and folded evaluations as the per-layer launches.
constexpr uint32_t kTailLog = 10;
constexpr uint32_t kTailMax = 16;
constexpr uint32_t kTailThreads = 1024;

struct FriTailLayer {
  const uint32_t *src[4];  // the layer's committed evaluation (log lg)
  uint32_t *tree;          // its packed lifted tree
  uint32_t *dst[4];        // the fold's output (log lg - step)
  uint32_t lg, step, root_slot, pad;
};
struct FriTail {
  FriTailLayer l[kTailMax];
  uint32_t n;
  const uint32_t *itw_end;
  uint32_t *chan, *roots, *alpha;
};

// Output i of the layer's fold (a run-time step of 1..4 at the call site).
template <int STEP>
__device__ __forceinline__ void tail_fold(const FriTailLayer &L, const uint32_t *itw_end, qm31 alpha, uint32_t i) {
  Qm31ColsConst src;
#pragma unroll
  for (int k = 0; k < 4; k++) src.c[k] = L.src[k];
  const qm31 r = fold_line_at<STEP>(src, itw_end, L.lg, alpha, i);
#pragma unroll
  for (int k = 0; k < 4; k++) L.dst[k][i] = r.v[k];
}

// Node levels top_log - 1 .. 0 of a packed tree, by one workgroup of
// kTailThreads lanes: H(left || right), a quad per hash on levels of at most
// kTailThreads / 4 hashes (their latency is the critical path), a lane per
// hash above.  Every level is written to the tree and kept in LDS (lv[0] and
// lv[1] in turn, at most kTailThreads digests each), and the next level reads
// it from there: the global stores are never waited for.  Level top_log is
// read from lv[0] when `in_lds` (written there by the caller), else from the
// tree.  Starts with a barrier; returns the root (LDS, or the tree when
// top_log = 0 and not in_lds) after a closing barrier.
__device__ __forceinline__ const uint32_t *tree_levels_to_root(uint32_t *tree, uint32_t top_log,
                                                               uint32_t (&lv)[2][kTailThreads * 8], bool in_lds) {
  const uint32_t tid = threadIdx.x, q = tid >> 2, j = tid & 3;
  uint32_t c = 0;  // lv[c] holds the level below the one being hashed
  auto level = [&](const uint32_t *child, uint32_t l, uint32_t *keep) {
    uint32_t *out = tree + tree_layer_off(l);
    if ((1u << l) <= kTailThreads / 4) {
      if (q < (1u << l)) {
        uint32_t m[16];
#pragma unroll
        for (int k = 0; k < 16; k++) m[k] = child[16 * (size_t)q + k];
        uint32_t lo = b2_iv_lo(j), hi = b2_iv_hi(j);
        b2_compress_quad(lo, hi, m, 64, true);
        out[8 * (size_t)q + j] = lo;
        out[8 * (size_t)q + 4 + j] = hi;
        keep[8 * q + j] = lo;
        keep[8 * q + 4 + j] = hi;
      }
      return;
    }
    for (uint32_t i = tid; i < (1u << l); i += kTailThreads) {
      uint32_t m[16];
#pragma unroll
      for (int k = 0; k < 16; k++) m[k] = child[16 * (size_t)i + k];
      b2_state s;
      b2_init(s);
      b2_compress(s, m, 64, true);
#pragma unroll
      for (int k = 0; k < 8; k++) {
        out[8 * (size_t)i + k] = s.h[k];
        keep[8 * i + k] = s.h[k];
      }
    }
  };
  for (uint32_t l = top_log; l-- > 0;) {
    __syncthreads();
    if (!in_lds && l + 1 == top_log) {
      level(tree + tree_layer_off(l + 1), l, lv[0]);
      continue;  // the level is now in lv[0] = lv[c]
    }
    level(lv[c], l, lv[c ^ 1]);
    c ^= 1;
  }
  __syncthreads();
  return (top_log == 0 && !in_lds) ? tree : lv[c];
}

__global__ void __launch_bounds__(kTailThreads) fri_tail_kernel(FriTail t) {
  __shared__ uint32_t lv[2][kTailThreads * 8];
  const uint32_t tid = threadIdx.x;
  for (uint32_t li = 0; li < t.n; li++) {
    const FriTailLayer &L = t.l[li];
    // lifted leaves: H(le32 of the 4 coordinates), one 16-byte block
    // Levels of at most kTailThreads / 4 hashes use a quad per hash (their
    // latency is the layer's critical path), larger ones a lane per hash.
    const uint32_t n = 1u << L.lg;
    const uint32_t q = tid >> 2, j = tid & 3;
    uint32_t *leaves = L.tree + tree_layer_off(L.lg);
    if (n <= kTailThreads / 4) {
      if (q < n) {
        uint32_t m[16];
#pragma unroll
        for (int k = 0; k < 4; k++) m[k] = L.src[k][q];
#pragma unroll
        for (int k = 4; k < 16; k++) m[k] = 0;
        uint32_t lo = b2_iv_lo(j), hi = b2_iv_hi(j);
        b2_compress_quad(lo, hi, m, 16, true);
        leaves[8 * (size_t)q + j] = lo;
        leaves[8 * (size_t)q + 4 + j] = hi;
        lv[0][8 * q + j] = lo;
        lv[0][8 * q + 4 + j] = hi;
      }
    } else {
      for (uint32_t i = tid; i < n; i += kTailThreads) {
        uint32_t m[16];
#pragma unroll
        for (int k = 0; k < 4; k++) m[k] = L.src[k][i];
#pragma unroll
        for (int k = 4; k < 16; k++) m[k] = 0;
        b2_state s;
        b2_init(s);
        b2_compress(s, m, 16, true);
#pragma unroll
        for (int k = 0; k < 8; k++) {
          leaves[8 * (size_t)i + k] = s.h[k];
          lv[0][8 * i + k] = s.h[k];
        }
      }
    }
    // node layers lg - 1 .. 0
    const uint32_t *root = tree_levels_to_root(L.tree, L.lg, lv, true);
    if (tid < 4) channel_step(t.chan, root, t.roots + 8 * L.root_slot, t.alpha);
    __syncthreads();
    const qm31 alpha = qm31_make(t.alpha[0], t.alpha[1], t.alpha[2], t.alpha[3]);
    const uint32_t n_out = 1u << (L.lg - L.step);
    for (uint32_t i = tid; i < n_out; i += kTailThreads) {
      switch (L.step) {
        case 1: tail_fold<1>(L, t.itw_end, alpha, i); break;
        case 2: tail_fold<2>(L, t.itw_end, alpha, i); break;
        case 3: tail_fold<3>(L, t.itw_end, alpha, i); break;
        default: tail_fold<4>(L, t.itw_end, alpha, i); break;
      }
    }
    __syncthreads();
  }
}

// ---- one FRI layer in two launches --------------------------------------
//
// Above kTailLog a layer is committed by fri_layer_kernel, which also runs
// the fold that produces the layer (from the previous layer, with alpha from
// the device channel), then hashes its lifted leaves and the first node
// levels of every 256-leaf subtree; the remaining levels and the channel
// step on the root run in fri_top_kernel (after merkle_nodes launches for
// levels wider than 2^kTopLog).  The trees, roots, channel state and folded
// evaluations written are the ones the separate fold, leaf, node and channel
// launches write.
constexpr int kFoldCircle = 5;    // fri_layer_kernel's FOLD for the circle fold
constexpr uint32_t kTopLog = 10;  // fri_top_kernel starts at a level of at most 2^kTopLog nodes

// FOLD = 0: the layer's values are src at i.  FOLD = 1..4: line folds of
// that many steps from a layer of log lg + FOLD.  kFoldCircle: the circle
// fold from log lg + 1.  A fold writes the layer to dst.  n_sub <= min(8, lg)
// node levels follow the leaves in each workgroup.
template <int FOLD>
__global__ void __launch_bounds__(256) fri_layer_kernel(Qm31ColsConst src, Qm31Cols dst, const uint32_t *itw_end,
                                                        const uint32_t *alpha_dev, uint32_t lg, uint32_t *tree,
                                                        uint32_t n_sub) {
  __shared__ uint32_t hs[256][9];
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i < (1u << lg)) {
    qm31 v;
    if constexpr (FOLD == 0) {
      v = qm31_make(as_global(src.c[0])[i], as_global(src.c[1])[i], as_global(src.c[2])[i], as_global(src.c[3])[i]);
    } else {
      const qm31 alpha = load_qm31(alpha_dev);
      if constexpr (FOLD == kFoldCircle)
        v = fold_circle_at(src, itw_end, lg + 1, alpha, i);
      else
        v = fold_line_at<FOLD>(src, itw_end, lg + FOLD, alpha, i);
#pragma unroll
      for (int k = 0; k < 4; k++) as_global(dst.c[k])[i] = v.v[k];
    }
    // lifted leaf: H(le32 of the 4 coordinates), one 16-byte block
    uint32_t m[16];
#pragma unroll
    for (int k = 0; k < 4; k++) m[k] = v.v[k];
#pragma unroll
    for (int k = 4; k < 16; k++) m[k] = 0;
    b2_state s;
    b2_init(s);
    b2_compress(s, m, 16, true);
    store_hash(tree + tree_layer_off(lg) + 8 * (size_t)i, s, 0);
#pragma unroll
    for (int k = 0; k < 8; k++) hs[threadIdx.x][k] = s.h[k];
  }
  subtree_levels(hs, tree, lg, n_sub, 0);
}

// Levels top_log - 1 .. 0 (level top_log - 1 of at most 2^kTopLog nodes),
// then the channel step on the root.
__global__ void __launch_bounds__(kTailThreads) fri_top_kernel(uint32_t *tree, uint32_t top_log, uint32_t *chan,
                                                               uint32_t *root_out, uint32_t *alpha_out) {
  __shared__ uint32_t lv[2][kTailThreads * 8];
  const uint32_t *root = tree_levels_to_root(tree, top_log, lv, false);
  if (threadIdx.x < 4) channel_step(chan, root, root_out, alpha_out);
}

// chan = [digest, n_draws, 0] (the device channel's starting state).
struct ChanInit {
  uint32_t w[10];
};
__global__ void chan_init_kernel(ChanInit init, uint32_t *chan) {
  if (threadIdx.x < 10) chan[threadIdx.x] = init.w[threadIdx.x];
}

// One Fiat-Shamir step of the prove on the device: chan (optionally first
// set to [init digest, 0, 0]) mixes `root` and draws a secure felt t into
// draw_out (channel_step).  When point_out is set, the draw is the OODS
// point's parameter (CirclePoint::get_random_point, STWO/core/circle.rs:169-181):
// p = ((1 - t^2) / (1 + t^2), 2t / (1 + t^2)); point_out receives the point's
// factors for eval_at_point, g_0 = p.y, g_1 = p.x, g_k = pi^(k-1)(p.x) for
// k <= point_log and 0 above (pi(x) = 2x^2 - 1; 128 words), then p.x, p.y and
// vn = pi^(point_log - 1)(p.x) (12 words) for the barycentric weights.
// One wave; the step runs on lanes 0-3.
__global__ void __launch_bounds__(64) prove_channel_kernel(uint32_t *chan, ChanInit init, uint32_t has_init,
                                                           const uint32_t *root, uint32_t *root_out,
                                                           uint32_t *draw_out, uint32_t *point_out,
                                                           uint32_t point_log) {
  const uint32_t tid = threadIdx.x;
  if (has_init) {
    if (tid < 10) chan[tid] = init.w[tid];
    __syncthreads();
  }
  if (tid < 4) channel_step(chan, root, root_out, draw_out);
  if (!point_out) return;
  __syncthreads();
  if (tid != 0) return;
  const qm31 t = load_qm31(draw_out);
  const qm31 one = qm31_make(1, 0, 0, 0);
  const qm31 t2 = qm31_mul(t, t);
  const qm31 inv = qm31_inv(qm31_add(t2, one));
  const qm31 px = qm31_mul(qm31_sub(one, t2), inv);
  const qm31 py = qm31_mul(qm31_add(t, t), inv);
  qm31 vn = px;
  for (int k = 0; k < 4; k++) point_out[k] = py.v[k];
  // g_1 .. g_point_log (the samples read g_k only for k below their log
  // size, which is at most point_log); zeros above
  qm31 x = px;
  for (uint32_t g = 1; g < 32; g++) {
    if (g == point_log) vn = x;
    for (int k = 0; k < 4; k++) point_out[4 * g + k] = g <= point_log ? x.v[k] : 0u;
    if (g < point_log) {
      const qm31 sq = qm31_mul(x, x);
      x = qm31_sub(qm31_add(sq, sq), one);
    }
  }
  for (int k = 0; k < 4; k++) {
    point_out[128 + k] = px.v[k];
    point_out[132 + k] = py.v[k];
    point_out[136 + k] = vn.v[k];
  }
}

int prove_channel_step(stwo_hip_ctx *ctx, uint32_t *chan, const uint8_t *init_digest, const uint32_t *root,
                       uint32_t *root_out, uint32_t *draw_out, uint32_t *point_out, uint32_t point_log) {
  STWO_HIP_CHECK_ARG(ctx, point_log >= 1 && point_log < 32, "internal: OODS point log %u", point_log);
  ChanInit init{};
  if (init_digest) memcpy(init.w, init_digest, 32);
  return launch(ctx, "prove_channel", prove_channel_kernel, dim3(1), dim3(64), 0, chan, init,
                (uint32_t)(init_digest != nullptr), root, root_out, draw_out, point_out, point_log);
}

// ---- the end of the FRI commit and the grind's prefix on the device ------
//
// For a last layer of at most 2^kDevLastMax values: LinePoly::interpolate of
// the last evaluation (fri.rs:274-291; line.rs:131-160: into natural order,
// the line IFFT with x^-1 twiddles of each doubling of the line domain,
// 1/n), into_ordered_coefficients (bit reverse), the degree check (every
// coefficient at or above 2^bound_log must be zero, else flags[0] = 1),
// from_ordered_coefficients (bit reverse of the kept ones) into poly, then
// mix_felts(poly) into the channel (blake2s.rs:58-68: H(digest || poly)),
// then the proof-of-work prefix H(le32(POW_PREFIX) || 0^12 || digest ||
// le32(pow_bits)) (blake2s.rs:122-139) into prefixed, and the grind's
// best-nonce word reset.  One workgroup; the hashing on lanes 0-3.
// The natural-order butterfly i of a line domain of log m uses the inverse x
// of point i, held at itw_end - 2^m + bitrev_{m-1}(i) (the tree keeps each
// doubling's x's of the first half in bit-reversed order).
constexpr uint32_t kDevLastMax = 6;
struct LastLayerArgs {
  const uint32_t *src[4];
  const uint32_t *itw_end;
  uint32_t lg, bound_log, pow_bits, n_inv;
  uint32_t *chan, *poly, *prefixed, *flags;
  unsigned long long *best;
};
__global__ void __launch_bounds__(256) fri_last_layer_kernel(LastLayerArgs a) {
  __shared__ uint32_t v[4][1u << kDevLastMax];
  __shared__ uint32_t msg[8 + (4u << kDevLastMax) + 16];
  const uint32_t tid = threadIdx.x, n = 1u << a.lg;
  if (tid == 0) a.flags[0] = 0;
  for (uint32_t j = tid; j < n; j += blockDim.x)
    for (int c = 0; c < 4; c++) v[c][j] = a.src[c][bitrev32(j, a.lg)];
  __syncthreads();
  for (uint32_t m = a.lg; m >= 1; m--) {
    const uint32_t half = 1u << (m - 1);
    const uint32_t *lvl = a.itw_end - ((size_t)1 << m);
    for (uint32_t b = tid; b < n / 2; b += blockDim.x) {
      const uint32_t i = b & (half - 1), base = (b >> (m - 1)) << m;
      const uint32_t xinv = lvl[bitrev32(i, m - 1)];
      for (int c = 0; c < 4; c++) {
        const uint32_t x = v[c][base + i], y = v[c][base + i + half];
        v[c][base + i] = m31_add(x, y);
        v[c][base + i + half] = m31_mul(m31_sub(x, y), xinv);
      }
    }
    __syncthreads();
  }
  const uint32_t bound = 1u << a.bound_log;
  if (tid < 8) msg[tid] = a.chan[tid];
  for (uint32_t j = tid; j < n; j += blockDim.x) {
    // ordered coefficient j = natural-order IFFT output at bitrev(j), times 1/n
    uint32_t cf[4];
    bool nz = false;
    for (int c = 0; c < 4; c++) {
      cf[c] = m31_mul(v[c][bitrev32(j, a.lg)], a.n_inv);
      nz = nz || cf[c] != 0;
    }
    if (j >= bound) {
      if (nz) a.flags[0] = 1;
    } else {
      const uint32_t o = 4 * bitrev32(j, a.bound_log);
      for (int c = 0; c < 4; c++) {
        a.poly[o + c] = cf[c];
        msg[8 + o + c] = cf[c];
      }
    }
  }
  __syncthreads();
  if (tid >= 4) return;
  const uint32_t j = tid;
  // digest = H(digest || poly): 32 + 16 bound bytes, streamed as the blake2
  // crate buffers it (the last block, full or partial, is the finalized one)
  const uint32_t total = 32 + 16 * bound, n_blocks = (total + 63) / 64;
  uint32_t lo = b2_iv_lo(j), hi = b2_iv_hi(j);
  for (uint32_t k = 0; k < n_blocks; k++) {
    uint32_t m[16];
    for (uint32_t w = 0; w < 16; w++) {
      const uint32_t idx = 16 * k + w;
      m[w] = idx < total / 4 ? msg[idx] : 0u;
    }
    const bool last = k + 1 == n_blocks;
    b2_compress_quad(lo, hi, m, last ? total : 64 * (k + 1), last);
  }
  uint32_t d[8];
  b2_quad_gather(lo, hi, d);
  a.chan[j] = d[j];
  a.chan[4 + j] = d[4 + j];
  if (j == 0) a.chan[8] = 0;  // mixing resets the draw counter
  uint32_t m[16];
  m[0] = kPowPrefix;
  m[1] = m[2] = m[3] = 0;
  for (int w = 0; w < 8; w++) m[4 + w] = d[w];
  m[12] = a.pow_bits;
  m[13] = m[14] = m[15] = 0;
  lo = b2_iv_lo(j);
  hi = b2_iv_hi(j);
  b2_compress_quad(lo, hi, m, 52, true);
  a.prefixed[j] = lo;
  a.prefixed[4 + j] = hi;
  if (j == 0) *a.best = ~0ull;
}

static int check_line_domain(stwo_hip_ctx *ctx, stwo_hip_coset dom, const stwo_hip_twiddles *tw) {
  STWO_HIP_CHECK_ARG(ctx, tw != nullptr && tw->device == ctx->device, "twiddles missing or from another device");
  STWO_HIP_CHECK_ARG(ctx, dom.log_size <= tw->root.log_size, "domain log %u exceeds twiddle root log %u",
                     dom.log_size, tw->root.log_size);
  uint32_t k = tw->root.log_size - dom.log_size;
  uint32_t expect = (uint32_t)(((uint64_t)tw->root.initial_index << k) & IDX_MASK);
  STWO_HIP_CHECK_ARG(ctx, (dom.initial_index & IDX_MASK) == expect,
                     "domain (initial %u, log %u) is not a doubling of the twiddle root (initial %u, log %u)",
                     dom.initial_index, dom.log_size, tw->root.initial_index, tw->root.log_size);
  return STWO_HIP_OK;
}

static int fold_circle(stwo_hip_ctx *ctx, uint32_t *const dst[4], const uint32_t *const src[4], uint32_t N,
                       const qm31 &alpha, const stwo_hip_twiddles *tw, bool accumulate,
                       const uint32_t *alpha_dev = nullptr) {
  Qm31Cols d;
  Qm31ColsConst s;
  for (int k = 0; k < 4; k++) {
    d.c[k] = dst[k];
    s.c[k] = src[k];
  }
  uint32_t n_out = 1u << (N - 1);
  return launch(ctx, accumulate ? "fold_circle_into_line" : "fold_circle_first_layer", fold_circle_kernel,
                dim3((n_out + 255) / 256), dim3(256), 0, d, s, (const uint32_t *)(tw->d_itw + tw->len), N, alpha,
                qm31_mul(alpha, alpha), (uint32_t)accumulate, alpha_dev);
}

// fold_step folds of a line evaluation of log size n (domain checked by caller).
// With `alpha_dev` (alpha in device memory), `alpha_tmp` is a 4-word device
// slot that receives alpha^(2^4k) for fold steps beyond the first 4.
static int fold_line_steps(stwo_hip_ctx *ctx, const uint32_t *const src[4], uint32_t n, qm31 alpha, uint32_t fold_step,
                           const stwo_hip_twiddles *tw, uint32_t *const dst[4], const uint32_t *alpha_dev = nullptr,
                           uint32_t *alpha_tmp = nullptr) {
  const uint32_t *itw_end = tw->d_itw + tw->len;
  const uint32_t *cur[4] = {src[0], src[1], src[2], src[3]};
  uint32_t iter = 0;
  while (fold_step > 0) {
    uint32_t step = std::min<uint32_t>(fold_step, 4);
    uint32_t n_out_log = n - step;
    uint32_t *out[4];
    if (fold_step == step) {
      for (int k = 0; k < 4; k++) out[k] = dst[k];
    } else {
      // Intermediate results alternate between two scratch buffers.
      void *p = nullptr;
      int rc = get_scratch(ctx, (iter & 1) ? "fold_line_b" : "fold_line_a", ((size_t)4 << n_out_log) * 4, &p);
      if (rc) return rc;
      for (int k = 0; k < 4; k++) out[k] = (uint32_t *)p + ((size_t)k << n_out_log);
    }
    Qm31Cols d;
    Qm31ColsConst s;
    for (int k = 0; k < 4; k++) {
      d.c[k] = out[k];
      s.c[k] = cur[k];
    }
    uint32_t blocks = ((1u << n_out_log) + 255) / 256;
    int rc;
    switch (step) {
      case 1: rc = launch(ctx, "fold_line", fold_line_kernel<1>, dim3(blocks), dim3(256), 0, d, s, itw_end, n, alpha, alpha_dev); break;
      case 2: rc = launch(ctx, "fold_line", fold_line_kernel<2>, dim3(blocks), dim3(256), 0, d, s, itw_end, n, alpha, alpha_dev); break;
      case 3: rc = launch(ctx, "fold_line", fold_line_kernel<3>, dim3(blocks), dim3(256), 0, d, s, itw_end, n, alpha, alpha_dev); break;
      default: rc = launch(ctx, "fold_line", fold_line_kernel<4>, dim3(blocks), dim3(256), 0, d, s, itw_end, n, alpha, alpha_dev); break;
    }
    if (rc) return rc;
    for (uint32_t k = 0; k < step; k++) alpha = qm31_mul(alpha, alpha);
    if (alpha_dev && fold_step > step) {
      rc = launch(ctx, "alpha_pow2k", alpha_pow2k_kernel, dim3(1), dim3(64), 0, alpha_dev, alpha_tmp, step);
      if (rc) return rc;
      alpha_dev = alpha_tmp;
    }
    for (int k = 0; k < 4; k++) cur[k] = out[k];
    n = n_out_log;
    fold_step -= step;
    iter++;
  }
  return STWO_HIP_OK;
}

// LineEvaluation::interpolate on the host (STWO/prover/line.rs:54-99): the
// last FRI layer has at most 2^(10 + 16) values and usually 2^1..2^4.
static void host_line_interpolate(std::vector<qm31> &vals, uint32_t line_initial, uint32_t log_n) {
  size_t n = (size_t)1 << log_n;
  for (size_t i = 0; i < n; i++) {
    size_t j = bitrev32((uint32_t)i, log_n);
    if (j > i) std::swap(vals[i], vals[j]);
  }
  uint32_t init = line_initial, lg = log_n;
  while (lg >= 1) {
    size_t size = (size_t)1 << lg;
    cpoint c = host_index_to_point(init), st = host_index_to_point(coset_step(lg));
    for (size_t i = 0; i < size / 2; i++) {
      uint32_t xinv = m31_inv(c.x);
      for (size_t base = 0; base < n; base += size) {
        qm31 &a = vals[base + i], &b = vals[base + i + size / 2];
        qm31 s = qm31_add(a, b), d = qm31_mul_m31(qm31_sub(a, b), xinv);
        a = s;
        b = d;
      }
      c = cp_add(c, st);
    }
    init = (uint32_t)(((uint64_t)init * 2) & IDX_MASK);
    lg--;
  }
  uint32_t inv = m31_inv((uint32_t)(n % P));
  for (auto &v : vals) v = qm31_mul_m31(v, inv);
}

static inline size_t layer_off(uint32_t l) { return (((size_t)1 << l) - 1) * 8; }

}  // namespace stwo_hip

using namespace stwo_hip;

extern "C" {

int stwo_hip_fold_circle_into_line(stwo_hip_ctx *ctx, uint32_t *const dst[4], const uint32_t *const src[4],
                                   stwo_hip_coset src_domain, const uint32_t alpha[4], const stwo_hip_twiddles *tw) {
  if (!ctx) return STWO_HIP_EINVAL;
  STWO_HIP_CHECK_ARG(ctx, dst && src && alpha, "fold_circle_into_line: NULL argument");
  for (int k = 0; k < 4; k++) STWO_HIP_CHECK_ARG(ctx, dst[k] && src[k], "fold_circle_into_line: NULL coordinate");
  int rc = check_line_domain(ctx, src_domain, tw);
  if (rc) return rc;
  uint32_t N = src_domain.log_size + 1;
  STWO_HIP_CHECK_ARG(ctx, N >= 3, "fold_circle_into_line: domain of log size %u is below 3", N);
  return fold_circle(ctx, dst, src, N, qm31_make(alpha[0], alpha[1], alpha[2], alpha[3]), tw, true);
}

int stwo_hip_fold_line(stwo_hip_ctx *ctx, const uint32_t *const src[4], stwo_hip_coset line_domain,
                       const uint32_t alpha[4], uint32_t fold_step, const stwo_hip_twiddles *tw,
                       uint32_t *const dst[4]) {
  if (!ctx) return STWO_HIP_EINVAL;
  STWO_HIP_CHECK_ARG(ctx, dst && src && alpha, "fold_line: NULL argument");
  for (int k = 0; k < 4; k++) STWO_HIP_CHECK_ARG(ctx, dst[k] && src[k], "fold_line: NULL coordinate");
  STWO_HIP_CHECK_ARG(ctx, fold_step >= 1 && fold_step <= line_domain.log_size,
                     "fold_line: fold_step %u invalid for log size %u", fold_step, line_domain.log_size);
  int rc = check_line_domain(ctx, line_domain, tw);
  if (rc) return rc;
  return fold_line_steps(ctx, src, line_domain.log_size, qm31_make(alpha[0], alpha[1], alpha[2], alpha[3]), fold_step,
                         tw, dst);
}

}  // extern "C"

namespace stwo_hip {
// The layer schedule of FriProver::commit (fri.rs:188-262): the first
// (circle) layer of log N folds once into the line domain of log N - 1;
// inner layers fold line_fold_step at a time, the last one exactly down to
// the last layer of log (log_last_layer_degree_bound + log_blowup_factor).
std::vector<FriLayerShape> fri_schedule(const stwo_hip_fri_config &cfg, uint32_t log_size) {
  const uint32_t last_log = cfg.log_last_layer_degree_bound + cfg.log_blowup_factor;
  std::vector<FriLayerShape> out;
  out.push_back({log_size, 1});
  uint32_t lg = log_size - 1;
  if (lg != last_log) {
    for (;;) {
      bool last = !(lg > last_log + cfg.line_fold_step);
      uint32_t step = last ? lg - last_log : cfg.line_fold_step;
      out.push_back({lg, step});
      lg -= step;
      if (last) break;
    }
  }
  return out;
}

static int check_fri_args(stwo_hip_ctx *ctx, const stwo_hip_fri_config &cfg, uint32_t log_size,
                          const stwo_hip_twiddles *tw) {
  const uint32_t last_log = cfg.log_last_layer_degree_bound + cfg.log_blowup_factor;
  STWO_HIP_CHECK_ARG(ctx, cfg.line_fold_step >= 1, "Line fold step must be positive.");
  // A step above the largest line domain (log 27) can only overshoot; larger
  // values would also wrap the layer-size arithmetic of the schedule.
  STWO_HIP_CHECK_ARG(ctx, cfg.line_fold_step <= 28, "line_fold_step %u exceeds 28", cfg.line_fold_step);
  STWO_HIP_CHECK_ARG(ctx, cfg.log_last_layer_degree_bound <= 10, "log_last_layer_degree_bound must be <= 10");
  STWO_HIP_CHECK_ARG(ctx, cfg.log_blowup_factor >= 1 && cfg.log_blowup_factor <= 16, "log_blowup_factor not in 1..=16");
  STWO_HIP_CHECK_ARG(ctx, log_size >= 3 && log_size - 1 >= last_log && log_size <= 28,
                     "fri_commit: log size %u incompatible with last layer log %u", log_size, last_log);
  stwo_hip_coset dom = {half_odds_initial(log_size - 1), log_size - 1};
  return check_line_domain(ctx, dom, tw);
}

// FriProver::commit.  With `keep`, every committed layer's tree and every
// folded evaluation stays in `keep->mem` (laid out by the caller from
// fri_schedule); without, one scratch tree and two scratch layer buffers
// are reused.
// grind: NULL, or (prove) the end of the commit and the proof of work's first
// search range run on the device when the last layer allows
// (fri_last_layer_kernel): grind->found then says whether that range held
// the nonce.
struct FriGrind {
  uint32_t pow_bits;
  bool found;
  uint64_t nonce;
};
static int fri_commit_impl(stwo_hip_ctx *ctx, uint8_t channel_digest[32], uint32_t *channel_n_draws,
                           const stwo_hip_fri_config &cfg, const uint32_t *const column[4], uint32_t log_size,
                           const stwo_hip_twiddles *tw, uint32_t *roots, uint32_t *n_roots, uint32_t *last_layer_poly,
                           stwo_hip_fri_prover *keep, FriGrind *grind = nullptr) {
  int rc = check_fri_args(ctx, cfg, log_size, tw);
  if (rc) return rc;
  const uint32_t last_log = cfg.log_last_layer_degree_bound + cfg.log_blowup_factor;
  const std::vector<FriLayerShape> sched = fri_schedule(cfg, log_size);
  if (keep && keep->layers.size() != sched.size())
    return set_error(ctx, STWO_HIP_EINVAL, "internal: FRI prover storage does not match the schedule");
  // The layer loop runs on the stream without host round trips: after each
  // layer's Merkle commit, the kernel that hashes its root (fri_top_kernel,
  // or the tail) mixes the root into the device copy of the channel and
  // draws the next alpha, which the folds read from device memory.  The host reads the roots, the channel
  // and the last layer once, at the end.
  constexpr uint32_t kMaxRoots = 32;
  // chan[0..10) | alpha[10..14) | alpha_tmp[14..18) | pad | roots[32 ..)
  constexpr uint32_t kRootsOff = 32;
  // One root per layer and at most log_size layers (log_size <= 28).
  static_assert(kMaxRoots >= 28, "root slots");
  void *tree_p = nullptr, *lay_a = nullptr, *lay_b = nullptr, *state_p = nullptr;
  if (!keep) {
    rc = get_scratch(ctx, "fri_tree", layer_off(log_size + 1) * 4, &tree_p);
    if (rc) return rc;
    rc = get_scratch(ctx, "fri_layer_a", ((size_t)4 << (log_size - 1)) * 4, &lay_a);
    if (rc) return rc;
    rc = get_scratch(ctx, "fri_layer_b", ((size_t)4 << (log_size - 1)) * 4, &lay_b);
    if (rc) return rc;
  }
  // then flags[kFlagsOff], the grind's best nonce (8-byte aligned), the
  // grind prefix and the last layer's coefficients (device end only)
  constexpr uint32_t kFlagsOff = kRootsOff + 8 * kMaxRoots, kBestOff = kFlagsOff + 2, kPrefixOff = kBestOff + 2,
                     kPolyOff = kPrefixOff + 8;
  static_assert(kBestOff % 2 == 0, "best nonce alignment");
  rc = get_scratch(ctx, "fri_state", (kPolyOff + (4u << kDevLastMax)) * 4, &state_p);
  if (rc) return rc;
  uint32_t *scratch_tree = (uint32_t *)tree_p;
  uint32_t *state = (uint32_t *)state_p;
  uint32_t *d_chan = state, *d_alpha = state + 10, *d_alpha_tmp = state + 14, *d_roots = state + kRootsOff;
  {
    // set by a launch, in stream order: no synchronization
    ChanInit init{};
    memcpy(init.w, channel_digest, 32);
    init.w[8] = *channel_n_draws;
    init.w[9] = 0;
    rc = launch(ctx, "fri_chan_init", chan_init_kernel, dim3(1), dim3(64), 0, init, d_chan);
    if (rc) return rc;
  }
  uint32_t nr = 0;

  // A fold not launched yet: the circle fold after the first layer or a line
  // fold of `step` steps.  The next layer's commit runs it in its own launch
  // (fri_layer_kernel) when the step is at most 4; otherwise, and before the
  // tail or the end, it runs on its own.
  struct PendingFold {
    bool on, circle;
    uint32_t step, src_log;
    const uint32_t *src[4];
    uint32_t *dst[4];
  } pend{};
  const qm31 unused = qm31_make(0, 0, 0, 0);
  auto flush = [&]() -> int {
    if (!pend.on) return STWO_HIP_OK;
    pend.on = false;
    if (pend.circle) return fold_circle(ctx, pend.dst, pend.src, pend.src_log, unused, tw, false, d_alpha);
    return fold_line_steps(ctx, pend.src, pend.src_log, unused, pend.step, tw, pend.dst, d_alpha, d_alpha_tmp);
  };

  // Commits the layer `cols` of log lg (running the pending fold that writes
  // it, if it fuses), then mixes its root into the channel and draws alpha.
  auto commit_layer = [&](const uint32_t *const cols[4], uint32_t lg) -> int {
    if (keep && nr >= keep->layers.size()) return set_error(ctx, STWO_HIP_EINVAL, "internal: FRI layer overflow");
    uint32_t *tree = keep ? keep->layers[nr].tree : scratch_tree;
    if (keep)
      for (int k = 0; k < 4; k++) keep->layers[nr].cols[k] = cols[k];
    Qm31ColsConst src;
    Qm31Cols dst{};
    int fold = 0;
    if (pend.on && (pend.circle || pend.step <= 4)) {
      fold = pend.circle ? kFoldCircle : (int)pend.step;
      for (int k = 0; k < 4; k++) {
        src.c[k] = pend.src[k];
        dst.c[k] = pend.dst[k];
      }
      pend.on = false;
    } else {
      int r = flush();
      if (r) return r;
      for (int k = 0; k < 4; k++) src.c[k] = cols[k];
    }
    const uint32_t n_sub = std::min<uint32_t>(8, lg);
    const dim3 grid((uint32_t)(((1ull << lg) + 255) / 256)), block(256);
    const uint32_t *itw_end = tw->d_itw + tw->len;
    int r;
    switch (fold) {
      case 0: r = launch(ctx, "fri_layer", fri_layer_kernel<0>, grid, block, 0, src, dst, itw_end, (const uint32_t *)d_alpha, lg, tree, n_sub); break;
      case 1: r = launch(ctx, "fri_layer", fri_layer_kernel<1>, grid, block, 0, src, dst, itw_end, (const uint32_t *)d_alpha, lg, tree, n_sub); break;
      case 2: r = launch(ctx, "fri_layer", fri_layer_kernel<2>, grid, block, 0, src, dst, itw_end, (const uint32_t *)d_alpha, lg, tree, n_sub); break;
      case 3: r = launch(ctx, "fri_layer", fri_layer_kernel<3>, grid, block, 0, src, dst, itw_end, (const uint32_t *)d_alpha, lg, tree, n_sub); break;
      case 4: r = launch(ctx, "fri_layer", fri_layer_kernel<4>, grid, block, 0, src, dst, itw_end, (const uint32_t *)d_alpha, lg, tree, n_sub); break;
      default: r = launch(ctx, "fri_layer", fri_layer_kernel<kFoldCircle>, grid, block, 0, src, dst, itw_end, (const uint32_t *)d_alpha, lg, tree, n_sub); break;
    }
    if (r) return r;
    uint32_t top = lg - n_sub;
    if (top > kTopLog + 1) {
      r = hash_node_layers(ctx, tree, top, kTopLog + 1, STWO_HIP_HASH_BLAKE2S);
      if (r) return r;
      top = kTopLog + 1;
    }
    r = launch(ctx, "fri_top", fri_top_kernel, dim3(1), dim3(kTailThreads), 0, tree, top, d_chan, d_roots + 8 * nr,
               d_alpha);
    if (r) return r;
    nr++;
    return STWO_HIP_OK;
  };

  // First layer (fri.rs:176-183, 360-369).
  rc = commit_layer(column, log_size);
  if (rc) return rc;
  // Inner layers (fri.rs:188-262).
  uint32_t lg = log_size - 1;
  uint32_t line_init = half_odds_initial(lg);
  uint32_t *bufs[2] = {(uint32_t *)lay_a, (uint32_t *)lay_b};
  uint32_t *cur[4];
  // Evaluation buffer of the layer of log `l` that fold number `f` writes.
  auto layer_buf = [&](int f, uint32_t l, int k) -> uint32_t * {
    return keep ? keep->evals[f] + ((size_t)k << l) : bufs[f & 1] + ((size_t)k << l);
  };
  for (int k = 0; k < 4; k++) cur[k] = layer_buf(0, lg, k);
  int n_folds = 1;
  pend.on = true;
  pend.circle = true;
  pend.step = 1;
  pend.src_log = log_size;
  for (int k = 0; k < 4; k++) {
    pend.src[k] = column[k];
    pend.dst[k] = cur[k];
  }
  if (lg != last_log) {
    for (;;) {
      bool last = !(lg > last_log + cfg.line_fold_step);
      uint32_t step = last ? lg - last_log : cfg.line_fold_step;
      if (lg <= kTailLog && cfg.line_fold_step <= 4) {
        rc = flush();
        if (rc) return rc;
        // Every remaining inner layer (at most kTailLog of them) in one launch.
        FriTail t{};
        t.itw_end = tw->d_itw + tw->len;
        t.chan = d_chan;
        t.roots = d_roots;
        t.alpha = d_alpha;
        for (;;) {
          last = !(lg > last_log + cfg.line_fold_step);
          step = last ? lg - last_log : cfg.line_fold_step;
          if (t.n >= kTailMax || (keep && nr >= keep->layers.size()))
            return set_error(ctx, STWO_HIP_EINVAL, "internal: FRI tail overflow");
          FriTailLayer &L = t.l[t.n++];
          L.tree = keep ? keep->layers[nr].tree : scratch_tree;
          for (int k = 0; k < 4; k++) {
            L.src[k] = cur[k];
            if (keep) keep->layers[nr].cols[k] = cur[k];
            L.dst[k] = layer_buf(n_folds, lg - step, k);
            cur[k] = L.dst[k];
          }
          L.lg = lg;
          L.step = step;
          L.root_slot = nr++;
          n_folds++;
          line_init = (uint32_t)(((uint64_t)line_init << step) & IDX_MASK);
          lg -= step;
          if (last) break;
        }
        rc = launch(ctx, "fri_tail", fri_tail_kernel, dim3(1), dim3(kTailThreads), 0, t);
        if (rc) return rc;
        break;
      }
      rc = commit_layer(cur, lg);
      if (rc) return rc;
      pend.on = true;
      pend.circle = false;
      pend.step = step;
      pend.src_log = lg;
      for (int k = 0; k < 4; k++) {
        pend.src[k] = cur[k];
        pend.dst[k] = layer_buf(n_folds, lg - step, k);
        cur[k] = pend.dst[k];
      }
      n_folds++;
      line_init = (uint32_t)(((uint64_t)line_init << step) & IDX_MASK);
      lg -= step;
      if (last) break;
    }
  }
  // The fold into the last layer when no tail ran (or the circle fold alone).
  rc = flush();
  if (rc) return rc;
  if (grind) grind->found = false;
  if (grind && lg <= kDevLastMax && grind->pow_bits + 6 < 20) {
    // The last layer, the channel mix and the grind's first range on the
    // device, then one synchronization.
    const uint32_t bound_log = cfg.log_last_layer_degree_bound;
    LastLayerArgs la{};
    for (int k = 0; k < 4; k++) la.src[k] = cur[k];
    la.itw_end = tw->d_itw + tw->len;
    la.lg = lg;
    la.bound_log = bound_log;
    la.pow_bits = grind->pow_bits;
    la.n_inv = m31_inv((uint32_t)(((uint64_t)1 << lg) % P));
    la.chan = d_chan;
    la.poly = state + kPolyOff;
    la.prefixed = state + kPrefixOff;
    la.flags = state + kFlagsOff;
    la.best = (unsigned long long *)(state + kBestOff);
    rc = launch(ctx, "fri_last_layer", fri_last_layer_kernel, dim3(1), dim3(256), 0, la);
    if (rc) return rc;
    rc = grind_first_range(ctx, state + kPrefixOff, grind->pow_bits, la.best);
    if (rc) return rc;
    const size_t bound = (size_t)1 << bound_log;
    std::vector<uint32_t> st(kPolyOff + 4 * bound);
    {
      const DownloadReq req = {st.data(), state, st.size()};
      rc = download_many(ctx, &req, 1);
      if (rc) return rc;
    }
    if (st[9]) return set_error(ctx, STWO_HIP_EDEVICE, "fri_commit: channel draw did not terminate");
    if (nr != sched.size()) return set_error(ctx, STWO_HIP_EDEVICE, "internal: FRI schedule mismatch");
    if (st[kFlagsOff]) return set_error(ctx, STWO_HIP_EDEGREE, "invalid degree");
    memcpy(roots, st.data() + kRootsOff, 32 * nr);
    if (keep)
      for (uint32_t k = 0; k < nr; k++) memcpy(keep->layers[k].root, roots + 8 * k, 32);
    memcpy(last_layer_poly, st.data() + kPolyOff, 16 * bound);
    memcpy(channel_digest, st.data(), 32);
    *channel_n_draws = st[8];
    uint64_t best;
    memcpy(&best, st.data() + kBestOff, 8);
    if (best != ~0ull) {
      grind->found = true;
      grind->nonce = best;
    }
    *n_roots = nr;
    if (keep) {
      keep->last_layer_poly.assign(last_layer_poly, last_layer_poly + 4 * bound);
      keep->last_log = lg;
      for (int k = 0; k < 4; k++) keep->last_layer_eval[k] = cur[k];
    }
    return STWO_HIP_OK;
  }
  // One synchronization for the channel state, the roots and the last layer.
  std::vector<uint32_t> st(kRootsOff + 8 * nr);
  size_t n = (size_t)1 << lg;
  std::vector<uint32_t> h(4 * n);
  {
    const DownloadReq reqs[5] = {{st.data(), state, st.size()},
                                 {h.data(), cur[0], n},
                                 {h.data() + n, cur[1], n},
                                 {h.data() + 2 * n, cur[2], n},
                                 {h.data() + 3 * n, cur[3], n}};
    rc = download_many(ctx, reqs, 5);
    if (rc) return rc;
  }
  if (st[9]) return set_error(ctx, STWO_HIP_EDEVICE, "fri_commit: channel draw did not terminate");
  if (nr != sched.size()) return set_error(ctx, STWO_HIP_EDEVICE, "internal: FRI schedule mismatch");
  memcpy(roots, st.data() + kRootsOff, 32 * nr);
  if (keep)
    for (uint32_t k = 0; k < nr; k++) memcpy(keep->layers[k].root, roots + 8 * k, 32);
  HostChannel ch;
  memcpy(ch.digest, st.data(), 32);
  ch.n_draws = st[8];
  // Last layer (fri.rs:274-291).
  std::vector<qm31> vals(n);
  for (size_t i = 0; i < n; i++) vals[i] = qm31_make(h[i], h[n + i], h[2 * n + i], h[3 * n + i]);
  host_line_interpolate(vals, line_init, lg);
  // into_ordered_coefficients: bit reverse (line.rs:157-160)
  for (size_t i = 0; i < n; i++) {
    size_t j = bitrev32((uint32_t)i, lg);
    if (j > i) std::swap(vals[i], vals[j]);
  }
  size_t bound = (size_t)1 << cfg.log_last_layer_degree_bound;
  for (size_t i = bound; i < n; i++)
    for (int k = 0; k < 4; k++)
      if (vals[i].v[k]) return set_error(ctx, STWO_HIP_EDEGREE, "invalid degree");
  // from_ordered_coefficients bit-reverses the kept coefficients (line.rs:167-170)
  for (size_t i = 0; i < bound; i++) {
    size_t j = bitrev32((uint32_t)i, cfg.log_last_layer_degree_bound);
    for (int k = 0; k < 4; k++) last_layer_poly[4 * j + k] = vals[i].v[k];
  }
  ch.mix_felts(last_layer_poly, bound);
  memcpy(channel_digest, ch.digest, 32);
  *channel_n_draws = ch.n_draws;
  *n_roots = nr;
  if (keep) {
    keep->last_layer_poly.assign(last_layer_poly, last_layer_poly + 4 * bound);
    keep->last_log = lg;
    for (int k = 0; k < 4; k++) keep->last_layer_eval[k] = cur[k];
  }
  return STWO_HIP_OK;
}

}  // namespace stwo_hip

extern "C" {

int stwo_hip_fri_commit(stwo_hip_ctx *ctx, uint8_t channel_digest[32], uint32_t *channel_n_draws,
                        stwo_hip_fri_config cfg, const uint32_t *const column[4], uint32_t log_size,
                        const stwo_hip_twiddles *tw, uint32_t *roots, uint32_t *n_roots, uint32_t *last_layer_poly) {
  if (!ctx) return STWO_HIP_EINVAL;
  STWO_HIP_CHECK_ARG(ctx, channel_digest && channel_n_draws && column && roots && n_roots && last_layer_poly,
                     "fri_commit: NULL argument");
  return fri_commit_impl(ctx, channel_digest, channel_n_draws, cfg, column, log_size, tw, roots, n_roots,
                         last_layer_poly, nullptr);
}

int stwo_hip_fri_prover_commit(stwo_hip_ctx *ctx, uint8_t channel_digest[32], uint32_t *channel_n_draws,
                               stwo_hip_fri_config cfg, const uint32_t *const column[4], uint32_t log_size,
                               const stwo_hip_twiddles *tw, uint32_t *roots, uint32_t *n_roots,
                               uint32_t *last_layer_poly, stwo_hip_fri_prover **out) {
  return fri_prover_commit_mem(ctx, channel_digest, channel_n_draws, cfg, column, log_size, tw, roots, n_roots,
                               last_layer_poly, nullptr, out, -1, nullptr, nullptr);
}

}  // extern "C"

int stwo_hip::fri_prover_commit_mem(stwo_hip_ctx *ctx, uint8_t channel_digest[32], uint32_t *channel_n_draws,
                                    stwo_hip_fri_config cfg, const uint32_t *const column[4], uint32_t log_size,
                                    const stwo_hip_twiddles *tw, uint32_t *roots, uint32_t *n_roots,
                                    uint32_t *last_layer_poly, const char *scratch_key, stwo_hip_fri_prover **out,
                                    int pow_bits, uint64_t *nonce, bool *nonce_found) {
  if (!ctx) return STWO_HIP_EINVAL;
  STWO_HIP_CHECK_ARG(ctx, channel_digest && channel_n_draws && column && roots && n_roots && last_layer_poly && out,
                     "fri_prover_commit: NULL argument");
  *out = nullptr;
  for (int k = 0; k < 4; k++) STWO_HIP_CHECK_ARG(ctx, column[k] != nullptr, "fri_prover_commit: NULL column");
  int rc = check_fri_args(ctx, cfg, log_size, tw);
  if (rc) return rc;
  auto *pr = new stwo_hip_fri_prover();
  pr->device = ctx->device;
  pr->cfg = cfg;
  pr->log_size = log_size;
  const std::vector<FriLayerShape> sched = fri_schedule(cfg, log_size);
  // One allocation: a packed tree per committed layer, then the evaluation
  // written by each fold (the last one is the uncommitted last layer).
  size_t words = 0;
  std::vector<size_t> tree_off, eval_off;
  for (const FriLayerShape &l : sched) {
    tree_off.push_back(words);
    words += layer_off(l.log_size + 1);
  }
  uint32_t lg = log_size;
  for (size_t f = 0; f < sched.size(); f++) {
    lg -= sched[f].fold_step;
    eval_off.push_back(words);
    words += (size_t)4 << lg;
  }
  if (scratch_key) {
    void *m = nullptr;
    rc = get_scratch(ctx, scratch_key, words * 4, &m);
    pr->mem = (uint32_t *)m;
    pr->owns_mem = false;
  } else {
    rc = stwo_hip_alloc(ctx, words, &pr->mem);
  }
  if (rc) {
    delete pr;
    return rc;
  }
  pr->mem_words = words;
  pr->layers.resize(sched.size());
  for (size_t k = 0; k < sched.size(); k++) {
    pr->layers[k].shape = sched[k];
    pr->layers[k].tree = pr->mem + tree_off[k];
  }
  for (size_t f = 0; f < sched.size(); f++) pr->evals.push_back(pr->mem + eval_off[f]);
  FriGrind g{};
  if (pow_bits >= 0) g.pow_bits = (uint32_t)pow_bits;
  rc = fri_commit_impl(ctx, channel_digest, channel_n_draws, cfg, column, log_size, tw, roots, n_roots,
                       last_layer_poly, pr, pow_bits >= 0 ? &g : nullptr);
  if (!rc && nonce_found) {
    *nonce_found = g.found;
    if (g.found) *nonce = g.nonce;
  }
  if (rc) {
    if (pr->owns_mem) stwo_hip_free(ctx, pr->mem);
    delete pr;
    return rc;
  }
  *out = pr;
  return STWO_HIP_OK;
}

extern "C" {

// FriProver::decommit (fri.rs:294-347): draw the query positions from the
// channel (draw_queries, core/queries.rs:11-28; Queries::new sorts and
// deduplicates), then decommit every committed layer:
// compute_decommitment_positions_and_witness_evals (fri.rs:455-489) with the
// layer's fold step (1 for the circle layer), and the Merkle decommitment
// of the layer's tree on those positions; the queries fold by the layer's
// step between layers (Queries::fold, queries.rs:56-63).  Every value and
// hash read of all layers is one device gather.
int stwo_hip_fri_prover_decommit(stwo_hip_ctx *ctx, stwo_hip_fri_prover *pr, uint8_t channel_digest[32],
                                 uint32_t *channel_n_draws, uint32_t *query_positions, uint32_t *n_positions,
                                 uint32_t *unsorted_queries) {
  if (!ctx) return STWO_HIP_EINVAL;
  STWO_HIP_CHECK_ARG(ctx, query_positions && n_positions && unsorted_queries, "fri_prover_decommit: NULL argument");
  STWO_HIP_CHECK_ARG(ctx, channel_digest && channel_n_draws, "fri_prover_decommit: NULL argument");
  // The channel is written back only on success.
  uint8_t digest[32];
  memcpy(digest, channel_digest, 32);
  uint32_t n_draws = *channel_n_draws;
  Gatherer g;
  FriDecommitPlan plan;
  int rc = fri_decommit_plan(ctx, pr, digest, &n_draws, g, plan);
  if (rc) return rc;
  rc = g.run(ctx);
  if (rc) return rc;
  memcpy(channel_digest, digest, 32);
  *channel_n_draws = n_draws;
  memcpy(query_positions, plan.queries.data(), plan.queries.size() * 4);
  *n_positions = (uint32_t)plan.queries.size();
  memcpy(unsorted_queries, plan.raw.data(), plan.raw.size() * 4);
  fri_decommit_finish(pr, g, plan);
  return STWO_HIP_OK;
}

}  // extern "C"

extern "C" {

int stwo_hip_fri_prover_layer_proof(stwo_hip_ctx *ctx, const stwo_hip_fri_prover *pr, uint32_t layer,
                                    uint32_t commitment[8], uint32_t *fri_witness, size_t fri_cap, size_t *n_fri,
                                    uint32_t *hash_witness, size_t hash_cap, size_t *n_hash, uint32_t *positions,
                                    uint32_t *values, size_t pos_cap, size_t *n_pos) {
  if (!ctx) return STWO_HIP_EINVAL;
  STWO_HIP_CHECK_ARG(ctx, pr && commitment && n_fri && n_hash && n_pos, "fri_prover_layer_proof: NULL argument");
  STWO_HIP_CHECK_ARG(ctx, pr->decommitted, "fri_prover_layer_proof: not decommitted yet");
  STWO_HIP_CHECK_ARG(ctx, layer < pr->proofs.size(), "fri_prover_layer_proof: layer %u of %zu", layer,
                     pr->proofs.size());
  const FriLayerDecommit &d = pr->proofs[layer];
  memcpy(commitment, pr->layers[layer].root, 32);
  *n_fri = d.fri_witness.size() / 4;
  *n_hash = d.hash_witness.size() / 8;
  *n_pos = d.positions.size();
  if (!fri_witness && !hash_witness && !positions) return STWO_HIP_OK;  // size query
  STWO_HIP_CHECK_ARG(ctx, *n_fri <= fri_cap && *n_hash <= hash_cap && (!positions || *n_pos <= pos_cap),
                     "fri_prover_layer_proof: capacity too small (%zu, %zu, %zu needed)", *n_fri, *n_hash, *n_pos);
  STWO_HIP_CHECK_ARG(ctx, (fri_witness || !*n_fri) && (hash_witness || !*n_hash),
                     "fri_prover_layer_proof: NULL output");
  if (*n_fri) memcpy(fri_witness, d.fri_witness.data(), d.fri_witness.size() * 4);
  if (*n_hash) memcpy(hash_witness, d.hash_witness.data(), d.hash_witness.size() * 4);
  if (positions) {
    memcpy(positions, d.positions.data(), d.positions.size() * 4);
    if (values) memcpy(values, d.values.data(), d.values.size() * 4);
  }
  return STWO_HIP_OK;
}

int stwo_hip_fri_prover_destroy(stwo_hip_ctx *ctx, stwo_hip_fri_prover *prover) {
  if (!ctx) return STWO_HIP_EINVAL;
  if (!prover) return STWO_HIP_OK;
  int rc = stwo_hip_sync(ctx);
  int rc2 = prover->owns_mem ? stwo_hip_free(ctx, prover->mem) : STWO_HIP_OK;
  delete prover;
  return rc ? rc : rc2;
}

}  // extern "C"

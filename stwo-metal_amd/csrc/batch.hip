// Batched LDE + Merkle commitment of columns of uneven sizes.
//
// The plain path dispatches per column (the reference's evaluate_polynomials
// submits one RFFT per column, MB/poly.rs:715-779, and its IFFT one per
// column, MB/poly.rs:364-391).  Here a plan covers a whole column set:
//   - columns are grouped by log size; each group's FFT is a list of passes
//     (radix-16, see fft.hip); passes of the same level from every group run
//     in ONE launch (fft_pass_multi_kernel), so the IFFT and the RFFT take
//     2-3 launches each whatever the number of sizes;
//   - the Merkle tree is built layer by layer: mixed degree injects the
//     evaluations of each size at their own layer (MerkleProver::commit,
//     STWO/prover/vcs/prover.rs:39-65), lifted hashes all columns into the
//     leaves (MerkleProverLifted::commit, STWO/prover/vcs_lifted/prover.rs:33-44);
//   - every table (column pointers, pass descriptors) is uploaded once at
//     plan time into plan-owned device memory, so the launch sequence has no
//     host-to-device copies and can be captured into a hipGraph and replayed.
#include <string.h>

#include <algorithm>
#include <functional>
#include <map>
#include <vector>

#include "common.hpp"
#include "field.hpp"
#include "internal.hpp"

struct stwo_hip_batch {
  int device = 0;
  void *d_mem = nullptr;
  std::function<int(stwo_hip_ctx *)> issue;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  size_t n_launches = 0;
};

namespace stwo_hip {

namespace {

// Bump allocator over one host staging vector mirrored by one device buffer.
struct Staging {
  std::vector<uint8_t> host;
  size_t put(const void *src, size_t bytes) {
    size_t off = (host.size() + 255) & ~(size_t)255;
    host.resize(off + bytes);
    memcpy(host.data() + off, src, bytes);
    return off;
  }
};

struct LevelLaunch {
  bool inverse;
  bool fused;  // the fused IFFT-last + RFFT-first pass (groups_off: IFFT descriptors, fwd_off: RFFT ones)
  int kind;  // launch_fft_multi's (or launch_lde_multi's) schedule kind for every group of the launch
  size_t groups_off, fwd_off, wgs_off;  // offsets into the device buffer
  uint32_t n_groups, total_wgs, wg_log;
};

}  // namespace

}  // namespace stwo_hip

using namespace stwo_hip;

extern "C" {

int stwo_hip_batch_create(stwo_hip_ctx *ctx, uint32_t *const *cols, const uint32_t *log_sizes, size_t n_cols,
                          uint32_t log_blowup, const stwo_hip_twiddles *tw, uint32_t *const *outs, uint32_t *layers,
                          int tree_kind, int hash_mode, int flags, stwo_hip_batch **out) {
  const int use_graph = flags & STWO_HIP_BATCH_GRAPH;
  if (!ctx || !out) return STWO_HIP_EINVAL;
  *out = nullptr;
  STWO_HIP_CHECK_ARG(ctx, n_cols > 0 && cols && log_sizes && outs && layers, "batch_create: NULL or empty arguments");
  STWO_HIP_CHECK_ARG(ctx, tw && tw->device == ctx->device, "batch_create: twiddles missing or from another device");
  STWO_HIP_CHECK_ARG(ctx, tree_kind == STWO_HIP_TREE_MIXED || tree_kind == STWO_HIP_TREE_LIFTED,
                     "batch_create: unknown tree kind %d", tree_kind);
  STWO_HIP_CHECK_ARG(ctx, hash_mode == STWO_HIP_HASH_BLAKE2S || hash_mode == STWO_HIP_HASH_BLAKE2S_M31,
                     "batch_create: unknown hash mode %d", hash_mode);
  uint32_t max_log = 0;
  for (size_t i = 0; i < n_cols; i++) {
    STWO_HIP_CHECK_ARG(ctx, cols[i] && outs[i], "batch_create: NULL column %zu", i);
    STWO_HIP_CHECK_ARG(ctx, log_sizes[i] >= 4 && log_sizes[i] <= 28 && log_blowup <= 24 && log_sizes[i] + log_blowup <= 28,
                       "batch_create: column %zu log size %u outside [4, 28 - blowup]", i, log_sizes[i]);
    max_log = std::max(max_log, log_sizes[i] + log_blowup);
  }
  // Every CanonicCoset(k).circle_domain(), k <= max_log, must be a doubling of the root.
  uint32_t rl = tw->root.log_size;
  STWO_HIP_CHECK_ARG(ctx, max_log - 1 <= rl && tw->root.initial_index == half_odds_initial(rl),
                     "batch_create: twiddle root must be half_odds(r) with r >= %u", max_log - 1);
  // The radix-16 passes take the doubled trees.
  const uint32_t *itw_end = tw->d_itw2 + tw->len, *tw_end = tw->d_tw2 + tw->len;

  // Groups by log size, largest first (their workgroups start first).
  std::map<uint32_t, std::vector<size_t>, std::greater<uint32_t>> groups;
  for (size_t i = 0; i < n_cols; i++) groups[log_sizes[i]].push_back(i);

  Staging st;
  struct GroupTabs {
    uint32_t n;
    size_t ifft_off, rfft_off, inplace_off;
    uint32_t count;
    bool aligned16;  // every column pointer of the group is 16-byte aligned
  };
  std::vector<GroupTabs> gt;
  for (auto &g : groups) {
    std::vector<FftCol> a, b, c;
    for (size_t i : g.second) {
      a.push_back({cols[i], cols[i]});
      b.push_back({cols[i], outs[i]});
      c.push_back({outs[i], outs[i]});
    }
    GroupTabs t;
    t.aligned16 = true;
    for (size_t i : g.second)
      if ((((uintptr_t)cols[i]) | ((uintptr_t)outs[i])) & 15) t.aligned16 = false;
    t.n = g.first;
    t.count = (uint32_t)g.second.size();
    t.ifft_off = st.put(a.data(), a.size() * sizeof(FftCol));
    t.rfft_off = st.put(b.data(), b.size() * sizeof(FftCol));
    t.inplace_off = st.put(c.data(), c.size() * sizeof(FftCol));
    gt.push_back(t);
  }
  // Merkle column tables.
  struct LayerTab {
    uint32_t log_size;
    size_t off;
    uint32_t count;
  };
  std::vector<LayerTab> lt;
  if (tree_kind == STWO_HIP_TREE_MIXED) {
    for (int l = (int)max_log; l >= 0; l--) {
      std::vector<ColRef> refs;
      for (size_t i = 0; i < n_cols; i++)
        if (log_sizes[i] + log_blowup == (uint32_t)l) refs.push_back({outs[i], 0, 0});
      if (!refs.empty()) lt.push_back({(uint32_t)l, st.put(refs.data(), refs.size() * sizeof(ColRef)),
                                       (uint32_t)refs.size()});
    }
  } else {
    std::vector<size_t> order(n_cols);
    for (size_t i = 0; i < n_cols; i++) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) { return log_sizes[a] < log_sizes[b]; });
    std::vector<ColRef> refs;
    for (size_t i : order) refs.push_back({outs[i], max_log - (log_sizes[i] + log_blowup), 0});
    lt.push_back({max_log, st.put(refs.data(), refs.size() * sizeof(ColRef)), (uint32_t)refs.size()});
  }
  // Reserve the pass descriptor area after the tables: the descriptors hold
  // device addresses of the tables, known once the buffer exists.
  size_t tables_bytes = (st.host.size() + 255) & ~(size_t)255;
  void *dmem = nullptr;
  // Upper bound on descriptors: one per group and pass (2 directions, at most
  // 4 passes), plus the 256-byte alignment of two tables per launch (2
  // directions x 4 levels x 2 workgroup sizes x at most 14 schedule kinds).
  size_t desc_bytes = 2 * 4 * gt.size() * (sizeof(FftPass2) + 8) + 2 * 4 * 2 * 14 * 2 * 256 + 8192;
  if (hipMalloc(&dmem, tables_bytes + desc_bytes) != hipSuccess) {
    (void)hipGetLastError();
    return set_error(ctx, STWO_HIP_ENOMEM, "batch_create: out of device memory for tables");
  }
  uint8_t *dbase = (uint8_t *)dmem;
  // Both transforms' passes of every group; where the IFFT's last pass and
  // the RFFT's first pass tile alike they become one fused pass.
  struct GroupPasses {
    std::vector<PassLaunch> inv, fwd;
    bool fused = false;
  };
  std::vector<GroupPasses> gp(gt.size());
  for (size_t gi = 0; gi < gt.size(); gi++) {
    const GroupTabs &t = gt[gi];
    for (int dir = 0; dir < 2; dir++) {
      const bool inverse = dir == 0;
      const uint32_t n = t.n, N = t.n + (inverse ? 0 : log_blowup);
      const FftCol *first = (const FftCol *)(dbase + (inverse ? t.ifft_off : t.rfft_off));
      const FftCol *inpl = (const FftCol *)(dbase + t.inplace_off);
      int rc = make_group_passes(ctx, inverse ? N : n, N, inverse, first, inpl, t.count, inverse ? itw_end : tw_end,
                                 t.aligned16, inverse ? gp[gi].inv : gp[gi].fwd);
      if (rc) {
        (void)hipFree(dmem);
        return rc;
      }
    }
    gp[gi].fused = lde_fusable(gp[gi].inv.back(), gp[gi].fwd.front(), log_blowup);
  }
  std::vector<LevelLaunch> levels;
  auto add_level = [&](bool inverse, uint32_t wg_log, std::vector<PassLaunch> &v) -> int {
    // Groups whose pass matches a compile-time schedule get a launch per
    // schedule with that schedule's kernel; the rest share one.
    std::vector<uint32_t> kinds;
    for (auto &pl : v)
      if (std::find(kinds.begin(), kinds.end(), pl.p.fixed) == kinds.end()) kinds.push_back(pl.p.fixed);
    for (uint32_t kd : kinds) {
      std::vector<FftPass2> descs;
      std::vector<uint32_t> starts;
      uint64_t acc = 0;
      bool uniform = true;
      for (auto &pl : v) {
        if (pl.p.fixed != kd) continue;
        descs.push_back(pl.p);
        starts.push_back((uint32_t)acc);
        acc += pl.wgs;
        if (pl.p.T < 10) uniform = false;
      }
      if (descs.empty()) continue;
      if (acc >= (1ull << 31)) return set_error(ctx, STWO_HIP_EINVAL, "batch_create: grid too large");
      LevelLaunch ll{};
      ll.inverse = inverse;
      ll.fused = false;
      ll.kind = kd != 0 ? (int)kd : (uniform ? -1 : 0);
      ll.groups_off = st.put(descs.data(), descs.size() * sizeof(FftPass2));
      ll.wgs_off = st.put(starts.data(), starts.size() * sizeof(uint32_t));
      ll.n_groups = (uint32_t)descs.size();
      ll.total_wgs = (uint32_t)acc;
      ll.wg_log = wg_log;
      levels.push_back(ll);
    }
    return STWO_HIP_OK;
  };
  for (int dir = 0; dir < 2; dir++) {
    const bool inverse = dir == 0;
    // (level, workgroup size) -> passes of every group (the fused ones aside)
    std::map<std::pair<uint32_t, uint32_t>, std::vector<PassLaunch>> by_level;
    for (const GroupPasses &g : gp) {
      const std::vector<PassLaunch> &passes = inverse ? g.inv : g.fwd;
      const size_t k0 = (!inverse && g.fused) ? 1 : 0, k1 = passes.size() - ((inverse && g.fused) ? 1 : 0);
      for (size_t k = k0; k < k1; k++) {
        uint32_t level = inverse ? (uint32_t)k : (uint32_t)(passes.size() - 1 - k);
        by_level[{level, passes[k].wg_log}].push_back(passes[k]);
      }
    }
    if (!inverse) {
      // the fused passes run after every IFFT level and before every RFFT one
      std::map<std::pair<uint32_t, int>, std::vector<size_t>> fused;  // (wg_log, kind) -> groups
      for (size_t gi = 0; gi < gp.size(); gi++)
        if (gp[gi].fused)
          fused[{gp[gi].inv.back().wg_log, lde_kind(gp[gi].inv.back(), gp[gi].fwd.front())}].push_back(gi);
      for (auto &kv : fused) {
        std::vector<FftPass2> pis, pfs;
        std::vector<uint32_t> starts;
        uint64_t acc = 0;
        for (size_t gi : kv.second) {
          pis.push_back(gp[gi].inv.back().p);
          pfs.push_back(gp[gi].fwd.front().p);
          starts.push_back((uint32_t)acc);
          acc += gp[gi].inv.back().wgs;
        }
        if (acc >= (1ull << 31)) {
          (void)hipFree(dmem);
          return set_error(ctx, STWO_HIP_EINVAL, "batch_create: grid too large");
        }
        LevelLaunch ll{};
        ll.inverse = false;
        ll.fused = true;
        ll.kind = kv.first.second;
        ll.groups_off = st.put(pis.data(), pis.size() * sizeof(FftPass2));
        ll.fwd_off = st.put(pfs.data(), pfs.size() * sizeof(FftPass2));
        ll.wgs_off = st.put(starts.data(), starts.size() * sizeof(uint32_t));
        ll.n_groups = (uint32_t)pis.size();
        ll.total_wgs = (uint32_t)acc;
        ll.wg_log = kv.first.first;
        levels.push_back(ll);
      }
    }
    std::vector<std::pair<uint32_t, uint32_t>> order;
    for (auto &kv : by_level) order.push_back(kv.first);
    if (!inverse) std::reverse(order.begin(), order.end());  // forward: highest level first
    for (const auto &key : order) {
      int rc = add_level(inverse, key.second, by_level[key]);
      if (rc) {
        (void)hipFree(dmem);
        return rc;
      }
    }
  }
  if (st.host.size() > tables_bytes + desc_bytes) {
    (void)hipFree(dmem);
    return set_error(ctx, STWO_HIP_EINVAL, "internal: batch descriptor area overflow");
  }
  if (hipMemcpy(dmem, st.host.data(), st.host.size(), hipMemcpyHostToDevice) != hipSuccess) {
    (void)hipFree(dmem);
    return set_error(ctx, STWO_HIP_EDEVICE, "batch_create: table upload failed");
  }

  stwo_hip_batch *b = new (std::nothrow) stwo_hip_batch();
  if (!b) {
    (void)hipFree(dmem);
    return set_error(ctx, STWO_HIP_ENOMEM, "host allocation failed");
  }
  b->device = ctx->device;
  b->d_mem = dmem;
  const int mode = hash_mode;
  const int kind = tree_kind;
  const uint32_t keep_coeffs = (flags & STWO_HIP_BATCH_DROP_COEFFS) ? 0u : 1u;
  b->issue = [levels, lt, dbase, layers, max_log, mode, kind, keep_coeffs](stwo_hip_ctx *c) -> int {
    for (const LevelLaunch &ll : levels) {
      int rc = ll.fused ? launch_lde_multi(c, (const FftPass2 *)(dbase + ll.groups_off),
                                           (const FftPass2 *)(dbase + ll.fwd_off),
                                           (const uint32_t *)(dbase + ll.wgs_off), ll.n_groups, ll.total_wgs,
                                           ll.wg_log, ll.kind, keep_coeffs)
                        : launch_fft_multi(c, ll.inverse, (const FftPass2 *)(dbase + ll.groups_off),
                                           (const uint32_t *)(dbase + ll.wgs_off), ll.n_groups, ll.total_wgs,
                                           ll.wg_log, ll.kind);
      if (rc) return rc;
    }
    if (kind == STWO_HIP_TREE_LIFTED) {
      const LayerTab &t = lt[0];
      int rc = launch_merkle_layer(c, nullptr, (const ColRef *)(dbase + t.off), t.count, max_log,
                                   layers + tree_layer_off(max_log), mode, true);
      if (rc) return rc;
      return hash_node_layers(c, layers, max_log, 0, mode);
    }
    // mixed degree: layers max_log .. 0; runs of pure node layers are fused.
    int l = (int)max_log;
    size_t ti = 0;
    while (l >= 0) {
      const uint32_t *prev = (l == (int)max_log) ? nullptr : layers + tree_layer_off((uint32_t)l + 1);
      if (ti < lt.size() && lt[ti].log_size == (uint32_t)l) {
        const LayerTab &t = lt[ti++];
        int rc = launch_merkle_layer(c, prev, (const ColRef *)(dbase + t.off), t.count, (uint32_t)l,
                                     layers + tree_layer_off((uint32_t)l), mode, false);
        if (rc) return rc;
        l--;
        continue;
      }
      int next = ti < lt.size() ? (int)lt[ti].log_size : -1;
      int rc = hash_node_layers(c, layers, (uint32_t)l + 1, (uint32_t)(next + 1), mode);
      if (rc) return rc;
      l = next;
    }
    return STWO_HIP_OK;
  };
  b->n_launches = levels.size();
  if (use_graph) {
    // Tables of earlier calls still pending are copied outside the graph.
    int frc = flush_uploads(ctx);
    if (frc) {
      (void)hipFree(dmem);
      delete b;
      return frc;
    }
    bool prof = ctx->prof;
    ctx->prof = false;  // no event records inside the captured sequence
    hipError_t e = hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal);
    int rc = e == hipSuccess ? b->issue(ctx) : STWO_HIP_EDEVICE;
    hipGraph_t g = nullptr;
    hipError_t e2 = hipStreamEndCapture(ctx->stream, &g);
    ctx->prof = prof;
    if (rc == STWO_HIP_OK && e2 == hipSuccess && g) {
      if (hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0) != hipSuccess) rc = STWO_HIP_EDEVICE;
      b->graph = g;
    } else if (rc == STWO_HIP_OK) {
      rc = STWO_HIP_EDEVICE;
    }
    if (rc) {
      if (g) (void)hipGraphDestroy(g);
      (void)hipFree(dmem);
      delete b;
      (void)hipGetLastError();
      return set_error(ctx, rc, "batch_create: graph capture failed (%s)",
                       ctx->last_error.empty() ? "hip" : ctx->last_error.c_str());
    }
  }
  *out = b;
  return STWO_HIP_OK;
}

int stwo_hip_batch_run(stwo_hip_ctx *ctx, stwo_hip_batch *b) {
  if (!ctx || !b) return STWO_HIP_EINVAL;
  STWO_HIP_CHECK_ARG(ctx, b->device == ctx->device, "batch_run: plan belongs to another device");
  if (b->exec) {
    hipEvent_t start = nullptr;
    prof_begin(ctx, "batch_graph", &start);
    STWO_HIP_TRY(ctx, hipGraphLaunch(b->exec, ctx->stream));
    prof_end(ctx, "batch_graph", start);
    return STWO_HIP_OK;
  }
  return b->issue(ctx);
}

int stwo_hip_batch_destroy(stwo_hip_ctx *ctx, stwo_hip_batch *b) {
  if (!ctx || !b) return STWO_HIP_EINVAL;
  (void)hipStreamSynchronize(ctx->stream);
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  if (b->d_mem) (void)hipFree(b->d_mem);
  delete b;
  return STWO_HIP_OK;
}

}  // extern "C"

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
#include <tuple>
#include <vector>

#include "common.hpp"
#include "field.hpp"
#include "internal.hpp"

struct stwo_hip_batch {
  int device = 0;
  void *d_mem = nullptr;
  bool mem_from_job = false;  // tables carved from a serving job's memory block
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

// One tree's Merkle launches: its layer tables and shape.
struct TreeLaunches {
  struct LayerTab {
    uint32_t log_size;
    size_t off;
    uint32_t count;
  };
  std::vector<LayerTab> lt;
  uint32_t *layers;
  uint32_t max_log;
  int mode, kind;
};

}  // namespace

int batch_create_trees(stwo_hip_ctx *ctx, const BatchTree *trees, size_t n_trees, const stwo_hip_twiddles *tw,
                       int flags, stwo_hip_batch **out) {
  const int use_graph = flags & STWO_HIP_BATCH_GRAPH;
  if (!ctx || !out) return STWO_HIP_EINVAL;
  *out = nullptr;
  STWO_HIP_CHECK_ARG(ctx, n_trees > 0 && trees, "batch_create: no column sets");
  STWO_HIP_CHECK_ARG(ctx, tw && tw->device == ctx->device, "batch_create: twiddles missing or from another device");
  uint32_t all_max = 0;
  for (size_t t = 0; t < n_trees; t++) {
    const BatchTree &T = trees[t];
    STWO_HIP_CHECK_ARG(ctx, T.stage == BATCH_LDE || T.stage == BATCH_IFFT || T.stage == BATCH_RFFT,
                       "batch_create: unknown stage %d (set %zu)", T.stage, t);
    const bool has_outs = T.stage != BATCH_IFFT, has_tree = T.stage == BATCH_LDE || T.layers;
    STWO_HIP_CHECK_ARG(ctx, T.n_cols > 0 && T.cols && T.log_sizes && (T.outs || !has_outs) &&
                                (T.layers || T.stage != BATCH_LDE),
                       "batch_create: NULL or empty arguments (set %zu)", t);
    if (has_tree) {
      STWO_HIP_CHECK_ARG(ctx, T.tree_kind == STWO_HIP_TREE_MIXED || T.tree_kind == STWO_HIP_TREE_LIFTED,
                         "batch_create: unknown tree kind %d", T.tree_kind);
      STWO_HIP_CHECK_ARG(ctx, T.hash_mode == STWO_HIP_HASH_BLAKE2S || T.hash_mode == STWO_HIP_HASH_BLAKE2S_M31,
                         "batch_create: unknown hash mode %d", T.hash_mode);
    }
    const uint32_t b = T.stage == BATCH_IFFT ? 0 : T.log_blowup;
    for (size_t i = 0; i < T.n_cols; i++) {
      STWO_HIP_CHECK_ARG(ctx, T.cols[i] && (!has_outs || T.outs[i]), "batch_create: NULL column %zu", i);
      STWO_HIP_CHECK_ARG(ctx, T.log_sizes[i] >= 4 && T.log_sizes[i] <= 28 && b <= 24 && T.log_sizes[i] + b <= 28,
                         "batch_create: column %zu log size %u outside [4, 28 - blowup]", i, T.log_sizes[i]);
      all_max = std::max(all_max, T.log_sizes[i] + b);
    }
  }
  // Every CanonicCoset(k).circle_domain(), k <= max_log, must be a doubling of the root.
  uint32_t rl = tw->root.log_size;
  STWO_HIP_CHECK_ARG(ctx, all_max - 1 <= rl && tw->root.initial_index == half_odds_initial(rl),
                     "batch_create: twiddle root must be half_odds(r) with r >= %u", all_max - 1);
  // The radix-16 passes take the doubled trees.
  const uint32_t *itw_end = tw->d_itw2 + tw->len, *tw_end = tw->d_tw2 + tw->len;

  // Groups by (log size, blowup, stage) over every set, largest first (their
  // workgroups start first): columns of different sets share launches.
  std::map<std::tuple<uint32_t, uint32_t, int>, std::vector<std::pair<size_t, size_t>>,
           std::greater<std::tuple<uint32_t, uint32_t, int>>> groups;
  for (size_t t = 0; t < n_trees; t++)
    for (size_t i = 0; i < trees[t].n_cols; i++)
      groups[{trees[t].log_sizes[i], trees[t].stage == BATCH_IFFT ? 0u : trees[t].log_blowup, trees[t].stage}]
          .push_back({t, i});

  Staging st;
  struct GroupTabs {
    uint32_t n, b;
    int stage;
    size_t ifft_off, rfft_off, inplace_off;
    uint32_t count;
    bool aligned16;  // every column pointer of the group is 16-byte aligned
  };
  std::vector<GroupTabs> gt;
  for (auto &g : groups) {
    std::vector<FftCol> a, bb, c;
    GroupTabs t;
    t.aligned16 = true;
    t.stage = std::get<2>(g.first);
    for (auto &ti : g.second) {
      uint32_t *col = trees[ti.first].cols[ti.second];
      uint32_t *o = t.stage == BATCH_IFFT ? col : trees[ti.first].outs[ti.second];
      a.push_back({col, col});
      bb.push_back({col, o});
      c.push_back({o, o});
      if ((((uintptr_t)col) | ((uintptr_t)o)) & 15) t.aligned16 = false;
    }
    t.n = std::get<0>(g.first);
    t.b = std::get<1>(g.first);
    t.count = (uint32_t)g.second.size();
    t.ifft_off = st.put(a.data(), a.size() * sizeof(FftCol));
    t.rfft_off = st.put(bb.data(), bb.size() * sizeof(FftCol));
    t.inplace_off = st.put(c.data(), c.size() * sizeof(FftCol));
    gt.push_back(t);
  }
  // Merkle column tables, per set.
  std::vector<TreeLaunches> tl;
  for (size_t t = 0; t < n_trees; t++) {
    const BatchTree &T = trees[t];
    if (T.stage == BATCH_IFFT || !T.layers) continue;
    tl.emplace_back();
    TreeLaunches &L = tl.back();
    L.layers = T.layers;
    L.mode = T.hash_mode;
    L.kind = T.tree_kind;
    L.max_log = 0;
    for (size_t i = 0; i < T.n_cols; i++) L.max_log = std::max(L.max_log, T.log_sizes[i] + T.log_blowup);
    if (T.tree_kind == STWO_HIP_TREE_MIXED) {
      for (int l = (int)L.max_log; l >= 0; l--) {
        std::vector<ColRef> refs;
        for (size_t i = 0; i < T.n_cols; i++)
          if (T.log_sizes[i] + T.log_blowup == (uint32_t)l) refs.push_back({T.outs[i], 0, 0});
        if (!refs.empty())
          L.lt.push_back({(uint32_t)l, st.put(refs.data(), refs.size() * sizeof(ColRef)), (uint32_t)refs.size()});
      }
    } else {
      std::vector<size_t> order(T.n_cols);
      for (size_t i = 0; i < T.n_cols; i++) order[i] = i;
      std::stable_sort(order.begin(), order.end(),
                       [&](size_t a, size_t b) { return T.log_sizes[a] < T.log_sizes[b]; });
      std::vector<ColRef> refs;
      for (size_t i : order) refs.push_back({T.outs[i], L.max_log - (T.log_sizes[i] + T.log_blowup), 0});
      L.lt.push_back({L.max_log, st.put(refs.data(), refs.size() * sizeof(ColRef)), (uint32_t)refs.size()});
    }
  }
  // Reserve the pass descriptor area after the tables: the descriptors hold
  // device addresses of the tables, known once the buffer exists.
  size_t tables_bytes = (st.host.size() + 255) & ~(size_t)255;
  void *dmem = nullptr;
  bool from_job = false;
  // Upper bound on descriptors: one per group and pass (2 directions, at most
  // 4 passes), plus the 256-byte alignment of two or three tables per launch
  // (2 directions x 4 levels x 2 workgroup sizes x at most 14 schedule kinds,
  // and the fused launches).
  size_t desc_bytes = 2 * 4 * gt.size() * (sizeof(FftPass2) + 8) + 2 * 4 * 2 * 14 * 2 * 256 + 8192;
  int arc = device_alloc(ctx, tables_bytes + desc_bytes, &dmem, &from_job);
  if (arc) return arc;
  auto release = [&]() { device_free(dmem, from_job); };
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
      if ((inverse && t.stage == BATCH_RFFT) || (!inverse && t.stage == BATCH_IFFT)) continue;
      const uint32_t n = t.n, N = t.n + (inverse ? 0 : t.b);
      const FftCol *first = (const FftCol *)(dbase + (inverse ? t.ifft_off : t.rfft_off));
      const FftCol *inpl = (const FftCol *)(dbase + t.inplace_off);
      int rc = make_group_passes(ctx, inverse ? N : n, N, inverse, first, inpl, t.count, inverse ? itw_end : tw_end,
                                 t.aligned16, inverse ? gp[gi].inv : gp[gi].fwd);
      if (rc) {
        release();
        return rc;
      }
    }
    gp[gi].fused = t.stage == BATCH_LDE && lde_fusable(gp[gi].inv.back(), gp[gi].fwd.front(), t.b);
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
          release();
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
        release();
        return rc;
      }
    }
  }
  if (st.host.size() > tables_bytes + desc_bytes) {
    release();
    return set_error(ctx, STWO_HIP_EINVAL, "internal: batch descriptor area overflow");
  }
  if (from_job) {
    // within a job: an ordered copy on the job's stream
    int rc = flush_uploads(ctx);
    if (rc) return rc;
    STWO_HIP_TRY(ctx, hipMemcpyAsync(dmem, st.host.data(), st.host.size(), hipMemcpyHostToDevice, ctx->stream));
    STWO_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  } else if (hipMemcpy(dmem, st.host.data(), st.host.size(), hipMemcpyHostToDevice) != hipSuccess) {
    release();
    return set_error(ctx, STWO_HIP_EDEVICE, "batch_create: table upload failed");
  }

  stwo_hip_batch *b = new (std::nothrow) stwo_hip_batch();
  if (!b) {
    release();
    return set_error(ctx, STWO_HIP_ENOMEM, "host allocation failed");
  }
  b->device = ctx->device;
  b->d_mem = dmem;
  b->mem_from_job = from_job;
  const uint32_t keep_coeffs = (flags & STWO_HIP_BATCH_DROP_COEFFS) ? 0u : 1u;
  b->issue = [levels, tl, dbase, keep_coeffs](stwo_hip_ctx *c) -> int {
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
    for (const TreeLaunches &L : tl) {
      uint32_t *layers = L.layers;
      const uint32_t max_log = L.max_log;
      if (L.kind == STWO_HIP_TREE_LIFTED) {
        const auto &t = L.lt[0];
        int rc = launch_merkle_layer(c, nullptr, (const ColRef *)(dbase + t.off), t.count, max_log,
                                     layers + tree_layer_off(max_log), L.mode, true);
        if (rc) return rc;
        rc = hash_node_layers(c, layers, max_log, 0, L.mode);
        if (rc) return rc;
        continue;
      }
      // mixed degree: layers max_log .. 0; runs of pure node layers are fused.
      int l = (int)max_log;
      size_t ti = 0;
      while (l >= 0) {
        const uint32_t *prev = (l == (int)max_log) ? nullptr : layers + tree_layer_off((uint32_t)l + 1);
        if (ti < L.lt.size() && L.lt[ti].log_size == (uint32_t)l) {
          const auto &t = L.lt[ti++];
          int rc = launch_merkle_layer(c, prev, (const ColRef *)(dbase + t.off), t.count, (uint32_t)l,
                                       layers + tree_layer_off((uint32_t)l), L.mode, false);
          if (rc) return rc;
          l--;
          continue;
        }
        int next = ti < L.lt.size() ? (int)L.lt[ti].log_size : -1;
        int rc = hash_node_layers(c, layers, (uint32_t)l + 1, (uint32_t)(next + 1), L.mode);
        if (rc) return rc;
        l = next;
      }
    }
    return STWO_HIP_OK;
  };
  b->n_launches = levels.size();
  if (use_graph) {
    // Tables of earlier calls still pending are copied outside the graph.
    int frc = flush_uploads(ctx);
    if (frc) {
      release();
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
      release();
      delete b;
      (void)hipGetLastError();
      return set_error(ctx, rc, "batch_create: graph capture failed (%s)",
                       ctx->last_error.empty() ? "hip" : ctx->last_error.c_str());
    }
  }
  *out = b;
  return STWO_HIP_OK;
}

}  // namespace stwo_hip

using namespace stwo_hip;

extern "C" {

int stwo_hip_batch_create(stwo_hip_ctx *ctx, uint32_t *const *cols, const uint32_t *log_sizes, size_t n_cols,
                          uint32_t log_blowup, const stwo_hip_twiddles *tw, uint32_t *const *outs, uint32_t *layers,
                          int tree_kind, int hash_mode, int flags, stwo_hip_batch **out) {
  const BatchTree t = {cols, log_sizes, n_cols, log_blowup, outs, layers, tree_kind, hash_mode};
  return batch_create_trees(ctx, &t, 1, tw, flags, out);
}

int stwo_hip_batch_create_multi(stwo_hip_ctx *ctx, const stwo_hip_batch_set *sets, size_t n_sets,
                                const stwo_hip_twiddles *tw, int flags, stwo_hip_batch **out) {
  if (!ctx) return STWO_HIP_EINVAL;
  STWO_HIP_CHECK_ARG(ctx, sets && n_sets, "batch_create_multi: no column sets");
  std::vector<BatchTree> t(n_sets);
  for (size_t i = 0; i < n_sets; i++)
    t[i] = {sets[i].cols,   sets[i].log_sizes, sets[i].n_cols,    sets[i].log_blowup, sets[i].outs,
            sets[i].layers, sets[i].tree_kind, sets[i].hash_mode, sets[i].stage};
  return batch_create_trees(ctx, t.data(), n_sets, tw, flags, out);
}

int stwo_hip_lde_commit_per_column(stwo_hip_ctx *ctx, uint32_t *const *cols, const uint32_t *log_sizes,
                                   size_t n_cols, uint32_t log_blowup, const stwo_hip_twiddles *tw,
                                   uint32_t *const *outs, uint32_t *layers, int tree_kind, int hash_mode) {
  if (!ctx) return STWO_HIP_EINVAL;
  STWO_HIP_CHECK_ARG(ctx, n_cols > 0 && cols && log_sizes && outs && layers, "lde_commit_per_column: NULL arguments");
  STWO_HIP_CHECK_ARG(ctx, tree_kind == STWO_HIP_TREE_MIXED || tree_kind == STWO_HIP_TREE_LIFTED,
                     "lde_commit_per_column: unknown tree kind %d", tree_kind);
  uint32_t max_log = 0;
  std::vector<uint32_t> out_logs(n_cols);
  for (size_t i = 0; i < n_cols; i++) {
    STWO_HIP_CHECK_ARG(ctx, log_sizes[i] >= 1 && log_sizes[i] + log_blowup <= 28,
                       "lde_commit_per_column: column %zu log size %u", i, log_sizes[i]);
    const uint32_t n = log_sizes[i], N = n + log_blowup;
    // one interpolate and one evaluate call per column, as the reference's
    // plain path submits them
    int rc = stwo_hip_interpolate(ctx, cols[i], {half_odds_initial(n - 1), n - 1}, tw);
    if (rc) return rc;
    rc = stwo_hip_evaluate(ctx, cols[i], n, {half_odds_initial(N - 1), N - 1}, tw, outs[i]);
    if (rc) return rc;
    out_logs[i] = N;
    max_log = std::max(max_log, N);
  }
  std::vector<const uint32_t *> oc(outs, outs + n_cols);
  if (tree_kind == STWO_HIP_TREE_MIXED) return stwo_hip_merkle_commit(ctx, oc.data(), out_logs.data(), n_cols, layers,
                                                                      hash_mode);
  return stwo_hip_merkle_lifted_commit(ctx, oc.data(), out_logs.data(), n_cols, max_log, layers, hash_mode);
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
  device_free(b->d_mem, b->mem_from_job);
  delete b;
  return STWO_HIP_OK;
}

}  // extern "C"

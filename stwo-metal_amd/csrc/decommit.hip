// Merkle decommitment of a lifted tree (MerkleProverLifted::decommit,
// STWO/prover/vcs_lifted/prover.rs:58-201) and the batched device gather it
// and the FRI decommitment share.
//
// The traversal plan depends only on the query positions, so the host plans
// every layer first (a few integer operations per query and layer), then one
// kernel gathers every needed value and hash (queried column values and both
// children of every visited parent, from all layers at once), and one copy
// brings them back.  This is the reference's per-layer `batch_at` reads
// (and the Metal backend's single bulk gather, MB/decommit.rs:96-220) as one
// launch and one synchronisation per decommitment.
#include <string.h>

#include <algorithm>
#include <vector>

#include "channel.hpp"
#include "common.hpp"
#include "field.hpp"
#include "internal.hpp"

namespace stwo_hip {

// Gather op: copy n words from src to out[dst..].  src already points at the
// first word (device addresses are computed on the host).
struct GatherOp {
  const uint32_t *src;
  uint32_t n;
  uint32_t dst;
};
static_assert(sizeof(GatherOp) == 16, "GatherOp layout");

__global__ void __launch_bounds__(256) gather_kernel(const GatherOp *ops, uint32_t n_ops, uint32_t *out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_ops) return;
  const GatherOp o = ops[i];
  for (uint32_t w = 0; w < o.n; w++) as_global(out)[o.dst + w] = as_global(o.src)[w];
}

void Gatherer::reserve(size_t more_ops) { ops_.reserve(ops_.size() + more_ops); }

size_t Gatherer::add(const uint32_t *src, uint32_t n) {
  size_t off = words;
  ops_.push_back({(uintptr_t)src, n, (uint32_t)off});
  words += n;
  return off;
}

int Gatherer::run(stwo_hip_ctx *ctx) {
  out.assign(words, 0);
  if (ops_.empty()) return STWO_HIP_OK;
  STWO_HIP_CHECK_ARG(ctx, words < (1ull << 32) && ops_.size() < (1ull << 32), "decommit: gather too large");
  const size_t ops_bytes = ops_.size() * sizeof(GatherOp), out_off = (ops_bytes + 255) & ~(size_t)255;
  if (ctx->h_gather && out_off + words * 4 <= ctx->gather_cap) {
    // The kernel reads the op table from pinned host memory and writes the
    // result there: no copy on the stream in either direction.
    GatherOp *h_ops = (GatherOp *)ctx->h_gather;
    for (size_t i = 0; i < ops_.size(); i++) h_ops[i] = {(const uint32_t *)ops_[i].src, ops_[i].n, ops_[i].dst};
    const uint32_t n_ops = (uint32_t)ops_.size();
    int rc = launch(ctx, "decommit_gather", gather_kernel, dim3((n_ops + 255) / 256), dim3(256), 0,
                    (const GatherOp *)ctx->d_gather, n_ops, (uint32_t *)((uint8_t *)ctx->d_gather + out_off));
    if (rc) return rc;
    STWO_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    memcpy(out.data(), ctx->h_gather + out_off, words * 4);
    return STWO_HIP_OK;
  }
  std::vector<GatherOp> ops(ops_.size());
  for (size_t i = 0; i < ops_.size(); i++)
    ops[i] = {(const uint32_t *)ops_[i].src, ops_[i].n, ops_[i].dst};
  // The op table goes through the pinned upload ring when it fits (stream
  // order keeps it valid for the kernel), the result through the pinned
  // staging buffer: one synchronization per gather.
  void *d_ops = nullptr, *d_out = nullptr;
  int rc;
  if (ops_bytes <= ctx->arena_cap / 2) {
    rc = upload_table(ctx, ops.data(), ops_bytes, &d_ops);
    if (rc) return rc;
  } else {
    rc = get_scratch(ctx, "gather_ops", ops_bytes, &d_ops);
    if (rc) return rc;
    STWO_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    STWO_HIP_TRY(ctx, hipMemcpy(d_ops, ops.data(), ops_bytes, hipMemcpyHostToDevice));
  }
  rc = get_scratch(ctx, "gather_out", words * 4, &d_out);
  if (rc) return rc;
  const uint32_t n_ops = (uint32_t)ops.size();
  rc = launch(ctx, "decommit_gather", gather_kernel, dim3((n_ops + 255) / 256), dim3(256), 0,
              (const GatherOp *)d_ops, n_ops, (uint32_t *)d_out);
  if (rc) return rc;
  const DownloadReq r = {out.data(), (const uint32_t *)d_out, words};
  return download_many(ctx, &r, 1);
}

// The hash-witness traversal of MerkleProverLifted::decommit
// (vcs_lifted/prover.rs:117-186): start from the deduplicated queries at the
// leaf layer L; per layer, group the sorted queries into sibling pairs
// (chunk_by a ^ 1 == b); each group's parent is queried on the next layer;
// a group with one member needs its sibling as a witness.  Both children of
// every parent are gathered (they are the aux node values; the witness is
// one of them).
void plan_merkle_decommit(Gatherer &g, const uint32_t *tree, uint32_t L, const std::vector<uint32_t> &queries,
                          MerkleDecommitPlan &plan) {
  plan.parents.clear();
  std::vector<uint32_t> prev = queries, cur;
  prev.erase(std::unique(prev.begin(), prev.end()), prev.end());
  // at most one parent per query and level: no reallocation below
  plan.parents.reserve((size_t)L * prev.size());
  cur.reserve(prev.size());
  g.reserve((size_t)L * prev.size());
  for (uint32_t child_layer = L; child_layer >= 1; child_layer--) {
    cur.clear();
    const uint32_t *layer = tree + tree_layer_off(child_layer);
    for (size_t i = 0; i < prev.size();) {
      const uint32_t first = prev[i];
      const bool pair = i + 1 < prev.size() && (first ^ 1) == prev[i + 1];
      const uint32_t parent = first >> 1;
      MerkleDecommitPlan::Parent p;
      p.child_layer = child_layer;
      p.parent = parent;
      p.witness_child = pair ? -1 : (int)((first ^ 1) & 1);
      p.off = g.add(layer + (size_t)16 * parent, 16);  // children 2p and 2p + 1 are adjacent
      plan.parents.push_back(p);
      cur.push_back(parent);
      i += pair ? 2 : 1;
    }
    prev.swap(cur);
  }
}

void finish_merkle_decommit(const Gatherer &g, const MerkleDecommitPlan &plan, std::vector<uint32_t> &witness,
                            std::vector<uint32_t> *aux) {
  witness.clear();
  if (aux) aux->clear();
  for (const MerkleDecommitPlan::Parent &p : plan.parents) {
    const uint32_t *children = g.out.data() + p.off;
    if (p.witness_child >= 0) witness.insert(witness.end(), children + 8 * p.witness_child,
                                             children + 8 * p.witness_child + 8);
    if (aux) {
      for (uint32_t c = 0; c < 2; c++) {
        aux->push_back(p.child_layer);
        aux->push_back(2 * p.parent + c);
        aux->insert(aux->end(), children + 8 * c, children + 8 * c + 8);
      }
    }
  }
}

// The FRI layers' decommitment (FriProver::decommit, fri.rs:294-347): draw
// the query positions from the channel (draw_queries, core/queries.rs:11-28;
// Queries::new sorts and deduplicates), then for every committed layer
// compute_decommitment_positions_and_witness_evals (fri.rs:455-489) with the
// layer's fold step (1 for the circle layer) and the Merkle decommitment of
// the layer's tree on those positions; the queries fold by the layer's step
// between layers (Queries::fold, queries.rs:56-63).
int fri_decommit_plan(stwo_hip_ctx *ctx, stwo_hip_fri_prover *pr, uint8_t channel_digest[32],
                                uint32_t *channel_n_draws, Gatherer &g, FriDecommitPlan &plan) {
  if (!ctx) return STWO_HIP_EINVAL;
  STWO_HIP_CHECK_ARG(ctx, pr && channel_digest && channel_n_draws, "fri_prover_decommit: NULL argument");
  STWO_HIP_CHECK_ARG(ctx, pr->device == ctx->device, "fri_prover_decommit: prover of another device");
  STWO_HIP_CHECK_ARG(ctx, !pr->decommitted, "fri_prover_decommit: already decommitted");
  const uint32_t n_queries = pr->cfg.n_queries;
  STWO_HIP_CHECK_ARG(ctx, n_queries >= 1 && n_queries <= (1u << 20), "fri_prover_decommit: n_queries %u invalid",
                     n_queries);
  HostChannel ch;
  memcpy(ch.digest, channel_digest, 32);
  ch.n_draws = *channel_n_draws;
  const uint32_t N = pr->log_size;
  const uint32_t mask = (uint32_t)((1ull << N) - 1);
  std::vector<uint32_t> &raw = plan.raw;
  raw.clear();
  while (raw.size() < n_queries) {
    uint32_t w[8];
    ch.draw_u32s(w);
    for (int k = 0; k < 8 && raw.size() < n_queries; k++) raw.push_back(w[k] & mask);
  }
  std::vector<uint32_t> &queries = plan.queries;
  queries = raw;
  std::sort(queries.begin(), queries.end());
  queries.erase(std::unique(queries.begin(), queries.end()), queries.end());

  plan.layers.assign(pr->layers.size(), FriDecommitPlan::Layer());
  std::vector<uint32_t> lq = queries, folded;
  folded.reserve(lq.size());
  for (size_t k = 0; k < pr->layers.size(); k++) {
    const FriLayerKeep &layer = pr->layers[k];
    const uint32_t step = layer.shape.fold_step;
    FriDecommitPlan::Layer &pl = plan.layers[k];
    // at most 2^step positions per query
    const size_t max_pos = lq.size() << step;
    pl.positions.reserve(max_pos);
    pl.val_off.reserve(4 * max_pos);
    pl.witness.reserve(max_pos);
    g.reserve(4 * max_pos);
    for (size_t i = 0; i < lq.size();) {
      size_t j = i;
      while (j < lq.size() && (lq[j] >> step) == (lq[i] >> step)) j++;
      const uint32_t start = (lq[i] >> step) << step;
      size_t next = i;
      for (uint32_t pos = start; pos < start + (1u << step); pos++) {
        pl.positions.push_back(pos);
        for (int c = 0; c < 4; c++) pl.val_off.push_back(g.add(layer.cols[c] + pos, 1));
        const bool queried = next < j && lq[next] == pos;
        if (queried) next++;
        pl.witness.push_back(queried ? 0 : 1);
      }
      i = j;
    }
    plan_merkle_decommit(g, layer.tree, layer.shape.log_size, pl.positions, pl.merkle);
    folded.clear();
    for (uint32_t q : lq) {
      uint32_t f = q >> step;
      if (folded.empty() || folded.back() != f) folded.push_back(f);
    }
    lq.swap(folded);
  }
  memcpy(channel_digest, ch.digest, 32);
  *channel_n_draws = ch.n_draws;
  return STWO_HIP_OK;
}

void fri_decommit_finish(stwo_hip_fri_prover *pr, const Gatherer &g, FriDecommitPlan &plan) {
  pr->proofs.assign(pr->layers.size(), FriLayerDecommit());
  for (size_t k = 0; k < plan.layers.size(); k++) {
    const FriDecommitPlan::Layer &pl = plan.layers[k];
    FriLayerDecommit &d = pr->proofs[k];
    d.positions = pl.positions;
    for (size_t i = 0; i < pl.positions.size(); i++) {
      for (int c = 0; c < 4; c++) d.values.push_back(g.out[pl.val_off[4 * i + c]]);
      if (pl.witness[i])
        for (int c = 0; c < 4; c++) d.fri_witness.push_back(g.out[pl.val_off[4 * i + c]]);
    }
    finish_merkle_decommit(g, pl.merkle, d.hash_witness, nullptr);
  }
  pr->query_positions = plan.queries;
  pr->unsorted_queries = plan.raw;
  pr->decommitted = true;
}

}  // namespace stwo_hip

using namespace stwo_hip;

extern "C" {

int stwo_hip_merkle_lifted_decommit(stwo_hip_ctx *ctx, const uint32_t *layers, uint32_t lifting_log_size,
                                    const uint32_t *const *cols, const uint32_t *log_sizes, size_t n_cols,
                                    const uint32_t *queries, size_t n_queries, uint32_t *values,
                                    uint32_t *hash_witness, size_t witness_cap, size_t *n_witness,
                                    uint32_t *aux_nodes, size_t aux_cap, size_t *n_aux) {
  if (!ctx) return STWO_HIP_EINVAL;
  STWO_HIP_CHECK_ARG(ctx, layers && n_witness && (queries || !n_queries) && (cols || !n_cols) &&
                              (log_sizes || !n_cols) && (values || !(n_cols && n_queries)) &&
                              (hash_witness || !witness_cap) && (!aux_nodes || n_aux),
                     "merkle_lifted_decommit: NULL argument");
  const uint32_t L = lifting_log_size;
  STWO_HIP_CHECK_ARG(ctx, L <= 30, "merkle_lifted_decommit: lifting log size %u too large", L);
  for (size_t q = 0; q < n_queries; q++) {
    STWO_HIP_CHECK_ARG(ctx, queries[q] < (1ull << L), "merkle_lifted_decommit: query %u outside 2^%u", queries[q],
                       L);
    STWO_HIP_CHECK_ARG(ctx, q == 0 || queries[q - 1] <= queries[q],
                       "merkle_lifted_decommit: query positions must be in increasing order");
  }
  for (size_t c = 0; c < n_cols; c++)
    STWO_HIP_CHECK_ARG(ctx, cols[c] && log_sizes[c] >= 1 && log_sizes[c] <= L,
                       "merkle_lifted_decommit: column %zu (log %u) invalid for lifting log %u", c, log_sizes[c], L);
  Gatherer g;
  // Queried values (prover.rs:72-110): column c at lift(pos, L - log_size(c)).
  std::vector<size_t> val_off(n_cols * n_queries);
  for (size_t c = 0; c < n_cols; c++) {
    const uint32_t shift = L - log_sizes[c];
    for (size_t q = 0; q < n_queries; q++) {
      const uint32_t pos = queries[q];
      const uint32_t idx = ((pos >> (shift + 1)) << 1) + (pos & 1);
      val_off[c * n_queries + q] = g.add(cols[c] + idx, 1);
    }
  }
  MerkleDecommitPlan plan;
  plan_merkle_decommit(g, layers, L, std::vector<uint32_t>(queries, queries + n_queries), plan);
  int rc = g.run(ctx);
  if (rc) return rc;
  for (size_t i = 0; i < val_off.size(); i++) values[i] = g.out[val_off[i]];
  std::vector<uint32_t> witness, aux;
  finish_merkle_decommit(g, plan, witness, aux_nodes ? &aux : nullptr);
  *n_witness = witness.size() / 8;
  STWO_HIP_CHECK_ARG(ctx, *n_witness <= witness_cap, "merkle_lifted_decommit: %zu witness hashes exceed capacity %zu",
                     *n_witness, witness_cap);
  if (!witness.empty()) memcpy(hash_witness, witness.data(), witness.size() * 4);
  if (aux_nodes) {
    *n_aux = aux.size() / 10;
    STWO_HIP_CHECK_ARG(ctx, *n_aux <= aux_cap, "merkle_lifted_decommit: %zu aux nodes exceed capacity %zu", *n_aux,
                       aux_cap);
    if (!aux.empty()) memcpy(aux_nodes, aux.data(), aux.size() * 4);
  }
  return STWO_HIP_OK;
}

}  // extern "C"

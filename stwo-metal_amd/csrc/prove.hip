// End-to-end STARK prove of the wide-Fibonacci AIR on the device kernels: the
// host flow of the reference's benchmark
// (fixtures/standalone-benchmarks/src/bin/wide_fibonacci_simd_prove.rs
// run_one_sample: Blake2sChannel, Blake2sMerkleChannel, an empty preprocessed
// tree, the trace tree, then stwo's prove) with every bulk step on the GPU.
//
// Steps, in channel order (STWO = vendor/stwo-upstream-dev-62b228e/crates/stwo/src):
//   1. preprocessed tree: no columns, root = Blake2s("") (CPUB/merkle_lifted.rs:44-48),
//      mixed into the channel (CommitmentSchemeProver::commit, STWO/prover/pcs/mod.rs:368-380);
//   2. trace tree: generate_trace, interpolate, evaluate at the blowup, lifted
//      Merkle commit, mix the root (TreeBuilder::extend_evals + commit, :707-735);
//   3. prove_ex (STWO/prover/mod.rs:41-144): draw the composition coefficient,
//      evaluate the constraint quotients on CanonicCoset(log_n + 1), interpolate
//      (DomainEvaluationAccumulator::finalize, STWO/prover/air/accumulation.rs:96-120),
//      split the 4 coordinate polynomials at the middle (CpuBackend::split_at_mid,
//      CPUB/circle.rs:283-291), commit the 8 halves, draw the OODS point
//      (CirclePoint::get_random_point, STWO/core/circle.rs:169-181);
//   4. prove_values (STWO/prover/pcs/mod.rs:528-698): every trace column and
//      composition half sampled at the OODS point, values mixed into the
//      channel, FRI quotients with a fresh coefficient, FRI commit, grind, mix
//      the nonce, FRI decommit (query positions), then each tree decommitted
//      at the query positions (prepare_query_positions_for_tree is the identity
//      for trees at the lifting size and empty for the preprocessed tree,
//      STWO/core/pcs/utils.rs:195-213);
//   5. prove_ex's self-check: the composition value rebuilt from the sampled
//      halves (StarkProof::extract_composition_oods_eval, STWO/core/proof.rs:26-56)
//      must equal the constraints evaluated at the OODS point from the sampled
//      trace values (FrameworkComponent::evaluate_constraint_quotients_at_point,
//      constraint-framework/src/component.rs:271-300; PointEvaluationAccumulator,
//      STWO/core/air/accumulation.rs).
//
// Two sampling paths, as the PCS has them (STWO/prover/pcs/mod.rs:540-655):
// with STWO_HIP_PROVE_STORE_COEFFS (set_store_polynomials_coefficients, what
// the reference's wide-Fibonacci benchmark sets, MB/benchmark.rs:576) the
// coefficients are kept and sampled with eval_at_point; without it (the PCS
// default, store_polynomials_coefficients = false, :347) the trace's LDE pass
// does not store them and every column is sampled from its committed
// evaluation with the barycentric weights of CanonicCoset(log_n + blowup) at
// the OODS point (build_weights_hash_map, :455-510; component_prover.rs:50-67).
// Both give the same field elements, so the proof bytes are the same.  As in
// the reference, the default path needs the blowup to be 1: a larger blowup
// extends the trace to the constraint domain from its coefficients
// (constraint-framework/src/prover/component_prover.rs:56-77, which panics
// without them; here an error).
//
// Outside a serving job the trace commit (trace LDE + lifted tree) and the
// composition commit (interpolation of the 4 coordinate columns, evaluation of
// their 8 halves, lifted tree) each run as a batch plan replayed as a
// hipGraph; the plans are kept on the context for the next prove of the same
// shape and buffers.  Within a job they are issued eagerly.
#include <string.h>

#include <array>
#include <chrono>
#include <memory>
#include <vector>

#include "channel.hpp"
#include "common.hpp"
#include "field.hpp"
#include "internal.hpp"

namespace stwo_hip {

// ---- host secure-field helpers ------------------------------------------------

static qm31 q_from_m31(uint32_t a) { return qm31_make(a, 0, 0, 0); }

static qm31 qm31_inv_host(const qm31 &v) { return qm31_inv(v); }

// CirclePoint::double_x: 2x^2 - 1.
static qm31 q_double_x(const qm31 &x) {
  const qm31 sq = qm31_mul(x, x);
  return qm31_sub(qm31_add(sq, sq), q_from_m31(1));
}

struct SecurePoint {
  qm31 x, y;
};

// CirclePoint::<SecureField>::get_random_point (STWO/core/circle.rs:169-181).
static SecurePoint random_point(HostChannel &ch) {
  const qm31 t = ch.draw_secure_felt();
  const qm31 t2 = qm31_mul(t, t);
  const qm31 inv = qm31_inv_host(qm31_add(t2, q_from_m31(1)));
  SecurePoint p;
  p.x = qm31_mul(qm31_sub(q_from_m31(1), t2), inv);
  p.y = qm31_mul(qm31_add(t, t), inv);
  return p;
}

// SecureField::from_partial_evals (STWO/core/fields/qm31.rs:51-57).
static qm31 from_partial_evals(const qm31 e[4]) {
  qm31 r = e[0];
  r = qm31_add(r, qm31_mul(e[1], qm31_make(0, 1, 0, 0)));
  r = qm31_add(r, qm31_mul(e[2], qm31_make(0, 0, 1, 0)));
  r = qm31_add(r, qm31_mul(e[3], qm31_make(0, 0, 0, 1)));
  return r;
}

// ---- proof layout (bincode 1 defaults: fixed-width little-endian integers,
// u64 sequence lengths, Option as a one-byte tag) ----------------------------------

struct ByteWriter {
  std::vector<uint8_t> b;
  void u8(uint8_t v) { b.push_back(v); }
  void u32(uint32_t v) {
    for (int i = 0; i < 4; i++) b.push_back((uint8_t)(v >> (8 * i)));
  }
  void u64(uint64_t v) {
    for (int i = 0; i < 8; i++) b.push_back((uint8_t)(v >> (8 * i)));
  }
  void words(const uint32_t *w, size_t n) {
    for (size_t i = 0; i < n; i++) u32(w[i]);
  }
};

// Per-column word runs of one tree, stored flat (column k's words follow
// column k - 1's): the proof's Vec<Vec<_>> fields without an allocation per
// column.
struct ColumnRuns {
  std::vector<uint32_t> len;   // words per column
  std::vector<uint32_t> data;  // every column's words, in column order
  // n columns of `words` words each, from src.
  void assign_uniform(size_t n, uint32_t words, const uint32_t *src) {
    len.assign(n, words);
    data.assign(src, src + n * (size_t)words);
  }
};

struct ProofLayer {
  std::vector<uint32_t> fri_witness;   // 4 words per QM31
  std::vector<uint32_t> hash_witness;  // 8 words per hash
  uint32_t commitment[8];
};

}  // namespace stwo_hip

struct stwo_hip_stark_proof {
  stwo_hip_pcs_config config{};
  std::vector<std::array<uint32_t, 8>> commitments;
  std::vector<stwo_hip::ColumnRuns> sampled_values;  // [tree]: per column 4 words per sample
  std::vector<std::vector<uint32_t>> decommitments;   // [tree] -> 8 words per hash
  std::vector<stwo_hip::ColumnRuns> queried_values;  // [tree]: per column one word per query
  uint64_t proof_of_work = 0;
  stwo_hip::ProofLayer first_layer;
  std::vector<stwo_hip::ProofLayer> inner_layers;
  std::vector<uint32_t> last_layer_poly;  // 4 words per coefficient
  uint32_t last_layer_log_size = 0;
  std::vector<uint32_t> query_positions;
  double phase_ms[STWO_HIP_PROVE_PHASES] = {};
};

namespace stwo_hip {

// CommitmentSchemeProof field order (STWO/core/pcs/quotients.rs:26-34).
static void write_proof(const stwo_hip_stark_proof &p, ByteWriter &w) {
  // PcsConfig { pow_bits, fri_config { log_blowup_factor, log_last_layer_degree_bound,
  // n_queries: usize, line_fold_step }, lifting_log_size: None } (STWO/core/pcs/mod.rs:27-40,
  // STWO/core/fri.rs:31-36)
  w.u32(p.config.pow_bits);
  w.u32(p.config.fri.log_blowup_factor);
  w.u32(p.config.fri.log_last_layer_degree_bound);
  w.u64(p.config.fri.n_queries);
  w.u32(p.config.fri.line_fold_step);
  w.u8(0);
  w.u64(p.commitments.size());
  for (const auto &c : p.commitments) w.words(c.data(), 8);
  w.u64(p.sampled_values.size());
  for (const ColumnRuns &tree : p.sampled_values) {
    w.u64(tree.len.size());
    size_t off = 0;
    for (uint32_t n : tree.len) {
      w.u64(n / 4);
      w.words(tree.data.data() + off, n);
      off += n;
    }
  }
  w.u64(p.decommitments.size());
  for (const auto &d : p.decommitments) {
    w.u64(d.size() / 8);
    w.words(d.data(), d.size());
  }
  w.u64(p.queried_values.size());
  for (const ColumnRuns &tree : p.queried_values) {
    w.u64(tree.len.size());
    size_t off = 0;
    for (uint32_t n : tree.len) {
      w.u64(n);
      w.words(tree.data.data() + off, n);
      off += n;
    }
  }
  w.u64(p.proof_of_work);
  // FriProof { first_layer, inner_layers, last_layer_poly } (STWO/core/fri.rs:360-397)
  auto layer = [&](const ProofLayer &l) {
    w.u64(l.fri_witness.size() / 4);
    w.words(l.fri_witness.data(), l.fri_witness.size());
    w.u64(l.hash_witness.size() / 8);
    w.words(l.hash_witness.data(), l.hash_witness.size());
    w.words(l.commitment, 8);
  };
  layer(p.first_layer);
  w.u64(p.inner_layers.size());
  for (const auto &l : p.inner_layers) layer(l);
  // LinePoly { coeffs, log_size } (STWO/core/poly/line.rs:113-123)
  w.u64(p.last_layer_poly.size() / 4);
  w.words(p.last_layer_poly.data(), p.last_layer_poly.size());
  w.u32(p.last_layer_log_size);
}

using Clock = std::chrono::steady_clock;
static double ms_since(Clock::time_point t0) {
  return std::chrono::duration<double, std::milli>(Clock::now() - t0).count();
}

// The FRI prover, destroyed on every exit path.  Its memory and the prove's
// buffers are context scratch (kept for the next prove, so a prove makes no
// device allocation once the context has seen one of its size).
struct ProveMem {
  stwo_hip_ctx *ctx = nullptr;
  stwo_hip_fri_prover *fri = nullptr;
  ~ProveMem() {
    if (fri) (void)stwo_hip_fri_prover_destroy(ctx, fri);
  }
};

// The trace-commit and composition-commit plans of one prove shape and
// buffer set (the key), cached on the context.
struct ProvePlans {
  const uint32_t *mem;
  uint32_t n_cols, log_n, B;
  const uint32_t *tw_d;
  uint32_t tw_initial, tw_log;
  bool store;
  stwo_hip_batch *trace = nullptr, *comp = nullptr;
  bool same(const ProvePlans &o) const {
    return mem == o.mem && n_cols == o.n_cols && log_n == o.log_n && B == o.B && tw_d == o.tw_d &&
           tw_initial == o.tw_initial && tw_log == o.tw_log && store == o.store;
  }
};

static void prove_plans_free(stwo_hip_ctx *ctx, void *p) {
  ProvePlans *pp = (ProvePlans *)p;
  if (!pp) return;
  if (pp->trace) (void)stwo_hip_batch_destroy(ctx, pp->trace);
  if (pp->comp) (void)stwo_hip_batch_destroy(ctx, pp->comp);
  delete pp;
}

// The context's plans for `key`, built (and the previous ones destroyed) when
// the key differs.  trace: LDE of the trace columns into the evaluations and
// their lifted tree (coefficients dropped unless key.store); comp: the
// composition columns interpolated in place, then their halves evaluated
// into the composition evaluations and its lifted tree.
static int prove_plans_get(stwo_hip_ctx *ctx, const ProvePlans &key, std::vector<uint32_t *> &coef,
                           std::vector<uint32_t *> &eval, uint32_t *tree1, uint32_t *const comp[4],
                           std::vector<uint32_t *> &halves, std::vector<uint32_t *> &ceval, uint32_t *tree2,
                           const stwo_hip_twiddles *tw, ProvePlans **out) {
  ProvePlans *cur = (ProvePlans *)ctx->prove_plans;
  if (cur && cur->same(key)) {
    *out = cur;
    return STWO_HIP_OK;
  }
  if (ctx->prove_plans_free) ctx->prove_plans_free(ctx, ctx->prove_plans);
  ctx->prove_plans = nullptr;
  ctx->prove_plans_free = nullptr;
  std::unique_ptr<ProvePlans> pp(new ProvePlans(key));
  const uint32_t C = key.log_n + 1;
  const std::vector<uint32_t> logs_n(key.n_cols, key.log_n), logs_c(4, C), logs_h(8, key.log_n);
  BatchTree t{coef.data(), logs_n.data(), key.n_cols, key.B, eval.data(), tree1, STWO_HIP_TREE_LIFTED,
              STWO_HIP_HASH_BLAKE2S, BATCH_LDE};
  int rc = batch_create_trees(ctx, &t, 1, tw, STWO_HIP_BATCH_GRAPH | (key.store ? 0 : STWO_HIP_BATCH_DROP_COEFFS),
                              &pp->trace);
  if (rc) return rc;
  BatchTree c[2] = {{comp, logs_c.data(), 4, 0, nullptr, nullptr, 0, 0, BATCH_IFFT},
                    {halves.data(), logs_h.data(), 8, key.B, ceval.data(), tree2, STWO_HIP_TREE_LIFTED,
                     STWO_HIP_HASH_BLAKE2S, BATCH_RFFT}};
  rc = batch_create_trees(ctx, c, 2, tw, STWO_HIP_BATCH_GRAPH, &pp->comp);
  if (rc) {
    prove_plans_free(ctx, pp.release());
    return rc;
  }
  ctx->prove_plans = pp.get();
  ctx->prove_plans_free = prove_plans_free;
  *out = pp.release();
  return STWO_HIP_OK;
}

static int fetch_layer(stwo_hip_ctx *ctx, const stwo_hip_fri_prover *fri, uint32_t k, ProofLayer &out) {
  size_t nf = 0, nh = 0, npos = 0;
  int rc = stwo_hip_fri_prover_layer_proof(ctx, fri, k, out.commitment, nullptr, 0, &nf, nullptr, 0, &nh, nullptr,
                                           nullptr, 0, &npos);
  if (rc) return rc;
  out.fri_witness.assign(4 * nf, 0);
  out.hash_witness.assign(8 * nh, 0);
  return stwo_hip_fri_prover_layer_proof(ctx, fri, k, out.commitment, out.fri_witness.data(), nf, &nf,
                                         out.hash_witness.data(), nh, &nh, nullptr, nullptr, 0, &npos);
}

// MerkleProverLifted::decommit of several trees (all columns of log size L)
// at the same query positions: plan_trees adds their reads to a gather,
// finish_trees reads them back after it ran.
struct TreeDecommit {
  const uint32_t *tree;
  std::vector<const uint32_t *> cols;
  ColumnRuns *values;                          // [column][query]
  std::vector<uint32_t> *witness;              // 8 words per hash
  std::vector<size_t> val_off;
  MerkleDecommitPlan plan;
};
static void plan_trees(Gatherer &g, uint32_t L, const std::vector<uint32_t> &queries,
                       std::vector<TreeDecommit> &trees) {
  for (TreeDecommit &t : trees) {
    t.val_off.clear();
    t.val_off.reserve(t.cols.size() * queries.size());
    g.reserve(t.cols.size() * queries.size());
    for (const uint32_t *c : t.cols)
      for (uint32_t pos : queries) t.val_off.push_back(g.add(c + pos, 1));  // log size L: no lifting
    plan_merkle_decommit(g, t.tree, L, queries, t.plan);
  }
}
static void finish_trees(const Gatherer &g, size_t nq, std::vector<TreeDecommit> &trees) {
  for (TreeDecommit &t : trees) {
    const size_t nc = t.cols.size();
    t.values->len.assign(nc, (uint32_t)nq);
    t.values->data.resize(nc * nq);
    for (size_t i = 0; i < nc * nq; i++) t.values->data[i] = g.out[t.val_off[i]];
    finish_merkle_decommit(g, t.plan, *t.witness, nullptr);
  }
}

// The device Fiat-Shamir block at the end of the prove's main buffer: the
// channel state, the composition coefficient, the OODS draw, copies of the
// two roots, the OODS point (prove_channel_step's point_out), then the
// sampled values (4 words per trace column and composition half).  One
// download reads it all.
constexpr size_t kFsChan = 0, kFsCompCoeff = 16, kFsOodsDraw = 20, kFsRoot1 = 24, kFsRoot2 = 32, kFsPoint = 64,
                 kFsValues = 256;
static_assert(kFsPoint + kOodsPointWords <= kFsValues, "Fiat-Shamir block layout");
static size_t prove_fs_words(uint32_t n_cols) { return kFsValues + 4 * ((size_t)n_cols + 8); }

// Device words of the prove's main buffer (the layout in stwo_hip_wide_fib_prove).
static size_t prove_layout_words(uint32_t n_cols, uint32_t log_n, uint32_t B) {
  const uint32_t E = log_n + B, C = log_n + 1;
  const size_t n = (size_t)1 << log_n, e = (size_t)1 << E, nc = (size_t)1 << C;
  const size_t tree_words = tree_layer_off(E + 1);
  return n_cols * n + n_cols * e + tree_words + (B > 1 ? n_cols * nc : 0) + 4 * nc + 8 * e + tree_words + 4 * e +
         prove_fs_words(n_cols);
}

size_t wide_fib_prove_device_bytes(uint32_t n_cols, uint32_t log_n, const stwo_hip_pcs_config &config) {
  const uint32_t B = config.fri.log_blowup_factor, E = log_n + B;
  size_t words = prove_layout_words(n_cols, log_n, B);
  // the FRI prover's trees and folded evaluations (fri_prover_commit_mem's layout)
  const std::vector<FriLayerShape> sched = fri_schedule(config.fri, E);
  uint32_t lg = E;
  for (const FriLayerShape &l : sched) {
    words += tree_layer_off(l.log_size + 1);
    lg -= l.fold_step;
    words += (size_t)4 << lg;
  }
  words += 32 + 8 * 32;  // the FRI channel state
  // OODS samples (eval_at_point from log_n, or barycentric from E): basis
  // table, block partials, sums; the constraint weights
  const size_t blocks = (size_t)(n_cols + 8) << (E > 11 ? E - 11 : 0);
  words += ((size_t)4 << 11) + 4 * blocks + 4 * (n_cols + 8) + 8 * (size_t)n_cols;
  // proof of work, decommitment gathers and tables: bounded by a margin
  return words * 4 + ((size_t)16 << 20);
}

}  // namespace stwo_hip

using namespace stwo_hip;

extern "C" {

int stwo_hip_wide_fib_prove_ex(stwo_hip_ctx *ctx, const uint32_t *a, const uint32_t *b, uint32_t n_cols,
                               uint32_t log_n, stwo_hip_pcs_config config, const stwo_hip_twiddles *tw, int flags,
                               uint8_t channel_digest[32], uint32_t *channel_n_draws, stwo_hip_stark_proof **out) {
  if (!ctx) return STWO_HIP_EINVAL;
  STWO_HIP_CHECK_ARG(ctx, a && b && tw && channel_digest && channel_n_draws && out, "wide_fib_prove: NULL argument");
  STWO_HIP_CHECK_ARG(ctx, (flags & ~STWO_HIP_PROVE_STORE_COEFFS) == 0, "wide_fib_prove: unknown flags 0x%x", flags);
  *out = nullptr;
  const bool store = flags & STWO_HIP_PROVE_STORE_COEFFS;
  const uint32_t B = config.fri.log_blowup_factor;
  STWO_HIP_CHECK_ARG(ctx, n_cols >= 3 && n_cols <= 4096, "wide_fib_prove: %u columns (3..4096)", n_cols);
  STWO_HIP_CHECK_ARG(ctx, B >= 1 && B <= 16 && log_n >= 4 && log_n <= 27 && log_n + B <= 27,
                     "wide_fib_prove: log_n %u with log blowup %u (log_n >= 4, 1 <= blowup <= 16, log_n + blowup <= 27)",
                     log_n, B);
  STWO_HIP_CHECK_ARG(ctx, config.pow_bits <= 32, "wide_fib_prove: pow_bits %u > 32", config.pow_bits);
  STWO_HIP_CHECK_ARG(ctx, store || B == 1,
                     "wide_fib_prove: log blowup %u without stored coefficients (extending the trace to the "
                     "constraint domain needs them; set STWO_HIP_PROVE_STORE_COEFFS)", B);
  const uint32_t E = log_n + B;  // committed evaluation (and lifting) log size
  const uint32_t C = log_n + 1;  // composition (constraint) domain log size
  const size_t n = (size_t)1 << log_n, e = (size_t)1 << E, nc = (size_t)1 << C;
  const size_t tree_words = tree_layer_off(E + 1);

  // Device layout (prove_layout_words): trace coefficients | trace
  // evaluations | trace tree | trace on CanonicCoset(C) (blowup > 1 only) |
  // composition coordinates | composition-half evaluations | composition
  // tree | quotient coordinates | Fiat-Shamir block.
  size_t off = 0;
  const size_t o_coef = off;
  off += n_cols * n;
  const size_t o_eval = off;
  off += n_cols * e;
  const size_t o_tree1 = off;
  off += tree_words;
  const size_t o_ext = off;
  if (B > 1) off += n_cols * nc;
  const size_t o_comp = off;
  off += 4 * nc;
  const size_t o_ceval = off;
  off += 8 * e;
  const size_t o_tree2 = off;
  off += tree_words;
  const size_t o_quot = off;
  off += 4 * e;
  const size_t o_fs = off;
  off += prove_fs_words(n_cols);

  ProveMem pm;
  pm.ctx = ctx;
  void *mem_p = nullptr;
  int rc = get_scratch(ctx, "wide_fib_prove", off * 4, &mem_p);
  if (rc) return rc;
  uint32_t *mem = (uint32_t *)mem_p;
  std::vector<uint32_t *> coef(n_cols), eval(n_cols), ext(n_cols);
  for (uint32_t k = 0; k < n_cols; k++) {
    coef[k] = mem + o_coef + k * n;
    eval[k] = mem + o_eval + k * e;
    ext[k] = B > 1 ? mem + o_ext + k * nc : eval[k];
  }
  uint32_t *comp[4], *quot[4];
  std::vector<uint32_t *> ceval(8);
  for (int k = 0; k < 4; k++) {
    comp[k] = mem + o_comp + k * nc;
    quot[k] = mem + o_quot + k * e;
  }
  for (int k = 0; k < 8; k++) ceval[k] = mem + o_ceval + k * e;
  uint32_t *tree1 = mem + o_tree1, *tree2 = mem + o_tree2;

  auto proof = std::make_unique<stwo_hip_stark_proof>();
  proof->config = config;
  HostChannel ch;
  memcpy(ch.digest, channel_digest, 32);
  ch.n_draws = *channel_n_draws;

  // 1. preprocessed tree: no columns.
  auto t0 = Clock::now();
  {
    b2_host h;
    std::array<uint32_t, 8> root;
    h.finalize((uint8_t *)root.data());
    proof->commitments.push_back(root);
    ch.mix_root(root.data());
  }

  // split_at_mid of the composition coordinates: left halves (coefficients
  // [0, n)), then right halves ([n, 2n)).
  std::vector<uint32_t *> halves(8);
  for (int k = 0; k < 4; k++) {
    halves[k] = comp[k];
    halves[4 + k] = comp[k] + n;
  }
  ProvePlans *plans = nullptr;
  if (!ctx->job_base) {
    ProvePlans key{mem, n_cols, log_n, B, tw->d_tw2, tw->root.initial_index, tw->root.log_size, store};
    rc = prove_plans_get(ctx, key, coef, eval, tree1, comp, halves, ceval, tree2, tw, &plans);
    if (rc) return rc;
  }

  // 2.-4a. Trace commit, composition commit and the OODS samples run on the
  // stream without a host round trip: after each commit a one-wave kernel
  // mixes the tree's root into the device copy of the channel and draws the
  // next value (the composition coefficient, then the OODS point), which the
  // following kernels read from device memory.  One download then brings the
  // roots, the channel state and the sampled values; the host replays the
  // same channel steps (two roots, two draws) and checks that it reaches the
  // device's state.
  uint32_t *fs = mem + o_fs;
  const std::vector<uint32_t> logs_n(n_cols, log_n), logs_e(n_cols + 8, E), logs_8(8, log_n);
  rc = stwo_hip_wide_fib_trace(ctx, coef.data(), n_cols, log_n, a, b);
  if (!rc && plans) {
    rc = stwo_hip_batch_run(ctx, plans->trace);
  } else if (!rc) {
    rc = stwo_hip_extend_evals(ctx, coef.data(), logs_n.data(), n_cols, B, tw, eval.data(), store ? 1 : 0);
    if (!rc)
      rc = stwo_hip_merkle_lifted_commit(ctx, eval.data(), logs_e.data(), n_cols, E, tree1, STWO_HIP_HASH_BLAKE2S);
  }
  if (!rc) rc = prove_channel_step(ctx, fs + kFsChan, ch.digest, tree1, fs + kFsRoot1, fs + kFsCompCoeff, nullptr, E);
  if (rc) return rc;
  proof->phase_ms[0] = ms_since(t0);

  // 3. composition polynomial.
  t0 = Clock::now();
  if (B > 1) rc = stwo_hip_evaluate_polynomials(ctx, coef.data(), logs_n.data(), n_cols, 1, tw, ext.data());
  if (!rc)
    rc = wide_fib_constraint_quotients_impl(ctx, ext.data(), n_cols, log_n, nullptr, comp, false, fs + kFsCompCoeff);
  const std::vector<uint32_t> logs_c(4, C);
  if (!rc && plans) {
    rc = stwo_hip_batch_run(ctx, plans->comp);
  } else if (!rc) {
    rc = stwo_hip_interpolate_columns(ctx, comp, logs_c.data(), 4, tw);
    std::vector<const uint32_t *> hc(halves.begin(), halves.end());
    if (!rc) rc = stwo_hip_evaluate_polynomials(ctx, hc.data(), logs_8.data(), 8, B, tw, ceval.data());
    if (!rc)
      rc = stwo_hip_merkle_lifted_commit(ctx, ceval.data(), logs_e.data(), 8, E, tree2, STWO_HIP_HASH_BLAKE2S);
  }
  if (!rc) rc = prove_channel_step(ctx, fs + kFsChan, nullptr, tree2, fs + kFsRoot2, fs + kFsOodsDraw, fs + kFsPoint, E);
  if (rc) return rc;
  proof->phase_ms[1] = ms_since(t0);

  // 4a. out-of-domain samples: every trace column and composition half at the OODS point.
  t0 = Clock::now();
  std::vector<const uint32_t *> sampled(n_cols + 8);
  uint32_t *d_values = fs + kFsValues;
  if (store) {
    // from the coefficients (log_n)
    std::vector<uint32_t> sample_logs(n_cols + 8, log_n);
    for (uint32_t k = 0; k < n_cols; k++) sampled[k] = coef[k];
    for (int k = 0; k < 8; k++) sampled[n_cols + k] = halves[k];
    rc = eval_at_point_dev(ctx, sampled.data(), sample_logs.data(), n_cols + 8, fs + kFsPoint, d_values);
  } else {
    // from the committed evaluations (log E, the lifting size: no doubling of
    // the point), one weight vector for every column; the FRI quotient
    // columns, written later, hold the weights meanwhile
    for (uint32_t k = 0; k < n_cols; k++) sampled[k] = eval[k];
    for (int k = 0; k < 8; k++) sampled[n_cols + k] = ceval[k];
    rc = barycentric_weights_dev(ctx, E, fs + kFsPoint + kOodsFactorWords, quot);
    if (!rc) rc = barycentric_eval_dev(ctx, sampled.data(), n_cols + 8, E, quot, d_values);
  }
  std::vector<uint32_t> fsh(prove_fs_words(n_cols));
  if (!rc) rc = stwo_hip_download(ctx, fsh.data(), fs, fsh.size());
  if (rc) return rc;
  std::array<uint32_t, 8> root1, root2;
  memcpy(root1.data(), &fsh[kFsRoot1], 32);
  memcpy(root2.data(), &fsh[kFsRoot2], 32);
  proof->commitments.push_back(root1);
  proof->commitments.push_back(root2);
  ch.mix_root(root1.data());
  const qm31 comp_coeff = ch.draw_secure_felt();
  ch.mix_root(root2.data());
  const SecurePoint oods = random_point(ch);
  if (fsh[kFsChan + 9] || memcmp(&fsh[kFsChan], ch.digest, 32) != 0 || fsh[kFsChan + 8] != ch.n_draws ||
      memcmp(&fsh[kFsCompCoeff], comp_coeff.v, 16) != 0)
    return set_error(ctx, STWO_HIP_EDEVICE, "wide_fib_prove: the device channel did not reach the host's state");
  stwo_hip_secure_point sp;
  memcpy(sp.x, oods.x.v, 16);
  memcpy(sp.y, oods.y.v, 16);
  const std::vector<uint32_t> values(fsh.begin() + kFsValues, fsh.end());
  proof->sampled_values.resize(3);
  proof->sampled_values[1].assign_uniform(n_cols, 4, values.data());
  proof->sampled_values[2].assign_uniform(8, 4, values.data() + 4 * n_cols);
  ch.mix_felts(values.data(), n_cols + 8);

  proof->phase_ms[2] = ms_since(t0);

  // 4b. FRI quotients over the flattened trees (trace columns, then composition halves).
  t0 = Clock::now();
  const qm31 quot_coeff = ch.draw_secure_felt();
  std::vector<const uint32_t *> qcols(n_cols + 8);
  for (uint32_t k = 0; k < n_cols; k++) qcols[k] = eval[k];
  for (int k = 0; k < 8; k++) qcols[n_cols + k] = ceval[k];
  std::vector<stwo_hip_point_sample> samples(n_cols + 8);
  std::vector<uint32_t> n_samples(n_cols + 8, 1);
  for (uint32_t k = 0; k < n_cols + 8; k++) {
    samples[k].point = sp;
    memcpy(samples[k].value, &values[4 * k], 16);
  }
  rc = stwo_hip_compute_fri_quotients(ctx, qcols.data(), logs_e.data(), n_cols + 8, samples.data(), n_samples.data(),
                                      quot_coeff.v, E, quot);
  if (rc) return rc;
  // prove_ex's self-check: composition value from the halves against the
  // constraints at the OODS point (on the host while the quotient kernel runs;
  // it changes no channel state, so the result is the reference's).
  {
    qm31 ev[8];
    for (int k = 0; k < 8; k++)
      ev[k] = qm31_make(values[4 * (n_cols + k)], values[4 * (n_cols + k) + 1], values[4 * (n_cols + k) + 2],
                        values[4 * (n_cols + k) + 3]);
    qm31 xr = oods.x;  // repeated_double(log_n - 1).x
    for (uint32_t d = 1; d < log_n; d++) xr = q_double_x(xr);
    const qm31 from_halves = qm31_add(from_partial_evals(ev), qm31_mul(xr, from_partial_evals(ev + 4)));
    // coset_vanishing(CanonicCoset(log_n).coset, p): the shift is zero for a
    // canonic coset, then log_n - 1 doublings of x.
    const qm31 denom_inv = qm31_inv_host(xr);
    qm31 acc = qm31_make(0, 0, 0, 0);
    auto tv = [&](uint32_t k) {
      return qm31_make(values[4 * k], values[4 * k + 1], values[4 * k + 2], values[4 * k + 3]);
    };
    qm31 va = tv(0), vb = tv(1);
    for (uint32_t k = 2; k < n_cols; k++) {
      const qm31 vc = tv(k);
      const qm31 con = qm31_sub(vc, qm31_add(qm31_mul(va, va), qm31_mul(vb, vb)));
      acc = qm31_add(qm31_mul(acc, comp_coeff), qm31_mul(denom_inv, con));
      va = vb;
      vb = vc;
    }
    if (memcmp(acc.v, from_halves.v, 16) != 0)
      return set_error(ctx, STWO_HIP_ECONSTRAINTS,
                       "wide_fib_prove: constraints not satisfied (the composition value at the OODS point does "
                       "not match the trace samples)");
  }
  proof->phase_ms[3] = ms_since(t0);

  // 4c. FRI commit.
  t0 = Clock::now();
  std::vector<uint32_t> roots(8 * (E + 1));
  uint32_t n_roots = 0;
  std::vector<uint32_t> last(4u << config.fri.log_last_layer_degree_bound);
  uint64_t nonce = 0;
  bool nonce_found = false;
  rc = fri_prover_commit_mem(ctx, ch.digest, &ch.n_draws, config.fri, quot, E, tw, roots.data(), &n_roots,
                             last.data(), "wide_fib_prove_fri", &pm.fri, (int)config.pow_bits, &nonce, &nonce_found);
  if (rc) return rc;
  proof->last_layer_poly = last;
  proof->last_layer_log_size = config.fri.log_last_layer_degree_bound;
  proof->phase_ms[4] = ms_since(t0);

  // 4d. proof of work.
  t0 = Clock::now();
  // (its first range ran with the FRI commit's end when that went on the device)
  if (nonce_found) {
    // the reference asserts both halves are below P (SIMD/grind.rs:50-57)
    if ((nonce >> 32) >= P || (uint32_t)nonce >= P)
      return set_error(ctx, STWO_HIP_EINVAL, "grind: nonce halves not below P");
    proof->proof_of_work = nonce;
  } else {
    rc = stwo_hip_grind(ctx, ch.digest, config.pow_bits, &proof->proof_of_work);
  }
  if (rc) return rc;
  ch.mix_u64(proof->proof_of_work);
  proof->phase_ms[5] = ms_since(t0);

  // 4e. FRI decommitment, then the trees at the query positions.
  t0 = Clock::now();
  // The FRI layers' and both trees' reads in one gather.
  proof->decommitments.resize(3);
  proof->queried_values.resize(3);
  {
    Gatherer g;
    FriDecommitPlan fplan;
    rc = fri_decommit_plan(ctx, pm.fri, ch.digest, &ch.n_draws, g, fplan);
    if (rc) return rc;
    proof->query_positions = fplan.queries;
    std::vector<TreeDecommit> trees(2);
    trees[0].tree = tree1;
    trees[0].cols.assign(eval.begin(), eval.end());
    trees[0].values = &proof->queried_values[1];
    trees[0].witness = &proof->decommitments[1];
    trees[1].tree = tree2;
    trees[1].cols.assign(ceval.begin(), ceval.end());
    trees[1].values = &proof->queried_values[2];
    trees[1].witness = &proof->decommitments[2];
    plan_trees(g, E, fplan.queries, trees);
    rc = g.run(ctx);
    if (rc) return rc;
    fri_decommit_finish(pm.fri, g, fplan);
    finish_trees(g, fplan.queries.size(), trees);
  }
  rc = fetch_layer(ctx, pm.fri, 0, proof->first_layer);
  if (rc) return rc;
  proof->inner_layers.resize(n_roots - 1);
  for (uint32_t k = 1; k < n_roots; k++) {
    rc = fetch_layer(ctx, pm.fri, k, proof->inner_layers[k - 1]);
    if (rc) return rc;
  }
  proof->phase_ms[6] = ms_since(t0);

  memcpy(channel_digest, ch.digest, 32);
  *channel_n_draws = ch.n_draws;
  *out = proof.release();
  return STWO_HIP_OK;
}

int stwo_hip_wide_fib_prove(stwo_hip_ctx *ctx, const uint32_t *a, const uint32_t *b, uint32_t n_cols,
                            uint32_t log_n, stwo_hip_pcs_config config, const stwo_hip_twiddles *tw,
                            uint8_t channel_digest[32], uint32_t *channel_n_draws, stwo_hip_stark_proof **out) {
  return stwo_hip_wide_fib_prove_ex(ctx, a, b, n_cols, log_n, config, tw, STWO_HIP_PROVE_STORE_COEFFS,
                                    channel_digest, channel_n_draws, out);
}

int stwo_hip_stark_proof_bytes(const stwo_hip_stark_proof *proof, uint8_t *buf, size_t cap, size_t *len) {
  if (!proof || !len) return STWO_HIP_EINVAL;
  ByteWriter w;
  write_proof(*proof, w);
  *len = w.b.size();
  if (buf) {
    if (cap < w.b.size()) return STWO_HIP_EINVAL;
    memcpy(buf, w.b.data(), w.b.size());
  }
  return STWO_HIP_OK;
}

int stwo_hip_stark_proof_info(const stwo_hip_stark_proof *proof, double phase_ms[STWO_HIP_PROVE_PHASES],
                              uint32_t *query_positions, size_t cap, size_t *n_positions) {
  if (!proof) return STWO_HIP_EINVAL;
  if (phase_ms) memcpy(phase_ms, proof->phase_ms, sizeof proof->phase_ms);
  if (n_positions) *n_positions = proof->query_positions.size();
  if (query_positions) {
    if (cap < proof->query_positions.size()) return STWO_HIP_EINVAL;
    memcpy(query_positions, proof->query_positions.data(), 4 * proof->query_positions.size());
  }
  return STWO_HIP_OK;
}

int stwo_hip_stark_proof_destroy(stwo_hip_stark_proof *proof) {
  delete proof;
  return STWO_HIP_OK;
}

}  // extern "C"

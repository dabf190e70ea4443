// Host code of the library built with AddressSanitizer and
// UndefinedBehaviorSanitizer (tests/test_sanitizers.py): the host Blake2s and
// channel, the Merkle and FRI decommitment planners and the gather
// bookkeeping they fill, on many shapes.  No GPU call is made.
//
// Prints "b2 <len> <hex digest>" lines (checked against hashlib by the test),
// one "chan <hex digest> <draws>" line, then "ok <parents> <ops>".
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <random>
#include <vector>

#include "channel.hpp"
#include "internal.hpp"

using namespace stwo_hip;

// runtime.hip's context creation sets up eval_point.hip's fixed scratch,
// which this program does not link (it creates no context through the API).
int stwo_hip::eval_point_scratch_setup(stwo_hip_ctx *) { return STWO_HIP_EDEVICE; }

static void fail(const char *what, long a, long b) {
  fprintf(stderr, "check failed: %s (%ld, %ld)\n", what, a, b);
  exit(1);
}

static void hex(const uint8_t *d, size_t n) {
  for (size_t i = 0; i < n; i++) printf("%02x", d[i]);
}

int main() {
  // Blake2s over every length 0..200 of a fixed byte pattern.
  std::vector<uint8_t> msg(200);
  for (size_t i = 0; i < msg.size(); i++) msg[i] = (uint8_t)(i * 7 + 3);
  for (size_t len = 0; len <= msg.size(); len++) {
    b2_host h;
    h.update(msg.data(), len);
    uint8_t d[32];
    h.finalize(d);
    printf("b2 %zu ", len);
    hex(d, 32);
    printf("\n");
  }
  // A channel transcript: roots, felts, u64s and draws.
  HostChannel ch;
  memset(ch.digest, 0, 32);
  ch.n_draws = 0;
  uint32_t root[8], felts[4 * 37], w[8];
  for (int i = 0; i < 8; i++) root[i] = 0x01020304u * (i + 1);
  for (int i = 0; i < 4 * 37; i++) felts[i] = (uint32_t)(i * 2654435761u) & 0x7fffffffu;
  ch.mix_root(root);
  ch.draw_u32s(w);
  ch.mix_felts(felts, 37);
  ch.draw_u32s(w);
  ch.draw_u32s(w);
  ch.mix_u64(0x123456789abcdefull);
  const qm31 a = ch.draw_secure_felt();
  printf("chan ");
  hex(ch.digest, 32);
  printf(" %u %u\n", ch.n_draws, a.v[0]);

  std::mt19937_64 rng(7);
  long parents = 0, ops = 0;
  // Merkle decommitment plans: sorted, deduplicated queries on trees of log 1..20.
  for (uint32_t L = 1; L <= 20; L++) {
    for (int rep = 0; rep < 20; rep++) {
      const uint32_t nq = 1 + (uint32_t)(rng() % 40);
      std::vector<uint32_t> q;
      for (uint32_t i = 0; i < nq; i++) q.push_back((uint32_t)(rng() & ((1u << L) - 1)));
      std::sort(q.begin(), q.end());
      q.erase(std::unique(q.begin(), q.end()), q.end());
      Gatherer g;
      MerkleDecommitPlan plan;
      const uint32_t *tree = (const uint32_t *)(uintptr_t)0x10000000ull;  // never dereferenced
      plan_merkle_decommit(g, tree, L, q, plan);
      size_t prev_off = 0;
      for (size_t i = 0; i < plan.parents.size(); i++) {
        const auto &p = plan.parents[i];
        if (p.child_layer < 1 || p.child_layer > L) fail("child layer", p.child_layer, L);
        if (p.parent >= (1u << (p.child_layer - 1))) fail("parent index", p.parent, p.child_layer);
        if (i && p.off != prev_off + 16) fail("gather offsets", (long)p.off, (long)prev_off);
        prev_off = p.off;
      }
      if (g.words != 16 * plan.parents.size()) fail("gather words", (long)g.words, (long)plan.parents.size());
      g.out.assign(g.words, 5);
      std::vector<uint32_t> witness, aux;
      finish_merkle_decommit(g, plan, witness, &aux);
      if (aux.size() != 20 * plan.parents.size()) fail("aux size", (long)aux.size(), (long)plan.parents.size());
      parents += (long)plan.parents.size();
      ops += (long)g.ops_.size();
    }
  }
  // FRI decommitment plans on synthetic provers: first (circle) layer of log
  // N with fold step 1, then line layers folding by `step` down to the last.
  stwo_hip_ctx ctx;
  for (uint32_t N = 3; N <= 22; N++) {
    for (uint32_t step = 1; step <= 4; step++) {
      for (uint32_t last = 0; last + step < N && last <= 3; last++) {
        stwo_hip_fri_prover pr;
        pr.device = ctx.device;
        pr.cfg.n_queries = 1 + (uint32_t)(rng() % 30);
        pr.cfg.line_fold_step = step;
        pr.log_size = N;
        uint32_t lg = N, k = 0;
        for (;;) {
          FriLayerKeep layer{};
          layer.shape.log_size = lg;
          layer.shape.fold_step = k == 0 ? 1 : std::min(step, lg - last);
          layer.tree = (uint32_t *)(uintptr_t)(0x20000000ull + ((uint64_t)k << 24));
          for (int c = 0; c < 4; c++) layer.cols[c] = (const uint32_t *)(uintptr_t)(0x40000000ull + ((uint64_t)k << 24) + ((uint64_t)c << 22));
          pr.layers.push_back(layer);
          lg -= layer.shape.fold_step;
          k++;
          if (lg <= last) break;
        }
        uint8_t digest[32] = {1, 2, 3, (uint8_t)N, (uint8_t)step};
        uint32_t n_draws = 0;
        Gatherer g;
        FriDecommitPlan plan;
        if (fri_decommit_plan(&ctx, &pr, digest, &n_draws, g, plan) != STWO_HIP_OK) fail("fri plan", N, step);
        if (plan.layers.size() != pr.layers.size()) fail("fri layers", (long)plan.layers.size(), (long)pr.layers.size());
        for (size_t li = 0; li < plan.layers.size(); li++) {
          const auto &pl = plan.layers[li];
          const uint32_t size = 1u << pr.layers[li].shape.log_size;
          for (uint32_t pos : pl.positions)
            if (pos >= size) fail("fri position", pos, size);
          if (pl.val_off.size() != 4 * pl.positions.size()) fail("fri values", (long)pl.val_off.size(), 0);
          if (pl.witness.size() != pl.positions.size()) fail("fri witness flags", (long)pl.witness.size(), 0);
        }
        for (uint32_t q : plan.queries)
          if (q >= (1u << N)) fail("query", q, N);
        g.out.assign(g.words, 9);
        fri_decommit_finish(&pr, g, plan);
        if (!pr.decommitted || pr.proofs.size() != pr.layers.size()) fail("fri finish", (long)pr.proofs.size(), 0);
        ops += (long)g.ops_.size();
      }
    }
  }
  printf("ok %ld %ld\n", parents, ops);
  return 0;
}

// Serving loop: continuous batching of commit + FRI and prove jobs on one
// device.
//
// Jobs are submitted from any thread and queued.  A scheduler thread admits
// them in arrival order whenever a worker slot is idle: the queue's head
// and the commit jobs right behind it (up to max_group) form a group if the
// group's device footprint fits in the HBM budget -- one slab allocated up
// front and sub-allocated first-fit, so admission never calls hipMalloc.
// Each slot owns a context (its own stream), so groups on different slots
// run concurrently and their kernels overlap on the device.  The twiddle
// tree (rooted at the largest domain the server accepts) is built once and
// shared read-only by every slot.
//
// A commit group, on its slot's stream: each job's columns are copied from
// the caller (device or host) or generated from the seed into the group's
// memory; ONE batched plan interpolates, evaluates and commits every job's
// columns -- columns of equal size and blowup from all jobs share each FFT
// pass launch -- into one lifted Blake2s tree per job (the PCS form,
// CommitmentTreeProver::new, STWO/prover/pcs/mod.rs:766-803); then per job
// the trace root is mixed into a fresh Blake2s channel and FRI commits the
// QM31 column formed by its first four evaluations (FriProver::commit,
// STWO/prover/fri.rs:133-291).  A prove group runs stwo_hip_wide_fib_prove.
// While a group runs, every device buffer the library takes for itself
// (plan tables, FRI trees, prove buffers) is carved from the group's block
// (job_mem_begin).  Results depend only on the job: each has its own
// buffers, tree and channel, and all arithmetic is exact.
#include <string.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "blake2s.hpp"
#include "common.hpp"
#include "field.hpp"
#include "internal.hpp"

namespace stwo_hip {
namespace {

// First-fit sub-allocator over one device slab (offsets, 256-byte granules).
class SlabPool {
 public:
  void init(uint8_t *base, size_t bytes) {
    base_ = base;
    free_.clear();
    free_[0] = bytes;
    cap_ = bytes;
  }
  uint8_t *alloc(size_t bytes) {
    bytes = (bytes + 255) & ~(size_t)255;
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second >= bytes) {
        size_t off = it->first, rest = it->second - bytes;
        free_.erase(it);
        if (rest) free_[off + bytes] = rest;
        used_ += bytes;
        return base_ + off;
      }
    }
    return nullptr;
  }
  void release(uint8_t *p, size_t bytes) {
    bytes = (bytes + 255) & ~(size_t)255;
    size_t off = (size_t)(p - base_);
    auto it = free_.emplace(off, bytes).first;
    auto nx = std::next(it);
    if (nx != free_.end() && it->first + it->second == nx->first) {
      it->second += nx->second;
      free_.erase(nx);
    }
    if (it != free_.begin()) {
      auto pv = std::prev(it);
      if (pv->first + pv->second == it->first) {
        pv->second += it->second;
        free_.erase(it);
      }
    }
    used_ -= bytes;
  }
  size_t used() const { return used_; }
  size_t largest_free() const {
    size_t m = 0;
    for (auto &kv : free_) m = std::max(m, kv.second);
    return m;
  }
  size_t capacity() const { return cap_; }

 private:
  uint8_t *base_ = nullptr;
  std::map<size_t, size_t> free_;
  size_t cap_ = 0, used_ = 0;
};

struct Job {
  uint64_t id;
  stwo_hip_job_desc desc;
  stwo_hip_job_result res;
  bool done = false;
};

struct Group {
  std::vector<std::shared_ptr<Job>> jobs;
  uint8_t *mem = nullptr;
  size_t bytes = 0;
};

inline size_t words_tree(uint32_t max_log) { return (((size_t)2 << max_log) - 1) * 8; }
inline size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

// A commit job's own data: columns, evaluations, tree.
size_t commit_data_bytes(const stwo_hip_job_desc &d) {
  const uint32_t L = d.log_size + d.log_blowup;
  return align256(((size_t)d.n_cols << d.log_size) * 4) + align256(((size_t)d.n_cols << L) * 4) +
         align256(words_tree(L) * 4);
}
// FriProver::commit's scratch for a first layer of 2^L (fri_commit_impl:
// tree, two layer buffers, channel state).
size_t fri_scratch_bytes(uint32_t L) {
  return align256(words_tree(L) * 4) + 2 * align256(((size_t)4 << (L - 1)) * 4) + align256((32 + 8 * 32) * 4);
}
// Device bytes of a group: every commit job's data, the batched plan's
// tables (column tables and pass descriptors), the largest FRI scratch (the
// FRI commits run largest first, so later ones reuse it); a prove job's
// buffers and its seeded inputs.
size_t group_bytes(const std::vector<const stwo_hip_job_desc *> &g) {
  if (g.size() == 1 && g[0]->kind == STWO_HIP_JOB_PROVE) {
    const stwo_hip_job_desc &d = *g[0];
    return align256(((size_t)2 << d.log_size) * 4) + wide_fib_prove_device_bytes(d.n_cols, d.log_size, d.pcs) +
           ((size_t)1 << 20);
  }
  size_t b = 0, cols = 0;
  uint32_t max_L = 0;
  for (const stwo_hip_job_desc *d : g) {
    b += commit_data_bytes(*d);
    cols += d->n_cols;
    max_L = std::max(max_L, d->log_size + d->log_blowup);
  }
  b += (size_t)256 * 1024 + 128 * cols;  // plan tables and descriptors
  b += fri_scratch_bytes(max_L) + (size_t)64 * 1024;
  return align256(b);
}

}  // namespace
}  // namespace stwo_hip

using namespace stwo_hip;

struct stwo_hip_server {
  int device = 0;
  uint32_t max_log = 0;  // max trace log + blowup accepted
  uint32_t max_group = 16;
  bool held = false;  // stwo_hip_server_hold: admit nothing until released
  stwo_hip_ctx *main_ctx = nullptr;
  stwo_hip_twiddles *tw = nullptr;
  uint8_t *slab = nullptr;
  SlabPool pool;
  std::vector<stwo_hip_ctx *> slot_ctx;
  std::vector<std::thread> workers;
  std::thread scheduler;
  std::mutex mu;
  std::condition_variable cv_sched, cv_work, cv_done;
  std::deque<std::shared_ptr<Job>> pending;
  std::deque<std::shared_ptr<Group>> admitted;  // handed to workers
  std::unordered_map<uint64_t, std::shared_ptr<Job>> jobs;
  uint64_t next_id = 1;
  uint32_t idle_slots = 0;
  uint64_t mem_epoch = 1, blocked_epoch = 0;  // releases seen / epoch the head group failed at
  size_t max_queue = 4096;
  bool stop = false;
  std::chrono::steady_clock::time_point t0;

  double now_ms() const {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  }
};

namespace stwo_hip {
namespace {

int check_job(const stwo_hip_server *s, const stwo_hip_job_desc &d) {
  if (d.kind == STWO_HIP_JOB_PROVE) {
    const uint32_t B = d.pcs.fri.log_blowup_factor;
    if (d.n_cols < 3 || d.n_cols > 4096 || d.log_size < 4 || B < 1 || B > 16 || d.log_size + B > 27 ||
        d.log_size + B > s->max_log || d.log_blowup != B || d.pcs.pow_bits > 32 || !d.proof_out ||
        d.pcs.fri.line_fold_step < 1 || d.pcs.fri.line_fold_step > 28 || d.pcs.fri.log_last_layer_degree_bound > 10)
      return STWO_HIP_EINVAL;
    return STWO_HIP_OK;
  }
  if (d.kind != STWO_HIP_JOB_COMMIT_FRI) return STWO_HIP_EINVAL;
  if (d.n_cols < 4 || d.log_size < 4 || d.log_blowup < 1 || d.log_size > 28 || d.log_blowup > 16 ||
      d.log_size + d.log_blowup > s->max_log || d.fri.log_blowup_factor != d.log_blowup ||
      d.fri.log_last_layer_degree_bound > 4 || d.fri.line_fold_step < 1 || d.fri.line_fold_step > 4 ||
      d.log_size + d.log_blowup - 1 < d.fri.log_last_layer_degree_bound + d.log_blowup)
    return STWO_HIP_EINVAL;
  return STWO_HIP_OK;
}

// Columns of a job into `dst` (n_cols x 2^log_size): the caller's device
// columns, its host data, or the seed.
int load_columns(stwo_hip_ctx *ctx, const stwo_hip_job_desc &d, uint32_t n_cols, uint32_t *dst) {
  const size_t n = (size_t)1 << d.log_size;
  for (uint32_t c = 0; c < n_cols; c++) {
    uint32_t *col = dst + c * n;
    if (d.cols) {
      STWO_HIP_CHECK_ARG(ctx, d.cols[c], "job column %u is NULL", c);
      STWO_HIP_TRY(ctx, hipMemcpyAsync(col, d.cols[c], n * 4, hipMemcpyDeviceToDevice, ctx->stream));
    } else if (d.host) {
      STWO_HIP_TRY(ctx, hipMemcpyAsync(col, d.host + c * n, n * 4, hipMemcpyHostToDevice, ctx->stream));
    } else {
      int rc = stwo_hip_fill_random(ctx, col, n, d.seed, c);
      if (rc) return rc;
    }
  }
  return STWO_HIP_OK;
}

int run_commit_group(stwo_hip_server *s, stwo_hip_ctx *ctx, Group &g) {
  uint8_t *p = g.mem;
  std::vector<std::vector<uint32_t *>> cols(g.jobs.size()), outs(g.jobs.size());
  std::vector<std::vector<uint32_t>> logs(g.jobs.size());
  std::vector<uint32_t *> trees(g.jobs.size());
  std::vector<BatchTree> sets(g.jobs.size());
  for (size_t k = 0; k < g.jobs.size(); k++) {
    const stwo_hip_job_desc &d = g.jobs[k]->desc;
    const uint32_t n = d.log_size, L = n + d.log_blowup;
    uint32_t *in = (uint32_t *)p;
    p += align256(((size_t)d.n_cols << n) * 4);
    uint32_t *out = (uint32_t *)p;
    p += align256(((size_t)d.n_cols << L) * 4);
    trees[k] = (uint32_t *)p;
    p += align256(words_tree(L) * 4);
    int rc = load_columns(ctx, d, d.n_cols, in);
    if (rc) return rc;
    for (uint32_t c = 0; c < d.n_cols; c++) {
      cols[k].push_back(in + ((size_t)c << n));
      outs[k].push_back(out + ((size_t)c << L));
    }
    logs[k].assign(d.n_cols, n);
    sets[k] = {cols[k].data(), logs[k].data(), d.n_cols, d.log_blowup, outs[k].data(), trees[k],
               STWO_HIP_TREE_LIFTED, STWO_HIP_HASH_BLAKE2S};
  }
  job_mem_begin(ctx, p, g.mem + g.bytes - p);
  stwo_hip_batch *plan = nullptr;
  int rc = batch_create_trees(ctx, sets.data(), sets.size(), s->tw, STWO_HIP_BATCH_DROP_COEFFS, &plan);
  if (rc == STWO_HIP_OK) rc = stwo_hip_batch_run(ctx, plan);
  if (rc == STWO_HIP_OK) {
    std::vector<DownloadReq> reqs;
    for (size_t k = 0; k < g.jobs.size(); k++) reqs.push_back({g.jobs[k]->res.trace_root, trees[k], 8});
    rc = download_many(ctx, reqs.data(), reqs.size());
  }
  if (plan) stwo_hip_batch_destroy(ctx, plan);
  // FRI per job, largest first (the FRI scratch carved for the first serves the others)
  std::vector<size_t> order(g.jobs.size());
  for (size_t k = 0; k < order.size(); k++) order[k] = k;
  std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) {
    return g.jobs[a]->desc.log_size + g.jobs[a]->desc.log_blowup > g.jobs[b]->desc.log_size + g.jobs[b]->desc.log_blowup;
  });
  for (size_t k : order) {
    if (rc) break;
    Job &j = *g.jobs[k];
    const stwo_hip_job_desc &d = j.desc;
    const uint32_t L = d.log_size + d.log_blowup;
    // Channel: mix the trace root, then FRI on (evals 0..3) as QM31 coordinates.
    b2_host h;
    uint8_t zero[32] = {0};
    h.update(zero, 32);
    h.update(j.res.trace_root, 32);
    uint8_t digest[32];
    h.finalize(digest);
    uint32_t n_draws = 0;
    const uint32_t *coords[4] = {outs[k][0], outs[k][1], outs[k][2], outs[k][3]};
    uint32_t n_roots = 0;
    rc = stwo_hip_fri_commit(ctx, digest, &n_draws, d.fri, coords, L, s->tw, j.res.fri_roots, &n_roots,
                             j.res.last_layer_poly);
    if (rc) break;
    j.res.n_fri_roots = n_roots;
    memcpy(j.res.final_digest, digest, 32);
  }
  if (rc == STWO_HIP_OK) rc = stwo_hip_sync(ctx);
  job_mem_end(ctx);
  return rc;
}

int run_prove_group(stwo_hip_server *s, stwo_hip_ctx *ctx, Group &g) {
  Job &j = *g.jobs[0];
  const stwo_hip_job_desc &d = j.desc;
  const size_t n = (size_t)1 << d.log_size;
  const uint32_t *a = nullptr, *b = nullptr;
  uint8_t *p = g.mem;
  int rc = STWO_HIP_OK;
  if (d.cols) {
    a = d.cols[0];
    b = d.cols[1];
  } else {
    uint32_t *ab = (uint32_t *)p;
    rc = load_columns(ctx, d, 2, ab);
    a = ab;
    b = ab + n;
  }
  p += align256(((size_t)2 << d.log_size) * 4);
  if (rc) return rc;
  job_mem_begin(ctx, p, g.mem + g.bytes - p);
  uint8_t digest[32] = {0};
  uint32_t n_draws = 0;
  stwo_hip_stark_proof *proof = nullptr;
  rc = stwo_hip_wide_fib_prove(ctx, a, b, d.n_cols, d.log_size, d.pcs, s->tw, digest, &n_draws, &proof);
  if (rc == STWO_HIP_OK) {
    size_t len = 0;
    rc = stwo_hip_stark_proof_bytes(proof, d.proof_out, d.proof_cap, &len);
    j.res.proof_len = len;
    memcpy(j.res.final_digest, digest, 32);
  }
  if (proof) stwo_hip_stark_proof_destroy(proof);
  if (rc == STWO_HIP_OK) rc = stwo_hip_sync(ctx);
  job_mem_end(ctx);
  return rc;
}

void worker_main(stwo_hip_server *s, uint32_t slot) {
  stwo_hip_ctx *ctx = s->slot_ctx[slot];
  (void)hipSetDevice(s->device);
  for (;;) {
    std::shared_ptr<Group> g;
    {
      std::unique_lock<std::mutex> lk(s->mu);
      s->cv_work.wait(lk, [&] { return !s->admitted.empty() || (s->stop && s->pending.empty()); });
      if (s->admitted.empty()) return;  // draining finished
      g = s->admitted.front();
      s->admitted.pop_front();
    }
    const double t_start = s->now_ms();
    for (auto &j : g->jobs) {
      j->res.t_start_ms = t_start;
      j->res.group_size = (uint32_t)g->jobs.size();
    }
    const bool prove = g->jobs[0]->desc.kind == STWO_HIP_JOB_PROVE;
    const int rc = prove ? run_prove_group(s, ctx, *g) : run_commit_group(s, ctx, *g);
    const double t_done = s->now_ms();
    {
      std::lock_guard<std::mutex> lk(s->mu);
      s->pool.release(g->mem, g->bytes);
      s->mem_epoch++;
      for (auto &j : g->jobs) {
        j->res.status = rc;
        j->res.t_done_ms = t_done;
        j->done = true;
      }
      s->idle_slots++;
    }
    s->cv_sched.notify_one();
    s->cv_done.notify_all();
  }
}

// Admits groups in arrival order while a slot is idle and memory fits.  A
// head group that does not fit (including by fragmentation) waits for the
// next release of memory; later jobs do not overtake it.
void scheduler_main(stwo_hip_server *s) {
  std::unique_lock<std::mutex> lk(s->mu);
  for (;;) {
    s->cv_sched.wait(lk, [&] {
      if (s->stop && s->pending.empty()) return true;
      if (s->held && !s->stop) return false;
      if (s->pending.empty() || s->idle_slots == 0) return false;
      return s->mem_epoch != s->blocked_epoch;
    });
    if (s->stop && s->pending.empty()) return;
    bool admitted_any = false;
    while (!s->pending.empty() && s->idle_slots > 0) {
      std::vector<const stwo_hip_job_desc *> descs = {&s->pending.front()->desc};
      size_t need = group_bytes(descs);
      if (need > s->pool.largest_free()) {
        s->blocked_epoch = s->mem_epoch;  // wait for a release
        break;
      }
      if (descs[0]->kind == STWO_HIP_JOB_COMMIT_FRI) {
        // the commit jobs right behind the head, while the group still fits
        for (size_t k = 1; k < s->pending.size() && descs.size() < s->max_group; k++) {
          const stwo_hip_job_desc *d = &s->pending[k]->desc;
          if (d->kind != STWO_HIP_JOB_COMMIT_FRI) break;
          descs.push_back(d);
          const size_t more = group_bytes(descs);
          if (more > s->pool.largest_free()) {
            descs.pop_back();
            break;
          }
          need = more;
        }
      }
      uint8_t *m = s->pool.alloc(need);
      if (!m) {
        s->blocked_epoch = s->mem_epoch;
        break;
      }
      auto g = std::make_shared<Group>();
      g->mem = m;
      g->bytes = need;
      const double t = s->now_ms();
      for (size_t k = 0; k < descs.size(); k++) {
        auto j = s->pending.front();
        s->pending.pop_front();
        j->res.t_admit_ms = t;
        g->jobs.push_back(j);
      }
      s->idle_slots--;
      s->admitted.push_back(g);
      admitted_any = true;
    }
    if (admitted_any) s->cv_work.notify_all();
  }
}

}  // namespace
}  // namespace stwo_hip

extern "C" {

int stwo_hip_server_create(int device, size_t hbm_budget_bytes, uint32_t max_log_size, uint32_t n_slots,
                           stwo_hip_server **out) {
  if (!out || n_slots == 0 || n_slots > 16 || max_log_size < 4 || max_log_size > 28) return STWO_HIP_EINVAL;
  *out = nullptr;
  stwo_hip_server *s = new (std::nothrow) stwo_hip_server();
  if (!s) return STWO_HIP_ENOMEM;
  s->device = device;
  s->max_log = max_log_size;
  s->t0 = std::chrono::steady_clock::now();
  int rc = stwo_hip_ctx_create(device, &s->main_ctx);
  if (rc) {
    delete s;
    return rc;
  }
  rc = stwo_hip_twiddles_create(s->main_ctx, {half_odds_initial(max_log_size - 1), max_log_size - 1}, &s->tw);
  if (rc == STWO_HIP_OK && hipMalloc((void **)&s->slab, hbm_budget_bytes) != hipSuccess) {
    (void)hipGetLastError();
    rc = STWO_HIP_ENOMEM;
    s->slab = nullptr;
  }
  for (uint32_t i = 0; rc == STWO_HIP_OK && i < n_slots; i++) {
    stwo_hip_ctx *c = nullptr;
    rc = stwo_hip_ctx_create(device, &c);
    if (rc == STWO_HIP_OK) s->slot_ctx.push_back(c);
  }
  if (rc) {
    for (auto c : s->slot_ctx) stwo_hip_ctx_destroy(c);
    if (s->slab) (void)hipFree(s->slab);
    if (s->tw) stwo_hip_twiddles_destroy(s->main_ctx, s->tw);
    stwo_hip_ctx_destroy(s->main_ctx);
    delete s;
    return rc;
  }
  s->pool.init(s->slab, hbm_budget_bytes);
  s->idle_slots = n_slots;
  for (uint32_t i = 0; i < n_slots; i++) s->workers.emplace_back(worker_main, s, i);
  s->scheduler = std::thread(scheduler_main, s);
  *out = s;
  return STWO_HIP_OK;
}

int stwo_hip_server_set_max_group(stwo_hip_server *s, uint32_t max_group) {
  if (!s || max_group < 1 || max_group > 64) return STWO_HIP_EINVAL;
  std::lock_guard<std::mutex> lk(s->mu);
  s->max_group = max_group;
  return STWO_HIP_OK;
}

int stwo_hip_server_hold(stwo_hip_server *s, int hold) {
  if (!s) return STWO_HIP_EINVAL;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->held = hold != 0;
  }
  s->cv_sched.notify_all();
  return STWO_HIP_OK;
}

int stwo_hip_server_job_bytes(stwo_hip_server *s, const stwo_hip_job_desc *desc, size_t *bytes) {
  if (!s || !desc || !bytes) return STWO_HIP_EINVAL;
  int rc = check_job(s, *desc);
  if (rc) return rc;
  std::vector<const stwo_hip_job_desc *> g = {desc};
  *bytes = group_bytes(g);
  return STWO_HIP_OK;
}

int stwo_hip_server_submit(stwo_hip_server *s, const stwo_hip_job_desc *desc, uint64_t *job_id) {
  if (!s || !desc || !job_id) return STWO_HIP_EINVAL;
  const stwo_hip_job_desc &d = *desc;
  int rc = check_job(s, d);
  if (rc) return rc;
  auto j = std::make_shared<Job>();
  j->desc = d;
  memset(&j->res, 0, sizeof j->res);
  j->res.status = STWO_HIP_EBUSY;
  std::vector<const stwo_hip_job_desc *> g = {&d};
  {
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->stop) return STWO_HIP_EINVAL;
    if (s->pending.size() >= s->max_queue) return STWO_HIP_EBUSY;
    if (group_bytes(g) > s->pool.capacity()) return STWO_HIP_ENOMEM;
    j->id = s->next_id++;
    j->res.t_submit_ms = s->now_ms();
    s->jobs[j->id] = j;
    s->pending.push_back(j);
    *job_id = j->id;
  }
  s->cv_sched.notify_one();
  return STWO_HIP_OK;
}

int stwo_hip_server_poll(stwo_hip_server *s, uint64_t job_id, stwo_hip_job_result *out) {
  if (!s) return STWO_HIP_EINVAL;
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->jobs.find(job_id);
  if (it == s->jobs.end()) return STWO_HIP_EINVAL;
  if (!it->second->done) return STWO_HIP_EBUSY;
  if (out) *out = it->second->res;
  return STWO_HIP_OK;
}

int stwo_hip_server_wait(stwo_hip_server *s, uint64_t job_id, stwo_hip_job_result *out) {
  if (!s) return STWO_HIP_EINVAL;
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->jobs.find(job_id);
  if (it == s->jobs.end()) return STWO_HIP_EINVAL;
  auto j = it->second;
  s->cv_done.wait(lk, [&] { return j->done; });
  if (out) *out = j->res;
  s->jobs.erase(job_id);
  return STWO_HIP_OK;
}

int stwo_hip_server_stats(stwo_hip_server *s, size_t *pool_used, size_t *pool_capacity, uint32_t *n_pending,
                          uint32_t *n_running) {
  if (!s) return STWO_HIP_EINVAL;
  std::lock_guard<std::mutex> lk(s->mu);
  if (pool_used) *pool_used = s->pool.used();
  if (pool_capacity) *pool_capacity = s->pool.capacity();
  if (n_pending) *n_pending = (uint32_t)s->pending.size();
  if (n_running) *n_running = (uint32_t)(s->slot_ctx.size() - s->idle_slots);
  return STWO_HIP_OK;
}

int stwo_hip_server_destroy(stwo_hip_server *s) {
  if (!s) return STWO_HIP_EINVAL;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
  }
  s->cv_sched.notify_all();
  s->scheduler.join();  // returns once every queued job is admitted
  s->cv_work.notify_all();
  for (auto &w : s->workers) w.join();  // return once admitted jobs are done
  for (auto c : s->slot_ctx) stwo_hip_ctx_destroy(c);
  if (s->slab) (void)hipFree(s->slab);
  if (s->tw) stwo_hip_twiddles_destroy(s->main_ctx, s->tw);
  stwo_hip_ctx_destroy(s->main_ctx);
  delete s;
  return STWO_HIP_OK;
}

}  // extern "C"

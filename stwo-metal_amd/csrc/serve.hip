// Serving loop: continuous batching of commit + FRI jobs on one device.
//
// Jobs (n columns of one trace size, a blowup, a FRI config) are submitted
// from any thread and queued.  A scheduler thread admits queued jobs in
// arrival order whenever a worker slot is idle and the job's device
// footprint fits in the HBM budget; the budget is one slab allocated up
// front and sub-allocated first-fit, so admission never calls hipMalloc.
// Each slot owns a context (its own stream); admitted jobs therefore run
// concurrently and their kernels overlap on the device.  The twiddle tree
// (rooted at the largest domain the server accepts) is built once and
// shared read-only by every slot.
//
// Per job, on its slot's stream: synthetic trace columns are generated on
// the device, one batched plan interpolates them, evaluates them at the
// blowup and commits them in a lifted Blake2s tree (the PCS form,
// CommitmentTreeProver::new, STWO/prover/pcs/mod.rs:766-803); the trace
// root is mixed into a fresh Blake2s channel and FRI commits the QM31
// column formed by the first four evaluations (FriProver::commit,
// STWO/prover/fri.rs:133-291).  Results depend only on the job: every job
// has its own buffers and channel, and all arithmetic is exact.
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
  uint8_t *mem = nullptr;
  size_t mem_bytes = 0;
};

inline size_t words_tree(uint32_t max_log) { return (((size_t)2 << max_log) - 1) * 8; }

// Device words a job needs: inputs, evaluations, tree, FRI scratch.
size_t job_bytes(const stwo_hip_job_desc &d) {
  size_t n_in = (size_t)d.n_cols << d.log_size;
  uint32_t L = d.log_size + d.log_blowup;
  size_t n_out = (size_t)d.n_cols << L;
  size_t w = n_in + n_out + words_tree(L) + 256;
  return (w * 4 + 255) & ~(size_t)255;
}

}  // namespace
}  // namespace stwo_hip

using namespace stwo_hip;

struct stwo_hip_server {
  int device = 0;
  uint32_t max_log = 0;  // max trace log + blowup accepted
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
  std::deque<std::shared_ptr<Job>> admitted;  // handed to workers
  std::unordered_map<uint64_t, std::shared_ptr<Job>> jobs;
  uint64_t next_id = 1;
  uint32_t idle_slots = 0;
  uint64_t mem_epoch = 1, blocked_epoch = 0;  // releases seen / epoch the head job failed at
  size_t max_queue = 4096;
  bool stop = false;
  std::chrono::steady_clock::time_point t0;

  double now_ms() const {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  }
};

namespace stwo_hip {
namespace {

int run_job(stwo_hip_server *s, stwo_hip_ctx *ctx, Job &j) {
  const stwo_hip_job_desc &d = j.desc;
  const uint32_t n = d.log_size, L = d.log_size + d.log_blowup;
  uint32_t *base = (uint32_t *)j.mem;
  uint32_t *in = base;
  uint32_t *out = in + ((size_t)d.n_cols << n);
  uint32_t *tree = out + ((size_t)d.n_cols << L);
  std::vector<uint32_t *> cols(d.n_cols), outs(d.n_cols);
  std::vector<uint32_t> logs(d.n_cols, n);
  for (uint32_t c = 0; c < d.n_cols; c++) {
    cols[c] = in + ((size_t)c << n);
    outs[c] = out + ((size_t)c << L);
    int rc = stwo_hip_fill_random(ctx, cols[c], (size_t)1 << n, d.seed, c);
    if (rc) return rc;
  }
  stwo_hip_batch *plan = nullptr;
  int rc = stwo_hip_batch_create(ctx, cols.data(), logs.data(), d.n_cols, d.log_blowup, s->tw, outs.data(), tree,
                                 STWO_HIP_TREE_LIFTED, STWO_HIP_HASH_BLAKE2S, STWO_HIP_BATCH_DROP_COEFFS, &plan);
  if (rc) return rc;
  rc = stwo_hip_batch_run(ctx, plan);
  if (rc == STWO_HIP_OK) rc = stwo_hip_download(ctx, j.res.trace_root, tree, 8);
  stwo_hip_batch_destroy(ctx, plan);
  if (rc) return rc;
  // Channel: mix the trace root, then FRI on (evals 0..3) as QM31 coordinates.
  b2_host h;
  uint8_t zero[32] = {0};
  h.update(zero, 32);
  h.update(j.res.trace_root, 32);
  uint8_t digest[32];
  h.finalize(digest);
  uint32_t n_draws = 0;
  const uint32_t *coords[4] = {outs[0], outs[1], outs[2], outs[3]};
  uint32_t n_roots = 0;
  rc = stwo_hip_fri_commit(ctx, digest, &n_draws, d.fri, coords, L, s->tw, j.res.fri_roots, &n_roots,
                           j.res.last_layer_poly);
  if (rc) return rc;
  j.res.n_fri_roots = n_roots;
  memcpy(j.res.final_digest, digest, 32);
  return STWO_HIP_OK;
}

void worker_main(stwo_hip_server *s, uint32_t slot) {
  stwo_hip_ctx *ctx = s->slot_ctx[slot];
  (void)hipSetDevice(s->device);
  for (;;) {
    std::shared_ptr<Job> j;
    {
      std::unique_lock<std::mutex> lk(s->mu);
      s->cv_work.wait(lk, [&] { return !s->admitted.empty() || (s->stop && s->pending.empty()); });
      if (s->admitted.empty()) return;  // draining finished
      j = s->admitted.front();
      s->admitted.pop_front();
    }
    j->res.t_start_ms = s->now_ms();
    int rc = run_job(s, ctx, *j);
    if (rc == STWO_HIP_OK) rc = stwo_hip_sync(ctx);
    j->res.status = rc;
    j->res.t_done_ms = s->now_ms();
    {
      std::lock_guard<std::mutex> lk(s->mu);
      s->pool.release(j->mem, j->mem_bytes);
      s->mem_epoch++;
      j->mem = nullptr;
      j->done = true;
      s->idle_slots++;
    }
    s->cv_sched.notify_one();
    s->cv_done.notify_all();
  }
}

// Admits pending jobs in arrival order while a slot is idle and memory fits.
// A head job that does not fit (including by fragmentation) waits for the
// next release of memory; later jobs do not overtake it.
void scheduler_main(stwo_hip_server *s) {
  std::unique_lock<std::mutex> lk(s->mu);
  for (;;) {
    s->cv_sched.wait(lk, [&] {
      if (s->stop && s->pending.empty()) return true;
      if (s->pending.empty() || s->idle_slots == 0) return false;
      return s->mem_epoch != s->blocked_epoch;
    });
    if (s->stop && s->pending.empty()) return;
    bool admitted_any = false;
    while (!s->pending.empty() && s->idle_slots > 0) {
      auto &j = s->pending.front();
      size_t need = job_bytes(j->desc);
      uint8_t *m = s->pool.alloc(need);
      if (!m) {
        s->blocked_epoch = s->mem_epoch;  // wait for a release
        break;
      }
      j->mem = m;
      j->mem_bytes = need;
      j->res.t_admit_ms = s->now_ms();
      s->idle_slots--;
      s->admitted.push_back(j);
      s->pending.pop_front();
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

int stwo_hip_server_submit(stwo_hip_server *s, const stwo_hip_job_desc *desc, uint64_t *job_id) {
  if (!s || !desc || !job_id) return STWO_HIP_EINVAL;
  const stwo_hip_job_desc &d = *desc;
  if (d.n_cols < 4 || d.log_size < 4 || d.log_blowup < 1 || d.log_size > 28 || d.log_blowup > 16 ||
      d.log_size + d.log_blowup > s->max_log ||
      d.fri.log_blowup_factor != d.log_blowup || d.fri.log_last_layer_degree_bound > 4 ||
      d.fri.line_fold_step < 1 || d.fri.line_fold_step > 28 || d.log_size + d.log_blowup - 1 < d.fri.log_last_layer_degree_bound + d.log_blowup)
    return STWO_HIP_EINVAL;
  auto j = std::make_shared<Job>();
  j->desc = d;
  memset(&j->res, 0, sizeof j->res);
  j->res.status = STWO_HIP_EBUSY;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->stop) return STWO_HIP_EINVAL;
    if (s->pending.size() >= s->max_queue) return STWO_HIP_EBUSY;
    if (job_bytes(d) > s->pool.capacity()) return STWO_HIP_ENOMEM;
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

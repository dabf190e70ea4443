// Context, device memory, table arena, profiling and seeded input fill.
//
// Replaces the reference's Obj-C runtime singleton and buffer entry points
// (SH/runtime.m:476-916, SYS/src/metal.rs:4442-4448) with an explicit
// per-device context that owns one HIP stream.
#include <stdint.h>
#include <string.h>

#include <algorithm>

#include "common.hpp"
#include "field.hpp"

namespace stwo_hip {

int set_error(stwo_hip_ctx *ctx, int code, const char *fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  if (ctx) ctx->last_error = buf;
  return code;
}

hipEvent_t take_event(stwo_hip_ctx *ctx) {
  if (!ctx->event_pool.empty()) {
    hipEvent_t e = ctx->event_pool.back();
    ctx->event_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

static void give_event(stwo_hip_ctx *ctx, hipEvent_t e) {
  if (e) ctx->event_pool.push_back(e);
}

int check_launch(stwo_hip_ctx *ctx, const char *name) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return set_error(ctx, STWO_HIP_EDEVICE, "launch of %s failed: %s", name, hipGetErrorString(e));
  return STWO_HIP_OK;
}

void prof_begin(stwo_hip_ctx *ctx, const char *name, hipEvent_t *start) {
  (void)name;
  *start = nullptr;
  if (!ctx->prof) return;
  *start = take_event(ctx);
  if (*start) (void)hipEventRecord(*start, ctx->stream);
}

void prof_end(stwo_hip_ctx *ctx, const char *name, hipEvent_t start) {
  if (!ctx->prof || !start) return;
  hipEvent_t stop = take_event(ctx);
  if (!stop) return;
  (void)hipEventRecord(stop, ctx->stream);
  ctx->prof_pending.push_back({name, start, stop});
}

static const size_t kArenaBytes = 8u << 20;
static const size_t kStageBytes = 1u << 20;
static const size_t kGatherBytes = 2u << 20;
static const int kEventPoolInit = 16;

static int arena_init(stwo_hip_ctx *ctx) {
  STWO_HIP_TRY(ctx, hipHostMalloc((void **)&ctx->h_arena, kArenaBytes, hipHostMallocDefault));
  STWO_HIP_TRY(ctx, hipMalloc((void **)&ctx->d_arena, kArenaBytes));
  ctx->arena_cap = kArenaBytes;
  ctx->arena_head = 0;
  STWO_HIP_TRY(ctx, hipHostMalloc((void **)&ctx->h_stage, kStageBytes, hipHostMallocDefault));
  ctx->stage_cap = kStageBytes;
  // coherent (fine-grained): the gather kernel's reads and writes of it are
  // not held in the GPU's caches across launches
  STWO_HIP_TRY(ctx, hipHostMalloc((void **)&ctx->h_gather, kGatherBytes, hipHostMallocCoherent));
  STWO_HIP_TRY(ctx, hipHostGetDevicePointer(&ctx->d_gather, ctx->h_gather, 0));
  ctx->gather_cap = kGatherBytes;
  return STWO_HIP_OK;
}

int download_many(stwo_hip_ctx *ctx, const DownloadReq *reqs, size_t n_reqs) {
  size_t total = 0;
  for (size_t i = 0; i < n_reqs; i++) total += reqs[i].n_words * 4;
  if (total > ctx->stage_cap) {
    STWO_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    for (size_t i = 0; i < n_reqs; i++)
      if (reqs[i].n_words)
        STWO_HIP_TRY(ctx, hipMemcpy(reqs[i].dst, reqs[i].src, reqs[i].n_words * 4, hipMemcpyDeviceToHost));
    return STWO_HIP_OK;
  }
  size_t off = 0;
  for (size_t i = 0; i < n_reqs; i++) {
    if (reqs[i].n_words)
      STWO_HIP_TRY(ctx, hipMemcpyAsync(ctx->h_stage + off, reqs[i].src, reqs[i].n_words * 4, hipMemcpyDeviceToHost,
                                       ctx->stream));
    off += reqs[i].n_words * 4;
  }
  STWO_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  off = 0;
  for (size_t i = 0; i < n_reqs; i++) {
    memcpy(reqs[i].dst, ctx->h_stage + off, reqs[i].n_words * 4);
    off += reqs[i].n_words * 4;
  }
  return STWO_HIP_OK;
}

// Makes host bytes [b, e) of the arena reusable.  Uploads complete in stream
// (= FIFO) order, so waiting for the newest overlapping upload retires it and
// every older one.  Completed uploads at the front are retired on the way.
static int arena_wait_range(stwo_hip_ctx *ctx, size_t b, size_t e) {
  size_t last = SIZE_MAX;
  for (size_t i = 0; i < ctx->inflight.size(); i++) {
    const auto &f = ctx->inflight[i];
    if (f.begin < e && b < f.end) last = i;
  }
  if (last != SIZE_MAX) {
    STWO_HIP_TRY(ctx, hipEventSynchronize(ctx->inflight[last].ev));
    for (size_t i = 0; i <= last; i++) {
      give_event(ctx, ctx->inflight.front().ev);
      ctx->inflight.pop_front();
    }
  }
  while (!ctx->inflight.empty() && hipEventQuery(ctx->inflight.front().ev) == hipSuccess) {
    give_event(ctx, ctx->inflight.front().ev);
    ctx->inflight.pop_front();
  }
  return STWO_HIP_OK;
}

int flush_uploads(stwo_hip_ctx *ctx) {
  if (ctx->pend_end <= ctx->pend_begin) return STWO_HIP_OK;
  const size_t b = ctx->pend_begin, e = ctx->pend_end;
  ctx->pend_begin = ctx->pend_end = 0;
  STWO_HIP_TRY(ctx, hipMemcpyAsync(ctx->d_arena + b, ctx->h_arena + b, e - b, hipMemcpyHostToDevice, ctx->stream));
  hipEvent_t ev = take_event(ctx);
  if (!ev) return set_error(ctx, STWO_HIP_EDEVICE, "hipEventCreate failed");
  STWO_HIP_TRY(ctx, hipEventRecord(ev, ctx->stream));
  ctx->inflight.push_back({b, e, ev});
  return STWO_HIP_OK;
}

static size_t table_span(size_t bytes) {
  size_t sz = (bytes + 255) & ~(size_t)255;
  return sz ? sz : 256;
}

int upload_tables(stwo_hip_ctx *ctx, size_t n, const void *const *srcs, const size_t *bytes, void **dptrs) {
  size_t total = 0;
  for (size_t i = 0; i < n; i++) total += table_span(bytes[i]);
  if (total > ctx->arena_cap / 2) {
    // Too large for the ring: a scratch buffer written synchronously (the
    // stream is drained first, since kernels of an earlier call may still
    // read the buffer's previous contents).
    void *d = nullptr;
    int rc = get_scratch(ctx, "large_tables", total, &d);
    if (rc) return rc;
    rc = flush_uploads(ctx);
    if (rc) return rc;
    STWO_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    size_t off = 0;
    for (size_t i = 0; i < n; i++) {
      if (bytes[i]) STWO_HIP_TRY(ctx, hipMemcpy((uint8_t *)d + off, srcs[i], bytes[i], hipMemcpyHostToDevice));
      dptrs[i] = (uint8_t *)d + off;
      off += table_span(bytes[i]);
    }
    return STWO_HIP_OK;
  }
  // One contiguous range of the ring for all of the call's tables: a table
  // handed out earlier is only safe from being overwritten once the kernels
  // that read it are enqueued (a later copy then follows them in stream
  // order), so the tables one launch sequence needs must not wrap onto each
  // other.
  size_t off = ctx->arena_head;
  if (off + total > ctx->arena_cap) off = 0;
  // A range that does not extend the pending one (the ring wrapped) starts a
  // new one: the pending bytes are copied first, so that the wait below also
  // covers them.
  int rc = STWO_HIP_OK;
  if (ctx->pend_end > ctx->pend_begin && off != ctx->pend_end) rc = flush_uploads(ctx);
  if (rc) return rc;
  rc = arena_wait_range(ctx, off, off + total);
  if (rc) return rc;
  if (ctx->pend_end <= ctx->pend_begin) ctx->pend_begin = off;
  for (size_t i = 0; i < n; i++) {
    if (bytes[i]) memcpy(ctx->h_arena + off, srcs[i], bytes[i]);
    dptrs[i] = ctx->d_arena + off;
    off += table_span(bytes[i]);
  }
  ctx->pend_end = off;
  ctx->arena_head = off;
  return STWO_HIP_OK;
}

int upload_tables_cached(stwo_hip_ctx *ctx, const char *key, size_t n, const void *const *srcs, const size_t *bytes,
                         void **dptrs) {
  if (ctx->job_base) return upload_tables(ctx, n, srcs, bytes, dptrs);
  size_t total = 0;
  for (size_t i = 0; i < n; i++) total += table_span(bytes[i]);
  const std::string skey = std::string("table_cache:") + key;
  std::vector<uint8_t> &shadow = ctx->table_shadow[key];
  bool same = shadow.size() == total;
  for (size_t i = 0, off = 0; same && i < n; off += table_span(bytes[i]), i++)
    same = !bytes[i] || memcmp(shadow.data() + off, srcs[i], bytes[i]) == 0;
  auto it = ctx->scratch.find(skey);
  if (same && it != ctx->scratch.end() && it->second.second >= total) {
    for (size_t i = 0, off = 0; i < n; off += table_span(bytes[i]), i++) dptrs[i] = (uint8_t *)it->second.first + off;
    return STWO_HIP_OK;
  }
  // New tables: through the ring for this call, then into the kept buffer.
  shadow.clear();
  int rc = upload_tables(ctx, n, srcs, bytes, dptrs);
  if (rc) return rc;
  void *dev = nullptr;
  rc = get_scratch(ctx, skey.c_str(), total, &dev);
  if (rc) return rc;
  rc = flush_uploads(ctx);
  if (rc) return rc;
  STWO_HIP_TRY(ctx, hipMemcpyAsync(dev, dptrs[0], total, hipMemcpyDeviceToDevice, ctx->stream));
  shadow.assign(total, 0);
  for (size_t i = 0, off = 0; i < n; off += table_span(bytes[i]), i++)
    if (bytes[i]) memcpy(shadow.data() + off, srcs[i], bytes[i]);
  return STWO_HIP_OK;
}

int arena_quiesce(stwo_hip_ctx *ctx) {
  int rc = flush_uploads(ctx);
  if (rc) return rc;
  STWO_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  while (!ctx->inflight.empty()) {
    give_event(ctx, ctx->inflight.front().ev);
    ctx->inflight.pop_front();
  }
  return STWO_HIP_OK;
}

int upload_table(stwo_hip_ctx *ctx, const void *src, size_t bytes, void **dptr) {
  return upload_tables(ctx, 1, &src, &bytes, dptr);
}

static int job_carve(stwo_hip_ctx *ctx, size_t bytes, void **dptr, const char *what) {
  const size_t off = (ctx->job_used + 255) & ~(size_t)255;
  if (off + bytes > ctx->job_cap)
    return set_error(ctx, STWO_HIP_ENOMEM, "%s of %zu bytes exceeds the job's memory (%zu of %zu used)", what, bytes,
                     ctx->job_used, ctx->job_cap);
  *dptr = ctx->job_base + off;
  ctx->job_used = off + bytes;
  return STWO_HIP_OK;
}

void job_mem_begin(stwo_hip_ctx *ctx, void *base, size_t cap) {
  ctx->job_base = (uint8_t *)base;
  ctx->job_cap = cap;
  ctx->job_used = 0;
  ctx->job_keys.clear();
}

void job_mem_end(stwo_hip_ctx *ctx) {
  for (const std::string &k : ctx->job_keys) ctx->scratch.erase(k);
  ctx->job_keys.clear();
  ctx->job_base = nullptr;
  ctx->job_cap = ctx->job_used = 0;
}

hipError_t ctx_malloc(stwo_hip_ctx *ctx, void **dptr, size_t bytes) {
  if (ctx->fail_allocs_after == 0) {
    *dptr = nullptr;
    return hipErrorOutOfMemory;
  }
  if (ctx->fail_allocs_after > 0) ctx->fail_allocs_after--;
  return hipMalloc(dptr, bytes);
}

int device_alloc(stwo_hip_ctx *ctx, size_t bytes, void **dptr, bool *from_job) {
  *from_job = ctx->job_base != nullptr;
  if (*from_job) return job_carve(ctx, bytes, dptr, "device table");
  if (ctx_malloc(ctx, dptr, bytes ? bytes : 256) != hipSuccess) {
    (void)hipGetLastError();
    *dptr = nullptr;
    return set_error(ctx, STWO_HIP_ENOMEM, "device allocation of %zu bytes failed", bytes);
  }
  return STWO_HIP_OK;
}

void device_free(void *p, bool from_job) {
  if (p && !from_job) (void)hipFree(p);
}

int get_scratch(stwo_hip_ctx *ctx, const char *key, size_t bytes, void **dptr) {
  if (ctx->job_base) {
    // Within a job: entries carved from the job block (a larger request for
    // a key carves a new piece; the block is released with the job).
    auto it = ctx->scratch.find(key);
    const bool carved = std::find(ctx->job_keys.begin(), ctx->job_keys.end(), key) != ctx->job_keys.end();
    if (it != ctx->scratch.end() && carved && it->second.second >= bytes) {
      *dptr = it->second.first;
      return STWO_HIP_OK;
    }
    void *p = nullptr;
    int rc = job_carve(ctx, bytes ? bytes : 256, &p, key);
    if (rc) return rc;
    if (it != ctx->scratch.end() && !carved) {
      // an entry from before the job: freed, the job's piece takes its key
      STWO_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
      (void)hipFree(it->second.first);
      ctx->scratch.erase(it);
    }
    ctx->scratch[key] = {p, bytes};
    if (!carved) ctx->job_keys.push_back(key);
    *dptr = p;
    return STWO_HIP_OK;
  }
  auto it = ctx->scratch.find(key);
  if (it != ctx->scratch.end() && it->second.second >= bytes) {
    *dptr = it->second.first;
    return STWO_HIP_OK;
  }
  if (it != ctx->scratch.end()) {
    STWO_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    (void)hipFree(it->second.first);
    ctx->scratch.erase(it);
  }
  void *p = nullptr;
  if (ctx_malloc(ctx, &p, bytes ? bytes : 256) != hipSuccess) {
    (void)hipGetLastError();
    return set_error(ctx, STWO_HIP_ENOMEM, "scratch '%s' of %zu bytes: out of device memory", key, bytes);
  }
  ctx->scratch[key] = {p, bytes};
  *dptr = p;
  return STWO_HIP_OK;
}

const stwo_hip_twiddles *check_twiddles(stwo_hip_ctx *ctx, const stwo_hip_twiddles *tw) {
  if (!tw || tw->device != ctx->device) return nullptr;
  return tw;
}

// ---- seeded synthetic inputs -------------------------------------------

__device__ __forceinline__ uint64_t splitmix64_step(uint64_t s) {
  uint64_t z = s + 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

__global__ void __launch_bounds__(256) fill_random_kernel(uint32_t *out, uint64_t n, uint64_t seed, uint64_t stream) {
  uint64_t base = (seed << 32) ^ (stream * 0xD1B54A32D192ED03ull);
  for (uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; j < n; j += (uint64_t)gridDim.x * blockDim.x) {
    uint64_t z = splitmix64_step(base ^ (j * 0x9E3779B97F4A7C15ull));
    out[j] = (uint32_t)(z % P);
  }
}

}  // namespace stwo_hip

using namespace stwo_hip;

extern "C" {

int stwo_hip_ctx_create(int device, stwo_hip_ctx **out) {
  if (!out) return STWO_HIP_EINVAL;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n) return STWO_HIP_EDEVICE;
  if (hipSetDevice(device) != hipSuccess) return STWO_HIP_EDEVICE;
  stwo_hip_ctx *ctx = new (std::nothrow) stwo_hip_ctx();
  if (!ctx) return STWO_HIP_ENOMEM;
  ctx->device = device;
  if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
    delete ctx;
    return STWO_HIP_EDEVICE;
  }
  if (arena_init(ctx) != STWO_HIP_OK) {
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return STWO_HIP_ENOMEM;
  }
  // Events for the arena's copies, created up front so that calls which
  // must not allocate find the pool stocked.
  for (int i = 0; i < kEventPoolInit; i++) {
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) break;
    ctx->event_pool.push_back(e);
  }
  if (eval_point_scratch_setup(ctx) != STWO_HIP_OK) {
    stwo_hip_ctx_destroy(ctx);
    return STWO_HIP_EDEVICE;
  }
  *out = ctx;
  return STWO_HIP_OK;
}

int stwo_hip_ctx_destroy(stwo_hip_ctx *ctx) {
  if (!ctx) return STWO_HIP_EINVAL;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  for (auto &f : ctx->inflight) (void)hipEventDestroy(f.ev);
  for (auto &p : ctx->prof_pending) {
    (void)hipEventDestroy(p.start);
    (void)hipEventDestroy(p.stop);
  }
  if (ctx->prove_plans_free) ctx->prove_plans_free(ctx, ctx->prove_plans);
  for (auto e : ctx->event_pool) (void)hipEventDestroy(e);
  for (auto &kv : ctx->scratch) (void)hipFree(kv.second.first);
  if (ctx->h_arena) (void)hipHostFree(ctx->h_arena);
  if (ctx->h_stage) (void)hipHostFree(ctx->h_stage);
  if (ctx->h_gather) (void)hipHostFree(ctx->h_gather);
  if (ctx->d_arena) (void)hipFree(ctx->d_arena);
  (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return STWO_HIP_OK;
}

int stwo_hip_last_error(const stwo_hip_ctx *ctx, char *buf, size_t len) {
  if (!ctx || !buf || !len) return STWO_HIP_EINVAL;
  size_t n = std::min(len - 1, ctx->last_error.size());
  memcpy(buf, ctx->last_error.data(), n);
  buf[n] = 0;
  return STWO_HIP_OK;
}

int stwo_hip_sync(stwo_hip_ctx *ctx) {
  if (!ctx) return STWO_HIP_EINVAL;
  int rc = flush_uploads(ctx);
  if (rc) return rc;
  STWO_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  return STWO_HIP_OK;
}

void *stwo_hip_stream(stwo_hip_ctx *ctx) { return ctx ? (void *)ctx->stream : nullptr; }

int stwo_hip_ctx_fail_allocs_after(stwo_hip_ctx *ctx, long long after) {
  if (!ctx) return STWO_HIP_EINVAL;
  ctx->fail_allocs_after = after < 0 ? -1 : after;
  return STWO_HIP_OK;
}

int stwo_hip_alloc(stwo_hip_ctx *ctx, size_t n_words, uint32_t **out) {
  if (!ctx || !out) return STWO_HIP_EINVAL;
  *out = nullptr;
  if (hipSetDevice(ctx->device) != hipSuccess) return set_error(ctx, STWO_HIP_EDEVICE, "hipSetDevice failed");
  hipError_t e = ctx_malloc(ctx, (void **)out, (n_words ? n_words : 1) * sizeof(uint32_t));
  if (e != hipSuccess) {
    (void)hipGetLastError();
    *out = nullptr;
    return set_error(ctx, STWO_HIP_ENOMEM, "hipMalloc(%zu words): %s", n_words, hipGetErrorString(e));
  }
  return STWO_HIP_OK;
}

int stwo_hip_free(stwo_hip_ctx *ctx, uint32_t *ptr) {
  if (!ctx) return STWO_HIP_EINVAL;
  if (!ptr) return STWO_HIP_OK;
  STWO_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  STWO_HIP_TRY(ctx, hipFree(ptr));
  return STWO_HIP_OK;
}

int stwo_hip_upload(stwo_hip_ctx *ctx, uint32_t *dst, const uint32_t *src, size_t n_words) {
  if (!ctx || (!dst && n_words) || (!src && n_words)) return STWO_HIP_EINVAL;
  STWO_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  if (n_words * 4 <= ctx->stage_cap) {
    memcpy(ctx->h_stage, src, n_words * 4);
    STWO_HIP_TRY(ctx, hipMemcpyAsync(dst, ctx->h_stage, n_words * 4, hipMemcpyHostToDevice, ctx->stream));
    STWO_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return STWO_HIP_OK;
  }
  STWO_HIP_TRY(ctx, hipMemcpy(dst, src, n_words * 4, hipMemcpyHostToDevice));
  return STWO_HIP_OK;
}

int stwo_hip_download(stwo_hip_ctx *ctx, uint32_t *dst, const uint32_t *src, size_t n_words) {
  if (!ctx || (!dst && n_words) || (!src && n_words)) return STWO_HIP_EINVAL;
  const DownloadReq r = {dst, src, n_words};
  return download_many(ctx, &r, 1);
}

int stwo_hip_copy(stwo_hip_ctx *ctx, uint32_t *dst, const uint32_t *src, size_t n_words) {
  if (!ctx || (!dst && n_words) || (!src && n_words)) return STWO_HIP_EINVAL;
  STWO_HIP_TRY(ctx, hipMemcpyAsync(dst, src, n_words * 4, hipMemcpyDeviceToDevice, ctx->stream));
  return STWO_HIP_OK;
}

int stwo_hip_memset_zero(stwo_hip_ctx *ctx, uint32_t *dst, size_t n_words) {
  if (!ctx || (!dst && n_words)) return STWO_HIP_EINVAL;
  STWO_HIP_TRY(ctx, hipMemsetAsync(dst, 0, n_words * 4, ctx->stream));
  return STWO_HIP_OK;
}

int stwo_hip_fill_random(stwo_hip_ctx *ctx, uint32_t *dst, size_t n, uint64_t seed, uint64_t stream) {
  if (!ctx || (!dst && n)) return STWO_HIP_EINVAL;
  uint64_t blocks = std::min<uint64_t>((n + 255) / 256, 65536);
  return launch(ctx, "fill_random", fill_random_kernel, dim3((uint32_t)blocks), dim3(256), 0, dst, (uint64_t)n, seed,
                stream);
}

int stwo_hip_prof_enable(stwo_hip_ctx *ctx, int enable) {
  if (!ctx) return STWO_HIP_EINVAL;
  ctx->prof = enable != 0;
  return STWO_HIP_OK;
}

static int prof_collect(stwo_hip_ctx *ctx) {
  if (ctx->prof_pending.empty()) return STWO_HIP_OK;
  STWO_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  for (auto &p : ctx->prof_pending) {
    float ms = 0;
    (void)hipEventElapsedTime(&ms, p.start, p.stop);
    auto &t = ctx->prof_totals[p.name];
    t.first += 1;
    t.second += ms;
    give_event(ctx, p.start);
    give_event(ctx, p.stop);
  }
  ctx->prof_pending.clear();
  return STWO_HIP_OK;
}

int stwo_hip_prof_reset(stwo_hip_ctx *ctx) {
  if (!ctx) return STWO_HIP_EINVAL;
  int rc = prof_collect(ctx);
  ctx->prof_totals.clear();
  return rc;
}

int stwo_hip_prof_read(stwo_hip_ctx *ctx, char *name_buf, size_t name_buf_len, uint64_t *counts, double *total_ms,
                       size_t max_entries, size_t *n_entries) {
  if (!ctx) return STWO_HIP_EINVAL;
  int rc = prof_collect(ctx);
  if (rc) return rc;
  size_t k = 0, pos = 0;
  for (auto &kv : ctx->prof_totals) {
    if (k >= max_entries) break;
    size_t need = kv.first.size() + 1;
    if (name_buf && pos + need <= name_buf_len) {
      memcpy(name_buf + pos, kv.first.c_str(), need);
      pos += need;
    }
    if (counts) counts[k] = kv.second.first;
    if (total_ms) total_ms[k] = kv.second.second;
    k++;
  }
  if (n_entries) *n_entries = k;
  return STWO_HIP_OK;
}

}  // extern "C"

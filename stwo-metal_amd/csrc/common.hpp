// Context, error handling, per-call table arena and launch profiling shared
// by every translation unit of libstwo_hip.
#pragma once
#include <hip/hip_runtime.h>
#include <stdarg.h>
#include <stdint.h>
#include <stdio.h>

#include <deque>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "stwo_hip.h"

struct stwo_hip_twiddles {
  int device;
  stwo_hip_coset root;
  size_t len;          // root size (2^root.log_size)
  uint32_t *d_tw;      // twiddle tree, CpuBackend layout
  uint32_t *d_itw;     // elementwise inverses
  uint32_t *d_tw2;     // both trees doubled (2 * t < 2^32), the radix-16 FFT's operands
  uint32_t *d_itw2;
};

namespace stwo_hip {

// Planner choices for one FFT shape (interpolate or evaluate of 2^n values on
// a domain of 2^N), kept per context; stwo_hip_fft_autotune sets them from
// device timings and every plan made afterwards on that context uses them.
struct FftTune {
  uint8_t t0 = 12;   // tile log of the contiguous low pass (10..13)
  uint8_t smax = 8;  // most layers per strided pass (5..8)
  uint8_t wg = 0;    // workgroup log: 0 = max(8, T - 4); 9 = 512 lanes for every tile size
  uint8_t pad = 0;
};
inline uint64_t fft_tune_key(uint32_t n, uint32_t N, bool inverse) {
  return ((uint64_t)n << 16) | ((uint64_t)N << 1) | (inverse ? 1u : 0u);
}

struct ProfPending {
  const char *name;
  hipEvent_t start, stop;
};

}  // namespace stwo_hip

struct stwo_hip_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string last_error;
  std::mutex mu;

  // Ring arena for small per-call tables (column pointer lists, per-column
  // parameters).  Host side is pinned so the H2D copy is a true async DMA;
  // the device side is read by kernels on the same stream, so stream order
  // alone protects it.  `inflight` keeps, per copy, its host byte range and
  // an event after it; host bytes are only rewritten once the copy that read
  // them has completed.  Tables written back to back form one pending range
  // [pend_begin, pend_end), copied by a single hipMemcpyAsync when the next
  // kernel is launched (flush_uploads).
  uint8_t *h_arena = nullptr;
  uint8_t *d_arena = nullptr;
  size_t arena_cap = 0;
  size_t arena_head = 0;
  size_t pend_begin = 0, pend_end = 0;
  struct Inflight {
    size_t begin, end;
    hipEvent_t ev;
  };
  std::deque<Inflight> inflight;
  std::vector<hipEvent_t> event_pool;

  // Launch profiling.
  bool prof = false;
  std::vector<stwo_hip::ProfPending> prof_pending;
  std::map<std::string, std::pair<uint64_t, double>> prof_totals;

  // Pinned staging for small synchronous copies (a pageable hipMemcpy costs
  // tens of microseconds more per call).
  uint8_t *h_stage = nullptr;
  size_t stage_cap = 0;

  // Pinned host buffer that the decommitment gather reads its op table from
  // and writes its result to directly (no copy either way), and its device
  // address.
  uint8_t *h_gather = nullptr;
  void *d_gather = nullptr;
  size_t gather_cap = 0;

  // Scratch device buffers kept across calls, keyed by purpose.
  std::map<std::string, std::pair<void *, size_t>> scratch;
  // Host copies of the tables upload_tables_cached last placed in its
  // device buffers, by call site.
  std::map<std::string, std::vector<uint8_t>> table_shadow;

  // FFT planner choices by shape (fft_tune_key); absent shapes use the defaults.
  std::map<uint64_t, stwo_hip::FftTune> fft_tune;

  // Job memory (serving): while set, every device buffer the library would
  // allocate for its own use (get_scratch, batch-plan tables) is carved from
  // [job_base, job_base + job_cap) instead, and dropped when the job ends
  // (job_mem_end), so a job's device use stays inside its budget block.
  uint8_t *job_base = nullptr;
  size_t job_cap = 0, job_used = 0;

  // Allocation-failure injection (stwo_hip_ctx_fail_allocs_after): device
  // allocations left before they fail, or -1 (off).
  long long fail_allocs_after = -1;
  std::vector<std::string> job_keys;  // scratch entries carved from the job block

  // Launch plans of the last prove shape, kept for the next prove of that
  // shape; prove_plans_free destroys them (shape change, context destruction).
  void *prove_plans = nullptr;
  void (*prove_plans_free)(stwo_hip_ctx *, void *) = nullptr;
};

namespace stwo_hip {

int set_error(stwo_hip_ctx *ctx, int code, const char *fmt, ...);

#define STWO_HIP_TRY(ctx, expr)                                                                   \
  do {                                                                                            \
    hipError_t _e = (expr);                                                                       \
    if (_e != hipSuccess)                                                                         \
      return ::stwo_hip::set_error((ctx), STWO_HIP_EDEVICE, "%s: %s (%s:%d)", #expr,              \
                                   hipGetErrorString(_e), __FILE__, __LINE__);                    \
  } while (0)

#define STWO_HIP_CHECK_ARG(ctx, cond, ...)                                   \
  do {                                                                       \
    if (!(cond)) return ::stwo_hip::set_error((ctx), STWO_HIP_EINVAL, __VA_ARGS__); \
  } while (0)

// Copies `bytes` of host data into the arena and returns the device address
// (16-byte aligned).  The host-to-device copy is enqueued on the context
// stream by the next flush_uploads (every launch() makes one), together with
// the other tables written since the last one.
int upload_table(stwo_hip_ctx *ctx, const void *src, size_t bytes, void **dptr);
// Several tables for one launch sequence, from one contiguous range of the
// ring (they never overwrite each other before their kernels are enqueued).
// Every table a sequence of launches reads must come from one such call
// made before the sequence's first launch.  Sets larger than half the ring
// go through a scratch buffer with a synchronous copy.
// hipMalloc through the context's allocation-failure injection: every
// device allocation of the library goes through here.
hipError_t ctx_malloc(stwo_hip_ctx *ctx, void **dptr, size_t bytes);
int upload_tables(stwo_hip_ctx *ctx, size_t n, const void *const *srcs, const size_t *bytes, void **dptrs);
// upload_tables for tables that repeat from call to call (column lists of
// buffers a caller reuses): a call whose tables equal the previous call's
// under `key` gets the device copy kept from then, with no copy on the
// stream; otherwise the tables go through the ring and are copied into the
// kept buffer (stream-ordered) for the next call.  Within a serving job it
// is upload_tables.
int upload_tables_cached(stwo_hip_ctx *ctx, const char *key, size_t n, const void *const *srcs, const size_t *bytes,
                         void **dptrs);

// Enqueues the copy of the pending tables (none: nothing to do).
int flush_uploads(stwo_hip_ctx *ctx);
// Copies the pending tables, drains the stream and returns every copy's
// event to the pool (which context creation stocks), so that the next
// upload and launch take pooled events and create none.
int arena_quiesce(stwo_hip_ctx *ctx);
// Resolves, once per device, the address of the fixed scratch array of
// stwo_hip_eval_at_point_noalloc (context creation calls it).
int eval_point_scratch_setup(stwo_hip_ctx *ctx);

// Returns a device scratch buffer of at least `bytes` for `key` (grown on
// demand; growing synchronizes the stream before freeing the old buffer).
// Within a job (job_mem_begin), carved from the job's memory block instead.
int get_scratch(stwo_hip_ctx *ctx, const char *key, size_t bytes, void **dptr);
// Device memory owned by the caller of device_alloc (a batch plan's tables):
// from the job block within a job (device_free then does nothing), else
// hipMalloc.  *from_job tells which.
int device_alloc(stwo_hip_ctx *ctx, size_t bytes, void **dptr, bool *from_job);
void device_free(void *p, bool from_job);
// Starts / ends a job whose library-internal device buffers come from
// [base, base + cap).  job_mem_end drops every scratch entry carved from it.
void job_mem_begin(stwo_hip_ctx *ctx, void *base, size_t cap);
void job_mem_end(stwo_hip_ctx *ctx);
// Several device-to-host reads with one stream synchronization (through the
// pinned staging buffer when they fit).
struct DownloadReq {
  uint32_t *dst;
  const uint32_t *src;
  size_t n_words;
};
int download_many(stwo_hip_ctx *ctx, const DownloadReq *reqs, size_t n_reqs);

hipEvent_t take_event(stwo_hip_ctx *ctx);
void prof_begin(stwo_hip_ctx *ctx, const char *name, hipEvent_t *start);
void prof_end(stwo_hip_ctx *ctx, const char *name, hipEvent_t start);
int check_launch(stwo_hip_ctx *ctx, const char *name);

// Launches `kernel` on the context stream, bracketed by profiling events
// when profiling is enabled, and reports launch errors.
template <typename... KArgs, typename... Args>
int launch(stwo_hip_ctx *ctx, const char *name, void (*kernel)(KArgs...), dim3 grid, dim3 block, size_t shmem,
           Args... args) {
  if (grid.x == 0 || grid.y == 0 || grid.z == 0) return STWO_HIP_OK;
  int frc = flush_uploads(ctx);
  if (frc) return frc;
  hipEvent_t start = nullptr;
  prof_begin(ctx, name, &start);
  hipLaunchKernelGGL(kernel, grid, block, shmem, ctx->stream, static_cast<KArgs>(args)...);
  int rc = check_launch(ctx, name);
  prof_end(ctx, name, start);
  return rc;
}

const stwo_hip_twiddles *check_twiddles(stwo_hip_ctx *ctx, const stwo_hip_twiddles *tw);

}  // namespace stwo_hip

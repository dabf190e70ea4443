// Circle IFFT / RFFT (PolyOps::interpolate / evaluate) and bit reversal.
//
// Semantics (CpuBackend, CPUB/circle.rs:28-80, 199-250, 313-354): a column of
// log size N has N butterfly layers.  Layer s >= 1 (a "line" layer) pairs
// elements 2^s apart; the pair starting at index i0 uses twiddle
// line_tw[s-1][i0 >> (s+1)], where line layer j of the domain's twiddles is
// the tree slice [len - 2^(N-1-j), len - 2^(N-2-j)) (domain_line_twiddles_from_tree,
// STWO/core/poly/utils.rs:98-114).  Layer 0 (the "circle" layer) pairs
// neighbours with twiddles [y, -y, -x, x] built from consecutive pairs (x, y)
// of line layer 0 (circle_twiddles_from_line_twiddles).  RFFT runs layers
// N-1 .. 0 with `butterfly`; IFFT runs 0 .. N-1 with `ibutterfly` on the
// inverse tree and scales by 1/2^N (STWO/core/fft.rs:5-21).
//
// Evaluating 2^n coefficients on a domain of 2^N points zero-pads them; the
// top N-n layers then only copy (butterfly(v, 0, t) = (v, v)), so the first
// pass reads coefficient index (g mod 2^n) for output index g and starts at
// layer n-1 -- the 2^(N-n) independent sub-FFTs of SimdBackend's
// evaluate_into (SIMD/circle.rs:413-440), fused into one pass over the
// output.
//
// MI355X mapping: a pass moves a 4096-element tile (16 KB) of one column
// through LDS.  The tile owns up to 12 consecutive layers: the lowest pass
// covers layers [0, 12) on a contiguous chunk; higher layers run in strided
// passes of <= 6 layers whose tile rows are 64 consecutive elements (256 B
// coalesced runs).  Columns up to 2^12 take one pass, up to 2^18 two, up to
// 2^24 three.  blockIdx.y selects the column, so one launch covers a whole
// group of same-size columns.
#include <algorithm>
#include <map>
#include <vector>

#include "common.hpp"
#include "field.hpp"
#include "internal.hpp"

namespace stwo_hip {

constexpr uint32_t kTileLog = 12;
constexpr uint32_t kTile = 1u << kTileLog;
constexpr uint32_t kFftThreads = 256;
constexpr uint32_t kStridedMaxLayers = 6;


struct FftPass {
  const FftCol *cols;     // one entry per blockIdx.y
  const uint32_t *tw_end; // one past the end of the (inverse) twiddle tree
  uint32_t N;             // domain log size
  uint32_t lo;            // lowest layer of the tile's butterfly bits
  uint32_t m;             // tile butterfly bits: layers [lo, lo + m)
  uint32_t w;             // tile inner bits (consecutive elements per row)
  uint32_t s_begin, s_end;// layers processed, subset of [lo, lo + m)
  uint32_t in_log;        // input index = g mod 2^in_log
  uint32_t inverse;
  uint32_t scale;         // multiply outputs by this when nonzero
};

__device__ __forceinline__ uint32_t line_twiddle(const uint32_t *tw_end, uint32_t N, uint32_t s, uint32_t h) {
  return tw_end[(int64_t)h - ((int64_t)1 << (N - s))];
}

// Circle twiddle for pair h (elements 2h, 2h+1): [y, -y, -x, x] from the pair
// (x, y) = line layer 0 entries 2(h/4), 2(h/4)+1 (CPUB/circle.rs:332-354).
__device__ __forceinline__ uint32_t circle_twiddle(const uint32_t *tw_end, uint32_t N, uint32_t h) {
  const uint32_t *l0 = tw_end - ((int64_t)1 << (N - 1));
  uint32_t x = l0[2 * (h >> 2)], y = l0[2 * (h >> 2) + 1];
  uint32_t v = (h & 2) ? x : y;
  bool neg = ((h & 3) == 1) || ((h & 3) == 2);
  return neg ? m31_neg(v) : v;
}

__device__ __forceinline__ uint32_t tile_global(const FftPass &p, uint32_t tile, uint32_t t) {
  uint32_t b = t >> p.w;
  uint32_t il = t & ((1u << p.w) - 1);
  uint32_t ih_bits = p.lo - p.w;
  uint32_t ih = tile & ((1u << ih_bits) - 1);
  uint32_t o = tile >> ih_bits;
  return (o << (p.lo + p.m)) | (b << p.lo) | (ih << p.w) | il;
}

__global__ void __launch_bounds__(kFftThreads) fft_tile_kernel(FftPass p) {
  __shared__ uint32_t lds[kTile];
  const FftCol col = p.cols[blockIdx.y];
  const uint32_t tile = blockIdx.x;
  const uint32_t T = p.m + p.w;
  const uint32_t tsize = 1u << T;
  const uint32_t in_mask = (uint32_t)((1ull << p.in_log) - 1);
  for (uint32_t t = threadIdx.x; t < tsize; t += kFftThreads) lds[t] = col.in[tile_global(p, tile, t) & in_mask];
  __syncthreads();
  const uint32_t n_layers = p.s_end - p.s_begin;
  for (uint32_t k = 0; k < n_layers; k++) {
    const uint32_t s = p.inverse ? p.s_begin + k : p.s_end - 1 - k;
    const uint32_t qb = s - p.lo + p.w;  // tile bit of this layer
    for (uint32_t pr = threadIdx.x; pr < tsize / 2; pr += kFftThreads) {
      uint32_t i0 = ((pr >> qb) << (qb + 1)) | (pr & ((1u << qb) - 1));
      uint32_t i1 = i0 | (1u << qb);
      uint32_t h = tile_global(p, tile, i0) >> (s + 1);
      uint32_t t = s ? line_twiddle(p.tw_end, p.N, s, h) : circle_twiddle(p.tw_end, p.N, h);
      uint32_t v0 = lds[i0], v1 = lds[i1];
      if (p.inverse) {
        lds[i0] = m31_add(v0, v1);
        lds[i1] = m31_mul(m31_sub(v0, v1), t);
      } else {
        uint32_t tmp = m31_mul(v1, t);
        lds[i0] = m31_add(v0, tmp);
        lds[i1] = m31_sub(v0, tmp);
      }
    }
    __syncthreads();
  }
  for (uint32_t t = threadIdx.x; t < tsize; t += kFftThreads) {
    uint32_t v = lds[t];
    if (p.scale) v = m31_mul(v, p.scale);
    col.out[tile_global(p, tile, t)] = v;
  }
}

// Domains of 2 and 4 points: the closed forms of CPUB/circle.rs:35-59
// (interpolate) and :215-235 (evaluate).  c[] holds host-computed constants.
struct SmallFft {
  const FftCol *cols;
  uint32_t N, in_log, inverse;
  uint32_t c0, c1, c2, c3;
};

__global__ void fft_small_kernel(SmallFft a) {
  const FftCol col = a.cols[blockIdx.x];
  if (threadIdx.x != 0) return;
  uint32_t mask = (1u << a.in_log) - 1;
  uint32_t n = 1u << a.N;
  uint32_t v[4];
  for (uint32_t i = 0; i < n; i++) v[i] = (i <= mask) ? col.in[i] : 0;
  auto bf = [](uint32_t &x0, uint32_t &x1, uint32_t t) {
    uint32_t tmp = m31_mul(x1, t);
    x1 = m31_sub(x0, tmp);
    x0 = m31_add(x0, tmp);
  };
  auto ibf = [](uint32_t &x0, uint32_t &x1, uint32_t t) {
    uint32_t tmp = x0;
    x0 = m31_add(tmp, x1);
    x1 = m31_mul(m31_sub(tmp, x1), t);
  };
  if (!a.inverse) {
    if (a.N == 1) {
      bf(v[0], v[1], a.c1);  // y
    } else {
      bf(v[0], v[2], a.c0);  // x
      bf(v[1], v[3], a.c0);
      bf(v[0], v[1], a.c1);  // y
      bf(v[2], v[3], m31_neg(a.c1));
    }
  } else {
    if (a.N == 1) {
      ibf(v[0], v[1], a.c1);  // y_inv
      v[0] = m31_mul(v[0], a.c2);
      v[1] = m31_mul(v[1], a.c2);  // n_inv
    } else {
      ibf(v[0], v[1], a.c1);  // y_inv
      ibf(v[2], v[3], m31_neg(a.c1));
      ibf(v[0], v[2], a.c0);  // x_inv
      ibf(v[1], v[3], a.c0);
      for (int i = 0; i < 4; i++) v[i] = m31_mul(v[i], a.c2);
    }
  }
  for (uint32_t i = 0; i < n; i++) col.out[i] = v[i];
}

// ---- bit reversal -----------------------------------------------------------
// In-place v[i] <-> v[bitrev(i)] (STWO/core/utils.rs:91-101).  A tile of
// 64 x 64 elements, rows a, columns b, is exchanged with the tile at
// (rev(b-block), rev(a-block)) through LDS so both reads and writes are
// 256-byte coalesced runs; tiles on the diagonal swap within themselves.
// Index split for log n = 2k + r bits: i = [hi: 6 | mid: r' | lo: 6].
constexpr uint32_t kBrLog = 6;

__global__ void __launch_bounds__(256) bit_reverse_small_kernel(uint32_t *v, uint32_t log_n) {
  uint32_t n = 1u << log_n;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    uint32_t j = bitrev32(i, log_n);
    if (i < j) {
      uint32_t a = v[i], b = v[j];
      v[i] = b;
      v[j] = a;
    }
  }
}

// For log_n >= 2 * kBrLog: i = (hi << (log_n - 6)) | (mid << 6) | lo with
// hi, lo 6 bits; bitrev(i) = (rev6(lo) << (log_n - 6)) | (rev(mid) << 6) | rev6(hi).
// Block (mid) handles the 64x64 tile {hi, lo} at `mid` and its partner at rev(mid).
// 512 lanes: the two tiles' 33 KB of LDS allow 4 workgroups per CU, and the
// kernel is bound by memory latency, so each workgroup carries 8 waves.
// VEC (16-byte aligned column): each lane moves 4 consecutive words of a row
// with one 16-byte load or store.
constexpr uint32_t kBrThreads = 512;
template <bool VEC>
__global__ void __launch_bounds__(kBrThreads) bit_reverse_tiled_kernel(uint32_t *v, uint32_t log_n) {
  __shared__ uint32_t ta[64][65];
  __shared__ uint32_t tb[64][65];
  const uint32_t mid_bits = log_n - 2 * kBrLog;
  const uint32_t mid = blockIdx.x;
  const uint32_t rmid = bitrev32(mid, mid_bits);
  if (rmid < mid) return;  // the partner block does the swap
  const uint32_t hs = log_n - kBrLog;
  if (VEC) {
    for (uint32_t e = threadIdx.x; e < 1024; e += kBrThreads) {
      const uint32_t hi = e >> 4, lo = (e & 15) << 2;
      const u32x4 a = *(gconst_u4 *)as_global(v + ((hi << hs) | (mid << kBrLog) | lo));
      const u32x4 b = *(gconst_u4 *)as_global(v + ((hi << hs) | (rmid << kBrLog) | lo));
      ta[hi][lo] = a.x;
      ta[hi][lo + 1] = a.y;
      ta[hi][lo + 2] = a.z;
      ta[hi][lo + 3] = a.w;
      tb[hi][lo] = b.x;
      tb[hi][lo + 1] = b.y;
      tb[hi][lo + 2] = b.z;
      tb[hi][lo + 3] = b.w;
    }
  } else {
    for (uint32_t e = threadIdx.x; e < 4096; e += kBrThreads) {
      uint32_t hi = e >> 6, lo = e & 63;
      ta[hi][lo] = v[(hi << hs) | (mid << kBrLog) | lo];
      tb[hi][lo] = v[(hi << hs) | (rmid << kBrLog) | lo];
    }
  }
  __syncthreads();
  // new v[(hi, mid, lo)] = old v[bitrev] = old v[(rev(lo), rmid, rev(hi))] = tb[rev lo][rev hi]
  if (VEC) {
    for (uint32_t e = threadIdx.x; e < 1024; e += kBrThreads) {
      const uint32_t hi = e >> 4, lo = (e & 15) << 2, rh = bitrev32(hi, kBrLog);
      u32x4 a, b;
      a.x = tb[bitrev32(lo, kBrLog)][rh];
      a.y = tb[bitrev32(lo + 1, kBrLog)][rh];
      a.z = tb[bitrev32(lo + 2, kBrLog)][rh];
      a.w = tb[bitrev32(lo + 3, kBrLog)][rh];
      *(g_u4 *)as_global(v + ((hi << hs) | (mid << kBrLog) | lo)) = a;
      if (rmid != mid) {
        b.x = ta[bitrev32(lo, kBrLog)][rh];
        b.y = ta[bitrev32(lo + 1, kBrLog)][rh];
        b.z = ta[bitrev32(lo + 2, kBrLog)][rh];
        b.w = ta[bitrev32(lo + 3, kBrLog)][rh];
        *(g_u4 *)as_global(v + ((hi << hs) | (rmid << kBrLog) | lo)) = b;
      }
    }
  } else {
    for (uint32_t e = threadIdx.x; e < 4096; e += kBrThreads) {
      uint32_t hi = e >> 6, lo = e & 63;
      uint32_t rh = bitrev32(hi, kBrLog), rl = bitrev32(lo, kBrLog);
      v[(hi << hs) | (mid << kBrLog) | lo] = tb[rl][rh];
      if (rmid != mid) v[(hi << hs) | (rmid << kBrLog) | lo] = ta[rl][rh];
    }
  }
}

// ---- register-resident radix-16 passes -------------------------------------
//
// A pass owns a tile of 2^T elements (T <= 12) of one column: tile bits
// [w, w + m) are butterfly layers [lo, lo + m) (global bit = tile bit - w + lo)
// and tile bits [0, w) are consecutive elements ("inner").  2^(T-4) lanes
// share a tile, each holding 16 elements in VGPRs.  A stage works on a
// 4-bit register group at tile bits [b0, b0 + 4): lane `t` holds tile
// indices base(t) | (j << b0), j < 16, and applies the layers at group bits
// [r_lo, r_hi) with unrolled register butterflies.  Between stages the tile
// is exchanged through LDS (one padding word per 16 avoids most bank
// conflicts).  The first stage reads HBM directly and the last writes HBM
// directly.
//
// Twiddles: for the group of lane t the pair (j0, j0 | 1 << r) of layer s
// uses h = (g(base) >> (s + 1)) | (j0 >> (r + 1)), so a stage needs 2^(3-r)
// consecutive words of tree level s per layer -- 15 words per lane per
// stage, loaded once.  The circle layer (s = 0) reads 4 words (two (x, y)
// pairs of line layer 0) for its 8 pairs.

// `t2` is the twiddle doubled (the passes read the doubled trees).
template <bool INV>
__device__ __forceinline__ void bfly(uint32_t &a, uint32_t &b, uint32_t t2) {
  if (INV) {
    uint32_t s = m31_add(a, b);
    b = m31_mul_dbl(m31_sub(a, b), t2);
    a = s;
  } else {
    uint32_t tmp = m31_mul_dbl(b, t2);
    b = m31_sub(a, tmp);
    a = m31_add(a, tmp);
  }
}

// Twiddles of one stage: for register bit r, 8 >> r consecutive words of
// tree level s = sh + r starting at gb >> (s + 1); the circle layer (s = 0)
// is expanded from two (x, y) pairs of line layer 0 into [y, -y, -x, x].
struct StageTw {
  uint32_t t0[8], t1[4], t2[2], t3[1];
};

// Word `byte_off / 4` of a table whose base is wave-uniform: the base stays
// in SGPRs and the lane part is a 32-bit offset (global_load ... saddr).
__device__ __forceinline__ uint32_t gld(const uint32_t *base, uint32_t byte_off) {
  return *(gconst_u32 *)((gconst_u8 *)as_global(base) + byte_off);
}
__device__ __forceinline__ void gst(uint32_t *base, uint32_t byte_off, uint32_t v) {
  *(g_u32 *)((g_u8 *)as_global(base) + byte_off) = v;
}
__device__ __forceinline__ u32x4 gld4(const uint32_t *base, uint32_t byte_off) {
  return *(gconst_u4 *)((gconst_u8 *)as_global(base) + byte_off);
}
__device__ __forceinline__ void gst4(uint32_t *base, uint32_t byte_off, u32x4 v) {
  *(g_u4 *)((g_u8 *)as_global(base) + byte_off) = v;
}

__device__ __forceinline__ void load_level(uint32_t *dst, int cnt, const uint32_t *tw_end, uint32_t N, uint32_t s,
                                           uint32_t gb) {
  const uint32_t hs = gb >> (s + 1);
  if (s == 0) {
    const uint32_t *l0 = tw_end - ((size_t)1 << (N - 1));
    const uint32_t o = (hs >> 1) << 2;
    uint32_t w4[4];
#pragma unroll
    for (int k = 0; k < 4; k++) w4[k] = gld(l0, o + 4 * k);
#pragma unroll
    for (int u = 0; u < 8; u++) {
      uint32_t xx = w4[2 * (u >> 2)], yy = w4[2 * (u >> 2) + 1];
      uint32_t v = (u & 2) ? xx : yy;
      // doubled negation: 2 (P - v / 2) = 2P - v < 2^32 (for v = 0 it is 2P, still 0 mod P)
      dst[u] = ((u & 3) == 1 || (u & 3) == 2) ? 2u * P - v : v;
    }
    return;
  }
  const uint32_t *lv = tw_end - ((size_t)1 << (N - s));
  const uint32_t o = hs << 2;
#pragma unroll
  for (int u = 0; u < 8; u++)
    if (u < cnt) dst[u] = gld(lv, o + 4 * u);
}

// Loads only the levels of the stage's active register bits [r_lo, r_hi).
__device__ __forceinline__ void load_stage_twiddles(StageTw &t, const uint32_t *tw_end, uint32_t N, uint32_t sh,
                                                    uint32_t gb, uint32_t r_lo, uint32_t r_hi) {
  if (r_lo <= 0 && 0 < r_hi) load_level(t.t0, 8, tw_end, N, sh + 0, gb);
  if (r_lo <= 1 && 1 < r_hi) load_level(t.t1, 4, tw_end, N, sh + 1, gb);
  if (r_lo <= 2 && 2 < r_hi) load_level(t.t2, 2, tw_end, N, sh + 2, gb);
  if (r_lo <= 3 && 3 < r_hi) load_level(t.t3, 1, tw_end, N, sh + 3, gb);
}

template <int R>
__device__ __forceinline__ uint32_t stage_tw(const StageTw &t, int u) {
  return R == 0 ? t.t0[u] : R == 1 ? t.t1[u] : R == 2 ? t.t2[u] : t.t3[u];
}

// Layers at register bits [RLO, RHI) of the group (forward: high to low).
template <bool INV, int RLO, int RHI>
__device__ __forceinline__ void stage_layers(uint32_t (&x)[16], const StageTw &t) {
#pragma unroll
  for (int rr = 0; rr < RHI - RLO; rr++) {
    const int r = INV ? RLO + rr : RHI - 1 - rr;
#pragma unroll
    for (int j0 = 0; j0 < 16; j0++)
      if (!((j0 >> r) & 1)) {
        const int u = j0 >> (r + 1);
        const uint32_t tw = r == 0 ? stage_tw<0>(t, u) : r == 1 ? stage_tw<1>(t, u) : r == 2 ? stage_tw<2>(t, u)
                                                                                           : stage_tw<3>(t, u);
        bfly<INV>(x[j0], x[j0 | (1 << r)], tw);
      }
  }
}

template <bool INV>
__device__ __forceinline__ void stage_dispatch(uint32_t (&x)[16], const StageTw &t, uint32_t r_lo, uint32_t r_hi) {
  switch (r_lo * 8 + r_hi) {
    case 4: stage_layers<INV, 0, 4>(x, t); break;
    case 3: stage_layers<INV, 0, 3>(x, t); break;
    case 2: stage_layers<INV, 0, 2>(x, t); break;
    case 1: stage_layers<INV, 0, 1>(x, t); break;
    case 10: stage_layers<INV, 1, 2>(x, t); break;
    case 11: stage_layers<INV, 1, 3>(x, t); break;
    case 12: stage_layers<INV, 1, 4>(x, t); break;
    case 19: stage_layers<INV, 2, 3>(x, t); break;
    case 20: stage_layers<INV, 2, 4>(x, t); break;
    case 28: stage_layers<INV, 3, 4>(x, t); break;
    default: break;  // (0, 0): copy-only stage; every other shape is rejected on the host
  }
}

// LDS exchange of a lane's 16 group elements.  Element j of the group at
// tile bit B0 with base index `base` (group bits zero) lives at word
// pad(base | j << B0) = (base + (base >> 4)) + (j << B0) + ((j << B0) >> 4)
// (pad(i) = i + (i >> 4): one padding word per 16),
// since the two parts have disjoint bits: the per-element part is a
// compile-time offset, selected by a wave-uniform switch on B0.
template <int B0>
__device__ __forceinline__ constexpr uint32_t lds_off(int j) {
  return ((uint32_t)j << B0) + (((uint32_t)j << B0) >> 4);
}
template <int B0>
__device__ __forceinline__ void lds_put16(uint32_t *Lv, const uint32_t (&x)[16]) {
#pragma unroll
  for (int j = 0; j < 16; j++) Lv[lds_off<B0>(j)] = x[j];
}
template <int B0>
__device__ __forceinline__ void lds_get16(const uint32_t *Lv, uint32_t (&x)[16]) {
#pragma unroll
  for (int j = 0; j < 16; j++) x[j] = Lv[lds_off<B0>(j)];
}
// Group bit b0 is at most kMaxTileLog - 4 = 9.
__device__ __forceinline__ void lds_put(uint32_t *L, uint32_t b0, uint32_t base, const uint32_t (&x)[16]) {
  uint32_t *Lv = L + base + (base >> 4);
  switch (b0) {
    case 0: lds_put16<0>(Lv, x); break;
    case 1: lds_put16<1>(Lv, x); break;
    case 2: lds_put16<2>(Lv, x); break;
    case 3: lds_put16<3>(Lv, x); break;
    case 4: lds_put16<4>(Lv, x); break;
    case 5: lds_put16<5>(Lv, x); break;
    case 6: lds_put16<6>(Lv, x); break;
    case 7: lds_put16<7>(Lv, x); break;
    case 8: lds_put16<8>(Lv, x); break;
    default: lds_put16<9>(Lv, x); break;
  }
}
__device__ __forceinline__ void lds_get(const uint32_t *L, uint32_t b0, uint32_t base, uint32_t (&x)[16]) {
  const uint32_t *Lv = L + base + (base >> 4);
  switch (b0) {
    case 0: lds_get16<0>(Lv, x); break;
    case 1: lds_get16<1>(Lv, x); break;
    case 2: lds_get16<2>(Lv, x); break;
    case 3: lds_get16<3>(Lv, x); break;
    case 4: lds_get16<4>(Lv, x); break;
    case 5: lds_get16<5>(Lv, x); break;
    case 6: lds_get16<6>(Lv, x); break;
    case 7: lds_get16<7>(Lv, x); break;
    case 8: lds_get16<8>(Lv, x); break;
    default: lds_get16<9>(Lv, x); break;
  }
}

// One workgroup's share of a pass: workgroup `wg` of the pass's grid, with
// 2^WG_LOG lanes (a tile of 2^T elements takes 2^(T-4) of them).
//
// Data movement: global loads and stores use a lane-consecutive layout --
// the top register group (b0 = T - 4, lane `local` holds tile indices
// local + j * 2^(T-4)) or any group with b0 >= 5 (runs of >= 32 lanes).
// Between stages whose groups differ, the tile is exchanged through LDS.
// Each stage's twiddles are loaded before the exchange that precedes it.
//
// With rep_log > 0 (the first RFFT pass of a blowup), the tile grid covers
// the coefficient domain 2^(N - rep_log) and each workgroup writes the tile
// of every one of the 2^rep_log output blocks: the coefficients are read
// once from HBM (repeats hit L2) and the blocks differ only in twiddles.
// One stage of a compile-time schedule: the register group at tile bit B0,
// its register bits [RLO, RHI) (all 4 layers, or some of them).  PREV is
// the previous stage's group, or -1 when it is only known at run time (the
// first stage follows the load layout).
template <bool INV, int B0, int PREV, int RHI = 4, int RLO = 0, typename GB, typename GI>
__device__ __forceinline__ void fixed_stage(const FftPass2 &p, uint32_t (&x)[16], uint32_t *L, uint32_t &cur,
                                            uint32_t local, GB group_base, GI gidx) {
  // Keep the scheduler from hoisting later stages' twiddle loads above this
  // point: with the whole schedule visible it would otherwise keep several
  // stages' twiddles live at once and spill.
  __builtin_amdgcn_sched_barrier(0);
  const uint32_t base = ((local >> B0) << (B0 + 4)) | (local & ((1u << B0) - 1));
  const uint32_t gb = gidx(base);
  const uint32_t sh = B0 - p.w + p.lo;
  StageTw t;
  load_stage_twiddles(t, p.tw_end, p.N, sh, gb, RLO, RHI);
  if constexpr (PREV >= 0) {
    if constexpr (PREV != B0) {
      const uint32_t cb = ((local >> PREV) << (PREV + 4)) | (local & ((1u << PREV) - 1));
      __syncthreads();
      lds_put16<PREV>(L + cb + (cb >> 4), x);
      __syncthreads();
      lds_get16<B0>(L + base + (base >> 4), x);
    }
  } else {
    if (cur != B0) {
      __syncthreads();
      lds_put(L, cur, group_base(cur), x);
      __syncthreads();
      lds_get16<B0>(L + base + (base >> 4), x);
    }
  }
  cur = B0;
  stage_layers<INV, RLO, RHI>(x, t);
}

// The fixed schedules need about 125 VGPRs (4 waves per SIMD); asking for 5
// makes them spill ~30 VGPRs and run at half speed.  Measured on one box,
// 256 x 2^20 -> 2^22: 9.06 ms per step against 9.48 ms with the run-time
// schedule at 5 waves (the 8-layer strided RFFT pass 1.95 ms against 2.28).
//
constexpr uint32_t T_FIXED12 = 12;  // tile bits of the fixed three-stage passes

// NS > 0: the pass's stages are the compile-time schedule B0A, B0B, B0C
// (first NS of them), each applying all 4 layers of its group; the host only
// picks such a kernel for passes whose plan is exactly that schedule.
// NS = 0: the schedule is read from p.st at run time.
template <bool INV, int WG_LOG, int NS, int B0A, int B0B, int B0C>
__device__ __forceinline__ void fft_pass_body(const FftPass2 &p, uint32_t wg, uint32_t *lds) {
  const uint32_t tpt_log = p.T - 4;  // log2 lanes per tile
  const uint32_t local = threadIdx.x & ((1u << tpt_log) - 1);
  const uint32_t slot = threadIdx.x >> tpt_log;
  const uint64_t gtile = ((uint64_t)wg << (WG_LOG - tpt_log)) | slot;
  const uint32_t dom_log = p.N - p.rep_log;
  const uint32_t tpc_log = dom_log - p.T;  // log2 tiles per column
  // Tiles are numbered column-major: concurrent workgroups stream through
  // consecutive 16 KB chunks of one column, spread over the HBM channels.
  // (Numbering them column-fastest, so that an XCD's L2 serves one tile's
  // twiddles to many columns, cut the low pass's fetched bytes by a third
  // but measured 2% slower over the whole step: the 256 columns' same offsets
  // were then read at once.)
  const bool active = gtile < ((uint64_t)p.n_cols << tpc_log);
  // Lanes past the last tile (a workgroup holds several small tiles) take
  // tile 0 of column 0: they load no data and store nothing, but they do
  // read twiddles, and tile 0's twiddle reads stay inside the table.
  uint32_t tile = 0, col_idx = 0;
  if (active) {
    tile = (uint32_t)(gtile & ((1ull << tpc_log) - 1));
    col_idx = (uint32_t)(gtile >> tpc_log);
  }
  const uint32_t *in = nullptr;
  uint32_t *out = nullptr;
  if (active) {
    in = p.cols[col_idx].in;
    out = p.cols[col_idx].out;
  }
  // A tile spans whole waves: the column is wave-uniform.  The fixed
  // schedules (NS > 0) only run tiles of 2^12 or more and NS = -1 marks a
  // run-time schedule launched only for tiles of 2^10 or more, so for them
  // this is known at compile time and the pointers stay in SGPRs (loads and
  // stores then take an SGPR base and a 32-bit lane offset instead of 64-bit
  // lane addresses, which also frees half the kernel's VGPRs).
  if (NS != 0 || tpt_log >= 6) {
    in = (const uint32_t *)readfirstlane_u64((uint64_t)in);
    out = (uint32_t *)readfirstlane_u64((uint64_t)out);
  }
  uint32_t *L = lds + slot * ((1u << p.T) + (1u << (p.T - 4)));
  const uint32_t ih_bits = p.lo - p.w;
  const uint32_t ih = tile & ((1u << ih_bits) - 1);
  const uint32_t o = tile >> ih_bits;
  const uint32_t in_mask = (uint32_t)((1ull << p.in_log) - 1);
  const uint32_t wmask = (1u << p.w) - 1;
  const uint32_t top_b0 = p.T - 4;
  const uint32_t n_rep = 1u << p.rep_log;
  uint32_t x[16];
  for (uint32_t rep = 0; rep < n_rep; rep++) {
    const uint32_t tile_gb = (rep << dom_log) | (o << (p.lo + p.m)) | (ih << p.w);
    // global index of tile index `b` (group bits zero) and the shift of group bit 0
    auto gidx = [&](uint32_t b) { return tile_gb | ((b >> p.w) << p.lo) | (b & wmask); };
    auto group_base = [&](uint32_t b0) {
      return ((local >> b0) << (b0 + 4)) | (local & ((1u << b0) - 1));
    };
    uint32_t cur = top_b0;
    if (p.vec_io & 1) {
      // --- load straight into layout b0 = 0: 16 consecutive words per lane
      // (in_log >= 4, so the input mask keeps them consecutive)
      if (rep > 0) __syncthreads();
      if (active) {
        const uint32_t vo = (gidx(local << 4) & in_mask) << 2;
#pragma unroll
        for (int k = 0; k < 4; k++) {
          const u32x4 v = gld4(in, vo + 16 * k);
          x[4 * k] = v.x;
          x[4 * k + 1] = v.y;
          x[4 * k + 2] = v.z;
          x[4 * k + 3] = v.w;
        }
      }
      cur = 0;
    } else {
      // --- load in the top layout
      // (gb | j << sh) & in_mask = (gb & in_mask) | ((j << sh) & in_mask):
      // a lane offset plus a wave-uniform one.
      // One SGPR base and a lane byte offset per element: (j << sh) & in_mask
      // = (j & (in_mask >> sh)) << sh, since in_mask is a low-bit mask.
      // (Sixteen 64-bit element bases in SGPRs overflowed the SGPR file and
      // spilled into VGPR lanes.)
      // The fixed 12-layer passes (NS == 3) have w = lo = 0 and an input of
      // at least one tile (both checked at launch), so there the shift is the
      // compile-time top_b0 = 8 and the element offsets are constants.
      const uint32_t gb = gidx(local), sh = top_b0 - p.w + p.lo;
      if (rep > 0) __syncthreads();
      if (active) {
        const uint32_t vo = (gb & in_mask) << 2, jmask = in_mask >> sh;
        if (NS == 3) {
#pragma unroll
          for (int j = 0; j < 16; j++) x[j] = gld(in, vo + ((uint32_t)j << (T_FIXED12 - 4 + 2)));
        } else {
#pragma unroll
          for (int j = 0; j < 16; j++) x[j] = gld(in, vo + ((((uint32_t)j) & jmask) << (sh + 2)));
        }
      }
    }
    if constexpr (NS == 1) {
      // One stage: register bits [B0B, B0C) of the group at tile bit B0A.
      fixed_stage<INV, B0A, -1, B0C, B0B>(p, x, L, cur, local, group_base, gidx);
    } else if constexpr (NS == 2 && B0C != 0) {
      // A strided pass of 4 + B0C layers: the group at tile bit 6 (= w)
      // applies only its low B0C register bits (first stage when inverse).
      fixed_stage<INV, B0A, -1, INV ? B0C : 4>(p, x, L, cur, local, group_base, gidx);
      fixed_stage<INV, B0B, B0A, INV ? 4 : B0C>(p, x, L, cur, local, group_base, gidx);
    } else if constexpr (NS > 0) {
      fixed_stage<INV, B0A, -1>(p, x, L, cur, local, group_base, gidx);
      if constexpr (NS > 1) fixed_stage<INV, B0B, B0A>(p, x, L, cur, local, group_base, gidx);
      if constexpr (NS > 2) fixed_stage<INV, B0C, B0B>(p, x, L, cur, local, group_base, gidx);
    } else
    for (uint32_t si = 0; si < p.n_stages; si++) {
      const FftStage S = p.st[si];
      const uint32_t base = group_base(S.b0);
      const uint32_t gb = gidx(base);
      const uint32_t sh = S.b0 - p.w + p.lo;
      StageTw t;
      load_stage_twiddles(t, p.tw_end, p.N, sh, gb, S.r_lo, S.r_hi);
      if (S.b0 != cur) {
        __syncthreads();
        lds_put(L, cur, group_base(cur), x);
        __syncthreads();
        lds_get(L, S.b0, base, x);
        cur = S.b0;
      }
      stage_dispatch<INV>(x, t, S.r_lo, S.r_hi);
    }
    if (p.scale) {
#pragma unroll
      for (int j = 0; j < 16; j++) x[j] = m31_mul(x[j], p.scale);
    }
    if ((p.vec_io & 2) && cur == 0) {
      // --- store straight from layout b0 = 0
      if (active) {
        const uint32_t vo = gidx(local << 4) << 2;
#pragma unroll
        for (int k = 0; k < 4; k++) {
          const u32x4 v = {x[4 * k], x[4 * k + 1], x[4 * k + 2], x[4 * k + 3]};
          gst4(out, vo + 16 * k, v);
        }
      }
      continue;
    }
    // --- store in a lane-consecutive layout
    if (cur != top_b0 && cur < 5) {
      __syncthreads();
      lds_put(L, cur, group_base(cur), x);
      __syncthreads();
      lds_get(L, top_b0, local, x);
      cur = top_b0;
    }
    if (active) {
      const uint32_t gb = gidx(group_base(cur)), sh = cur - p.w + p.lo;
      const uint32_t vo = gb << 2;
      if (NS == 3) {
        // fixed 12-layer passes end in the top layout (B0C is 0 or 8; 0 was
        // moved there above) with w = lo = 0: sh = 8
#pragma unroll
        for (int j = 0; j < 16; j++) gst(out, vo + ((uint32_t)j << (T_FIXED12 - 4 + 2)), x[j]);
      } else {
#pragma unroll
        for (int j = 0; j < 16; j++) gst(out, vo + ((uint32_t)j << (sh + 2)), x[j]);
      }
    }
  }
}

template <bool INV, int WG_LOG, int NS = 0, int B0A = 0, int B0B = 0, int B0C = 0>
__global__ void __launch_bounds__(1 << WG_LOG) __attribute__((amdgpu_waves_per_eu(NS > 0 ? 4 : 5))) fft_pass_kernel(FftPass2 p) {
  __shared__ uint32_t lds[(16u << WG_LOG) + (1u << WG_LOG)];
  fft_pass_body<INV, WG_LOG, NS, B0A, B0B, B0C>(p, blockIdx.x, lds);
}

// Several passes of different column groups (sizes) in one launch: group g
// owns workgroups [wg_start[g], wg_start[g + 1]).  Used by the batched path,
// where every column size contributes its pass of the same level.  The
// schedule parameters are those of fft_pass_kernel: every group of a launch
// matches them (NS = 3 or 2: the compile-time schedules; NS = -1: tiles of
// 2^10 or more; NS = 0: any).
template <bool INV, int WG_LOG, int NS, int B0A = 0, int B0B = 0, int B0C = 0>
__global__ void __launch_bounds__(1 << WG_LOG) __attribute__((amdgpu_waves_per_eu(NS > 0 ? 4 : 5))) fft_pass_multi_kernel(const FftPass2 *groups,
                                                                    const uint32_t *wg_start, uint32_t n_groups) {
  __shared__ uint32_t lds[(16u << WG_LOG) + (1u << WG_LOG)];
  // The group whose range holds this workgroup: the last g with
  // wg_start[g] <= blockIdx.x (wg_start[0] = 0, increasing), by bisection.
  uint32_t g = 0, hi = n_groups;
  while (hi - g > 1) {
    const uint32_t mid = (g + hi) >> 1;
    if (wg_start[mid] <= blockIdx.x) g = mid; else hi = mid;
  }
  // Read in place (wave-uniform scalar loads): a private copy would put the
  // run-time stage list in scratch memory.
  const FftPass2 &p = groups[g];
  fft_pass_body<INV, WG_LOG, NS, B0A, B0B, B0C>(p, blockIdx.x - wg_start[g], lds);
}

// ---- fused LDE pass ----------------------------------------------------------
//
// The IFFT's last pass and the RFFT's first (blowup) pass of the same column
// cover the same layers on the same tile geometry (interpolate ends with the
// top layers, evaluate starts with them), so one kernel runs both on a tile
// held in registers: load the tile of evaluations, run the inverse stages,
// scale by 1/2^n (the coefficients; stored only when the caller keeps them),
// then for each of the 2^b output blocks run the forward stages with that
// block's twiddles and store.  The coefficients never go to HBM and back:
// TreeBuilder::extend_evals (STWO/prover/pcs/mod.rs:718-735) drops them
// unless store_polynomials_coefficients is set (:347).
//
// K is the compile-time schedule both passes share (fixed_schedule's kind):
// 2 = the 8-layer pair on a 2^13 tile (inverse stages at tile bits 5 then 9,
// forward 9 then 5); 3 / 4 / 5 = 5 / 6 / 7 strided layers (inverse: the
// partial group at tile bit 6, then the full one at T - 4; forward the
// reverse); 6..12 = one group at tile bit 6 with register bits kShort[K - 6].
// The inverse stages end in the group the forward stages start in, so each
// output block needs no LDS exchange before its first stage.  Both passes
// tile the coefficient domain (pf.rep_log = b), and pi.T >= 10 (wave-uniform
// columns).  A pass pair without such a schedule is not fused (the run-time
// stage lists made the fused pass slower than the two passes it replaces).
template <int K>
struct LdeSchedule {
  // full group, partial group, partial register bits [0, PART) (kinds 2..5);
  // short kinds: register bits [RLO, RHI) of the group at tile bit 6
  static constexpr int FULL = K == 2 ? 9 : K == 3 ? 7 : K == 4 ? 8 : K == 5 ? 9 : 6;
  static constexpr int PARTB0 = K == 2 ? 5 : 6;
  static constexpr int PART = K == 2 ? 4 : K == 3 ? 1 : K == 4 ? 2 : K == 5 ? 3 : 0;
  static constexpr int RLO = K >= 6 ? (K == 7 ? 1 : K == 8 ? 2 : K == 9 ? 3 : 0) : 0;
  static constexpr int RHI = K >= 6 ? (K <= 9 ? 4 : K == 10 ? 1 : K == 11 ? 2 : 3) : 4;
};

template <int WG_LOG, int K>
__device__ __forceinline__ void lde_pass_body(const FftPass2 &pi, const FftPass2 &pf, uint32_t store_coeffs,
                                              uint32_t wg, uint32_t *lds) {
  using S = LdeSchedule<K>;
  const uint32_t tpt_log = pi.T - 4;
  const uint32_t local = threadIdx.x & ((1u << tpt_log) - 1);
  const uint32_t slot = threadIdx.x >> tpt_log;
  const uint64_t gtile = ((uint64_t)wg << (WG_LOG - tpt_log)) | slot;
  const uint32_t dom_log = pi.N;  // the coefficient domain; pf.N - pf.rep_log is the same
  const uint32_t tpc_log = dom_log - pi.T;
  const bool active = gtile < ((uint64_t)pi.n_cols << tpc_log);
  uint32_t tile = 0, col_idx = 0;
  if (active) {
    tile = (uint32_t)(gtile & ((1ull << tpc_log) - 1));
    col_idx = (uint32_t)(gtile >> tpc_log);
  }
  // Tiles of 2^10 or more span whole waves: the column is wave-uniform.
  const uint32_t *ev = (const uint32_t *)readfirstlane_u64(active ? (uint64_t)pi.cols[col_idx].in : 0);
  uint32_t *co = (uint32_t *)readfirstlane_u64(active ? (uint64_t)pi.cols[col_idx].out : 0);
  uint32_t *out = (uint32_t *)readfirstlane_u64(active ? (uint64_t)pf.cols[col_idx].out : 0);
  uint32_t *L = lds + slot * ((1u << pi.T) + (1u << (pi.T - 4)));
  const uint32_t ih_bits = pi.lo - pi.w;
  const uint32_t ih = tile & ((1u << ih_bits) - 1);
  const uint32_t o = tile >> ih_bits;
  const uint32_t wmask = (1u << pi.w) - 1;
  const uint32_t top_b0 = pi.T - 4;
  const uint32_t tile_gb0 = (o << (pi.lo + pi.m)) | (ih << pi.w);
  auto gidx0 = [&](uint32_t b) { return tile_gb0 | ((b >> pi.w) << pi.lo) | (b & wmask); };
  auto group_base = [&](uint32_t b0) { return ((local >> b0) << (b0 + 4)) | (local & ((1u << b0) - 1)); };
  uint32_t x[16];
  const uint32_t sh_top = top_b0 - pi.w + pi.lo;
  if (active) {
    const uint32_t vo = gidx0(local) << 2;
#pragma unroll
    for (int j = 0; j < 16; j++) x[j] = gld(ev, vo + ((uint32_t)j << (sh_top + 2)));
  }
  uint32_t cur = top_b0;
  if constexpr (K >= 6) {
    fixed_stage<true, 6, -1, S::RHI, S::RLO>(pi, x, L, cur, local, group_base, gidx0);
  } else {
    fixed_stage<true, S::PARTB0, -1, S::PART>(pi, x, L, cur, local, group_base, gidx0);
    fixed_stage<true, S::FULL, S::PARTB0>(pi, x, L, cur, local, group_base, gidx0);
  }
#pragma unroll
  for (int j = 0; j < 16; j++) x[j] = m31_mul(x[j], pi.scale);
  if (store_coeffs && active) {
    // store in a lane-consecutive layout (group bit >= 5 or the top layout)
    uint32_t y[16];
#pragma unroll
    for (int j = 0; j < 16; j++) y[j] = x[j];
    uint32_t c2 = cur;
    if (c2 != top_b0 && c2 < 5) {
      __syncthreads();
      lds_put(L, c2, group_base(c2), y);
      __syncthreads();
      lds_get(L, top_b0, local, y);
      c2 = top_b0;
    }
    const uint32_t vo = gidx0(group_base(c2)) << 2, sh = c2 - pi.w + pi.lo;
#pragma unroll
    for (int j = 0; j < 16; j++) gst(co, vo + ((uint32_t)j << (sh + 2)), y[j]);
  }
  const uint32_t last_inv = cur;
  const uint32_t n_rep = 1u << pf.rep_log;
  for (uint32_t rep = 0; rep < n_rep; rep++) {
    const uint32_t tile_gb = (rep << dom_log) | tile_gb0;
    auto gidx = [&](uint32_t b) { return tile_gb | ((b >> pi.w) << pi.lo) | (b & wmask); };
    uint32_t y[16];
#pragma unroll
    for (int j = 0; j < 16; j++) y[j] = x[j];
    cur = last_inv;
    if constexpr (K >= 6) {
      fixed_stage<false, 6, 6, S::RHI, S::RLO>(pf, y, L, cur, local, group_base, gidx);
    } else {
      fixed_stage<false, S::FULL, S::FULL>(pf, y, L, cur, local, group_base, gidx);
      fixed_stage<false, S::PARTB0, S::FULL, S::PART>(pf, y, L, cur, local, group_base, gidx);
    }
    if (cur != top_b0 && cur < 5) {
      __syncthreads();
      lds_put(L, cur, group_base(cur), y);
      __syncthreads();
      lds_get(L, top_b0, local, y);
      cur = top_b0;
    }
    if (active) {
      const uint32_t vo = gidx(group_base(cur)) << 2, sh = cur - pi.w + pi.lo;
#pragma unroll
      for (int j = 0; j < 16; j++) gst(out, vo + ((uint32_t)j << (sh + 2)), y[j]);
    }
  }
}

template <int WG_LOG, int K>
__global__ void __launch_bounds__(1 << WG_LOG) __attribute__((amdgpu_waves_per_eu(4)))
lde_pass_kernel(FftPass2 pi, FftPass2 pf, uint32_t store_coeffs) {
  __shared__ uint32_t lds[(16u << WG_LOG) + (1u << WG_LOG)];
  lde_pass_body<WG_LOG, K>(pi, pf, store_coeffs, blockIdx.x, lds);
}

// The fused pass of several column groups in one launch (the batched path):
// group g owns workgroups [wg_start[g], wg_start[g + 1]) and the pair
// (pis[g], pfs[g]); every group matches the kernel's schedule (K).
template <int WG_LOG, int K>
__global__ void __launch_bounds__(1 << WG_LOG) __attribute__((amdgpu_waves_per_eu(4)))
lde_pass_multi_kernel(const FftPass2 *pis, const FftPass2 *pfs, const uint32_t *wg_start, uint32_t n_groups,
                      uint32_t store_coeffs) {
  __shared__ uint32_t lds[(16u << WG_LOG) + (1u << WG_LOG)];
  uint32_t g = 0, hi = n_groups;
  while (hi - g > 1) {
    const uint32_t mid = (g + hi) >> 1;
    if (wg_start[mid] <= blockIdx.x) g = mid; else hi = mid;
  }
  const FftPass2 &pi = pis[g], &pf = pfs[g];
  lde_pass_body<WG_LOG, K>(pi, pf, store_coeffs, blockIdx.x - wg_start[g], lds);
}

// Launches of the fused pass by schedule kind (wg9: kinds 2 and 5; wg8: 3, 4
// and the short kinds).  Profiling names: lde_pass_wg9 / lde_pass_wg8
// (lde_pass_batched for the multi-group form).
#define STWO_LDE_KINDS(X) X(9, 2) X(8, 3) X(8, 4) X(9, 5) X(8, 6) X(8, 7) X(8, 8) X(8, 9) X(8, 10) X(8, 11) X(8, 12)

static int launch_lde(stwo_hip_ctx *ctx, int kind, const FftPass2 &pi, const FftPass2 &pf, uint32_t wgs,
                      uint32_t store_coeffs) {
#define STWO_LDE_ONE(WG, K) \
  if (kind == K)            \
    return launch(ctx, "lde_pass_wg" #WG, lde_pass_kernel<WG, K>, dim3(wgs), dim3(1u << WG), 0, pi, pf, store_coeffs);
  STWO_LDE_KINDS(STWO_LDE_ONE)
#undef STWO_LDE_ONE
  return set_error(ctx, STWO_HIP_EINVAL, "internal: no fused schedule of kind %d", kind);
}

int launch_lde_multi(stwo_hip_ctx *ctx, const FftPass2 *d_pis, const FftPass2 *d_pfs, const uint32_t *d_wg_start,
                     uint32_t n_groups, uint32_t total_wgs, uint32_t wg_log, int kind, uint32_t store_coeffs) {
  (void)wg_log;  // implied by the kind
#define STWO_LDE_MULTI(WG, K)                                                                                    \
  if (kind == K)                                                                                                 \
    return launch(ctx, "lde_pass_batched", lde_pass_multi_kernel<WG, K>, dim3(total_wgs), dim3(1u << WG), 0, d_pis, \
                  d_pfs, d_wg_start, n_groups, store_coeffs);
  STWO_LDE_KINDS(STWO_LDE_MULTI)
#undef STWO_LDE_MULTI
  return set_error(ctx, STWO_HIP_EINVAL, "internal: no fused schedule of kind %d", kind);
}

int lde_kind(const PassLaunch &last_inv, const PassLaunch &first_fwd) {
  const uint32_t k = last_inv.p.fixed;
  if (k < 2 || k > 12 || first_fwd.p.fixed != k || last_inv.wg_log != first_fwd.wg_log) return 0;
  const uint32_t wg = (k == 2 || k == 5) ? 9 : 8;
  return last_inv.wg_log == wg ? (int)k : 0;
}

bool lde_fusable(const PassLaunch &last_inv, const PassLaunch &first_fwd, uint32_t log_blowup) {
  const FftPass2 &li = last_inv.p, &ff = first_fwd.p;
  return li.T == ff.T && li.lo == ff.lo && li.w == ff.w && li.m == ff.m && li.lo > 0 && li.T >= 10 &&
         ff.rep_log == log_blowup && last_inv.wgs == first_fwd.wgs && lde_kind(last_inv, first_fwd) != 0;
}

// Stages covering layer tile bits [A, B) of a tile of T bits, forward order.
static std::vector<FftStage> plan_stages(uint32_t A, uint32_t B, uint32_t T) {
  std::vector<FftStage> st;
  uint32_t hi = B;
  while (hi > A) {
    uint32_t lo_s = (hi >= A + 4) ? hi - 4 : A;
    uint32_t b0 = std::min(lo_s, T - 4);
    st.push_back({b0, lo_s - b0, hi - b0});
    hi = lo_s;
  }
  return st;
}

// ---- planning ---------------------------------------------------------------

struct PassPlan {
  uint32_t lo, m, w, s_begin, s_end;
};

constexpr uint32_t kStridedMaxLayers2 = 8;
constexpr uint32_t kMaxTileLog = 13;  // 8192-element tiles: 34 KB of LDS, 512 lanes

// Passes of the radix-16 kernel covering layers [0, L) of a domain of log
// size N (N >= 4), forward order: strided passes of <= tune.smax (default 8)
// layers, then the contiguous low pass of min(N, tune.t0) (default 12) tile
// bits.
static std::vector<PassPlan> plan_passes2(uint32_t N, uint32_t L, const FftTune &tune) {
  std::vector<PassPlan> fwd;
  const uint32_t smax = std::min<uint32_t>(kStridedMaxLayers2, std::max<uint32_t>(5, tune.smax));
  uint32_t t0 = std::min<uint32_t>(N, std::min<uint32_t>(kMaxTileLog, std::max<uint32_t>(10, tune.t0)));
  if (L > t0) {
    uint32_t rest = L - t0;
    uint32_t npass = (rest + smax - 1) / smax;
    uint32_t hi = L;
    // Full 8-layer passes from the top (the compile-time two-stage schedule
    // of a 2^13 tile, which also serves a blowup's first pass), the
    // remainder in the last strided pass.
    for (uint32_t k = 0; k < npass; k++) {
      uint32_t left = hi - t0;
      uint32_t m = k + 1 < npass ? smax : left;
      uint32_t lo = hi - m;
      // The tile's row bits must hold a whole 4-bit register group: a pass of
      // fewer than 4 layers takes 4 rows, the extra ones (batch bits) being
      // the global bits just above the layers, or below them at the top.
      uint32_t rows = std::max<uint32_t>(m, 4);
      uint32_t row_lo = std::min(lo, L - rows);
      // 2^w consecutive elements per tile row: 256 B runs, 128 B for 8-row tiles.
      uint32_t w = std::min<uint32_t>(6, kMaxTileLog - rows);
      fwd.push_back({row_lo, rows, w, lo, hi});
      hi = lo;
    }
  }
  // With a blowup (L < N) and no strided pass, the low pass tiles the
  // coefficient domain so that its output blocks can repeat the tile.
  uint32_t tl = (L < N && L <= t0 && L >= 4) ? L : t0;
  fwd.push_back({0, tl, 0, 0, std::min(L, tl)});
  return fwd;
}

// Passes covering layers [0, L) of a domain of log size N (N >= 3), in
// forward (top layers first) order.
static std::vector<PassPlan> plan_passes(uint32_t N, uint32_t L) {
  std::vector<PassPlan> fwd;
  uint32_t t0 = std::min(N, kTileLog);
  if (L > t0) {
    uint32_t rest = L - t0;
    uint32_t npass = (rest + kStridedMaxLayers - 1) / kStridedMaxLayers;
    uint32_t hi = L;
    // Full 8-layer passes from the top (the compile-time two-stage schedule
    // of a 2^13 tile, which also serves a blowup's first pass), the
    // remainder in the last strided pass.
    for (uint32_t k = 0; k < npass; k++) {
      uint32_t left = hi - t0;
      uint32_t m = k + 1 < npass ? kStridedMaxLayers2 : left;
      uint32_t lo = hi - m;
      fwd.push_back({lo, m, kTileLog - m, lo, hi});
      hi = lo;
    }
  }
  fwd.push_back({0, t0, 0, 0, std::min(L, t0)});
  return fwd;
}

// Validates that the circle domain with half coset `dom` is a doubling of
// the tree's root (CPUB/circle.rs:205 asserts is_doubling_of).
static int check_domain(stwo_hip_ctx *ctx, stwo_hip_coset dom, const stwo_hip_twiddles *tw) {
  STWO_HIP_CHECK_ARG(ctx, tw != nullptr && tw->device == ctx->device, "twiddles missing or from another device");
  STWO_HIP_CHECK_ARG(ctx, dom.log_size <= tw->root.log_size, "domain half coset log %u exceeds twiddle root log %u",
                     dom.log_size, tw->root.log_size);
  uint32_t k = tw->root.log_size - dom.log_size;
  uint32_t expect = (uint32_t)(((uint64_t)tw->root.initial_index << k) & IDX_MASK);
  STWO_HIP_CHECK_ARG(ctx, (dom.initial_index & IDX_MASK) == expect,
                     "domain (initial %u, log %u) is not a doubling of the twiddle root (initial %u, log %u)",
                     dom.initial_index, dom.log_size, tw->root.initial_index, tw->root.log_size);
  return STWO_HIP_OK;
}

// The compile-time schedules: every stage applies its group's 4 layers, and
// the groups are T-4, T-8, ... (forward) or the reverse (inverse).  These are
// the passes of columns of log size >= 16 split 12 + 8 + ..., the low pass
// (T = 12) and the 8-layer strided passes (T = 13).
static bool full_schedule(const FftPass2 &p, const uint32_t *b0s, uint32_t ns) {
  if (p.n_stages != ns) return false;
  for (uint32_t i = 0; i < ns; i++)
    if (p.st[i].b0 != b0s[i] || p.st[i].r_lo != 0 || p.st[i].r_hi != 4) return false;
  return true;
}

// A strided pass of m = 5, 6 or 7 layers on a tile of 2^(m + 6) (w = 6): the
// full group at tile bit T - 4 and the low m - 4 register bits of the group
// at tile bit 6 (forward: full first; inverse: partial first).
static bool partial_schedule(const FftPass2 &p, bool inverse, uint32_t m) {
  if (p.n_stages != 2 || p.w != 6 || p.T != m + 6) return false;
  const FftStage full = {p.T - 4, 0, 4}, part = {6, 0, m - 4};
  const FftStage &a = p.st[inverse ? 1 : 0], &b = p.st[inverse ? 0 : 1];
  return a.b0 == full.b0 && a.r_lo == full.r_lo && a.r_hi == full.r_hi && b.b0 == part.b0 &&
         b.r_lo == part.r_lo && b.r_hi == part.r_hi;
}

// The single-stage passes of 1 to 4 strided layers on a 2^10 tile (w = 6,
// the group at tile bit 6, which is also the load and store layout): the
// stage's register bits [r_lo, r_hi) for kinds 6 + k.
static const uint8_t kShort[7][2] = {{0, 4}, {1, 4}, {2, 4}, {3, 4}, {0, 1}, {0, 2}, {0, 3}};

// FftPass2::fixed for a pass: 1 = 12 layers in 3 full stages (wg8), 2 = 8
// strided layers in 2 full stages on a 2^13 tile (wg9), 3 / 4 / 5 = 5 / 6 / 7
// strided layers (partial_schedule; wg8, wg8, wg9), 6..12 = kShort (wg8),
// 0 = run-time schedule.
static uint32_t fixed_schedule(const FftPass2 &p, bool inverse, uint32_t wg_log) {
  static const uint32_t kFwd12[3] = {8, 4, 0}, kInv12[3] = {0, 4, 8}, kFwd13[2] = {9, 5}, kInv13[2] = {5, 9};
  if (wg_log == 8 && p.T == T_FIXED12 && p.w == 0 && p.lo == 0 && p.in_log >= T_FIXED12 &&
      full_schedule(p, inverse ? kInv12 : kFwd12, 3))
    return 1;
  if (wg_log == 9 && p.T == 13 && full_schedule(p, inverse ? kInv13 : kFwd13, 2)) return 2;
  if (wg_log == 8 && partial_schedule(p, inverse, 5)) return 3;
  if (wg_log == 8 && partial_schedule(p, inverse, 6)) return 4;
  if (wg_log == 9 && partial_schedule(p, inverse, 7)) return 5;
  if (wg_log == 8 && p.T == 10 && p.w == 6 && p.n_stages == 1 && p.st[0].b0 == 6)
    for (uint32_t k = 0; k < 7; k++)
      if (p.st[0].r_lo == kShort[k][0] && p.st[0].r_hi == kShort[k][1]) return 6 + k;
  return 0;
}

// Pass parameters (in execution order) of one group of same-shape columns on
// the radix-16 path (N >= 4).  `d_tab` is the group's FftCol table for the
// first pass; later forward passes run in place on the outputs and use
// `d_tab_inplace` (entries {out, out}).
int make_group_passes(stwo_hip_ctx *ctx, uint32_t n, uint32_t N, bool inverse, const FftCol *d_tab,
                      const FftCol *d_tab_inplace, uint32_t n_cols, const uint32_t *tw_end, bool aligned16,
                      std::vector<PassLaunch> &out, const FftTune *tune) {
  out.clear();
  STWO_HIP_CHECK_ARG(ctx, N >= 4 && N <= 28 && n <= N, "internal: radix-16 path needs 4 <= N <= 28 (N = %u)", N);
  FftTune tn;
  if (tune) {
    tn = *tune;
  } else {
    auto it = ctx->fft_tune.find(fft_tune_key(n, N, inverse));
    if (it != ctx->fft_tune.end()) tn = it->second;
  }
  uint32_t L = inverse ? N : n;
  std::vector<PassPlan> plan = plan_passes2(N, L, tn);
  if (inverse) std::reverse(plan.begin(), plan.end());
  uint32_t n_inv = m31_inv((uint32_t)(((uint64_t)1 << N) % P));
  for (size_t k = 0; k < plan.size(); k++) {
    const PassPlan &pp = plan[k];
    FftPass2 p{};
    p.cols = (!inverse && k > 0) ? d_tab_inplace : d_tab;
    p.n_cols = n_cols;
    p.tw_end = tw_end;
    p.N = N;
    p.T = pp.m + pp.w;
    p.lo = pp.lo;
    p.w = pp.w;
    p.m = pp.m;
    // layer tile bits [s_begin - lo + w, s_end - lo + w)
    std::vector<FftStage> st = plan_stages(pp.s_begin - pp.lo + pp.w, pp.s_end - pp.lo + pp.w, p.T);
    if (inverse) std::reverse(st.begin(), st.end());
    if (st.empty()) st.push_back({0, 0, 0});  // pure copy (e.g. evaluating one coefficient)
    STWO_HIP_CHECK_ARG(ctx, st.size() <= kMaxStages, "internal: %zu stages", st.size());
    for (const FftStage &sg : st)
      STWO_HIP_CHECK_ARG(ctx, (sg.r_lo < sg.r_hi && sg.r_hi <= 4) || (sg.r_lo == 0 && sg.r_hi == 0),
                         "internal: unsupported register stage [%u, %u)", sg.r_lo, sg.r_hi);
    p.n_stages = (uint32_t)st.size();
    for (size_t i = 0; i < st.size(); i++) p.st[i] = st[i];
    p.in_log = (!inverse && k == 0) ? n : N;
    p.scale = (inverse && k + 1 == plan.size()) ? n_inv : 0;
    // First RFFT pass of a blowup whose tile rows stay inside the coefficient
    // domain: one workgroup writes all 2^(N-n) output blocks of its tile.
    p.rep_log = (!inverse && k == 0 && N > n && pp.lo + pp.m <= n) ? N - n : 0;
    uint32_t wg_log = std::max<uint32_t>(8, p.T - 4);
    if (tn.wg == 9) wg_log = 9;
    uint64_t total_tiles = (uint64_t)n_cols << (N - p.rep_log - p.T);
    uint32_t tiles_per_wg = 1u << (wg_log + 4 - p.T);
    uint64_t wgs = (total_tiles + tiles_per_wg - 1) / tiles_per_wg;
    STWO_HIP_CHECK_ARG(ctx, wgs * tiles_per_wg <= (1ull << 32), "FFT group of %u columns of log %u too large",
                       n_cols, N);
    p.vec_io = 0;
    if (aligned16 && p.w == 0 && p.st[0].r_hi > 0) {
      if (p.st[0].b0 == 0 && p.in_log >= 4) p.vec_io |= 1;
      if (p.st[p.n_stages - 1].b0 == 0) p.vec_io |= 2;
    }
    p.fixed = fixed_schedule(p, inverse, wg_log);
    out.push_back({p, (uint32_t)wgs, wg_log});
  }
  return STWO_HIP_OK;
}

int launch_fft_pass(stwo_hip_ctx *ctx, bool inverse, const FftPass2 &p, uint32_t wgs, uint32_t wg_log) {
  if (wg_log == 8 && p.fixed == 1)
    return inverse ? launch(ctx, "ifft_pass_wg8", fft_pass_kernel<true, 8, 3, 0, 4, 8>, dim3(wgs), dim3(256), 0, p)
                   : launch(ctx, "rfft_pass_wg8", fft_pass_kernel<false, 8, 3, 8, 4, 0>, dim3(wgs), dim3(256), 0, p);
  if (p.fixed == 3)
    return inverse ? launch(ctx, "ifft_pass_wg8", fft_pass_kernel<true, 8, 2, 6, 7, 1>, dim3(wgs), dim3(256), 0, p)
                   : launch(ctx, "rfft_pass_wg8", fft_pass_kernel<false, 8, 2, 7, 6, 1>, dim3(wgs), dim3(256), 0, p);
  if (p.fixed == 4)
    return inverse ? launch(ctx, "ifft_pass_wg8", fft_pass_kernel<true, 8, 2, 6, 8, 2>, dim3(wgs), dim3(256), 0, p)
                   : launch(ctx, "rfft_pass_wg8", fft_pass_kernel<false, 8, 2, 8, 6, 2>, dim3(wgs), dim3(256), 0, p);
  if (p.fixed == 5)
    return inverse ? launch(ctx, "ifft_pass_wg9", fft_pass_kernel<true, 9, 2, 6, 9, 3>, dim3(wgs), dim3(512), 0, p)
                   : launch(ctx, "rfft_pass_wg9", fft_pass_kernel<false, 9, 2, 9, 6, 3>, dim3(wgs), dim3(512), 0, p);
  if (wg_log == 9 && p.fixed == 2)
    return inverse ? launch(ctx, "ifft_pass_wg9", fft_pass_kernel<true, 9, 2, 5, 9>, dim3(wgs), dim3(512), 0, p)
                   : launch(ctx, "rfft_pass_wg9", fft_pass_kernel<false, 9, 2, 9, 5>, dim3(wgs), dim3(512), 0, p);
#define STWO_SHORT(K, RLO, RHI)                                                                                  \
  if (wg_log == 8 && p.fixed == K)                                                                              \
    return inverse ? launch(ctx, "ifft_pass_wg8", fft_pass_kernel<true, 8, 1, 6, RLO, RHI>, dim3(wgs), dim3(256), 0, p) \
                   : launch(ctx, "rfft_pass_wg8", fft_pass_kernel<false, 8, 1, 6, RLO, RHI>, dim3(wgs), dim3(256), 0, p);
  STWO_SHORT(6, 0, 4) STWO_SHORT(7, 1, 4) STWO_SHORT(8, 2, 4) STWO_SHORT(9, 3, 4)
  STWO_SHORT(10, 0, 1) STWO_SHORT(11, 0, 2) STWO_SHORT(12, 0, 3)
#undef STWO_SHORT
  // Profiling names follow the instantiation: *_wg9 = fft_pass_kernel<INV, 9, ...>.
  // Tiles of 2^10 or more span whole waves (NS = -1: wave-uniform columns).
  const bool uni = p.T >= 10;
  if (wg_log == 9) {
    if (uni)
      return inverse ? launch(ctx, "ifft_pass_wg9", fft_pass_kernel<true, 9, -1>, dim3(wgs), dim3(512), 0, p)
                     : launch(ctx, "rfft_pass_wg9", fft_pass_kernel<false, 9, -1>, dim3(wgs), dim3(512), 0, p);
    return inverse ? launch(ctx, "ifft_pass_wg9", fft_pass_kernel<true, 9>, dim3(wgs), dim3(512), 0, p)
                   : launch(ctx, "rfft_pass_wg9", fft_pass_kernel<false, 9>, dim3(wgs), dim3(512), 0, p);
  }
  if (uni)
    return inverse ? launch(ctx, "ifft_pass_wg8", fft_pass_kernel<true, 8, -1>, dim3(wgs), dim3(256), 0, p)
                   : launch(ctx, "rfft_pass_wg8", fft_pass_kernel<false, 8, -1>, dim3(wgs), dim3(256), 0, p);
  return inverse ? launch(ctx, "ifft_pass_wg8", fft_pass_kernel<true, 8>, dim3(wgs), dim3(256), 0, p)
                 : launch(ctx, "rfft_pass_wg8", fft_pass_kernel<false, 8>, dim3(wgs), dim3(256), 0, p);
}

int launch_fft_multi(stwo_hip_ctx *ctx, bool inverse, const FftPass2 *d_groups, const uint32_t *d_wg_start,
                     uint32_t n_groups, uint32_t total_wgs, uint32_t wg_log, int kind) {
  const dim3 grid(total_wgs), b9(512), b8(256);
  const char *name = inverse ? "ifft_pass_batched" : "rfft_pass_batched";
#define STWO_MULTI(WG, ...)                                                                                    \
  (inverse ? launch(ctx, name, fft_pass_multi_kernel<true, WG, __VA_ARGS__>, grid, WG == 9 ? b9 : b8, 0, d_groups, \
                    d_wg_start, n_groups)                                                                      \
           : launch(ctx, name, fft_pass_multi_kernel<false, WG, __VA_ARGS__>, grid, WG == 9 ? b9 : b8, 0,       \
                    d_groups, d_wg_start, n_groups))
  if (wg_log == 9) {
    if (kind == 2) return inverse ? launch(ctx, name, fft_pass_multi_kernel<true, 9, 2, 5, 9>, grid, b9, 0, d_groups, d_wg_start, n_groups)
                                  : launch(ctx, name, fft_pass_multi_kernel<false, 9, 2, 9, 5>, grid, b9, 0, d_groups, d_wg_start, n_groups);
    if (kind == 5) return inverse ? launch(ctx, name, fft_pass_multi_kernel<true, 9, 2, 6, 9, 3>, grid, b9, 0, d_groups, d_wg_start, n_groups)
                                  : launch(ctx, name, fft_pass_multi_kernel<false, 9, 2, 9, 6, 3>, grid, b9, 0, d_groups, d_wg_start, n_groups);
    if (kind == -1) return STWO_MULTI(9, -1);
    return STWO_MULTI(9, 0);
  }
  if (kind == 1) return inverse ? launch(ctx, name, fft_pass_multi_kernel<true, 8, 3, 0, 4, 8>, grid, b8, 0, d_groups, d_wg_start, n_groups)
                                : launch(ctx, name, fft_pass_multi_kernel<false, 8, 3, 8, 4, 0>, grid, b8, 0, d_groups, d_wg_start, n_groups);
  if (kind == 3) return inverse ? launch(ctx, name, fft_pass_multi_kernel<true, 8, 2, 6, 7, 1>, grid, b8, 0, d_groups, d_wg_start, n_groups)
                                : launch(ctx, name, fft_pass_multi_kernel<false, 8, 2, 7, 6, 1>, grid, b8, 0, d_groups, d_wg_start, n_groups);
  if (kind == 4) return inverse ? launch(ctx, name, fft_pass_multi_kernel<true, 8, 2, 6, 8, 2>, grid, b8, 0, d_groups, d_wg_start, n_groups)
                                : launch(ctx, name, fft_pass_multi_kernel<false, 8, 2, 8, 6, 2>, grid, b8, 0, d_groups, d_wg_start, n_groups);
  switch (kind) {
    case 6: return STWO_MULTI(8, 1, 6, 0, 4);
    case 7: return STWO_MULTI(8, 1, 6, 1, 4);
    case 8: return STWO_MULTI(8, 1, 6, 2, 4);
    case 9: return STWO_MULTI(8, 1, 6, 3, 4);
    case 10: return STWO_MULTI(8, 1, 6, 0, 1);
    case 11: return STWO_MULTI(8, 1, 6, 0, 2);
    case 12: return STWO_MULTI(8, 1, 6, 0, 3);
    default: break;
  }
  if (kind == -1) return STWO_MULTI(8, -1);
  return STWO_MULTI(8, 0);
#undef STWO_MULTI
}

// Radix-16 path for N >= 4: one launch per pass over all columns of the group.
static int run_fft_group2(stwo_hip_ctx *ctx, const std::vector<FftCol> &cols, void *dtab, void *dtab_out,
                          uint32_t n, uint32_t N, const uint32_t *tw_end, bool inverse) {
  bool aligned16 = true;
  for (const FftCol &c : cols)
    if ((((uintptr_t)c.in) | ((uintptr_t)c.out)) & 15) aligned16 = false;
  std::vector<PassLaunch> passes;
  int rc = make_group_passes(ctx, n, N, inverse, (const FftCol *)dtab, (const FftCol *)dtab_out,
                             (uint32_t)cols.size(), tw_end, aligned16, passes);
  if (rc) return rc;
  for (const PassLaunch &pl : passes) {
    rc = launch_fft_pass(ctx, inverse, pl.p, pl.wgs, pl.wg_log);
    if (rc) return rc;
  }
  return STWO_HIP_OK;
}

// interpolate + evaluate of a group of same-size columns (log n, n >= 4)
// with the IFFT's last pass and the RFFT's first pass fused when their tiles
// match; otherwise the two transforms one after the other.
static int run_lde_group(stwo_hip_ctx *ctx, const std::vector<FftCol> &cols, const std::vector<uint32_t *> &outs,
                         uint32_t n, uint32_t b, const stwo_hip_twiddles *tw, bool store_coeffs) {
  const uint32_t N = n + b;
  // tables: IFFT {c, c}; first RFFT pass {c, out}; later RFFT passes {out, out}
  std::vector<FftCol> tf(cols.size()), to(cols.size());
  bool aligned16 = true;
  for (size_t i = 0; i < cols.size(); i++) {
    tf[i] = {cols[i].in, outs[i]};
    to[i] = {outs[i], outs[i]};
    if ((((uintptr_t)cols[i].in) | ((uintptr_t)outs[i])) & 15) aligned16 = false;
  }
  void *d[3];
  {
    const void *srcs[3] = {cols.data(), tf.data(), to.data()};
    const size_t bytes[3] = {cols.size() * sizeof(FftCol), tf.size() * sizeof(FftCol), to.size() * sizeof(FftCol)};
    int rc = upload_tables(ctx, 3, srcs, bytes, d);
    if (rc) return rc;
  }
  std::vector<PassLaunch> inv, fwd;
  int rc = make_group_passes(ctx, n, n, true, (const FftCol *)d[0], (const FftCol *)d[0], (uint32_t)cols.size(),
                             tw->d_itw2 + tw->len, aligned16, inv);
  if (rc) return rc;
  rc = make_group_passes(ctx, n, N, false, (const FftCol *)d[1], (const FftCol *)d[2], (uint32_t)cols.size(),
                         tw->d_tw2 + tw->len, aligned16, fwd);
  if (rc) return rc;
  const FftPass2 &li = inv.back().p, &ff = fwd.front().p;
  const bool fuse = lde_fusable(inv.back(), fwd.front(), b);
  for (size_t k = 0; k + (fuse ? 1 : 0) < inv.size(); k++) {
    rc = launch_fft_pass(ctx, true, inv[k].p, inv[k].wgs, inv[k].wg_log);
    if (rc) return rc;
  }
  if (fuse) {
    const uint32_t wgs = inv.back().wgs;
    const uint32_t sc = store_coeffs ? 1 : 0;
    rc = launch_lde(ctx, lde_kind(inv.back(), fwd.front()), li, ff, wgs, sc);
    if (rc) return rc;
  }
  for (size_t k = fuse ? 1 : 0; k < fwd.size(); k++) {
    rc = launch_fft_pass(ctx, false, fwd[k].p, fwd[k].wgs, fwd[k].wg_log);
    if (rc) return rc;
  }
  return STWO_HIP_OK;
}

// Runs the transform of a group of same-shape columns.  Forward: `in` holds
// 2^n coefficients, `out` receives 2^N evaluations.  Inverse: in == out holds
// 2^N evaluations replaced by coefficients (n == N).
static int run_fft_group(stwo_hip_ctx *ctx, const std::vector<FftCol> &cols, uint32_t n, uint32_t N,
                         const stwo_hip_twiddles *tw, bool inverse, stwo_hip_coset dom) {
  if (cols.empty()) return STWO_HIP_OK;
  STWO_HIP_CHECK_ARG(ctx, cols.size() <= 65535, "too many columns in one group (%zu)", cols.size());
  // The group's table and, forward, the in-place table of the passes after
  // the first (entries {out, out}), uploaded together.
  void *dtab = nullptr, *dtab_out = nullptr;
  int rc;
  {
    std::vector<FftCol> c2(cols);
    for (auto &c : c2) c.in = c.out;
    const void *srcs[2] = {cols.data(), c2.data()};
    const size_t bytes[2] = {cols.size() * sizeof(FftCol), cols.size() * sizeof(FftCol)};
    void *d[2];
    rc = upload_tables(ctx, inverse ? 1 : 2, srcs, bytes, d);
    if (rc) return rc;
    dtab = d[0];
    dtab_out = inverse ? d[0] : d[1];
  }
  if (N <= 2) {
    SmallFft a{};
    a.cols = (const FftCol *)dtab;
    a.N = N;
    a.in_log = n;
    a.inverse = inverse;
    cpoint pt = host_index_to_point(dom.initial_index);
    if (!inverse) {
      a.c0 = pt.x;
      a.c1 = pt.y;
    } else if (N == 1) {
      uint32_t yn_inv = m31_inv(m31_mul(pt.y, 2));
      a.c1 = m31_mul(yn_inv, 2);      // y^-1
      a.c2 = m31_mul(yn_inv, pt.y);   // 1/2
    } else {
      uint32_t xyn_inv = m31_inv(m31_mul(m31_mul(pt.x, pt.y), 4));
      a.c0 = m31_mul(m31_mul(xyn_inv, pt.y), 4);  // x^-1
      a.c1 = m31_mul(m31_mul(xyn_inv, pt.x), 4);  // y^-1
      a.c2 = m31_mul(m31_mul(xyn_inv, pt.x), pt.y);  // 1/4
    }
    return launch(ctx, "fft_small", fft_small_kernel, dim3((uint32_t)cols.size()), dim3(64), 0, a);
  }
  if (N >= 4)
    return run_fft_group2(ctx, cols, dtab, dtab_out, n, N, (inverse ? tw->d_itw2 : tw->d_tw2) + tw->len, inverse);
  const uint32_t *tree = inverse ? tw->d_itw : tw->d_tw;
  const uint32_t *tw_end = tree + tw->len;
  uint32_t L = inverse ? N : n;
  std::vector<PassPlan> plan = plan_passes(N, L);
  if (inverse) std::reverse(plan.begin(), plan.end());
  uint32_t n_inv = m31_inv((uint32_t)(((uint64_t)1 << N) % P));
  for (size_t k = 0; k < plan.size(); k++) {
    const PassPlan &pp = plan[k];
    FftPass p{};
    p.cols = (const FftCol *)((!inverse && k > 0) ? dtab_out : dtab);
    p.tw_end = tw_end;
    p.N = N;
    p.lo = pp.lo;
    p.m = pp.m;
    p.w = pp.w;
    p.s_begin = pp.s_begin;
    p.s_end = pp.s_end;
    p.in_log = (!inverse && k == 0) ? n : N;
    p.inverse = inverse;
    p.scale = (inverse && k + 1 == plan.size()) ? n_inv : 0;
    uint32_t tiles = 1u << (N - pp.m - pp.w);
    rc = launch(ctx, inverse ? "ifft_tile" : "rfft_tile", fft_tile_kernel, dim3(tiles, (uint32_t)cols.size()),
                dim3(kFftThreads), 0, p);
    if (rc) return rc;
  }
  return STWO_HIP_OK;
}

static int bit_reverse_one(stwo_hip_ctx *ctx, uint32_t *col, uint32_t log_size) {
  if (log_size >= 2 * kBrLog) {
    uint32_t mids = 1u << (log_size - 2 * kBrLog);
    return launch(ctx, "bit_reverse_tiled",
                  ((uintptr_t)col & 15) ? bit_reverse_tiled_kernel<false> : bit_reverse_tiled_kernel<true>,
                  dim3(mids), dim3(kBrThreads), 0, col, log_size);
  }
  uint32_t n = 1u << log_size;
  return launch(ctx, "bit_reverse_small", bit_reverse_small_kernel, dim3((n + 255) / 256), dim3(256), 0, col,
                log_size);
}

}  // namespace stwo_hip

using namespace stwo_hip;

extern "C" {

int stwo_hip_bit_reverse_column(stwo_hip_ctx *ctx, uint32_t *col, uint32_t log_size) {
  if (!ctx) return STWO_HIP_EINVAL;
  STWO_HIP_CHECK_ARG(ctx, col != nullptr && log_size <= 31, "bit_reverse_column: bad column");
  return bit_reverse_one(ctx, col, log_size);
}

int stwo_hip_bit_reverse_secure_column(stwo_hip_ctx *ctx, uint32_t *const coords[4], uint32_t log_size) {
  if (!ctx) return STWO_HIP_EINVAL;
  STWO_HIP_CHECK_ARG(ctx, coords != nullptr && log_size <= 31, "bit_reverse_secure_column: bad column");
  for (int k = 0; k < 4; k++) {
    STWO_HIP_CHECK_ARG(ctx, coords[k] != nullptr, "bit_reverse_secure_column: NULL coordinate");
    int rc = bit_reverse_one(ctx, coords[k], log_size);
    if (rc) return rc;
  }
  return STWO_HIP_OK;
}

int stwo_hip_interpolate(stwo_hip_ctx *ctx, uint32_t *col, stwo_hip_coset domain, const stwo_hip_twiddles *tw) {
  if (!ctx) return STWO_HIP_EINVAL;
  STWO_HIP_CHECK_ARG(ctx, col != nullptr, "interpolate: NULL column");
  int rc = check_domain(ctx, domain, tw);
  if (rc) return rc;
  uint32_t N = domain.log_size + 1;
  STWO_HIP_CHECK_ARG(ctx, N <= 28, "interpolate: log size %u too large", N);
  std::vector<FftCol> cols{{col, col}};
  return run_fft_group(ctx, cols, N, N, tw, true, domain);
}

int stwo_hip_interpolate_columns(stwo_hip_ctx *ctx, uint32_t *const *cols, const uint32_t *log_sizes, size_t n_cols,
                                 const stwo_hip_twiddles *tw) {
  if (!ctx) return STWO_HIP_EINVAL;
  STWO_HIP_CHECK_ARG(ctx, n_cols == 0 || (cols && log_sizes), "interpolate_columns: NULL arrays");
  std::map<uint32_t, std::vector<FftCol>> groups;
  for (size_t i = 0; i < n_cols; i++) {
    STWO_HIP_CHECK_ARG(ctx, cols[i] != nullptr, "interpolate_columns: NULL column %zu", i);
    STWO_HIP_CHECK_ARG(ctx, log_sizes[i] >= 1 && log_sizes[i] <= 28, "interpolate_columns: bad log size %u",
                       log_sizes[i]);
    groups[log_sizes[i]].push_back({cols[i], cols[i]});
  }
  for (auto &g : groups) {
    uint32_t N = g.first;
    stwo_hip_coset dom = {half_odds_initial(N - 1), N - 1};
    int rc = check_domain(ctx, dom, tw);
    if (rc) return rc;
    for (size_t off = 0; off < g.second.size(); off += 65535) {
      std::vector<FftCol> part(g.second.begin() + off,
                               g.second.begin() + std::min(g.second.size(), off + (size_t)65535));
      rc = run_fft_group(ctx, part, N, N, tw, true, dom);
      if (rc) return rc;
    }
  }
  return STWO_HIP_OK;
}

int stwo_hip_evaluate(stwo_hip_ctx *ctx, const uint32_t *coeffs, uint32_t coeffs_log_size, stwo_hip_coset domain,
                      const stwo_hip_twiddles *tw, uint32_t *out) {
  if (!ctx) return STWO_HIP_EINVAL;
  STWO_HIP_CHECK_ARG(ctx, coeffs != nullptr && out != nullptr, "evaluate: NULL buffer");
  int rc = check_domain(ctx, domain, tw);
  if (rc) return rc;
  uint32_t N = domain.log_size + 1;
  STWO_HIP_CHECK_ARG(ctx, coeffs_log_size <= N, "evaluate: can only evaluate on larger domains (%u > %u)",
                     coeffs_log_size, N);
  STWO_HIP_CHECK_ARG(ctx, N <= 28, "evaluate: log size %u too large", N);
  std::vector<FftCol> cols{{coeffs, out}};
  return run_fft_group(ctx, cols, coeffs_log_size, N, tw, false, domain);
}

int stwo_hip_evaluate_polynomials(stwo_hip_ctx *ctx, const uint32_t *const *coeffs, const uint32_t *log_sizes,
                                  size_t n_cols, uint32_t log_blowup, const stwo_hip_twiddles *tw,
                                  uint32_t *const *outs) {
  if (!ctx) return STWO_HIP_EINVAL;
  STWO_HIP_CHECK_ARG(ctx, n_cols == 0 || (coeffs && log_sizes && outs), "evaluate_polynomials: NULL arrays");
  std::map<uint32_t, std::vector<FftCol>> groups;
  for (size_t i = 0; i < n_cols; i++) {
    STWO_HIP_CHECK_ARG(ctx, coeffs[i] && outs[i], "evaluate_polynomials: NULL column %zu", i);
    STWO_HIP_CHECK_ARG(ctx, log_sizes[i] <= 28 && log_blowup <= 28 && log_sizes[i] + log_blowup >= 1 &&
                                log_sizes[i] + log_blowup <= 28,
                       "evaluate_polynomials: bad log size %u + %u", log_sizes[i], log_blowup);
    groups[log_sizes[i]].push_back({coeffs[i], outs[i]});
  }
  for (auto &g : groups) {
    uint32_t n = g.first, N = n + log_blowup;
    stwo_hip_coset dom = {half_odds_initial(N - 1), N - 1};
    int rc = check_domain(ctx, dom, tw);
    if (rc) return rc;
    for (size_t off = 0; off < g.second.size(); off += 65535) {
      std::vector<FftCol> part(g.second.begin() + off,
                               g.second.begin() + std::min(g.second.size(), off + (size_t)65535));
      rc = run_fft_group(ctx, part, n, N, tw, false, dom);
      if (rc) return rc;
    }
  }
  return STWO_HIP_OK;
}

int stwo_hip_extend_evals(stwo_hip_ctx *ctx, uint32_t *const *cols, const uint32_t *log_sizes, size_t n_cols,
                          uint32_t log_blowup, const stwo_hip_twiddles *tw, uint32_t *const *outs, int store_coeffs) {
  if (!ctx) return STWO_HIP_EINVAL;
  STWO_HIP_CHECK_ARG(ctx, n_cols == 0 || (cols && log_sizes && outs), "extend_evals: NULL arrays");
  std::map<uint32_t, std::vector<size_t>> groups;
  for (size_t i = 0; i < n_cols; i++) {
    STWO_HIP_CHECK_ARG(ctx, cols[i] && outs[i], "extend_evals: NULL column %zu", i);
    STWO_HIP_CHECK_ARG(ctx, log_sizes[i] >= 1 && log_sizes[i] <= 28 && log_blowup <= 28 &&
                                log_sizes[i] + log_blowup <= 28,
                       "extend_evals: bad log size %u + %u", log_sizes[i], log_blowup);
    groups[log_sizes[i]].push_back(i);
  }
  for (auto &g : groups) {
    const uint32_t n = g.first, N = n + log_blowup;
    const stwo_hip_coset dom_i = {half_odds_initial(n - 1), n - 1}, dom_f = {half_odds_initial(N - 1), N - 1};
    int rc = check_domain(ctx, dom_i, tw);
    if (rc) return rc;
    rc = check_domain(ctx, dom_f, tw);
    if (rc) return rc;
    for (size_t off = 0; off < g.second.size(); off += 65535) {
      const size_t e = std::min(g.second.size(), off + (size_t)65535);
      std::vector<FftCol> part;
      std::vector<uint32_t *> po;
      for (size_t k = off; k < e; k++) {
        part.push_back({cols[g.second[k]], cols[g.second[k]]});
        po.push_back(outs[g.second[k]]);
      }
      if (n >= 4) {
        rc = run_lde_group(ctx, part, po, n, log_blowup, tw, store_coeffs != 0);
      } else {
        rc = run_fft_group(ctx, part, n, n, tw, true, dom_i);
        if (rc) return rc;
        std::vector<FftCol> pf;
        for (size_t k = 0; k < part.size(); k++) pf.push_back({part[k].in, po[k]});
        rc = run_fft_group(ctx, pf, n, N, tw, false, dom_f);
      }
      if (rc) return rc;
    }
  }
  return STWO_HIP_OK;
}

// ---- autotuning ----------------------------------------------------------------
//
// For each FFT shape of a column set (interpolate of 2^n, evaluate of 2^n at
// the blowup), every distinct plan among the candidate low-pass tiles
// (2^10..2^13), strided-pass lengths (5..8 layers) and workgroup sizes (the
// tile's own, or 512 lanes) is run on synthetic columns of that size on the
// device -- as many columns as the set has of that size, up to 2^26 words --
// and timed with HIP events (the fastest of 3 runs after a warm-up); the
// fastest plan becomes the context's choice for the shape.
}  // extern "C"

namespace stwo_hip {
namespace {

struct PlanKey {
  std::vector<uint32_t> v;
  bool operator<(const PlanKey &o) const { return v < o.v; }
};
PlanKey plan_key(const std::vector<PassLaunch> &ps) {
  PlanKey k;
  for (const PassLaunch &pl : ps) {
    const FftPass2 &p = pl.p;
    k.v.insert(k.v.end(), {p.T, p.lo, p.w, p.m, pl.wg_log, p.n_stages, p.fixed});
    for (uint32_t i = 0; i < p.n_stages; i++) k.v.insert(k.v.end(), {p.st[i].b0, p.st[i].r_lo, p.st[i].r_hi});
  }
  return k;
}

int time_passes(stwo_hip_ctx *ctx, bool inverse, const std::vector<PassLaunch> &ps, float *best_ms) {
  hipEvent_t e0 = take_event(ctx), e1 = take_event(ctx);
  if (!e0 || !e1) return set_error(ctx, STWO_HIP_EDEVICE, "autotune: no events");
  int rc = STWO_HIP_OK;
  float best = 1e30f;
  const bool prof = ctx->prof;
  ctx->prof = false;
  for (int it = 0; it < 4 && rc == STWO_HIP_OK; it++) {
    if (hipEventRecord(e0, ctx->stream) != hipSuccess) rc = STWO_HIP_EDEVICE;
    for (const PassLaunch &pl : ps)
      if (rc == STWO_HIP_OK) rc = launch_fft_pass(ctx, inverse, pl.p, pl.wgs, pl.wg_log);
    if (rc == STWO_HIP_OK && hipEventRecord(e1, ctx->stream) != hipSuccess) rc = STWO_HIP_EDEVICE;
    if (rc == STWO_HIP_OK && hipEventSynchronize(e1) != hipSuccess) rc = STWO_HIP_EDEVICE;
    float ms = 0;
    if (rc == STWO_HIP_OK && hipEventElapsedTime(&ms, e0, e1) != hipSuccess) rc = STWO_HIP_EDEVICE;
    if (rc == STWO_HIP_OK && it > 0) best = std::min(best, ms);
  }
  ctx->prof = prof;
  ctx->event_pool.push_back(e0);
  ctx->event_pool.push_back(e1);
  if (rc) return set_error(ctx, rc, "autotune: timing failed");
  *best_ms = best;
  return STWO_HIP_OK;
}

}  // namespace
}  // namespace stwo_hip

extern "C" {

int stwo_hip_fft_autotune(stwo_hip_ctx *ctx, const uint32_t *log_sizes, size_t n_cols, uint32_t log_blowup,
                          const stwo_hip_twiddles *tw, uint32_t *report, size_t max_report, size_t *n_report) {
  if (!ctx) return STWO_HIP_EINVAL;
  STWO_HIP_CHECK_ARG(ctx, (log_sizes || !n_cols) && tw && tw->device == ctx->device, "fft_autotune: bad arguments");
  std::map<uint32_t, size_t> count;
  for (size_t i = 0; i < n_cols; i++) {
    STWO_HIP_CHECK_ARG(ctx, log_sizes[i] <= 28 && log_sizes[i] + log_blowup <= 28, "fft_autotune: bad log size %u",
                       log_sizes[i]);
    count[log_sizes[i]]++;
  }
  size_t nr = 0;
  for (auto &kv : count) {
    const uint32_t n = kv.first, N = n + log_blowup;
    if (n < 4) continue;
    for (int dir = 0; dir < 2; dir++) {
      const bool inverse = dir == 0;
      const uint32_t Nd = inverse ? n : N;
      int rc = check_domain(ctx, {half_odds_initial(Nd - 1), Nd - 1}, tw);
      if (rc) return rc;
      const size_t c = std::max<size_t>(1, std::min(kv.second, ((size_t)1 << 26) >> Nd));
      void *mem = nullptr;
      rc = get_scratch(ctx, "fft_autotune", (c << n) * 4 + (inverse ? 0 : (c << Nd) * 4), &mem);
      if (rc) return rc;
      uint32_t *x = (uint32_t *)mem, *y = x + (c << n);
      rc = stwo_hip_fill_random(ctx, x, c << n, 0x7u, n);
      if (rc) return rc;
      std::vector<FftCol> ta(c), tb(c), tc(c);
      for (size_t i = 0; i < c; i++) {
        ta[i] = {x + (i << n), x + (i << n)};
        tb[i] = {x + (i << n), y + (i << Nd)};
        tc[i] = {y + (i << Nd), y + (i << Nd)};
      }
      void *d[3];
      const void *srcs[3] = {ta.data(), tb.data(), tc.data()};
      const size_t bytes[3] = {c * sizeof(FftCol), c * sizeof(FftCol), c * sizeof(FftCol)};
      rc = upload_tables(ctx, 3, srcs, bytes, d);
      if (rc) return rc;
      const FftCol *first = (const FftCol *)(inverse ? d[0] : d[1]);
      const FftCol *inpl = (const FftCol *)(inverse ? d[0] : d[2]);
      const uint32_t *tw_end = (inverse ? tw->d_itw2 : tw->d_tw2) + tw->len;
      std::map<PlanKey, FftTune> cands;
      FftTune def;
      std::vector<PassLaunch> ps;
      rc = make_group_passes(ctx, n, Nd, inverse, first, inpl, (uint32_t)c, tw_end, true, ps, &def);
      if (rc) return rc;
      const PlanKey def_key = plan_key(ps);
      cands[def_key] = def;
      for (uint32_t t0 = 10; t0 <= 13; t0++)
        for (uint32_t sm = 5; sm <= 8; sm++)
          for (uint32_t wg : {0u, 9u}) {
            FftTune t;
            t.t0 = (uint8_t)t0;
            t.smax = (uint8_t)sm;
            t.wg = (uint8_t)wg;
            rc = make_group_passes(ctx, n, Nd, inverse, first, inpl, (uint32_t)c, tw_end, true, ps, &t);
            if (rc) return rc;
            PlanKey k = plan_key(ps);
            if (!cands.count(k)) cands[k] = t;
          }
      float best = 1e30f, def_ms = 0;
      FftTune best_t = def;
      for (auto &ck : cands) {
        rc = make_group_passes(ctx, n, Nd, inverse, first, inpl, (uint32_t)c, tw_end, true, ps, &ck.second);
        if (rc) return rc;
        float ms = 0;
        rc = time_passes(ctx, inverse, ps, &ms);
        if (rc) return rc;
        if (!(ck.first < def_key) && !(def_key < ck.first)) def_ms = ms;
        if (ms < best) {
          best = ms;
          best_t = ck.second;
        }
      }
      // Another plan must beat the default by kTuneMargin: timed alone a shape
      // gains a few percent from a run-time schedule, but inside a batched plan
      // the default schedules share compile-time kernels with the other sizes
      // of their level and run as fewer launches.
      constexpr float kTuneMargin = 0.05f;
      if (def_ms > 0 && best > def_ms * (1.0f - kTuneMargin)) {
        best = def_ms;
        best_t = def;
      }
      ctx->fft_tune[fft_tune_key(n, Nd, inverse)] = best_t;
      if (report && nr < max_report) {
        uint32_t *o = report + 8 * nr;
        o[0] = n; o[1] = Nd; o[2] = inverse; o[3] = best_t.t0; o[4] = best_t.smax; o[5] = best_t.wg;
        o[6] = (uint32_t)(best * 1e3f + 0.5f);    // microseconds
        o[7] = (uint32_t)(def_ms * 1e3f + 0.5f);  // the default plan's microseconds
      }
      nr++;
    }
  }
  if (n_report) *n_report = nr;
  return STWO_HIP_OK;
}

int stwo_hip_fft_tune_set(stwo_hip_ctx *ctx, const uint32_t *entries, size_t n_entries) {
  if (!ctx) return STWO_HIP_EINVAL;
  STWO_HIP_CHECK_ARG(ctx, entries || !n_entries, "fft_tune_set: NULL entries");
  for (size_t i = 0; i < n_entries; i++) {
    const uint32_t *e = entries + 6 * i;
    STWO_HIP_CHECK_ARG(ctx, e[1] >= 4 && e[1] <= 28 && e[0] <= e[1] && e[2] <= 1 && e[3] >= 10 && e[3] <= 13 &&
                                e[4] >= 5 && e[4] <= 8 && (e[5] == 0 || e[5] == 9),
                       "fft_tune_set: entry %zu invalid", i);
    FftTune t;
    t.t0 = (uint8_t)e[3];
    t.smax = (uint8_t)e[4];
    t.wg = (uint8_t)e[5];
    ctx->fft_tune[fft_tune_key(e[0], e[1], e[2] != 0)] = t;
  }
  return STWO_HIP_OK;
}

int stwo_hip_fft_tune_clear(stwo_hip_ctx *ctx) {
  if (!ctx) return STWO_HIP_EINVAL;
  ctx->fft_tune.clear();
  return STWO_HIP_OK;
}

// Describes the passes the radix-16 path would run for a transform of log
// size N (coefficient log size n), in execution order, 16 words per pass:
// T, lo, w, m, rep_log, wg_log, in_log, scaled, n_stages, 4 packed stages
// (b0 << 16 | r_lo << 8 | r_hi), workgroups for one column, 0, 0.
// Host-only (launches nothing); lets tests check the planner without a device.
int stwo_hip_debug_fft_plan_tuned(uint32_t n, uint32_t N, int inverse, uint32_t t0, uint32_t smax, uint32_t wg,
                                  uint32_t *out, uint32_t max_passes, uint32_t *n_passes) {
  if (!out || !n_passes || N < 4 || N > 28 || n > N) return STWO_HIP_EINVAL;
  std::vector<PassLaunch> passes;
  stwo_hip_ctx dummy;
  FftTune tn;
  if (t0) tn.t0 = (uint8_t)t0;
  if (smax) tn.smax = (uint8_t)smax;
  tn.wg = (uint8_t)wg;
  int rc = make_group_passes(&dummy, inverse ? N : n, N, inverse != 0, nullptr, nullptr, 1, nullptr, true, passes,
                             &tn);
  if (rc) return rc;
  if (passes.size() > max_passes) return STWO_HIP_EINVAL;
  for (size_t k = 0; k < passes.size(); k++) {
    const FftPass2 &p = passes[k].p;
    uint32_t *o = out + 16 * k;
    o[0] = p.T; o[1] = p.lo; o[2] = p.w; o[3] = p.m; o[4] = p.rep_log; o[5] = passes[k].wg_log;
    o[6] = p.in_log; o[7] = p.scale ? 1 : 0; o[8] = p.n_stages;
    for (uint32_t i = 0; i < kMaxStages; i++)
      o[9 + i] = i < p.n_stages ? (p.st[i].b0 << 16) | (p.st[i].r_lo << 8) | p.st[i].r_hi : 0;
    o[13] = passes[k].wgs;
    o[14] = p.vec_io;
    o[15] = p.fixed;
  }
  *n_passes = (uint32_t)passes.size();
  return STWO_HIP_OK;
}

int stwo_hip_debug_fft_plan(uint32_t n, uint32_t N, int inverse, uint32_t *out, uint32_t max_passes,
                            uint32_t *n_passes) {
  return stwo_hip_debug_fft_plan_tuned(n, N, inverse, 0, 0, 0, out, max_passes, n_passes);
}

}  // extern "C"

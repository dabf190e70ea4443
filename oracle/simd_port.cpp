// SimdBackend-equivalent CPU port of the LDE + Merkle commitment step: the
// CPU baseline timed beside the GPU.  TEST AND BASELINE INFRASTRUCTURE ONLY:
// nothing in the library links or calls it; only tests/ and bench.py's
// cpu_baseline leg load it.
//
// What it restates (algorithms, not code) from SimdBackend
// (vendor/stwo-upstream-dev-62b228e/crates/stwo/src/prover/backend/simd):
//   - packed M31: one vector of W lanes (W = 16 with AVX-512, the
//     PackedM31 width; W = 8 with AVX2), canonical values, add/sub as
//     min(c, c -/+ P), multiplication by a doubled twiddle (2t) through
//     32x32 -> 64-bit even/odd lane products, hi + lo/2 (m31.rs:579-627,
//     fft/mod.rs mul_twiddle);
//   - the circle FFT (fft/rfft.rs:35-80, fft/ifft.rs:33-57): the layers
//     below log2(W) inside a vector ("vecwise", with lane permutes), the
//     others as fused 3-layer (radix-8) loops over whole vectors with
//     broadcast twiddles; above 2^16 elements (CACHED_FFT_LOG_SIZE) the
//     layers at and above 16 run as sweeps over the whole column and the
//     rest on 2^16-element chunks that stay in cache (SimdBackend transposes
//     instead; both keep the low layers' working set in cache);
//   - evaluate at a blowup as 2^b independent sub-FFTs of the coefficients,
//     one per output block (circle.rs:413-440);
//   - Blake2s commit_on_layer on W leaves at once, one leaf per lane, the
//     message words loaded straight from the columns (blake2s.rs:60-144,
//     369-417); node layers the same over gathered child hashes;
//   - threads: columns (and output blocks) in parallel for the FFTs, leaf
//     and node chunks in parallel for the tree (rayon's par_iter in the
//     reference, OpenMP here), on all host cores.
//
// Twiddles are the CpuBackend/SimdBackend tree doubled (2t), built on the
// host before the timed region like SimdBackend's precompute_twiddles.
// Built twice by oracle/Makefile: -march=x86-64-v3 (AVX2, W = 8) and
// -march=x86-64-v4 (AVX-512, W = 16).
#include <immintrin.h>
#include <omp.h>
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <vector>

namespace {

constexpr uint32_t P = 0x7fffffffu;
constexpr uint32_t kCachedLog = 16;

#if defined(__AVX512F__)
typedef __m512i V;
constexpr uint32_t LOGW = 4;
static inline V vset1(uint32_t x) { return _mm512_set1_epi32((int)x); }
static inline V vload(const uint32_t *p) { return _mm512_loadu_si512((const void *)p); }
static inline void vstore(uint32_t *p, V v) { _mm512_storeu_si512((void *)p, v); }
static inline V vaddu(V a, V b) { return _mm512_add_epi32(a, b); }
static inline V vsubu(V a, V b) { return _mm512_sub_epi32(a, b); }
static inline V vminu(V a, V b) { return _mm512_min_epu32(a, b); }
static inline V vxor(V a, V b) { return _mm512_xor_si512(a, b); }
static inline V vperm(V v, V idx) { return _mm512_permutexvar_epi32(idx, v); }
static inline V vblend_lanes(V a, V b, uint32_t mask) { return _mm512_mask_blend_epi32((__mmask16)mask, a, b); }
static inline V vror(V x, int n) {
  switch (n) {
    case 16: return _mm512_ror_epi32(x, 16);
    case 12: return _mm512_ror_epi32(x, 12);
    case 8: return _mm512_ror_epi32(x, 8);
    default: return _mm512_ror_epi32(x, 7);
  }
}
// a * t2 (t2 = 2t): hi + (lo >> 1) of the 64-bit product, then one
// conditional subtraction.  Even lanes from the low halves of the 64-bit
// lanes, odd lanes from the high halves shifted down.
static inline V vmul_dbl(V a, V t2) {
  const V pe = _mm512_mul_epu32(a, t2);
  const V po = _mm512_mul_epu32(_mm512_srli_epi64(a, 32), _mm512_srli_epi64(t2, 32));
  const V lo = _mm512_mask_blend_epi32((__mmask16)0xAAAA, pe, _mm512_slli_epi64(po, 32));
  const V hi = _mm512_mask_blend_epi32((__mmask16)0xAAAA, _mm512_srli_epi64(pe, 32), po);
  const V r = _mm512_add_epi32(hi, _mm512_srli_epi32(lo, 1));
  return _mm512_min_epu32(r, _mm512_sub_epi32(r, _mm512_set1_epi32((int)P)));
}
static inline V vgather(const uint32_t *base, V idx) { return _mm512_i32gather_epi32(idx, (const void *)base, 4); }
static inline V vsrl(V v, uint32_t n) { return _mm512_srl_epi32(v, _mm_cvtsi32_si128((int)n)); }
static inline V vsll(V v, uint32_t n) { return _mm512_sll_epi32(v, _mm_cvtsi32_si128((int)n)); }
#else
typedef __m256i V;
constexpr uint32_t LOGW = 3;
static inline V vset1(uint32_t x) { return _mm256_set1_epi32((int)x); }
static inline V vload(const uint32_t *p) { return _mm256_loadu_si256((const __m256i *)p); }
static inline void vstore(uint32_t *p, V v) { _mm256_storeu_si256((__m256i *)p, v); }
static inline V vaddu(V a, V b) { return _mm256_add_epi32(a, b); }
static inline V vsubu(V a, V b) { return _mm256_sub_epi32(a, b); }
static inline V vminu(V a, V b) { return _mm256_min_epu32(a, b); }
static inline V vxor(V a, V b) { return _mm256_xor_si256(a, b); }
static inline V vperm(V v, V idx) { return _mm256_permutevar8x32_epi32(v, idx); }
static inline V vblend_lanes(V a, V b, uint32_t mask) {
  // lane l from b where bit l of mask is set
  const V bits = _mm256_setr_epi32(1, 2, 4, 8, 16, 32, 64, 128);
  const V sel = _mm256_cmpeq_epi32(_mm256_and_si256(_mm256_set1_epi32((int)mask), bits), bits);
  return _mm256_blendv_epi8(a, b, sel);
}
static inline V vror(V x, int n) {
  if (n == 16)
    return _mm256_shuffle_epi8(x, _mm256_setr_epi8(2, 3, 0, 1, 6, 7, 4, 5, 10, 11, 8, 9, 14, 15, 12, 13, 2, 3, 0, 1, 6,
                                                   7, 4, 5, 10, 11, 8, 9, 14, 15, 12, 13));
  if (n == 8)
    return _mm256_shuffle_epi8(x, _mm256_setr_epi8(1, 2, 3, 0, 5, 6, 7, 4, 9, 10, 11, 8, 13, 14, 15, 12, 1, 2, 3, 0, 5,
                                                   6, 7, 4, 9, 10, 11, 8, 13, 14, 15, 12));
  return _mm256_or_si256(_mm256_srli_epi32(x, n), _mm256_slli_epi32(x, 32 - n));
}
static inline V vmul_dbl(V a, V t2) {
  const V pe = _mm256_mul_epu32(a, t2);
  const V po = _mm256_mul_epu32(_mm256_srli_epi64(a, 32), _mm256_srli_epi64(t2, 32));
  const V lo = _mm256_blend_epi32(pe, _mm256_slli_epi64(po, 32), 0xAA);
  const V hi = _mm256_blend_epi32(_mm256_srli_epi64(pe, 32), po, 0xAA);
  const V r = _mm256_add_epi32(hi, _mm256_srli_epi32(lo, 1));
  return _mm256_min_epu32(r, _mm256_sub_epi32(r, _mm256_set1_epi32((int)P)));
}
static inline V vgather(const uint32_t *base, V idx) { return _mm256_i32gather_epi32((const int *)base, idx, 4); }
static inline V vsrl(V v, uint32_t n) { return _mm256_srl_epi32(v, _mm_cvtsi32_si128((int)n)); }
static inline V vsll(V v, uint32_t n) { return _mm256_sll_epi32(v, _mm_cvtsi32_si128((int)n)); }
#endif
constexpr uint32_t W = 1u << LOGW;

static inline V vadd(V a, V b) {
  const V c = vaddu(a, b);
  return vminu(c, vsubu(c, vset1(P)));
}
static inline V vsub(V a, V b) {
  const V c = vsubu(a, b);
  return vminu(c, vaddu(c, vset1(P)));
}
static inline V lane_index() {
  alignas(64) uint32_t l[16];
  for (uint32_t i = 0; i < 16; i++) l[i] = i;
  return vload(l);
}

static inline uint32_t s_mul(uint32_t a, uint32_t b) {
  uint64_t p = (uint64_t)a * b;
  uint32_t s = (uint32_t)(p & P) + (uint32_t)(p >> 31);
  return std::min(s, s - P);
}

// Twiddles of one transform: layer s >= 1 of the domain (log size N) is the
// doubled tree slice starting 2^(N - s) words before its end; the circle
// layer (s = 0) takes [y, -y, -x, x] from the pairs (x, y) of layer 1's
// slice.  `k` = output block (blowup sub-FFT): its pairs' twiddle indices
// are offset by k << (n - 1 - s) within the slice.
struct Tw {
  const uint32_t *end;  // one past the doubled tree
  uint32_t N;           // domain log size
  const uint32_t *layer(uint32_t s) const { return end - ((size_t)1 << (N - s)); }
};

// Forward butterfly on vectors: (a + b t, a - b t).
static inline void bf(V &a, V &b, V t2) {
  const V m = vmul_dbl(b, t2);
  b = vsub(a, m);
  a = vadd(a, m);
}
// Inverse butterfly: (a + b, (a - b) t).
static inline void ibf(V &a, V &b, V t2) {
  const V s = vadd(a, b);
  b = vmul_dbl(vsub(a, b), t2);
  a = s;
}

// Twiddle of pair h of layer s (global pair index within the domain).
static inline uint32_t tw_at(const Tw &tw, uint32_t s, size_t h) {
  if (s) return tw.layer(s)[h];
  const uint32_t *l0 = tw.layer(1);
  const uint32_t x = l0[2 * (h >> 2)], y = l0[2 * (h >> 2) + 1];
  const uint32_t v = (h & 2) ? x : y;
  return ((h & 3) == 1 || (h & 3) == 2) ? 2 * P - v : v;
}


// ---- FFT ----------------------------------------------------------------

struct Consts {
  V lanes;
  V cidx;          // circle layer: per-lane word of the (x, y) window
  uint32_t cneg;   // circle layer: lanes whose twiddle is negated
  V partner[LOGW];
  uint32_t hi_mask[LOGW];  // lanes whose bit s is set (the b element of the pair)
  Consts() {
    alignas(64) uint32_t l[16], c[16];
    cneg = 0;
    for (uint32_t i = 0; i < 16; i++) l[i] = i;
    for (uint32_t i = 0; i < W; i++) {
      const uint32_t hl = i >> 1;
      c[i] = 2 * (hl >> 2) + ((hl & 2) ? 0 : 1);
      if ((hl & 3) == 1 || (hl & 3) == 2) cneg |= 1u << i;
    }
    lanes = vload(l);
    cidx = vload(c);
    for (uint32_t s = 0; s < LOGW; s++) {
      alignas(64) uint32_t p[16];
      hi_mask[s] = 0;
      for (uint32_t i = 0; i < 16; i++) p[i] = i ^ (1u << s);
      for (uint32_t i = 0; i < W; i++)
        if ((i >> s) & 1) hi_mask[s] |= 1u << i;
      partner[s] = vload(p);
    }
  }
};
static const Consts &K() {
  static const Consts k;
  return k;
}

// Per-lane twiddles of vecwise layer s on the vector whose first element has
// global index e0 (a multiple of W): lane l belongs to pair (e0 + l) >> (s + 1).
static inline V vecwise_tw(const Tw &tw, uint32_t s, size_t e0) {
  const Consts &k = K();
  const size_t h0 = e0 >> (s + 1);
  if (s) return vgather(tw.layer(s) + h0, vsrl(k.lanes, s + 1));
  // circle layer: h0 is a multiple of 4; pairs h0 + (l >> 1) take [y, -y, -x, x]
  // from the (x, y) pairs at layer 1's words 2 (h0 >> 2) ...
  const V t = vgather(tw.layer(1) + 2 * (h0 >> 2), k.cidx);
  return vblend_lanes(t, vsubu(vset1(2 * P), t), k.cneg);
}

template <bool INV>
static inline void vecwise_layer(uint32_t *x, size_t n_elems, size_t g0, const Tw &tw, uint32_t s) {
  const Consts &k = K();
  const uint32_t mask = k.hi_mask[s];
  for (size_t i = 0; i < n_elems; i += W) {
    const V v = vload(x + i);
    const V pv = vperm(v, k.partner[s]);
    const V a = vblend_lanes(v, pv, mask), b = vblend_lanes(pv, v, mask);
    const V t = vecwise_tw(tw, s, g0 + i);
    V out;
    if (INV) {
      out = vblend_lanes(vadd(a, b), vmul_dbl(vsub(a, b), t), mask);
    } else {
      const V m = vmul_dbl(b, t);
      out = vblend_lanes(vadd(a, m), vsub(a, m), mask);
    }
    vstore(x + i, out);
  }
}

// One fused group of R layers [t - R + 1, t] (forward: top first; inverse:
// bottom first) on the 2^m elements at x (global index of x[0] = g0, a
// multiple of 2^m); the bottom layer is at least log2(W), so every vector
// holds W consecutive elements of one pair side.  Reads from `src` (may be
// x), writes x; multiplies by `scale` before storing when nonzero.
template <bool INV, int R>
static void group(const uint32_t *src, uint32_t *x, uint32_t m, size_t g0, const Tw &tw, uint32_t t,
                  uint32_t scale) {
  const uint32_t sb = t + 1 - R;
  const size_t stride = (size_t)1 << sb;
  const size_t n_hi = (size_t)1 << (m - 1 - t);
  const V vs = vset1(scale);
  for (size_t hi = 0; hi < n_hi; hi++) {
    const size_t base = hi << (t + 1);
    // twiddles: layer sb + j holds 2^(R - 1 - j) pairs of this group
    V tws[R][4];
    for (int j = 0; j < R; j++) {
      const uint32_t s = sb + j;
      const size_t h0 = (g0 + base) >> (s + 1);
      for (int u = 0; u < (1 << (R - 1 - j)); u++) tws[j][u] = vset1(tw_at(tw, s, h0 + u));
    }
    for (size_t lo = 0; lo < stride; lo += W) {
      V v[1 << R];
      for (int q = 0; q < (1 << R); q++) v[q] = vload(src + base + q * stride + lo);
      for (int jj = 0; jj < R; jj++) {
        const int j = INV ? jj : R - 1 - jj;
        for (int q = 0; q < (1 << R); q++)
          if (!((q >> j) & 1)) {
            const V tv = tws[j][q >> (j + 1)];
            if (INV)
              ibf(v[q], v[q | (1 << j)], tv);
            else
              bf(v[q], v[q | (1 << j)], tv);
          }
      }
      if (scale)
        for (int q = 0; q < (1 << R); q++) v[q] = vmul_dbl(v[q], vs);
      for (int q = 0; q < (1 << R); q++) vstore(x + base + q * stride + lo, v[q]);
    }
  }
}

template <bool INV>
static void run_group(const uint32_t *src, uint32_t *x, uint32_t m, size_t g0, const Tw &tw, uint32_t t, int r,
                      uint32_t scale) {
  if (r == 3) group<INV, 3>(src, x, m, g0, tw, t, scale);
  else if (r == 2) group<INV, 2>(src, x, m, g0, tw, t, scale);
  else group<INV, 1>(src, x, m, g0, tw, t, scale);
}

// Layers [lo, hi) (lo >= log2(W)) as fused groups of up to 3, on all 2^m
// elements at x: forward from the top, inverse from the bottom.  The first
// group reads `src`; the last one scales (scale2 = 2 * scale, as a doubled
// twiddle; 0 = none).
template <bool INV>
static void vertical_layers(const uint32_t *src, uint32_t *x, uint32_t m, size_t g0, const Tw &tw, uint32_t lo,
                            uint32_t hi, uint32_t scale2) {
  if (INV) {
    uint32_t s = lo;
    while (s < hi) {
      const uint32_t r = std::min<uint32_t>(3, hi - s);
      run_group<true>(src, x, m, g0, tw, s + r - 1, (int)r, s + r == hi ? scale2 : 0);
      src = x;
      s += r;
    }
  } else {
    uint32_t s = hi;
    while (s > lo) {
      const uint32_t r = std::min<uint32_t>(3, s - lo);
      run_group<false>(src, x, m, g0, tw, s - 1, (int)r, 0);
      src = x;
      s -= r;
    }
  }
}

// Forward FFT of one block: layers m-1 .. 0 of the 2^m elements at x read
// from src (global index of element 0 = g0).
static void fft_block(const uint32_t *src, uint32_t *x, uint32_t m, size_t g0, const Tw &tw) {
  if (m <= kCachedLog) {
    vertical_layers<false>(src, x, m, g0, tw, LOGW, m, 0);
    for (int s = (int)LOGW - 1; s >= 0; s--) vecwise_layer<false>(x, (size_t)1 << m, g0, tw, (uint32_t)s);
    return;
  }
  vertical_layers<false>(src, x, m, g0, tw, kCachedLog, m, 0);
  for (size_t c = 0; c < ((size_t)1 << (m - kCachedLog)); c++) {
    uint32_t *xc = x + (c << kCachedLog);
    const size_t gc = g0 + (c << kCachedLog);
    vertical_layers<false>(xc, xc, kCachedLog, gc, tw, LOGW, kCachedLog, 0);
    for (int s = (int)LOGW - 1; s >= 0; s--) vecwise_layer<false>(xc, (size_t)1 << kCachedLog, gc, tw, (uint32_t)s);
  }
}

// Inverse FFT of one column in place (layers 0 .. m-1), scaled by 2^-m.
static void ifft_column(uint32_t *x, uint32_t m, const Tw &tw, uint32_t scale2) {
  if (m <= kCachedLog) {
    for (uint32_t s = 0; s < LOGW; s++) vecwise_layer<true>(x, (size_t)1 << m, 0, tw, s);
    vertical_layers<true>(x, x, m, 0, tw, LOGW, m, scale2);
    return;
  }
  for (size_t c = 0; c < ((size_t)1 << (m - kCachedLog)); c++) {
    uint32_t *xc = x + (c << kCachedLog);
    const size_t gc = c << kCachedLog;
    for (uint32_t s = 0; s < LOGW; s++) vecwise_layer<true>(xc, (size_t)1 << kCachedLog, gc, tw, s);
    vertical_layers<true>(xc, xc, kCachedLog, gc, tw, LOGW, kCachedLog, 0);
  }
  vertical_layers<true>(x, x, m, 0, tw, kCachedLog, m, scale2);
}

static uint32_t s_pow(uint32_t a, uint64_t e) {
  uint32_t r = 1;
  while (e) {
    if (e & 1) r = s_mul(r, a);
    a = s_mul(a, a);
    e >>= 1;
  }
  return r;
}

// ---- Blake2s --------------------------------------------------------------

static const uint32_t IV[8] = {0x6A09E667u, 0xBB67AE85u, 0x3C6EF372u, 0xA54FF53Au,
                               0x510E527Fu, 0x9B05688Cu, 0x1F83D9ABu, 0x5BE0CD19u};
static const uint8_t SIGMA[10][16] = {
    {0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15}, {14, 10, 4, 8, 9, 15, 13, 6, 1, 12, 0, 2, 11, 7, 5, 3},
    {11, 8, 12, 0, 5, 2, 15, 13, 10, 14, 3, 6, 7, 1, 9, 4}, {7, 9, 3, 1, 13, 12, 11, 14, 2, 6, 5, 10, 4, 0, 15, 8},
    {9, 0, 5, 7, 2, 4, 10, 15, 14, 1, 11, 12, 6, 8, 3, 13}, {2, 12, 6, 10, 0, 11, 8, 3, 4, 13, 7, 5, 15, 14, 1, 9},
    {12, 5, 1, 15, 14, 13, 4, 10, 0, 7, 6, 3, 9, 2, 8, 11}, {13, 11, 7, 14, 12, 1, 3, 9, 5, 0, 15, 4, 8, 6, 2, 10},
    {6, 15, 14, 9, 11, 3, 0, 8, 12, 2, 13, 7, 1, 4, 10, 5}, {10, 2, 8, 4, 7, 6, 1, 5, 15, 11, 9, 14, 3, 12, 13, 0}};

// One compression of W independent states (h[w] = word w of every lane).
static inline void compress_w(V h[8], const V m[16], uint32_t t, bool last) {
  V v[16];
  for (int i = 0; i < 8; i++) {
    v[i] = h[i];
    v[i + 8] = vset1(IV[i]);
  }
  v[12] = vxor(v[12], vset1(t));
  if (last) v[14] = vxor(v[14], vset1(0xFFFFFFFFu));
#define GV(a, b, c, d, x, y)                 \
  v[a] = vaddu(vaddu(v[a], v[b]), (x));      \
  v[d] = vror(vxor(v[d], v[a]), 16);         \
  v[c] = vaddu(v[c], v[d]);                  \
  v[b] = vror(vxor(v[b], v[c]), 12);         \
  v[a] = vaddu(vaddu(v[a], v[b]), (y));      \
  v[d] = vror(vxor(v[d], v[a]), 8);          \
  v[c] = vaddu(v[c], v[d]);                  \
  v[b] = vror(vxor(v[b], v[c]), 7);
  for (int r = 0; r < 10; r++) {
    const uint8_t *s = SIGMA[r];
    GV(0, 4, 8, 12, m[s[0]], m[s[1]]);
    GV(1, 5, 9, 13, m[s[2]], m[s[3]]);
    GV(2, 6, 10, 14, m[s[4]], m[s[5]]);
    GV(3, 7, 11, 15, m[s[6]], m[s[7]]);
    GV(0, 5, 10, 15, m[s[8]], m[s[9]]);
    GV(1, 6, 11, 12, m[s[10]], m[s[11]]);
    GV(2, 7, 8, 13, m[s[12]], m[s[13]]);
    GV(3, 4, 9, 14, m[s[14]], m[s[15]]);
  }
#undef GV
  for (int i = 0; i < 8; i++) h[i] = vxor(h[i], vxor(v[i], v[i + 8]));
}

static inline uint32_t rotr(uint32_t x, int n) { return (x >> n) | (x << (32 - n)); }
static void compress_1(uint32_t h[8], const uint32_t m[16], uint32_t t, bool last) {
  uint32_t v[16];
  for (int i = 0; i < 8; i++) {
    v[i] = h[i];
    v[i + 8] = IV[i];
  }
  v[12] ^= t;
  if (last) v[14] ^= 0xFFFFFFFFu;
#define GS(a, b, c, d, x, y)                               \
  v[a] = v[a] + v[b] + (x); v[d] = rotr(v[d] ^ v[a], 16);  \
  v[c] = v[c] + v[d]; v[b] = rotr(v[b] ^ v[c], 12);        \
  v[a] = v[a] + v[b] + (y); v[d] = rotr(v[d] ^ v[a], 8);   \
  v[c] = v[c] + v[d]; v[b] = rotr(v[b] ^ v[c], 7);
  for (int r = 0; r < 10; r++) {
    const uint8_t *s = SIGMA[r];
    GS(0, 4, 8, 12, m[s[0]], m[s[1]]);
    GS(1, 5, 9, 13, m[s[2]], m[s[3]]);
    GS(2, 6, 10, 14, m[s[4]], m[s[5]]);
    GS(3, 7, 11, 15, m[s[6]], m[s[7]]);
    GS(0, 5, 10, 15, m[s[8]], m[s[9]]);
    GS(1, 6, 11, 12, m[s[10]], m[s[11]]);
    GS(2, 7, 8, 13, m[s[12]], m[s[13]]);
    GS(3, 4, 9, 14, m[s[14]], m[s[15]]);
  }
#undef GS
  for (int i = 0; i < 8; i++) h[i] ^= v[i] ^ v[i + 8];
}

constexpr uint32_t kParam = 0x01010020u;  // digest 32 bytes, no key, fanout 1, depth 1

// Stores lane j's state as the 8-word hash out + 8 j.
static inline void store_states(const V h[8], uint32_t *out) {
  alignas(64) uint32_t tmp[8][16];
  for (int w = 0; w < 8; w++) vstore(tmp[w], h[w]);
  for (uint32_t j = 0; j < W; j++)
    for (int w = 0; w < 8; w++) out[8 * j + w] = tmp[w][j];
}

// Leaf layer of a mixed-degree tree whose columns all have 2^L values:
// leaf i = H(le32 c_0[i] || ... || le32 c_{C-1}[i]) (STWO/core/vcs/blake2_merkle.rs:17-38).
static void leaves(const uint32_t *const *cols, size_t C, uint32_t L, uint32_t *out) {
  const size_t n = (size_t)1 << L;
  const size_t nblk = std::max<size_t>(1, (4 * C + 63) / 64);
  const uint32_t total = (uint32_t)(4 * C);
  const V zero = vset1(0);
#pragma omp parallel for schedule(static)
  for (size_t i = 0; i < n; i += W) {
    V h[8];
    for (int w = 0; w < 8; w++) h[w] = vset1(IV[w] ^ (w == 0 ? kParam : 0));
    for (size_t b = 0; b < nblk; b++) {
      V m[16];
      for (int k = 0; k < 16; k++) {
        const size_t c = 16 * b + k;
        m[k] = c < C ? vload(cols[c] + i) : zero;
      }
      const bool last = b + 1 == nblk;
      compress_w(h, m, last ? total : (uint32_t)(64 * (b + 1)), last);
    }
    store_states(h, out + 8 * i);
  }
}

// Node layer: node i = H(prev[2i] || prev[2i + 1]) over 2^l nodes.
static void nodes(const uint32_t *prev, uint32_t l, uint32_t *out) {
  const size_t n = (size_t)1 << l;
  if (n < W) {
    for (size_t i = 0; i < n; i++) {
      uint32_t h[8];
      for (int w = 0; w < 8; w++) h[w] = IV[w] ^ (w == 0 ? kParam : 0);
      compress_1(h, prev + 16 * i, 64, true);
      memcpy(out + 8 * i, h, 32);
    }
    return;
  }
  const V idx = vsll(K().lanes, 4);
#pragma omp parallel for schedule(static) if (n >= 4096)
  for (size_t i = 0; i < n; i += W) {
    V h[8], m[16];
    for (int w = 0; w < 8; w++) h[w] = vset1(IV[w] ^ (w == 0 ? kParam : 0));
    for (int k = 0; k < 16; k++) m[k] = vgather(prev + 16 * i + k, idx);
    compress_w(h, m, 64, true);
    store_states(h, out + 8 * i);
  }
}

}  // namespace

extern "C" {

int sp_simd_width(void) { return (int)W; }

int sp_num_threads(void) { return omp_get_max_threads(); }

// PolyOps::interpolate_columns: n_cols columns of 2^log_n bit-reversed
// evaluations on CanonicCoset(log_n) replaced in place by their coefficients
// (natural order).  itw2: the doubled inverse twiddle tree (tw_len words) of
// a root whose doublings include the domain's half coset.
int sp_interpolate_columns(uint32_t *const *cols, size_t n_cols, uint32_t log_n, const uint32_t *itw2,
                           size_t tw_len) {
  if (log_n <= LOGW || log_n > 28 || ((size_t)1 << (log_n - 1)) > tw_len) return -1;
  const Tw tw = {itw2 + tw_len, log_n};
  const uint32_t scale = s_pow(s_pow(2, log_n), P - 2);
  const uint32_t scale2 = 2 * scale;
#pragma omp parallel for schedule(dynamic, 1)
  for (size_t c = 0; c < n_cols; c++) ifft_column(cols[c], log_n, tw, scale2);
  return 0;
}

// PolyOps::evaluate_polynomials at log_blowup: column c's 2^log_n
// coefficients evaluated on CanonicCoset(log_n + log_blowup), bit-reversed,
// into outs[c]; each of the 2^log_blowup output blocks is its own sub-FFT.
int sp_evaluate_columns(const uint32_t *const *coeffs, size_t n_cols, uint32_t log_n, uint32_t log_blowup,
                        const uint32_t *tw2, size_t tw_len, uint32_t *const *outs) {
  const uint32_t N = log_n + log_blowup;
  if (log_n <= LOGW || N > 28 || ((size_t)1 << (N - 1)) > tw_len) return -1;
  const Tw tw = {tw2 + tw_len, N};
  const size_t n_blocks = (size_t)1 << log_blowup;
#pragma omp parallel for schedule(dynamic, 1)
  for (size_t task = 0; task < n_cols * n_blocks; task++) {
    const size_t c = task / n_blocks, k = task % n_blocks;
    fft_block(coeffs[c], outs[c] + (k << log_n), log_n, k << log_n, tw);
  }
  return 0;
}

// MerkleProver::commit of n_cols columns of 2^log_size values (one size):
// the packed tree, layer l at word (2^l - 1) * 8, root first.
int sp_merkle_commit(const uint32_t *const *cols, size_t n_cols, uint32_t log_size, uint32_t *layers) {
  if (log_size < LOGW || log_size > 28 || n_cols == 0) return -1;
  leaves(cols, n_cols, log_size, layers + ((((size_t)1 << log_size) - 1) * 8));
  for (int l = (int)log_size - 1; l >= 0; l--)
    nodes(layers + ((((size_t)2 << l) - 1) * 8), (uint32_t)l, layers + ((((size_t)1 << l) - 1) * 8));
  return 0;
}

}  // extern "C"

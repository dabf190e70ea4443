// M31 / CM31 / QM31 arithmetic for host and gfx950 device code.
//
// Field definitions follow STWO/core/fields/{m31,cm31,qm31}.rs (P = 2^31 - 1,
// CM31 = M31[i]/(i^2 + 1), QM31 = CM31[u]/(u^2 - (2 + i)), limb order
// [a.re, a.im, b.re, b.im]).  Every function takes and returns canonical
// values in [0, P).
//
// Device cost (VALU slots, wave64): add/sub 2 each (v_add/v_sub + v_min_u32);
// mul = one 32x32->64 product (v_mad_u64_u32) + v_alignbit + v_and + v_add +
// v_sub + v_min.  The reduction uses 2^31 = 1 (mod P): for p = a*b < 2^62,
// p = (p >> 31) * 2^31 + (p & P) = (p >> 31) + (p & P) (mod P), and the sum is
// at most 2P - 3, so one conditional subtraction (min(s, s - P), which wraps
// for s < P) canonicalises it.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace stwo_hip {

constexpr uint32_t P = 0x7fffffffu;

__host__ __device__ __forceinline__ uint32_t umin32(uint32_t a, uint32_t b) { return a < b ? a : b; }

__host__ __device__ __forceinline__ uint32_t m31_add(uint32_t a, uint32_t b) {
  uint32_t s = a + b;
  return umin32(s, s - P);
}
__host__ __device__ __forceinline__ uint32_t m31_sub(uint32_t a, uint32_t b) {
  uint32_t s = a - b;
  return umin32(s, s + P);
}
__host__ __device__ __forceinline__ uint32_t m31_neg(uint32_t a) { return m31_sub(0, a); }
__host__ __device__ __forceinline__ uint32_t m31_mul(uint32_t a, uint32_t b) {
  uint64_t p = (uint64_t)a * b;
  uint32_t lo = (uint32_t)p, hi = (uint32_t)(p >> 32);
#ifdef __HIP_DEVICE_COMPILE__
  // (hi:lo) >> 31 in one v_alignbit_b32 (the shift-or form compiles to three).
  uint32_t s = (lo & P) + __builtin_amdgcn_alignbit(hi, lo, 31);
#else
  uint32_t s = (lo & P) + ((hi << 1) | (lo >> 31));
#endif
  return umin32(s, s - P);
}
// a * t for a twiddle given doubled, t2 = 2t (t canonical, so t2 < 2^32), as
// SimdBackend stores its twiddles (SIMD/m31.rs mul_doubled_twiddle): with
// a * t2 = hi * 2^32 + lo, a * t = hi * 2^31 + lo / 2 = hi + (lo >> 1) (mod P),
// hi < P - 2 and lo >> 1 <= P, so one conditional subtraction canonicalises.
// Five VALU slots against m31_mul's six.
__host__ __device__ __forceinline__ uint32_t m31_mul_dbl(uint32_t a, uint32_t t2) {
  uint64_t p = (uint64_t)a * t2;
  uint32_t s = (uint32_t)(p >> 32) + ((uint32_t)p >> 1);
  return umin32(s, s - P);
}
// Lazy accumulation of products of canonical values (each below 2^62):
// m31_fold64 keeps x mod P and brings any 64-bit x below 2^33.6, so three
// more products fit before the next fold; m31_reduce64 gives the canonical
// value of any 64-bit x (two folds leave less than P + 8).
__host__ __device__ __forceinline__ uint64_t m31_fold64(uint64_t x) { return (x & P) + (x >> 31); }
__host__ __device__ __forceinline__ uint32_t m31_reduce64(uint64_t x) {
  x = m31_fold64(m31_fold64(x));
  uint32_t r = (uint32_t)x;
  return umin32(r, r - P);
}
__host__ __device__ inline uint32_t m31_pow(uint32_t a, uint64_t e) {
  uint32_t r = 1;
  while (e) {
    if (e & 1) r = m31_mul(r, a);
    a = m31_mul(a, a);
    e >>= 1;
  }
  return r;
}
// a^(P-2) (Fermat), as M31::inverse (STWO/core/fields/m31.rs:66-69).
__host__ __device__ inline uint32_t m31_inv(uint32_t a) { return m31_pow(a, P - 2); }

struct qm31 {
  uint32_t v[4];
};

__host__ __device__ __forceinline__ qm31 qm31_make(uint32_t a, uint32_t b, uint32_t c, uint32_t d) {
  qm31 r;
  r.v[0] = a; r.v[1] = b; r.v[2] = c; r.v[3] = d;
  return r;
}
__host__ __device__ __forceinline__ qm31 qm31_add(const qm31 &x, const qm31 &y) {
  return qm31_make(m31_add(x.v[0], y.v[0]), m31_add(x.v[1], y.v[1]), m31_add(x.v[2], y.v[2]), m31_add(x.v[3], y.v[3]));
}
__host__ __device__ __forceinline__ qm31 qm31_sub(const qm31 &x, const qm31 &y) {
  return qm31_make(m31_sub(x.v[0], y.v[0]), m31_sub(x.v[1], y.v[1]), m31_sub(x.v[2], y.v[2]), m31_sub(x.v[3], y.v[3]));
}
__host__ __device__ __forceinline__ qm31 qm31_mul_m31(const qm31 &x, uint32_t m) {
  return qm31_make(m31_mul(x.v[0], m), m31_mul(x.v[1], m), m31_mul(x.v[2], m), m31_mul(x.v[3], m));
}
// (a + bi)(c + di) = (ac - bd) + (ad + bc)i (STWO/core/fields/cm31.rs:47-53), with
// the Karatsuba form ad + bc = (a + b)(c + d) - ac - bd: 3 products.
__host__ __device__ __forceinline__ void cm31_mul(uint32_t a, uint32_t b, uint32_t c, uint32_t d, uint32_t &re,
                                                  uint32_t &im) {
  uint32_t ac = m31_mul(a, c), bd = m31_mul(b, d);
  uint32_t s = m31_mul(m31_add(a, b), m31_add(c, d));
  re = m31_sub(ac, bd);
  im = m31_sub(m31_sub(s, ac), bd);
}
// (a + bu)(c + du) = (ac + (2 + i) bd) + (ad + bc)u (STWO/core/fields/qm31.rs:14,81-87),
// ad + bc by Karatsuba over CM31: 3 CM31 products = 9 M31 products.
__host__ __device__ __forceinline__ qm31 qm31_mul(const qm31 &x, const qm31 &y) {
  uint32_t ac0, ac1, bd0, bd1, s0, s1;
  cm31_mul(x.v[0], x.v[1], y.v[0], y.v[1], ac0, ac1);
  cm31_mul(x.v[2], x.v[3], y.v[2], y.v[3], bd0, bd1);
  cm31_mul(m31_add(x.v[0], x.v[2]), m31_add(x.v[1], x.v[3]), m31_add(y.v[0], y.v[2]), m31_add(y.v[1], y.v[3]), s0,
           s1);
  // (2 + i)(bd0 + bd1 i) = (2 bd0 - bd1) + (bd0 + 2 bd1) i
  uint32_t rb0 = m31_sub(m31_add(bd0, bd0), bd1);
  uint32_t rb1 = m31_add(bd0, m31_add(bd1, bd1));
  qm31 r;
  r.v[0] = m31_add(ac0, rb0);
  r.v[1] = m31_add(ac1, rb1);
  r.v[2] = m31_sub(m31_sub(s0, ac0), bd0);
  r.v[3] = m31_sub(m31_sub(s1, ac1), bd1);
  return r;
}

// CM31 inverse: (a + bi)^-1 = (a - bi) / (a^2 + b^2) (STWO/core/fields/cm31.rs).
__host__ __device__ __forceinline__ void cm31_inv(uint32_t a, uint32_t b, uint32_t &ra, uint32_t &rb) {
  const uint32_t n = m31_inv(m31_add(m31_mul(a, a), m31_mul(b, b)));
  ra = m31_mul(a, n);
  rb = m31_mul(m31_neg(b), n);
}

// QM31 inverse: (x + yu)^-1 = (x - yu) / (x^2 - (2 + i) y^2), the denominator
// in CM31 (STWO/core/fields/qm31.rs).
__host__ __device__ __forceinline__ qm31 qm31_inv(const qm31 &v) {
  uint32_t x2r, x2i, y2r, y2i;
  cm31_mul(v.v[0], v.v[1], v.v[0], v.v[1], x2r, x2i);
  cm31_mul(v.v[2], v.v[3], v.v[2], v.v[3], y2r, y2i);
  // (2 + i)(y2r + y2i i) = (2 y2r - y2i) + (y2r + 2 y2i) i
  const uint32_t ry = m31_sub(m31_add(y2r, y2r), y2i), iy = m31_add(y2r, m31_add(y2i, y2i));
  uint32_t dr, di;
  cm31_inv(m31_sub(x2r, ry), m31_sub(x2i, iy), dr, di);
  uint32_t ar, ai, br, bi;
  cm31_mul(v.v[0], v.v[1], dr, di, ar, ai);
  cm31_mul(m31_neg(v.v[2]), m31_neg(v.v[3]), dr, di, br, bi);
  return qm31_make(ar, ai, br, bi);
}

// ---- circle group (STWO/core/circle.rs) ----------------------------------

struct cpoint {
  uint32_t x, y;
};
// (x1 x2 - y1 y2, x1 y2 + y1 x2), STWO/core/circle.rs:123-131.
__host__ __device__ __forceinline__ cpoint cp_add(cpoint p, cpoint q) {
  cpoint r;
  r.x = m31_sub(m31_mul(p.x, q.x), m31_mul(p.y, q.y));
  r.y = m31_add(m31_mul(p.x, q.y), m31_mul(p.y, q.x));
  return r;
}
constexpr uint32_t CIRCLE_LOG_ORDER = 31;
constexpr uint32_t IDX_MASK = (1u << CIRCLE_LOG_ORDER) - 1;

// M31_CIRCLE_GEN * index (STWO/core/circle.rs:202-205,231-233), host side.
inline cpoint host_index_to_point(uint32_t idx) {
  cpoint res = {1, 0}, cur = {2, 1268011823u};
  idx &= IDX_MASK;
  while (idx) {
    if (idx & 1) res = cp_add(res, cur);
    cur = cp_add(cur, cur);
    idx >>= 1;
  }
  return res;
}
// Coset step index 2^(31 - log), STWO/core/circle.rs:295-305.
__host__ __device__ __forceinline__ uint32_t coset_step(uint32_t log_size) {
  return (uint32_t)((1ull << (CIRCLE_LOG_ORDER - log_size)) & IDX_MASK);
}
__host__ __device__ __forceinline__ uint32_t bitrev32(uint32_t i, uint32_t log_size) {
#ifdef __HIP_DEVICE_COMPILE__
  return log_size ? (__brev(i) >> (32 - log_size)) : i;
#else
  uint32_t r = 0;
  for (uint32_t k = 0; k < log_size; k++) { r = (r << 1) | (i & 1); i >>= 1; }
  return log_size ? r : i;
#endif
}
// Coset::half_odds(n).initial_index = 2^(29 - n) (STWO/core/circle.rs:328-330);
// CanonicCoset(n).circle_domain() = CircleDomain(half_odds(n - 1)).
__host__ __device__ __forceinline__ uint32_t half_odds_initial(uint32_t n) { return 1u << (29 - n); }

}  // namespace stwo_hip

namespace stwo_hip {
// Device pointers read from tables are generic (flat) pointers to the
// compiler; these casts state that they point to global memory so loads and
// stores use global_* instructions (SGPR base + 32-bit VGPR offset when the
// base is wave-uniform) instead of flat ones with 64-bit per-lane addresses.
typedef __attribute__((address_space(1))) const uint32_t gconst_u32;
typedef __attribute__((address_space(1))) uint32_t g_u32;
typedef __attribute__((address_space(1))) const uint8_t gconst_u8;
typedef __attribute__((address_space(1))) uint8_t g_u8;
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(1))) const u32x4 gconst_u4;
typedef __attribute__((address_space(1))) u32x4 g_u4;
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));
typedef __attribute__((address_space(1))) const u32x2 gconst_u2;
typedef __attribute__((address_space(1))) u32x2 g_u2;
__device__ __forceinline__ gconst_u32 *as_global(const uint32_t *p) { return (gconst_u32 *)p; }
__device__ __forceinline__ g_u32 *as_global(uint32_t *p) { return (g_u32 *)p; }
// 8- and 16-byte global loads and stores (p aligned to their size).  Flat
// loads would also count against the LDS/scalar-load counter, so every wait
// for a scalar load (a table entry) would wait for them too.
__device__ __forceinline__ uint2 ld_global2(const uint32_t *p) {
  const u32x2 v = *(gconst_u2 *)p;
  return make_uint2(v.x, v.y);
}
__device__ __forceinline__ uint4 ld_global4(const uint32_t *p) {
  const u32x4 v = *(gconst_u4 *)p;
  return make_uint4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ void st_global2(uint32_t *p, uint32_t a, uint32_t b) { *(g_u2 *)p = u32x2{a, b}; }
}  // namespace stwo_hip

namespace stwo_hip {
// Wave-uniform copy of a 64-bit value (e.g. a pointer) into SGPRs.
// __builtin_amdgcn_readfirstlane is a 32-bit operation: each half is read
// separately so the upper half of a device address is preserved.
__device__ __forceinline__ uint64_t readfirstlane_u64(uint64_t v) {
  uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
  uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
  return ((uint64_t)hi << 32) | lo;
}
}  // namespace stwo_hip

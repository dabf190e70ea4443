// FRI quotients: compute_fri_quotients (STWO/prover/pcs/quotient_ops.rs:131-232)
// for the QuotientOps + AccumulationOps path of the reference
// (accumulate_numerators, lift_and_accumulate, compute_quotients_and_combine:
// CPUB/quotients.rs, CPUB/accumulation.rs; Metal MB/quotient.rs:388,
// SH/quotients.metal).
//
// For a row p of CanonicCoset(L) (bit-reversed order) the quotient is
//   sum over sample points z of
//     ( sum over the entries (column i, sample at z) of (c_i * f~_i(p) - b_i)
//       - (sum of a_i) * p.y ) * 1 / ((Re z.x - p.x) Im z.y - (Re z.y - p.y) Im z.x)
// evaluated as (sum of c_i f~_i) - (sum of b_i) - (sum of a_i) p.y, the first
// sum accumulated lazily in 64 bits and the other two precomputed per point;
// with f~_i(p) column i read at its lifted index lift(row, L - log_i), and
// (a_i, b_i, c_i) = alpha^k_i times complex_conjugate_line_coeffs
// (STWO/core/constraints.rs:119-134).  The reference materialises one
// partial-numerator column per (column size, sample point), lifts and adds
// them, then combines; lifting composes (lift(lift(i, r), s) = lift(i, r + s)),
// so reading every column at its lifted index from one kernel gives the same
// field elements with one pass over the columns and no intermediate buffers.
// The host does the per-sample work: randomness powers, the periodicity
// sample (STWO/core/pcs/quotients.rs:289-336), line coefficients, grouping by
// sample point.
#include <string.h>

#include <algorithm>
#include <vector>

#include "common.hpp"
#include "field.hpp"

namespace stwo_hip {

struct QuotEntry {
  const uint32_t *col;
  uint32_t shift;  // L - log size of the column
  uint32_t pad;
  uint32_t c[4];
};
struct QuotPoint {
  uint32_t prx[2], pix[2], pry[2], piy[2];  // Re/Im (CM31) of z.x and z.y
  uint32_t first[4];                        // sum of a_i over the point's entries
  uint32_t bsum[4];                         // sum of b_i over the point's entries
  uint32_t e_begin, e_end;
};
struct GenTable {
  cpoint pw[31];  // M31_CIRCLE_GEN * 2^k
};

__device__ __forceinline__ uint32_t cm_re_mul(uint32_t a, uint32_t b, uint32_t c, uint32_t d) {
  return m31_sub(m31_mul(a, c), m31_mul(b, d));
}

__global__ void __launch_bounds__(256) quotient_kernel(const QuotPoint *pts, uint32_t n_pts, const QuotEntry *ents,
                                                       uint32_t L, GenTable gen, uint32_t *o0, uint32_t *o1,
                                                       uint32_t *o2, uint32_t *o3) {
  const uint32_t row = blockIdx.x * blockDim.x + threadIdx.x;
  if (row >> L) return;
  // CanonicCoset(L).circle_domain().at(bit_reverse(row)): half coset
  // half_odds(L - 1), the second half conjugated (domain.rs:62-68).
  const uint32_t i = bitrev32(row, L);
  const uint32_t half = 1u << (L - 1);
  const uint32_t hstep = coset_step(L - 1);
  uint32_t idx = (half_odds_initial(L - 1) + hstep * (i & (half - 1))) & IDX_MASK;
  if (i >= half) idx = (0u - idx) & IDX_MASK;
  cpoint p = {1, 0};
  for (uint32_t k = 0; k < 31; k++)
    if (idx >> k & 1) p = cp_add(p, gen.pw[k]);
  qm31 q = qm31_make(0, 0, 0, 0);
  for (uint32_t s = 0; s < n_pts; s++) {
    const QuotPoint z = pts[s];
    // (Re z.x - p.x) Im z.y - (Re z.y - p.y) Im z.x, in CM31
    const uint32_t dx0 = m31_sub(z.prx[0], p.x), dx1 = z.prx[1];
    const uint32_t dy0 = m31_sub(z.pry[0], p.y), dy1 = z.pry[1];
    const uint32_t d0 = m31_sub(cm_re_mul(dx0, dx1, z.piy[0], z.piy[1]), cm_re_mul(dy0, dy1, z.pix[0], z.pix[1]));
    const uint32_t d1 = m31_sub(m31_add(m31_mul(dx0, z.piy[1]), m31_mul(dx1, z.piy[0])),
                                m31_add(m31_mul(dy0, z.pix[1]), m31_mul(dy1, z.pix[0])));
    // 1 / (d0 + d1 i) = (d0 - d1 i) / (d0^2 + d1^2)
    const uint32_t ninv = m31_inv(m31_add(m31_mul(d0, d0), m31_mul(d1, d1)));
    const uint32_t i0 = m31_mul(d0, ninv), i1 = m31_mul(m31_neg(d1), ninv);
    // sum (c_i f~_i - b_i) = sum c_i f~_i - sum b_i: the products go into
    // 64-bit accumulators, folded after every third entry.  Six entries'
    // loads are issued together.
    uint64_t acc[4] = {0, 0, 0, 0};
    uint32_t e = z.e_begin;
    for (; e + 6 <= z.e_end; e += 6) {
      uint32_t v[6];
#pragma unroll
      for (int u = 0; u < 6; u++) {
        const QuotEntry &en = ents[e + u];
        v[u] = as_global(en.col)[((row >> (en.shift + 1)) << 1) + (row & 1)];
      }
#pragma unroll
      for (int u = 0; u < 6; u++) {
#pragma unroll
        for (int k = 0; k < 4; k++) acc[k] += (uint64_t)v[u] * ents[e + u].c[k];
        if (u == 2 || u == 5)
#pragma unroll
          for (int k = 0; k < 4; k++) acc[k] = m31_fold64(acc[k]);
      }
    }
    for (; e + 3 <= z.e_end; e += 3) {
#pragma unroll
      for (int u = 0; u < 3; u++) {
        const QuotEntry en = ents[e + u];
        const uint32_t v = as_global(en.col)[((row >> (en.shift + 1)) << 1) + (row & 1)];
#pragma unroll
        for (int k = 0; k < 4; k++) acc[k] += (uint64_t)v * en.c[k];
      }
#pragma unroll
      for (int k = 0; k < 4; k++) acc[k] = m31_fold64(acc[k]);
    }
    for (; e < z.e_end; e++) {
      const QuotEntry en = ents[e];
      const uint32_t v = as_global(en.col)[((row >> (en.shift + 1)) << 1) + (row & 1)];
#pragma unroll
      for (int k = 0; k < 4; k++) acc[k] += (uint64_t)v * en.c[k];
    }
    qm31 num = qm31_make(m31_reduce64(acc[0]), m31_reduce64(acc[1]), m31_reduce64(acc[2]), m31_reduce64(acc[3]));
    num = qm31_sub(num, qm31_make(z.bsum[0], z.bsum[1], z.bsum[2], z.bsum[3]));
    num = qm31_sub(num, qm31_mul_m31(qm31_make(z.first[0], z.first[1], z.first[2], z.first[3]), p.y));
    // QM31 x CM31: each CM31 half times (i0 + i1 i)
    uint32_t r0, r1, r2, r3;
    cm31_mul(num.v[0], num.v[1], i0, i1, r0, r1);
    cm31_mul(num.v[2], num.v[3], i0, i1, r2, r3);
    q = qm31_add(q, qm31_make(r0, r1, r2, r3));
  }
  as_global(o0)[row] = q.v[0];
  as_global(o1)[row] = q.v[1];
  as_global(o2)[row] = q.v[2];
  as_global(o3)[row] = q.v[3];
}

// Every column at the lifting size (no lifting) and 8-byte aligned: a lane
// takes rows 2t and 2t + 1, one 8-byte load per entry.  Their bit-reversed
// indices are i and i + 2^(L-1), so the second point is the conjugate
// (p.x, -p.y) of the first: one point computation, and one M31 inversion per
// sample point for both denominators (1/a = b/(ab), 1/b = a/(ab)).
__global__ void __launch_bounds__(256) quotient_pair_kernel(const QuotPoint *pts, uint32_t n_pts,
                                                            const QuotEntry *ents, uint32_t L, GenTable gen,
                                                            uint32_t *o0, uint32_t *o1, uint32_t *o2, uint32_t *o3) {
  const uint32_t row = 2 * (blockIdx.x * blockDim.x + threadIdx.x);
  if (row >> L) return;
  const uint32_t i = bitrev32(row, L);  // below 2^(L-1): row is even
  const uint32_t half = 1u << (L - 1);
  const uint32_t hstep = coset_step(L - 1);
  const uint32_t idx = (half_odds_initial(L - 1) + hstep * (i & (half - 1))) & IDX_MASK;
  cpoint p = {1, 0};
  for (uint32_t k = 0; k < 31; k++)
    if (idx >> k & 1) p = cp_add(p, gen.pw[k]);
  const uint32_t py[2] = {p.y, m31_neg(p.y)};
  qm31 q[2] = {qm31_make(0, 0, 0, 0), qm31_make(0, 0, 0, 0)};
  for (uint32_t s = 0; s < n_pts; s++) {
    const QuotPoint z = pts[s];
    uint32_t d0[2], d1[2], nrm[2];
#pragma unroll
    for (int rr = 0; rr < 2; rr++) {
      const uint32_t dx0 = m31_sub(z.prx[0], p.x), dx1 = z.prx[1];
      const uint32_t dy0 = m31_sub(z.pry[0], py[rr]), dy1 = z.pry[1];
      d0[rr] = m31_sub(cm_re_mul(dx0, dx1, z.piy[0], z.piy[1]), cm_re_mul(dy0, dy1, z.pix[0], z.pix[1]));
      d1[rr] = m31_sub(m31_add(m31_mul(dx0, z.piy[1]), m31_mul(dx1, z.piy[0])),
                       m31_add(m31_mul(dy0, z.pix[1]), m31_mul(dy1, z.pix[0])));
      nrm[rr] = m31_add(m31_mul(d0[rr], d0[rr]), m31_mul(d1[rr], d1[rr]));
    }
    const uint32_t inv01 = m31_inv(m31_mul(nrm[0], nrm[1]));
    const uint32_t ninv[2] = {m31_mul(inv01, nrm[1]), m31_mul(inv01, nrm[0])};
    uint64_t acc[2][4] = {{0, 0, 0, 0}, {0, 0, 0, 0}};
    uint32_t e = z.e_begin;
    for (; e + 6 <= z.e_end; e += 6) {
      uint2 v[6];
#pragma unroll
      for (int u = 0; u < 6; u++) v[u] = ld_global2(ents[e + u].col + row);
#pragma unroll
      for (int u = 0; u < 6; u++) {
#pragma unroll
        for (int k = 0; k < 4; k++) {
          acc[0][k] += (uint64_t)v[u].x * ents[e + u].c[k];
          acc[1][k] += (uint64_t)v[u].y * ents[e + u].c[k];
        }
        if (u == 2 || u == 5)
#pragma unroll
          for (int k = 0; k < 4; k++) {
            acc[0][k] = m31_fold64(acc[0][k]);
            acc[1][k] = m31_fold64(acc[1][k]);
          }
      }
    }
    for (; e < z.e_end; e++) {
      const uint2 v = ld_global2(ents[e].col + row);
#pragma unroll
      for (int k = 0; k < 4; k++) {
        acc[0][k] = m31_fold64(acc[0][k] + (uint64_t)v.x * ents[e].c[k]);
        acc[1][k] = m31_fold64(acc[1][k] + (uint64_t)v.y * ents[e].c[k]);
      }
    }
#pragma unroll
    for (int rr = 0; rr < 2; rr++) {
      qm31 num =
          qm31_make(m31_reduce64(acc[rr][0]), m31_reduce64(acc[rr][1]), m31_reduce64(acc[rr][2]), m31_reduce64(acc[rr][3]));
      num = qm31_sub(num, qm31_make(z.bsum[0], z.bsum[1], z.bsum[2], z.bsum[3]));
      num = qm31_sub(num, qm31_mul_m31(qm31_make(z.first[0], z.first[1], z.first[2], z.first[3]), py[rr]));
      const uint32_t i0 = m31_mul(d0[rr], ninv[rr]), i1 = m31_mul(m31_neg(d1[rr]), ninv[rr]);
      uint32_t r0, r1, r2, r3;
      cm31_mul(num.v[0], num.v[1], i0, i1, r0, r1);
      cm31_mul(num.v[2], num.v[3], i0, i1, r2, r3);
      q[rr] = qm31_add(q[rr], qm31_make(r0, r1, r2, r3));
    }
  }
  st_global2(o0 + row, q[0].v[0], q[1].v[0]);
  st_global2(o1 + row, q[0].v[1], q[1].v[1]);
  st_global2(o2 + row, q[0].v[2], q[1].v[2]);
  st_global2(o3 + row, q[0].v[3], q[1].v[3]);
}

static qm31 conj(const qm31 &v) { return qm31_make(v.v[0], v.v[1], m31_neg(v.v[2]), m31_neg(v.v[3])); }
static bool eq(const qm31 &a, const qm31 &b) { return memcmp(a.v, b.v, 16) == 0; }
static qm31 from(const uint32_t v[4]) { return qm31_make(v[0], v[1], v[2], v[3]); }

}  // namespace stwo_hip

using namespace stwo_hip;

extern "C" {

int stwo_hip_compute_fri_quotients(stwo_hip_ctx *ctx, const uint32_t *const *cols, const uint32_t *log_sizes,
                                   size_t n_cols, const stwo_hip_point_sample *samples, const uint32_t *n_samples,
                                   const uint32_t random_coeff[4], uint32_t lifting_log_size, uint32_t *const out[4]) {
  if (!ctx) return STWO_HIP_EINVAL;
  const uint32_t L = lifting_log_size;
  STWO_HIP_CHECK_ARG(ctx, out && random_coeff && (cols || !n_cols) && (log_sizes || !n_cols) && (n_samples || !n_cols),
                     "compute_fri_quotients: NULL argument");
  for (int k = 0; k < 4; k++) STWO_HIP_CHECK_ARG(ctx, out[k] != nullptr, "compute_fri_quotients: NULL output");
  STWO_HIP_CHECK_ARG(ctx, L >= 1 && L <= 28, "compute_fri_quotients: lifting log size %u not in 1..=28", L);
  size_t n_total = 0;
  for (size_t c = 0; c < n_cols; c++) {
    STWO_HIP_CHECK_ARG(ctx, cols[c] && log_sizes[c] >= 1 && log_sizes[c] <= L,
                       "compute_fri_quotients: column %zu (log %u) invalid for lifting log %u", c, log_sizes[c], L);
    n_total += n_samples[c];
  }
  STWO_HIP_CHECK_ARG(ctx, samples || !n_total, "compute_fri_quotients: NULL samples");
  // Samples with randomness and periodicity (quotient_ops.rs:80-127).
  struct Ent {
    qm31 px, py, value, rand;
    size_t col;
  };
  std::vector<Ent> all;
  all.reserve(n_total + n_cols);
  const qm31 alpha = from(random_coeff);
  qm31 pw = qm31_make(1, 0, 0, 0);
  size_t si = 0;
  for (size_t c = 0; c < n_cols; c++) {
    const stwo_hip_point_sample *sm = samples + si;
    if (n_samples[c] == 2) {
      // CanonicCoset(L).step().repeated_double(log size): index 2^(31 - L + log)
      const cpoint g = host_index_to_point((uint32_t)((1ull << (31 - L + log_sizes[c])) & IDX_MASK));
      const qm31 x1 = from(sm[1].point.x), y1 = from(sm[1].point.y);
      Ent e;
      e.px = qm31_sub(qm31_mul_m31(x1, g.x), qm31_mul_m31(y1, g.y));
      e.py = qm31_add(qm31_mul_m31(y1, g.x), qm31_mul_m31(x1, g.y));
      e.value = from(sm[1].value);
      e.rand = pw;
      e.col = c;
      all.push_back(e);
      pw = qm31_mul(pw, alpha);
    }
    for (uint32_t j = 0; j < n_samples[c]; j++) {
      Ent e;
      e.px = from(sm[j].point.x);
      e.py = from(sm[j].point.y);
      e.value = from(sm[j].value);
      e.rand = pw;
      e.col = c;
      all.push_back(e);
      pw = qm31_mul(pw, alpha);
    }
    si += n_samples[c];
  }
  // Group by sample point; line coefficients (constraints.rs:119-134): for an
  // entry with value v at point z, a = conj(v) - v, c = conj(z.y) - z.y and
  // b = v c - a z.y, each times its randomness r.  Per point the sums are
  // taken as sum r a = S_c - S_v and sum r b = c S_v - z.y sum r a, with
  // S_v = sum r v and S_c = sum r conj(v) (the same field elements, with
  // three products per entry instead of five).
  std::vector<QuotPoint> pts;
  std::vector<std::vector<QuotEntry>> per;
  std::vector<qm31> sv, sc, ccs;
  std::vector<std::pair<qm31, qm31>> keys;
  for (const Ent &e : all) {
    STWO_HIP_CHECK_ARG(ctx, !eq(e.py, conj(e.py)), "Cannot evaluate a line with a single point.");
    size_t s = 0;
    while (s < keys.size() && !(eq(keys[s].first, e.px) && eq(keys[s].second, e.py))) s++;
    if (s == keys.size()) {
      keys.push_back({e.px, e.py});
      per.emplace_back();
      per.back().reserve(all.size());
      sv.push_back(qm31_make(0, 0, 0, 0));
      sc.push_back(qm31_make(0, 0, 0, 0));
      ccs.push_back(qm31_sub(conj(e.py), e.py));
    }
    sv[s] = qm31_add(sv[s], qm31_mul(e.rand, e.value));
    sc[s] = qm31_add(sc[s], qm31_mul(e.rand, conj(e.value)));
    const qm31 rc = qm31_mul(e.rand, ccs[s]);
    QuotEntry q{};
    q.col = cols[e.col];
    q.shift = L - log_sizes[e.col];
    memcpy(q.c, rc.v, 16);
    per[s].push_back(q);
  }
  std::vector<qm31> firsts(keys.size()), bsums(keys.size());
  for (size_t s = 0; s < keys.size(); s++) {
    firsts[s] = qm31_sub(sc[s], sv[s]);
    bsums[s] = qm31_sub(qm31_mul(ccs[s], sv[s]), qm31_mul(keys[s].second, firsts[s]));
  }
  std::vector<QuotEntry> ents;
  ents.reserve(all.size());
  for (size_t s = 0; s < keys.size(); s++) {
    QuotPoint z{};
    z.prx[0] = keys[s].first.v[0];
    z.prx[1] = keys[s].first.v[1];
    z.pix[0] = keys[s].first.v[2];
    z.pix[1] = keys[s].first.v[3];
    z.pry[0] = keys[s].second.v[0];
    z.pry[1] = keys[s].second.v[1];
    z.piy[0] = keys[s].second.v[2];
    z.piy[1] = keys[s].second.v[3];
    memcpy(z.first, firsts[s].v, 16);
    memcpy(z.bsum, bsums[s].v, 16);
    z.e_begin = (uint32_t)ents.size();
    ents.insert(ents.end(), per[s].begin(), per[s].end());
    z.e_end = (uint32_t)ents.size();
    pts.push_back(z);
  }
  void *d_pts = nullptr, *d_ents = nullptr;
  {
    const void *srcs[2] = {pts.data(), ents.data()};
    const size_t bytes[2] = {pts.size() * sizeof(QuotPoint), ents.size() * sizeof(QuotEntry)};
    void *d[2];
    int rc = upload_tables(ctx, 2, srcs, bytes, d);
    if (rc) return rc;
    d_pts = d[0];
    d_ents = d[1];
  }
  GenTable gen;
  cpoint cur = {2, 1268011823u};
  for (int k = 0; k < 31; k++) {
    gen.pw[k] = cur;
    cur = cp_add(cur, cur);
  }
  // Two rows per lane when no column is lifted and every column and output is
  // 8-byte aligned.
  bool pairs = L >= 1;
  for (const QuotEntry &q : ents) pairs = pairs && q.shift == 0 && ((uintptr_t)q.col & 7) == 0;
  for (int k = 0; k < 4; k++) pairs = pairs && ((uintptr_t)out[k] & 7) == 0;
  const uint64_t n = (1ull << L) >> (pairs ? 1 : 0);
  return launch(ctx, "fri_quotients", pairs ? quotient_pair_kernel : quotient_kernel,
                dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, (const QuotPoint *)d_pts, (uint32_t)pts.size(),
                (const QuotEntry *)d_ents, L, gen, out[0], out[1], out[2], out[3]);
}

}  // extern "C"

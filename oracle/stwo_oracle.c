/*
 * CPU oracle for the stwo Circle-STARK commitment hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file restates, in plain C, the semantics of
 * stwo's CpuBackend/SimdBackend for the commitment path so the HIP library can
 * be checked bit-for-bit.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it, and only as the checker (or the timed CPU
 * baseline), never as a product code path.
 *
 * Every function names the reference file:line it restates.  Paths are
 * relative to the reference checkout; STWO = vendor/stwo-upstream-dev-62b228e/
 * crates/stwo/src, CPUB = STWO/prover/backend/cpu, SIMD = STWO/prover/backend/simd.
 *
 * Parity pins (what independently checks each function) live in tests/:
 * hashlib for Blake2s, brute-force polynomial evaluation for the FFTs and
 * folds, the reference's own known-answer tests, and round-trip invariants.
 *
 * Multithreading: the batch entry points use OpenMP over columns / nodes so
 * that the same code doubles as the multi-core CPU baseline.
 */
#include <stdint.h>
#include <stddef.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define P 0x7fffffffu

/* ------------------------------------------------------------------------ */
/* M31 / CM31 / QM31 (STWO/core/fields/{m31,cm31,qm31}.rs)                   */
/* ------------------------------------------------------------------------ */

/* M31::reduce, STWO/core/fields/m31.rs:58-60 (valid for v < P^2). */
static inline uint32_t m31_reduce(uint64_t v) {
    return (uint32_t)((((((v >> 31) + v + 1) >> 31) + v)) & (uint64_t)P);
}
/* M31::partial_reduce + Add/Sub/Neg, STWO/core/fields/m31.rs:46-48,76-100. */
static inline uint32_t m31_add(uint32_t a, uint32_t b) { uint32_t s = a + b; return s >= P ? s - P : s; }
static inline uint32_t m31_sub(uint32_t a, uint32_t b) { uint32_t s = a + P - b; return s >= P ? s - P : s; }
static inline uint32_t m31_neg(uint32_t a) { uint32_t s = P - a; return s >= P ? s - P : s; }
static inline uint32_t m31_mul(uint32_t a, uint32_t b) { return m31_reduce((uint64_t)a * b); }
static uint32_t m31_pow(uint32_t a, uint64_t e) {
    uint32_t r = 1;
    while (e) { if (e & 1) r = m31_mul(r, a); a = m31_mul(a, a); e >>= 1; }
    return r;
}
/* M31::inverse = a^(P-2), STWO/core/fields/m31.rs:66-69. */
static inline uint32_t m31_inv(uint32_t a) { return m31_pow(a, P - 2); }

typedef struct { uint32_t a, b; } cm31;           /* a + b i, i^2 = -1 */
typedef struct { cm31 a, b; } qm31;                /* a + b u, u^2 = 2 + i */

static inline cm31 cm31_add(cm31 x, cm31 y) { cm31 r = {m31_add(x.a, y.a), m31_add(x.b, y.b)}; return r; }
static inline cm31 cm31_sub(cm31 x, cm31 y) { cm31 r = {m31_sub(x.a, y.a), m31_sub(x.b, y.b)}; return r; }
/* CM31 Mul, STWO/core/fields/cm31.rs:47-53. */
static inline cm31 cm31_mul(cm31 x, cm31 y) {
    cm31 r = {m31_sub(m31_mul(x.a, y.a), m31_mul(x.b, y.b)), m31_add(m31_mul(x.a, y.b), m31_mul(x.b, y.a))};
    return r;
}
static inline qm31 qm31_add(qm31 x, qm31 y) { qm31 r = {cm31_add(x.a, y.a), cm31_add(x.b, y.b)}; return r; }
static inline qm31 qm31_sub(qm31 x, qm31 y) { qm31 r = {cm31_sub(x.a, y.a), cm31_sub(x.b, y.b)}; return r; }
/* QM31 Mul: (a+bu)(c+du) = (ac + R bd) + (ad + bc)u with R = 2+i, STWO/core/fields/qm31.rs:14,81-87. */
static inline qm31 qm31_mul(qm31 x, qm31 y) {
    const cm31 R = {2, 1};
    qm31 r;
    r.a = cm31_add(cm31_mul(x.a, y.a), cm31_mul(R, cm31_mul(x.b, y.b)));
    r.b = cm31_add(cm31_mul(x.a, y.b), cm31_mul(x.b, y.a));
    return r;
}
static inline qm31 qm31_mul_m31(qm31 x, uint32_t m) {
    qm31 r = {{m31_mul(x.a.a, m), m31_mul(x.a.b, m)}, {m31_mul(x.b.a, m), m31_mul(x.b.b, m)}};
    return r;
}
static inline qm31 qm31_from4(const uint32_t v[4]) { qm31 r = {{v[0], v[1]}, {v[2], v[3]}}; return r; }
static inline void qm31_to4(qm31 x, uint32_t v[4]) { v[0] = x.a.a; v[1] = x.a.b; v[2] = x.b.a; v[3] = x.b.b; }
static inline int qm31_is_zero(qm31 x) { return (x.a.a | x.a.b | x.b.a | x.b.b) == 0; }

/* CM31/QM31 inverse, STWO/core/fields/cm31.rs (conjugate / norm) and qm31.rs inverse:
 * (a + bu)^-1 = (a - bu) / (a^2 - (2+i) b^2). */
static cm31 cm31_inv(cm31 x) {
    uint32_t n = m31_add(m31_mul(x.a, x.a), m31_mul(x.b, x.b));
    uint32_t ni = m31_inv(n);
    cm31 r = {m31_mul(x.a, ni), m31_mul(m31_neg(x.b), ni)};
    return r;
}
static qm31 qm31_inv(qm31 x) {
    const cm31 R = {2, 1};
    cm31 b2 = cm31_mul(x.b, x.b);
    cm31 denom = cm31_sub(cm31_mul(x.a, x.a), cm31_mul(R, b2));
    cm31 di = cm31_inv(denom);
    cm31 nb = {m31_neg(x.b.a), m31_neg(x.b.b)};
    qm31 r = {cm31_mul(x.a, di), cm31_mul(nb, di)};
    return r;
}

uint32_t or_m31_mul(uint32_t a, uint32_t b) { return m31_mul(a, b); }
uint32_t or_m31_add(uint32_t a, uint32_t b) { return m31_add(a, b); }
uint32_t or_m31_sub(uint32_t a, uint32_t b) { return m31_sub(a, b); }
uint32_t or_m31_inv(uint32_t a) { return m31_inv(a); }
void or_qm31_mul(const uint32_t a[4], const uint32_t b[4], uint32_t out[4]) { qm31_to4(qm31_mul(qm31_from4(a), qm31_from4(b)), out); }
void or_qm31_inv(const uint32_t a[4], uint32_t out[4]) { qm31_to4(qm31_inv(qm31_from4(a)), out); }

/* ------------------------------------------------------------------------ */
/* Circle group (STWO/core/circle.rs)                                        */
/* ------------------------------------------------------------------------ */

typedef struct { uint32_t x, y; } cpoint;
#define CIRCLE_LOG_ORDER 31
#define IDX_MASK ((1u << CIRCLE_LOG_ORDER) - 1)

/* CirclePoint Add: (x1x2 - y1y2, x1y2 + y1x2), STWO/core/circle.rs:123-131. */
static inline cpoint cp_add(cpoint p, cpoint q) {
    cpoint r = {m31_sub(m31_mul(p.x, q.x), m31_mul(p.y, q.y)), m31_add(m31_mul(p.x, q.y), m31_mul(p.y, q.x))};
    return r;
}
/* M31_CIRCLE_GEN = (2, 1268011823), STWO/core/circle.rs:202-205;
 * CirclePointIndex::to_point = GEN.mul(index), STWO/core/circle.rs:231-233,80-91. */
static cpoint index_to_point(uint32_t idx) {
    cpoint res = {1, 0};
    cpoint cur = {2, 1268011823u};
    idx &= IDX_MASK;
    while (idx) {
        if (idx & 1) res = cp_add(res, cur);
        cur = cp_add(cur, cur);
        idx >>= 1;
    }
    return res;
}
void or_index_to_point(uint32_t idx, uint32_t out[2]) { cpoint p = index_to_point(idx); out[0] = p.x; out[1] = p.y; }

/* Coset::new(initial_index, log_size): step = subgroup_gen(log) = 2^(31-log), STWO/core/circle.rs:295-305. */
static inline uint32_t coset_step(uint32_t log_size) { return (uint32_t)(1u << (CIRCLE_LOG_ORDER - log_size)) & IDX_MASK; }
/* Coset::index_at, STWO/core/circle.rs:382-384. */
static inline uint32_t coset_index_at(uint32_t initial, uint32_t log_size, uint64_t i) {
    return (uint32_t)((initial + (uint64_t)coset_step(log_size) * i) & IDX_MASK);
}
/* CircleDomain::index_at, STWO/core/poly/circle/domain.rs:62-68 (half coset `initial`, `half_log`). */
static inline uint32_t domain_index_at(uint32_t initial, uint32_t half_log, uint64_t i) {
    uint64_t half = (uint64_t)1 << half_log;
    if (i < half) return coset_index_at(initial, half_log, i);
    uint32_t idx = coset_index_at(initial, half_log, i - half);
    return (uint32_t)(((1u << CIRCLE_LOG_ORDER) - idx) & IDX_MASK);
}
void or_coset_at(uint32_t initial, uint32_t log_size, uint64_t i, uint32_t out[2]) {
    cpoint p = index_to_point(coset_index_at(initial, log_size, i)); out[0] = p.x; out[1] = p.y;
}

/* bit_reverse_index, STWO/core/utils.rs:79-84. */
static inline uint64_t bitrev(uint64_t i, uint32_t log_size) {
    if (log_size == 0) return i;
    uint64_t r = 0;
    for (uint32_t k = 0; k < log_size; k++) { r = (r << 1) | (i & 1); i >>= 1; }
    return r;
}
/* bit_reverse, STWO/core/utils.rs:91-101. */
void or_bit_reverse(uint32_t *v, uint32_t log_size) {
    uint64_t n = (uint64_t)1 << log_size;
    for (uint64_t i = 0; i < n; i++) {
        uint64_t j = bitrev(i, log_size);
        if (j > i) { uint32_t t = v[i]; v[i] = v[j]; v[j] = t; }
    }
}

/* ------------------------------------------------------------------------ */
/* Twiddle tree (CPUB/circle.rs:252-311)                                     */
/* ------------------------------------------------------------------------ */

/* slow_precompute_twiddles, CPUB/circle.rs:294-311: for each doubling of the
 * root coset, the x-coordinates of the first half of its points in
 * bit-reversed order; a final pad of 1.  Length = root size.  Coset points
 * are produced by repeated addition of the step as CosetIterator does
 * (STWO/core/circle.rs:346-357, 420-440). */
void or_precompute_twiddles(uint32_t initial, uint32_t log_size, uint32_t *tw, uint32_t *itw) {
    size_t pos = 0;
    uint32_t init = initial, lg = log_size;
    for (uint32_t level = 0; level < log_size; level++) {
        uint64_t half = (uint64_t)1 << (lg - 1);
        cpoint cur = index_to_point(init);
        cpoint step = index_to_point(coset_step(lg));
        for (uint64_t i = 0; i < half; i++) { tw[pos + i] = cur.x; cur = cp_add(cur, step); }
        or_bit_reverse(tw + pos, lg - 1);
        pos += half;
        init = (uint32_t)(((uint64_t)init * 2) & IDX_MASK); /* Coset::double, circle.rs:359-368 */
        lg -= 1;
    }
    tw[pos++] = 1;
    if (itw) for (size_t i = 0; i < pos; i++) itw[i] = m31_inv(tw[i]);
}

/* ------------------------------------------------------------------------ */
/* Circle FFT (CPUB/circle.rs:28-80 interpolate, 199-250 evaluate_into,      */
/* 313-354 fft_layer_loop / circle_twiddles_from_line_twiddles)              */
/* ------------------------------------------------------------------------ */

/* butterfly / ibutterfly, STWO/core/fft.rs:5-21. */
static inline void bfly(uint32_t *v0, uint32_t *v1, uint32_t t) {
    uint32_t tmp = m31_mul(*v1, t); *v1 = m31_sub(*v0, tmp); *v0 = m31_add(*v0, tmp);
}
static inline void ibfly(uint32_t *v0, uint32_t *v1, uint32_t it) {
    uint32_t tmp = *v0; *v0 = m31_add(tmp, *v1); *v1 = m31_mul(m31_sub(tmp, *v1), it);
}

/* Line twiddle layer j of a domain with log size N (j = 0 is the largest
 * layer): domain_line_twiddles_from_tree, STWO/core/poly/utils.rs:98-114. */
static inline const uint32_t *line_tw(const uint32_t *tree, size_t tree_len, uint32_t N, uint32_t j) {
    return tree + tree_len - ((size_t)1 << (N - 1 - j));
}
/* circle_twiddles_from_line_twiddles: [y, -y, -x, x] per (x, y) pair, CPUB/circle.rs:332-354. */
static inline uint32_t circle_tw(const uint32_t *l0, uint64_t h) {
    uint32_t x = l0[2 * (h >> 2)], y = l0[2 * (h >> 2) + 1];
    switch (h & 3) { case 0: return y; case 1: return m31_neg(y); case 2: return m31_neg(x); default: return x; }
}

/* fft_layer_loop over the whole column for line layer j (stride 2^(j+1)). */
static void line_layer(uint32_t *v, uint32_t N, uint32_t j, const uint32_t *t, int inverse) {
    uint64_t stride = (uint64_t)1 << (j + 1);
    uint64_t nh = (uint64_t)1 << (N - 2 - j);
    for (uint64_t h = 0; h < nh; h++) {
        uint32_t tw = t[h];
        uint32_t *a = v + (h << (j + 2)), *b = a + stride;
        if (inverse) for (uint64_t l = 0; l < stride; l++) ibfly(&a[l], &b[l], tw);
        else for (uint64_t l = 0; l < stride; l++) bfly(&a[l], &b[l], tw);
    }
}

/* PolyOps::interpolate (CpuBackend), CPUB/circle.rs:28-80.  `values` is a
 * bit-reversed evaluation on CircleDomain(half coset `initial`, log N-1);
 * replaced in place by the FFT-basis coefficients. */
int or_interpolate(uint32_t *v, uint32_t initial, uint32_t N, const uint32_t *itw, size_t tw_len) {
    if (N == 0) return -1;
    if (N == 1) {
        cpoint p = index_to_point(initial);
        uint32_t n = 2, yn_inv = m31_inv(m31_mul(p.y, n));
        uint32_t y_inv = m31_mul(yn_inv, n), n_inv = m31_mul(yn_inv, p.y);
        ibfly(&v[0], &v[1], y_inv);
        v[0] = m31_mul(v[0], n_inv); v[1] = m31_mul(v[1], n_inv);
        return 0;
    }
    if (N == 2) {
        cpoint p = index_to_point(initial);
        uint32_t n = 4, xyn_inv = m31_inv(m31_mul(m31_mul(p.x, p.y), n));
        uint32_t x_inv = m31_mul(m31_mul(xyn_inv, p.y), n);
        uint32_t y_inv = m31_mul(m31_mul(xyn_inv, p.x), n);
        uint32_t n_inv = m31_mul(m31_mul(xyn_inv, p.x), p.y);
        ibfly(&v[0], &v[1], y_inv); ibfly(&v[2], &v[3], m31_neg(y_inv));
        ibfly(&v[0], &v[2], x_inv); ibfly(&v[1], &v[3], x_inv);
        for (int i = 0; i < 4; i++) v[i] = m31_mul(v[i], n_inv);
        return 0;
    }
    if (((size_t)1 << (N - 1)) > tw_len) return -2; /* "Not enough twiddles!" utils.rs:103-106 */
    uint64_t n = (uint64_t)1 << N;
    const uint32_t *l0 = line_tw(itw, tw_len, N, 0);
    for (uint64_t h = 0; h < n / 2; h++) ibfly(&v[2 * h], &v[2 * h + 1], circle_tw(l0, h));
    for (uint32_t j = 0; j + 1 < N; j++) line_layer(v, N, j, line_tw(itw, tw_len, N, j), 1);
    uint32_t inv = m31_inv((uint32_t)(n % P));
    for (uint64_t i = 0; i < n; i++) v[i] = m31_mul(v[i], inv);
    return 0;
}

/* PolyOps::evaluate_into (CpuBackend), CPUB/circle.rs:199-250, with the
 * zero-padded top layers computed as copies of the coefficients the way
 * SimdBackend's evaluate_into evaluates 2^b independent sub-FFTs
 * (SIMD/circle.rs:413-440); the results are identical because
 * butterfly(v, 0, t) = (v, v). */
int or_evaluate(const uint32_t *coeffs, uint32_t n_log, uint32_t initial, uint32_t N,
                const uint32_t *tw, size_t tw_len, uint32_t *out) {
    if (n_log > N || N == 0) return -1;
    uint64_t n = (uint64_t)1 << N, m = (uint64_t)1 << n_log;
    if (N <= 2) {
        memcpy(out, coeffs, m * 4);
        for (uint64_t i = m; i < n; i++) out[i] = 0;
        cpoint p = index_to_point(initial);
        if (N == 1) { bfly(&out[0], &out[1], p.y); return 0; }
        bfly(&out[0], &out[2], p.x); bfly(&out[1], &out[3], p.x);
        bfly(&out[0], &out[1], p.y); bfly(&out[2], &out[3], m31_neg(p.y));
        return 0;
    }
    if (((size_t)1 << (N - 1)) > tw_len) return -2;
    for (uint64_t b = 0; b < n / m; b++) memcpy(out + b * m, coeffs, m * 4);
    /* Line layers j = N-2 .. 0 (stride 2^(j+1)); layers with stride >= m are the copies above. */
    for (int j = (int)N - 2; j >= 0; j--) {
        if (((uint64_t)1 << (j + 1)) >= m) continue;
        line_layer(out, N, (uint32_t)j, line_tw(tw, tw_len, N, (uint32_t)j), 0);
    }
    if (m >= 2) {
        const uint32_t *l0 = line_tw(tw, tw_len, N, 0);
        for (uint64_t h = 0; h < n / 2; h++) bfly(&out[2 * h], &out[2 * h + 1], circle_tw(l0, h));
    }
    return 0;
}

/* interpolate_columns over canonic domains, STWO/prover/poly/circle/ops.rs:29-40
 * (CanonicCoset(log).circle_domain() = CircleDomain(half_odds(log-1)),
 * STWO/core/poly/circle/canonic.rs:41-48; half_odds(n) initial = 2^(29-n), circle.rs:328-330). */
static inline uint32_t half_odds_initial(uint32_t n) { return (uint32_t)(1u << (29 - n)); }
uint32_t or_half_odds_initial(uint32_t n) { return half_odds_initial(n); }

int or_interpolate_columns(uint32_t **cols, const uint32_t *log_sizes, size_t n_cols,
                           const uint32_t *itw, size_t tw_len) {
    int err = 0;
    #pragma omp parallel for schedule(dynamic, 1) reduction(|:err)
    for (size_t c = 0; c < n_cols; c++)
        err |= or_interpolate(cols[c], half_odds_initial(log_sizes[c] - 1), log_sizes[c], itw, tw_len) != 0;
    return err ? -1 : 0;
}
/* evaluate_polynomials at blowup 2^b onto CanonicCoset(log + b), ops.rs:110-141. */
int or_evaluate_polynomials(uint32_t **coeffs, const uint32_t *log_sizes, size_t n_cols, uint32_t log_blowup,
                            const uint32_t *tw, size_t tw_len, uint32_t **outs) {
    int err = 0;
    #pragma omp parallel for schedule(dynamic, 1) reduction(|:err)
    for (size_t c = 0; c < n_cols; c++) {
        uint32_t N = log_sizes[c] + log_blowup;
        err |= or_evaluate(coeffs[c], log_sizes[c], half_odds_initial(N - 1), N, tw, tw_len, outs[c]) != 0;
    }
    return err ? -1 : 0;
}

/* ------------------------------------------------------------------------ */
/* Blake2s-256 (RFC 7693; the `blake2` 0.10.6 crate used by                   */
/* STWO/core/vcs/blake2_hash.rs:73-110, and SIMD/blake2s.rs:23-50,369-417)    */
/* ------------------------------------------------------------------------ */

static const uint32_t B2_IV[8] = {0x6A09E667u, 0xBB67AE85u, 0x3C6EF372u, 0xA54FF53Au,
                                  0x510E527Fu, 0x9B05688Cu, 0x1F83D9ABu, 0x5BE0CD19u};
static const uint8_t B2_SIGMA[10][16] = {
    {0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15}, {14, 10, 4, 8, 9, 15, 13, 6, 1, 12, 0, 2, 11, 7, 5, 3},
    {11, 8, 12, 0, 5, 2, 15, 13, 10, 14, 3, 6, 7, 1, 9, 4}, {7, 9, 3, 1, 13, 12, 11, 14, 2, 6, 5, 10, 4, 0, 15, 8},
    {9, 0, 5, 7, 2, 4, 10, 15, 14, 1, 11, 12, 6, 8, 3, 13}, {2, 12, 6, 10, 0, 11, 8, 3, 4, 13, 7, 5, 15, 14, 1, 9},
    {12, 5, 1, 15, 14, 13, 4, 10, 0, 7, 6, 3, 9, 2, 8, 11}, {13, 11, 7, 14, 12, 1, 3, 9, 5, 0, 15, 4, 8, 6, 2, 10},
    {6, 15, 14, 9, 11, 3, 0, 8, 12, 2, 13, 7, 1, 4, 10, 5}, {10, 2, 8, 4, 7, 6, 1, 5, 15, 11, 9, 14, 3, 12, 13, 0}};

static inline uint32_t rotr(uint32_t x, int n) { return (x >> n) | (x << (32 - n)); }

static void b2_compress(uint32_t h[8], const uint32_t m[16], uint32_t t, int last) {
    uint32_t v[16];
    for (int i = 0; i < 8; i++) { v[i] = h[i]; v[i + 8] = B2_IV[i]; }
    v[12] ^= t;           /* t >> 32 is 0 for messages < 4 GiB */
    if (last) v[14] ^= 0xFFFFFFFFu;
    for (int r = 0; r < 10; r++) {
        const uint8_t *s = B2_SIGMA[r];
#define G(a, b, c, d, x, y)                                   \
        v[a] = v[a] + v[b] + (x); v[d] = rotr(v[d] ^ v[a], 16); \
        v[c] = v[c] + v[d];       v[b] = rotr(v[b] ^ v[c], 12); \
        v[a] = v[a] + v[b] + (y); v[d] = rotr(v[d] ^ v[a], 8);  \
        v[c] = v[c] + v[d];       v[b] = rotr(v[b] ^ v[c], 7);
        G(0, 4, 8, 12, m[s[0]], m[s[1]]);
        G(1, 5, 9, 13, m[s[2]], m[s[3]]);
        G(2, 6, 10, 14, m[s[4]], m[s[5]]);
        G(3, 7, 11, 15, m[s[6]], m[s[7]]);
        G(0, 5, 10, 15, m[s[8]], m[s[9]]);
        G(1, 6, 11, 12, m[s[10]], m[s[11]]);
        G(2, 7, 8, 13, m[s[12]], m[s[13]]);
        G(3, 4, 9, 14, m[s[14]], m[s[15]]);
#undef G
    }
    for (int i = 0; i < 8; i++) h[i] ^= v[i] ^ v[i + 8];
}

/* Streaming hasher with the crate's buffering rule: a full block is only
 * compressed once more input arrives, so the last block is always the
 * finalized one (and the empty message is one zero block with t = 0). */
typedef struct { uint32_t h[8]; uint8_t buf[64]; uint32_t buflen; uint32_t t; } b2s;
static void b2_init(b2s *s) {
    for (int i = 0; i < 8; i++) s->h[i] = B2_IV[i];
    s->h[0] ^= 0x01010020u; /* digest 32, key 0, fanout 1, depth 1: SIMD/blake2s.rs:29 */
    s->buflen = 0; s->t = 0;
}
static void b2_block(b2s *s, int last) {
    uint32_t m[16];
    memcpy(m, s->buf, 64); /* little-endian host */
    b2_compress(s->h, m, s->t, last);
}
static void b2_update(b2s *s, const uint8_t *d, size_t len) {
    while (len) {
        if (s->buflen == 64) { s->t += 64; b2_block(s, 0); s->buflen = 0; }
        size_t take = 64 - s->buflen; if (take > len) take = len;
        memcpy(s->buf + s->buflen, d, take);
        s->buflen += (uint32_t)take; d += take; len -= take;
    }
}
static void b2_final(b2s *s, uint8_t out[32]) {
    s->t += s->buflen;
    memset(s->buf + s->buflen, 0, 64 - s->buflen);
    b2_block(s, 1);
    memcpy(out, s->h, 32);
}
void or_blake2s(const uint8_t *data, size_t len, uint8_t out[32]) {
    b2s s; b2_init(&s); b2_update(&s, data, len); b2_final(&s, out);
}
/* reduce_to_m31, STWO/core/vcs/blake2_hash.rs:111-119. */
static void reduce_to_m31_hash(uint8_t h[32]) {
    uint32_t w[8]; memcpy(w, h, 32);
    for (int i = 0; i < 8; i++) w[i] = m31_reduce(w[i]);
    memcpy(h, w, 32);
}

/* ------------------------------------------------------------------------ */
/* Mixed-degree Merkle (MerkleOps::commit_on_layer / MerkleProver::commit)    */
/* ------------------------------------------------------------------------ */

/* Blake2sMerkleHasher::hash_node, STWO/core/vcs/blake2_merkle.rs:17-38 applied
 * per node as in CPUB/blake2s.rs:13-26.  Hashes are 8 LE u32 words. */
int or_commit_on_layer(uint32_t log_size, const uint32_t *prev, const uint32_t *const *cols, size_t n_cols,
                       uint32_t *out, int m31_output) {
    int64_t n = (int64_t)1 << log_size;
    #pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) {
        b2s s; b2_init(&s);
        if (prev) b2_update(&s, (const uint8_t *)(prev + 16 * i), 64);
        for (size_t c = 0; c < n_cols; c++) b2_update(&s, (const uint8_t *)&cols[c][i], 4);
        uint8_t h[32]; b2_final(&s, h);
        if (m31_output) reduce_to_m31_hash(h);
        memcpy(out + 8 * i, h, 32);
    }
    return 0;
}

/* MerkleProver::commit, STWO/prover/vcs/prover.rs:39-65: columns stable-sorted
 * by descending length; layer log_size takes the columns of exactly that size.
 * Output: packed layers, layer l (2^l hashes) at word offset (2^l - 1) * 8, so
 * the root is at offset 0 (layers[0] is the root, as in prover.rs:176-178).
 * Returns the max log size, or -1 when `n_cols == 0` (then out[0..8] = H("")). */
int or_merkle_commit(const uint32_t *const *cols, const uint32_t *log_sizes, size_t n_cols, uint32_t *layers,
                     int m31_output) {
    if (n_cols == 0) {
        uint8_t h[32]; or_blake2s(NULL, 0, h);
        if (m31_output) reduce_to_m31_hash(h);
        memcpy(layers, h, 32);
        return -1;
    }
    uint32_t max_log = 0;
    for (size_t c = 0; c < n_cols; c++) if (log_sizes[c] > max_log) max_log = log_sizes[c];
    const uint32_t **sel = (const uint32_t **)malloc(sizeof(*sel) * n_cols);
    for (int l = (int)max_log; l >= 0; l--) {
        size_t k = 0;
        /* stable order within equal lengths == input order */
        for (size_t c = 0; c < n_cols; c++) if (log_sizes[c] == (uint32_t)l) sel[k++] = cols[c];
        const uint32_t *prev = (l == (int)max_log) ? NULL : layers + (((size_t)2 << l) - 1) * 8;
        or_commit_on_layer((uint32_t)l, prev, sel, k, layers + (((size_t)1 << l) - 1) * 8, m31_output);
    }
    free(sel);
    return (int)max_log;
}

/* ------------------------------------------------------------------------ */
/* Lifted Merkle (MerkleOpsLifted, CPUB/merkle_lifted.rs:44-85)               */
/* ------------------------------------------------------------------------ */

/* build_leaves, CPUB/merkle_lifted.rs:44-78.  Columns are taken in ascending
 * stable length order (vcs_lifted/ops.rs:37).  Leaf i hashes, per column c of
 * log size s, the value c[lift(i, L - s)] with lift(i, r) = ((i >> (r+1)) << 1)
 * + (i & 1); the per-size cloning in the reference composes to this index
 * (lift(lift(i, r1), r2) = lift(i, r1 + r2)). */
int or_lifted_build_leaves(const uint32_t *const *cols, const uint32_t *log_sizes, size_t n_cols,
                           uint32_t lifting_log, uint32_t *out, int m31_output) {
    if (n_cols == 0) {
        uint8_t h[32]; or_blake2s(NULL, 0, h);
        if (m31_output) reduce_to_m31_hash(h);
        memcpy(out, h, 32);
        return 0;
    }
    size_t *order = (size_t *)malloc(sizeof(size_t) * n_cols);
    size_t k = 0;
    for (uint32_t s = 0; s <= 31; s++)
        for (size_t c = 0; c < n_cols; c++) if (log_sizes[c] == s) order[k++] = c;
    for (size_t c = 0; c < n_cols; c++) if (log_sizes[c] < 1 || log_sizes[c] > lifting_log) { free(order); return -1; }
    int64_t n = (int64_t)1 << lifting_log;
    #pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) {
        b2s s; b2_init(&s);
        for (size_t q = 0; q < n_cols; q++) {
            size_t c = order[q];
            uint32_t r = lifting_log - log_sizes[c];
            uint64_t idx = (((uint64_t)i >> (r + 1)) << 1) + ((uint64_t)i & 1);
            b2_update(&s, (const uint8_t *)&cols[c][idx], 4);
        }
        uint8_t h[32]; b2_final(&s, h);
        if (m31_output) reduce_to_m31_hash(h);
        memcpy(out + 8 * i, h, 32);
    }
    free(order);
    return 0;
}
/* build_next_layer, CPUB/merkle_lifted.rs:80-85 with hash_children,
 * STWO/core/vcs_lifted/blake2_merkle.rs:18-27.  `out_log` is the output size. */
void or_lifted_build_next_layer(const uint32_t *prev, uint32_t out_log, uint32_t *out, int m31_output) {
    int64_t n = (int64_t)1 << out_log;
    #pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) {
        uint8_t h[32]; or_blake2s((const uint8_t *)(prev + 16 * i), 64, h);
        if (m31_output) reduce_to_m31_hash(h);
        memcpy(out + 8 * i, h, 32);
    }
}
/* build_merkle_layers, STWO/prover/vcs_lifted/ops.rs:29-45, packed like or_merkle_commit. */
int or_lifted_commit(const uint32_t *const *cols, const uint32_t *log_sizes, size_t n_cols, uint32_t lifting_log,
                     uint32_t *layers, int m31_output) {
    if (n_cols == 0) return or_lifted_build_leaves(cols, log_sizes, 0, 0, layers, m31_output);
    int rc = or_lifted_build_leaves(cols, log_sizes, n_cols, lifting_log, layers + (((size_t)1 << lifting_log) - 1) * 8, m31_output);
    if (rc) return rc;
    for (int l = (int)lifting_log - 1; l >= 0; l--)
        or_lifted_build_next_layer(layers + (((size_t)2 << l) - 1) * 8, (uint32_t)l, layers + (((size_t)1 << l) - 1) * 8, m31_output);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* FRI folds (CPUB/fri.rs:15-30,67-123)                                      */
/* ------------------------------------------------------------------------ */

/* Coset points in natural order by repeated addition (CosetIterator). */
static void coset_points(uint32_t initial, uint32_t log_size, cpoint *pts) {
    uint64_t n = (uint64_t)1 << log_size;
    cpoint cur = index_to_point(initial), step = index_to_point(coset_step(log_size));
    for (uint64_t i = 0; i < n; i++) { pts[i] = cur; cur = cp_add(cur, step); }
}

/* fold_line_cpu, CPUB/fri.rs:67-92, repeated fold_step times with alpha <- alpha^2
 * and the domain doubled (CpuBackend::fold_line, CPUB/fri.rs:15-30).  `src` and
 * `dst` are 4 SoA coordinate columns (SecureColumnByCoords). */
int or_fold_line(const uint32_t *const src[4], uint32_t line_initial, uint32_t log_n, const uint32_t alpha[4],
                 uint32_t fold_step, uint32_t *const dst[4]) {
    if (fold_step < 1 || fold_step > log_n) return -1;
    uint64_t n = (uint64_t)1 << log_n;
    uint32_t *cur[4], *nxt[4];
    for (int k = 0; k < 4; k++) { cur[k] = (uint32_t *)malloc(n * 4); memcpy(cur[k], src[k], n * 4); nxt[k] = (uint32_t *)malloc(n * 4); }
    qm31 a = qm31_from4(alpha);
    uint32_t init = line_initial, lg = log_n;
    cpoint *pts = (cpoint *)malloc(sizeof(cpoint) * (n / 2 > 0 ? n / 2 : 1));
    for (uint32_t step = 0; step < fold_step; step++) {
        uint64_t half = (uint64_t)1 << (lg - 1);
        /* x = domain.at(bitrev(2i, lg)) = first-half coset point bitrev(i, lg-1) */
        {
            uint64_t nn = half; cpoint c = index_to_point(init), st = index_to_point(coset_step(lg));
            for (uint64_t i = 0; i < nn; i++) { pts[i] = c; c = cp_add(c, st); }
        }
        #pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < (int64_t)half; i++) {
            uint32_t xinv = m31_inv(pts[bitrev((uint64_t)i, lg - 1)].x);
            qm31 f0, f1;
            uint32_t v0[4], v1[4];
            for (int k = 0; k < 4; k++) { v0[k] = cur[k][2 * i]; v1[k] = cur[k][2 * i + 1]; ibfly(&v0[k], &v1[k], xinv); }
            f0 = qm31_from4(v0); f1 = qm31_from4(v1);
            qm31 r = qm31_add(f0, qm31_mul(a, f1));
            uint32_t o[4]; qm31_to4(r, o);
            for (int k = 0; k < 4; k++) nxt[k][i] = o[k];
        }
        for (int k = 0; k < 4; k++) { uint32_t *t = cur[k]; cur[k] = nxt[k]; nxt[k] = t; }
        a = qm31_mul(a, a);
        init = (uint32_t)(((uint64_t)init * 2) & IDX_MASK);
        lg -= 1;
    }
    uint64_t out_n = (uint64_t)1 << lg;
    for (int k = 0; k < 4; k++) { memcpy(dst[k], cur[k], out_n * 4); free(cur[k]); free(nxt[k]); }
    free(pts);
    return 0;
}

/* fold_circle_into_line_cpu, CPUB/fri.rs:95-123: for i < N/2, p =
 * domain.at(bitrev(2i, log N)), (f0, f1) = ibutterfly(f[2i], f[2i+1], 1/p.y),
 * dst[i] = dst[i] * alpha^2 + (alpha * f1 + f0). */
int or_fold_circle_into_line(uint32_t *const dst[4], const uint32_t *const src[4], uint32_t domain_initial,
                             uint32_t log_N, const uint32_t alpha[4]) {
    if (log_N < 1) return -1;
    uint32_t half_log = log_N - 1;
    uint64_t half = (uint64_t)1 << half_log;
    qm31 a = qm31_from4(alpha), a2 = qm31_mul(a, a);
    cpoint *pts = (cpoint *)malloc(sizeof(cpoint) * half);
    coset_points(domain_initial, half_log, pts);
    #pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)half; i++) {
        /* bitrev(2i, log N) < N/2, so the point lies in the half coset. */
        uint64_t j = bitrev(2 * (uint64_t)i, log_N);
        uint32_t yinv = m31_inv(pts[j].y);
        uint32_t v0[4], v1[4], d[4];
        for (int k = 0; k < 4; k++) { v0[k] = src[k][2 * i]; v1[k] = src[k][2 * i + 1]; ibfly(&v0[k], &v1[k], yinv); d[k] = dst[k][i]; }
        qm31 fp = qm31_add(qm31_mul(a, qm31_from4(v1)), qm31_from4(v0));
        qm31 r = qm31_add(qm31_mul(qm31_from4(d), a2), fp);
        qm31_to4(r, d);
        for (int k = 0; k < 4; k++) dst[k][i] = d[k];
    }
    free(pts);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Blake2sChannel (STWO/core/channel/blake2s.rs)                             */
/* ------------------------------------------------------------------------ */

typedef struct { uint8_t digest[32]; uint32_t n_draws; } or_channel;

/* update_digest resets n_draws, blake2s.rs:31-34. */
static void ch_update(or_channel *ch, const uint8_t d[32]) { memcpy(ch->digest, d, 32); ch->n_draws = 0; }
/* MerkleChannel::mix_root: digest = H(digest || root), STWO/core/vcs_lifted/blake2_merkle.rs:49-55. */
void or_channel_mix_root(or_channel *ch, const uint8_t root[32]) {
    b2s s; b2_init(&s); b2_update(&s, ch->digest, 32); b2_update(&s, root, 32);
    uint8_t h[32]; b2_final(&s, h); ch_update(ch, h);
}
/* mix_u32s, blake2s.rs:70-78. */
void or_channel_mix_u32s(or_channel *ch, const uint32_t *data, size_t n) {
    b2s s; b2_init(&s); b2_update(&s, ch->digest, 32); b2_update(&s, (const uint8_t *)data, 4 * n);
    uint8_t h[32]; b2_final(&s, h); ch_update(ch, h);
}
/* mix_felts, blake2s.rs:58-68 (QM31 as 4 LE u32 limbs each). */
void or_channel_mix_felts(or_channel *ch, const uint32_t *felts4, size_t n_felts) {
    b2s s; b2_init(&s); b2_update(&s, ch->digest, 32); b2_update(&s, (const uint8_t *)felts4, 16 * n_felts);
    uint8_t h[32]; b2_final(&s, h); ch_update(ch, h);
}
/* draw_u32s: H(digest || le32(n_draws) || 0x00), n_draws += 1, blake2s.rs:103-117. */
void or_channel_draw_u32s(or_channel *ch, uint32_t out[8]) {
    uint8_t in[37]; memcpy(in, ch->digest, 32); memcpy(in + 32, &ch->n_draws, 4); in[36] = 0;
    ch->n_draws += 1;
    uint8_t h[32]; or_blake2s(in, 37, h); memcpy(out, h, 32);
}
/* draw_secure_felt via draw_base_felts (retry unless all 8 words < 2P), blake2s.rs:36-51,85-88. */
void or_channel_draw_secure_felt(or_channel *ch, uint32_t out[4]) {
    uint32_t w[8];
    for (;;) {
        or_channel_draw_u32s(ch, w);
        int ok = 1;
        for (int i = 0; i < 8; i++) if (w[i] >= 2 * P) ok = 0;
        if (ok) break;
    }
    for (int i = 0; i < 4; i++) out[i] = m31_reduce(w[i]);
}

/* ------------------------------------------------------------------------ */
/* Line IFFT for the FRI last layer (STWO/prover/line.rs:54-99)               */
/* ------------------------------------------------------------------------ */

/* LineEvaluation::interpolate: bit-reverse, line_ifft over the natural-order
 * domain, scale by 1/len.  Coefficients come out in the LinePoly's internal
 * (bit-reversed) order.  QM31 values as [n][4]. */
static void line_interpolate(uint32_t *vals4, uint32_t line_initial, uint32_t log_n) {
    uint64_t n = (uint64_t)1 << log_n;
    /* bit_reverse the QM31 column */
    for (uint64_t i = 0; i < n; i++) {
        uint64_t j = bitrev(i, log_n);
        if (j > i) for (int k = 0; k < 4; k++) { uint32_t t = vals4[4 * i + k]; vals4[4 * i + k] = vals4[4 * j + k]; vals4[4 * j + k] = t; }
    }
    uint32_t init = line_initial, lg = log_n;
    while (lg >= 1) {
        uint64_t size = (uint64_t)1 << lg;
        cpoint c = index_to_point(init), st = index_to_point(coset_step(lg));
        for (uint64_t i = 0; i < size / 2; i++) {
            uint32_t xinv = m31_inv(c.x);
            for (uint64_t base = 0; base < n; base += size)
                for (int k = 0; k < 4; k++) ibfly(&vals4[4 * (base + i) + k], &vals4[4 * (base + i + size / 2) + k], xinv);
            c = cp_add(c, st);
        }
        init = (uint32_t)(((uint64_t)init * 2) & IDX_MASK);
        lg -= 1;
    }
    uint32_t inv = m31_inv((uint32_t)(n % P));
    for (uint64_t i = 0; i < 4 * n; i++) vals4[i] = m31_mul(vals4[i], inv);
}

/* ------------------------------------------------------------------------ */
/* FriProver::commit (STWO/prover/fri.rs:133-291)                             */
/* ------------------------------------------------------------------------ */

typedef struct { uint32_t log_blowup_factor, log_last_layer_degree_bound, n_queries, line_fold_step; } or_fri_config;

/* Commits a QM31 evaluation on CanonicCoset(log_size).circle_domain():
 *   first layer: lifted commit of the 4 coordinate columns, mix root;
 *   inner layers: draw alpha, fold (circle->line, then line folds of
 *   line_fold_step, the last one partial), commit each folded layer;
 *   last layer: interpolate, check the zero tail, mix the coefficients.
 * roots: 8 words per committed layer (first layer first).  last_poly:
 * 4 words per coefficient, 2^log_last_layer_degree_bound coefficients in
 * LinePoly order.  Returns the number of committed layers, -1 on invalid
 * config/size, -2 if the degree check fails (fri.rs:281-285 panics there). */
int or_fri_commit(or_channel *ch, const or_fri_config *cfg, const uint32_t *const col[4], uint32_t log_size,
                  uint32_t *roots, uint32_t *last_poly, uint32_t **layer_out /* nullable: per layer 4 SoA base */) {
    uint32_t last_log = cfg->log_last_layer_degree_bound + cfg->log_blowup_factor;
    if (log_size < 2 || log_size - 1 < last_log || cfg->line_fold_step < 1) return -1;
    int n_roots = 0;
    uint64_t N = (uint64_t)1 << log_size;
    /* first layer (fri.rs:176-183, 360-369) */
    {
        uint32_t ls[4] = {log_size, log_size, log_size, log_size};
        uint32_t *layers = (uint32_t *)malloc(((2 * N) - 1) * 32);
        or_lifted_commit(col, ls, 4, log_size, layers, 0);
        memcpy(roots, layers, 32); n_roots++;
        or_channel_mix_root(ch, (const uint8_t *)layers);
        free(layers);
    }
    /* inner layers (fri.rs:188-262) */
    uint32_t lg = log_size - 1;
    uint32_t line_init = half_odds_initial(lg); /* LineDomain(half_odds(log N - 1)), fri.rs:74-85 */
    uint32_t *cur[4];
    uint32_t alpha[4];
    or_channel_draw_secure_felt(ch, alpha);
    for (int k = 0; k < 4; k++) { cur[k] = (uint32_t *)calloc((size_t)1 << lg, 4); }
    or_fold_circle_into_line(cur, col, half_odds_initial(log_size - 1), log_size, alpha);
    if (lg != last_log) {
        for (;;) {
            uint32_t step = cfg->line_fold_step;
            int last = !(lg > last_log + cfg->line_fold_step);
            if (last) step = lg - last_log;
            /* commit this layer */
            uint64_t n = (uint64_t)1 << lg;
            uint32_t ls[4] = {lg, lg, lg, lg};
            uint32_t *layers = (uint32_t *)malloc(((2 * n) - 1) * 32);
            or_lifted_commit((const uint32_t *const *)cur, ls, 4, lg, layers, 0);
            memcpy(roots + 8 * n_roots, layers, 32); n_roots++;
            if (layer_out) for (int k = 0; k < 4; k++) memcpy(layer_out[n_roots - 2] + k * n, cur[k], n * 4);
            or_channel_mix_root(ch, (const uint8_t *)layers);
            free(layers);
            or_channel_draw_secure_felt(ch, alpha);
            uint32_t *nx[4];
            for (int k = 0; k < 4; k++) nx[k] = (uint32_t *)malloc((n >> step) * 4);
            or_fold_line((const uint32_t *const *)cur, line_init, lg, alpha, step, nx);
            for (int k = 0; k < 4; k++) { free(cur[k]); cur[k] = nx[k]; }
            for (uint32_t s = 0; s < step; s++) line_init = (uint32_t)(((uint64_t)line_init * 2) & IDX_MASK);
            lg -= step;
            if (last) break;
        }
    }
    /* last layer (fri.rs:274-291) */
    uint64_t n = (uint64_t)1 << lg;
    uint32_t *vals4 = (uint32_t *)malloc(n * 16);
    for (uint64_t i = 0; i < n; i++) for (int k = 0; k < 4; k++) vals4[4 * i + k] = cur[k][i];
    for (int k = 0; k < 4; k++) free(cur[k]);
    line_interpolate(vals4, line_init, lg);
    /* into_ordered_coefficients = bit_reverse (line.rs:157-160) */
    for (uint64_t i = 0; i < n; i++) {
        uint64_t j = bitrev(i, lg);
        if (j > i) for (int k = 0; k < 4; k++) { uint32_t t = vals4[4 * i + k]; vals4[4 * i + k] = vals4[4 * j + k]; vals4[4 * j + k] = t; }
    }
    uint64_t bound = (uint64_t)1 << cfg->log_last_layer_degree_bound;
    int ok = 1;
    for (uint64_t i = bound; i < n; i++) for (int k = 0; k < 4; k++) if (vals4[4 * i + k]) ok = 0;
    if (!ok) { free(vals4); return -2; }
    /* from_ordered_coefficients bit-reverses back (line.rs:167-170) */
    for (uint64_t i = 0; i < bound; i++) {
        uint64_t j = bitrev(i, cfg->log_last_layer_degree_bound);
        for (int k = 0; k < 4; k++) last_poly[4 * j + k] = vals4[4 * i + k];
    }
    free(vals4);
    or_channel_mix_felts(ch, last_poly, bound);
    return n_roots;
}

/* CpuBackend::eval_at_point at a secure point (CPUB/circle.rs:82-99): fold
 * (STWO/core/poly/utils.rs:27-38) with factors [pi^(n-2)(x), ..., x, y],
 * evaluated bottom-up: adjacent pairs with y, then x, then pi(x), ... */
void or_eval_at_point(const uint32_t *coeffs, uint32_t log_size, const uint32_t px[4], const uint32_t py[4],
                      uint32_t out[4]) {
    size_t n = (size_t)1 << log_size;
    qm31 *lv = (qm31 *)malloc(n * sizeof(qm31));
    for (size_t j = 0; j < n; j++) { qm31 v = {{coeffs[j], 0}, {0, 0}}; lv[j] = v; }
    qm31 x = qm31_from4(px), g = qm31_from4(py);
    for (uint32_t k = 0; k < log_size; k++) {
        if (k >= 1) {
            g = x;
            qm31 sq = qm31_mul(x, x);
            qm31 one = {{1, 0}, {0, 0}};
            x = qm31_sub(qm31_add(sq, sq), one);
        }
        size_t m = n >> (k + 1);
        for (size_t t = 0; t < m; t++) lv[t] = qm31_add(lv[2 * t], qm31_mul(lv[2 * t + 1], g));
    }
    qm31_to4(lv[0], out);
    free(lv);
}

/* ------------------------------------------------------------------------ */
/* Wide-Fibonacci AIR: generate_trace (examples/src/wide_fibonacci/mod.rs:27-50) */
/* and the single-component constraint quotients (constraint-framework/src/     */
/* prover/component_prover.rs:170-274, CpuDomainEvaluator::add_constraint,      */
/* DomainEvaluationAccumulator, STWO/prover/air/accumulation.rs:30-68;           */
/* coset_vanishing, STWO/core/constraints.rs:12-35).                             */
/* ------------------------------------------------------------------------ */
void or_wide_fib_trace(uint32_t **cols, uint32_t n_cols, uint32_t log_n, const uint32_t *a, const uint32_t *b) {
    size_t n = (size_t)1 << log_n;
    for (size_t i = 0; i < n; i++) {
        uint32_t x0 = a[i], x1 = b[i];
        cols[0][i] = x0; cols[1][i] = x1;
        for (uint32_t k = 2; k < n_cols; k++) {
            uint32_t c = m31_add(m31_mul(x0, x0), m31_mul(x1, x1));
            cols[k][i] = c; x0 = x1; x1 = c;
        }
    }
}

/* coset_vanishing(coset, p) for p in the base field: shift p by -initial +
 * step/2 (indices), then double its x log_size - 1 times. */
static uint32_t coset_vanishing_m31(uint32_t coset_initial, uint32_t coset_log, cpoint p) {
    uint32_t shift = (uint32_t)((((uint64_t)1 << 31) - coset_initial + ((1ull << (31 - coset_log)) >> 1)) & IDX_MASK);
    cpoint q = cp_add(p, index_to_point(shift));
    uint32_t x = q.x;
    for (uint32_t d = 1; d < coset_log; d++) x = m31_sub(m31_add(m31_mul(x, x), m31_mul(x, x)), 1);
    return x;
}

int or_wide_fib_constraint_quotients(uint32_t **cols, uint32_t n_cols, uint32_t log_n, const uint32_t alpha4[4],
                                     uint32_t **accum) {
    if (n_cols < 3) return -1;
    uint32_t eval_log = log_n + 1, n_cons = n_cols - 2;
    /* DomainEvaluationAccumulator::new: powers alpha^0 .. alpha^(total-1); the
     * component takes the last n_cons (all of them) and reverses them. */
    qm31 alpha = qm31_from4(alpha4);
    qm31 *pw = (qm31 *)malloc(n_cons * sizeof(qm31));
    qm31 p = {{1, 0}, {0, 0}};
    for (uint32_t e = 0; e < n_cons; e++) { pw[e] = p; p = qm31_mul(p, alpha); }
    for (uint32_t e = 0; e < n_cons / 2; e++) { qm31 t = pw[e]; pw[e] = pw[n_cons - 1 - e]; pw[n_cons - 1 - e] = t; }
    /* denominators at eval_domain.at(i), i < 2^(eval_log - log_n), bit-reversed */
    uint32_t log_expand = eval_log - log_n;
    size_t ne = (size_t)1 << log_expand;
    uint32_t *dinv = (uint32_t *)malloc(ne * 4);
    uint32_t coset_initial = (uint32_t)((1ull << (31 - (log_n + 1))) & IDX_MASK); /* Coset::odds(log_n) */
    uint32_t half_init = half_odds_initial(eval_log - 1);
    uint32_t half_step = (uint32_t)((1ull << (31 - (eval_log - 1))) & IDX_MASK);
    for (size_t i = 0; i < ne; i++) {
        cpoint dp = index_to_point((uint32_t)((half_init + (uint64_t)half_step * i) & IDX_MASK));
        dinv[bitrev(i, log_expand)] = m31_inv(coset_vanishing_m31(coset_initial, log_n, dp));
    }
    size_t n = (size_t)1 << eval_log;
    #pragma omp parallel for schedule(static)
    for (size_t r = 0; r < n; r++) {
        qm31 res = {{0, 0}, {0, 0}};
        uint32_t x0 = cols[0][r], x1 = cols[1][r];
        for (uint32_t k = 2; k < n_cols; k++) {
            uint32_t c = cols[k][r];
            uint32_t con = m31_sub(c, m31_add(m31_mul(x0, x0), m31_mul(x1, x1)));
            res = qm31_add(res, qm31_mul_m31(pw[k - 2], con));
            x0 = x1; x1 = c;
        }
        res = qm31_mul_m31(res, dinv[r >> log_n]);
        accum[0][r] = m31_add(accum[0][r], res.a.a); accum[1][r] = m31_add(accum[1][r], res.a.b);
        accum[2][r] = m31_add(accum[2][r], res.b.a); accum[3][r] = m31_add(accum[3][r], res.b.b);
    }
    free(pw); free(dinv);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* FRI quotients: compute_fri_quotients (STWO/prover/pcs/quotient_ops.rs:131-232) */
/* with CpuBackend's QuotientOps (CPUB/quotients.rs) and AccumulationOps::       */
/* lift_and_accumulate (CPUB/accumulation.rs), helpers from                      */
/* STWO/core/pcs/quotients.rs and STWO/core/constraints.rs:119-134.              */
/* ------------------------------------------------------------------------ */
typedef struct { uint32_t x[4], y[4]; } or_spoint;             /* CirclePoint<SecureField> */
typedef struct { or_spoint point; uint32_t value[4]; } or_sample; /* PointSample */

static qm31 qm31_conj(qm31 v) { qm31 r = {v.a, {m31_neg(v.b.a), m31_neg(v.b.b)}}; return r; }
static int qm31_eq(qm31 x, qm31 y) { return x.a.a == y.a.a && x.a.b == y.a.b && x.b.a == y.b.a && x.b.b == y.b.b; }

typedef struct { qm31 px, py; qm31 value; qm31 rand; int col; } q_entry;  /* (sample, alpha^k, column) */
typedef struct { qm31 px, py; size_t n; q_entry *e; } q_batch;
typedef struct { qm31 px, py; uint32_t log; qm31 *partial; qm31 first; } q_acc;

int or_compute_fri_quotients(const uint32_t *const *cols, const uint32_t *log_sizes, size_t n_cols,
                             const or_sample *samples, const uint32_t *n_samples, const uint32_t random_coeff[4],
                             uint32_t lifting_log, uint32_t *const out[4]) {
    if (lifting_log < 1 || lifting_log > 30) return -1;
    for (size_t c = 0; c < n_cols; c++) if (log_sizes[c] < 1 || log_sizes[c] > lifting_log || n_samples[c] > 64) return -1;
    /* 1. samples with randomness and periodicity (build_samples_with_randomness_and_periodicity,
     *    quotient_ops.rs:80-127): alpha^0, alpha^1, ... over columns in order; a column with
     *    exactly two samples first gets (second point + period generator, second value). */
    qm31 alpha = qm31_from4(random_coeff), pw = {{1, 0}, {0, 0}};
    size_t total = 0;
    for (size_t c = 0; c < n_cols; c++) total += n_samples[c] + (n_samples[c] == 2);
    q_entry *all = (q_entry *)calloc(total ? total : 1, sizeof(q_entry));
    size_t k = 0, si = 0;
    for (size_t c = 0; c < n_cols; c++) {
        const or_sample *sm = samples + si;
        if (n_samples[c] == 2) {
            /* CanonicCoset(L).step() repeated_double(log_size): index 2^(31 - L + log) */
            uint64_t idx = ((uint64_t)1 << (31 - lifting_log + log_sizes[c])) & IDX_MASK;
            cpoint g = index_to_point((uint32_t)idx);
            qm31 x1 = qm31_from4(sm[1].point.x), y1 = qm31_from4(sm[1].point.y);
            q_entry e;
            e.px = qm31_sub(qm31_mul_m31(x1, g.x), qm31_mul_m31(y1, g.y));
            e.py = qm31_add(qm31_mul_m31(y1, g.x), qm31_mul_m31(x1, g.y));
            e.value = qm31_from4(sm[1].value); e.rand = pw; e.col = (int)c;
            all[k++] = e; pw = qm31_mul(pw, alpha);
        }
        for (uint32_t j = 0; j < n_samples[c]; j++) {
            q_entry e;
            e.px = qm31_from4(sm[j].point.x); e.py = qm31_from4(sm[j].point.y);
            e.value = qm31_from4(sm[j].value); e.rand = pw; e.col = (int)c;
            all[k++] = e; pw = qm31_mul(pw, alpha);
        }
        si += n_samples[c];
    }
    /* 2-4. per log size ascending: batches by point in first-appearance order
     *      (ColumnSampleBatch::new_vec), line coefficients, partial numerators
     *      (accumulate_numerators, CPUB/quotients.rs:22-46). */
    q_acc *accs = (q_acc *)calloc(total ? total : 1, sizeof(q_acc));
    size_t n_accs = 0;
    for (uint32_t lg = 1; lg <= lifting_log; lg++) {
        q_batch *batches = (q_batch *)calloc(total ? total : 1, sizeof(q_batch));
        size_t nb = 0;
        for (size_t i = 0; i < total; i++) {
            if (log_sizes[all[i].col] != lg) continue;
            size_t b = 0;
            while (b < nb && !(qm31_eq(batches[b].px, all[i].px) && qm31_eq(batches[b].py, all[i].py))) b++;
            if (b == nb) {
                batches[nb].px = all[i].px; batches[nb].py = all[i].py; batches[nb].n = 0;
                batches[nb].e = (q_entry *)calloc(total, sizeof(q_entry)); nb++;
            }
            batches[b].e[batches[b].n++] = all[i];
        }
        size_t n = (size_t)1 << lg;
        for (size_t b = 0; b < nb; b++) {
            q_acc *A = &accs[n_accs++];
            A->px = batches[b].px; A->py = batches[b].py; A->log = lg;
            A->partial = (qm31 *)calloc(n, sizeof(qm31));
            qm31 first = {{0, 0}, {0, 0}};
            qm31 *la = (qm31 *)malloc(batches[b].n * sizeof(qm31)), *lb = (qm31 *)malloc(batches[b].n * sizeof(qm31)),
                 *lc = (qm31 *)malloc(batches[b].n * sizeof(qm31));
            for (size_t j = 0; j < batches[b].n; j++) {
                const q_entry *e = &batches[b].e[j];
                /* complex_conjugate_line_coeffs (constraints.rs:119-134) */
                qm31 a = qm31_sub(qm31_conj(e->value), e->value);
                qm31 cc = qm31_sub(qm31_conj(e->py), e->py);
                qm31 bb = qm31_sub(qm31_mul(e->value, cc), qm31_mul(a, e->py));
                la[j] = qm31_mul(e->rand, a); lb[j] = qm31_mul(e->rand, bb); lc[j] = qm31_mul(e->rand, cc);
                first = qm31_add(first, la[j]);
            }
            A->first = first;
            #pragma omp parallel for schedule(static)
            for (size_t r = 0; r < n; r++) {
                qm31 acc = {{0, 0}, {0, 0}};
                for (size_t j = 0; j < batches[b].n; j++) {
                    uint32_t v = cols[batches[b].e[j].col][r];
                    acc = qm31_add(acc, qm31_sub(qm31_mul_m31(lc[j], v), lb[j]));
                }
                A->partial[r] = acc;
            }
            free(la); free(lb); free(lc);
        }
        for (size_t b = 0; b < nb; b++) free(batches[b].e);
        free(batches);
    }
    /* 5. per sample point: sum the a-terms, lift_and_accumulate the partials (ascending sizes). */
    int *used = (int *)calloc(n_accs ? n_accs : 1, sizeof(int));
    q_acc *pts = (q_acc *)calloc(n_accs ? n_accs : 1, sizeof(q_acc));
    size_t n_pts = 0;
    for (size_t i = 0; i < n_accs; i++) {
        if (used[i]) continue;
        uint32_t clog = 1;
        qm31 *cur = (qm31 *)calloc(2, sizeof(qm31));
        qm31 first = {{0, 0}, {0, 0}};
        for (size_t j = i; j < n_accs; j++) {
            if (used[j] || !qm31_eq(accs[j].px, accs[i].px) || !qm31_eq(accs[j].py, accs[i].py)) continue;
            used[j] = 1;
            first = qm31_add(first, accs[j].first);
            uint32_t lr = accs[j].log - clog;
            size_t n = (size_t)1 << accs[j].log;
            qm31 *nx = (qm31 *)malloc(n * sizeof(qm31));
            for (size_t r = 0; r < n; r++) nx[r] = qm31_add(accs[j].partial[r], cur[((r >> (lr + 1)) << 1) + (r & 1)]);
            free(cur); cur = nx; clog = accs[j].log;
        }
        pts[n_pts].px = accs[i].px; pts[n_pts].py = accs[i].py; pts[n_pts].log = clog;
        pts[n_pts].partial = cur; pts[n_pts].first = first; n_pts++;
    }
    /* 6. compute_quotients_and_combine (CPUB/quotients.rs:48-76) with denominator_inverses
     *    (core/pcs/quotients.rs:262-287). */
    size_t N = (size_t)1 << lifting_log;
    uint32_t half_init = half_odds_initial(lifting_log - 1);
    uint32_t half_step = (uint32_t)(((uint64_t)1 << (31 - (lifting_log - 1))) & IDX_MASK);
    #pragma omp parallel for schedule(static)
    for (size_t row = 0; row < N; row++) {
        size_t i = bitrev(row, lifting_log);
        size_t h = N / 2;
        uint32_t idx = (uint32_t)((half_init + (uint64_t)half_step * (i < h ? i : i - h)) & IDX_MASK);
        if (i >= h) idx = (uint32_t)((((uint64_t)1 << 31) - idx) & IDX_MASK);
        cpoint dp = index_to_point(idx);
        qm31 q = {{0, 0}, {0, 0}};
        for (size_t s = 0; s < n_pts; s++) {
            cm31 prx = pts[s].px.a, pix = pts[s].px.b, pry = pts[s].py.a, piy = pts[s].py.b;
            cm31 dx = {m31_sub(prx.a, dp.x), prx.b}, dy = {m31_sub(pry.a, dp.y), pry.b};
            cm31 den = cm31_sub(cm31_mul(dx, piy), cm31_mul(dy, pix));
            cm31 dinv = cm31_inv(den);
            uint32_t lr = lifting_log - pts[s].log;
            qm31 full = qm31_sub(pts[s].partial[((row >> (lr + 1)) << 1) + (row & 1)], qm31_mul_m31(pts[s].first, dp.y));
            qm31 t = {cm31_mul(full.a, dinv), cm31_mul(full.b, dinv)};
            q = qm31_add(q, t);
        }
        out[0][row] = q.a.a; out[1][row] = q.a.b; out[2][row] = q.b.a; out[3][row] = q.b.b;
    }
    for (size_t s = 0; s < n_pts; s++) free(pts[s].partial);
    for (size_t i = 0; i < n_accs; i++) free(accs[i].partial);
    free(pts); free(used); free(accs); free(all);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Proof of work (GrindOps for the Blake2s channel)                          */
/* ------------------------------------------------------------------------ */
#define OR_POW_PREFIX 0x12345678u
/* H(POW_PREFIX || 0^12 || digest || le32(pow_bits)), STWO/core/channel/blake2s.rs:124-131. */
void or_pow_prefixed_digest(const uint8_t digest[32], uint32_t pow_bits, uint8_t out[32]) {
    uint8_t in[52];
    uint32_t pre = OR_POW_PREFIX;
    memcpy(in, &pre, 4);
    memset(in + 4, 0, 12);
    memcpy(in + 16, digest, 32);
    memcpy(in + 48, &pow_bits, 4);
    or_blake2s(in, 52, out);
}
/* Trailing zero bits of the first word of H(prefixed || le64(nonce))
 * (blake2s.rs:132-137 counts them over the first 16 bytes; for pow_bits <= 32
 * the first word decides, as SIMD/grind.rs:77-88 uses). */
static uint32_t pow_zeros(const uint8_t prefixed[32], uint64_t nonce) {
    uint8_t in[40], h[32];
    memcpy(in, prefixed, 32);
    memcpy(in + 32, &nonce, 8);
    or_blake2s(in, 40, h);
    uint32_t w; memcpy(&w, h, 4);
    return w ? (uint32_t)__builtin_ctz(w) : 32;
}
/* SimdBackend's grind (SIMD/grind.rs:22-62): the smallest nonce (hi << 32) | lo
 * with lo < 2^low_bits, hi = 0, 1, ..., whose hash has >= pow_bits trailing
 * zeros.  Scans hi in order, lo in parallel.  Returns 0 on success. */
int or_grind(const uint8_t digest[32], uint32_t pow_bits, uint32_t low_bits, uint64_t *nonce) {
    if (pow_bits > 32 || low_bits > 30) return -1;
    uint8_t pre[32];
    or_pow_prefixed_digest(digest, pow_bits, pre);
    for (uint64_t hi = 0; hi < 0x7fffffffull; hi++) {
        uint64_t best = UINT64_MAX;
        int64_t n = (int64_t)1 << low_bits;
        #pragma omp parallel for schedule(static) reduction(min:best)
        for (int64_t lo = 0; lo < n; lo++) {
            uint64_t cand = (hi << 32) | (uint64_t)lo;
            if (pow_zeros(pre, cand) >= pow_bits && cand < best) best = cand;
        }
        if (best != UINT64_MAX) { *nonce = best; return 0; }
    }
    return -1;
}

/* ------------------------------------------------------------------------ */
/* Seeded synthetic inputs (not from the reference): SplitMix64 -> u64 % P.  */
/* Shared with the library's bench generator so both sides see one input.    */
/* ------------------------------------------------------------------------ */
static inline uint64_t splitmix64(uint64_t *s) {
    uint64_t z = (*s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
/* Element j of (seed, stream) is one SplitMix64 step from a state mixed from
 * seed, stream and j, reduced as u64 % P (uniform in [0, P) up to a 2^-33
 * bias).  Random access, so columns fill in parallel on CPU and GPU alike. */
void or_fill_random(uint32_t *out, size_t n, uint64_t seed, uint64_t stream) {
    #pragma omp parallel for schedule(static)
    for (int64_t j = 0; j < (int64_t)n; j++) {
        uint64_t s = (seed << 32) ^ (stream * 0xD1B54A32D192ED03ull) ^ ((uint64_t)j * 0x9E3779B97F4A7C15ull);
        uint64_t z = splitmix64(&s);
        out[j] = (uint32_t)(z % P);
    }
}

int or_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

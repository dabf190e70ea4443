/*
 * stwo_hip.h -- C ABI of the MI355X (gfx950) implementation of the stwo
 * Circle-STARK commitment hot path.
 *
 * Each entry point replaces one operation of stwo's Backend traits; the
 * comment on each names the trait method it stands in for and the reference
 * interface it replaces.  Paths are relative to the reference checkout:
 *   STWO = vendor/stwo-upstream-dev-62b228e/crates/stwo/src
 *   SYS  = crates/stwo-metal-sys      (the reference's FFI crate)
 *   MB   = crates/stwo-metal/src/backend/metal
 *
 * Conventions
 *   - Device memory is plain `uint32_t *` (one M31 element or one hash word per
 *     u32).  A BaseField column of log size n is 2^n words in bit-reversed
 *     domain order, as in CircleEvaluation<_, BaseField, BitReversedOrder>.
 *     A SecureField (QM31) column is 4 separate coordinate columns
 *     (SecureColumnByCoords, STWO/prover/secure_column.rs).
 *   - A Blake2s hash is 8 words (its 32 bytes, little endian).  A Merkle tree
 *     is returned as packed layers: layer l (2^l hashes) starts at word
 *     (2^l - 1) * 8, so the root is words [0, 8) -- the reference's
 *     `layers[0]` -- and the tree of max layer L has (2^(L+1) - 1) * 8 words.
 *   - Arrays of columns are host arrays of device pointers.
 *   - Every call is enqueued on the context's stream and returns without
 *     waiting unless its comment says it synchronizes.  Results are ready
 *     after stwo_hip_sync().
 *   - Errors: every function returns STWO_HIP_OK (0) or a negative status;
 *     the message is kept per context (stwo_hip_last_error).  The library
 *     never aborts.  This replaces the reference's `bool` + 512-byte error
 *     buffer convention (SYS/src/metal.rs:8,4477-4485).
 *   - Values must be canonical M31 ([0, P), P = 2^31 - 1); every output is
 *     canonical.
 */
#ifndef STWO_HIP_H
#define STWO_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes --------------------------------------------------- */
#define STWO_HIP_OK 0
#define STWO_HIP_EINVAL (-1)   /* bad argument (size, order, NULL, domain)  */
#define STWO_HIP_ENOMEM (-2)   /* device or host allocation failed          */
#define STWO_HIP_EDEVICE (-3)  /* HIP runtime error                         */
#define STWO_HIP_EDEGREE (-4)  /* FRI last layer is not of low degree       */
#define STWO_HIP_EBUSY (-5)    /* serving queue full / budget exceeded      */
#define STWO_HIP_ECONSTRAINTS (-6) /* prove: constraints not satisfied     */

/* Hash output mode: Blake2sMerkleHasher or Blake2sM31MerkleHasher
 * (STWO/core/vcs/blake2_merkle.rs:8-10): the M31 variant reduces each output
 * word mod P (reduce_to_m31, STWO/core/vcs/blake2_hash.rs:111).  Any other
 * value is STWO_HIP_EINVAL. */
#define STWO_HIP_HASH_BLAKE2S 0
#define STWO_HIP_HASH_BLAKE2S_M31 1

typedef struct stwo_hip_ctx stwo_hip_ctx;
typedef struct stwo_hip_twiddles stwo_hip_twiddles;

/* A coset `initial_index + <2^(31 - log_size)>` of the circle group, by its
 * CirclePointIndex (STWO/core/circle.rs:286-305).  Circle domains are given
 * by their half coset (CircleDomain, STWO/core/poly/circle/domain.rs:22-32);
 * CanonicCoset(n).circle_domain() has half coset (2^(30-n), n-1). */
typedef struct {
  uint32_t initial_index;
  uint32_t log_size;
} stwo_hip_coset;

/* ---- context ------------------------------------------------------------
 * Replaces the process-global Metal runtime (SYS/src/metal.rs:4442-4448,
 * runtime creation SH/runtime.m:476-512).  One context = one device + one
 * stream; create several to run independent work concurrently. */
int stwo_hip_ctx_create(int device, stwo_hip_ctx **out);
int stwo_hip_ctx_destroy(stwo_hip_ctx *ctx);
/* Copies the last error message (NUL-terminated, truncated to len). */
int stwo_hip_last_error(const stwo_hip_ctx *ctx, char *buf, size_t len);
/* Blocks until all work enqueued on the context's stream is done. */
int stwo_hip_sync(stwo_hip_ctx *ctx);
/* The context's hipStream_t, for callers that order their own work. */
void *stwo_hip_stream(stwo_hip_ctx *ctx);
/* Test hook for the error paths; no counterpart in the reference.  After
 * this call the context's next `after` device allocations succeed and every
 * later one fails as an exhausted device would (STWO_HIP_ENOMEM with a
 * message), until it is called again with after < 0 (off, the default).
 * Counted: the library's own scratch, tables, plans and twiddle trees, and
 * stwo_hip_alloc; not allocations made before the call. */
int stwo_hip_ctx_fail_allocs_after(stwo_hip_ctx *ctx, long long after);

/* ---- device memory ------------------------------------------------------
 * Replaces stwo_metal_u32_buffer_{alloc_zeroed,alloc_uninitialized,destroy,
 * from_host,read,copy} (SH/runtime.m:523-916). */
int stwo_hip_alloc(stwo_hip_ctx *ctx, size_t n_words, uint32_t **out);
int stwo_hip_free(stwo_hip_ctx *ctx, uint32_t *ptr);
/* Synchronous host<->device copies (they wait for the stream). */
int stwo_hip_upload(stwo_hip_ctx *ctx, uint32_t *dst, const uint32_t *src, size_t n_words);
int stwo_hip_download(stwo_hip_ctx *ctx, uint32_t *dst, const uint32_t *src, size_t n_words);
int stwo_hip_copy(stwo_hip_ctx *ctx, uint32_t *dst, const uint32_t *src, size_t n_words);
int stwo_hip_memset_zero(stwo_hip_ctx *ctx, uint32_t *dst, size_t n_words);
/* Fills `n` words with seeded uniform values in [0, P) (SplitMix64 of
 * (seed, stream, index), u64 % P).  Synthetic benchmark inputs. */
int stwo_hip_fill_random(stwo_hip_ctx *ctx, uint32_t *dst, size_t n, uint64_t seed, uint64_t stream);

/* ---- twiddles -----------------------------------------------------------
 * PolyOps::precompute_twiddles(coset) (STWO/prover/poly/circle/ops.rs:144;
 * CpuBackend layout CPUB/circle.rs:294-311): for each doubling of the root
 * coset, the x-coordinates of its first half in bit-reversed order, then a
 * pad of 1; plus the elementwise inverses, and both trees doubled (2t as a
 * u32, the FFT's multiplication operands, as SimdBackend keeps them): 16 bytes
 * per root point.  Built on the device and kept resident (replaces
 * MB/poly.rs:185-224 and SH/runtime.m:1026).  A tree may be shared by every
 * context on its device. */
int stwo_hip_twiddles_create(stwo_hip_ctx *ctx, stwo_hip_coset root, stwo_hip_twiddles **out);
int stwo_hip_twiddles_destroy(stwo_hip_ctx *ctx, stwo_hip_twiddles *tw);
/* Device pointers and length (= root size) of the tree and its inverse. */
int stwo_hip_twiddles_buffers(const stwo_hip_twiddles *tw, const uint32_t **twiddles, const uint32_t **itwiddles,
                              size_t *len);

/* ---- ColumnOps ------------------------------------------------------------
 * ColumnOps<BaseField>::bit_reverse_column (STWO/prover/backend/mod.rs:37-40;
 * Metal MB/column.rs:24-32 -> SH/bit_reverse.metal:13). In place. */
int stwo_hip_bit_reverse_column(stwo_hip_ctx *ctx, uint32_t *col, uint32_t log_size);
/* ColumnOps<SecureField>::bit_reverse_column on the 4 coordinate columns
 * (MB/column.rs:34-42 -> SH/bit_reverse.metal:31). */
int stwo_hip_bit_reverse_secure_column(stwo_hip_ctx *ctx, uint32_t *const coords[4], uint32_t log_size);

/* ---- PolyOps ------------------------------------------------------------
 * PolyOps::interpolate (STWO/prover/poly/circle/ops.rs:24-27; CPUB/circle.rs:28-80;
 * Metal MB/poly.rs:364-391 -> stwo_metal_ifft_interpolate_u32,
 * SH/runtime.m:1572).  `col` holds the bit-reversed evaluation on the circle
 * domain with half coset `domain`; it is replaced in place by the
 * coefficients (natural FFT-basis order).  The domain must be a doubling of
 * the tree's root coset (or the root itself). */
int stwo_hip_interpolate(stwo_hip_ctx *ctx, uint32_t *col, stwo_hip_coset domain, const stwo_hip_twiddles *tw);
/* PolyOps::interpolate_columns (ops.rs:29-40) over canonic domains
 * CanonicCoset(log_sizes[i]).circle_domain(); one batched launch per pass. */
int stwo_hip_interpolate_columns(stwo_hip_ctx *ctx, uint32_t *const *cols, const uint32_t *log_sizes, size_t n_cols,
                                 const stwo_hip_twiddles *tw);
/* PolyOps::evaluate / evaluate_into (ops.rs:90-108; CPUB/circle.rs:190-250;
 * Metal MB/poly.rs:332-362 -> stwo_metal_rfft_evaluate_u32, SH/runtime.m:1218).
 * Coefficients of log size `coeffs_log_size` are evaluated on the circle
 * domain with half coset `domain` (log size domain.log_size + 1 >=
 * coeffs_log_size); `out` gets 2^(domain.log_size+1) bit-reversed values. */
int stwo_hip_evaluate(stwo_hip_ctx *ctx, const uint32_t *coeffs, uint32_t coeffs_log_size, stwo_hip_coset domain,
                      const stwo_hip_twiddles *tw, uint32_t *out);
/* PolyOps::evaluate_polynomials (ops.rs:110-141; Metal MB/poly.rs:658-793):
 * column i (log size log_sizes[i]) onto CanonicCoset(log_sizes[i] + log_blowup). */
int stwo_hip_evaluate_polynomials(stwo_hip_ctx *ctx, const uint32_t *const *coeffs, const uint32_t *log_sizes,
                                  size_t n_cols, uint32_t log_blowup, const stwo_hip_twiddles *tw,
                                  uint32_t *const *outs);

/* TreeBuilder::extend_evals (STWO/prover/pcs/mod.rs:707-716: interpolate_columns)
 * followed by CommitmentTreeProver::new's evaluate_polynomials at log_blowup
 * (:766-784), as one call: cols[i] holds 2^log_sizes[i] bit-reversed
 * evaluations on CanonicCoset(log_sizes[i]) and is consumed; outs[i] receives
 * the 2^(log_sizes[i] + log_blowup) evaluations on CanonicCoset(log_sizes[i] +
 * log_blowup), bit-reversed.  store_coeffs != 0 (store_polynomials_coefficients,
 * default false at :347) leaves the coefficients in cols[i] as interpolate_columns
 * would; with 0 their contents are unspecified.  The IFFT's last pass and the
 * RFFT's first pass run as one kernel when their tiles match, so the
 * coefficients are not written to HBM and read back.  Same values as
 * stwo_hip_interpolate_columns + stwo_hip_evaluate_polynomials. */
int stwo_hip_extend_evals(stwo_hip_ctx *ctx, uint32_t *const *cols, const uint32_t *log_sizes, size_t n_cols,
                          uint32_t log_blowup, const stwo_hip_twiddles *tw, uint32_t *const *outs, int store_coeffs);

/* ---- MerkleOps<Blake2sMerkleHasher> (mixed degree) -----------------------
 * MerkleOps::commit_on_layer (STWO/prover/vcs/ops.rs:24-28; SimdBackend
 * SIMD/blake2s.rs:60-144; no Metal implementation exists).  Node i =
 * Blake2s([prev[2i] || prev[2i+1]] || le32(cols[0][i]) || ...).  `prev` is
 * NULL for the leaf layer, else 2^(log_size+1) hashes; each column has
 * 2^log_size values; `out` receives 2^log_size hashes. */
int stwo_hip_merkle_commit_on_layer(stwo_hip_ctx *ctx, uint32_t log_size, const uint32_t *prev,
                                    const uint32_t *const *cols, size_t n_cols, uint32_t *out, int hash_mode);
/* MerkleProver::commit (STWO/prover/vcs/prover.rs:39-65): columns of any
 * power-of-two sizes, each injected at its own layer (stable descending
 * order).  `layers` receives the packed tree of max layer = max log size
 * (root first).  With n_cols == 0 the root is Blake2s("") (8 words). */
int stwo_hip_merkle_commit(stwo_hip_ctx *ctx, const uint32_t *const *cols, const uint32_t *log_sizes, size_t n_cols,
                           uint32_t *layers, int hash_mode);

/* ---- MerkleOpsLifted<Blake2sMerkleHasher> ----------------------------------
 * MerkleOpsLifted::build_leaves (STWO/prover/vcs_lifted/ops.rs:13-15;
 * CPUB/merkle_lifted.rs:44-78; Metal MB/blake2s.rs:203-392 ->
 * SH/blake2s.metal:282-370).  Columns are hashed in stable ascending length
 * order; column c of log size s contributes c[lift(i, L - s)] to leaf i.
 * `out` receives 2^lifting_log_size hashes. */
int stwo_hip_merkle_lifted_build_leaves(stwo_hip_ctx *ctx, const uint32_t *const *cols, const uint32_t *log_sizes,
                                        size_t n_cols, uint32_t lifting_log_size, uint32_t *out, int hash_mode);
/* build_next_layer (ops.rs:19-20; SH/blake2s.metal:459): out[i] = H(prev[2i] || prev[2i+1]);
 * `out_log_size` is the log size of the output layer. */
int stwo_hip_merkle_lifted_build_next_layer(stwo_hip_ctx *ctx, const uint32_t *prev, uint32_t out_log_size,
                                            uint32_t *out, int hash_mode);
/* build_merkle_layers (ops.rs:29-45) -> MerkleProverLifted::commit
 * (STWO/prover/vcs_lifted/prover.rs:33-44): packed tree of max layer
 * lifting_log_size. */
int stwo_hip_merkle_lifted_commit(stwo_hip_ctx *ctx, const uint32_t *const *cols, const uint32_t *log_sizes,
                                  size_t n_cols, uint32_t lifting_log_size, uint32_t *layers, int hash_mode);

/* ---- FriOps -----------------------------------------------------------------
 * FriOps::fold_circle_into_line (STWO/prover/fri.rs:61-66; CPUB/fri.rs:95-123;
 * Metal MB/fri.rs:347-365 -> SH/fold_circle_into_line.metal): for i < N/2,
 * dst[i] = dst[i] * alpha^2 + f0 + alpha * f1 with (f0, f1) =
 * ibutterfly(src[2i], src[2i+1], 1/p.y), p = domain.at(bitrev(2i)).
 * `src_domain` is the half coset of src's circle domain (log N - 1); dst has
 * N/2 values.  `alpha` is a QM31 as 4 limbs. */
int stwo_hip_fold_circle_into_line(stwo_hip_ctx *ctx, uint32_t *const dst[4], const uint32_t *const src[4],
                                   stwo_hip_coset src_domain, const uint32_t alpha[4], const stwo_hip_twiddles *tw);
/* FriOps::fold_line (fri.rs:39-44; CPUB/fri.rs:15-30,67-92; SimdBackend
 * SIMD/fri.rs:32-122; Metal MB/fri.rs:295-345): `fold_step` folds with alpha,
 * alpha^2, ... of the line evaluation on LineDomain(`line_domain`);
 * dst has 2^(line_domain.log_size - fold_step) values. */
int stwo_hip_fold_line(stwo_hip_ctx *ctx, const uint32_t *const src[4], stwo_hip_coset line_domain,
                       const uint32_t alpha[4], uint32_t fold_step, const stwo_hip_twiddles *tw,
                       uint32_t *const dst[4]);

/* ---- FriProver::commit --------------------------------------------------------
 * Runs the FRI commitment phase (STWO/prover/fri.rs:133-291) for one QM31
 * evaluation on CanonicCoset(log_size).circle_domain() with a Blake2s
 * channel (STWO/core/channel/blake2s.rs): commit the first layer (lifted
 * Merkle of the 4 coordinate columns), then per inner layer: mix root, draw
 * alpha, fold, commit; the last layer is interpolated on the host, its
 * degree checked and its coefficients mixed into the channel.
 * channel_digest (32 bytes) and channel_n_draws are read and updated.
 * roots: (number of committed layers) * 8 words, first layer first;
 * n_roots receives that number.  last_layer_poly: 4 words per coefficient,
 * 2^log_last_layer_degree_bound coefficients in LinePoly order.
 * Synchronizes.  Returns STWO_HIP_EDEGREE if the degree check fails. */
typedef struct {
  uint32_t log_blowup_factor;
  uint32_t log_last_layer_degree_bound;
  uint32_t n_queries;
  uint32_t line_fold_step;
} stwo_hip_fri_config;

int stwo_hip_fri_commit(stwo_hip_ctx *ctx, uint8_t channel_digest[32], uint32_t *channel_n_draws,
                        stwo_hip_fri_config config, const uint32_t *const column[4], uint32_t log_size,
                        const stwo_hip_twiddles *tw, uint32_t *roots, uint32_t *n_roots, uint32_t *last_layer_poly);

/* ---- decommitment ---------------------------------------------------------
 * MerkleProverLifted::decommit (STWO/prover/vcs_lifted/prover.rs:58-201;
 * Metal MB/decommit.rs:96-220) on a packed lifted tree `layers` of lifting
 * log size L (as written by stwo_hip_merkle_lifted_commit for the same
 * columns).  `queries`: n_queries positions in [0, 2^L), non-decreasing
 * (host memory).  Outputs, host memory:
 *   values[c * n_queries + q] = column c at lift(queries[q], L - log_sizes[c])
 *     (the reference's queried_values, column order as given);
 *   hash_witness: the witness hashes (8 words each) in the reference's order,
 *     *n_witness of them (at most witness_cap);
 *   aux_nodes (optional, MerkleDecommitmentLiftedAux::all_node_values): 10
 *     words per node [layer log size, index, hash], both children of every
 *     visited parent, from the leaf layer up, *n_aux nodes (<= aux_cap).
 * One device gather; synchronizes. */
int stwo_hip_merkle_lifted_decommit(stwo_hip_ctx *ctx, const uint32_t *layers, uint32_t lifting_log_size,
                                    const uint32_t *const *cols, const uint32_t *log_sizes, size_t n_cols,
                                    const uint32_t *queries, size_t n_queries, uint32_t *values,
                                    uint32_t *hash_witness, size_t witness_cap, size_t *n_witness,
                                    uint32_t *aux_nodes, size_t aux_cap, size_t *n_aux);

/* FriProver (STWO/prover/fri.rs:117-347) kept for decommitment:
 * stwo_hip_fri_prover_commit is stwo_hip_fri_commit that keeps every
 * committed layer's tree and evaluation (one device allocation, about
 * 8x the input column's size); the input column is borrowed and must stay
 * valid until the prover is destroyed, as the reference's prover borrows
 * its first-layer evaluation. */
typedef struct stwo_hip_fri_prover stwo_hip_fri_prover;
int stwo_hip_fri_prover_commit(stwo_hip_ctx *ctx, uint8_t channel_digest[32], uint32_t *channel_n_draws,
                               stwo_hip_fri_config config, const uint32_t *const column[4], uint32_t log_size,
                               const stwo_hip_twiddles *tw, uint32_t *roots, uint32_t *n_roots,
                               uint32_t *last_layer_poly, stwo_hip_fri_prover **out);
/* FriProver::decommit (fri.rs:294-347): draws config.n_queries positions
 * from the channel (draw_queries, STWO/core/queries.rs:11-28), updating
 * channel_digest / channel_n_draws; query_positions receives the sorted,
 * deduplicated positions (*n_positions of them, at most n_queries) and
 * unsorted_queries the n_queries raw draws.  Decommits every layer (one
 * device gather); the proofs are then read per layer with
 * stwo_hip_fri_prover_layer_proof.  Once per prover. */
int stwo_hip_fri_prover_decommit(stwo_hip_ctx *ctx, stwo_hip_fri_prover *prover, uint8_t channel_digest[32],
                                 uint32_t *channel_n_draws, uint32_t *query_positions, uint32_t *n_positions,
                                 uint32_t *unsorted_queries);
/* FriLayerProof of committed layer `layer` (0 = first/circle layer;
 * fri.rs:372-453): commitment (root, 8 words); fri_witness (4 words per
 * QM31, *n_fri of them); hash_witness (8 words per hash, *n_hash).  If
 * `positions` is not NULL it receives the layer's decommitment positions
 * (*n_pos <= pos_cap) and `values` (if not NULL) their QM31 values, 4 words
 * each (FriLayerProofAux::all_values).  With fri_witness, hash_witness and
 * positions all NULL the call only reports the three counts. */
int stwo_hip_fri_prover_layer_proof(stwo_hip_ctx *ctx, const stwo_hip_fri_prover *prover, uint32_t layer,
                                    uint32_t commitment[8], uint32_t *fri_witness, size_t fri_cap, size_t *n_fri,
                                    uint32_t *hash_witness, size_t hash_cap, size_t *n_hash, uint32_t *positions,
                                    uint32_t *values, size_t pos_cap, size_t *n_pos);
int stwo_hip_fri_prover_destroy(stwo_hip_ctx *ctx, stwo_hip_fri_prover *prover);

/* ---- FRI quotients ------------------------------------------------------------
 * compute_fri_quotients (STWO/prover/pcs/quotient_ops.rs:131-232) with the
 * QuotientOps/AccumulationOps semantics of the reference backends
 * (CPUB/quotients.rs, CPUB/accumulation.rs; Metal MB/quotient.rs:388,
 * SH/quotients.metal).  Columns are the flattened trees' evaluations
 * (bit-reversed, CanonicCoset(log_sizes[i]), 1 <= log <= lifting_log_size);
 * `samples` holds, column after column, n_samples[i] PointSamples of column
 * i (a column with exactly two gets the periodicity sample, as
 * build_samples_with_randomness_and_periodicity does,
 * STWO/core/pcs/quotients.rs:289-336); random powers alpha^0, alpha^1, ... run
 * over columns in order.  `out` receives the 4 coordinate columns of the
 * quotient evaluation on CanonicCoset(lifting_log_size), bit-reversed.  One
 * kernel; the sample tables go through the context's upload ring. */
typedef struct {
  uint32_t x[4];
  uint32_t y[4];
} stwo_hip_secure_point; /* CirclePoint<SecureField> */
typedef struct {
  stwo_hip_secure_point point;
  uint32_t value[4];
} stwo_hip_point_sample; /* PointSample, STWO/core/pcs/quotients.rs */
int stwo_hip_compute_fri_quotients(stwo_hip_ctx *ctx, const uint32_t *const *cols, const uint32_t *log_sizes,
                                   size_t n_cols, const stwo_hip_point_sample *samples, const uint32_t *n_samples,
                                   const uint32_t random_coeff[4], uint32_t lifting_log_size, uint32_t *const out[4]);

/* PolyOps::eval_at_point for a batch of polynomials at one secure point
 * (STWO/prover/poly/circle/ops.rs; CPUB/circle.rs:82-99; Metal
 * batch_eval_at_point, SH/eval_at_point.metal): the out-of-domain samples.
 * coeffs[i] holds 2^log_sizes[i] coefficients in natural FFT-basis order
 * (as interpolate returns them); values receives 4 words per polynomial
 * (host memory).  Synchronizes. */
int stwo_hip_eval_at_point(stwo_hip_ctx *ctx, const uint32_t *const *coeffs, const uint32_t *log_sizes, size_t n_cols,
                           const stwo_hip_secure_point *point, uint32_t *values);

/* stwo_hip_eval_at_point for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  No counterpart in the reference: its only purpose is to make
 * no allocating call (no hipMalloc / hipFree / allocating library routine;
 * with launch profiling off).  Same arguments and results as
 * stwo_hip_eval_at_point.  The kernels' scratch (basis table, block
 * partials, per-column sums) is a fixed 1 MiB array per device, resolved at
 * context creation and held until the process ends: the call is rejected
 * with STWO_HIP_EINVAL when 8192 + 4 * (n_cols + blocks) words exceed 2^18
 * (blocks = sum over columns of max(1, 2^(log_size - 11)); e.g. up to 256
 * columns of 2^18 coefficients, or 48 of 2^20) or its tables exceed half
 * the context's 8 MiB table arena.  Every call on one device shares the
 * array: calls that overlap on one GPU (from two contexts or threads) race
 * and must be serialized by the caller.  Synchronizes. */
int stwo_hip_eval_at_point_noalloc(stwo_hip_ctx *ctx, const uint32_t *const *coeffs, const uint32_t *log_sizes,
                                   size_t n_cols, const stwo_hip_secure_point *point, uint32_t *values);

/* ---- the other PolyOps / FriOps methods -----------------------------------------
 * PolyOps::barycentric_weights (STWO/prover/poly/circle/ops.rs:67-71; CpuBackend
 * CPUB/circle.rs:101-132; Metal SH/barycentric.metal): the weights of
 * CanonicCoset(log_size) (bit-reversed order) for the secure point p, which
 * must not lie in the domain; out = 4 device coordinate columns of
 * 2^log_size.  Asynchronous. */
int stwo_hip_barycentric_weights(stwo_hip_ctx *ctx, uint32_t log_size, const stwo_hip_secure_point *p,
                                 uint32_t *const out[4]);
/* PolyOps::barycentric_eval_at_point (ops.rs:73-80; CPUB/circle.rs:134-141) for a
 * batch of columns of one size: values (host, 4 words per column) = sum_i
 * evals[c][i] * weights[i], evals[c] holding 2^log_size bit-reversed
 * evaluations on CanonicCoset(log_size) and weights those of
 * stwo_hip_barycentric_weights for that size and point -- the PCS's sampling
 * path when coefficients are not stored (STWO/prover/pcs/mod.rs:481-497, 638-657).
 * Synchronizes. */
int stwo_hip_barycentric_eval_at_point(stwo_hip_ctx *ctx, const uint32_t *const *evals, size_t n_cols,
                                       uint32_t log_size, const uint32_t *const weights[4], uint32_t *values);
/* PolyOps::eval_at_point_by_folding (ops.rs:82-88; CPUB/circle.rs:143-180): the
 * evaluation (2^log_size bit-reversed values on CanonicCoset(log_size), device)
 * folded circle -> line and line -> point with get_folding_alphas(p)
 * (core/poly/utils.rs:45-70), divided by 2^log_size; value: 4 words (host).
 * log_size 3..28; tw must cover the domain.  Synchronizes. */
int stwo_hip_eval_at_point_by_folding(stwo_hip_ctx *ctx, const uint32_t *evals, uint32_t log_size,
                                      const stwo_hip_secure_point *p, const stwo_hip_twiddles *tw,
                                      uint32_t value[4]);
/* PolyOps::split_at_mid (ops.rs:163; CPUB/circle.rs:283-291): the first and the
 * second half of 2^log_size coefficients into left and right (each 2^(log_size-1)
 * words; left may be coeffs itself), so that p(z) = left(z) + pi^(L-2)(z.x) right(z).
 * Asynchronous. */
int stwo_hip_split_at_mid(stwo_hip_ctx *ctx, const uint32_t *coeffs, uint32_t log_size, uint32_t *left,
                          uint32_t *right);
/* FriOps::decompose (STWO/prover/fri.rs:86-94; CPUB/fri.rs:41-63,141-156): for a
 * secure evaluation (4 coordinate columns of 2^log_size, bit-reversed on
 * CanonicCoset(log_size)), lambda = (sum of the first half - sum of the second
 * half) / 2^log_size (host, 4 words) and g = eval - lambda on the first half,
 * eval + lambda on the second (4 device columns; may be eval).  Synchronizes. */
int stwo_hip_fri_decompose(stwo_hip_ctx *ctx, const uint32_t *const eval[4], uint32_t log_size, uint32_t *const g[4],
                           uint32_t lambda[4]);

/* ---- wide-Fibonacci AIR (the prove benchmark's component) ----------------------
 * vendor/stwo-upstream-dev-62b228e/crates/examples/src/wide_fibonacci/mod.rs.
 * stwo_hip_wide_fib_trace: generate_trace (mod.rs:27-50) -- row i of the n_cols
 * columns (2^log_n words each, device) is the sequence a[i], b[i], then
 * x_k = x_{k-2}^2 + x_{k-1}^2; a and b are device arrays of 2^log_n values.
 * stwo_hip_wide_fib_constraint_quotients: the component's
 * evaluate_constraint_quotients_on_domain (constraint-framework/src/prover/
 * component_prover.rs:170-274) for a single component: `cols` are the trace
 * columns evaluated on CanonicCoset(log_n + 1) (bit-reversed, the trace
 * committed at blowup 1); accum (4 coordinate columns of 2^(log_n + 1)) is
 * incremented by the random-linear combination of the N - 2 constraints
 * (constraint j weighted alpha^(N - 3 - j)) divided by the trace domain's
 * vanishing polynomial. */
int stwo_hip_wide_fib_trace(stwo_hip_ctx *ctx, uint32_t *const *cols, uint32_t n_cols, uint32_t log_n,
                            const uint32_t *a, const uint32_t *b);
int stwo_hip_wide_fib_constraint_quotients(stwo_hip_ctx *ctx, const uint32_t *const *cols, uint32_t n_cols,
                                           uint32_t log_n, const uint32_t random_coeff[4], uint32_t *const accum[4]);

/* ---- end-to-end prove (wide-Fibonacci AIR) -------------------------------------
 * stwo's prove (STWO/prover/mod.rs:32-144) of one wide-Fibonacci component
 * with a Blake2sChannel and Blake2sMerkleChannel, preceded by the commits the
 * reference's benchmark makes before calling it (an empty preprocessed tree,
 * then the trace tree): the flow of run_one_sample in
 * fixtures/standalone-benchmarks/src/bin/wide_fibonacci_simd_prove.rs, and of
 * the Metal V1 prove runtime (MB/prove_runtime_v1.rs:1079 execute_prove_core_v1,
 * :1181 commit_trace_with_breakdown) that the reference benchmarks against it.
 *
 * The trace is generate_trace (examples wide_fibonacci/mod.rs:27-50) of the
 * device inputs a, b (2^log_n values each): n_cols columns of 2^log_n rows.
 * config is PcsConfig (STWO/core/pcs/mod.rs:27-40) with lifting_log_size None;
 * PcsConfig::default() is {10, {1, 0, 3, 1}}.  tw must cover
 * CanonicCoset(log_n + log_blowup_factor).  channel_digest / channel_n_draws
 * are the channel's state, read and updated (zeros = Blake2sChannel::default()).
 * Returns STWO_HIP_ECONSTRAINTS when prove_ex's own OODS self-check fails (a
 * trace that does not satisfy the AIR).  Synchronizes.  The prove's device
 * buffers (4 + 4 * 2^log_blowup_factor bytes per trace cell, plus the trees
 * and the composition and FRI columns) and its launch plans (the trace and
 * composition commits as hipGraphs) are kept by the context and reused by
 * its next prove of the same shape; stwo_hip_ctx_destroy releases them.
 *
 * stwo_hip_wide_fib_prove keeps the coefficients
 * (set_store_polynomials_coefficients, as the reference's wide-Fibonacci
 * benchmark does, MB/benchmark.rs:576) and samples them with eval_at_point.
 * stwo_hip_wide_fib_prove_ex with flags 0 is the PCS default
 * (store_polynomials_coefficients = false, STWO/prover/pcs/mod.rs:347): the
 * trace LDE does not store coefficients and the OODS samples come from the
 * committed evaluations with barycentric weights (:540-655); the proof bytes
 * are the same.  Without STWO_HIP_PROVE_STORE_COEFFS the log blowup must be 1
 * (a larger one extends the trace from its coefficients; the reference
 * panics there, constraint-framework/src/prover/component_prover.rs:56-77).
 *
 * stwo_hip_stark_proof_bytes writes the StarkProof<Blake2sMerkleHasher> in the
 * layout bincode 1's default options give the reference's serde derives
 * (fixed-width little-endian integers, u64 lengths and usize, Option as a
 * one-byte tag, hashes as their 32 bytes, QM31 as 4 M31 words), fields in
 * declaration order: CommitmentSchemeProof (STWO/core/pcs/quotients.rs:26-34),
 * FriProof / FriLayerProof (STWO/core/fri.rs:360-397), LinePoly
 * (STWO/core/poly/line.rs:113-123).  With buf NULL only *len is set.
 * stwo_hip_stark_proof_info returns host wall times of the prove's phases
 * (trace commit, composition commit, OODS samples, quotients, FRI commit,
 * proof of work, decommitment) and the sorted query positions. */
typedef struct {
  uint32_t pow_bits;
  stwo_hip_fri_config fri;
} stwo_hip_pcs_config;
#define STWO_HIP_PROVE_PHASES 7
typedef struct stwo_hip_stark_proof stwo_hip_stark_proof;
int stwo_hip_wide_fib_prove(stwo_hip_ctx *ctx, const uint32_t *a, const uint32_t *b, uint32_t n_cols, uint32_t log_n,
                            stwo_hip_pcs_config config, const stwo_hip_twiddles *tw, uint8_t channel_digest[32],
                            uint32_t *channel_n_draws, stwo_hip_stark_proof **out);
#define STWO_HIP_PROVE_STORE_COEFFS 1
int stwo_hip_wide_fib_prove_ex(stwo_hip_ctx *ctx, const uint32_t *a, const uint32_t *b, uint32_t n_cols,
                               uint32_t log_n, stwo_hip_pcs_config config, const stwo_hip_twiddles *tw, int flags,
                               uint8_t channel_digest[32], uint32_t *channel_n_draws, stwo_hip_stark_proof **out);
int stwo_hip_stark_proof_bytes(const stwo_hip_stark_proof *proof, uint8_t *buf, size_t cap, size_t *len);
int stwo_hip_stark_proof_info(const stwo_hip_stark_proof *proof, double phase_ms[STWO_HIP_PROVE_PHASES],
                              uint32_t *query_positions, size_t cap, size_t *n_positions);
int stwo_hip_stark_proof_destroy(stwo_hip_stark_proof *proof);

/* ---- proof of work -----------------------------------------------------------
 * GrindOps::grind for the Blake2s channel as SimdBackend computes it
 * (STWO/prover/backend/simd/grind.rs:22-62; Metal MB/blake2s.rs:399-443,
 * SH/grind.metal): the smallest nonce (hi << 32) | lo with lo < 2^20, hi =
 * 0, 1, ..., such that H(H(POW_PREFIX || 0^12 || digest || le32(pow_bits)) ||
 * le64(nonce)) has at least pow_bits trailing zero bits
 * (Blake2sChannel::verify_pow_nonce, STWO/core/channel/blake2s.rs:122-139).
 * pow_bits <= 32.  Does not mix the nonce into the channel.  Synchronizes. */
int stwo_hip_grind(stwo_hip_ctx *ctx, const uint8_t channel_digest[32], uint32_t pow_bits, uint64_t *nonce);

/* ---- SimdBackend layout interop --------------------------------------------
 * No single counterpart in the reference: conversions so that columns can
 * move between this library and a SimdBackend-driven prover.
 * stwo_hip_coeffs_simd_transpose: SimdBackend's IFFT leaves coefficients of
 * log size > 16 with their 16-element vectors transposed (transpose_vecs,
 * STWO/prover/backend/simd/fft/mod.rs:22-56, from simd/fft/ifft.rs:42-56;
 * simd/circle.rs:503-512 transposes back before splitting); this library
 * keeps natural order.  The map is its own inverse, in place, a no-op at
 * log size <= 16; `col` must be 16-byte aligned.
 * stwo_hip_twiddles_simd_doubled: SimdBackend's twiddle tree (doubled values
 * 2t, same layout, simd/circle.rs:449-486) from a tree of this library;
 * outputs hold twiddles_buffers' `len` words each.
 * stwo_hip_canonicalize: maps the redundant value P (which SimdBackend may
 * emit, simd/m31.rs:58) to 0, in place. */
int stwo_hip_coeffs_simd_transpose(stwo_hip_ctx *ctx, uint32_t *col, uint32_t log_size);
int stwo_hip_twiddles_simd_doubled(stwo_hip_ctx *ctx, const stwo_hip_twiddles *tw, uint32_t *out_twiddles,
                                   uint32_t *out_itwiddles);
int stwo_hip_canonicalize(stwo_hip_ctx *ctx, uint32_t *col, size_t n);

/* ---- batched commitment of uneven columns -------------------------------
 * One plan for a whole set of columns of different log sizes, replacing the
 * per-column dispatch of the reference's TreeBuilder path
 * (TreeBuilder::extend_evals -> interpolate_columns, STWO/prover/pcs/mod.rs:707-716;
 * CommitmentTreeProver::new -> evaluate_polynomials + commit, :766-803;
 * Metal MB/poly.rs:658-793 and MB/prove_runtime_v1.rs:1181-1240):
 *   cols[i]  (2^log_sizes[i] bit-reversed evaluations on CanonicCoset(log))
 *            -> interpolated in place to coefficients;
 *   outs[i]  <- evaluations on CanonicCoset(log + log_blowup);
 *   layers   <- packed tree (max layer = max log + log_blowup) of the outs:
 *               STWO_HIP_TREE_MIXED  = MerkleProver::commit (each size at its layer),
 *               STWO_HIP_TREE_LIFTED = MerkleProverLifted::commit at the max size.
 * Columns must have log size >= 4.  Per-call tables are uploaded once; the
 * plan's launches (one per FFT pass level for all sizes, one per tree layer)
 * are captured into a hipGraph when flags has STWO_HIP_BATCH_GRAPH.  Each
 * column's IFFT's last pass and RFFT's first pass run as one fused pass
 * where their tiles match (above 2^12 values); with STWO_HIP_BATCH_DROP_COEFFS
 * (store_polynomials_coefficients = false) that pass does not store the
 * coefficients and cols[i] is left unspecified.  stwo_hip_batch_run
 * enqueues the whole sequence on the context stream; the buffers named at
 * creation must stay valid for the plan's lifetime. */
#define STWO_HIP_BATCH_GRAPH 1
#define STWO_HIP_BATCH_DROP_COEFFS 2
#define STWO_HIP_TREE_MIXED 0
#define STWO_HIP_TREE_LIFTED 1
typedef struct stwo_hip_batch stwo_hip_batch;
int stwo_hip_batch_create(stwo_hip_ctx *ctx, uint32_t *const *cols, const uint32_t *log_sizes, size_t n_cols,
                          uint32_t log_blowup, const stwo_hip_twiddles *tw, uint32_t *const *outs, uint32_t *layers,
                          int tree_kind, int hash_mode, int flags, stwo_hip_batch **out);
/* Several column sets in one plan, each with its own blowup and tree (e.g. the
 * jobs a server admits together): columns of the same log size and blowup
 * from every set share each FFT pass launch; each set's tree is built on its
 * own.  Same rules per set as stwo_hip_batch_create.  stage: LDE (interpolate
 * in place, evaluate into outs, commit -- stwo_hip_batch_create's work),
 * IFFT (interpolate in place only: outs, layers and blowup unused) or RFFT
 * (cols are coefficients: evaluate into outs, then commit if layers is not
 * NULL).  Every IFFT pass of the plan runs before any RFFT pass, so an RFFT
 * set may read coefficients an IFFT set of the same plan produces (a
 * composition polynomial's interpolation, then its halves' evaluations). */
#define STWO_HIP_BATCH_LDE 0
#define STWO_HIP_BATCH_IFFT 1
#define STWO_HIP_BATCH_RFFT 2
typedef struct {
  uint32_t *const *cols;
  const uint32_t *log_sizes;
  size_t n_cols;
  uint32_t log_blowup;
  uint32_t *const *outs;
  uint32_t *layers;
  int tree_kind, hash_mode;
  int stage;
} stwo_hip_batch_set;
int stwo_hip_batch_create_multi(stwo_hip_ctx *ctx, const stwo_hip_batch_set *sets, size_t n_sets,
                                const stwo_hip_twiddles *tw, int flags, stwo_hip_batch **out);
int stwo_hip_batch_run(stwo_hip_ctx *ctx, stwo_hip_batch *batch);
/* The plain form the batched plan replaces: the same work dispatched column by
 * column, as the reference's evaluate_polynomials and interpolate_columns submit
 * one transform per column (MB/poly.rs:715-779, 364-391) -- interpolate then
 * evaluate each column (its own launches, tables and argument checks), then one
 * commit of the evaluations (mixed or lifted).  Same results as a batch plan. */
int stwo_hip_lde_commit_per_column(stwo_hip_ctx *ctx, uint32_t *const *cols, const uint32_t *log_sizes,
                                   size_t n_cols, uint32_t log_blowup, const stwo_hip_twiddles *tw,
                                   uint32_t *const *outs, uint32_t *layers, int tree_kind, int hash_mode);
int stwo_hip_batch_destroy(stwo_hip_ctx *ctx, stwo_hip_batch *batch);

/* ---- serving: continuous batching of commit + FRI and prove jobs ----------
 * No counterpart in the reference (its prover handles one proof at a time;
 * the request loop follows its caller's per-job body, MB/prove_runtime_v1.rs:
 * 1181-1240); this is the north star's serving mode.  A server owns `n_slots`
 * contexts (streams), a resident twiddle tree for domains up to
 * 2^max_log_size, and an HBM budget allocated once and sub-allocated per
 * admitted group of jobs.
 *
 * Admission: whenever a slot is idle, the queue's head job and the commit
 * jobs queued right behind it (up to max_group, default 16, in arrival
 * order) are admitted together if their memory fits the budget; a head that
 * does not fit waits for memory to be released (later jobs do not overtake
 * it).  A group's commit jobs share one batched plan: their columns' FFT
 * passes of equal size and blowup run in the same launches, each job's tree
 * on its own; then each job's FRI.  A prove job runs alone in its group.
 * Groups on different slots run concurrently on the device.  Every device
 * buffer a job needs -- its columns, evaluations, trees, the plan's tables,
 * FRI and prove scratch -- comes from its group's block of the budget, so
 * running jobs never allocate device memory and never exceed the budget.
 *
 * STWO_HIP_JOB_COMMIT_FRI: n_cols columns of 2^log_size evaluations (the
 * caller's device columns `cols`, or host data `host` -- column after
 * column -- copied into the job's memory when it starts, or else column c =
 * stwo_hip_fill_random(seed, c)), LDE at log_blowup, lifted Blake2s commit
 * (trace_root), then FriProver::commit (fold steps <= 4) on the QM31 column
 * formed by the first four evaluations, with a channel that starts as
 * H(0^32 || trace_root).
 * STWO_HIP_JOB_PROVE: stwo_hip_wide_fib_prove of n_cols columns of
 * 2^log_size rows with `pcs`, fresh channel; the two first columns a, b from
 * cols[0], cols[1] (device), host (2 * 2^log_size words) or the seed; the
 * proof bytes go to proof_out (proof_cap bytes; proof_len is set, or
 * STWO_HIP_EINVAL if it does not fit).
 * Caller data (cols, host, proof_out) must stay valid until the job is done.
 * Times are milliseconds since the server was created. */
#define STWO_HIP_JOB_COMMIT_FRI 0
#define STWO_HIP_JOB_PROVE 1
typedef struct stwo_hip_server stwo_hip_server;
typedef struct {
  uint32_t kind;       /* STWO_HIP_JOB_COMMIT_FRI or STWO_HIP_JOB_PROVE */
  uint32_t n_cols;     /* commit: >= 4; prove: 3..4096 */
  uint32_t log_size;   /* trace log size, >= 4 */
  uint32_t log_blowup; /* commit: >= 1, log_size + log_blowup <= max_log_size; prove: pcs.fri.log_blowup_factor */
  uint64_t seed;
  const uint32_t *const *cols; /* optional device columns (see above) */
  const uint32_t *host;        /* optional host data (see above) */
  stwo_hip_fri_config fri;     /* commit: fri.log_blowup_factor == log_blowup, log_last_layer_degree_bound <= 4 */
  stwo_hip_pcs_config pcs;     /* prove */
  uint8_t *proof_out;          /* prove */
  size_t proof_cap;
} stwo_hip_job_desc;
typedef struct {
  int status; /* STWO_HIP_OK, or the error of the job */
  uint32_t trace_root[8];
  uint32_t n_fri_roots;
  uint32_t fri_roots[8 * 32];
  uint32_t last_layer_poly[4 * 16];
  uint8_t final_digest[32];
  size_t proof_len;    /* prove jobs */
  uint32_t group_size; /* jobs admitted and run together with this one (itself included) */
  double t_submit_ms, t_admit_ms, t_start_ms, t_done_ms;
} stwo_hip_job_result;
int stwo_hip_server_create(int device, size_t hbm_budget_bytes, uint32_t max_log_size, uint32_t n_slots,
                           stwo_hip_server **out);
/* Most jobs admitted as one group (1 = no co-batching; 1..64; default 16). */
int stwo_hip_server_set_max_group(stwo_hip_server *server, uint32_t max_group);
/* hold != 0: admit no job until called again with 0 (a caller's burst of
 * submissions is then grouped as one queue; destroying the server releases
 * the hold). */
int stwo_hip_server_hold(stwo_hip_server *server, int hold);
/* Bytes of the budget a job takes while it runs (its share of a group). */
int stwo_hip_server_job_bytes(stwo_hip_server *server, const stwo_hip_job_desc *desc, size_t *bytes);
/* Non-blocking.  STWO_HIP_EBUSY if the queue is full, STWO_HIP_ENOMEM if the
 * job can never fit the budget, STWO_HIP_EINVAL for a bad description. */
int stwo_hip_server_submit(stwo_hip_server *server, const stwo_hip_job_desc *desc, uint64_t *job_id);
/* STWO_HIP_OK and the result if done, STWO_HIP_EBUSY if not yet. */
int stwo_hip_server_poll(stwo_hip_server *server, uint64_t job_id, stwo_hip_job_result *out);
/* Blocks until the job is done, returns its result and forgets the job. */
int stwo_hip_server_wait(stwo_hip_server *server, uint64_t job_id, stwo_hip_job_result *out);
int stwo_hip_server_stats(stwo_hip_server *server, size_t *pool_used, size_t *pool_capacity, uint32_t *n_pending,
                          uint32_t *n_running);
/* Drains the queue (every submitted job runs), then releases everything. */
int stwo_hip_server_destroy(stwo_hip_server *server);

/* ---- introspection -----------------------------------------------------------
 * No counterpart in the reference: describes the FFT passes this library
 * runs for a transform (16 words per pass, see fft.hip).  Launches nothing. */
int stwo_hip_debug_fft_plan(uint32_t n, uint32_t N, int inverse, uint32_t *out, uint32_t max_passes,
                            uint32_t *n_passes);
/* The same for the planner choices (t0, smax, wg) of stwo_hip_fft_tune_set;
 * 0 = the default of each. */
int stwo_hip_debug_fft_plan_tuned(uint32_t n, uint32_t N, int inverse, uint32_t t0, uint32_t smax, uint32_t wg,
                                  uint32_t *out, uint32_t max_passes, uint32_t *n_passes);

/* ---- FFT autotuning --------------------------------------------------------------
 * No counterpart in the reference (whose Metal backend fixes its kernel
 * shapes): picks, per FFT shape of a column set, the fastest of the planner's
 * candidate plans by timing them on this device, and keeps the choice in the
 * context, where every later plan of that shape uses it (interpolate,
 * evaluate, extend_evals, batch_create).  For each distinct log size n of
 * log_sizes (n >= 4): interpolate of 2^n and evaluate of 2^n at log_blowup.
 * Candidates: low-pass tile 2^10..2^13 (default 2^12), strided passes of at
 * most 5..8 layers (default 8), the tile's own workgroup size or 512 lanes.
 * Each runs on as many synthetic columns as the set has of that size (at most
 * 2^26 words, in a context scratch buffer); the fastest of 3 timed runs after
 * a warm-up counts.  Another plan replaces the default only when it is at
 * least 5% faster (the default schedules batch into fewer launches).
 * report (optional): 8 words per tuned shape, up to
 * max_report: n, N, inverse, t0, smax, wg, best microseconds, the default
 * plan's microseconds; *n_report = shapes tuned.  Synchronizes. */
int stwo_hip_fft_autotune(stwo_hip_ctx *ctx, const uint32_t *log_sizes, size_t n_cols, uint32_t log_blowup,
                          const stwo_hip_twiddles *tw, uint32_t *report, size_t max_report, size_t *n_report);
/* Sets planner choices (6 words per entry: n, N, inverse, t0 10..13, smax
 * 5..8, wg 0 or 9), e.g. a table saved from an earlier stwo_hip_fft_autotune;
 * stwo_hip_fft_tune_clear returns every shape to the defaults. */
int stwo_hip_fft_tune_set(stwo_hip_ctx *ctx, const uint32_t *entries, size_t n_entries);
int stwo_hip_fft_tune_clear(stwo_hip_ctx *ctx);

/* ---- profiling ------------------------------------------------------------
 * When enabled, every kernel launch of the context is bracketed by HIP
 * events on the context's stream; stwo_hip_prof_read synchronizes and
 * reports, per kernel name, the number of launches and their summed device
 * time.  Names are reported through `name_buf` (NUL-separated). */
int stwo_hip_prof_enable(stwo_hip_ctx *ctx, int enable);
int stwo_hip_prof_reset(stwo_hip_ctx *ctx);
int stwo_hip_prof_read(stwo_hip_ctx *ctx, char *name_buf, size_t name_buf_len, uint64_t *counts, double *total_ms,
                       size_t max_entries, size_t *n_entries);

#ifdef __cplusplus
}
#endif
#endif /* STWO_HIP_H */

"""CPU restatement of the end-to-end wide-Fibonacci prove, of stwo's verifier
for it, and of the proof's byte layout.

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker and by bench.py's
CPU-baseline leg; the product never imports it.  The prover chains the C
oracle's pieces (oracle.py: FFT, lifted Merkle, constraint quotients,
eval_at_point, FRI quotients, FRI commit, grind), each pinned by its own
tests, in the channel order of the reference; the verifier is restated from
the reference's verifier code in plain Python integers (its field arithmetic
does not go through the C oracle), so a mistake shared by the prover and the
verifier would have to be made twice.

Restated from (STWO = vendor/stwo-upstream-dev-62b228e/crates/stwo/src):
  run_one_sample      fixtures/standalone-benchmarks/src/bin/wide_fibonacci_simd_prove.rs:118-166
  prove_ex            STWO/prover/mod.rs:41-144
  CommitmentSchemeProver::commit / prepare_prove_values / prepare_finish / finish
                      STWO/prover/pcs/mod.rs:368-380,528-698,87-316
  DomainEvaluationAccumulator::finalize   STWO/prover/air/accumulation.rs:96-120
  split_at_mid        STWO/prover/backend/cpu/circle.rs:283-291
  get_random_point    STWO/core/circle.rs:169-181
  verify_ex           STWO/core/verifier.rs:41-113
  CommitmentSchemeVerifier::commit / verify_values   STWO/core/pcs/verifier.rs:38-130
  fri_answers, accumulate_row_quotients, denominator_inverses,
  build_samples_with_randomness_and_periodicity      STWO/core/pcs/quotients.rs:98-336
  complex_conjugate_line_coeffs           STWO/core/constraints.rs:119-134
  coset_vanishing                         STWO/core/constraints.rs:12-36
  verify_pow_nonce                        STWO/core/channel/blake2s.rs:124-140
  extract_composition_oods_eval           STWO/core/proof.rs:26-56
  WideFibonacciEval::evaluate             examples/src/wide_fibonacci/mod.rs:88-104
  PointEvaluator / PointEvaluationAccumulator
                      constraint-framework/src/point.rs:20-63, STWO/core/air/accumulation.rs
Byte layout: the reference's serde derives (CommitmentSchemeProof,
STWO/core/pcs/quotients.rs:26-34; FriProof / FriLayerProof, STWO/core/fri.rs:360-397;
LinePoly, STWO/core/poly/line.rs:113-123; PcsConfig / FriConfig,
STWO/core/pcs/mod.rs:27-40, STWO/core/fri.rs:31-36) under bincode 1's default
options (fixed-width little-endian integers, u64 lengths and usize, Option as
a one-byte tag).
"""
import hashlib
import struct
from collections import OrderedDict

import numpy as np

import decommit as D
import oracle as O

P = O.P
IDX_MASK = (1 << 31) - 1
POW_PREFIX = 0x12345678


# --- secure field in plain integers (verifier side) ------------------------------

def cm_mul(a, b):
    return ((a[0] * b[0] - a[1] * b[1]) % P, (a[0] * b[1] + a[1] * b[0]) % P)


def cm_inv(a):
    n = pow((a[0] * a[0] + a[1] * a[1]) % P, P - 2, P)
    return (a[0] * n % P, (-a[1]) * n % P)


def q(v):
    return tuple(int(x) % P for x in v)


def q_add(x, y):
    return tuple((a + b) % P for a, b in zip(x, y))


def q_sub(x, y):
    return tuple((a - b) % P for a, b in zip(x, y))


def q_mul(x, y):
    """(a + bu)(c + du) = ac + (2 + i) bd + (ad + bc) u  (STWO/core/fields/qm31.rs)."""
    a, b, c, d = (x[0], x[1]), (x[2], x[3]), (y[0], y[1]), (y[2], y[3])
    ac, bd = cm_mul(a, c), cm_mul(b, d)
    rbd = cm_mul((2, 1), bd)
    ad, bc = cm_mul(a, d), cm_mul(b, c)
    return ((ac[0] + rbd[0]) % P, (ac[1] + rbd[1]) % P, (ad[0] + bc[0]) % P, (ad[1] + bc[1]) % P)


def q_inv(x):
    """(a + bu)^-1 = (a - bu) / (a^2 - (2 + i) b^2)."""
    a, b = (x[0], x[1]), (x[2], x[3])
    den = cm_mul(a, a)
    t = cm_mul((2, 1), cm_mul(b, b))
    di = cm_inv(((den[0] - t[0]) % P, (den[1] - t[1]) % P))
    ra = cm_mul(a, di)
    rb = cm_mul(((-b[0]) % P, (-b[1]) % P), di)
    return (ra[0], ra[1], rb[0], rb[1])


def q_m31(v):
    return (int(v) % P, 0, 0, 0)


def q_conj(x):
    """QM31 complex conjugate: (a, b) -> (a, -b) (STWO/core/fields/mod.rs:413-415)."""
    return (x[0], x[1], (-x[2]) % P, (-x[3]) % P)


ONE = (1, 0, 0, 0)
ZERO = (0, 0, 0, 0)


def pt_add(p1, p2):
    """Circle point addition (x1 x2 - y1 y2, x1 y2 + y1 x2) over QM31."""
    (x1, y1), (x2, y2) = p1, p2
    return (q_sub(q_mul(x1, x2), q_mul(y1, y2)), q_add(q_mul(x1, y2), q_mul(y1, x2)))


def pt_m31(idx):
    x, y = O.index_to_point(idx & IDX_MASK)
    return (q_m31(x), q_m31(y))


def double_x(x):
    return q_sub(q_add(q_mul(x, x), q_mul(x, x)), ONE)


def random_point(channel):
    """CirclePoint::<SecureField>::get_random_point."""
    t = q(channel.draw_secure_felt())
    t2 = q_mul(t, t)
    inv = q_inv(q_add(t2, ONE))
    return (q_mul(q_sub(ONE, t2), inv), q_mul(q_add(t, t), inv))


def from_partial_evals(e):
    r = q(e[0])
    r = q_add(r, q_mul(q(e[1]), (0, 1, 0, 0)))
    r = q_add(r, q_mul(q(e[2]), (0, 0, 1, 0)))
    return q_add(r, q_mul(q(e[3]), (0, 0, 0, 1)))


def coset_vanishing(log_size, p):
    """coset_vanishing(CanonicCoset(log_size).coset, p): shift by -initial + step/2,
    then log_size - 1 doublings of x."""
    initial = 1 << (30 - log_size)      # Coset::odds(log_size).initial_index
    half_step = 1 << (30 - log_size)    # step_size.half()
    p = pt_add(p, pt_m31(-initial))
    p = pt_add(p, pt_m31(half_step))
    x = p[0]
    for _ in range(1, log_size):
        x = double_x(x)
    return x


# --- wide-Fibonacci constraints at a point ---------------------------------------

def eval_composition_at_point(point, trace_mask, random_coeff, max_log_degree_bound):
    """Components::eval_composition_polynomial_at_point for one wide-Fibonacci
    component: acc = acc * alpha + c_j / v(point) over constraints
    c_j = x_{j+2} - (x_j^2 + x_{j+1}^2)."""
    den_inv = q_inv(coset_vanishing(max_log_degree_bound, point))
    acc = ZERO
    a, b = q(trace_mask[0][0]), q(trace_mask[1][0])
    for col in trace_mask[2:]:
        c = q(col[0])
        con = q_sub(c, q_add(q_mul(a, a), q_mul(b, b)))
        acc = q_add(q_mul(acc, random_coeff), q_mul(den_inv, con))
        a, b = b, c
    return acc


def extract_composition_oods_eval(sampled_values, point, max_log_degree_bound):
    last = sampled_values[-1]
    if len(last) != 8 or any(len(c) != 1 for c in last):
        return None
    e = [c[0] for c in last]
    x = point[0]
    for _ in range(max_log_degree_bound - 1):
        x = double_x(x)
    return q_add(from_partial_evals(e[:4]), q_mul(x, from_partial_evals(e[4:])))


# --- proof of work -------------------------------------------------------------

def pow_ok(digest, n_bits, nonce):
    prefixed = hashlib.blake2s(struct.pack("<I", POW_PREFIX) + bytes(12) + bytes(digest) +
                               struct.pack("<I", n_bits)).digest()
    res = hashlib.blake2s(prefixed + struct.pack("<Q", nonce)).digest()
    v = int.from_bytes(res[:16], "little")
    if v == 0:
        return True
    return (v & -v).bit_length() - 1 >= n_bits


# --- FRI answers (verifier side) -------------------------------------------------

def fri_answers(column_log_sizes, samples, random_coeff, queries, queried_values, lifting_log_size):
    """fri_answers: the quotient combination at each query position from the
    queried values.  column_log_sizes / samples / queried_values are per tree,
    per column; samples are (point, value)."""
    pw = [ONE]

    def next_pow():
        cur = pw[0]
        pw[0] = q_mul(cur, random_coeff)
        return cur

    step = 1 << (31 - lifting_log_size)  # CanonicCoset(lifting).step() as an index
    flat = []
    for tree_samples, tree_sizes in zip(samples, column_log_sizes):
        for col_samples, log_size in zip(tree_samples, tree_sizes):
            out = []
            if col_samples:
                if len(col_samples) == 2:
                    pt, val = col_samples[1]
                    out.append(((pt_add(pt, pt_m31(step << log_size)), val), next_pow()))
                for s in col_samples:
                    out.append((s, next_pow()))
            flat.append(out)
    batches = OrderedDict()
    for ci, col in enumerate(flat):
        for (pt, val), rp in col:
            key = (tuple(pt[0]), tuple(pt[1]))
            batches.setdefault(key, []).append((ci, q(val), rp))
    coeffs = {}
    for key, entries in batches.items():
        py = key[1]
        lc = []
        for ci, val, rp in entries:
            a = q_sub(q_conj(val), val)
            c = q_sub(q_conj(py), py)
            b = q_sub(q_mul(val, c), q_mul(a, py))
            lc.append((ci, q_mul(rp, a), q_mul(rp, b), q_mul(rp, c)))
        coeffs[key] = lc
    flat_vals = [col for tree in queried_values for col in tree]
    half_init, half_log = O.half_odds_initial(lifting_log_size - 1), lifting_log_size - 1
    res = []
    for qi, pos in enumerate(queries):
        idx = D.circle_domain_index_at(half_init, half_log, D.bit_reverse_index(pos, lifting_log_size))
        dx, dy = O.index_to_point(idx)
        row = ZERO
        for key, lc in coeffs.items():
            px, py = key
            den = cm_mul(((px[0] - dx) % P, px[1]), (py[2], py[3]))
            den2 = cm_mul(((py[0] - dy) % P, py[1]), (px[2], px[3]))
            di = cm_inv(((den[0] - den2[0]) % P, (den[1] - den2[1]) % P))
            num = ZERO
            for ci, a, b, c in lc:
                value = q_mul(c, q_m31(flat_vals[ci][qi]))
                num = q_add(num, q_sub(value, q_add(q_mul(a, q_m31(dy)), b)))
            lo, hi = cm_mul((num[0], num[1]), di), cm_mul((num[2], num[3]), di)
            row = q_add(row, (lo[0], lo[1], hi[0], hi[1]))
        res.append(row)
    return res


# --- the prover --------------------------------------------------------------------

class PcsConfig:
    """PcsConfig with lifting_log_size None; default() is {10, FriConfig::new(0, 1, 3, 1)}."""

    def __init__(self, pow_bits=10, log_blowup_factor=1, log_last_layer_degree_bound=0, n_queries=3,
                 line_fold_step=1):
        self.pow_bits = pow_bits
        self.fri = O.FriConfig(log_blowup_factor, log_last_layer_degree_bound, n_queries, line_fold_step)


def words(b):
    return tuple(struct.unpack("<8I", b))


def prepare_query_positions_for_tree(queries, max_log_size, tree_log_size):
    if tree_log_size == 0:
        return []
    if max_log_size < tree_log_size:
        return [(p >> 1 << (tree_log_size - max_log_size + 1)) + (p & 1) for p in queries]
    return [(p >> (max_log_size - tree_log_size + 1) << 1) + (p & 1) for p in queries]


def wide_fib_prove(a, b, n_cols, log_n, config=None, channel=None):
    """The benchmark's flow; returns the proof as a dict (see proof_bytes for its fields)."""
    cfg = config or PcsConfig()
    B = cfg.fri.log_blowup_factor
    E, n = log_n + B, 1 << log_n
    tw, itw = O.precompute_twiddles(O.half_odds_initial(E - 1), E - 1)
    ch = channel if channel is not None else O.Channel()
    # preprocessed tree (no columns) and trace tree
    root0 = words(hashlib.blake2s(b"").digest())
    ch.mix_root(list(root0))
    coef = O.wide_fib_trace(a, b, n_cols)
    O.interpolate_columns(coef, itw)
    evals = O.evaluate_polynomials(coef, B, tw)
    tree1 = O.lifted_commit(evals, E)
    root1 = tuple(int(x) for x in tree1[:8])
    ch.mix_root(list(root1))
    # composition polynomial
    comp_coeff = q(ch.draw_secure_felt())
    ext = evals if B == 1 else O.evaluate_polynomials(coef, 1, tw)
    accum = O.wide_fib_constraint_quotients(ext, log_n, list(comp_coeff))
    O.interpolate_columns(accum, itw)
    halves = [accum[k][:n].copy() for k in range(4)] + [accum[k][n:].copy() for k in range(4)]
    cevals = O.evaluate_polynomials(halves, B, tw)
    tree2 = O.lifted_commit(cevals, E)
    root2 = tuple(int(x) for x in tree2[:8])
    ch.mix_root(list(root2))
    # OODS samples
    oods = random_point(ch)
    pnt = (list(oods[0]), list(oods[1]))
    values = [O.eval_at_point(c, pnt) for c in list(coef) + halves]
    ch.mix_felts([list(v) for v in values])
    sampled = [[], [[v] for v in values[:n_cols]], [[v] for v in values[n_cols:]]]
    max_bound = E - B
    if extract_composition_oods_eval(sampled, oods, max_bound) != eval_composition_at_point(
            oods, sampled[1], comp_coeff, max_bound):
        raise ValueError("ConstraintsNotSatisfied")
    # FRI quotients, FRI commit, proof of work, decommitment
    quot_coeff = ch.draw_secure_felt()
    samples = [[(pnt[0], pnt[1], list(v))] for v in values]
    quot = O.compute_fri_quotients(list(evals) + list(cevals), samples, quot_coeff, E)
    layers, roots, last_poly = D.fri_layers(cfg.fri, quot, E, ch)
    nonce = O.grind(ch.digest_bytes(), cfg.pow_bits)
    ch.mix_u64(nonce)
    queries, raw, fri_proofs = D.fri_decommit(ch, cfg.fri, layers)
    v1, w1, _ = D.merkle_lifted_decommit(tree1, E, evals, prepare_query_positions_for_tree(queries, E, E))
    v2, w2, _ = D.merkle_lifted_decommit(tree2, E, cevals, prepare_query_positions_for_tree(queries, E, E))

    def layer(lp):
        return dict(fri_witness=[tuple(int(x) for x in v) for v in lp["fri_witness"]],
                    hash_witness=[tuple(h) for h in lp["hash_witness"]],
                    commitment=tuple(int(x) for x in lp["commitment"]))

    return dict(
        config=dict(pow_bits=cfg.pow_bits, log_blowup_factor=B,
                    log_last_layer_degree_bound=cfg.fri.log_last_layer_degree_bound,
                    n_queries=cfg.fri.n_queries, line_fold_step=cfg.fri.line_fold_step, lifting_log_size=None),
        commitments=[root0, root1, root2],
        sampled_values=sampled,
        decommitments=[[], [tuple(h) for h in w1], [tuple(h) for h in w2]],
        queried_values=[[], v1, v2],
        proof_of_work=nonce,
        fri=dict(first=layer(fri_proofs[0]), inner=[layer(lp) for lp in fri_proofs[1:]],
                 last_layer_poly=[tuple(int(x) for x in c) for c in last_poly],
                 last_log=cfg.fri.log_last_layer_degree_bound),
        query_positions=list(queries),
    )


# --- the verifier --------------------------------------------------------------------

class VerificationError(Exception):
    pass


def verify(proof, n_cols, log_n, channel=None):
    """The reference benchmark's verification: commit the two trace trees with
    the component's sizes, then verify_ex.  Raises VerificationError."""
    cfg = proof["config"]
    B = cfg["log_blowup_factor"]
    ch = channel if channel is not None else O.Channel()
    if len(proof["commitments"]) != 3:
        raise VerificationError("InvalidStructure: commitments")
    trees = []

    def commit(root, log_sizes):
        ch.mix_root(list(root))
        ext = [s + B for s in log_sizes]
        trees.append((tuple(root), ext, max(ext, default=0)))

    commit(proof["commitments"][0], [])
    commit(proof["commitments"][1], [log_n] * n_cols)
    # verify_ex
    split_bound = (log_n + 1) - 1
    lifting = split_bound + B
    max_bound = lifting - B
    comp_coeff = q(ch.draw_secure_felt())
    commit(proof["commitments"][2], [max_bound] * 8)
    oods = random_point(ch)
    sv = proof["sampled_values"]
    if len(sv) != 3 or len(sv[0]) != 0 or len(sv[1]) != n_cols or any(len(c) != 1 for c in sv[1]):
        raise VerificationError("InvalidStructure: sampled values")
    comp = extract_composition_oods_eval(sv, oods, max_bound)
    if comp is None:
        raise VerificationError("InvalidStructure: composition mask")
    if comp != eval_composition_at_point(oods, sv[1], comp_coeff, max_bound):
        raise VerificationError("OodsNotMatching")
    # verify_values
    ch.mix_felts([list(v) for tree in sv for col in tree for v in col])
    quot_coeff = q(ch.draw_secure_felt())
    lifting_log = trees[-1][2]
    samples = [[], [[(oods, v) for v in col] for col in sv[1]], [[(oods, v) for v in col] for col in sv[2]]]
    fri = proof["fri"]
    fcfg = O.FriConfig(B, cfg["log_last_layer_degree_bound"], cfg["n_queries"], cfg["line_fold_step"])

    def pow_check(channel_):
        if not pow_ok(channel_.digest_bytes(), cfg["pow_bits"], proof["proof_of_work"]):
            raise VerificationError("ProofOfWork")
        channel_.mix_u64(proof["proof_of_work"])

    def first_layer_evals(queries):
        for (root, ext, height), dec, qv in zip(trees, proof["decommitments"], proof["queried_values"]):
            tq = prepare_query_positions_for_tree(queries, lifting_log, height)
            if height == 0:
                if dec or any(len(c) for c in qv):
                    raise VerificationError("Merkle: empty tree with witness")
                continue
            if len(qv) != len(ext) or any(len(c) != len(tq) for c in qv):
                raise VerificationError("InvalidStructure: queried values")
            try:
                D.merkle_lifted_verify(root, ext, tq, qv, dec)
            except D.VerificationError as e:
                raise VerificationError("Merkle: %s" % e)
        sizes = [t[1] for t in trees]
        return fri_answers(sizes, samples, quot_coeff, queries, proof["queried_values"], lifting_log)

    layers = [fri["first"]] + list(fri["inner"])
    try:
        D.fri_verify(ch, fcfg, (layers, fri["last_layer_poly"]), lifting_log, first_layer_evals, pow_check)
    except D.VerificationError as e:
        raise VerificationError("Fri: %s" % e)


# --- byte layout -----------------------------------------------------------------------

def proof_bytes(proof):
    out = bytearray()

    def u8(v):
        out.extend(struct.pack("<B", v))

    def u32(v):
        out.extend(struct.pack("<I", v))

    def u64(v):
        out.extend(struct.pack("<Q", v))

    def ws(vals):
        for v in vals:
            u32(int(v))

    c = proof["config"]
    u32(c["pow_bits"])
    u32(c["log_blowup_factor"])
    u32(c["log_last_layer_degree_bound"])
    u64(c["n_queries"])
    u32(c["line_fold_step"])
    if c["lifting_log_size"] is None:
        u8(0)
    else:
        u8(1)
        u32(c["lifting_log_size"])
    u64(len(proof["commitments"]))
    for h in proof["commitments"]:
        ws(h)
    u64(len(proof["sampled_values"]))
    for tree in proof["sampled_values"]:
        u64(len(tree))
        for col in tree:
            u64(len(col))
            for v in col:
                ws(v)
    u64(len(proof["decommitments"]))
    for wit in proof["decommitments"]:
        u64(len(wit))
        for h in wit:
            ws(h)
    u64(len(proof["queried_values"]))
    for tree in proof["queried_values"]:
        u64(len(tree))
        for col in tree:
            u64(len(col))
            ws(col)
    u64(proof["proof_of_work"])

    def layer(lp):
        u64(len(lp["fri_witness"]))
        for v in lp["fri_witness"]:
            ws(v)
        u64(len(lp["hash_witness"]))
        for h in lp["hash_witness"]:
            ws(h)
        ws(lp["commitment"])

    f = proof["fri"]
    layer(f["first"])
    u64(len(f["inner"]))
    for lp in f["inner"]:
        layer(lp)
    u64(len(f["last_layer_poly"]))
    for v in f["last_layer_poly"]:
        ws(v)
    u32(f["last_log"])
    return bytes(out)


def parse_proof(data):
    """Inverse of proof_bytes."""
    off = [0]

    def take(fmt):
        v = struct.unpack_from(fmt, data, off[0])
        off[0] += struct.calcsize(fmt)
        return v

    def u8():
        return take("<B")[0]

    def u32():
        return take("<I")[0]

    def u64():
        return take("<Q")[0]

    def tup(k):
        return tuple(take("<%dI" % k))

    c = dict(pow_bits=u32(), log_blowup_factor=u32(), log_last_layer_degree_bound=u32(), n_queries=u64(),
             line_fold_step=u32())
    c["lifting_log_size"] = u32() if u8() else None
    commitments = [tup(8) for _ in range(u64())]
    sampled = [[[tup(4) for _ in range(u64())] for _ in range(u64())] for _ in range(u64())]
    decs = [[tup(8) for _ in range(u64())] for _ in range(u64())]
    queried = []
    for _ in range(u64()):
        tree = []
        for _ in range(u64()):
            nv = u64()
            tree.append(list(tup(nv)))
        queried.append(tree)
    pow_ = u64()

    def layer():
        fw = [tup(4) for _ in range(u64())]
        hw = [tup(8) for _ in range(u64())]
        return dict(fri_witness=fw, hash_witness=hw, commitment=tup(8))

    first = layer()
    inner = [layer() for _ in range(u64())]
    last = [tup(4) for _ in range(u64())]
    last_log = u32()
    if off[0] != len(data):
        raise ValueError("trailing bytes in proof")
    return dict(config=c, commitments=commitments, sampled_values=sampled, decommitments=decs,
                queried_values=queried, proof_of_work=pow_,
                fri=dict(first=first, inner=inner, last_layer_poly=last, last_log=last_log))

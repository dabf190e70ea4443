"""CPU restatement of the decommitment side of the path: lifted Merkle
decommitment, FRI query drawing and FRI decommitment, plus restatements of
the reference's verifiers for both.

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker; the product
never imports it.  Pure Python over small inputs (a few hundred queries at
most); field products and circle points go through the C oracle
(oracle.py), which its own tests pin.

Hashes are tuples of 8 u32 words (32 little-endian bytes).  QM31 values are
tuples of 4 limbs [a.re, a.im, b.re, b.im].

Restated from:
  MerkleProverLifted::decommit      STWO/prover/vcs_lifted/prover.rs:58-201
  MerkleVerifierLifted::verify      STWO/core/vcs_lifted/verifier.rs (verify)
  draw_queries / Queries            STWO/core/queries.rs:11-63
  FriProver::decommit, layer provers,
  compute_decommitment_positions_and_witness_evals
                                    STWO/prover/fri.rs:294-489
  FriVerifier (commit, decommit, layer verifiers, SparseEvaluation,
  fold_coset, fold_circle_into_line) STWO/core/fri.rs:82-790
  LinePoly::eval_at_point, fold     STWO/core/poly/line.rs:138-145,
                                    STWO/core/poly/utils.rs:27-38
"""
import hashlib
import struct

import oracle as O

P = O.P
IDX_MASK = (1 << 31) - 1


# --- small helpers -----------------------------------------------------------

def lift(pos, shift):
    """lift(i, r) = ((i >> (r + 1)) << 1) + (i & 1) (CPUB/merkle_lifted.rs:55-57)."""
    return ((pos >> (shift + 1)) << 1) + (pos & 1)


def dedup(xs):
    """Itertools dedup: drop consecutive repeats."""
    out = []
    for x in xs:
        if not out or out[-1] != x:
            out.append(x)
    return out


def chunk_by(xs, same):
    """slice::chunk_by: split where `same(prev, next)` is false."""
    out = []
    for x in xs:
        if out and same(out[-1][-1], x):
            out[-1].append(x)
        else:
            out.append([x])
    return out


def bit_reverse_index(i, log_size):
    if log_size == 0:
        return i
    return int(format(i, "0%db" % log_size)[::-1], 2)


def hash_bytes(h):
    return struct.pack("<8I", *h)


def blake2s_words(data):
    return struct.unpack("<8I", hashlib.blake2s(data).digest())


def hash_children(l, r):
    """Blake2sMerkleHasher::hash_children: H(left || right)."""
    return blake2s_words(hash_bytes(l) + hash_bytes(r))


def hash_leaf(row):
    """Blake2sMerkleHasher leaf: H(le32 values, columns ascending by size)."""
    return blake2s_words(b"".join(struct.pack("<I", int(v)) for v in row))


def tree_node(tree, layer, i):
    o = O.layer_offset(layer) + 8 * i
    return tuple(int(x) for x in tree[o: o + 8])


def qadd(a, b):
    return tuple((x + y) % P for x, y in zip(a, b))


def qsub(a, b):
    return tuple((x - y) % P for x, y in zip(a, b))


def qmul(a, b):
    return tuple(O.qm31_mul(list(a), list(b)))


def qmul_m31(a, m):
    return tuple(x * m % P for x in a)


def ibutterfly(v0, v1, itw):
    """(v0 + v1, (v0 - v1) * itw), STWO/core/fft.rs:14-21."""
    return qadd(v0, v1), qmul_m31(qsub(v0, v1), itw)


def m31_inv(a):
    return pow(a, P - 2, P)


def coset_index_at(initial, log_size, i):
    """Coset::index_at: initial + step * i, step = 2^(31 - log) (circle.rs:295-305,384-386)."""
    step = (1 << (31 - log_size)) & IDX_MASK
    return (initial + step * i) & IDX_MASK


def circle_domain_index_at(half_initial, half_log, i):
    """CircleDomain::index_at (poly/circle/domain.rs:62-68)."""
    half = 1 << half_log
    if i < half:
        return coset_index_at(half_initial, half_log, i)
    return (-coset_index_at(half_initial, half_log, i - half)) & IDX_MASK


def point(idx):
    return O.index_to_point(idx)


# --- lifted Merkle -------------------------------------------------------------

def merkle_lifted_decommit(tree, L, cols, queries):
    """MerkleProverLifted::decommit: (queried_values per column, hash_witness,
    aux all_node_values as a list of {index: hash}, leaf layer first)."""
    values = []
    for c in cols:
        shift = L - (len(c).bit_length() - 1)
        values.append([int(c[lift(p, shift)]) for p in queries])
    witness = []
    aux = []
    prev = dedup(list(queries))
    for layer_log_size in range(L - 1, -1, -1):
        child = layer_log_size + 1
        nodes = {}
        cur = []
        for chunk in chunk_by(prev, lambda a, b: a ^ 1 == b):
            first = chunk[0]
            parent = first >> 1
            if len(chunk) == 1:
                witness.append(tree_node(tree, child, first ^ 1))
            nodes[2 * parent] = tree_node(tree, child, 2 * parent)
            nodes[2 * parent + 1] = tree_node(tree, child, 2 * parent + 1)
            cur.append(parent)
        aux.append(nodes)
        prev = cur
    return values, witness, aux


class VerificationError(Exception):
    pass


def merkle_lifted_verify(root, column_log_sizes, queries, values, witness, lifting_log_size=None):
    """MerkleVerifierLifted::verify; raises VerificationError("WitnessTooShort" |
    "WitnessTooLong" | "RootMismatch")."""
    max_log = max(column_log_sizes, default=0)
    height = max_log if lifting_log_size is None else lifting_log_size
    assert max_log <= height
    if height == 0:
        return
    for i in range(len(queries) - 1):
        if queries[i] == queries[i + 1]:
            for col in values:
                assert col[i] == col[i + 1]
    order = sorted(range(len(values)), key=lambda c: column_log_sizes[c])
    dedup_idx = [i for i in range(len(queries)) if i == 0 or queries[i] != queries[i - 1]]
    prev = []
    for k, pos in enumerate(dedup(list(queries))):
        row = [values[c][dedup_idx[k]] for c in order]
        prev.append((pos, hash_leaf(row)))
    wit = list(witness)
    wi = 0
    for _ in range(height):
        cur = []
        for chunk in chunk_by(prev, lambda a, b: a[0] ^ 1 == b[0]):
            idx0, h0 = chunk[0]
            if len(chunk) == 1:
                if wi >= len(wit):
                    raise VerificationError("WitnessTooShort")
                w = wit[wi]
                wi += 1
                children = (h0, w) if idx0 & 1 == 0 else (w, h0)
            else:
                children = (h0, chunk[1][1])
            cur.append((idx0 >> 1, hash_children(*children)))
        prev = cur
    if wi != len(wit):
        raise VerificationError("WitnessTooLong")
    assert len(prev) == 1
    if prev[0][1] != tuple(root):
        raise VerificationError("RootMismatch")


# --- queries -----------------------------------------------------------------

def draw_queries(channel, log_domain_size, n_queries):
    raw = []
    mask = (1 << log_domain_size) - 1
    while True:
        for w in channel.draw_u32s():
            raw.append(w & mask)
            if len(raw) == n_queries:
                return raw


def queries_new(raw):
    return sorted(set(raw))


def queries_fold(queries, n_folds):
    return dedup([q >> n_folds for q in queries])


# --- FRI prover side -----------------------------------------------------------

def fri_schedule(cfg, log_size):
    """[(log size, fold step)] of the committed layers (fri.rs:188-262)."""
    last_log = cfg.log_last_layer_degree_bound + cfg.log_blowup_factor
    out = [(log_size, 1)]
    lg = log_size - 1
    if lg != last_log:
        while True:
            last = not (lg > last_log + cfg.line_fold_step)
            step = lg - last_log if last else cfg.line_fold_step
            out.append((lg, step))
            lg -= step
            if last:
                break
    return out


def fri_layers(cfg, col4, log_size, channel):
    """Commits with the C oracle and returns, per committed layer,
    (col4, log size, fold step, packed tree), plus roots and last layer poly."""
    roots, last_poly, bufs = O.fri_commit(channel, cfg, col4, log_size, keep_layers=True)
    sched = fri_schedule(cfg, log_size)
    layers = []
    for k, (lg, step) in enumerate(sched):
        if k == 0:
            c4 = [c for c in col4]
        else:
            n = 1 << lg
            c4 = [bufs[k - 1][i * n:(i + 1) * n].copy() for i in range(4)]
        tree = O.lifted_commit(c4, lg)
        assert tuple(int(x) for x in tree[:8]) == tuple(int(x) for x in roots[k])
        layers.append((c4, lg, step, tree))
    return layers, roots, last_poly


def decommitment_positions_and_witness(col4, queries, fold_step):
    """compute_decommitment_positions_and_witness_evals (fri.rs:455-489)."""
    positions, witness, values = [], [], []
    for chunk in chunk_by(queries, lambda a, b: a >> fold_step == b >> fold_step):
        start = (chunk[0] >> fold_step) << fold_step
        it = 0
        for pos in range(start, start + (1 << fold_step)):
            positions.append(pos)
            ev = tuple(int(col4[k][pos]) for k in range(4))
            values.append(ev)
            if it < len(chunk) and chunk[it] == pos:
                it += 1
            else:
                witness.append(ev)
    return positions, witness, values


def fri_decommit(channel, cfg, layers):
    """FriProver::decommit: (sorted queries, raw queries, per-layer proofs)."""
    raw = draw_queries(channel, layers[0][1], cfg.n_queries)
    queries = queries_new(raw)
    proofs = []
    lq = queries
    for c4, lg, step, tree in layers:
        positions, witness, values = decommitment_positions_and_witness(c4, lq, step)
        _, hash_witness, _ = merkle_lifted_decommit(tree, lg, c4, positions)
        proofs.append(dict(commitment=tree_node(tree, 0, 0), fri_witness=witness, hash_witness=hash_witness,
                           positions=positions, values=values))
        lq = queries_fold(lq, step)
    return queries, raw, proofs


# --- FRI verifier ---------------------------------------------------------------

def fold_coset(evals, initial, log_size, alpha):
    """fold_coset (core/fri.rs:720-742) on LineDomain(Coset(initial, log_size))."""
    ev = list(evals)
    n = log_size
    a = alpha
    init, lg = initial, log_size
    for i in range(n):
        for j in range(0, 1 << (n - i), 2):
            x = point(coset_index_at(init, lg, bit_reverse_index(j, lg)))[0]
            f0, f1 = ibutterfly(ev[j], ev[j + 1], m31_inv(x))
            ev[j >> 1] = qadd(f0, qmul(a, f1))
        a = qmul(a, a)
        init, lg = (init * 2) & IDX_MASK, lg - 1
    return ev[0]


def fold_circle_pair(f_p, f_neg_p, fold_initial, alpha):
    """fold_circle_into_line (core/fri.rs:744-772) of one pair on
    CircleDomain(Coset(fold_initial, 0)), accumulated into zero."""
    y = point(fold_initial)[1]
    f0, f1 = ibutterfly(f_p, f_neg_p, m31_inv(y))
    return qadd(qmul(alpha, f1), f0)


def line_poly_eval(coeffs, x):
    """LinePoly::eval_at_point with x in the base field (line.rs:138-145)."""
    log = len(coeffs).bit_length() - 1
    doublings = []
    for _ in range(log):
        doublings.append(x)
        x = (2 * x * x - 1) % P

    def fold(vals, factors):
        if len(vals) == 1:
            return tuple(vals[0])
        h = len(vals) // 2
        lhs = fold(vals[:h], factors[1:])
        rhs = fold(vals[h:], factors[1:])
        return qadd(lhs, qmul_m31(rhs, factors[0]))

    return fold([tuple(c) for c in coeffs], doublings)


def _rebuild(queries, query_evals, witness, fold_step, log_domain_size, err):
    """compute_decommitment_positions_and_rebuild_evals (core/fri.rs:592-628)."""
    qe = iter(query_evals)
    wi = [0]
    positions, subsets, initials = [], [], []
    for chunk in chunk_by(queries, lambda a, b: a >> fold_step == b >> fold_step):
        start = (chunk[0] >> fold_step) << fold_step
        positions.extend(range(start, start + (1 << fold_step)))
        it = 0
        sub = []
        for pos in range(start, start + (1 << fold_step)):
            if it < len(chunk) and chunk[it] == pos:
                it += 1
                sub.append(next(qe))
            else:
                if wi[0] >= len(witness):
                    raise VerificationError(err)
                sub.append(tuple(witness[wi[0]]))
                wi[0] += 1
        subsets.append(sub)
        initials.append(bit_reverse_index(start, log_domain_size))
    if wi[0] != len(witness):
        raise VerificationError(err)
    return positions, subsets, initials


def fri_verify(channel, cfg, proof, column_log_size, first_layer_query_evals_of, before_queries=None):
    """FriVerifier::commit + sample_query_positions + decommit.  `proof` is
    (layer proofs as from fri_decommit, last layer poly as [k][4]);
    `first_layer_query_evals_of(queries)` returns the committed column's
    QM31 values at the sorted query positions; `before_queries(channel)`, if
    given, runs between the commit phase and the query draw (the PCS
    verifier's proof-of-work check, STWO/core/pcs/verifier.rs:73-78).
    Raises VerificationError."""
    layer_proofs, last_poly = proof
    log_blowup = cfg.log_blowup_factor
    column_bound = column_log_size - log_blowup
    # commit phase (core/fri.rs:107-187)
    channel.mix_root(list(layer_proofs[0]["commitment"]))
    first_alpha = channel.draw_secure_felt()
    layer_bound = column_bound - 1
    dom_log = layer_bound + log_blowup
    dom_init = O.half_odds_initial(dom_log)
    inner = []
    n_inner = len(layer_proofs) - 1
    for li, lp in enumerate(layer_proofs[1:]):
        channel.mix_root(list(lp["commitment"]))
        if li != n_inner - 1:
            step = cfg.line_fold_step
        else:
            res = layer_bound - cfg.log_last_layer_degree_bound
            if not (1 <= res <= cfg.line_fold_step):
                raise VerificationError("InvalidNumFriLayers")
            step = res
        alpha = channel.draw_secure_felt()
        inner.append((dom_init, dom_log, alpha, lp, step))
        if layer_bound < step:
            raise VerificationError("InvalidNumFriLayers")
        layer_bound -= step
        dom_init, dom_log = (dom_init << step) & IDX_MASK, dom_log - step
    if layer_bound != cfg.log_last_layer_degree_bound:
        raise VerificationError("InvalidNumFriLayers")
    if len(last_poly) > (1 << cfg.log_last_layer_degree_bound):
        raise VerificationError("LastLayerDegreeInvalid")
    channel.mix_felts([list(c) for c in last_poly])
    if before_queries is not None:
        before_queries(channel)
    # queries (fri.rs:281-287)
    raw = draw_queries(channel, column_log_size, cfg.n_queries)
    queries = queries_new(raw)
    # first layer (fri.rs:438-493)
    fp = layer_proofs[0]
    positions, subsets, initials = _rebuild(queries, first_layer_query_evals_of(queries), fp["fri_witness"], 1,
                                            column_log_size, "FirstLayerEvaluationsInvalid")
    vals = [[e[k] for sub in subsets for e in sub] for k in range(4)]
    try:
        merkle_lifted_verify(fp["commitment"], [column_log_size] * 4, positions, vals, fp["hash_witness"])
    except VerificationError as e:
        raise VerificationError("FirstLayerCommitmentInvalid: %s" % e)
    half_init, half_log = O.half_odds_initial(column_log_size - 1), column_log_size - 1
    evals = [fold_circle_pair(sub[0], sub[1], circle_domain_index_at(half_init, half_log, ini), first_alpha)
             for sub, ini in zip(subsets, initials)]
    lq = queries_fold(queries, 1)
    # inner layers (fri.rs:518-590)
    for li, (d_init, d_log, alpha, lp, step) in enumerate(inner):
        positions, subsets, initials = _rebuild(lq, evals, lp["fri_witness"], step, d_log,
                                                "InnerLayerEvaluationsInvalid(%d)" % li)
        vals = [[e[k] for sub in subsets for e in sub] for k in range(4)]
        try:
            merkle_lifted_verify(lp["commitment"], [d_log] * 4, positions, vals, lp["hash_witness"])
        except VerificationError as e:
            raise VerificationError("InnerLayerCommitmentInvalid(%d): %s" % (li, e))
        evals = [fold_coset(sub, coset_index_at(d_init, d_log, ini), step, alpha)
                 for sub, ini in zip(subsets, initials)]
        lq = queries_fold(lq, step)
    # last layer (fri.rs:258-279)
    for q, ev in zip(lq, evals):
        x = point(coset_index_at(dom_init, dom_log, bit_reverse_index(q, dom_log)))[0]
        if tuple(ev) != line_poly_eval(last_poly, x):
            raise VerificationError("LastLayerEvaluationsInvalid")

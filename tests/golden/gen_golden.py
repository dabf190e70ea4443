#!/usr/bin/env python3
"""Writes the golden fixtures under tests/golden/ (pure Python ints + hashlib).

This generator is independent of both the C oracle and the HIP library: it
derives every value from the mathematical definitions the reference uses,
by methods that differ from the ones under test:

* circle points by scalar multiplication of M31_CIRCLE_GEN = (2, 1268011823)
  (STWO/core/circle.rs:202-205), not by the coset iterators the oracle uses;
* evaluations by direct evaluation of the circle polynomial in the FFT basis
  at every domain point, with the recursive `fold` of STWO/core/poly/utils.rs
  and the point mappings of CpuBackend::eval_at_point (CPUB/circle.rs:82-99)
  -- an O(N^2) algorithm, not an FFT;
* FRI folds from the even/odd split of the polynomial's coefficients
  (fold(f)(x) = 2 (f_even(x) + alpha f_odd(x))), not from ibutterfly;
* Blake2s from hashlib (RFC 7693), Merkle nodes from explicit byte streams
  (STWO/core/vcs/blake2_merkle.rs:17-38, CPUB/merkle_lifted.rs:44-85).

Paths: STWO = vendor/stwo-upstream-dev-62b228e/crates/stwo/src,
CPUB = STWO/prover/backend/cpu.

Run: python tests/golden/gen_golden.py   (a minute or so)
"""
import hashlib
import json
import os
import random

P = (1 << 31) - 1
HERE = os.path.dirname(os.path.abspath(__file__))

# ---------------------------------------------------------------- fields ---


def inv(a):
    return pow(a, P - 2, P)


def cm_mul(x, y):
    return ((x[0] * y[0] - x[1] * y[1]) % P, (x[0] * y[1] + x[1] * y[0]) % P)


def cm_add(x, y):
    return ((x[0] + y[0]) % P, (x[1] + y[1]) % P)


def qm_mul(x, y):
    a, b = (x[0], x[1]), (x[2], x[3])
    c, d = (y[0], y[1]), (y[2], y[3])
    r0 = cm_add(cm_mul(a, c), cm_mul((2, 1), cm_mul(b, d)))
    r1 = cm_add(cm_mul(a, d), cm_mul(b, c))
    return (r0[0], r0[1], r1[0], r1[1])


def qm_add(x, y):
    return tuple((a + b) % P for a, b in zip(x, y))


def qm_scale(x, m):
    return tuple((a * m) % P for a in x)


# ---------------------------------------------------------------- circle ---

GEN = (2, 1268011823)


def pt_add(p, q):
    return ((p[0] * q[0] - p[1] * q[1]) % P, (p[0] * q[1] + p[1] * q[0]) % P)


def to_point(idx):
    idx %= 1 << 31
    res, cur = (1, 0), GEN
    while idx:
        if idx & 1:
            res = pt_add(res, cur)
        cur = pt_add(cur, cur)
        idx >>= 1
    return res


def half_odds(n):
    return (1 << (29 - n)) % (1 << 31), n


def coset_index(initial, log_size, i):
    return (initial + (1 << (31 - log_size)) * i) % (1 << 31)


def domain_point(half_initial, half_log, i):
    """CircleDomain::at (STWO/core/poly/circle/domain.rs:56-68)."""
    h = 1 << half_log
    if i < h:
        return to_point(coset_index(half_initial, half_log, i))
    idx = coset_index(half_initial, half_log, i - h)
    return to_point((1 << 31) - idx)


def bitrev(i, n):
    return int(format(i, "0%db" % n)[::-1], 2) if n else i


def double_x(x):
    return (2 * x * x - 1) % P


# ---------------------------------------------------- polynomial evaluation -


def fold(values, factors):
    """STWO/core/poly/utils.rs:27-38: fold(lhs) + fold(rhs) * factors[0]."""
    n = len(values)
    if n == 1:
        return values[0]
    lhs = fold(values[: n // 2], factors[1:])
    rhs = fold(values[n // 2:], factors[1:])
    if isinstance(lhs, tuple):
        return qm_add(lhs, qm_mul(rhs, factors[0]) if isinstance(factors[0], tuple) else qm_scale(rhs, factors[0]))
    return (lhs + rhs * factors[0]) % P


def circle_eval(coeffs, pt):
    """CpuBackend::eval_at_point (CPUB/circle.rs:82-99) for M31 coefficients."""
    n = len(coeffs).bit_length() - 1
    if n == 0:
        return coeffs[0]
    x, y = pt
    maps = [y]
    for _ in range(1, n):
        maps.append(x)
        x = double_x(x)
    maps.reverse()
    return fold(coeffs, maps)


def line_eval_ordered(coeffs, x):
    """Line polynomial with ordered coefficients: basis element j is the product
    over set bits k of j of pi^k(x) (bit 0 <-> x, bit 1 <-> pi(x), ...)."""
    n = len(coeffs).bit_length() - 1
    pis = []
    for _ in range(n):
        pis.append(x)
        x = double_x(x)
    total = None
    for j, c in enumerate(coeffs):
        term = 1
        for k in range(n):
            if j >> k & 1:
                term = term * pis[k] % P
        v = qm_scale(c, term) if isinstance(c, tuple) else c * term % P
        total = v if total is None else (qm_add(total, v) if isinstance(v, tuple) else (total + v) % P)
    return total


# -------------------------------------------------------------- blake2s ----


def h(data: bytes) -> bytes:
    return hashlib.blake2s(data).digest()


def le32(vals):
    return b"".join(int(v).to_bytes(4, "little") for v in vals)


def words(digest: bytes):
    return [int.from_bytes(digest[i:i + 4], "little") for i in range(0, 32, 4)]


# ---------------------------------------------------------------- fixtures --


def gen_twiddles():
    """Twiddle trees (CPUB/circle.rs:294-311) for the reference's test cosets
    (crates/stwo-metal/tests/metal_twiddle_precompute_native.rs:17-49)."""
    out = []
    roots = [("canonic", *half_odds(l - 1)) for l in [1, 2, 3, 6, 10]]
    roots += [("half_odds", *half_odds(l)) for l in [4, 7, 9]]
    for kind, initial, log_size in roots:
        tw = []
        init, lg = initial, log_size
        for _ in range(log_size):
            xs = [to_point(coset_index(init, lg, i))[0] for i in range(1 << (lg - 1))]
            tw += [xs[bitrev(i, lg - 1)] for i in range(len(xs))]
            init, lg = init * 2 % (1 << 31), lg - 1
        tw.append(1)
        out.append({"kind": kind, "initial": initial, "log_size": log_size, "twiddles": tw,
                    "itwiddles": [inv(t) for t in tw]})
    return out


def gen_fft(rng):
    """Bit-reversed evaluations by brute force; coefficients random."""
    out = []
    cases = [(n, n, half_odds(n - 1)[0]) for n in range(1, 11)]          # canonic, blowup 0
    cases += [(n, n + b, half_odds(n + b - 1)[0]) for n, b in [(3, 1), (5, 2), (6, 1), (7, 3), (8, 1)]]
    # a domain that is not canonic: half coset initial index 3 * 2^(29 - n)
    cases += [(6, 6, 3 << (29 - 5)), (4, 5, 5 << (29 - 4))]
    for n_log, N, initial in cases:
        coeffs = [rng.randrange(P) for _ in range(1 << n_log)]
        padded = coeffs + [0] * ((1 << N) - (1 << n_log))
        ev = [circle_eval(padded, domain_point(initial, N - 1, bitrev(i, N))) for i in range(1 << N)]
        out.append({"coeff_log": n_log, "log_size": N, "domain_initial": initial, "coeffs": coeffs, "evals": ev})
    return out


def gen_merkle(rng):
    """Mixed-degree roots (STWO/prover/vcs/prover.rs:39-65) and lifted roots
    (STWO/prover/vcs_lifted/ops.rs:29-45) from explicit byte streams."""
    out = []
    shapes = [[4], [3, 3, 3], [5, 2, 5, 3, 2], [1], [4] * 17, [6, 6, 4, 3, 6, 2, 2, 1]]
    for shape in shapes:
        cols = [[rng.randrange(P) for _ in range(1 << s)] for s in shape]
        # mixed degree
        order = sorted(range(len(cols)), key=lambda c: -shape[c])
        max_log = max(shape)
        layers = {}
        prev = None
        for l in range(max_log, -1, -1):
            cl = [cols[c] for c in order if shape[c] == l]
            layer = []
            for i in range(1 << l):
                msg = b""
                if prev is not None:
                    msg += prev[2 * i] + prev[2 * i + 1]
                msg += le32([c[i] for c in cl])
                layer.append(h(msg))
            layers[l] = layer
            prev = layer
        mixed_root = words(layers[0][0])
        # lifted, at lifting size = max log
        if min(shape) >= 1:
            L = max_log
            asc = sorted(range(len(cols)), key=lambda c: shape[c])
            leaves = []
            for i in range(1 << L):
                vals = []
                for c in asc:
                    r = L - shape[c]
                    vals.append(cols[c][((i >> (r + 1)) << 1) + (i & 1)])
                leaves.append(h(le32(vals)))
            layer = leaves
            while len(layer) > 1:
                layer = [h(layer[2 * i] + layer[2 * i + 1]) for i in range(len(layer) // 2)]
            lifted_root = words(layer[0])
        else:
            lifted_root = None
        out.append({"shape": shape, "cols": cols, "mixed_root": mixed_root,
                    "mixed_leaf_layer": [words(x) for x in layers[max_log]], "lifted_root": lifted_root})
    return out


def rand_qm(rng):
    return tuple(rng.randrange(P) for _ in range(4))


def gen_folds(rng):
    """fold_circle_into_line / fold_line by the even-odd split (see module doc)."""
    circle = []
    for N in [2, 3, 4, 6, 8, 10]:
        n_coeffs = 1 << N
        coeffs = [rand_qm(rng) for _ in range(n_coeffs)]
        init = half_odds(N - 1)[0]
        src = [circle_eval(coeffs, domain_point(init, N - 1, bitrev(i, N))) for i in range(n_coeffs)]
        alpha = rand_qm(rng)
        prev = [rand_qm(rng) for _ in range(n_coeffs // 2)]
        # fold(f)(x) = 2 (f_even(x) + alpha f_odd(x)) where f = f_even(x) + y f_odd(x);
        # f_even / f_odd are line polynomials in the ordered basis (bit k <-> pi^k(x)),
        # and the coefficient index bits of f are [y, x, pi(x), ...] (CPUB/circle.rs:82-99).
        even = [coeffs[2 * j] for j in range(n_coeffs // 2)]
        odd = [coeffs[2 * j + 1] for j in range(n_coeffs // 2)]
        # the circle basis orders the remaining bits MSB-first by fold(): coefficient j of f_even
        # multiplies pi^{N-2}(x)^{b_{N-2}} ... x^{b_0} with bit 0 <-> x -- the ordered line basis.
        dst = []
        a2 = qm_mul(alpha, alpha)
        for i in range(n_coeffs // 2):
            x = domain_point(init, N - 1, bitrev(2 * i, N))[0]
            fe = line_eval_ordered(even, x) if len(even) > 0 else (0, 0, 0, 0)
            fo = line_eval_ordered(odd, x)
            val = qm_scale(qm_add(fe, qm_mul(alpha, fo)), 2)
            dst.append(qm_add(qm_mul(prev[i], a2), val))
        circle.append({"log_size": N, "src": src, "alpha": alpha, "prev": prev, "dst": dst})
    line = []
    for n, step in [(2, 1), (3, 1), (5, 1), (6, 2), (9, 2), (8, 3)]:
        coeffs = [rand_qm(rng) for _ in range(1 << n)]
        init = half_odds(n)[0]
        vals = [line_eval_ordered(coeffs, to_point(coset_index(init, n, bitrev(i, n)))[0]) for i in range(1 << n)]
        alpha = rand_qm(rng)
        cur, a = coeffs, alpha
        scale = 1
        for _ in range(step):
            cur = [qm_add(cur[2 * j], qm_mul(a, cur[2 * j + 1])) for j in range(len(cur) // 2)]
            a = qm_mul(a, a)
            scale *= 2
        m = n - step
        dinit = init * (1 << step) % (1 << 31)
        out_vals = [qm_scale(line_eval_ordered(cur, to_point(coset_index(dinit, m, bitrev(i, m)))[0]), scale)
                    for i in range(1 << m)]
        line.append({"log_size": n, "fold_step": step, "line_initial": init, "src": vals, "alpha": alpha,
                     "dst": out_vals})
    return {"circle": circle, "line": line}


class Channel:
    """Blake2sChannel (STWO/core/channel/blake2s.rs:20-141)."""

    def __init__(self):
        self.digest = bytes(32)
        self.n_draws = 0

    def update(self, d):
        self.digest, self.n_draws = d, 0

    def mix_root(self, root: bytes):
        self.update(h(self.digest + root))

    def mix_felts(self, felts):
        self.update(h(self.digest + b"".join(le32(f) for f in felts)))

    def draw_u32s(self):
        d = h(self.digest + self.n_draws.to_bytes(4, "little") + b"\x00")
        self.n_draws += 1
        return words(d)

    def draw_secure_felt(self):
        while True:
            w = self.draw_u32s()
            if all(x < 2 * P for x in w):
                return tuple(x % P for x in w[:4])


def lifted_root_qm(vals):
    """Lifted commit of the 4 coordinate columns of a QM31 column (all one size)."""
    layer = [h(le32(v)) for v in vals]
    while len(layer) > 1:
        layer = [h(layer[2 * i] + layer[2 * i + 1]) for i in range(len(layer) // 2)]
    return layer[0]


def gen_fri(rng):
    """FriProver::commit transcripts (STWO/prover/fri.rs:133-291) for a low-degree
    QM31 evaluation, with folds evaluated from the coefficient split."""
    out = []
    for log_size, (lld, lb, nq, step) in [(8, (1, 1, 3, 1)), (10, (3, 2, 3, 2)), (9, (0, 2, 3, 3))]:
        deg_log = log_size - lb
        coeffs = [rand_qm(rng) for _ in range(1 << deg_log)] + [(0, 0, 0, 0)] * ((1 << log_size) - (1 << deg_log))
        init = half_odds(log_size - 1)[0]
        evals = [circle_eval(coeffs, domain_point(init, log_size - 1, bitrev(i, log_size)))
                 for i in range(1 << log_size)]
        ch = Channel()
        ch.mix_root(bytes(range(32)))      # a prior transcript state
        roots = []
        r = lifted_root_qm(evals)
        roots.append(words(r))
        ch.mix_root(r)
        alpha = ch.draw_secure_felt()
        # circle -> line on the coefficient side: line coeffs = 2 (even + alpha odd)
        lc = [qm_scale(qm_add(coeffs[2 * j], qm_mul(alpha, coeffs[2 * j + 1])), 2) for j in range(len(coeffs) // 2)]
        lg = log_size - 1
        linit = half_odds(lg)[0]
        last_log = lld + lb

        def line_vals(c, li, l):
            return [line_eval_ordered(c, to_point(coset_index(li, l, bitrev(i, l)))[0]) for i in range(1 << l)]

        if lg != last_log:
            while True:
                last = not (lg > last_log + step)
                s = lg - last_log if last else step
                vals = line_vals(lc, linit, lg)
                r = lifted_root_qm(vals)
                roots.append(words(r))
                ch.mix_root(r)
                alpha = ch.draw_secure_felt()
                a = alpha
                for _ in range(s):
                    lc = [qm_scale(qm_add(lc[2 * j], qm_mul(a, lc[2 * j + 1])), 2) for j in range(len(lc) // 2)]
                    a = qm_mul(a, a)
                linit = linit * (1 << s) % (1 << 31)
                lg -= s
                if last:
                    break
        # last layer: LinePoly internal order is the bit reverse of the ordered coefficients;
        # the degree bound keeps the first 2^lld ordered coefficients.
        bound = 1 << lld
        assert all(c == (0, 0, 0, 0) for c in lc[bound:])
        ordered = lc[:bound]
        internal = [ordered[bitrev(i, lld)] for i in range(bound)]
        ch.mix_felts(internal)
        out.append({"log_size": log_size, "config": [lb, lld, nq, step], "evals": evals,
                    "prior_root": list(range(32)), "roots": roots, "last_layer_poly": internal,
                    "final_digest": list(ch.digest)})
    return out


def main():
    rng = random.Random(20241202)
    data = {
        "twiddles": gen_twiddles(),
        "fft": gen_fft(rng),
        "merkle": gen_merkle(rng),
        "folds": gen_folds(rng),
        "fri": gen_fri(rng),
    }
    for name, val in data.items():
        with open(os.path.join(HERE, name + ".json"), "w") as f:
            json.dump(val, f, separators=(",", ":"))
        print("wrote", name)


if __name__ == "__main__":
    main()

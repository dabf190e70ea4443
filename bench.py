#!/usr/bin/env python3
"""LDE + Merkle benchmark of the stwo commitment hot path on one MI355X.

Workload (one step): 256 M31 columns of 2^20 bit-reversed evaluations on
CanonicCoset(20), resident in HBM -> interpolate_columns (circle IFFT) ->
evaluate_polynomials at blowup 4 onto CanonicCoset(22) (RFFT) -> Blake2s
Merkle commit of the 256 evaluations.  The two transforms run through
stwo_hip_extend_evals (TreeBuilder::extend_evals with the PCS default of not
storing coefficients).  Then, once, the FRI chain on one QM31
evaluation of 2^22 points (fold_circle_into_line, fold_line down to 2^4, a
Merkle commit per layer, Blake2s channel), reported beside.

value = input trace elements (256 * 2^20) / step time, in Melem/s.  Inputs
are synthetic (seeded SplitMix64, u64 % P) and generated on the device.
Interpolation runs in place, so each step first restores the evaluations
from a pristine device copy; that copy is outside the timed span.

Prints ONE JSON line.  --gpus must be 1 (one GPU per box).
"""
import argparse
import json
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "stwo-metal_amd"))

import numpy as np  # noqa: E402

import stwo_hip as sh  # noqa: E402

N_COLS, LOG_N, LOG_BLOWUP = 256, 20, 2
FRI_LOG = 22
HBM_PEAK_GBS = 8000.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def run_step(ctx, work, pristine, outs, layers, tw, logs):
    for w, p in zip(work, pristine):
        ctx._check(sh.lib().stwo_hip_copy(ctx.h, w.ptr, p.ptr, w.n))
    ctx.sync()
    t0 = time.perf_counter()
    # interpolate_columns + evaluate_polynomials as TreeBuilder::extend_evals
    # runs them with the PCS default (store_polynomials_coefficients = false):
    # the IFFT's last pass and the RFFT's first pass fused, the coefficients
    # not stored.
    ctx.extend_evals(work, logs, LOG_BLOWUP, tw, outs, store_coeffs=False)
    ctx.merkle_commit(outs, [LOG_N + LOG_BLOWUP] * N_COLS, layers)
    ctx.sync()
    return time.perf_counter() - t0


def fri_chain(ctx, tw):
    """FRI on a low-degree QM31 evaluation of 2^22 points (degree < 2^21, blowup 2)."""
    cols = []
    for k in range(4):
        co = ctx.alloc(1 << (FRI_LOG - 1))
        ctx.fill_random(co, 1 << (FRI_LOG - 1), 2, 1000 + k)
        out = ctx.alloc(1 << FRI_LOG)
        ctx.evaluate(co, FRI_LOG - 1, sh.canonic_domain(FRI_LOG), tw, out)
        co.free()
        cols.append(out)
    cfg = sh.FriConfig(1, 3, 3, 1)  # last layer 2^4
    ctx.sync()
    best = None
    for _ in range(3):
        t0 = time.perf_counter()
        roots, last, digest, nd = ctx.fri_commit(bytes(32), 0, cfg, cols, FRI_LOG, tw)
        dt = time.perf_counter() - t0
        best = dt if best is None else min(best, dt)
    for c in cols:
        c.free()
    return best * 1e3, int(roots.shape[0])


def uneven_logs(n_cols=1024, lo=10, hi=22, seed=3):
    """log_size = 10 + j, P(j) proportional to 0.7^j, j in 0..12 (seeded)."""
    rng = np.random.default_rng(seed)
    w = np.array([0.7 ** j for j in range(hi - lo + 1)])
    return [int(lo + j) for j in rng.choice(len(w), size=n_cols, p=w / w.sum())]


def batched_vs_plain(ctx, steps=3, warm=6):
    """1,024 uneven columns (log 10..22), LDE at blowup 2, one mixed-degree tree:
    the plain per-column form (stwo_hip_lde_commit_per_column, a C loop of one
    interpolate and one evaluate per column, then the commit) against one
    batched plan replayed as a hipGraph (coefficients not stored, the PCS
    default), with the planner's defaults and after stwo_hip_fft_autotune
    timed every candidate plan of the set's 26 FFT shapes on the device; the
    tuned plan also issued launch by launch.  Each form: `warm` untimed runs,
    then the fastest of `steps`; batched_ms is the fastest batched form."""
    logs = uneven_logs()
    b = 1
    max_log = max(logs) + b
    tw = ctx.precompute_twiddles(sh.half_odds(max_log - 1))
    offs, acc = [], 0
    for l in logs:
        offs.append(acc)
        acc += 1 << l
    oouts, oacc = [], 0
    for l in logs:
        oouts.append(oacc)
        oacc += 1 << (l + b)
    pristine = ctx.alloc(acc)
    ctx.fill_random(pristine, acc, 3, 0)
    work = ctx.alloc(acc)
    outs = ctx.alloc(oacc)
    layers = ctx.alloc(sh.tree_words(max_log))
    cols = [work.at(o) for o in offs]
    out_cols = [outs.at(o) for o in oouts]
    elems = acc

    def restore():
        ctx._check(sh.lib().stwo_hip_copy(ctx.h, work.ptr, pristine.ptr, acc))
        ctx.sync()

    def timed(fn):
        # `warm` untimed runs first: the device's clocks take a few runs of
        # this size to settle, so whichever plan ran first would be slower
        ts = []
        for it in range(warm + steps):
            restore()
            t0 = time.perf_counter()
            fn()
            ctx.sync()
            if it >= warm:
                ts.append(time.perf_counter() - t0)
        return min(ts), layers.download(8)

    tp, root_plain = timed(lambda: ctx.lde_commit_per_column(cols, logs, b, tw, out_cols, layers))
    ctx.fft_tune_clear()
    plan = ctx.batch_create(cols, logs, b, tw, out_cols, layers, sh.TREE_MIXED, sh.HASH_BLAKE2S, True,
                            drop_coeffs=True)
    tb, root_batched = timed(plan.run)
    plan.destroy()
    t0 = time.perf_counter()
    tuned = ctx.fft_autotune(logs, b, tw)
    tune_s = time.perf_counter() - t0
    plan = ctx.batch_create(cols, logs, b, tw, out_cols, layers, sh.TREE_MIXED, sh.HASH_BLAKE2S, True,
                            drop_coeffs=True)
    tt, root_tuned = timed(plan.run)
    plan.destroy()
    # the same plan issued launch by launch instead of replayed as a graph
    plan = ctx.batch_create(cols, logs, b, tw, out_cols, layers, sh.TREE_MIXED, sh.HASH_BLAKE2S, False,
                            drop_coeffs=True)
    te, root_eager = timed(plan.run)
    plan.destroy()
    ctx.fft_tune_clear()
    for x in (pristine, work, outs, layers):
        x.free()
    tw.destroy()
    hist = {str(l): logs.count(l) for l in sorted(set(logs))}
    changed = [{"n": t[0], "N": t[1], "inverse": t[2], "t0": t[3], "smax": t[4], "wg": t[5], "best_us": t[6],
                "default_us": t[7]} for t in tuned if (t[3], t[4], t[5]) != (12, 8, 0)]
    best = min(tb, tt, te)
    return {
        "workload": "1,024 M31 columns, log_size 10..22 (P(10+j) ~ 0.7^j, seed 3), LDE blowup 2, one "
                    "mixed-degree Blake2s tree",
        "input_elems": elems,
        "log_size_histogram": hist,
        "plain_ms": round(tp * 1e3, 3),
        "plain_form": "C loop: per column stwo_hip_interpolate + stwo_hip_evaluate, then stwo_hip_merkle_commit",
        "timing": "each form: %d untimed runs, then the fastest of %d" % (warm, steps),
        "batched_ms": round(best * 1e3, 3),
        "batched_default_ms": round(tb * 1e3, 3),
        "batched_tuned_ms": round(tt * 1e3, 3),
        "batched_tuned_eager_ms": round(te * 1e3, 3),
        "autotune_s": round(tune_s, 2),
        "autotune_shapes": len(tuned),
        "autotune_changed": changed,
        "plain_melem_s": round(elems / tp / 1e6, 1),
        "batched_melem_s": round(elems / best / 1e6, 1),
        "batched_over_plain": round(tp / best, 2),
        "roots_equal": bool(np.array_equal(root_plain, root_batched) and np.array_equal(root_plain, root_tuned)
                            and np.array_equal(root_plain, root_eager)),
    }


def serving(rates=(10.0, 30.0, 60.0), n_jobs=200, budget=96 << 30, n_slots=4):
    """Poisson streams of commit + FRI jobs (8-512 columns of log 12-22, blowup 2 or 4,
    seeded values) through the continuous-batching server (jobs admitted together
    share one batched plan); p50/p99 latency, jobs/s and the mean group size per
    arrival rate.  Three jobs are re-run alone on a 1-slot server without
    co-batching and must give identical results; the two smallest are checked
    against the oracle."""
    import threading
    rng = np.random.default_rng(5)
    out = {"workload": "Poisson arrivals; per job 8-512 M31 columns of log_size 12-22, blowup 2 or 4, "
                       "lifted Blake2s commit + FRI (seeded values, seed 5)", "rates": []}
    all_jobs = []
    srv = sh.Server(0, budget, 24, n_slots)
    for rate in rates:
        jobs = []
        for k in range(n_jobs):
            n_cols = int(rng.integers(8, 513))
            lg = int(rng.integers(12, 23))
            b = int(rng.integers(1, 3))
            jobs.append((n_cols, lg, b, int(rng.integers(1, 1 << 40)), (b, 2, 3, 1)))
        gaps = rng.exponential(1.0 / rate, size=n_jobs)
        ids = [None] * n_jobs

        def producer():
            t_next = time.perf_counter()
            for k, job in enumerate(jobs):
                t_next += gaps[k]
                d = t_next - time.perf_counter()
                if d > 0:
                    time.sleep(d)
                n, lg_, b_, seed, f = job
                ids[k] = srv.submit(n, lg_, b_, seed, sh.FriConfig(*f))

        th = threading.Thread(target=producer)
        th.start()
        th.join()
        res = [srv.wait(i) for i in ids]
        assert all(r.status == 0 for r in res), [r.status for r in res]
        lat = np.array([r.t_done_ms - r.t_submit_ms for r in res])
        span_ms = max(r.t_done_ms for r in res) - min(r.t_submit_ms for r in res)
        elems = sum(n << lg for (n, lg, b, _, _) in jobs)
        out["rates"].append({
            "arrival_rate_jobs_s": rate, "jobs": n_jobs, "jobs_s": round(n_jobs / (span_ms / 1e3), 2),
            "p50_ms": round(float(np.percentile(lat, 50)), 2), "p99_ms": round(float(np.percentile(lat, 99)), 2),
            "mean_group_size": round(float(np.mean([r.group_size for r in res])), 2),
            "mean_queue_ms": round(float(np.mean([r.t_admit_ms - r.t_submit_ms for r in res])), 2),
            "input_melem_s": round(elems / (span_ms / 1e3) / 1e6, 1)})
        log("serving: rate %.0f jobs/s done: %s" % (rate, json.dumps(out["rates"][-1])))
        all_jobs += list(zip(jobs, res))
    srv.close()
    # alone vs mix, and the oracle on the two smallest jobs
    picks = sorted(all_jobs, key=lambda jr: jr[0][0] << jr[0][1])[:3]
    alone_ok = True
    with sh.Server(0, budget, 24, 1, max_group=1) as s1:
        for (job, r) in picks:
            n, lg_, b_, seed, f = job
            a = s1.wait(s1.submit(n, lg_, b_, seed, sh.FriConfig(*f)))
            alone_ok = alone_ok and a.key() == r.key()
    out["alone_equals_mix_checked"] = len(picks)
    out["alone_equals_mix"] = bool(alone_ok)
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import oracle
    ok = True
    for (job, r) in picks[:2]:
        n, lg_, b_, seed, f = job
        L = lg_ + b_
        tw_, itw_ = oracle.precompute_twiddles(oracle.half_odds_initial(L - 1), L - 1)
        evs = [oracle.fill_random(1 << lg_, seed, c) for c in range(n)]
        oracle.interpolate_columns(evs, itw_)
        outs = oracle.evaluate_polynomials(evs, b_, tw_)
        root = oracle.lifted_commit(outs, L)[:8]
        ch = oracle.Channel()
        ch.mix_root(root)
        roots, last, _ = oracle.fri_commit(ch, oracle.FriConfig(*f), outs[:4], L)
        ok = ok and list(r.trace_root) == root.tolist() and bytes(r.final_digest) == ch.digest_bytes() \
            and list(r.fri_roots[: 8 * len(roots)]) == roots.reshape(-1).tolist()
    out["oracle_checked"] = 2
    out["oracle_equal"] = bool(ok)
    return out


# Profiling names of the library's launches -> kernel symbols in rocprofv3 output.
ROCPROF_NAMES = {
    "rfft_pass_wg8": "void stwo_hip::fft_pass_kernel<false, 8, 3, 8, 4, 0>(stwo_hip::FftPass2)",
    "rfft_pass_wg9": "void stwo_hip::fft_pass_kernel<false, 9, 2, 9, 5, 0>(stwo_hip::FftPass2)",
    "ifft_pass_wg8": "void stwo_hip::fft_pass_kernel<true, 8, 3, 0, 4, 8>(stwo_hip::FftPass2)",
    "lde_pass_wg9": "void stwo_hip::lde_pass_kernel<9, 2>(stwo_hip::FftPass2, stwo_hip::FftPass2, unsigned int)",
    "ifft_pass_wg9": "void stwo_hip::fft_pass_kernel<true, 9, 2, 5, 9, 0>(stwo_hip::FftPass2)",
    "merkle_commit_on_layer": "void stwo_hip::merkle_layer_kernel<false>(...)",
    "merkle_nodes": "stwo_hip::merkle_nodes_kernel(...)",
}
TRAFFIC_FILE = os.path.join(REPO, "profiles", "traffic_lde_merkle_256x2e20_b4.json")


PROVE_LOG_N, PROVE_COLS = 18, 256


def prove_wide_fib(ctx, runs=9, cpu=True):
    """The end-to-end wide-Fibonacci prove: 256 columns x 2^18 rows, the first
    two columns seeded random, PcsConfig::default() (blowup 2, 3 queries,
    10 PoW bits), Blake2s channel -- trace generation, trace commit,
    composition, OODS samples, FRI quotients, FRI commit, grind and
    decommitment, with the inputs resident on the device.  Median of `runs`
    proves after two warm-ups, timed around the prove call (the proof object
    on the host; with its bincode bytes beside); one more prove with
    per-kernel device times.  Both sampling paths: coefficients stored (the
    reference benchmark's setting) and the PCS default.  The CPU leg runs the
    oracle's prove of the same inputs on the host."""
    n = 1 << PROVE_LOG_N
    a, b = ctx.alloc(n), ctx.alloc(n)
    ctx.fill_random(a, n, 4, 0)
    ctx.fill_random(b, n, 4, 1)
    cfg = sh.PcsConfig.default()
    tw = ctx.precompute_twiddles(sh.half_odds(PROVE_LOG_N + cfg.fri.log_blowup_factor - 1))

    def timed(store):
        for _ in range(2):  # warm-up: plans built, kept tables uploaded
            ctx.wide_fib_prove(a, b, PROVE_COLS, PROVE_LOG_N, tw, cfg, store_coeffs=store)
        times, walls, phases, data = [], [], [], None
        for _ in range(runs):
            ctx.sync()
            t0 = time.perf_counter()
            data, ph, _, _, _ = ctx.wide_fib_prove(a, b, PROVE_COLS, PROVE_LOG_N, tw, cfg, store_coeffs=store)
            walls.append((time.perf_counter() - t0) * 1e3)
            times.append(ctx.last_prove_ms)
            phases.append(ph)
        mid = sorted(range(runs), key=lambda i: times[i])[runs // 2]
        ctx.prof_enable(True)
        ctx.prof_reset()
        ctx.wide_fib_prove(a, b, PROVE_COLS, PROVE_LOG_N, tw, cfg, store_coeffs=store)
        prof = ctx.prof_read()
        ctx.prof_enable(False)
        return times, phases[mid], data, prof, times[mid], sorted(walls)[runs // 2]

    times, ph, data, prof, ms, wall = timed(True)
    times2, ph2, data2, prof2, ms2, wall2 = timed(False)
    res = {
        "workload": "wide-Fibonacci prove: 256 columns x 2^18 rows (first two seeded random), PcsConfig::default() "
                    "(log blowup 1, 3 queries, 10 PoW bits, fold step 1), Blake2s channel, bincode proof; "
                    "coefficients stored (set_store_polynomials_coefficients, as the reference's benchmark)",
        "ms": round(ms, 3),
        "ms_def": "wall time of the stwo_hip_wide_fib_prove call (inputs resident, proof object on the host)",
        "runs_ms": [round(t, 3) for t in times],
        "with_bincode_ms": round(wall, 3),
        "phases_ms": {k: round(v, 3) for k, v in ph.items()},
        "proof_bytes": len(data),
        "kernels_ms": {k: round(v[1], 3) for k, v in sorted(prof.items(), key=lambda kv: -kv[1][1])},
        "default_pcs": {
            "path": "store_polynomials_coefficients = false: LDE without the coefficient store, OODS samples "
                    "from the evaluations by barycentric weights",
            "ms": round(ms2, 3),
            "runs_ms": [round(t, 3) for t in times2],
            "with_bincode_ms": round(wall2, 3),
            "phases_ms": {k: round(v, 3) for k, v in ph2.items()},
            "kernels_ms": {k: round(v[1], 3) for k, v in sorted(prof2.items(), key=lambda kv: -kv[1][1])},
            "proof_equal_to_stored_path": data2 == data,
        },
    }
    a.free()
    b.free()
    tw.destroy()
    if cpu:
        sys.path.insert(0, os.path.join(REPO, "oracle"))
        import oracle
        import prove as oracle_prove
        ha, hb = oracle.fill_random(n, 4, 0), oracle.fill_random(n, 4, 1)
        ct = []
        for _ in range(2):
            t0 = time.perf_counter()
            want = oracle_prove.wide_fib_prove(ha, hb, PROVE_COLS, PROVE_LOG_N)
            ct.append((time.perf_counter() - t0) * 1e3)
        res["cpu_oracle_ms"] = round(min(ct), 1)
        res["cpu_oracle_cores"] = oracle.num_threads()
        res["proof_equal_to_oracle"] = oracle_prove.proof_bytes(want) == data
    return res


# Integer-VALU issue ceiling of the MI355X: 256 CUs x 4 SIMDs, one wave64
# VALU instruction per SIMD every 4 cycles at 2.4 GHz.
VALU_PEAK_WAVE_INSTR_S = 256 * 4 * 2.4e9 / 4


def measured_counters(kernel):
    """The committed rocprofv3 --pmc summary's entry for `kernel` (HBM bytes
    per launch from FETCH_SIZE and WRITE_SIZE, each scaled by the same
    counter's reading for a copy of known size; VALU wave-instructions per
    launch from SQ_INSTS_VALU), or {} when none covers it."""
    try:
        with open(TRAFFIC_FILE) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return {}
    return t.get("kernels", {}).get(kernel) or {}


def measured_traffic(kernel):
    return measured_counters(kernel).get("hbm_bytes_per_launch")


def compulsory_bytes():
    """Compulsory HBM bytes of one step, per transform (twiddle reads and the
    intermediate passes excluded): the IFFT reads and writes each column's
    2^n words once (8 * 2^n B), the RFFT reads 2^n coefficients and writes
    2^(n+b) evaluations once (4 * 2^n + 4 * 2^(n+b) B); Merkle leaves read
    4 B per element per column and write 32 B per leaf, node layers read
    64 B and write 32 B per node."""
    n, b = LOG_N, LOG_BLOWUP
    n_out = 1 << (n + b)
    ifft = 8.0 * (1 << n) * N_COLS
    rfft = (4.0 * (1 << n) + 4.0 * n_out) * N_COLS
    return {"ifft": ifft, "rfft": rfft, "lde": ifft + rfft,
            "merkle_leaves": 4.0 * N_COLS * n_out + 32.0 * n_out, "merkle_nodes": 96.0 * (n_out - 1)}


def transform_of(name):
    """The step's transform a launch (profiling name) belongs to."""
    if name.startswith("ifft_pass"):
        return "ifft"
    if name.startswith("rfft_pass"):
        return "rfft"
    if name.startswith("lde_pass"):  # IFFT and RFFT passes fused into one launch
        return "lde"
    return {"merkle_commit_on_layer": "merkle_leaves", "merkle_nodes": "merkle_nodes"}.get(name)


def compulsory_roofline(kernels, steps):
    """Roofline point of the dominant launch (most device time per step) on
    compulsory bytes: the compulsory bytes of the transform it belongs to
    over its mean launch time -- the rate the whole transform would reach if
    this launch alone moved its bytes.  Also the same rate for the IFFT, the
    RFFT and the whole LDE over their summed pass times.  Adds each kernel's
    own read + write per launch ("pass_io_GBps", from the library's pass
    planner) to `kernels`."""
    comp = compulsory_bytes()
    for inverse, n, N in ((True, LOG_N, LOG_N), (False, LOG_N, LOG_N + LOG_BLOWUP)):
        io = {}
        for p in sh.fft_plan(n, N, inverse):
            name = "%s_pass_wg%d" % ("ifft" if inverse else "rfft", p["wg"])
            io[name] = io.get(name, 0.0) + 4.0 * N_COLS * ((1 << min(p["in_log"], N - p["rep"])) + (1 << N))
        for k, v in io.items():
            if k in kernels:
                kernels[k]["pass_io_GBps"] = round(v / (kernels[k]["ms_total"] / steps * 1e-3) / 1e9, 1)
    ms = {t: 0.0 for t in comp}
    for k, v in kernels.items():
        t = transform_of(k)
        if t:
            ms[t] += v["ms_total"] / steps
    fused = ms["lde"] > 0
    ms["lde"] += ms["ifft"] + ms["rfft"]
    transforms = {}
    for t in (("lde",) if fused else ("ifft", "rfft", "lde")):
        if ms[t] > 0:
            a = comp[t] / (ms[t] * 1e-3) / 1e9
            transforms[t] = {"compulsory_bytes": int(comp[t]), "ms_per_step": ms[t], "achieved": a,
                             "frac": a / HBM_PEAK_GBS}
    dom = max((k for k in kernels if transform_of(k)), key=lambda k: kernels[k]["ms_total"])
    t = transform_of(dom)
    launches = kernels[dom]["launches"] / steps
    launch_ms = kernels[dom]["ms_total"] / kernels[dom]["launches"]
    dom_bytes = comp[t] / launches
    traffic = measured_traffic(dom)
    return {"kernel": dom, "algorithmic_of": t, "launch_ms": launch_ms, "algorithmic_bytes": dom_bytes,
            "achieved": dom_bytes / (launch_ms * 1e-3) / 1e9, "traffic": traffic, "transforms": transforms,
            "lde_ms_per_step": ms["lde"]}


def hbm_kernels(ctx, log_n=26, runs=3):
    """The kernels the HBM bandwidth bounds, at sizes far past the caches:
    bit_reverse_column of a 2^27-word column (8 B per element: one read and
    one write), fold_circle_into_line of a 2^26 QM31 evaluation into a 2^25
    line (read 32 B of pairs and 16 B of the accumulator, write 16 B per
    output) and one fold_line step 2^25 -> 2^24 (read 32 B, write 16 B per
    output); twiddle reads are not counted.  Device time from HIP events."""
    n_br = log_n + 1
    tw = ctx.precompute_twiddles(sh.half_odds(log_n - 1))
    col = ctx.alloc(1 << n_br)
    ctx.fill_random(col, 1 << n_br, 8, 0)
    src = [ctx.alloc(1 << log_n) for _ in range(4)]
    acc = [ctx.alloc(1 << (log_n - 1)) for _ in range(4)]
    line = [ctx.alloc(1 << (log_n - 2)) for _ in range(4)]
    for k, b in enumerate(src):
        ctx.fill_random(b, 1 << log_n, 8, 1 + k)
    for k, b in enumerate(acc):
        ctx.fill_random(b, 1 << (log_n - 1), 8, 5 + k)
    alpha = [1, 2, 3, 4]

    def once():
        ctx.bit_reverse_column(col, n_br)
        ctx.fold_circle_into_line(acc, src, sh.half_odds(log_n - 1), alpha, tw)
        ctx.fold_line(acc, sh.half_odds(log_n - 1), alpha, 1, tw, line)

    once()
    ctx.sync()
    ctx.prof_enable(True)
    ctx.prof_reset()
    for _ in range(runs):
        once()
    ctx.sync()
    prof = ctx.prof_read()
    ctx.prof_enable(False)
    out = {}
    for name, elems, per in (("bit_reverse_tiled", 1 << n_br, 8.0),
                             ("fold_circle_into_line", 1 << (log_n - 1), 64.0),
                             ("fold_line", 1 << (log_n - 2), 48.0)):
        cnt, ms = prof[name]
        t = ms / cnt
        gbps = elems * per / (t * 1e-3) / 1e9
        out[name] = {"outputs": elems, "bytes_per_output": per, "launch_ms": round(t, 4),
                     "algorithmic_GBps": round(gbps, 1), "frac": round(gbps / HBM_PEAK_GBS, 4)}
    for b in [col] + src + acc + line:
        b.free()
    tw.destroy()
    return out


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(gpu_root=None, runs=6):
    """The SimdBackend-equivalent C++ port (oracle/simd_port.cpp: packed M31
    on AVX-512 or AVX2, vecwise + radix-8 FFT layers cached below 2^16,
    W-way Blake2s; OpenMP over columns, blocks and leaf chunks on every host
    core) on the same workload: interpolate 256 x 2^20, evaluate at blowup
    4, Blake2s Merkle commit.  One warm-up, then the median of `runs`; the
    root must equal the GPU's."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import oracle
    import simd_port
    init = oracle.half_odds_initial(LOG_N + LOG_BLOWUP - 1)
    tw, itw = oracle.precompute_twiddles(init, LOG_N + LOG_BLOWUP - 1)
    port = simd_port.Port(tw, itw)
    pristine = [oracle.fill_random(1 << LOG_N, 2, c) for c in range(N_COLS)]
    work = [np.empty_like(p) for p in pristine]
    outs = [np.empty(1 << (LOG_N + LOG_BLOWUP), dtype=np.uint32) for _ in range(N_COLS)]
    layers = np.empty(((2 << (LOG_N + LOG_BLOWUP)) - 1) * 8, dtype=np.uint32)
    times = []
    for it in range(runs + 1):
        for w, p in zip(work, pristine):
            np.copyto(w, p)
        t0 = time.perf_counter()
        port.interpolate_columns(work, LOG_N)
        port.evaluate_columns(work, LOG_N, LOG_BLOWUP, outs)
        port.merkle_commit(outs, LOG_N + LOG_BLOWUP, layers)
        if it:
            times.append(time.perf_counter() - t0)
    dt = sorted(times)[len(times) // 2]
    res = {
        "value": round(N_COLS * (1 << LOG_N) / dt / 1e6, 3),
        "unit": "Melem/s",
        "cores": port.threads,
        "kind": "port",
        "label": "SimdBackend-equivalent C++ (not the Rust binary)",
        "simd_lanes": port.width,
        "cpu_model": cpu_model(),
        "sample": "the full workload (%d columns x 2^20 -> blowup 4 -> Blake2s Merkle commit), median of %d runs "
                  "after one warm-up, %.3f s each (runs %s)" % (N_COLS, runs, dt,
                                                                ", ".join("%.3f" % t for t in times)),
    }
    if gpu_root is not None:
        res["root_equal_to_gpu"] = bool(np.array_equal(layers[:8], gpu_root))
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fri", action="store_true")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-prove", action="store_true")
    ap.add_argument("--no-percol", action="store_true", help="skip the per-column form of the headline workload")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only (--gpus %d requested); "
              "the multi-GPU layout is a design note, not a measured path." % args.gpus, file=sys.stderr)
        sys.exit(2)

    ctx = sh.Context(0)
    root = sh.half_odds(LOG_N + LOG_BLOWUP - 1)
    tw = ctx.precompute_twiddles(root)
    logs = [LOG_N] * N_COLS
    log("allocating %d x 2^%d inputs, %d x 2^%d evaluations" % (N_COLS, LOG_N, N_COLS, LOG_N + LOG_BLOWUP))
    pristine = [ctx.alloc(1 << LOG_N) for _ in range(N_COLS)]
    for c, p in enumerate(pristine):
        ctx.fill_random(p, 1 << LOG_N, 2, c)
    work = [ctx.alloc(1 << LOG_N) for _ in range(N_COLS)]
    outs = [ctx.alloc(1 << (LOG_N + LOG_BLOWUP)) for _ in range(N_COLS)]
    layers = ctx.alloc(sh.tree_words(LOG_N + LOG_BLOWUP))
    ctx.sync()

    for i in range(args.warmup):
        dt = run_step(ctx, work, pristine, outs, layers, tw, logs)
        log("warmup %d: %.2f ms" % (i, dt * 1e3))
    ctx.prof_enable(True)
    ctx.prof_reset()
    times = []
    for i in range(args.steps):
        dt = run_step(ctx, work, pristine, outs, layers, tw, logs)
        times.append(dt)
        log("step %d: %.2f ms" % (i, dt * 1e3))
    prof = ctx.prof_read()
    ctx.prof_enable(False)
    gpu_root = layers.download(8)
    total = sum(times)
    ms_per_step = total / args.steps * 1e3
    elems = N_COLS * (1 << LOG_N)
    value = elems / (total / args.steps) / 1e6

    # Per-kernel device time (HIP events around every launch on the context stream).
    kernels = {k: {"launches": v[0], "ms_total": round(v[1], 3)} for k, v in prof.items() if k != "twiddle_tree"}
    roof = compulsory_roofline(kernels, args.steps)
    dom_name = roof["kernel"]
    dom_launch_ms = roof["launch_ms"]
    traffic = roof["traffic"]
    # The same kernel against the integer-VALU issue ceiling, which is what
    # bounds it: its VALU wave-instructions per launch (committed counter run)
    # over the live launch time.
    valu_instr = measured_counters(dom_name).get("valu_wave_instr_per_launch")
    valu_roofline = None
    if valu_instr:
        va = valu_instr / (dom_launch_ms * 1e-3)
        valu_roofline = {"kernel": dom_name, "wave_instr_per_launch": valu_instr, "achieved": round(va / 1e9, 1),
                         "peak": round(VALU_PEAK_WAVE_INSTR_S / 1e9, 1), "unit": "G wave-instr/s",
                         "frac": round(va / VALU_PEAK_WAVE_INSTR_S, 4)}
    log("kernels: %s" % json.dumps(kernels))
    # Side rates: evaluated (post-blowup) elements per second; the FFT
    # kernels' algorithmic bytes over their device time; Blake2s
    # compressions (2^L * ceil(4 C / 64) leaf blocks + 2^L - 1 nodes) per
    # second of step time and of Merkle kernel time.
    n_out = 1 << (LOG_N + LOG_BLOWUP)
    fft_ms = roof["lde_ms_per_step"]
    merkle_ms = sum(kernels[k]["ms_total"] for k in ("merkle_commit_on_layer", "merkle_nodes") if k in kernels) / args.steps
    n_compress = n_out * -(-4 * N_COLS // 64) + (n_out - 1)
    side = {
        "evaluated_melem_s": round(N_COLS * n_out / (total / args.steps) / 1e6, 1),
        "fft_ms_per_step": round(fft_ms, 3),
        "fft_compulsory_GBps": round(roof["transforms"]["lde"]["achieved"], 1),
        "merkle_ms_per_step": round(merkle_ms, 3),
        "compressions_per_step": n_compress,
        "compressions_per_s_step": round(n_compress / (total / args.steps), 1),
        "compressions_per_s_merkle_kernels": round(n_compress / (merkle_ms * 1e-3), 1) if merkle_ms else None,
    }
    log("side rates: %s" % json.dumps(side))

    # The same workload in the reference's plain per-column form (one
    # interpolate and one evaluate launch sequence per column, then the
    # commit: stwo_hip_lde_commit_per_column), timed beside the headline.
    pc = []
    for i in range(0 if args.no_percol else 3):
        for w, p in zip(work, pristine):
            ctx._check(sh.lib().stwo_hip_copy(ctx.h, w.ptr, p.ptr, w.n))
        ctx.sync()
        t0 = time.perf_counter()
        ctx.lde_commit_per_column(work, logs, LOG_BLOWUP, tw, outs, layers, sh.TREE_MIXED)
        ctx.sync()
        pc.append(time.perf_counter() - t0)
    per_column = None
    if pc:
        per_column = {"form": "per column: stwo_hip_interpolate + stwo_hip_evaluate per column, then the commit",
                      "ms": round(min(pc) * 1e3, 3), "melem_s": round(elems / min(pc) / 1e6, 1),
                      "root_equal": bool(np.array_equal(layers.download(8), gpu_root))}
        log("per-column form: %s" % json.dumps(per_column))

    serve = None
    batched = None
    if not args.no_batched:
        batched = batched_vs_plain(ctx)
        log("batched vs plain: %s" % json.dumps(batched))

    prove = None
    if not args.no_prove:
        prove = prove_wide_fib(ctx, cpu=not args.no_cpu)
        log("prove: %s" % json.dumps(prove))

    fri_ms, fri_layers = (None, None)
    if not args.no_fri:
        fri_ms, fri_layers = fri_chain(ctx, tw)
        log("fri chain: %.2f ms, %d committed layers" % (fri_ms, fri_layers))

    for x in pristine + work + outs + [layers]:
        x.free()
    pristine, work, outs, layers = [], [], [], None
    hbm = hbm_kernels(ctx)
    log("HBM-bound kernels: %s" % json.dumps(hbm))
    if not args.no_serve:
        serve = serving()
        log("serving: %s" % json.dumps(serve))

    cpu = None
    if not args.no_cpu:
        log("cpu baseline (SimdBackend-equivalent port) ...")
        cpu = cpu_baseline(gpu_root)
        log("cpu: %s" % json.dumps(cpu))

    tw.destroy()
    ctx.close()

    line = {
        "metric": "LDE+Merkle Melem/s (interpolate + evaluate + Blake2s Merkle commit, bit-exact)",
        "value": round(value, 2),
        "unit": "Melem/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32 (M31)",
        "data": "synthetic (seeded SplitMix64 u64 % P, generated on device)",
        "config": {
            "workload": "plain method, large: 256 M31 columns x 2^20 -> LDE blowup 4 (2^22 rows) -> Blake2s "
                        "Merkle commit; FRI chain on a 2^22 QM31 evaluation reported beside",
            "n_cols": N_COLS, "log_size": LOG_N, "log_blowup": LOG_BLOWUP,
            "parallelism": "single GPU",
        },
        "roofline": {
            "bound": "valu",
            "kernel": dom_name,
            "rocprof_name": ROCPROF_NAMES.get(dom_name, dom_name),
            "launch_ms": round(dom_launch_ms, 4),
            "algorithmic_bytes": int(roof["algorithmic_bytes"]),
            "algorithmic_of": roof["algorithmic_of"],
            "achieved": round(roof["achieved"], 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(roof["achieved"] / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "traffic_over_algorithmic": round(traffic / roof["algorithmic_bytes"], 3) if traffic else None,
            "transforms": {k: {kk: (round(vv, 4) if isinstance(vv, float) else vv) for kk, vv in v.items()}
                           for k, v in roof["transforms"].items()},
        },
        "valu_roofline": valu_roofline,
        "side_rates": side,
        "per_column_form": per_column,
        "hbm_kernels": hbm,
        "cpu_baseline": cpu,
        "kernels": kernels,
        "batched_uneven": batched,
        "serving": serve,
        "prove_wide_fib": prove,
        "fri_chain_ms": round(fri_ms, 3) if fri_ms is not None else None,
        "fri_committed_layers": fri_layers,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()

#!/usr/bin/env python3
"""Throughput of the --simulate --no-pp path on one MI355X.

Headline workload: the included girth-8 (3,12)-regular [[9216,4612]] code,
i.i.d. depolarizing noise at p = 0.04, max-iter 1000, FP32.  A step is a
batch of --trials (65,536) seeded trials: each is sampled on the device
(seed = splitmix64(base ^ idx*phi), mt19937_64), its syndromes formed, then
decoded by the learned min-sum decoder (T = 25 learned iterations; shipped
weights trained by this project's own kernels), with every trial it leaves
unconverged decoded again by plain FP32 BP from the prior up to max-iter
1000, and judged by the stabilizer-residual test.  The timed steps run as
one simulate call (--steps-per-call to split them), in launches of --chunk
(16,384) trials pipelined over four streams (sampling, two decode streams,
fallback/residual/copies) and eight buffer sets.  The same trial indices are then run
by plain batched FP32 BP: `value` is the learned decoder's trials/s, with
plain BP's trials/s, both FERs and their Wilson intervals ("matched FER":
the learned FER no higher than the upper end of plain BP's 95% interval;
"fer_below_plain_ci" when it is under the interval's lower end) beside it.

Prints one JSON line (last line of stdout).  `roofline` is for the learned
decoder's kernel (decode_qc_lms_kernel): it keeps every message on chip, so
HBM does not bind it; the vector ALU does.  Its per-trial-iteration issue
slots (wave64 VALU instructions at their issue cost) come from the rocprofv3
counter profile committed for this kernel source (profiles/, checked by the
source's SHA-256: a stale profile is not used); `achieved` is that count
times the kernel's trial-iterations per second, measured here with HIP
events over single-launch steps; `peak` is the MI355X issue rate (256 CUs x
4 SIMDs x 2.4 GHz / 4 cycles per wave64 instruction).  `cpu_baseline` times
the reference decoder (built from its own source under oracle/_ref/) on a
bounded sample of the same workload on this host's cores.
"""
import argparse
import ctypes as C
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
PARAMS = os.path.join(ROOT, "joint_bp_plus_pp_amd", "codes",
                      "H_P768_J3_L12_dmax3_nc0-3_1-2_seed11579811919164041.txt")
METRIC = "decoded trials/s at p=0.04 at matched FER (learned vs plain BP), % HBM peak"
EDGES_PER_ITER = 2 * 27648
BYTES_PER_ITER = 16 * EDGES_PER_ITER  # 884,736
HBM_PEAK_GBS = 8000.0
LEARNED_MS_FILE = os.path.join(ROOT, "joint_bp_plus_pp_amd", "codes", "learned_ms_T25_P768_J3_L12.npz")
LEARNED_LMS_FILE = os.path.join(ROOT, "joint_bp_plus_pp_amd", "codes", "learned_lms_T25_P768_J3_L12.npz")
LEARNED_SP_FILES = [os.path.join(ROOT, "joint_bp_plus_pp_amd", "codes", f)
                    for f in ("learned_T25_P768_J3_L12.npz", "learned_T25_nomlp_P768_J3_L12.npz")]
KERNEL_SOURCE = os.path.join(ROOT, "joint_bp_plus_pp_amd", "csrc", "jbp_decode_f32.hip")
VALU_PROFILE = os.path.join(ROOT, "profiles", "decode_qc_f32_kernels_p0.04_counters.json")
# MI355X: 256 compute units x 4 SIMDs, 2.4 GHz peak engine clock, one wave64
# VALU instruction per SIMD every 4 cycles
VALU_PEAK_GSLOTS = 256 * 4 * 2.4 / 4.0
MIXED_PS = (0.01, 0.025, 0.04, 0.05)
LEARNED_KINDS = {0: "learned sum-product", 1: "learned flooding min-sum",
                 2: "learned layered min-sum"}


def wilson(n, k, z=1.96):
    if n <= 0:
        return 0.0, 0.0
    phat = k / n
    denom = 1 + z * z / n
    center = (phat + z * z / (2 * n)) / denom
    half = z * math.sqrt((phat * (1 - phat) + z * z / (4 * n)) / n) / denom
    return max(0.0, center - half), min(1.0, center + half)


def cpu_threads():
    """Host threads for the CPU baseline: every core of this process's
    affinity mask, capped at the job's CPU share when the environment states
    one (OMP_NUM_THREADS; 16 per GPU on the GPU hosts).  Returns (threads,
    facts about the host)."""
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    share = os.environ.get("OMP_NUM_THREADS", "")
    threads = avail
    if share.isdigit() and int(share) > 0:
        threads = min(threads, int(share))
    facts = {"host_cpu_count": os.cpu_count(), "affinity_cpus": avail,
             "cpu_share": int(share) if share.isdigit() else None}
    return max(1, threads), facts


def _run_cpu(count, first_trial, args, threads):
    """(seconds, iterations, successes, kind) of the reference decoder (or,
    without it, the oracle restatement) on trials first..first+count-1."""
    harness = os.path.join(ROOT, "oracle", "_ref", "libref_harness.so")
    n = count
    it = (C.c_int * n)()
    ok = (C.c_uint8 * n)()
    su = (C.c_uint8 * n)()
    ex = (C.c_uint8 * n)()
    if os.path.exists(harness):
        L = C.CDLL(harness)
        L.ref_load_params.argtypes = [C.c_char_p]
        L.ref_simulate.restype = C.c_double
        L.ref_simulate.argtypes = [C.c_uint64, C.c_longlong, C.c_longlong, C.c_double, C.c_int,
                                   C.c_double, C.c_int] + [C.c_void_p] * 4
        if L.ref_load_params(PARAMS.encode()) != 0:
            raise RuntimeError("reference could not read the params file")
        secs = L.ref_simulate(args.seed, first_trial, count, args.p, args.max_iter, 0.0, threads,
                              it, ok, su, ex)
        kind = "reference"
    else:
        L = C.CDLL(os.path.join(ROOT, "oracle", "liboracle.so"))
        L.orc_code_from_params.restype = C.c_void_p
        L.orc_code_from_params.argtypes = [C.c_char_p]
        L.orc_simulate.restype = C.c_double
        L.orc_simulate.argtypes = [C.c_void_p, C.c_uint64, C.c_longlong, C.c_longlong, C.c_double,
                                   C.c_int, C.c_double, C.c_int] + [C.c_void_p] * 5
        h = L.orc_code_from_params(PARAMS.encode())
        secs = L.orc_simulate(h, args.seed, first_trial, count, args.p, args.max_iter, 0.0, threads,
                              it, ok, su, ex, None)
        kind = "port"
    return secs, sum(it), sum(su), kind


def cpu_baseline(args, first_trial):
    """Reference decoder (or, without it, the oracle restatement) on a bounded
    sample of the workload: the first `cpu_trials` trial indices on every
    host thread of the job, and the first `cpu_trials_1core` on one thread
    (the reference program itself is single-threaded, jointbp_ets.cpp:5931)."""
    threads, facts = cpu_threads()
    count = args.cpu_trials
    secs, iters, succ, kind = _run_cpu(count, first_trial, args, threads)
    out = {"value": round(count / secs, 3), "unit": "trials/s", "cores": threads, "kind": kind,
           "sample": "%d trials (indices %d..%d) of the same workload, p=%g, max-iter %d, FP64 "
                     "reference decoder, one trial per thread on %d threads; %.1f s wall, "
                     "avg_iter %.2f, failures %d" % (count, first_trial, first_trial + count - 1,
                                                    args.p, args.max_iter, threads, secs,
                                                    iters / count, count - succ)}
    out.update(facts)
    m = args.cpu_trials_1core
    if m > 0:
        s1, it1, _, _ = _run_cpu(m, first_trial, args, 1)
        out["one_core"] = {"value": round(m / s1, 3), "unit": "trials/s",
                           "sample": "%d trials (indices %d..%d), one thread, %.1f s, avg_iter %.2f"
                                     % (m, first_trial, first_trial + m - 1, s1, it1 / m)}
    return out


def pp_workload(jbp, dec, args, first):
    """Post-processing on: the reference's default mode (ETS / flip / OSD),
    on a workload where most trials need it (p=0.04, max-iter 8: BP alone
    fails on about 70% of the trials).  The reference's own PP loop (built
    from its source, oracle/_ref/) runs a bounded sample beside it."""
    n, max_iter, p = args.pp_trials, args.pp_max_iter, args.p
    dec.set_pp_tables(jbp.PPTables.from_dir(dec.code, jbp.default_ets_dir(PARAMS)))
    dec.simulate(args.seed, 1024, p, max_iter=max_iter, first_trial=first, pp=True, per_trial=False)
    done = fails = pp_trials = pp_ok = bp_fail = 0
    pp_ms = 0.0
    t0 = time.perf_counter()
    for k in range(args.pp_calls):
        st, _ = dec.simulate(args.seed, n, p, max_iter=max_iter, first_trial=first + k * n, pp=True,
                             per_trial=False)
        done += st["done"]
        fails += st["failures"]
        bp_fail += st["bp_failures"]
        pp_trials += st["pp_trials"]
        pp_ok += st["pp_success"]
        pp_ms += st["pp_ms"]
    el = time.perf_counter() - t0
    dec.set_pp_tables(None)
    out = {"workload": "pp on, p=%g, max-iter %d, %d trials per call, FP32 first pass + FP64 "
                       "stage" % (p, max_iter, n),
           "trials_per_s": round(done / el, 2), "trials": done, "bp_failures": bp_fail,
           "pp_trials": pp_trials, "pp_success": pp_ok, "fer": fails / done,
           "stage_share": round(pp_ms / 1e3 / el, 3)}
    harness = os.path.join(ROOT, "oracle", "_ref", "libref_harness.so")
    if os.path.exists(harness) and args.pp_cpu_trials > 0:
        L = C.CDLL(harness)
        L.ref_load_params.argtypes = [C.c_char_p]
        L.ref_ets_clear()
        L.ref_ets_add.argtypes = [C.c_int, C.c_char_p, C.c_char_p, C.c_char_p, C.c_int, C.c_void_p]
        L.ref_simulate_pp.restype = C.c_double
        L.ref_simulate_pp.argtypes = [C.c_uint64, C.c_longlong, C.c_longlong, C.c_double, C.c_int,
                                      C.c_double, C.c_int, C.c_int] + [C.c_void_p] * 7
        if L.ref_load_params(PARAMS.encode()) != 0:
            raise RuntimeError("reference could not read the params file")
        counts = (C.c_int * 2)()
        for size, label, name in jbp.ETS_GROUPS:
            path = os.path.join(jbp.default_ets_dir(PARAMS), name).encode()
            L.ref_ets_add(size, label.encode(), path, path, dec.code.P, counts)
        m = args.pp_cpu_trials
        bufs = [(C.c_int * m)()] + [(C.c_uint8 * m)() for _ in range(4)]
        ctr = (C.c_longlong * 10)()
        threads, _ = cpu_threads()
        secs = L.ref_simulate_pp(args.seed, first, m, p, max_iter, 0.0, 5, threads, *bufs, None, ctr)
        L.ref_ets_clear()
        out["cpu_reference"] = {"trials_per_s": round(m / secs, 2), "cores": threads,
                                "sample": "%d trials (indices %d..%d), %.1f s wall" %
                                          (m, first, first + m - 1, secs)}
        out["speedup_vs_cpu_reference"] = round(out["trials_per_s"] / (m / secs), 1)
    return out


def sp_learned_leg(jbp, dec, args, first):
    """The sum-product learned models (weights on the reference's joint BP),
    for comparison: trials/s and FER over --sp-learned-steps steps."""
    from joint_bp_plus_pp_amd import learned as jl
    out = {}
    for path in LEARNED_SP_FILES:
        if not os.path.exists(path) or args.sp_learned_steps <= 0:
            continue
        model = jl.load(dec, path)
        nxt, done, fails, iters = first, 0, 0, 0
        t0 = time.perf_counter()
        for _ in range(args.sp_learned_steps):
            st, _ = dec.simulate_learned(model, args.seed, args.trials, args.p,
                                         max_iter=args.max_iter, first_trial=nxt, per_trial=False,
                                         chunk=args.chunk)
            nxt += args.trials
            done += st["done"]
            fails += st["failures"]
            iters += st["total_iters"]
        el = time.perf_counter() - t0
        key = "sum_product_mlp_damping" if model.use_mlp else "sum_product_weights_only"
        out[key] = {"weights": os.path.relpath(path, ROOT), "trials": done,
                    "trials_per_s": round(done / el, 2), "fer": fails / done,
                    "avg_iter": round(iters / done, 4)}
        model.destroy()
    return out or None


def training_leg(jbp, dec, args):
    """The training workload: --train-steps steps (after one untimed) of the
    min-sum model with its MLP damping at T = 25, batch --train-batch, p
    uniform in [0.03, 0.05], each a device-sampled batch, the forward pass
    with its tape, the adjoint pass and the Adam update.  The forward tape
    (the check messages of every iteration, (T + 1) x 55,296 FP32 values per
    trial) is written once and read once per step: that is the step's
    streaming traffic.  The same number of steps of the headline decoder's
    model (layered min-sum, no damping) is timed after it."""
    T, B = 25, args.train_batch
    model = jbp.LearnedModel(dec, T, kind=jbp.LEARNED_MIN_SUM)
    params = model.get_params()
    params[:4 * model.nb * T] = 1.0
    params[:2 * model.nb * T] = 0.75
    model.set_params(params, use_mlp=True)
    model.train_step(args.seed, 1, B, 0.03, 0.05, lr=1e-3)
    t0 = time.perf_counter()
    losses = []
    for k in range(args.train_steps):
        losses.append(model.train_step(args.seed, 1 + (k + 1) * B, B, 0.03, 0.05, lr=1e-3))
    el = (time.perf_counter() - t0) / args.train_steps
    tape = B * (T + 1) * EDGES_PER_ITER * 4
    model.destroy()
    # the headline decoder's model (layered min-sum, no damping), same steps
    lm = jbp.LearnedModel(dec, T, kind=jbp.LEARNED_LAYERED_MIN_SUM)
    lp = lm.get_params()
    lp[:4 * lm.nb * T] = 1.0
    lp[:2 * lm.nb * T] = 0.75
    lm.set_params(lp, use_mlp=False)
    lm.train_step(args.seed, 1, B, 0.03, 0.05, lr=1e-3)
    t1 = time.perf_counter()
    llosses = []
    for k in range(args.train_steps):
        llosses.append(lm.train_step(args.seed, 1 + (k + 1) * B, B, 0.03, 0.05, lr=1e-3))
    lel = (time.perf_counter() - t1) / args.train_steps
    lm.destroy()
    return {"workload": "min-sum model with MLP damping, T=%d, batch %d, p ~ U[0.03, 0.05], FP32 forward "
                        "+ adjoint + Adam" % (T, B),
            "layered_min_sum": {"ms_per_step": round(1e3 * lel, 2),
                                "trial_iterations_per_s": round(B * T / lel, 1),
                                "loss": [round(x, 6) for x in llosses]},
            "steps": args.train_steps, "ms_per_step": round(1e3 * el, 2),
            "trial_iterations_per_s": round(B * T / el, 1),
            "tape_bytes_per_step": tape,
            "tape_traffic_gbs": round(2 * tape / el / 1e9, 1),
            "tape_traffic_frac_hbm": round(2 * tape / el / 1e9 / HBM_PEAK_GBS, 4),
            "loss": [round(x, 6) for x in losses]}


def source_sha256(path):
    import hashlib
    with open(path, "rb") as f:
        return hashlib.sha256(f.read()).hexdigest()


def kernel_profile():
    """Counter profile committed for the current kernel source, or None when
    there is none or it was taken on another version of the source."""
    if not os.path.exists(VALU_PROFILE):
        return None, "no counter profile"
    with open(VALU_PROFILE) as f:
        prof = json.load(f)
    if prof.get("source_sha256") != source_sha256(KERNEL_SOURCE):
        return None, "counter profile taken on another version of %s" % os.path.relpath(KERNEL_SOURCE, ROOT)
    return prof, None


def mixed_workload(jbp, qc_dec, args):
    """Interleaved p over the included code (trial idx uses p[idx % 4]) and
    the seeded hypergraph-product code HGP(150 x 200, (3,4)-regular),
    n = 62,500, on the HBM-resident decoder (slots autotuned first): each as
    its own call, both codes in one engine, and both codes' own engines at
    once from two host threads."""
    out = {}
    n_qc = 4 * args.mixed_trials
    t0 = time.perf_counter()
    st, r = qc_dec.simulate(args.seed, n_qc, list(MIXED_PS), max_iter=args.mixed_max_iter,
                            first_trial=1, precision=32)
    el = time.perf_counter() - t0
    import numpy as np
    idx = np.arange(1, n_qc + 1)
    per_p = {}
    for j, p in enumerate(MIXED_PS):
        sel = idx % 4 == j
        k = int(np.sum(1 - r["success"][sel]))
        per_p["%g" % p] = {"trials": int(sel.sum()), "failures": k,
                           "avg_iter": round(float(r["iterations"][sel].mean()), 4)}
    out["qc_interleaved"] = {"trials": n_qc, "max_iter": args.mixed_max_iter,
                             "trials_per_s": round(n_qc / el, 2),
                             "decode_ms": round(st["device_ms"], 4), "per_p": per_p}
    code = jbp.Code.hgp(150, 200, 3, 4, seed=1)
    hdec = jbp.Decoder(code, 0)
    n_h = args.hgp_trials
    slots, _, cand = hdec.tune(args.seed + 1, n_h, args.hgp_p, max_iter=args.mixed_max_iter,
                               precision=32)
    t0 = time.perf_counter()
    st, _ = hdec.simulate(args.seed, n_h, args.hgp_p, max_iter=args.mixed_max_iter, first_trial=1,
                          precision=32, per_trial=False)
    el = time.perf_counter() - t0
    bytes_iter = 16 * (code.x_edges + code.z_edges)
    ach = st["total_iters"] * bytes_iter / (st["device_ms"] / 1e3) / 1e9
    lo, hi = wilson(st["done"], st["failures"])
    out["hgp"] = {"code": "HGP(150x200,(3,4)-regular, seed 1): n=%d, %d+%d checks, %d edges/side"
                          % (code.n, code.mx, code.mz, code.x_edges),
                  "p": args.hgp_p, "trials": st["done"], "max_iter": args.mixed_max_iter,
                  "trials_per_s": round(st["done"] / el, 2),
                  "trial_iterations_per_s": round(st["total_iters"] / el, 1),
                  "avg_iter": round(st["total_iters"] / st["done"], 4),
                  "fer": st["failures"] / st["done"], "fer_ci95": [lo, hi],
                  "bp_failures": st["bp_failures"], "tuned_slots": slots,
                  "tune_ms": [round(x, 3) for x in cand],
                  "roofline": {"bound": "hbm", "achieved": round(ach, 2), "peak": HBM_PEAK_GBS,
                               "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4),
                               "bytes_per_trial_iteration": bytes_iter,
                               "kernels": "gen_var_kernel + gen_check_kernel (+ book/emit/load)"}}
    sep_ms = 1e3 * (out["qc_interleaved"]["trials"] / out["qc_interleaved"]["trials_per_s"] +
                    n_h / out["hgp"]["trials_per_s"])
    # both codes in one engine (jbp_simulate_mixed): every step's kernels
    # cover the included code's trials (HBM engine, --mixed-qc-slots) and
    # the HGP trials (tuned slots); one untimed run first
    runs = [(qc_dec, dict(base_seed=args.seed, n_trials=n_qc, p=list(MIXED_PS),
                          max_iter=args.mixed_max_iter, slots=args.mixed_qc_slots)),
            (hdec, dict(base_seed=args.seed, n_trials=n_h, p=args.hgp_p, max_iter=args.mixed_max_iter,
                        slots=slots))]
    jbp.simulate_mixed(runs, per_trial=False)
    t0 = time.perf_counter()
    (sq, _), (sh, _) = jbp.simulate_mixed(runs, per_trial=False)
    el = time.perf_counter() - t0
    out["one_engine"] = {"trials": n_qc + n_h, "trials_per_s": round((n_qc + n_h) / el, 2),
                         "ms": round(1e3 * el, 2), "device_ms": round(sq["device_ms"], 2),
                         "qc_slots": args.mixed_qc_slots, "hgp_slots": slots,
                         "qc_failures": sq["failures"], "hgp_failures": sh["failures"],
                         "separate_calls_ms": round(sep_ms, 2)}
    # the two codes' own engines at once: one host thread per context
    import threading
    box = {}

    def go(key, fn):
        t = time.perf_counter()
        fn()
        box[key] = time.perf_counter() - t
    ths = [threading.Thread(target=go, args=("qc", lambda: qc_dec.simulate(
               args.seed, n_qc, list(MIXED_PS), max_iter=args.mixed_max_iter, first_trial=1,
               precision=32, per_trial=False))),
           threading.Thread(target=go, args=("hgp", lambda: hdec.simulate(
               args.seed, n_h, args.hgp_p, max_iter=args.mixed_max_iter, first_trial=1, precision=32,
               per_trial=False)))]
    t0 = time.perf_counter()
    for t in ths:
        t.start()
    for t in ths:
        t.join()
    el = time.perf_counter() - t0
    out["concurrent_engines"] = {"trials": n_qc + n_h, "trials_per_s": round((n_qc + n_h) / el, 2),
                                 "ms": round(1e3 * el, 2),
                                 "qc_ms": round(1e3 * box["qc"], 2), "hgp_ms": round(1e3 * box["hgp"], 2)}
    hdec.destroy()
    code.close()
    return out


def bench_hgp(jbp, args):
    """Steps of --trials trials of the hypergraph-product code at --hgp-p,
    max-iter --mixed-max-iter, on the generic decoder after one untimed
    tuning pass (resident slots)."""
    code = jbp.Code.hgp(150, 200, 3, 4, seed=1)
    dec = jbp.Decoder(code, 0)
    trials = args.hgp_trials
    slots, _, cand = dec.tune(args.seed + 1, trials, args.hgp_p, max_iter=args.mixed_max_iter,
                              precision=32)
    nxt = 1
    for _ in range(args.warmup):
        dec.simulate(args.seed, trials, args.hgp_p, max_iter=args.mixed_max_iter, first_trial=nxt,
                     precision=32, per_trial=False)
        nxt += trials
    done = fails = iters = 0
    dev = 0.0
    t0 = time.perf_counter()
    for _ in range(args.steps):
        st, _ = dec.simulate(args.seed, trials, args.hgp_p, max_iter=args.mixed_max_iter,
                             first_trial=nxt, precision=32, per_trial=False)
        nxt += trials
        done += st["done"]
        fails += st["failures"]
        iters += st["total_iters"]
        dev += st["device_ms"]
    el = time.perf_counter() - t0
    bytes_iter = 16 * (code.x_edges + code.z_edges)
    ach = iters * bytes_iter / (dev / 1e3) / 1e9
    lo, hi = wilson(done, fails)
    line = {
        "metric": "decoded trials/s, HGP(150x200) code, p=%g" % args.hgp_p,
        "value": round(done / el, 2), "unit": "trials/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(1e3 * el / args.steps, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "fp32",
        "data": "synthetic",
        "config": {"workload": "plain joint BP, HGP(150x200,(3,4)-regular, seed 1), n=%d, p=%g, "
                               "%d trials per step, max-iter %d, generic HBM-resident decoder, "
                               "%d resident slots (tuned)" % (code.n, args.hgp_p, trials,
                                                              args.mixed_max_iter, slots),
                   "avg_iter": round(iters / done, 4), "fer": fails / done, "fer_ci95": [lo, hi],
                   "trial_iterations_per_s": round(iters / el, 1), "tune_ms": cand},
        "roofline": {"bound": "hbm", "achieved": round(ach, 2), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": None,
                     "kernel": "gen_check_kernel + gen_var_kernel (one engine step each)",
                     "bytes_per_trial_iteration": bytes_iter},
        "cpu_baseline": None,
    }
    dec.destroy()
    print(json.dumps(line), flush=True)


def copy_calibration(mib):
    """One hipMemcpy device-to-device of `mib` MiB (the runtime's copy kernel)."""
    hip = C.CDLL("libamdhip64.so")
    nbytes = mib << 20
    a, b = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(a), C.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(C.byref(b), C.c_size_t(nbytes)) == 0
    assert hip.hipMemset(a, 1, C.c_size_t(nbytes)) == 0
    assert hip.hipMemcpy(b, a, C.c_size_t(nbytes), 3) == 0  # device to device
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)
    print("copy calibration: %d MiB device-to-device" % mib, file=sys.stderr)


def ms_kernel_name(model):
    """Kernel instance that runs a learned min-sum model on the included code:
    decode_qc_lms_kernel<P, J, L> (layered) or decode_qc_ms_kernel<P, J, L,
    DAMPED> (flooding)."""
    if model.kind == 2:
        return "decode_qc_lms_kernel<768, 3, 12>"
    return "decode_qc_ms_kernel<768, 3, 12, %s>" % ("true" if model.use_mlp else "false")


def profile_run(jbp, args):
    """The workload rocprofv3 counter passes run (--profile-run FILE):
    --roofline-steps single launches of --chunk trials of the learned
    decoder (--weights), of the second learned decoder (--compare-weights)
    and of plain BP, each on the same indices.  Writes the trial iterations
    each kernel did, which the counter totals are divided by."""
    from joint_bp_plus_pp_amd import learned as jl
    code = jbp.Code.from_params(PARAMS)
    dec = jbp.Decoder(code, 0)
    models = [("ms", jl.load(dec, args.weights), args.weights)]
    if args.compare_weights:
        models.append(("ms_other", jl.load(dec, args.compare_weights), args.compare_weights))
    n, R = args.chunk, args.roofline_steps
    its = {key: 0 for key, _, _ in models}
    fb_it = pl_it = 0
    for k in range(R):
        for key, model, _ in models:
            _, r = dec.simulate_learned(model, args.seed, n, args.p, max_iter=args.max_iter,
                                        first_trial=1 + k * n, chunk=n)
            its[key] += int(np.minimum(r["iterations"], model.T).sum())
            fb_it += int(np.maximum(r["iterations"] - model.T, 0).sum())
        st, _ = dec.simulate(args.seed, n, args.p, max_iter=args.max_iter, first_trial=1 + k * n,
                             precision=32, per_trial=False, chunk=n)
        pl_it += st["total_iters"]
    out = {"launches": R, "trials_per_launch": n, "p": args.p, "max_iter": args.max_iter,
           "plain_trial_iterations": pl_it + fb_it, "fallback_trial_iterations": fb_it,
           "learned": {key: {"weights": os.path.relpath(path, ROOT), "kind": int(model.kind),
                             "kernel": ms_kernel_name(model), "trial_iterations": its[key]}
                       for key, model, path in models}}
    with open(args.profile_run, "w") as f:
        json.dump(out, f)
    for _, model, _ in models:
        model.destroy()
    dec.destroy()
    print("profile run: %s" % json.dumps(out), flush=True)


def summarize_counters(args):
    """Per-kernel totals of rocprofv3 counter CSVs over a --profile-run
    workload, per trial-iteration, into a JSON summary (--summarize OUT)
    stamped with the kernel source's SHA-256.  FETCH_SIZE and WRITE_SIZE are
    scaled by what they read for a device-to-device copy of known size (a
    dispatch of the runtime's copy kernel in the same run)."""
    import csv
    import collections
    with open(args.profile_log) as f:
        log = json.load(f)
    kernels = {key: d["kernel"] for key, d in log["learned"].items()}
    kernels["plain"] = "decode_qc_plain_kernel<768, 3, 12>"
    iters = {key: d["trial_iterations"] for key, d in log["learned"].items()}
    iters["plain"] = log["plain_trial_iterations"]
    tot = {key: collections.defaultdict(float) for key in kernels}
    passes = collections.defaultdict(set)  # the CSVs that hold each counter
    disp = {key: set() for key in kernels}
    copy = collections.defaultdict(list)
    for path in args.counters:
        for r in csv.DictReader(open(path)):
            name, c, v = r["Kernel_Name"], r["Counter_Name"], float(r["Counter_Value"])
            key = next((k for k, kn in kernels.items() if kn in name), None)
            if key is None:
                if "copyBuffer" in name:
                    copy[c].append(v)
                continue
            tot[key][c] += v
            passes[c].add(path)
            disp[key].add((path, r["Dispatch_Id"]))
    out = {"source": os.path.relpath(KERNEL_SOURCE, ROOT), "source_sha256": source_sha256(KERNEL_SOURCE),
           "workload": log, "command": args.command}
    # copy calibration: the largest copy dispatch moved copy_bytes each way
    scale = {}
    cb = args.copy_calibration_mb << 20
    for c in ("FETCH_SIZE", "WRITE_SIZE"):
        if copy.get(c) and cb:
            scale[c] = cb / (max(copy[c]) * 1024.0)
    for key, its in iters.items():
        # a counter collected in several passes: the mean over those passes
        t = {c: v / len(passes[c]) for c, v in tot[key].items()}
        d = {"kernel": kernels[key],
             "dispatches_per_pass": len(disp[key]) // max(1, len(args.counters)),
             "trial_iterations": its}
        for c, v in sorted(t.items()):
            d[c + "_per_trial_iteration"] = v / its
        slots = t.get("SQ_INSTS_VALU", 0.0) + t.get("SQ_INSTS_VALU_TRANS_F32", 0.0)
        d["valu_slots_per_trial_iteration"] = round(slots / its, 1)
        launches = log["launches"]
        if "FETCH_SIZE" in t and "WRITE_SIZE" in t and scale:
            # counters in KiB; scaled by the copy calibration
            d["hbm_bytes_per_launch"] = round(1024.0 * (t["FETCH_SIZE"] * scale["FETCH_SIZE"] +
                                                        t["WRITE_SIZE"] * scale["WRITE_SIZE"]) / launches)
        out[key] = d
    out["copy_scale"] = scale
    with open(args.summarize, "w") as f:
        json.dump(out, f, indent=1)
    print("wrote %s" % args.summarize, flush=True)


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--workload", choices=("qc", "hgp"), default="qc",
                    help="qc: the included code (default, the headline); hgp: the seeded "
                         "hypergraph-product code HGP(150x200) on the HBM-resident decoder")
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--trials", type=int, default=65536, help="trials per step")
    ap.add_argument("--steps-per-call", type=int, default=0,
                    help="steps per simulate call in the timed region (0: all steps in one call)")
    ap.add_argument("--weights", default=LEARNED_LMS_FILE,
                    help="learned decoder of the headline (default: the layered min-sum model)")
    ap.add_argument("--compare-weights", default=LEARNED_MS_FILE,
                    help="second learned decoder timed on the same trials ('' to skip)")
    ap.add_argument("--chunk", type=int, default=16384,
                    help="trials per persistent launch; launches alternate two streams")
    ap.add_argument("--roofline-steps", type=int, default=20,
                    help="single-launch steps of each decoder timed for the roofline")
    ap.add_argument("--p", type=float, default=0.04)
    ap.add_argument("--max-iter", type=int, default=1000)
    ap.add_argument("--seed", type=int, default=106)
    ap.add_argument("--cpu-trials", type=int, default=2048)
    ap.add_argument("--cpu-trials-1core", type=int, default=128,
                    help="trials of the CPU baseline also timed on one thread (0: skip)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--sp-learned-steps", type=int, default=0,
                    help="steps of the sum-product learned models, for comparison (0: skip)")
    ap.add_argument("--train-steps", type=int, default=2,
                    help="timed training steps of the min-sum model (0: skip)")
    ap.add_argument("--train-batch", type=int, default=4096)
    ap.add_argument("--train-only", action="store_true",
                    help="run only the training leg (for profiling its kernels)")
    ap.add_argument("--mixed-trials", type=int, default=4096, help="trials per p of the mixed run")
    ap.add_argument("--mixed-max-iter", type=int, default=200)
    ap.add_argument("--mixed-qc-slots", type=int, default=1024,
                    help="resident trials of the included code in the one-engine mixed run")
    ap.add_argument("--hgp-trials", type=int, default=2048)
    ap.add_argument("--hgp-p", type=float, default=0.03)
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--pp-trials", type=int, default=8192,
                    help="trials per call of the post-processing workload (0: skip it)")
    ap.add_argument("--pp-calls", type=int, default=3)
    ap.add_argument("--pp-max-iter", type=int, default=8)
    ap.add_argument("--pp-cpu-trials", type=int, default=128,
                    help="trials of the reference's PP loop timed beside it (0: skip)")
    ap.add_argument("--profile-run", default="",
                    help="run only the counter-profiling workload and write its trial iterations "
                         "to this JSON file")
    ap.add_argument("--summarize", default="",
                    help="write a per-kernel counter summary (with --counters, --profile-log)")
    ap.add_argument("--counters", nargs="*", default=[], help="rocprofv3 counter_collection CSVs")
    ap.add_argument("--profile-log", default="", help="the --profile-run JSON of those passes")
    ap.add_argument("--command", default="", help="command recorded in the summary")
    ap.add_argument("--copy-calibration-mb", type=int, default=0,
                    help="also run one device-to-device copy of this many MiB (a plain copy "
                         "kernel of known size, to scale profiler byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: trials shard over GPUs by index range, and the "
              "figure here is one GPU's share (got --gpus %d)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    if args.summarize:
        return summarize_counters(args)

    import joint_bp_plus_pp_amd as jbp

    if args.copy_calibration_mb > 0:
        copy_calibration(args.copy_calibration_mb)
    if args.profile_run:
        return profile_run(jbp, args)
    if args.train_only:
        code = jbp.Code.from_params(PARAMS)
        dec = jbp.Decoder(code, 0)
        print(json.dumps(training_leg(jbp, dec, args)), flush=True)
        dec.destroy()
        return
    if args.workload == "hgp":
        return bench_hgp(jbp, args)
    from joint_bp_plus_pp_amd import learned as jl
    code = jbp.Code.from_params(PARAMS)
    dec = jbp.Decoder(code, 0)
    info = dec.info()
    model = jl.load(dec, args.weights)
    other = jl.load(dec, args.compare_weights) if args.compare_weights else None
    T = model.T
    N = args.trials
    # warm-up: both decoders, indices outside the timed range
    wfirst = 1 + (args.steps + 8) * N
    for w in range(args.warmup):
        dec.simulate_learned(model, args.seed, N, args.p, max_iter=args.max_iter,
                             first_trial=wfirst + w * N, per_trial=False, chunk=args.chunk)
        dec.simulate(args.seed, N, args.p, max_iter=args.max_iter, first_trial=wfirst + w * N,
                     precision=32, per_trial=False, chunk=args.chunk)

    # the timed steps run as calls of --steps-per-call steps each (default:
    # one call over all of them, as the reference's --simulate --trials runs
    # its whole trial count in one loop)
    per_call = args.steps if args.steps_per_call <= 0 else min(args.steps_per_call, args.steps)

    def run(learned, m=None):
        done = fails = iters = 0
        nxt = 1
        t0 = time.perf_counter()
        left = args.steps
        while left > 0:
            k = min(per_call, left)
            if learned:
                st, _ = dec.simulate_learned(m or model, args.seed, k * N, args.p,
                                             max_iter=args.max_iter, first_trial=nxt,
                                             per_trial=False, chunk=args.chunk)
            else:
                st, _ = dec.simulate(args.seed, k * N, args.p, max_iter=args.max_iter, first_trial=nxt,
                                     precision=32, per_trial=False, chunk=args.chunk)
            nxt += k * N
            left -= k
            done += st["done"]
            fails += st["failures"]
            iters += st["total_iters"]
        return done, fails, iters, time.perf_counter() - t0

    l_done, l_fails, l_iters, l_el = run(True)
    p_done, p_fails, p_iters, p_el = run(False)
    compare = None
    if other is not None:
        for w in range(args.warmup):
            dec.simulate_learned(other, args.seed, N, args.p, max_iter=args.max_iter,
                                 first_trial=wfirst + w * N, per_trial=False, chunk=args.chunk)
        o_done, o_fails, o_iters, o_el = run(True, other)
        compare = {"weights": os.path.relpath(args.compare_weights, ROOT),
                   "decoder": LEARNED_KINDS[other.kind], "trials": o_done,
                   "trials_per_s": round(o_done / o_el, 2), "avg_iter": round(o_iters / o_done, 4),
                   "fer": o_fails / o_done, "fer_ci95": list(wilson(o_done, o_fails)),
                   "speedup_vs_plain": round((o_done / o_el) / (p_done / p_el), 4)}
        print("bench: compare %s" % json.dumps(compare), file=sys.stderr)
    value = l_done / l_el
    llo, lhi = wilson(l_done, l_fails)
    plo, phi = wilson(p_done, p_fails)
    # matched: the learned FER is no worse than plain BP's (at most the upper
    # end of plain's 95% interval); below: under the interval's lower end
    matched = l_fails / l_done <= phi
    below = l_fails / l_done < plo
    print("bench: learned %d trials %.3f s (%.0f/s, avg_iter %.3f, FER %.3g [%.3g, %.3g]); plain "
          "%.3f s (%.0f/s, avg_iter %.3f, FER %.3g [%.3g, %.3g]); matched %s"
          % (l_done, l_el, value, l_iters / l_done, l_fails / l_done, llo, lhi, p_el, p_done / p_el,
             p_iters / p_done, p_fails / p_done, plo, phi, matched), file=sys.stderr)

    # roofline leg: single launches (one chunk per call, no stream overlap)
    # so that each kernel's HIP-event duration is its own
    rfirst = wfirst + args.warmup * N
    ms_ms, ms_iters, pl_ms, pl_iters, n_r = [], [], [], [], args.roofline_steps
    for k in range(n_r):
        st, r = dec.simulate_learned(model, args.seed, args.chunk, args.p, max_iter=args.max_iter,
                                     first_trial=rfirst + k * args.chunk, chunk=args.chunk)
        ms_ms.append(dec.last_kernel_ms(4)[0])
        ms_iters.append(int(np.minimum(r["iterations"], T).sum()))
        st, _ = dec.simulate(args.seed, args.chunk, args.p, max_iter=args.max_iter,
                             first_trial=rfirst + k * args.chunk, precision=32, per_trial=False,
                             chunk=args.chunk)
        pl_ms.append(dec.last_kernel_ms(1)[0])
        pl_iters.append(st["total_iters"])
    ms_rate = sum(ms_iters) / (sum(ms_ms) / 1e3)     # learned-kernel trial-iterations/s
    pl_rate = sum(pl_iters) / (sum(pl_ms) / 1e3)     # plain-kernel trial-iterations/s
    prof, why = kernel_profile()
    # the profile's entry for the headline decoder's kernel instance
    mine = None
    if prof:
        mine = next((d for k, d in prof.items() if isinstance(d, dict) and
                     d.get("kernel") == ms_kernel_name(model)), None)
        if mine is None:
            prof, why = None, "counter profile has no entry for %s" % ms_kernel_name(model)
    roofline = {"bound": "valu", "unit": "Gslots/s", "peak": VALU_PEAK_GSLOTS,
                "kernel": "%s (%s, FP16 pairs)" % (ms_kernel_name(model), LEARNED_KINDS[model.kind]),
                "kernel_ms_per_launch": round(sum(ms_ms) / n_r, 4), "launches": n_r,
                "trial_iterations_per_s": round(ms_rate, 1),
                "streaming_equivalent_gbs": round(ms_rate * BYTES_PER_ITER / 1e9, 1)}
    # the sustained (pipelined) plain rate against one launch at the median
    # launch time: what stragglers and the sampling/residual work around the
    # decoder cost the long run
    med_rate = args.chunk / (float(np.median(pl_ms)) / 1e3)
    plain_roof = {"kernel": "decode_qc_plain_kernel<768, 3, 12> (plain sum-product, FP32)",
                  "kernel_ms_per_launch": round(sum(pl_ms) / n_r, 4),
                  "median_launch_ms": round(float(np.median(pl_ms)), 4),
                  "max_launch_ms": round(float(np.max(pl_ms)), 4),
                  "median_launch_trials_per_s": round(med_rate, 1),
                  "sustained_over_median_launch": round((p_done / p_el) / med_rate, 4),
                  "trial_iterations_per_s": round(pl_rate, 1),
                  "streaming_equivalent_gbs": round(pl_rate * BYTES_PER_ITER / 1e9, 1)}
    if prof:
        v = mine["valu_slots_per_trial_iteration"]
        roofline.update({"achieved": round(v * ms_rate / 1e9, 2),
                         "frac": round(v * ms_rate / 1e9 / VALU_PEAK_GSLOTS, 4),
                         "valu_slots_per_trial_iteration": v,
                         "traffic": mine.get("hbm_bytes_per_launch"),
                         "profile": os.path.relpath(VALU_PROFILE, ROOT)})
        vp = prof["plain"]["valu_slots_per_trial_iteration"]
        plain_roof.update({"valu_achieved": round(vp * pl_rate / 1e9, 2),
                           "valu_frac": round(vp * pl_rate / 1e9 / VALU_PEAK_GSLOTS, 4),
                           "valu_slots_per_trial_iteration": vp,
                           "traffic": prof["plain"].get("hbm_bytes_per_launch")})
    else:
        roofline.update({"achieved": None, "frac": None, "traffic": None, "note": why})
    print("bench: roofline %s; plain %s" % (json.dumps(roofline), json.dumps(plain_roof)),
          file=sys.stderr)

    after = 1 + (args.steps + 8 + args.warmup) * N + n_r * args.chunk
    sp_learned = sp_learned_leg(jbp, dec, args, after) if args.sp_learned_steps > 0 else None
    training = None
    if args.train_steps > 0:
        training = training_leg(jbp, dec, args)
        print("bench: training %s" % json.dumps(training), file=sys.stderr)
    mixed = None
    if not args.no_mixed:
        mixed = mixed_workload(jbp, dec, args)
        print("bench: mixed %s" % json.dumps(mixed), file=sys.stderr)
    pp = None
    if args.pp_trials > 0:
        pp = pp_workload(jbp, dec, args, after)
        print("bench: pp %s" % json.dumps(pp), file=sys.stderr)
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(args, 1)
    line = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "trials/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * l_el / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp16",
        "data": "synthetic",
        "config": {
            "workload": "%s BP (T=%d, shipped weights) + plain FP32 BP fallback to "
                        "max-iter %d, included [[9216,4612]] code, p=%g, %d trials per step, "
                        "%d steps per simulate call, in launches of %d trials pipelined over "
                        "three streams, seed %d; plain FP32 BP on the same trial indices beside "
                        "it" % (LEARNED_KINDS[model.kind], T, args.max_iter, args.p, N, per_call,
                                args.chunk, args.seed),
            "decoder": "%s, FP16 (X, Z) message pairs, LDS-resident; fallback plain "
                       "FP32 sum-product" % LEARNED_KINDS[model.kind],
            "weights": os.path.relpath(args.weights, ROOT),
            "trials_per_step": N, "steps_per_call": per_call, "launch_trials": args.chunk, "p": args.p,
            "max_iter": args.max_iter, "T": T,
            "trials": l_done, "avg_iter": round(l_iters / l_done, 4),
            "fer": l_fails / l_done, "fer_ci95": [llo, lhi],
            "trial_iterations_per_s": round(l_iters / l_el, 1),
            "edge_updates_per_s": round(l_iters * EDGES_PER_ITER / l_el, 1),
            "plain_same_trials": {"trials": p_done, "trials_per_s": round(p_done / p_el, 2),
                                  "avg_iter": round(p_iters / p_done, 4), "fer": p_fails / p_done,
                                  "fer_ci95": [plo, phi],
                                  "trial_iterations_per_s": round(p_iters / p_el, 1),
                                  "roofline": plain_roof},
            "speedup_vs_plain": round(value / (p_done / p_el), 4),
            "other_learned_same_trials": compare,
            "fer_matched": matched,
            "fer_below_plain_ci": below,
            "sp_learned": sp_learned,
            "training": training,
            "mixed": mixed,
            "pp": pp,
        },
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    model.destroy()
    if other is not None:
        other.destroy()
    dec.destroy()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()

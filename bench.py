#!/usr/bin/env python3
"""Throughput of the --simulate --no-pp path on one MI355X.

Workload (the batched GPU configuration): the included girth-8 (3,12)-regular
[[9216,4612]] code, i.i.d. depolarizing noise at p = 0.04, 8192 seeded trials
per step = one persistent decoder launch (each compute unit pulls the next
trial when its current one stops), max-iter 1000 with per-trial early
termination, FP32 messages.  A step samples its trials on the device
(seed = splitmix64(base ^ idx*phi), mt19937_64), computes syndromes,
decodes, and runs the stabilizer-residual test; consecutive steps take
consecutive trial indices.  config.pipelined reports calls of 65,536 trials
(eight launches alternating two streams, the next launch's trials sampled
while the current one decodes); their launches overlap, so the roofline is
taken from the one-launch steps.

Prints one JSON line (last line of stdout).  `value` is decoded trials/s over
the timed steps.  `roofline` is for the decoder kernel: algorithmic bytes
(16 B per edge per trial-iteration, 884,736 B per trial-iteration) over its
HIP-event duration.  `cpu_baseline` times the reference decoder (built from
its own source under oracle/_ref/) on a bounded sample of the same workload
on this host's cores.
"""
import argparse
import ctypes as C
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
PARAMS = os.path.join(ROOT, "joint_bp_plus_pp_amd", "codes",
                      "H_P768_J3_L12_dmax3_nc0-3_1-2_seed11579811919164041.txt")
METRIC = "decoded trials/s at p=0.04 at matched FER (learned vs plain BP), % HBM peak"
EDGES_PER_ITER = 2 * 27648
BYTES_PER_ITER = 16 * EDGES_PER_ITER  # 884,736
HBM_PEAK_GBS = 8000.0
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "decode_qc_f32_p0.04_traffic.json")
LEARNED_FILES = [os.path.join(ROOT, "joint_bp_plus_pp_amd", "codes", f)
                 for f in ("learned_T25_P768_J3_L12.npz", "learned_T25_nomlp_P768_J3_L12.npz")]
MIXED_PS = (0.01, 0.025, 0.04, 0.05)


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


def learned_vs_plain(jbp, dec, args, first, plain):
    """Each shipped learned model (T learned iterations, then plain BP
    restarted from the prior) on the first timed trial indices, beside plain
    BP on the same indices."""
    from joint_bp_plus_pp_amd import learned as jl
    p_done, p_fails, p_el = plain
    plo, phi = wilson(p_done, p_fails)
    out = {"plain_same_trials": {"trials": p_done, "trials_per_s": round(p_done / p_el, 2),
                                 "fer": p_fails / p_done, "fer_ci95": [plo, phi]}}
    for path in LEARNED_FILES:
        if not os.path.exists(path):
            continue
        model = jl.load(dec, path)
        nxt, done, fails, iters, dev = first, 0, 0, 0, 0.0
        l0 = dec.last_kernel_ms(1)[1]
        t0 = time.perf_counter()
        for _ in range(args.learned_steps):
            st, _ = dec.simulate_learned(model, args.seed, args.trials, args.p,
                                         max_iter=args.max_iter, first_trial=nxt, per_trial=False)
            nxt += args.trials
            done += st["done"]
            fails += st["failures"]
            iters += st["total_iters"]
            dev += st["device_ms"]
        el = time.perf_counter() - t0
        nl = max(1, dec.last_kernel_ms(1)[1] - l0)
        lo, hi = wilson(done, fails)
        key = "mlp_damping" if model.use_mlp else "weights_only"
        out[key] = {"weights": os.path.relpath(path, ROOT), "T": model.T, "trials": done,
                    "trials_per_s": round(done / el, 2), "fer": fails / done, "fer_ci95": [lo, hi],
                    "avg_iter": round(iters / done, 4), "decode_ms_per_launch": round(dev / nl, 4),
                    "speedup_vs_plain": round((done / el) / (p_done / p_el), 4)}
        model.destroy()
    return out if len(out) > 1 else None


def mixed_workload(jbp, qc_dec, args):
    """Interleaved p over the included code (trial idx uses p[idx % 4]) and
    the seeded hypergraph-product code HGP(150 x 200, (3,4)-regular),
    n = 62,500, on the HBM-resident decoder (slots autotuned first)."""
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


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--workload", choices=("qc", "hgp"), default="qc",
                    help="qc: the included code (default, the headline); hgp: the seeded "
                         "hypergraph-product code HGP(150x200) on the HBM-resident decoder")
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--trials", type=int, default=8192,
                    help="trials per step: one persistent decoder launch")
    ap.add_argument("--pipelined-steps", type=int, default=8,
                    help="also time this many calls of 65,536 trials (eight launches each, "
                         "pipelined over two streams); 0 = skip")
    ap.add_argument("--p", type=float, default=0.04)
    ap.add_argument("--max-iter", type=int, default=1000)
    ap.add_argument("--seed", type=int, default=106)
    ap.add_argument("--precision", type=int, default=32, choices=(32, 64))
    ap.add_argument("--cpu-trials", type=int, default=2048)
    ap.add_argument("--cpu-trials-1core", type=int, default=128,
                    help="trials of the CPU baseline also timed on one thread (0: skip)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--learned-steps", type=int, default=40,
                    help="steps of the learned decoder on the first timed trial indices (0: skip)")
    ap.add_argument("--mixed-trials", type=int, default=4096, help="trials per p of the mixed run")
    ap.add_argument("--mixed-max-iter", type=int, default=200)
    ap.add_argument("--hgp-trials", type=int, default=2048)
    ap.add_argument("--hgp-p", type=float, default=0.03)
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--pp-trials", type=int, default=8192,
                    help="trials per call of the post-processing workload (0: skip it)")
    ap.add_argument("--pp-calls", type=int, default=3)
    ap.add_argument("--pp-max-iter", type=int, default=8)
    ap.add_argument("--pp-cpu-trials", type=int, default=128,
                    help="trials of the reference's PP loop timed beside it (0: skip)")
    ap.add_argument("--copy-calibration-mb", type=int, default=0,
                    help="also run one device-to-device copy of this many MiB (a plain copy "
                         "kernel of known size, to scale profiler byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: the decoder shards trials, and this build "
              "measures the single-GPU share (got --gpus %d)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import joint_bp_plus_pp_amd as jbp

    if args.copy_calibration_mb > 0:
        copy_calibration(args.copy_calibration_mb)
    if args.workload == "hgp":
        return bench_hgp(jbp, args)
    code = jbp.Code.from_params(PARAMS)
    dec = jbp.Decoder(code, 0)
    info = dec.info()
    nxt = 1
    for _ in range(args.warmup):
        dec.simulate(args.seed, args.trials, args.p, max_iter=args.max_iter, first_trial=nxt,
                     precision=args.precision, per_trial=False)
        nxt += args.trials
    first_timed = nxt
    launches0 = dec.last_kernel_ms(1)[1]
    done = fails = iters = 0
    dev_ms = []
    step_iters = []
    head = [0, 0, 0.0]  # plain trials, failures, seconds over the first learned_steps steps
    t0 = time.perf_counter()
    for step in range(args.steps):
        ts = time.perf_counter()
        st, _ = dec.simulate(args.seed, args.trials, args.p, max_iter=args.max_iter, first_trial=nxt,
                             precision=args.precision, per_trial=False)
        nxt += args.trials
        done += st["done"]
        fails += st["failures"]
        iters += st["total_iters"]
        dev_ms.append(st["device_ms"])
        step_iters.append(st["total_iters"])
        if step < args.learned_steps:
            head[0] += st["done"]
            head[1] += st["failures"]
            head[2] += time.perf_counter() - ts
    elapsed = time.perf_counter() - t0
    value = done / elapsed
    launches = max(1, dec.last_kernel_ms(1)[1] - launches0)
    kern_s = sum(dev_ms) / 1e3  # sum of the decoder launches' HIP-event durations
    # FP64: 64 B of message traffic per edge per iteration (pair sums)
    per_iter = BYTES_PER_ITER if args.precision == 32 else 64 * EDGES_PER_ITER
    achieved = sum(step_iters) * per_iter / kern_s / 1e9
    traffic = None
    if args.precision == 32 and os.path.exists(TRAFFIC_FILE):
        with open(TRAFFIC_FILE) as f:
            t = json.load(f)
        traffic = t.get("hbm_bytes_per_launch")
    lo, hi = wilson(done, fails)
    print("bench: %d trials in %.3f s; avg_iter %.3f; FER %.3g ci95 [%.3g, %.3g]; decode kernel "
          "%.3f ms/launch (%d launches); %.3g trial-iter/s; info %s"
          % (done, elapsed, iters / done, fails / done, lo, hi, sum(dev_ms) / launches, launches,
             iters / elapsed, info), file=sys.stderr)
    pipelined = None
    if args.pipelined_steps > 0 and args.precision == 32:
        pn = 65536
        pdone = pfails = piters = 0
        tp0 = time.perf_counter()
        for _ in range(args.pipelined_steps):
            st, _ = dec.simulate(args.seed, pn, args.p, max_iter=args.max_iter, first_trial=nxt,
                                 precision=args.precision, per_trial=False)
            nxt += pn
            pdone += st["done"]
            pfails += st["failures"]
            piters += st["total_iters"]
        pel = time.perf_counter() - tp0
        plo, phi = wilson(pdone, pfails)
        pipelined = {"trials_per_call": pn, "calls": args.pipelined_steps,
                     "trials_per_s": round(pdone / pel, 2),
                     "trial_iterations_per_s": round(piters / pel, 1),
                     "fer": pfails / pdone, "fer_ci95": [plo, phi]}
        print("bench: pipelined %s" % json.dumps(pipelined), file=sys.stderr)
    learned = None
    if args.learned_steps > 0 and args.precision == 32:
        learned = learned_vs_plain(jbp, dec, args, first_timed, head)
        if learned:
            print("bench: learned %s" % json.dumps(learned), file=sys.stderr)
    mixed = None
    if not args.no_mixed and args.precision == 32:
        mixed = mixed_workload(jbp, dec, args)
        print("bench: mixed %s" % json.dumps(mixed), file=sys.stderr)
    pp = None
    if args.pp_trials > 0 and args.precision == 32:
        pp = pp_workload(jbp, dec, args, nxt)
        print("bench: pp %s" % json.dumps(pp), file=sys.stderr)
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(args, first_timed)
    line = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "trials/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * elapsed / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32" if args.precision == 32 else "f64",
        "data": "synthetic",
        "config": {
            "workload": "plain joint BP, batched, included [[9216,4612]] code, p=%g, %d trials per "
                        "step (one persistent launch), max-iter %d, early termination, seed %d"
                        % (args.p, args.trials, args.max_iter, args.seed),
            "decoder": "plain FP32 LDS-resident" if args.precision == 32 else "plain FP64 parity",
            "trials_per_step": args.trials,
            "p": args.p,
            "max_iter": args.max_iter,
            "avg_iter": round(iters / done, 4),
            "fer": fails / done,
            "fer_ci95": [lo, hi],
            "trial_iterations_per_s": round(iters / elapsed, 1),
            "edge_updates_per_s": round(iters * EDGES_PER_ITER / elapsed, 1),
            "pipelined": pipelined,
            "learned": learned,
            "mixed": mixed,
            "pp": pp,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 2),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "kernel": "decode_qc_f32_kernel" if args.precision == 32 else "decode_f64_kernel",
            "kernel_ms_per_launch": round(sum(dev_ms) / launches, 4),
            "launches": launches,
        },
        "cpu_baseline": cpu,
    }
    dec.destroy()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()

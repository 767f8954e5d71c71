#!/usr/bin/env python3
"""Throughput of the --simulate --no-pp path on one MI355X.

Workload (the batched GPU configuration): the included girth-8 (3,12)-regular
[[9216,4612]] code, i.i.d. depolarizing noise at p = 0.04, 8192 seeded trials
per step (one persistent decoder launch; each compute unit pulls the next
trial when its current one stops), max-iter 1000 with per-trial early
termination, FP32 messages.  A step samples its trials on the device
(seed = splitmix64(base ^ idx*phi), mt19937_64), computes syndromes,
decodes, and runs the stabilizer-residual test; consecutive steps take
consecutive trial indices.

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


def wilson(n, k, z=1.96):
    if n <= 0:
        return 0.0, 0.0
    phat = k / n
    denom = 1 + z * z / n
    center = (phat + z * z / (2 * n)) / denom
    half = z * math.sqrt((phat * (1 - phat) + z * z / (4 * n)) / n) / denom
    return max(0.0, center - half), min(1.0, center + half)


def cpu_threads():
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    return max(1, min(16, avail))


def cpu_baseline(args, first_trial):
    """Reference decoder (or, without it, the oracle restatement) on a bounded
    sample of the workload: the first `cpu_trials` trial indices."""
    threads = cpu_threads()
    count = args.cpu_trials
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
    iters = sum(it)
    return {"value": round(count / secs, 3), "unit": "trials/s", "cores": threads, "kind": kind,
            "sample": "%d trials (indices %d..%d) of the same workload, p=%g, max-iter %d, FP64 "
                      "reference decoder, one trial per thread; %.1f s wall, avg_iter %.2f, "
                      "failures %d" % (count, first_trial, first_trial + count - 1, args.p,
                                       args.max_iter, secs, iters / count, count - sum(su))}


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
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--trials", type=int, default=8192, help="trials per step")
    ap.add_argument("--p", type=float, default=0.04)
    ap.add_argument("--max-iter", type=int, default=1000)
    ap.add_argument("--seed", type=int, default=106)
    ap.add_argument("--precision", type=int, default=32, choices=(32, 64))
    ap.add_argument("--cpu-trials", type=int, default=2048)
    ap.add_argument("--no-cpu", action="store_true")
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
    code = jbp.Code.from_params(PARAMS)
    dec = jbp.Decoder(code, 0)
    info = dec.info()
    nxt = 1
    for _ in range(args.warmup):
        dec.simulate(args.seed, args.trials, args.p, max_iter=args.max_iter, first_trial=nxt,
                     precision=args.precision, per_trial=False)
        nxt += args.trials
    first_timed = nxt
    done = fails = iters = 0
    dev_ms = []
    step_iters = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        st, _ = dec.simulate(args.seed, args.trials, args.p, max_iter=args.max_iter, first_trial=nxt,
                             precision=args.precision, per_trial=False)
        nxt += args.trials
        done += st["done"]
        fails += st["failures"]
        iters += st["total_iters"]
        dev_ms.append(st["device_ms"])
        step_iters.append(st["total_iters"])
    elapsed = time.perf_counter() - t0
    value = done / elapsed
    # decoder kernel: one launch per step (trials per step <= library chunk)
    kern_s = sum(dev_ms) / 1e3
    achieved = sum(step_iters) * BYTES_PER_ITER / kern_s / 1e9
    traffic = None
    if os.path.exists(TRAFFIC_FILE):
        with open(TRAFFIC_FILE) as f:
            t = json.load(f)
        traffic = t.get("hbm_bytes_per_launch")
    lo, hi = wilson(done, fails)
    print("bench: %d trials in %.3f s; avg_iter %.3f; FER %.3g ci95 [%.3g, %.3g]; decode kernel "
          "%.3f ms/launch; %.3g trial-iter/s; info %s" % (done, elapsed, iters / done, fails / done,
                                                        lo, hi, sum(dev_ms) / len(dev_ms),
                                                        iters / elapsed, info), file=sys.stderr)
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
                        "persistent launch, max-iter %d, early termination, seed %d"
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
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 2),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "kernel": "decode_qc_f32_kernel",
            "kernel_ms_per_launch": round(sum(dev_ms) / len(dev_ms), 4),
        },
        "cpu_baseline": cpu,
    }
    dec.destroy()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()

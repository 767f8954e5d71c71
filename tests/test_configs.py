"""GPU: the production FP32 paths at the sizes of the benchmarked workloads.

* plain FP32 on the included code at p = 0.04, max-iter 1000, 8,192-trial
  persistent launches (the throughput headline), judged against the FP64
  decoder, which is per-trial identical to the reference's CPU decoder, on the
  same 262,144 trial indices, and against the reference program's own
  aggregate over the first 2,000 of them (jointbp_ets.cpp:5931-6207; Wilson
  interval 4388-4399);
* the mixed-noise workload: p in {0.01, 0.025, 0.04, 0.05} interleaved over
  4 x 4,096 trials at max-iter 200, and the seeded hypergraph-product code
  HGP(150 x 200) at p = 0.03 with 2,048 trials at max-iter 200.

Tolerances: FP32 messages round differently from FP64 probabilities, so a
trial near a decision boundary can take another path; FER must lie inside the
FP64 run's 95% Wilson interval and per-trial success agree on at least 99.9%
of the trials (the HGP code: see its test)."""
import math

import numpy as np
import pytest

import joint_bp_plus_pp_amd as jbp
from conftest import golden_json

pytestmark = pytest.mark.gpu


def wilson(n, k, z=1.96):
    # jointbp_ets.cpp:4388-4399
    phat = k / n
    denom = 1 + z * z / n
    center = (phat + z * z / (2 * n)) / denom
    half = z * math.sqrt((phat * (1 - phat) + z * z / (4 * n)) / n) / denom
    return max(0.0, center - half), min(1.0, center + half)


def test_fp32_headline_workload_against_fp64_and_reference(main_decoder):
    n = 32 * 8192  # 262,144 trials: 32 persistent launches of 8,192
    print("fp32 p=0.04 max-iter 1000, %d trials" % n, flush=True)
    s32, r32 = main_decoder.simulate(106, n, 0.04, max_iter=1000, precision=jbp.FP32)
    print("fp64 same indices", flush=True)
    s64, r64 = main_decoder.simulate(106, n, 0.04, max_iter=1000, precision=jbp.FP64)
    assert s32["done"] == s64["done"] == n
    lo, hi = wilson(n, s64["failures"])
    fer32 = s32["failures"] / n
    agree = float(np.mean(r32["success"] == r64["success"]))
    same_it = float(np.mean(r32["iterations"] == r64["iterations"]))
    print("FER fp32 %.3g fp64 %.3g ci95 [%.3g, %.3g]; success agreement %.6f; same iterations "
          "%.4f; avg_iter fp32 %.4f fp64 %.4f"
          % (fer32, s64["failures"] / n, lo, hi, agree, same_it, s32["total_iters"] / n,
             s64["total_iters"] / n), flush=True)
    assert lo <= fer32 <= hi
    assert agree >= 0.999
    assert same_it >= 0.95
    assert abs(s32["total_iters"] - s64["total_iters"]) <= 0.01 * s64["total_iters"]
    # the reference program's aggregate over the first 2,000 indices: FP64
    # equals it exactly, FP32 within the tolerances above
    agg = golden_json("aggregates.json")["p0.04_it1000_seed106_2000"]
    m = agg["count"]
    assert (agg["base"], agg["first"], agg["p"], agg["max_iter"]) == (106, 1, 0.04, 1000)
    assert int(np.sum(1 - r64["success"][:m])) == agg["failures"]
    assert int(np.sum(r64["iterations"][:m])) == agg["iters"]
    hist = np.bincount(r64["iterations"][:m], minlength=1001)
    assert {str(k): int(v) for k, v in enumerate(hist) if v} == agg["iter_hist"]
    assert abs(int(np.sum(1 - r32["success"][:m])) - agg["failures"]) <= 1
    assert abs(int(np.sum(r32["iterations"][:m])) - agg["iters"]) <= 0.01 * agg["iters"]


PS = [0.01, 0.025, 0.04, 0.05]


def test_mixed_workload_interleaved_full_size(main_decoder):
    # 4 x 4,096 trials at max-iter 200, trial idx at p = PS[idx % 4]: every
    # trial equals the same index of a single-p run at its p (both through
    # the pipelined 8,192-trial launches)
    n, it = 4 * 4096, 200
    print("mixed interleaved", n, flush=True)
    _, mixed = main_decoder.simulate(106, n, PS, max_iter=it, first_trial=1, precision=jbp.FP32)
    idx = np.arange(1, n + 1)
    for j, p in enumerate(PS):
        _, single = main_decoder.simulate(106, n, p, max_iter=it, first_trial=1,
                                          precision=jbp.FP32)
        sel = idx % 4 == j
        for k in ("iterations", "syn_match", "success", "exact"):
            assert np.array_equal(mixed[k][sel], single[k][sel]), (p, k)
        print("p=%g: %d trials, failures %d, avg_iter %.3f" % (
            p, sel.sum(), int(np.sum(1 - mixed["success"][sel])),
            mixed["iterations"][sel].mean()), flush=True)


@pytest.fixture(scope="module")
def hgp150():
    code = jbp.Code.hgp(150, 200, 3, 4, seed=1)
    dec = jbp.Decoder(code, 0)
    yield code, dec
    dec.destroy()


def test_hgp_workload_full_size(hgp150):
    # HGP(150 x 200), n = 62,500, p = 0.03, 2,048 trials at max-iter 200 on
    # the HBM-resident FP32 engine, against FP64 on the first 512 indices.
    # On this girth-4 code the two arithmetics part ways (measured: FP32 fails
    # on about a third as many trials).  They differ at saturation: FP32
    # log-ratio messages are clamped at 64 bits, while the reference's FP64
    # probabilities keep growing until they underflow, and its normalisation
    # resets a message whose sum is zero to uniform (jointbp_ets.cpp:650-657).
    # Asserted: FP32 is never worse than the FP64 interval, FP32 decodes
    # nearly every trial FP64 decodes, and the engine's own invariants.
    code, dec = hgp150
    n, m, it = 2048, 512, 200
    print("hgp fp32", n, flush=True)
    s32, r32 = dec.simulate(1, n, 0.03, max_iter=it, precision=jbp.FP32)
    assert s32["done"] == n
    assert np.all(r32["iterations"][r32["syn_match"] == 0] == it)
    assert np.all(r32["success"] <= r32["syn_match"])
    print("hgp fp64", m, flush=True)
    s64, r64 = dec.simulate(1, m, 0.03, max_iter=it, precision=jbp.FP64)
    lo, hi = wilson(m, s64["failures"])
    fer32 = float(np.mean(1 - r32["success"][:m]))
    agree = float(np.mean(r32["success"][:m] == r64["success"]))
    both = float(np.mean(r32["success"][:m][r64["success"] == 1]))
    print("hgp FER fp32 %.4f (all %d: %.4f) fp64 %.4f ci95 [%.4f, %.4f]; agreement %.4f; fp32 "
          "success where fp64 succeeds %.4f" % (fer32, n, s32["failures"] / n, s64["failures"] / m,
                                               lo, hi, agree, both), flush=True)
    assert s32["failures"] / n <= hi
    assert both >= 0.98
    assert agree >= 0.9


def test_mixed_codes_in_one_engine_full_size(main_decoder, hgp150):
    # cfg3 in one engine: the included code's 4 x 4,096 interleaved-p trials
    # and HGP(150 x 200)'s 2,048 trials at p = 0.03, max-iter 200, decoded
    # together (every engine step's kernels cover both codes); every trial
    # equals the same trial of that code's own HBM-engine run
    code, hdec = hgp150
    n_qc, n_h, it = 4 * 4096, 2048, 200
    print("mixed codes, one engine", flush=True)
    (sq, rq), (sh, rh) = jbp.simulate_mixed([
        (main_decoder, dict(base_seed=106, n_trials=n_qc, p=PS, max_iter=it)),
        (hdec, dict(base_seed=1, n_trials=n_h, p=0.03, max_iter=it))])
    print("separate runs", flush=True)
    _, aq = main_decoder.simulate(106, n_qc, PS, max_iter=it, engine=jbp.ENGINE_HBM)
    _, ah = hdec.simulate(1, n_h, 0.03, max_iter=it, engine=jbp.ENGINE_HBM)
    for k in ("iterations", "syn_match", "success", "exact"):
        assert np.array_equal(rq[k], aq[k]), ("qc", k)
        assert np.array_equal(rh[k], ah[k]), ("hgp", k)
    assert sq["done"] == n_qc and sh["done"] == n_h
    assert sq["failures"] == int(np.sum(1 - rq["success"])) and sh["failures"] == int(np.sum(1 - rh["success"]))
    print("included code: failures %d avg_iter %.3f; HGP: failures %d avg_iter %.3f; one engine %.1f ms"
          % (sq["failures"], sq["total_iters"] / n_qc, sh["failures"], sh["total_iters"] / n_h,
             sq["device_ms"]), flush=True)

"""GPU: the layered learned min-sum decoder (kind LEARNED_LAYERED_MIN_SUM).

* training kernels (device forward + adjoint, layered schedule) against the
  FP64 oracle (oracle/learned_oracle.cpp kind 2, itself pinned by its
  one-layer reduction to the flooding min-sum kind, by brute-force
  max-product on trees and by central differences on smooth pieces), in
  FP64 and FP32;
* the packed-FP16 layered inference phase against the training forward
  pass (same model) on the same errors: the iteration each declares
  converged;
* trials the layered phase leaves unconverged finish as plain BP does;
* the shipped trained model on the benchmarked workload (p = 0.04, T = 25,
  16,384 trials per launch, plain FP32 BP as fallback): FER no worse than
  plain BP's 95% interval on the same trial indices;
* what training buys over the same decoder with hand-set weights (one
  constant normalisation; the schedule training starts from)."""
import math
import os

import numpy as np
import pytest

import joint_bp_plus_pp_amd as jbp
from joint_bp_plus_pp_amd import learned as jl

pytestmark = pytest.mark.gpu

SHIPPED = os.path.join(os.path.dirname(jbp.__file__), "codes", "learned_lms_T25_P768_J3_L12.npz")
LMS = jbp.LEARNED_LAYERED_MIN_SUM


def wilson(n, k, z=1.96):
    # jointbp_ets.cpp:4388-4399
    phat = k / n
    denom = 1 + z * z / n
    center = (phat + z * z / (2 * n)) / denom
    half = z * math.sqrt((phat * (1 - phat) + z * z / (4 * n)) / n) / denom
    return max(0.0, center - half), min(1.0, center + half)


@pytest.fixture(scope="module")
def lo(main_code_oracle):
    from _oracle import LearnedOracle
    return LearnedOracle(main_code_oracle)


def schedule_params(model, a0=0.45, a1=0.65, a=0.75):
    p = model.get_params()
    T, nb = model.T, model.nb
    wc = np.full((2, nb, T), a)
    wc[:, :, 0] = a0
    if T > 1:
        wc[:, :, 1] = a1
    p[:2 * nb * T] = wc.ravel()
    p[2 * nb * T:4 * nb * T] = 1.0
    return p


def test_fp64_loss_and_gradient_match_oracle(main_decoder, main_code_oracle, lo):
    T = 4
    model = jbp.LearnedModel(main_decoder, T, kind=LMS)
    rng = np.random.default_rng(5)
    params = model.get_params()
    k = 2 * model.nb * T
    params[:2 * k] = 0.8 + 0.1 * rng.standard_normal(2 * k)
    model.set_params(params, use_mlp=False)
    errs = np.stack([main_code_oracle.sample(106, t, 0.05) for t in (1, 2, 3, 4)])
    p = np.array([0.05, 0.04, 0.03, 0.045])
    loss, grad, conv = model.loss_grad(errs, p, precision=jbp.FP64)
    wc, wv, mlp = model.split(params)
    o_loss, (o_wc, o_wv, _), o_conv, _ = lo.loss(T, wc, wv, mlp, 0, errs, p, grad=True, kind=2)
    want = np.concatenate([o_wc.ravel(), o_wv.ravel(), np.zeros(65)])
    print("layered fp64 loss %.15g oracle %.15g; gradient relative difference %.2e; conv %s / %s"
          % (loss, o_loss, np.linalg.norm(grad - want) / np.linalg.norm(want), conv, o_conv), flush=True)
    assert abs(loss - o_loss) <= 1e-9 * abs(o_loss)
    assert np.linalg.norm(grad - want) <= 1e-7 * np.linalg.norm(want)
    assert np.array_equal(conv, o_conv)
    l32, g32, c32 = model.loss_grad(errs, p, precision=jbp.FP32)
    cos = g32 @ grad / (np.linalg.norm(g32) * np.linalg.norm(grad))
    assert abs(l32 - loss) <= 1e-4 * abs(loss) and cos > 0.9999
    with pytest.raises(jbp.JbpError):
        model.set_params(params, use_mlp=True)
    model.destroy()


def test_fp32_trainer_follows_fp64_at_t25(main_decoder, main_code_oracle):
    # the training workload's depth (T = 25) on 24 trials at p in
    # [0.03, 0.05], weights spread around the shipped schedule: the FP32
    # trainer (messages on chip, for this code) against the FP64 one, which
    # the oracle pins
    T = 25
    model = jbp.LearnedModel(main_decoder, T, kind=LMS)
    rng = np.random.default_rng(17)
    params = schedule_params(model)
    k = 2 * model.nb * T
    params[:2 * k] *= 1.0 + 0.05 * rng.standard_normal(2 * k)
    model.set_params(params, use_mlp=False)
    n = 24
    p = np.linspace(0.03, 0.05, n)
    errs = np.stack([main_code_oracle.sample(2024, 1 + t, float(p[t])) for t in range(n)])
    l64, g64, c64 = model.loss_grad(errs, p, precision=jbp.FP64)
    l32, g32, c32 = model.loss_grad(errs, p, precision=jbp.FP32)
    cos = g32 @ g64 / (np.linalg.norm(g32) * np.linalg.norm(g64))
    print("layered T=25: loss fp32 %.9g fp64 %.9g; gradient cosine %.7f; convergence agreement %.3f"
          % (l32, l64, cos, float(np.mean(c32 == c64))), flush=True)
    assert np.isfinite(l32) and np.all(np.isfinite(g32))
    assert abs(l32 - l64) <= 1e-4 * abs(l64)
    assert cos > 0.9999
    assert np.mean(c32 == c64) >= 0.9
    model.destroy()


def test_fp16_inference_follows_the_training_forward_pass(main_decoder):
    # the inference phase (FP16 pairs, base 2) and the trainer's forward pass
    # (FP32, natural log) are the same model with the same convergence rule:
    # the iteration each declares converged agrees on almost every trial
    T = 25
    model = jbp.LearnedModel(main_decoder, T, kind=LMS)
    model.set_params(schedule_params(model), use_mlp=False)
    n = 512
    err, _, _ = main_decoder.sample(106, 1, n, 0.04)
    _, _, conv = model.loss_grad(err, 0.04, precision=jbp.FP32, grad=False)
    _, r = main_decoder.simulate_learned(model, 106, n, 0.04, max_iter=200, want_err=True)
    assert np.array_equal(r["err"], err)
    it = r["iterations"]
    learned_conv = np.where(it <= T, it, 0)
    agree = float(np.mean(learned_conv == conv))
    print("convergence iteration agreement %.4f (mean %.3f vs %.3f)" % (
        agree, learned_conv.mean(), conv.mean()), flush=True)
    assert agree >= 0.97
    model.destroy()


@pytest.mark.parametrize("n,chunk", [(3 * 8192 + 5, 2048), (3000, 0)])
def test_trials_left_unconverged_finish_as_plain_bp(main_decoder, n, chunk):
    # T = 6 layered iterations leave about three quarters of the trials
    # unconverged at p = 0.04 (none converges before iteration 5 there, nor
    # before 7 at p = 0.05): each is decoded again by plain BP from the prior
    # and reports T plus plain BP's own count and plain BP's outcome
    T = 6
    model = jbp.LearnedModel(main_decoder, T, kind=LMS)
    model.set_params(schedule_params(model), use_mlp=False)
    print("layered fallback case n=%d chunk=%d" % (n, chunk), flush=True)
    _, rl = main_decoder.simulate_learned(model, 106, n, 0.04, max_iter=60, chunk=chunk)
    _, rp = main_decoder.simulate(106, n, 0.04, max_iter=60, precision=jbp.FP32)
    fb = rl["iterations"] > T
    print("%d of %d trials handed to plain BP" % (int(fb.sum()), n), flush=True)
    assert n // 10 < fb.sum() < n - n // 10
    assert np.array_equal(rl["iterations"][fb] - T, rp["iterations"][fb])
    for k in ("success", "syn_match", "exact"):
        assert np.array_equal(rl[k][fb], rp[k][fb]), k
    assert np.all(rl["syn_match"][~fb] == 1) and np.all(rl["iterations"][~fb] >= 1)
    model.destroy()


def test_training_step_at_full_size(main_decoder):
    # one step of the training workload: T = 25, batch 4,096, p ~ U[0.03, 0.05]
    model = jbp.LearnedModel(main_decoder, 25, kind=LMS)
    params = schedule_params(model)
    model.set_params(params, use_mlp=False)
    loss0 = model.train_step(2024, 1, 4096, 0.03, 0.05, lr=1e-2)
    after = model.get_params()
    print("layered full-size step loss %.6f, mean weight change %.3e" % (
        loss0, np.abs(after - params).mean()), flush=True)
    assert np.isfinite(loss0) and 0 < loss0 < 1
    assert np.all(np.isfinite(after)) and np.abs(after - params).max() > 0
    model.destroy()


@pytest.fixture(scope="module")
def shipped(main_decoder):
    if not os.path.exists(SHIPPED):
        pytest.skip("no shipped layered model")
    m = jl.load(main_decoder, SHIPPED)
    yield m
    m.destroy()


def test_shipped_model_fer_at_16384_per_launch(main_decoder, shipped):
    # p = 0.04, T = 25 learned iterations, then plain FP32 BP from the prior
    # up to max-iter 1000; 16,384 trials per launch, 16 launches
    model = shipped
    assert model.kind == LMS and model.T == 25
    n, per = 16 * 16384, 16384
    sl, rl = main_decoder.simulate_learned(model, 106, n, 0.04, max_iter=1000, chunk=per)
    sp, rp = main_decoder.simulate(106, n, 0.04, max_iter=1000, precision=jbp.FP32)
    lo_, hi_ = wilson(n, sp["failures"])
    fb = rl["iterations"] > 25
    print("layered: FER %.3g avg_iter %.3f; plain: FER %.3g ci95 [%.3g, %.3g] avg_iter %.3f; "
          "handed to plain BP: %d" % (sl["failures"] / n, sl["total_iters"] / n, sp["failures"] / n,
                                      lo_, hi_, sp["total_iters"] / n, int(fb.sum())), flush=True)
    assert sl["failures"] / n <= hi_
    assert np.array_equal(rl["iterations"][fb] - 25, rp["iterations"][fb])
    assert np.array_equal(rl["success"][fb], rp["success"][fb])
    assert np.all(rl["success"] <= rl["syn_match"])


def test_training_gain_over_hand_set_weights(main_decoder, shipped):
    # on held-out trials, the trained weights against (a) normalised
    # min-sum with one hand-set weight (wc 0.75, wv 1 everywhere) and (b)
    # the schedule training started from (wc 0.45, 0.65, then 0.75): training
    # must cut the mean iteration count well below (a) and not lose to (b)
    n = 16384
    first = 50_000_001
    const = jbp.LearnedModel(main_decoder, 25, kind=LMS)
    const.set_params(schedule_params(const, 0.75, 0.75, 0.75), use_mlp=False)
    sched = jbp.LearnedModel(main_decoder, 25, kind=LMS)
    sched.set_params(schedule_params(sched), use_mlp=False)
    out = {}
    for tag, m in (("constant 0.75", const), ("schedule", sched), ("trained", shipped)):
        s, r = main_decoder.simulate_learned(m, 106, n, 0.04, max_iter=1000, first_trial=first)
        out[tag] = s["total_iters"] / n
        print("%s: avg_iter %.4f, %d handed over, %d failures" % (
            tag, out[tag], int(np.sum(r["iterations"] > 25)), s["failures"]), flush=True)
    assert out["trained"] < 0.95 * out["constant 0.75"]
    assert out["trained"] <= 1.002 * out["schedule"]
    const.destroy()
    sched.destroy()


@pytest.mark.parametrize("p", [0.0, 0.01, 0.05, 0.07])
def test_layered_stops_on_estimates_that_match_the_syndrome(main_decoder, main_code_oracle, shipped, p):
    # The layered phase stops a trial only after a verification pass has
    # found every check accepting the estimate it returns: recompute the
    # syndromes of those estimates with the oracle.  Trials it hands over
    # finish as plain BP does; without noise every trial stops in its first
    # iteration with the exact (identity) estimate.
    n = 512
    print("layered decoder at p=%g" % p, flush=True)
    sl, rl = main_decoder.simulate_learned(shipped, 106, n, p, max_iter=200, want_err=True, want_est=True)
    sp, rp = main_decoder.simulate(106, n, p, max_iter=200, precision=jbp.FP32)
    fb = rl["iterations"] > shipped.T
    stopped = np.flatnonzero(~fb)
    print("stopped in the layered phase: %d of %d" % (stopped.size, n), flush=True)
    assert np.all(rl["syn_match"][stopped] == 1)
    for t in stopped:
        sx, sz = main_code_oracle.syndrome(rl["est"][t])
        ex, ez = main_code_oracle.syndrome(rl["err"][t])
        assert np.array_equal(sx, ex) and np.array_equal(sz, ez), t
    assert np.array_equal(rl["iterations"][fb] - shipped.T, rp["iterations"][fb])
    for k in ("success", "syn_match", "exact"):
        assert np.array_equal(rl[k][fb], rp[k][fb]), k
    if p == 0.0:
        assert np.all(rl["iterations"] == 1) and np.all(rl["exact"] == 1)

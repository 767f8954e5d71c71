"""GPU: the learned min-sum decoder (kind LEARNED_MIN_SUM).

* training kernels (device forward + adjoint, kind 1) against the FP64
  oracle (oracle/learned_oracle.cpp, itself pinned by brute-force max-product
  on a tree and by central differences on smooth pieces), in FP64 and FP32;
* the packed-FP16 inference phase of the LDS decoder against the training
  forward pass (same model, FP32) on the same errors;
* the shipped trained model on its benchmarked workload: p = 0.04, T = 25,
  16,384 trials per launch, plain FP32 BP as fallback: FER inside plain BP's
  95% interval on the same trial indices, and every trial it hands over
  finishing exactly as plain BP does on that trial;
* one training step at the training workload's size (T = 25, batch 4,096,
  p uniform in [0.03, 0.05]);
* what training buys: the trained model against the same decoder with unit
  weights (plain min-sum) on held-out trials."""
import math
import os

import numpy as np
import pytest

import joint_bp_plus_pp_amd as jbp
from joint_bp_plus_pp_amd import learned as jl

pytestmark = pytest.mark.gpu

SHIPPED = os.path.join(os.path.dirname(jbp.__file__), "codes", "learned_ms_T25_P768_J3_L12.npz")


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


def ms_params(model, seed, w0=0.8):
    rng = np.random.default_rng(seed)
    k = 2 * model.nb * model.T
    p = model.get_params()
    p[:2 * k] = w0 + 0.1 * rng.standard_normal(2 * k)
    p[2 * k:] = 0.3 * rng.standard_normal(65)
    p[-1] = -1.0
    return p


@pytest.mark.parametrize("use_mlp", [False, True])
def test_fp64_loss_and_gradient_match_oracle(main_decoder, main_code_oracle, lo, use_mlp):
    T = 5
    model = jbp.LearnedModel(main_decoder, T, kind=jbp.LEARNED_MIN_SUM)
    params = ms_params(model, 3)
    model.set_params(params, use_mlp=use_mlp)
    errs = np.stack([main_code_oracle.sample(106, t, 0.05) for t in (1, 2, 3, 4)])
    p = np.array([0.05, 0.04, 0.03, 0.045])
    loss, grad, conv = model.loss_grad(errs, p, precision=jbp.FP64)
    wc, wv, mlp = model.split(params)
    o_loss, (o_wc, o_wv, o_mlp), o_conv, _ = lo.loss(T, wc, wv, mlp, int(use_mlp), errs, p, grad=True,
                                                      kind=1)
    want = np.concatenate([o_wc.ravel(), o_wv.ravel(), o_mlp])
    print("min-sum fp64 loss %.15g oracle %.15g; gradient relative difference %.2e"
          % (loss, o_loss, np.linalg.norm(grad - want) / np.linalg.norm(want)), flush=True)
    assert abs(loss - o_loss) <= 1e-9 * abs(o_loss)
    assert np.linalg.norm(grad - want) <= 1e-7 * np.linalg.norm(want)
    assert np.array_equal(conv, o_conv)
    l32, g32, c32 = model.loss_grad(errs, p, precision=jbp.FP32)
    cos = g32 @ grad / (np.linalg.norm(g32) * np.linalg.norm(grad))
    assert abs(l32 - loss) <= 1e-4 * abs(loss) and cos > 0.9999
    model.destroy()


@pytest.mark.parametrize("use_mlp", [False, True])
def test_fp32_trainer_follows_fp64_at_t25(main_decoder, main_code_oracle, use_mlp):
    # the training workload's depth (T = 25, with and without the MLP
    # damping) on 24 trials at p in [0.03, 0.05]: the FP32 trainer (messages
    # on chip, for this code) against the FP64 one, which the oracle pins
    T = 25
    model = jbp.LearnedModel(main_decoder, T, kind=jbp.LEARNED_MIN_SUM)
    params = ms_params(model, 11)
    model.set_params(params, use_mlp=use_mlp)
    n = 24
    p = np.linspace(0.03, 0.05, n)
    errs = np.stack([main_code_oracle.sample(2024, 1 + t, float(p[t])) for t in range(n)])
    l64, g64, c64 = model.loss_grad(errs, p, precision=jbp.FP64)
    l32, g32, c32 = model.loss_grad(errs, p, precision=jbp.FP32)
    cos = g32 @ g64 / (np.linalg.norm(g32) * np.linalg.norm(g64))
    print("T=25 mlp=%d: loss fp32 %.9g fp64 %.9g; gradient cosine %.7f; convergence agreement %.3f"
          % (use_mlp, l32, l64, cos, float(np.mean(c32 == c64))), flush=True)
    assert np.isfinite(l32) and np.all(np.isfinite(g32))
    assert abs(l32 - l64) <= 1e-4 * abs(l64)
    assert cos > 0.9999
    assert np.mean(c32 == c64) >= 0.9
    model.destroy()


@pytest.mark.parametrize("kind,T", [(jbp.LEARNED_MIN_SUM, 1), (jbp.LEARNED_MIN_SUM, 64),
                                    (jbp.LEARNED_LAYERED_MIN_SUM, 1), (jbp.LEARNED_LAYERED_MIN_SUM, 64)])
def test_fp32_trainer_at_the_ends_of_its_range(main_decoder, main_code_oracle, kind, T):
    # the on-chip FP32 trainer at T = 1 and at the largest T (its LDS
    # accumulators are sized by T), on 5 trials (fewer trials than
    # workgroups), against the FP64 trainer
    model = jbp.LearnedModel(main_decoder, T, kind=kind)
    params = ms_params(model, 23)
    use_mlp = kind == jbp.LEARNED_MIN_SUM
    model.set_params(params, use_mlp=use_mlp)
    p = np.array([0.03, 0.035, 0.04, 0.045, 0.05])
    errs = np.stack([main_code_oracle.sample(77, 1 + t, float(p[t])) for t in range(len(p))])
    l64, g64, c64 = model.loss_grad(errs, p, precision=jbp.FP64)
    l32, g32, c32 = model.loss_grad(errs, p, precision=jbp.FP32)
    cos = g32 @ g64 / (np.linalg.norm(g32) * np.linalg.norm(g64))
    print("kind %d T=%d: loss fp32 %.9g fp64 %.9g; gradient cosine %.7f; conv %s / %s"
          % (kind, T, l32, l64, cos, c32, c64), flush=True)
    assert np.isfinite(l32) and np.all(np.isfinite(g32))
    assert abs(l32 - l64) <= 1e-4 * abs(l64)
    assert cos > 0.9999
    assert np.array_equal(c32, c64)
    model.destroy()


@pytest.mark.parametrize("kind,eps,weights_only", [(jbp.LEARNED_SUM_PRODUCT, 1e-3, True),
                                                   (jbp.LEARNED_SUM_PRODUCT, 1e-6, False),
                                                   (jbp.LEARNED_MIN_SUM, 1e-6, False)])
def test_fp64_gradient_central_differences_at_t25(main_decoder, main_code_oracle, kind, eps, weights_only):
    # T = 25 with the MLP damping: the FP64 adjoint against central
    # differences, relative error < 1e-3.  Both losses are piecewise smooth:
    # the MLP's second input (the unsatisfied fraction of a hard-decision
    # estimate) moves in steps, and the min-sum loss also where an argument
    # minimum changes.  A parameter counts only where its two one-sided
    # differences agree within 1% (curvature moves them apart by under 0.5%
    # at eps = 1e-3); at least 8 of those probed must.  eps = 1e-3 on the
    # sum-product model's weights (8 of them: early and late iterations, both
    # sides and kinds); eps = 1e-6 on those and on MLP parameters of both
    # kinds (an MLP parameter shifts the damping of every iteration, and the
    # steps around it are too dense for eps = 1e-3).
    T = 25
    model = jbp.LearnedModel(main_decoder, T, kind=kind)
    params = ms_params(model, 29)
    model.set_params(params, use_mlp=True)
    p = np.array([0.04, 0.05])
    errs = np.stack([main_code_oracle.sample(5, t, float(p[t - 1])) for t in (1, 2)])
    l0, grad, _ = model.loss_grad(errs, p, precision=jbp.FP64)
    nb = model.nb
    k = 2 * nb * T
    picks = [0, 7 * T + 3, nb * T + 1, 2 * nb * T - 2, k, k + 5 * T, k + nb * T + 2, 2 * k - 1]
    if not weights_only:
        picks += [2 * k, 2 * k + 1, 2 * k + 17, 2 * k + 33, 2 * k + 40, 2 * k + 48, 2 * k + 55, model.count - 1]
    smooth = 0
    for i in picks:
        q = params.copy()
        q[i] += eps
        model.set_params(q, use_mlp=True)
        lp = model.loss_grad(errs, p, precision=jbp.FP64, grad=False)[0]
        q[i] -= 2 * eps
        model.set_params(q, use_mlp=True)
        lm = model.loss_grad(errs, p, precision=jbp.FP64, grad=False)[0]
        fp, fm, fd = (lp - l0) / eps, (l0 - lm) / eps, (lp - lm) / (2 * eps)
        agree = abs(fp - fm) <= 0.01 * abs(fd)
        rel = abs(grad[i] - fd) / max(abs(fd), 1e-12)
        print("kind %d eps %g param %d: adjoint %.9g central %.9g (one-sided %.9g, %.9g) rel %.2e %s"
              % (kind, eps, i, grad[i], fd, fp, fm, rel, "smooth" if agree else "step"), flush=True)
        if agree:
            smooth += 1
            assert rel < 1e-3 or abs(grad[i] - fd) < 1e-9
    assert smooth >= 8
    model.destroy()


def test_fp16_inference_follows_the_training_forward_pass(main_decoder):
    # the inference phase (FP16 pairs, base 2) and the trainer's forward
    # pass (FP32, natural log) are the same model: the first iteration whose
    # estimate satisfies both syndromes agrees on almost every trial
    T = 25
    model = jbp.LearnedModel(main_decoder, T, kind=jbp.LEARNED_MIN_SUM)
    params = model.get_params()
    params[:4 * model.nb * T] = 1.0
    params[:2 * model.nb * T] = 0.75
    model.set_params(params, use_mlp=False)
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


@pytest.mark.parametrize("n,chunk", [(3 * 8192 + 5, 0), (3 * 8192 + 5, 2048), (3000, 0)])
def test_trials_left_unconverged_finish_as_plain_bp(main_decoder, n, chunk):
    # unit weights (plain min-sum) leave most trials unconverged after T
    # iterations: each is decoded again by plain BP from the prior, in a
    # second launch over the list of such trials, and reports T plus plain
    # BP's own count and plain BP's outcome (pipelined chunks, the fallback
    # launches on the finishing stream; 13 chunks over the eight buffer
    # sets; one chunk)
    T = 25
    model = jbp.LearnedModel(main_decoder, T, kind=jbp.LEARNED_MIN_SUM)
    params = model.get_params()
    params[:4 * model.nb * T] = 1.0
    model.set_params(params, use_mlp=False)
    print("fallback case n=%d chunk=%d" % (n, chunk), flush=True)
    _, rl = main_decoder.simulate_learned(model, 106, n, 0.05, max_iter=60, chunk=chunk)
    _, rp = main_decoder.simulate(106, n, 0.05, max_iter=60, precision=jbp.FP32)
    fb = rl["iterations"] > T
    print("%d of %d trials handed to plain BP" % (int(fb.sum()), n), flush=True)
    assert fb.sum() > n // 4
    assert np.array_equal(rl["iterations"][fb] - T, rp["iterations"][fb])
    for k in ("success", "syn_match", "exact"):
        assert np.array_equal(rl[k][fb], rp[k][fb]), k
    assert np.all(rl["syn_match"][~fb] == 1) and np.all(rl["iterations"][~fb] >= 1)
    model.destroy()


@pytest.fixture(scope="module")
def shipped(main_decoder):
    if not os.path.exists(SHIPPED):
        pytest.skip("no shipped min-sum model")
    m = jl.load(main_decoder, SHIPPED)
    yield m
    m.destroy()


def test_shipped_model_matches_plain_fer_at_16384_per_launch(main_decoder, shipped):
    # p = 0.04, T = 25 learned iterations, then plain FP32 BP from the prior
    # up to max-iter 1000; 16,384 trials per launch, 16 launches
    model = shipped
    assert model.kind == jbp.LEARNED_MIN_SUM and model.T == 25
    n, per = 16 * 16384, 16384
    sl = {"done": 0, "failures": 0, "total_iters": 0}
    rl = {k: [] for k in ("iterations", "success", "syn_match")}
    for j in range(n // per):
        s, r = main_decoder.simulate_learned(model, 106, per, 0.04, max_iter=1000,
                                             first_trial=1 + j * per)
        for k in sl:
            sl[k] += s[k]
        for k in rl:
            rl[k].append(r[k])
    rl = {k: np.concatenate(v) for k, v in rl.items()}
    sp, rp = main_decoder.simulate(106, n, 0.04, max_iter=1000, precision=jbp.FP32)
    lo_, hi_ = wilson(n, sp["failures"])
    print("learned: FER %.3g avg_iter %.3f; plain: FER %.3g ci95 [%.3g, %.3g] avg_iter %.3f; "
          "handed to plain BP: %d" % (sl["failures"] / n, sl["total_iters"] / n, sp["failures"] / n,
                                      lo_, hi_, sp["total_iters"] / n,
                                      int(np.sum(rl["iterations"] > 25))), flush=True)
    assert lo_ <= sl["failures"] / n <= hi_
    # a trial the learned phase leaves unconverged is decoded by plain BP
    # from the prior: T + plain's own count, plain's outcome
    fb = rl["iterations"] > 25
    assert np.array_equal(rl["iterations"][fb] - 25, rp["iterations"][fb])
    assert np.array_equal(rl["success"][fb], rp["success"][fb])
    # converged decodes are judged by the residual test like any other
    assert np.all(rl["success"] <= rl["syn_match"])


def test_training_step_at_full_size(main_decoder):
    # one step of the training workload: T = 25, batch 4,096, p ~ U[0.03, 0.05]
    model = jbp.LearnedModel(main_decoder, 25, kind=jbp.LEARNED_MIN_SUM)
    params = model.get_params()
    params[:4 * model.nb * 25] = 1.0
    params[:2 * model.nb * 25] = 0.75
    model.set_params(params, use_mlp=False)
    loss0 = model.train_step(2024, 1, 4096, 0.03, 0.05, lr=1e-2)
    after = model.get_params()
    print("full-size step loss %.6f, mean weight change %.3e" % (
        loss0, np.abs(after - params).mean()), flush=True)
    assert np.isfinite(loss0) and 0 < loss0 < 1
    assert np.all(np.isfinite(after)) and np.abs(after - params).max() > 0
    model.destroy()


def test_training_gain_over_untrained_min_sum(main_decoder, shipped):
    # the same decoder with unit weights (plain min-sum) against the trained
    # weights, on held-out trials: training must cut the trials handed to
    # plain BP and the mean iteration count
    n = 16384
    first = 50_000_001
    base = jbp.LearnedModel(main_decoder, 25, kind=jbp.LEARNED_MIN_SUM)
    params = base.get_params()
    params[:4 * base.nb * 25] = 1.0
    base.set_params(params, use_mlp=False)
    sb, rb = main_decoder.simulate_learned(base, 106, n, 0.04, max_iter=1000, first_trial=first)
    st, rt = main_decoder.simulate_learned(shipped, 106, n, 0.04, max_iter=1000, first_trial=first)
    fb_b, fb_t = int(np.sum(rb["iterations"] > 25)), int(np.sum(rt["iterations"] > 25))
    print("untrained: %d handed over, avg_iter %.3f; trained: %d handed over, avg_iter %.3f"
          % (fb_b, sb["total_iters"] / n, fb_t, st["total_iters"] / n), flush=True)
    assert fb_t * 10 < fb_b
    assert st["total_iters"] < 0.6 * sb["total_iters"]
    base.destroy()

"""Learned BP on the device: forward and adjoint kernels against the FP64
oracle and against central finite differences, the reduction to plain BP,
and training."""
import numpy as np
import pytest

import joint_bp_plus_pp_amd as jbp

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def lo(main_code_oracle):
    from _oracle import LearnedOracle
    return LearnedOracle(main_code_oracle)


def make_params(model, seed=0):
    rng = np.random.default_rng(seed)
    k = 2 * model.nb * model.T
    p = np.empty(model.count)
    p[:2 * k] = 1 + 0.1 * rng.standard_normal(2 * k)
    p[2 * k:] = 0.3 * rng.standard_normal(65)
    p[-1] = -1.0
    return p


def test_fp64_loss_and_gradient_match_oracle(main_decoder, main_code_oracle, lo):
    T = 4
    model = jbp.LearnedModel(main_decoder, T)
    params = make_params(model)
    model.set_params(params)
    errs = np.stack([main_code_oracle.sample(106, t, 0.05) for t in (1, 2, 3)])
    p = np.array([0.05, 0.04, 0.03])
    loss, grad, conv = model.loss_grad(errs, p, precision=jbp.FP64)
    wc, wv, mlp = model.split(params)
    o_loss, (o_wc, o_wv, o_mlp), o_conv, _ = lo.loss(T, wc, wv, mlp, 1, errs, p, grad=True)
    print("fp64 loss", loss, o_loss, flush=True)
    assert abs(loss - o_loss) <= 1e-9 * abs(o_loss)
    want = np.concatenate([o_wc.ravel(), o_wv.ravel(), o_mlp])
    assert np.linalg.norm(grad - want) <= 1e-7 * np.linalg.norm(want)
    assert np.array_equal(conv, o_conv)
    model.destroy()


def test_fp64_gradient_central_differences(main_decoder, main_code_oracle):
    # eps = 1e-3, relative error < 1e-3 on a spread of parameters
    T = 4
    model = jbp.LearnedModel(main_decoder, T)
    params = make_params(model, seed=1)
    model.set_params(params)
    errs = np.stack([main_code_oracle.sample(7, t, 0.04) for t in (1, 2)])
    _, grad, _ = model.loss_grad(errs, 0.04, precision=jbp.FP64)
    k = 2 * model.nb * model.T
    picks = [0, 17, 143, k // 2, k - 1, k, k + 29, 2 * k - 5, 2 * k, 2 * k + 33, 2 * k + 50,
             model.count - 1]
    eps = 1e-3
    for i in picks:
        q = params.copy()
        q[i] += eps
        model.set_params(q)
        lp = model.loss_grad(errs, 0.04, precision=jbp.FP64, grad=False)[0]
        q[i] -= 2 * eps
        model.set_params(q)
        lm = model.loss_grad(errs, 0.04, precision=jbp.FP64, grad=False)[0]
        fd = (lp - lm) / (2 * eps)
        rel = abs(grad[i] - fd) / max(abs(fd), 1e-12)
        print("param", i, grad[i], fd, rel, flush=True)
        assert rel < 1e-3 or abs(grad[i] - fd) < 1e-9
    model.destroy()


def test_fp32_gradient_close_to_fp64(main_decoder, main_code_oracle):
    T = 6
    model = jbp.LearnedModel(main_decoder, T)
    model.set_params(make_params(model, seed=2))
    errs = np.stack([main_code_oracle.sample(11, t, 0.045) for t in range(1, 9)])
    l64, g64, c64 = model.loss_grad(errs, 0.045, precision=jbp.FP64)
    l32, g32, c32 = model.loss_grad(errs, 0.045, precision=jbp.FP32)
    print("fp32 vs fp64 loss", l32, l64, flush=True)
    assert abs(l32 - l64) <= 1e-4 * abs(l64)
    cos = g32 @ g64 / (np.linalg.norm(g32) * np.linalg.norm(g64))
    assert cos > 0.9999 and np.linalg.norm(g32 - g64) <= 1e-2 * np.linalg.norm(g64)
    model.destroy()


def test_unit_weights_reduce_to_plain_bp(main_decoder):
    T = 8
    model = jbp.LearnedModel(main_decoder, T)
    k = 2 * model.nb * T
    params = np.zeros(model.count)
    params[:2 * k] = 1.0
    model.set_params(params, use_mlp=False)
    n = 512
    _, plain = main_decoder.simulate(106, n, 0.05, max_iter=50, precision=jbp.FP32, want_est=True)
    _, learned = main_decoder.simulate_learned(model, 106, n, 0.05, max_iter=50, want_est=True)
    it_p = plain["iterations"]
    conv_early = (plain["syn_match"] == 1) & (it_p <= T)
    want = np.where(conv_early, it_p, T + it_p)
    assert np.array_equal(learned["iterations"], want)
    assert np.array_equal(learned["success"], plain["success"])
    assert np.array_equal(learned["est"], plain["est"])
    model.destroy()


def test_training_repairs_detuned_weights(main_decoder):
    # Start with every message weight at 0.6 (under-confident messages) and
    # train on one fixed batch: Adam on the device-computed gradient must move
    # the weights back up and lower that batch's loss.
    T = 8
    model = jbp.LearnedModel(main_decoder, T)
    params = model.get_params()
    k = 2 * model.nb * T
    params[:2 * k] = 0.6
    model.set_params(params)
    losses = [model.train_step(31, 1, 128, 0.04, 0.05, lr=5e-2) for _ in range(20)]
    w_after = model.get_params()[:2 * k].mean()
    print("detuned fixed-batch loss", np.round(losses[::4], 6), "mean weight", w_after, flush=True)
    assert np.all(np.isfinite(losses))
    assert losses[-1] < 0.9 * losses[0]
    assert w_after > 0.65
    model.destroy()


def test_ablation_trained_vs_untrained(main_decoder, main_code_oracle):
    # What training buys on held-out trials, against the same decoder with
    # unit weights and its initial damping (reported; training must not
    # make the held-out loss worse by more than noise).
    T = 10
    trained = jbp.LearnedModel(main_decoder, T)
    base = jbp.LearnedModel(main_decoder, T)
    for i in range(40):
        trained.train_step(2024, 1 + 512 * i, 512, 0.03, 0.05, lr=1e-2)
    held = np.stack([main_code_oracle.sample(999, t, 0.045) for t in range(1, 257)])
    lt, _, ct = trained.loss_grad(held, 0.045, precision=jbp.FP32, grad=False)
    lb, _, cb = base.loss_grad(held, 0.045, precision=jbp.FP32, grad=False)
    it_t = np.where(ct > 0, ct, 2 * T).mean()
    it_b = np.where(cb > 0, cb, 2 * T).mean()
    print("held-out loss trained %.6f untrained %.6f; mean convergence iteration %.3f vs %.3f"
          % (lt, lb, it_t, it_b), flush=True)
    assert lt <= lb * 1.01
    trained.destroy()
    base.destroy()

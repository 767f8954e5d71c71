"""Training driver for the learned (neural-weighted) joint BP decoder.

The forward and adjoint passes, the Adam update and the sampling of the
training syndromes all run on the device (libjbp: jbp_learned_train_step);
this module only sequences the steps, evaluates held-out batches, and stores
parameters as .npz files (plain arrays, loaded with allow_pickle=False).

No counterpart in the reference program: its only learned-ish knob is the
fixed --damping (jointbp_ets.cpp:1903-1911).
"""
import os
import time

import numpy as np

from . import FP32, LearnedModel


def train(decoder, T=25, steps=200, batch=4096, lr=1e-2, p_lo=0.03, p_hi=0.05, seed=2024,
          first_trial=1, log_every=20, lr_decay=0.0, log=None, use_mlp=True, kind=0, init=None):
    """Train a LearnedModel on device-sampled trials (p uniform in
    [p_lo, p_hi] per trial, consecutive trial indices per step).  Returns
    (model, history) with one (step, loss, seconds) entry per logged step.
    use_mlp=False trains the message weights alone (damping 0), which the
    inference decoder runs without the damping state."""
    model = LearnedModel(decoder, T, kind=kind)
    params = model.get_params()
    if init is not None:
        params = np.asarray(init, dtype=np.float64).copy()
    model.set_params(params, use_mlp=use_mlp and kind != 2)  # the layered kind has no damping
    history = []
    t0 = time.time()
    nxt = first_trial
    for step in range(1, steps + 1):
        rate = lr / (1.0 + lr_decay * (step - 1))
        loss = model.train_step(seed, nxt, batch, p_lo, p_hi, lr=rate)
        nxt += batch
        if step % log_every == 0 or step == 1 or step == steps:
            history.append((step, loss, time.time() - t0))
            if log:
                log("train step %d loss %.6f lr %.2e (%.1f s)" % (step, loss, rate, time.time() - t0))
    return model, history


def save(model, path, meta=None):
    d = {"params": model.get_params(), "T": np.array(model.T), "nb": np.array(model.nb),
         "use_mlp": np.array(int(getattr(model, "use_mlp", True))),
         "kind": np.array(int(getattr(model, "kind", 0)))}
    for k, v in (meta or {}).items():
        d["meta_" + k] = np.asarray(v)
    np.savez(path, **d)


def load(decoder, path, use_mlp=None):
    """use_mlp=None takes the setting stored with the parameters."""
    with np.load(path, allow_pickle=False) as z:
        T = int(z["T"])
        if use_mlp is None:
            use_mlp = bool(int(z["use_mlp"])) if "use_mlp" in z.files else True
        kind = int(z["kind"]) if "kind" in z.files else 0
        model = LearnedModel(decoder, T, kind=kind)
        params = z["params"]
        if params.size != model.count:
            raise ValueError("%s holds %d parameters, the code needs %d" % (path, params.size,
                                                                            model.count))
        model.set_params(params, use_mlp=use_mlp)
    return model


def evaluate(decoder, model, base_seed, n_trials, p, max_iter=1000, first_trial=1):
    """Learned decoder with plain fallback against plain BP on the same
    trials: (learned stats, plain stats, per-trial dicts)."""
    sl, rl = decoder.simulate_learned(model, base_seed, n_trials, p, max_iter=max_iter,
                                      first_trial=first_trial)
    sp, rp = decoder.simulate(base_seed, n_trials, p, max_iter=max_iter, first_trial=first_trial,
                              precision=FP32)
    return sl, sp, rl, rp


def main(argv=None):
    """Train on the device and write the parameters:
    python -m joint_bp_plus_pp_amd.learned --params CODE.txt --out W.npz"""
    import argparse
    import sys

    from . import Code, Decoder

    ap = argparse.ArgumentParser(description=main.__doc__.split("\n")[0])
    ap.add_argument("--params", required=True, help="params file of the code")
    ap.add_argument("--out", required=True)
    ap.add_argument("--T", type=int, default=25)
    ap.add_argument("--steps", type=int, default=300)
    ap.add_argument("--batch", type=int, default=4096)
    ap.add_argument("--lr", type=float, default=1e-2)
    ap.add_argument("--lr-decay", type=float, default=0.01)
    ap.add_argument("--p-lo", type=float, default=0.03)
    ap.add_argument("--p-hi", type=float, default=0.05)
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--device", type=int, default=0)
    ap.add_argument("--no-mlp", action="store_true", help="message weights only, no damping MLP")
    ap.add_argument("--kind", choices=("sum-product", "min-sum", "layered-min-sum"),
                    default="layered-min-sum",
                    help="updates and schedule of the learned decoder (layered: no damping)")
    ap.add_argument("--init-weight", type=float, default=1.0,
                    help="starting value of every message weight")
    ap.add_argument("--init-var-weight", type=float, default=None,
                    help="starting value of the variable-message weights (default: --init-weight)")
    ap.add_argument("--first-trial", type=int, default=1,
                    help="trial index of the first training sample")
    ap.add_argument("--init-schedule", default="",
                    help="comma-separated starting check-message weights of the first "
                         "iterations (the rest start at --init-weight)")
    a = ap.parse_args(argv)
    dec = Decoder(Code.from_params(a.params), a.device)
    kind = {"sum-product": 0, "min-sum": 1, "layered-min-sum": 2}[a.kind]
    probe = LearnedModel(dec, a.T, kind=kind)
    init = probe.get_params()
    probe.destroy()
    nw = 2 * probe.nb * a.T
    init[:nw] = a.init_weight
    init[nw:2 * nw] = a.init_weight if a.init_var_weight is None else a.init_var_weight
    sched = [float(x) for x in a.init_schedule.split(",") if x.strip()]
    if len(sched) > a.T:
        ap.error("--init-schedule has more entries than --T")
    wc = init[:2 * probe.nb * a.T].reshape(2, probe.nb, a.T)  # check weights, a view
    for k, w in enumerate(sched):
        wc[:, :, k] = w
    model, hist = train(dec, T=a.T, steps=a.steps, batch=a.batch, lr=a.lr, p_lo=a.p_lo,
                        p_hi=a.p_hi, seed=a.seed, first_trial=a.first_trial, lr_decay=a.lr_decay,
                        log=lambda m: print(m, flush=True), use_mlp=not a.no_mlp and kind != 2,
                        kind=kind,
                        init=init)
    loss = hist[-1][1] if hist else float("nan")
    save(model, a.out, meta={"steps": a.steps, "batch": a.batch, "lr": a.lr, "p_lo": a.p_lo,
                             "p_hi": a.p_hi, "seed": a.seed, "loss": loss,
                             "init_weight": a.init_weight, "init_schedule": a.init_schedule,
                             "init_var_weight": (a.init_weight if a.init_var_weight is None
                                                 else a.init_var_weight),
                             "first_trial": a.first_trial})
    print("wrote %s (%d steps, final loss %.6f)" % (a.out, a.steps, loss), flush=True)
    model.destroy()
    dec.destroy()
    return 0


if __name__ == "__main__":
    raise SystemExit(main())


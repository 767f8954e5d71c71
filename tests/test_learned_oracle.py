"""Pins of the FP64 learned-BP oracle (oracle/learned_oracle.cpp): its
hand-derived adjoint against central finite differences of its own loss,
and its reduction to the reference's plain joint BP at unit weights with
damping off."""
import numpy as np
import pytest

from conftest import golden_npz


@pytest.fixture(scope="module")
def lo(main_code_oracle):
    from _oracle import LearnedOracle
    return LearnedOracle(main_code_oracle)


def random_params(nb, T, seed=0):
    rng = np.random.default_rng(seed)
    wc = 1 + 0.1 * rng.standard_normal((2, nb, T))
    wv = 1 + 0.1 * rng.standard_normal((2, nb, T))
    mlp = 0.3 * rng.standard_normal(65)
    mlp[64] = -1.0
    return wc, wv, mlp


def test_adjoint_matches_central_differences(lo, main_code_oracle):
    T = 4
    wc, wv, mlp = random_params(lo.nb, T)
    errs = np.stack([main_code_oracle.sample(106, t, 0.05) for t in (1, 2)])
    p = np.array([0.05, 0.04])
    _, (gwc, gwv, gm), _, _ = lo.loss(T, wc, wv, mlp, 1, errs, p, grad=True)
    eps = 1e-3
    picks = [("wc", (0, 3, 1)), ("wc", (1, 20, 3)), ("wc", (0, 0, 0)), ("wv", (0, 5, 0)),
             ("wv", (1, 30, 2)), ("wv", (1, 35, 3)), ("mlp", (64,)), ("mlp", (3,)), ("mlp", (40,)),
             ("mlp", (50,))]
    an, fd = [], []
    for name, ix in picks:
        arr = {"wc": wc, "wv": wv, "mlp": mlp}[name]
        g = {"wc": gwc, "wv": gwv, "mlp": gm}[name]
        old = arr[ix]
        arr[ix] = old + eps
        lp = lo.loss(T, wc, wv, mlp, 1, errs, p)[0]
        arr[ix] = old - eps
        lm = lo.loss(T, wc, wv, mlp, 1, errs, p)[0]
        arr[ix] = old
        an.append(g[ix])
        fd.append((lp - lm) / (2 * eps))
    an, fd = np.array(an), np.array(fd)
    rel = np.abs(an - fd) / np.maximum(np.abs(fd), 1e-30)
    assert rel.max() < 1e-3, rel


def test_unit_weights_without_damping_are_plain_bp(lo, main_code_oracle):
    z = golden_npz("trials_p0.05_it50.npz")
    T, N = 50, 24
    ones = np.ones((2, lo.nb, T))
    errs = np.stack([main_code_oracle.sample(106, t, 0.05) for t in range(1, N + 1)])
    _, conv, est = lo.loss(T, ones, ones, np.zeros(65), 0, errs, 0.05)
    want = np.where(z["syn_ok"][:N] == 1, z["iterations"][:N], 0)
    assert np.array_equal(conv, want)

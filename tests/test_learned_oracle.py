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


# ---------------------------------------------------------------------------
# Min-sum kind (kind = 1): max-product BP with learned weights.

def tree_code(oracle):
    """A cycle-free joint factor graph: X checks {0,1,2}, {2,3}; Z checks
    {1,4,5}, {3,6,7} (qubits 1, 2, 3 join the checks into one tree)."""
    xs = np.array([[0, 1, 2], [2, 3, -1]], np.int32)
    zs = np.array([[1, 4, 5], [3, 6, 7]], np.int32)
    return oracle.code_from_checks(8, xs, zs), xs, zs


def brute_force_max_marginals(n, xs, zs, sx, sz, prior):
    """Per qubit: max log P(e) over errors with the syndromes and x_v = 0,
    minus the same with x_v = 1 (and for z_v); states 0 I, 1 X, 2 Z, 3 Y."""
    lp = np.log(prior)
    states = np.array(np.meshgrid(*[np.arange(4)] * n, indexing="ij")).reshape(n, -1).T
    xb, zb = states & 1, states >> 1
    ok = np.ones(len(states), bool)
    for c, row in enumerate(xs):
        q = row[row >= 0]
        ok &= (xb[:, q].sum(1) & 1) == sx[c]
    for c, row in enumerate(zs):
        q = row[row >= 0]
        ok &= (zb[:, q].sum(1) & 1) == sz[c]
    w = lp[states[ok]].sum(1)
    out = np.zeros(2 * n)
    for v in range(n):
        for side, bits in ((0, xb[ok, v]), (1, zb[ok, v])):
            out[side * n + v] = w[bits == 0].max() - w[bits == 1].max()
    return out


def test_minsum_unit_weights_are_exact_max_product_on_a_tree(oracle):
    # max-product BP is exact on a cycle-free graph: after enough iterations
    # the min-sum marginals equal the brute-force max-marginal log ratios
    code, xs, zs = tree_code(oracle)
    offs = [code.csr(s)[0] for s in (0, 1)]
    blk = [np.zeros(off[-1], np.int32) for off in offs]
    lo = LearnedOracle_(code, blk=blk, nb=1)
    T = 8
    ones = np.ones((2, 1, T))
    rng = np.random.default_rng(5)
    for trial in range(12):
        err = rng.integers(0, 4, size=8).astype(np.int32)
        err[rng.random(8) < 0.5] = 0
        p = [0.05, 0.12, 0.2][trial % 3]
        sx, sz = code.syndrome(err)
        _, conv, est, ex = lo.loss(T, ones, ones, np.zeros(65), 0, err[None], p, kind=1, extra=True)
        want = brute_force_max_marginals(8, xs, zs, sx, sz, oracle.prior(p))
        assert np.allclose(ex["marginals"][0], want, rtol=0, atol=1e-9), (trial, err)


def LearnedOracle_(code, **kw):
    from _oracle import LearnedOracle
    return LearnedOracle(code, **kw)


def test_minsum_adjoint_matches_central_differences_on_smooth_pieces(oracle):
    # The min-sum loss is piecewise smooth.  Central differences (eps = 1e-3)
    # are compared where the forward pass makes the same discrete choices at
    # theta - eps, theta and theta + eps (the oracle's signature hash), which
    # must happen for at least 4 parameters of each group (check weights,
    # variable weights, MLP).  The code is a small hypergraph product, so
    # that such parameters are common.
    import joint_bp_plus_pp_amd as jbp
    c = jbp.Code.hgp(15, 20, 3, 4, seed=2)
    oc = oracle.code_from_checks(c.n, c.checks(0), c.checks(1))
    blk = []
    for s in (0, 1):
        off = oc.csr(s)[0]
        b = np.zeros(off[-1], np.int32)
        for k in range(len(off) - 1):
            b[off[k]:off[k + 1]] = np.arange(off[k + 1] - off[k])
        blk.append(b)
    nb = int(max(b.max() for b in blk)) + 1
    lo = LearnedOracle_(oc, blk=blk, nb=nb)
    T = 5
    rng = np.random.default_rng(1)
    wc = 0.8 + 0.1 * rng.standard_normal((2, nb, T))
    wv = 1 + 0.1 * rng.standard_normal((2, nb, T))
    mlp = 0.3 * rng.standard_normal(65)
    mlp[64] = -1.0
    errs = np.stack([oc.sample(7, t, 0.05) for t in (1, 2, 3)])
    p = np.array([0.05, 0.04, 0.06])
    _, (gwc, gwv, gm), _, _, ex = lo.loss(T, wc, wv, mlp, 1, errs, p, grad=True, kind=1, extra=True)
    eps = 1e-3
    found = {"wc": 0, "wv": 0, "mlp": 0}
    worst = 0.0
    for attempt in range(300):
        name = ("wc", "wv", "mlp")[attempt % 3]
        if found[name] >= 4:
            continue
        arr = {"wc": wc, "wv": wv, "mlp": mlp}[name]
        g = {"wc": gwc, "wv": gwv, "mlp": gm}[name]
        ix = tuple(int(rng.integers(0, s)) for s in arr.shape)
        old = arr[ix]
        arr[ix] = old + eps
        lp, _, _, ep = lo.loss(T, wc, wv, mlp, 1, errs, p, kind=1, extra=True)
        arr[ix] = old - eps
        lm, _, _, em = lo.loss(T, wc, wv, mlp, 1, errs, p, kind=1, extra=True)
        arr[ix] = old
        if not (ep["signature"] == em["signature"] == ex["signature"]):
            continue
        fd = (lp - lm) / (2 * eps)
        if abs(fd) < 1e-12 and abs(g[ix]) < 1e-12:
            continue  # parameters that do not reach the loss
        rel = abs(g[ix] - fd) / abs(fd)
        worst = max(worst, rel)
        found[name] += 1
        assert rel < 1e-3, (name, ix, g[ix], fd)
    print("smooth pieces checked", found, "worst relative error %.2e" % worst)
    assert all(v >= 4 for v in found.values()), found


# ---------------------------------------------------------------------------
# Layered min-sum kind (kind = 2).

def small_hgp_oracle(oracle, nlay=3):
    import joint_bp_plus_pp_amd as jbp
    c = jbp.Code.hgp(15, 20, 3, 4, seed=2)
    oc = oracle.code_from_checks(c.n, c.checks(0), c.checks(1))
    blk = []
    for s in (0, 1):
        off = oc.csr(s)[0]
        b = np.zeros(off[-1], np.int32)
        for k in range(len(off) - 1):
            b[off[k]:off[k + 1]] = np.arange(off[k + 1] - off[k])
        blk.append(b)
    nb = int(max(b.max() for b in blk)) + 1
    # layers: contiguous ranges of check indices
    lay = [(np.arange(m) * nlay // m).astype(np.int32) for m in (oc.mx, oc.mz)]
    return oc, LearnedOracle_(oc, blk=blk, nb=nb, lay=lay, nlay=nlay), nb


def test_layered_with_one_layer_is_the_flooding_min_sum(oracle):
    # one layer: every check updates from the same state, i.e. the flooding
    # schedule of kind 1 (whose adjoint and tree exactness are pinned above):
    # loss and gradient agree to rounding
    oc, lo, nb = small_hgp_oracle(oracle, nlay=1)
    T = 5
    rng = np.random.default_rng(3)
    wc = 0.8 + 0.1 * rng.standard_normal((2, nb, T))
    wv = 1 + 0.1 * rng.standard_normal((2, nb, T))
    errs = np.stack([oc.sample(7, t, 0.05) for t in (1, 2, 3, 4)])
    p = np.array([0.05, 0.04, 0.06, 0.05])
    l1, (gc1, gv1, _), _, _ = lo.loss(T, wc, wv, np.zeros(65), 0, errs, p, grad=True, kind=1)
    l2, (gc2, gv2, _), _, _ = lo.loss(T, wc, wv, np.zeros(65), 0, errs, p, grad=True, kind=2)
    assert abs(l1 - l2) <= 1e-12 * abs(l1)
    assert np.allclose(gc1, gc2, rtol=1e-10, atol=1e-15) and np.allclose(gv1, gv2, rtol=1e-10, atol=1e-15)


@pytest.mark.parametrize("nlay", [2, 3])
def test_layered_unit_weights_are_exact_max_product_on_a_tree(oracle, nlay):
    # on a cycle-free graph layered max-product also reaches the exact
    # max-marginals
    code, xs, zs = tree_code(oracle)
    offs = [code.csr(s)[0] for s in (0, 1)]
    blk = [np.zeros(off[-1], np.int32) for off in offs]
    lay = [np.arange(2, dtype=np.int32) % nlay for _ in range(2)]
    lo = LearnedOracle_(code, blk=blk, nb=1, lay=lay, nlay=nlay)
    T = 8
    ones = np.ones((2, 1, T))
    rng = np.random.default_rng(6)
    for trial in range(12):
        err = rng.integers(0, 4, size=8).astype(np.int32)
        err[rng.random(8) < 0.5] = 0
        p = [0.05, 0.12, 0.2][trial % 3]
        sx, sz = code.syndrome(err)
        _, conv, est, ex = lo.loss(T, ones, ones, np.zeros(65), 0, err[None], p, kind=2, extra=True)
        want = brute_force_max_marginals(8, xs, zs, sx, sz, oracle.prior(p))
        assert np.allclose(ex["marginals"][0], want, rtol=0, atol=1e-9), (trial, err)


def test_layered_adjoint_matches_central_differences_on_smooth_pieces(oracle):
    # as for kind 1: central differences where the forward pass makes the
    # same discrete choices at theta - eps, theta and theta + eps, for at
    # least 5 check weights and 5 variable weights, over 3 layers (a variable
    # weight moves every message it scales, so fewer of them sit on a smooth
    # piece at this eps: more draws)
    oc, lo, nb = small_hgp_oracle(oracle, nlay=3)
    T = 5
    rng = np.random.default_rng(11)
    wc = 0.8 + 0.1 * rng.standard_normal((2, nb, T))
    wv = 1 + 0.1 * rng.standard_normal((2, nb, T))
    mlp = np.zeros(65)
    errs = np.stack([oc.sample(7, t, 0.05) for t in (1, 2, 3)])
    p = np.array([0.05, 0.04, 0.06])
    _, (gwc, gwv, _), _, _, ex = lo.loss(T, wc, wv, mlp, 0, errs, p, grad=True, kind=2, extra=True)
    eps = 1e-3
    found = {"wc": 0, "wv": 0}
    worst = 0.0
    for attempt in range(2000):
        name = ("wc", "wv")[attempt % 2]
        if found[name] >= 5:
            continue
        arr = {"wc": wc, "wv": wv}[name]
        g = {"wc": gwc, "wv": gwv}[name]
        ix = tuple(int(rng.integers(0, s)) for s in arr.shape)
        old = arr[ix]
        arr[ix] = old + eps
        lp, _, _, ep = lo.loss(T, wc, wv, mlp, 0, errs, p, kind=2, extra=True)
        arr[ix] = old - eps
        lm, _, _, em = lo.loss(T, wc, wv, mlp, 0, errs, p, kind=2, extra=True)
        arr[ix] = old
        if not (ep["signature"] == em["signature"] == ex["signature"]):
            continue
        fd = (lp - lm) / (2 * eps)
        if abs(fd) < 1e-9:
            # weights whose effect is below the differences' rounding (a
            # loss of ~0.05 differenced at eps 1e-3 resolves ~1e-13)
            continue
        rel = abs(g[ix] - fd) / abs(fd)
        worst = max(worst, rel)
        found[name] += 1
        assert rel < 1e-3, (name, ix, g[ix], fd)
    print("smooth pieces checked", found, "worst relative error %.2e" % worst)
    assert all(v >= 5 for v in found.values()), found

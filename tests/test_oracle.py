"""Pins of the CPU oracle (oracle/oracle.cpp) before it is trusted as the
checker: published known answers, the reference decoder's golden vectors
(tests/golden/, written by make_golden.py from the reference's own code),
the reference README's code summary, and brute force on small inputs."""
import hashlib

import numpy as np
import pytest

from conftest import PARAMS, fnv1a, golden_json, golden_npz, unpack2

GOLDEN_SETS = [
    ("trials_p0.05_it50.npz", 0.0),
    ("trials_p0.04_it1000.npz", 0.0),
    ("trials_p0.025_it200.npz", 0.0),
    ("trials_p0.05_it50_damp0.3.npz", 0.3),
]


def test_splitmix64_published_vectors(oracle):
    # SplitMix64 reference outputs (Vigna's splitmix64.c): seed 0 first
    # output, and the first five outputs from seed 1234567.
    assert oracle.splitmix64(0) == 0xE220A8397B1DCDAF
    g, M = 0x9E3779B97F4A7C15, 2 ** 64
    got = [oracle.splitmix64((1234567 + i * g) % M) for i in range(5)]
    assert got == [6457827717110365317, 3203168211198807973, 9817491932198370423,
                   4593380528125082431, 16408922859458223821]


def test_trial_seeds_match_reference(oracle):
    z = golden_npz("trials_p0.05_it50.npz")
    base, first, count, _ = z["meta"]
    got = [oracle.trial_seed(int(base), int(first) + i) for i in range(int(count))]
    assert np.array_equal(np.array(got, dtype=np.uint64), z["seed"])


def test_mt19937_64_standard_known_answer(oracle):
    # C++ standard [rand.predef]: the 10000th output of a default-constructed
    # mt19937_64 (seed 5489) is 9981545732273789042.
    assert int(oracle.mt_outputs(5489, 10000)[-1]) == 9981545732273789042


def test_code_construction_matches_reference(main_code_oracle):
    c = main_code_oracle
    summary = golden_json("code_summary.json")
    # README.md:124-132 known answers
    assert (c.n, c.mx, c.mz, c.ex, c.ez) == (9216, 2304, 2304, 27648, 27648)
    assert (c.rank_x, c.rank_z) == (2302, 2302) == (summary["rank_x"], summary["rank_z"])
    assert c.n - c.rank_x - c.rank_z == 4612
    for side, key in ((0, "checks_x_sha256"), (1, "checks_z_sha256")):
        off, ev = c.csr(side)
        assert np.all(np.diff(off) == 12)
        assert hashlib.sha256(ev.astype(np.int32).tobytes()).hexdigest() == summary[key]


def dense(code, side):
    off, ev = code.csr(side)
    m = len(off) - 1
    H = np.zeros((m, code.n), np.uint8)
    for c in range(m):
        H[c, ev[off[c]:off[c + 1]]] ^= 1
    return H


def test_hx_hz_orthogonal_by_dense_product(main_code_oracle):
    hx = dense(main_code_oracle, 0).astype(np.int64)
    hz = dense(main_code_oracle, 1).astype(np.int64)
    assert not np.any((hx @ hz.T) & 1)


@pytest.mark.parametrize("name,damping", GOLDEN_SETS[:3])
def test_sampler_matches_reference(oracle, main_code_oracle, name, damping):
    z = golden_npz(name)
    base, first, count, _ = map(int, z["meta"])
    p = float(z["p"][0])
    for i in range(count):
        err = main_code_oracle.sample(base, first + i, p)
        assert fnv1a(err) == int(z["err_hash"][i]), (name, i)
        assert np.count_nonzero(err) == int(z["err_weight"][i])
        if "err_packed" in z and i < len(z["err_packed"]):
            assert np.array_equal(err, unpack2(z["err_packed"][i]))


def test_depolarizing_frequencies(main_code_oracle):
    # each of X, Z, Y with probability p/3 (jointbp_ets.cpp:5938-5945)
    p = 0.3
    counts = np.zeros(4)
    for t in range(1, 41):
        counts += np.bincount(main_code_oracle.sample(7, t, p), minlength=4)
    frac = counts / counts.sum()
    assert abs(frac[0] - (1 - p)) < 0.01
    for k in (1, 2, 3):
        assert abs(frac[k] - p / 3) < 0.01


def test_syndromes_match_dense_product_and_reference(main_code_oracle):
    z = golden_npz("trials_p0.05_it50.npz")
    base, first, count, _ = map(int, z["meta"])
    hx = dense(main_code_oracle, 0).astype(np.int64)
    hz = dense(main_code_oracle, 1).astype(np.int64)
    for i in range(count):
        err = main_code_oracle.sample(base, first + i, 0.05)
        sx, sz = main_code_oracle.syndrome(err)
        assert fnv1a(sx) == int(z["sx_hash"][i]) and fnv1a(sz) == int(z["sz_hash"][i])
        if i < 16:
            xb = (err == 1) | (err == 3)
            zb = (err == 2) | (err == 3)
            assert np.array_equal(sx, (hx @ xb) & 1)
            assert np.array_equal(sz, (hz @ zb) & 1)


@pytest.mark.parametrize("name,damping", GOLDEN_SETS)
def test_decode_matches_reference_per_trial(main_code_oracle, name, damping):
    z = golden_npz(name)
    base, first, count, max_iter = map(int, z["meta"])
    p = float(z["p"][0])
    assert float(z["p"][1]) == damping
    r = main_code_oracle.simulate(base, first, count, p, max_iter, damping=damping)
    assert np.array_equal(r["iterations"], z["iterations"])
    assert np.array_equal(r["syn_ok"], z["syn_ok"].astype(np.uint8))
    assert np.array_equal(r["success"], z["success"].astype(np.uint8))
    assert np.array_equal(r["exact"], z["exact"].astype(np.uint8))
    assert np.array_equal(r["est_hash"], z["est_hash"])


def test_cfg1_aggregate(main_code_oracle):
    agg = golden_json("aggregates.json")["cfg1_p0.05_it50_seed106_1000"]
    r = main_code_oracle.simulate(agg["base"], agg["first"], agg["count"], agg["p"], agg["max_iter"])
    assert int(agg["count"] - r["success"].sum()) == agg["failures"]
    assert int(r["iterations"].sum()) == agg["iters"]
    hist = {int(k): int(v) for k, v in zip(*np.unique(r["iterations"], return_counts=True))}
    assert hist == {int(k): v for k, v in agg["iter_hist"].items()}


def test_prior_normalisation(oracle):
    pr = oracle.prior(0.04)
    m = [1 - 0.04, 0.04 / 3, 0.04 / 3, 0.04 / 3]
    s = ((m[0] + m[1]) + m[2]) + m[3]
    assert list(pr) == [v / s for v in m]


def span_brute(rows, v):
    """Is v in the GF(2) span of rows?  Enumerates all 2^k combinations."""
    k = rows.shape[0]
    for mask in range(1 << k):
        acc = np.zeros(rows.shape[1], np.uint8)
        for i in range(k):
            if mask >> i & 1:
                acc ^= rows[i]
        if np.array_equal(acc, v):
            return True
    return False


def test_row_space_membership_brute_force(oracle):
    rng = np.random.default_rng(5)
    n = 12
    for trial in range(6):
        m = 5
        x_checks = np.array([np.sort(rng.choice(n, 4, replace=False)) for _ in range(m)], np.int32)
        z_checks = np.array([np.sort(rng.choice(n, 3, replace=False)) for _ in range(m)], np.int32)
        code = oracle.code_from_checks(n, x_checks, z_checks)
        for side, checks in ((0, x_checks), (1, z_checks)):
            H = np.zeros((m, n), np.uint8)
            for c, row in enumerate(checks):
                H[c, row] ^= 1
            vecs = [rng.integers(0, 2, n).astype(np.uint8) for _ in range(10)]
            combo = rng.integers(0, 2, m).astype(np.uint8)
            vecs.append(((combo[:, None] * H).sum(0) & 1).astype(np.uint8))  # a member
            for v in vecs:
                assert code.in_row_space(side, v) == span_brute(H, v)


def test_planted_stabilizer_is_success(main_code_oracle):
    c = main_code_oracle
    err = c.sample(106, 3, 0.05)
    off, ev = c.csr(1)  # a row of Hz as the X part of the residual
    row = ev[off[5]:off[6]]
    est = err.copy()
    est[row] ^= 1  # flip the X bit: I<->X, Z<->Y
    assert c.stab_equiv(err, est)
    est2 = err.copy()
    est2[row[:-1]] ^= 1  # drop one qubit: not a stabilizer
    assert not c.stab_equiv(err, est2)


def test_freeze_syn_matches_reference(oracle, main_code_oracle):
    # --freeze-syn (jointbp_ets.cpp:1884-1901, 2526-2535) on X-only errors:
    # per decode the reference's iterations, match and estimate, with and
    # without freezing (the fixture holds decodes where freezing changes them)
    z = golden_npz("decodes_xonly_p0.04_it60_freeze.npz")
    base, _, count, max_iter = map(int, z["meta"])
    p, damping = z["p"]
    prior = oracle.prior(p)
    changed = 0
    for i in range(count):
        err = main_code_oracle.sample(base, 1 + i, p) & 1
        assert fnv1a(err) == int(z["err_hash"][i])
        sx, sz = main_code_oracle.syndrome(err)
        est, it, ok = main_code_oracle.decode(sx, sz, prior, max_iter, damping, freeze_syn=True)
        assert (it, int(ok), fnv1a(est)) == (int(z["iterations"][i]), int(z["syn_ok"][i]),
                                             int(z["est_hash"][i])), i
        est, it, ok = main_code_oracle.decode(sx, sz, prior, max_iter, damping, freeze_syn=False)
        assert (it, fnv1a(est)) == (int(z["iterations_nofreeze"][i]), int(z["est_hash_nofreeze"][i]))
        changed += int(z["iterations"][i] != z["iterations_nofreeze"][i] or
                       z["est_hash"][i] != z["est_hash_nofreeze"][i])
    assert changed >= 2

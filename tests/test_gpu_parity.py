"""GPU parity: the HIP path through the C ABI against the oracle and the
reference's golden vectors.

Bars: sampling, syndromes and residual tests bit-exact; the FP64 decoder
bit-exact per trial (iterations, syndrome match, success, est); the FP32
decoder's frame error rate inside the reference's 95% Wilson interval and
its per-trial agreement high (FP32 messages round differently, so a trial
near a decision boundary may take another path)."""
import math

import numpy as np
import pytest

import joint_bp_plus_pp_amd as jbp
from conftest import fnv1a, golden_json, golden_npz, unpack2

pytestmark = pytest.mark.gpu

GOLDEN_SETS = [
    ("trials_p0.05_it50.npz", 0.0),
    ("trials_p0.04_it1000.npz", 0.0),
    ("trials_p0.025_it200.npz", 0.0),
    ("trials_p0.05_it50_damp0.3.npz", 0.3),
]


def wilson(n, k, z=1.96):
    # jointbp_ets.cpp:4388-4399
    phat = k / n
    denom = 1 + z * z / n
    center = (phat + z * z / (2 * n)) / denom
    half = z * math.sqrt((phat * (1 - phat) + z * z / (4 * n)) / n) / denom
    return max(0.0, center - half), min(1.0, center + half)


def test_device_info(main_decoder):
    info = main_decoder.info()
    print("device info", info, flush=True)
    assert info["qc"] == 1 and (info["P"], info["J"], info["L"]) == (768, 3, 12)
    assert info["basis"] == 1


@pytest.mark.parametrize("p,base", [(0.05, 106), (0.04, 106), (0.3, 7), (0.001, 5), (0.999, 5),
                                    (1.0, 9), (0.0, 9)])
def test_sampler_bit_exact(main_decoder, main_code_oracle, p, base):
    print("sampler", p, base, flush=True)
    err, sx, sz = main_decoder.sample(base, 1, 48, p)
    for i in range(48):
        e = main_code_oracle.sample(base, 1 + i, p)
        assert np.array_equal(err[i], e), (p, i)
        ox, oz = main_code_oracle.syndrome(e)
        assert np.array_equal(sx[i], ox) and np.array_equal(sz[i], oz)


def test_sampler_matches_golden(main_decoder):
    z = golden_npz("trials_p0.05_it50.npz")
    base, first, count, _ = map(int, z["meta"])
    err, sx, sz = main_decoder.sample(base, first, count, 0.05)
    assert [fnv1a(e) for e in err] == [int(h) for h in z["err_hash"]]
    assert [fnv1a(s) for s in sx] == [int(h) for h in z["sx_hash"]]
    assert [fnv1a(s) for s in sz] == [int(h) for h in z["sz_hash"]]
    for i, packed in enumerate(z["err_packed"]):
        assert np.array_equal(err[i], unpack2(packed))


def test_sampler_far_trial_indices(main_decoder, main_code_oracle):
    # large trial indices and another base seed (64-bit seed arithmetic)
    first = 10 ** 12
    err, _, _ = main_decoder.sample(2 ** 63 + 12345, first, 8, 0.04)
    for i in range(8):
        assert np.array_equal(err[i], main_code_oracle.sample(2 ** 63 + 12345, first + i, 0.04))


@pytest.mark.parametrize("name,damping", GOLDEN_SETS)
def test_fp64_simulate_matches_reference_per_trial(main_decoder, name, damping):
    z = golden_npz(name)
    base, first, count, max_iter = map(int, z["meta"])
    p = float(z["p"][0])
    print("fp64 simulate", name, flush=True)
    stats, r = main_decoder.simulate(base, count, p, max_iter=max_iter, first_trial=first,
                                     damping=damping, precision=jbp.FP64, want_est=True)
    assert stats["done"] == count
    assert np.array_equal(r["iterations"], z["iterations"])
    assert np.array_equal(r["syn_match"], z["syn_ok"].astype(np.uint8))
    assert np.array_equal(r["success"], z["success"].astype(np.uint8))
    assert np.array_equal(r["exact"], z["exact"].astype(np.uint8))
    assert [fnv1a(e) for e in r["est"]] == [int(h) for h in z["est_hash"]]


def test_fp64_cfg1_aggregate(main_decoder):
    agg = golden_json("aggregates.json")["cfg1_p0.05_it50_seed106_1000"]
    stats, r = main_decoder.simulate(agg["base"], agg["count"], agg["p"], max_iter=agg["max_iter"],
                                     first_trial=agg["first"], precision=jbp.FP64)
    assert stats["failures"] == agg["failures"]
    assert stats["total_iters"] == agg["iters"]
    assert stats["exact"] == agg["exact"]


def test_fp64_decode_single_matches_oracle(main_decoder, main_code_oracle):
    # the cuda_bp_decode-shaped entry point: one syndrome pair
    prior = jbp.prior_from_p(0.05)
    for t in (7, 11, 15):  # trial 15 does not converge within 50 iterations
        e = main_code_oracle.sample(106, t, 0.05)
        sx, sz = main_code_oracle.syndrome(e)
        want_est, want_it, want_ok = main_code_oracle.decode(sx, sz, prior, 50)
        est, it, ok, _ = main_decoder.decode(sx, sz, prior, max_iter=50, precision=jbp.FP64)
        assert (it, ok) == (want_it, want_ok)
        assert np.array_equal(est, want_est)


def test_fp64_decode_batch_matches_oracle(main_decoder, main_code_oracle):
    prior = jbp.prior_from_p(0.04)
    errs = [main_code_oracle.sample(3, t, 0.04) for t in range(1, 9)]
    syn = [main_code_oracle.syndrome(e) for e in errs]
    sx = np.array([s[0] for s in syn])
    sz = np.array([s[1] for s in syn])
    it, ok, est = main_decoder.decode_batch(sx, sz, prior, max_iter=200, precision=jbp.FP64)
    for i in range(8):
        w_est, w_it, w_ok = main_code_oracle.decode(sx[i], sz[i], prior, 200)
        assert (it[i], bool(ok[i])) == (w_it, w_ok)
        assert np.array_equal(est[i], w_est)


def test_empty_and_zero_syndrome(main_decoder):
    prior = jbp.prior_from_p(0.05)
    zx = np.zeros(2304, np.int32)
    for prec in (jbp.FP64, jbp.FP32):
        est, it, ok, _ = main_decoder.decode(zx, zx, prior, max_iter=10, precision=prec)
        assert ok and it == 1 and not est.any()
    stats, r = main_decoder.simulate(1, 0, 0.05, precision=jbp.FP32)
    assert stats["done"] == 0


def test_row_space_membership(main_decoder, main_code_oracle):
    rng = np.random.default_rng(1)
    code = main_decoder.code
    vecs = []
    for side in (0, 1):
        H = code.checks(side)
        for k in range(4):
            v = np.zeros(code.n, np.int32)
            for c in rng.choice(H.shape[0], 1 + 3 * k, replace=False):
                v[H[c]] ^= 1
            vecs.append((side, v))              # member
            w = v.copy()
            w[rng.integers(code.n)] ^= 1
            vecs.append((side, w))              # one flip off
    for side, v in vecs:
        got = main_decoder.in_row_space(side, v)[0]
        assert got == main_code_oracle.in_row_space(side, v)


def test_min_fails_prefix(main_decoder):
    # --min-fails stops at the first trial index where failures reach N
    # (jointbp_ets.cpp:6409-6411); trials beyond it are not counted.
    z = golden_npz("trials_p0.05_it50.npz")
    fails = np.cumsum(1 - z["success"])
    stop = int(np.argmax(fails >= 3)) + 1
    stats, r = main_decoder.simulate(106, 256, 0.05, max_iter=50, precision=jbp.FP64, min_fails=3)
    assert stats["done"] == stop and stats["failures"] == 3


def test_fp32_matches_reference_fer_and_trials(main_decoder):
    agg = golden_json("aggregates.json")["cfg1_p0.05_it50_seed106_1000"]
    stats, r = main_decoder.simulate(agg["base"], agg["count"], agg["p"], max_iter=agg["max_iter"],
                                     precision=jbp.FP32)
    lo, hi = wilson(agg["count"], agg["failures"])
    fer = stats["failures"] / stats["done"]
    print("fp32 fer", fer, "ref", agg["failures"] / agg["count"], (lo, hi), flush=True)
    assert lo <= fer <= hi
    z = golden_npz("trials_p0.05_it50.npz")
    n = len(z["iterations"])
    same_it = np.mean(r["iterations"][:n] == z["iterations"])
    same_ok = np.mean(r["success"][:n] == z["success"])
    print("fp32 agreement: iterations %.3f success %.3f" % (same_it, same_ok), flush=True)
    assert same_ok >= 0.98 and same_it >= 0.8
    assert abs(stats["total_iters"] - agg["iters"]) / agg["iters"] < 0.03


def test_fp32_decode_single_agrees_with_oracle(main_decoder, main_code_oracle):
    prior = jbp.prior_from_p(0.05)
    agree = 0
    for t in range(1, 21):
        e = main_code_oracle.sample(106, t, 0.05)
        sx, sz = main_code_oracle.syndrome(e)
        w_est, w_it, w_ok = main_code_oracle.decode(sx, sz, prior, 50)
        est, it, ok, _ = main_decoder.decode(sx, sz, prior, max_iter=50, precision=jbp.FP32)
        agree += int(np.array_equal(est, w_est) and it == w_it)
    assert agree >= 16


def test_fp32_check_interval_semantics(main_decoder):
    # with a check interval the reported iteration is the first tested one
    stats1, r1 = main_decoder.simulate(106, 64, 0.05, max_iter=50, precision=jbp.FP32)
    stats5, r5 = main_decoder.simulate(106, 64, 0.05, max_iter=50, precision=jbp.FP32,
                                       check_interval=5, check_warmup=2)
    conv = r1["syn_match"].astype(bool) & r5["syn_match"].astype(bool)
    assert np.all(r5["iterations"][conv] >= r1["iterations"][conv])
    assert np.all(((r5["iterations"][conv] - 2) % 5 == 0) | (r5["iterations"][conv] == 50))


@pytest.mark.parametrize("precision", [jbp.FP32, jbp.FP64])
def test_single_decode_graph_replay(main_decoder, main_code_oracle, precision):
    # use_graph replays a captured graph (syndrome copies, counter fill,
    # decode, result copies).  Every replay must equal the eager decode of the
    # same syndromes, with eager work and a re-capture (new max_iter) between
    # replays; FP64 must also equal the oracle.
    prior = jbp.prior_from_p(0.05)
    print("graph replay", precision, flush=True)
    for rep, t in enumerate((7, 11, 15, 7, 23, 11)):
        e = main_code_oracle.sample(106, t, 0.05)
        sx, sz = main_code_oracle.syndrome(e)
        max_iter = 50 if rep < 4 else 40
        g_est, g_it, g_ok, _ = main_decoder.decode(sx, sz, prior, max_iter=max_iter,
                                                   precision=precision, use_graph=True)
        main_decoder.simulate(5, 64, 0.03, max_iter=20, precision=jbp.FP32, per_trial=False)
        est, it, ok, _ = main_decoder.decode(sx, sz, prior, max_iter=max_iter, precision=precision)
        assert (g_it, g_ok) == (it, ok) and np.array_equal(g_est, est), (rep, t)
        if precision == jbp.FP64:
            w_est, w_it, w_ok = main_code_oracle.decode(sx, sz, prior, max_iter)
            assert (g_it, g_ok) == (w_it, w_ok) and np.array_equal(g_est, w_est)


@pytest.mark.parametrize("n,chunk,warmup,interval", [(3 * 8192 + 77, 0, 0, 1), (3 * 8192 + 77, 1024, 0, 1),
                                                    (3 * 8192 + 77, 0, 4, 3), (2 * 8192 + 77, 0, 0, 1)])
def test_pipelined_run_equals_serial_run(main_decoder, n, chunk, warmup, interval):
    # Runs longer than one launch are pipelined: a sampling stream, a decode
    # stream and a finishing stream over eight buffer sets (chunk 1024: 25
    # chunks, so the buffer sets are reused three times over).  From four
    # chunks on, the decode launch stops at the first scheduled syndrome
    # test from 32 iterations on (34 with warmup 4, interval 3) and the
    # finishing stream decodes what is left from the prior to max-iter;
    # three chunks decode whole on the two streams in turn.  The per-trial
    # outcome must equal the serial path (one launch, taken when estimates
    # are requested), and min_fails must cut at the same trial index.
    print("pipelined vs serial", n, chunk, warmup, interval, flush=True)
    kw = dict(check_warmup=warmup, check_interval=interval)
    sa, a = main_decoder.simulate(106, n, 0.05, max_iter=100, precision=jbp.FP32, want_est=True, **kw)
    sb, b = main_decoder.simulate(106, n, 0.05, max_iter=100, precision=jbp.FP32, chunk=chunk, **kw)
    # trials handed to the finishing stream, some of them running to max-iter
    assert np.sum(b["iterations"] > 34) > 10 and np.sum(b["iterations"] == 100) > 0
    for k in ("iterations", "syn_match", "success", "exact"):
        assert np.array_equal(a[k], b[k]), k
    assert (sa["done"], sa["failures"], sa["total_iters"]) == (sb["done"], sb["failures"], sb["total_iters"])
    limit = max(2, int(sa["failures"]) // 2)
    sa, a = main_decoder.simulate(106, n, 0.05, max_iter=100, precision=jbp.FP32, want_est=True,
                                  min_fails=limit, **kw)
    sb, b = main_decoder.simulate(106, n, 0.05, max_iter=100, precision=jbp.FP32, min_fails=limit,
                                  chunk=chunk, **kw)
    assert sa["done"] == sb["done"] < n and sa["failures"] == sb["failures"] == limit
    assert np.array_equal(a["success"], b["success"])


def test_fp64_freeze_syn_matches_reference(main_decoder, main_code_oracle):
    # --freeze-syn (jointbp_ets.cpp:1884-1901, 2526-2535) on the FP64 decoder:
    # per decode the reference's iterations, match and estimate, on X-only
    # errors (the Z side matches and freezes first); freezing changes some
    # of these decodes, and FP32 refuses it
    z = golden_npz("decodes_xonly_p0.04_it60_freeze.npz")
    base, _, count, max_iter = map(int, z["meta"])
    p, damping = z["p"]
    prior = jbp.prior_from_p(p)
    for i in range(count):
        err = main_code_oracle.sample(base, 1 + i, p) & 1
        sx, sz = main_code_oracle.syndrome(err)
        print("freeze decode", i, flush=True)
        est, it, ok, _ = main_decoder.decode(sx, sz, prior, max_iter=max_iter, precision=jbp.FP64,
                                             freeze_syn=True, damping=damping)
        assert (it, int(ok), fnv1a(est)) == (int(z["iterations"][i]), int(z["syn_ok"][i]),
                                             int(z["est_hash"][i])), i
        est, it, ok, _ = main_decoder.decode(sx, sz, prior, max_iter=max_iter, precision=jbp.FP64,
                                             damping=damping)
        assert (it, fnv1a(est)) == (int(z["iterations_nofreeze"][i]), int(z["est_hash_nofreeze"][i]))
    with pytest.raises(jbp.JbpError):
        main_decoder.decode(sx, sz, prior, max_iter=10, precision=jbp.FP32, freeze_syn=True)


def test_fp64_freeze_syn_simulate_matches_oracle(main_decoder, main_code_oracle):
    # the --simulate path with --freeze-syn: per trial the oracle's decoder
    # with the same flag
    n, p, it = 48, 0.05, 50
    _, r = main_decoder.simulate(106, n, p, max_iter=it, precision=jbp.FP64, freeze_syn=True,
                                 want_est=True)
    prior = jbp.prior_from_p(p)
    for t in range(n):
        err = main_code_oracle.sample(106, 1 + t, p)
        sx, sz = main_code_oracle.syndrome(err)
        est, k, ok = main_code_oracle.decode(sx, sz, prior, it, 0.0, freeze_syn=True)
        assert (r["iterations"][t], bool(r["syn_match"][t])) == (k, ok)
        assert np.array_equal(r["est"][t], est)

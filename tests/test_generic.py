"""GPU: interleaved mixed-noise runs, the generic HBM-resident FP32 decoder,
and hypergraph-product codes (success judged by logical operators)."""
import numpy as np
import pytest

import joint_bp_plus_pp_amd as jbp

pytestmark = pytest.mark.gpu

PS = [0.01, 0.025, 0.04, 0.05]


@pytest.fixture(scope="module")
def small_hgp():
    code = jbp.Code.hgp(15, 20, 3, 4, seed=2)
    dec = jbp.Decoder(code, 0)
    yield code, dec
    dec.destroy()


@pytest.fixture(scope="module")
def big_hgp():
    code = jbp.Code.hgp(150, 200, 3, 4, seed=1)
    dec = jbp.Decoder(code, 0)
    yield code, dec
    dec.destroy()


def test_sampling_runs_in_pieces_past_the_generator_scratch(small_hgp):
    # the sampler seeds at most 65,536 generator states at a time; a longer
    # call runs in pieces, and every trial is still drawn from its own
    # index's stream (the trials on both sides of the piece boundary equal a
    # short call over the same indices, which the oracle tests pin)
    code, dec = small_hgp
    n = 65536 + 40
    print("sampling %d trials in pieces" % n, flush=True)
    err, sx, sz = dec.sample(7, 3, n, 0.05)
    e1, x1, z1 = dec.sample(7, 3 + 65536 - 20, 60, 0.05)
    assert np.array_equal(err[65516:], e1) and np.array_equal(sx[65516:], x1) and np.array_equal(sz[65516:], z1)
    e0, _, _ = dec.sample(7, 3, 16, 0.05)
    assert np.array_equal(err[:16], e0)
    assert 0.02 < np.mean(err != 0) < 0.08


@pytest.mark.parametrize("order", [0, 1])
def test_mixed_codes_equal_separate_hbm_runs(main_decoder, small_hgp, order):
    # two codes in one engine (either code first; the HGP code with damping
    # 0.3, the included one with an interleaved p list and a check schedule):
    # per trial the same as each code's own HBM-engine run
    code, hdec = small_hgp
    qc = (main_decoder, dict(base_seed=106, n_trials=1500, p=PS, max_iter=60, first_trial=7,
                             check_warmup=2, check_interval=3))
    hg = (hdec, dict(base_seed=3, n_trials=300, p=0.06, max_iter=40, damping=0.3, slots=128))
    runs = [qc, hg] if order == 0 else [hg, qc]
    print("mixed codes, order", order, flush=True)
    res = jbp.simulate_mixed(runs)
    for (dec, kw), (st, r) in zip(runs, res):
        _, a = dec.simulate(kw["base_seed"], kw["n_trials"], kw["p"], max_iter=kw["max_iter"],
                            first_trial=kw.get("first_trial", 1), damping=kw.get("damping", 0.0),
                            check_warmup=kw.get("check_warmup", 0),
                            check_interval=kw.get("check_interval", 1), engine=jbp.ENGINE_HBM)
        assert per_trial_equal(r, a), kw
        assert st["done"] == kw["n_trials"]
    # a code with no trials takes no part
    (s0, r0), (s1, r1) = jbp.simulate_mixed([(main_decoder, dict(base_seed=106, n_trials=0, p=0.04)),
                                             (hdec, dict(base_seed=3, n_trials=64, p=0.06, max_iter=40))])
    assert s0["done"] == 0 and s1["done"] == 64


def per_trial_equal(a, b, keys=("iterations", "syn_match", "success", "exact")):
    return all(np.array_equal(a[k], b[k]) for k in keys)


@pytest.mark.parametrize("precision,engine", [(jbp.FP64, 0), (jbp.FP32, jbp.ENGINE_LDS),
                                              (jbp.FP32, jbp.ENGINE_HBM)])
def test_mixed_p_matches_single_p_runs(main_decoder, precision, engine):
    # trial index idx uses PS[idx % 4]: each trial must equal the same index
    # of a single-p run at its p (same seed, same draws, same decode)
    first, n, it = 5, 512 if precision == jbp.FP64 else 2048, 60
    print("mixed p", precision, engine, flush=True)
    _, mixed = main_decoder.simulate(106, n, PS, max_iter=it, first_trial=first,
                                     precision=precision, engine=engine, want_est=True)
    idx = np.arange(first, first + n)
    for j, p in enumerate(PS):
        _, single = main_decoder.simulate(106, n, p, max_iter=it, first_trial=first,
                                          precision=precision, engine=engine, want_est=True)
        sel = idx % 4 == j
        for k in ("iterations", "syn_match", "success", "exact", "est"):
            assert np.array_equal(mixed[k][sel], single[k][sel]), (p, k)


def test_mixed_p_fp64_matches_oracle(main_decoder, main_code_oracle):
    n = 96
    _, r = main_decoder.simulate(106, n, PS, max_iter=50, first_trial=1, precision=jbp.FP64)
    for t in range(n):
        idx = 1 + t
        p = PS[idx % 4]
        err = main_code_oracle.sample(106, idx, p)
        sx, sz = main_code_oracle.syndrome(err)
        est, it, ok = main_code_oracle.decode(sx, sz, jbp.prior_from_p(p), 50)
        assert r["iterations"][t] == it and bool(r["syn_match"][t]) == ok
        assert bool(r["success"][t]) == (ok and main_code_oracle.stab_equiv(err, est))


def test_hbm_engine_is_slot_invariant(main_decoder):
    # a trial's decode does not depend on the slot it ran in or on S
    print("slot invariance", flush=True)
    runs = [main_decoder.simulate(106, 3000, 0.05, max_iter=80, precision=jbp.FP32,
                                  engine=jbp.ENGINE_HBM, slots=s, want_est=True)[1]
            for s in (64, 640, 2048)]
    for r in runs[1:]:
        assert per_trial_equal(runs[0], r) and np.array_equal(runs[0]["est"], r["est"])


def test_hbm_engine_agrees_with_lds_engine(main_decoder):
    # two FP32 implementations of the same arithmetic in different operation
    # order: per-trial outcomes agree except for rare trials near a decision
    n = 8192
    print("hbm vs lds", flush=True)
    sh, rh = main_decoder.simulate(106, n, 0.05, max_iter=100, precision=jbp.FP32,
                                   engine=jbp.ENGINE_HBM)
    sl, rl = main_decoder.simulate(106, n, 0.05, max_iter=100, precision=jbp.FP32,
                                   engine=jbp.ENGINE_LDS)
    assert np.mean(rh["success"] == rl["success"]) > 0.995
    assert np.mean(rh["iterations"] == rl["iterations"]) > 0.98
    assert abs(sh["failures"] - sl["failures"]) <= 3 * np.sqrt(max(sl["failures"], 1)) + 3
    print("failures hbm %d lds %d" % (sh["failures"], sl["failures"]), flush=True)


def test_small_hgp_fp64_matches_oracle(small_hgp, oracle):
    # FP64 decoder on a hypergraph-product code, success judged by logical
    # operators on the device, against the oracle (its own row-space test)
    code, dec = small_hgp
    assert dec.info()["basis"] == 2  # logical operators installed
    oc = oracle.code_from_checks(code.n, code.checks(0), code.checks(1))
    n, p, it = 400, 0.06, 40
    _, r = dec.simulate(11, n, p, max_iter=it, precision=jbp.FP64, want_est=True)
    o = oc.simulate(11, 1, n, p, it, threads=8)
    assert np.array_equal(r["iterations"], o["iterations"])
    assert np.array_equal(r["syn_match"], o["syn_ok"])
    assert np.array_equal(r["success"], o["success"])
    assert np.array_equal(r["exact"], o["exact"])
    assert 0 < int(np.sum(r["syn_match"] & (1 - r["exact"])))  # residual tests exercised


def test_small_hgp_logicals_equal_row_basis(small_hgp):
    code, dec = small_hgp
    _, a = dec.simulate(12, 600, 0.07, max_iter=30, precision=jbp.FP64)
    dec.set_logicals(None, None)
    assert dec.info()["basis"] == 1
    _, b = dec.simulate(12, 600, 0.07, max_iter=30, precision=jbp.FP64)
    dec.set_logicals(code.logicals(0), code.logicals(1))
    assert per_trial_equal(a, b)
    assert np.any(a["syn_match"] & (1 - a["success"]))  # some logical failures seen


def test_small_hgp_fp32_hbm_vs_fp64(small_hgp):
    _, dec = small_hgp
    n = 4096
    s32, r32 = dec.simulate(13, n, 0.05, max_iter=60, precision=jbp.FP32)
    s64, r64 = dec.simulate(13, n, 0.05, max_iter=60, precision=jbp.FP64)
    assert np.mean(r32["success"] == r64["success"]) > 0.99
    assert abs(s32["failures"] - s64["failures"]) <= 3 * np.sqrt(max(s64["failures"], 1)) + 3


def test_big_hgp_hbm_engine(big_hgp):
    # the 62,500-qubit product code at p = 0.03, max-iter 200
    code, dec = big_hgp
    assert dec.info()["basis"] == 2
    n = 1024
    print("big hgp fp32", flush=True)
    s32, r32 = dec.simulate(21, n, 0.03, max_iter=200, precision=jbp.FP32)
    assert s32["done"] == n
    assert np.all((r32["iterations"] >= 0) & (r32["iterations"] <= 200))
    assert np.all(r32["success"] <= r32["syn_match"])
    assert np.all(r32["iterations"][r32["syn_match"] == 0] == 200)
    print("big hgp fp64", flush=True)
    m = 96
    _, r64 = dec.simulate(21, m, 0.03, max_iter=200, precision=jbp.FP64)
    agree = np.mean(r64["success"] == r32["success"][:m])
    print("big hgp: FER fp32 %.4f, agreement with fp64 on %d trials %.3f"
          % (s32["failures"] / n, m, agree), flush=True)
    assert agree > 0.95


def test_tune_picks_a_candidate_and_keeps_results(small_hgp):
    _, dec = small_hgp
    print("tune fp32", flush=True)
    slots, block, ms = dec.tune(31, 1024, 0.05, max_iter=60, precision=jbp.FP32)
    assert slots % 64 == 0 and 64 <= slots <= 1024 and block == 0
    assert len(ms) >= 2 and min(ms) > 0
    _, a = dec.simulate(31, 1024, 0.05, max_iter=60, precision=jbp.FP32, want_est=True)
    _, b = dec.simulate(31, 1024, 0.05, max_iter=60, precision=jbp.FP32, slots=64, want_est=True)
    assert per_trial_equal(a, b) and np.array_equal(a["est"], b["est"])
    print("tune fp64", flush=True)
    slots, block, ms = dec.tune(32, 256, 0.05, max_iter=40, precision=jbp.FP64)
    assert block in (64, 128, 256, 512) and slots > 0 and len(ms) >= 4
    _, a = dec.simulate(32, 256, 0.05, max_iter=40, precision=jbp.FP64)
    _, b = dec.simulate(32, 256, 0.05, max_iter=40, precision=jbp.FP64, block=64)
    assert per_trial_equal(a, b)


def test_fp64_workgroup_sizes_keep_results(main_decoder):
    """On the included (quasi-cyclic) code the FP64 decoder stages its
    variable pass through LDS, each thread copying its own rows in and out;
    every workgroup size it accepts decodes every trial the same."""
    dec = main_decoder
    _, a = dec.simulate(41, 96, 0.05, max_iter=30, precision=jbp.FP64, want_est=True)
    for bs in (64, 128, 512, 256):  # (the context keeps the last size: 256, its default)
        print("fp64 block", bs, flush=True)
        _, b = dec.simulate(41, 96, 0.05, max_iter=30, precision=jbp.FP64, block=bs, want_est=True)
        assert per_trial_equal(a, b) and np.array_equal(a["est"], b["est"])


def test_tune_lds_engine_has_nothing_to_tune(main_decoder):
    assert main_decoder.tune(1, 256, 0.05, max_iter=20, precision=jbp.FP32)[:2] == (0, 0)


def irregular_hgp_checks(m, n, seed):
    """Hypergraph product of a random irregular m x n matrix (column weights
    2..4, so check degrees and variable degrees vary), as check lists."""
    rng = np.random.default_rng(seed)
    H = np.zeros((m, n), np.uint8)
    for j in range(n):
        H[rng.choice(m, size=2 + j % 3, replace=False), j] = 1
    rows = [np.nonzero(H[a])[0] for a in range(m)]
    cols = [np.nonzero(H[:, j])[0] for j in range(n)]
    xs, zs = [], []
    for a in range(m):          # X check (a, j): (i, j) for i in row a; (a, b) for b in column j
        for j in range(n):
            xs.append([i * n + j for i in rows[a]] + [n * n + a * m + b for b in cols[j]])
    for i in range(n):          # Z check (i, b): (i, j) for j in row b; (a, b) for a in column i
        for b in range(m):
            zs.append([i * n + j for j in rows[b]] + [n * n + a * m + b for a in cols[i]])
    width = max(max(map(len, xs)), max(map(len, zs)))
    pad = lambda L: np.array([c + [-1] * (width - len(c)) for c in L], np.int32)
    return n * n + m * m, pad(xs), pad(zs)


def test_irregular_code_hbm_engine_vs_fp64(oracle):
    # variable check degrees (the generic check pass without a fixed degree)
    # and variable degrees up to 8 per side
    nq, xs, zs = irregular_hgp_checks(8, 12, seed=3)
    code = jbp.Code.from_checks(nq, xs, zs)
    dec = jbp.Decoder(code, 0)
    try:
        info = code.diagnostics(girth=False)
        assert info["x_check_deg_min"] < info["x_check_deg_max"]
        print("irregular code", nq, info["x_check_deg_min"], info["x_check_deg_max"], flush=True)
        n = 3000
        s32, r32 = dec.simulate(41, n, 0.04, max_iter=60, precision=jbp.FP32, engine=jbp.ENGINE_HBM,
                                want_est=True)
        s64, r64 = dec.simulate(41, n, 0.04, max_iter=60, precision=jbp.FP64, want_est=True)
        assert np.mean(r32["success"] == r64["success"]) > 0.99
        assert np.mean(np.all(r32["est"] == r64["est"], axis=1)) > 0.97
        assert abs(s32["failures"] - s64["failures"]) <= 3 * np.sqrt(max(s64["failures"], 1)) + 3
        oc = oracle.code_from_checks(nq, xs, zs)
        o = oc.simulate(41, 1, 200, 0.04, 60, threads=8)
        assert np.array_equal(r64["iterations"][:200], o["iterations"])
        assert np.array_equal(r64["success"][:200], o["success"])
    finally:
        dec.destroy()


@pytest.mark.parametrize("precision,engine,n", [(jbp.FP32, jbp.ENGINE_LDS, 3 * 8192 + 5),
                                                (jbp.FP32, jbp.ENGINE_LDS, 3000),
                                                (jbp.FP32, jbp.ENGINE_HBM, 3000),
                                                (jbp.FP64, 0, 600)])
def test_decode_order_does_not_change_results(main_decoder, precision, engine, n):
    # trials are handed out heaviest syndrome first by default; index order
    # must give the same per-trial outcome (the first case runs pipelined)
    print("order", precision, engine, n, flush=True)
    _, a = main_decoder.simulate(106, n, 0.05, max_iter=80, precision=precision, engine=engine,
                                 schedule=0)
    _, b = main_decoder.simulate(106, n, 0.05, max_iter=80, precision=precision, engine=engine,
                                 schedule=1)
    assert per_trial_equal(a, b)


def test_hbm_engine_graph_replay_equals_eager(main_decoder, small_hgp):
    # With graph=True the engine replays a captured step batch; every such
    # call replays it many times, with eager calls (and their
    # synchronizations) in between, and must equal the launched steps per
    # trial.
    _, dec = small_hgp
    for d, n, p in ((main_decoder, 3000, 0.05), (dec, 3000, 0.06)):
        for rep in range(2):
            print("graph vs eager", n, p, rep, flush=True)
            _, g = d.simulate(51 + rep, n, p, max_iter=80, precision=jbp.FP32,
                              engine=jbp.ENGINE_HBM, slots=128, want_est=True, graph=True)
            _, e = d.simulate(51 + rep, n, p, max_iter=80, precision=jbp.FP32,
                              engine=jbp.ENGINE_HBM, slots=128, want_est=True)
            assert per_trial_equal(g, e) and np.array_equal(g["est"], e["est"])


def test_fp32_damping_matches_fp64_fer(main_decoder):
    # damping 0.3 (the reference's scaled_mix, here on the HBM engine's
    # binary message pairs) against the FP64 decoder, which is per-trial the
    # reference's (trials_p0.05_it50_damp0.3 fixture), on the same indices
    import math
    from conftest import golden_npz
    z = golden_npz("trials_p0.05_it50_damp0.3.npz")
    n = 4096
    print("fp32 damping 0.3", flush=True)
    s32, r32 = main_decoder.simulate(106, n, 0.05, max_iter=50, damping=0.3, precision=jbp.FP32)
    s64, r64 = main_decoder.simulate(106, n, 0.05, max_iter=50, damping=0.3, precision=jbp.FP64)
    m = len(z["success"])
    assert np.array_equal(r64["success"][:m], z["success"])
    assert np.array_equal(r64["iterations"][:m], z["iterations"])
    k = s64["failures"]
    phat, zz = k / n, 1.96
    den = 1 + zz * zz / n
    c = (phat + zz * zz / (2 * n)) / den
    h = zz * math.sqrt((phat * (1 - phat) + zz * zz / (4 * n)) / n) / den
    fer32 = s32["failures"] / n
    agree = float(np.mean(r32["success"] == r64["success"]))
    same_it = float(np.mean(r32["iterations"] == r64["iterations"]))
    print("damping 0.3: FER fp32 %.4f fp64 %.4f ci95 [%.4f, %.4f]; agreement %.4f, same "
          "iterations %.4f; avg_iter %.3f vs %.3f" % (fer32, k / n, c - h, c + h, agree, same_it,
                                                      s32["total_iters"] / n, s64["total_iters"] / n),
          flush=True)
    assert c - h <= fer32 <= c + h
    assert agree >= 0.98 and same_it >= 0.8
    # the LDS engine keeps no previous messages: it refuses damping
    with pytest.raises(jbp.JbpError):
        main_decoder.simulate(106, 64, 0.05, max_iter=50, damping=0.3, precision=jbp.FP32,
                              engine=jbp.ENGINE_LDS)

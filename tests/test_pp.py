"""Post-processing on the device path: jbp_simulate with pp=1 against the
reference's simulate loop with PP on (jointbp_ets.cpp:5931-6228, decoder
1943-3079), per trial.  The fixtures come from the reference's own code
(tests/golden/make_pp_golden.py)."""
import hashlib

import numpy as np
import pytest

from conftest import PARAMS, golden_npz, hip_device_available

jbp = pytest.importorskip("joint_bp_plus_pp_amd")

WORKLOADS = ["p0.04_it8", "p0.03_it5", "p0.04_it8_hist0", "p0.02_it1", "p0.05_it115",
             "p0.04_it10_damp0.3"]
COUNTERS = ["pp_success", "flip_pp_success", "osd_pp_success", "stab_success", "ets_used",
            "ets_saved", "ets6_used", "ets6_saved", "ets12_used", "ets12_saved"]


def digest(est):
    return hashlib.sha1(np.asarray(est, dtype=np.uint8).tobytes()).hexdigest().encode()


@pytest.fixture(scope="module")
def pp_decoder(main_decoder):
    tables = jbp.PPTables.from_dir(main_decoder.code, jbp.default_ets_dir(PARAMS))
    main_decoder.set_pp_tables(tables)
    yield main_decoder
    main_decoder.set_pp_tables(None)


def run(dec, g, precision, **kw):
    return dec.simulate(int(g["seed"]), len(g["iterations"]), float(g["p"]),
                        max_iter=int(g["max_iter"]), first_trial=int(g["first_trial"]),
                        precision=precision, pp=True, flip_hist=int(g["flip_hist"]), want_est=True,
                        damping=float(g["damping"]), **kw)


def check_equal(stats, per, g):
    n_bad = int(np.count_nonzero(per["iterations"] != g["iterations"]))
    assert n_bad == 0, "iterations differ on %d trials" % n_bad
    np.testing.assert_array_equal(per["syn_match"], g["syn_match"])
    np.testing.assert_array_equal(per["success"], g["success"])
    np.testing.assert_array_equal(per["exact"], g["exact"])
    np.testing.assert_array_equal(per["pp_flags"], g["flags"])
    got = np.array([digest(e) for e in per["est"]])
    np.testing.assert_array_equal(got, g["est_sha1"])
    assert [int(stats[k]) for k in COUNTERS] == g["counters"].tolist()
    assert int(stats["failures"]) == int(np.count_nonzero(g["success"] == 0))
    assert int(stats["bp_failures"]) == int(np.count_nonzero((g["flags"] & 1) == 0))


@pytest.mark.gpu
@pytest.mark.parametrize("name", WORKLOADS)
def test_pp_fp64_equals_reference_per_trial(pp_decoder, name):
    g = golden_npz("pp_trials_%s.npz" % name)
    print("pp fp64", name, flush=True)
    stats, per = run(pp_decoder, g, jbp.FP64)
    check_equal(stats, per, g)


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["hgp15x20", "hgp21x28"])
def test_pp_fp64_hgp_equals_reference_per_trial(name):
    """Small HGP codes without tables; HGP(21x28) has 19 syndrome words per
    side, an odd count (the stage copies its syndromes trial by trial)."""
    if not hip_device_available():
        pytest.skip("no HIP device")
    g = golden_npz("pp_trials_%s.npz" % name)
    m, n, cw, rw, seed = (int(v) for v in g["hgp"])
    code = jbp.Code.hgp(m, n, cw, rw, seed)
    with jbp.Decoder(code, 0) as dec:
        print("pp fp64 hgp", m, n, flush=True)
        stats, per = run(dec, g, jbp.FP64)
        check_equal(stats, per, g)


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["p0.04_it8", "p0.03_it5"])
def test_pp_fp32_first_pass(pp_decoder, name):
    """FP32 decodes the first min(99, max_iter - 1) iterations; the trials it
    leaves unconverged go through the same FP64 stage.  Outcomes agree with
    the reference except where FP32 and FP64 rounding end a decode at a
    different iteration."""
    g = golden_npz("pp_trials_%s.npz" % name)
    print("pp fp32", name, flush=True)
    stats, per = run(pp_decoder, g, jbp.FP32)
    agree = np.mean((per["success"] == g["success"]) & (per["pp_flags"] == g["flags"]))
    assert agree >= 0.97, agree
    assert stats["pp_trials"] > 0
    want = g["counters"][0]
    assert abs(int(stats["pp_success"]) - int(want)) <= max(3, 0.05 * want)


@pytest.mark.gpu
def test_pp_on_leaves_early_convergers_alone(pp_decoder):
    """A trial that converges before iteration 100 never reaches a PP
    trigger: its outcome with pp on is the outcome with pp off."""
    kw = dict(max_iter=300, precision=jbp.FP32, want_est=True)
    print("pp on/off p0.045", flush=True)
    s_off, off = pp_decoder.simulate(31, 4096, 0.045, **kw)
    s_on, on = pp_decoder.simulate(31, 4096, 0.045, pp=True, **kw)
    early = off["iterations"] < 100
    assert early.mean() > 0.99
    for k in ("iterations", "syn_match", "success", "exact"):
        np.testing.assert_array_equal(on[k][early], off[k][early])
    np.testing.assert_array_equal(on["est"][early], off["est"][early])
    assert s_on["pp_trials"] == int(np.count_nonzero(~early))


@pytest.mark.gpu
def test_pp_min_fails_and_argument_errors(pp_decoder):
    g = golden_npz("pp_trials_p0.04_it8.npz")
    fails = np.cumsum(g["success"] == 0)
    limit = 5
    stop = int(np.argmax(fails >= limit)) + 1
    stats, per = pp_decoder.simulate(int(g["seed"]), len(g["iterations"]), float(g["p"]),
                                     max_iter=int(g["max_iter"]), precision=jbp.FP64, pp=True,
                                     min_fails=limit)
    assert int(stats["done"]) == stop
    np.testing.assert_array_equal(per["success"], g["success"][:stop])
    with pytest.raises(jbp.JbpError):
        pp_decoder.simulate(1, 8, 0.04, max_iter=8, pp=True, flip_hist=-1)


@pytest.mark.gpu
def test_decode_batch_pp_equals_simulate(pp_decoder):
    """The same syndromes through jbp_decode_batch_pp give the simulate
    loop's per-trial outcome (iterations, match, estimate, repair flags)."""
    g = golden_npz("pp_trials_p0.04_it8.npz")
    count = 64
    err, sx, sz = pp_decoder.sample(int(g["seed"]), 1, count, float(g["p"]))
    print("decode_batch_pp", count, flush=True)
    it, ok, est, flags = pp_decoder.decode_batch(sx, sz, jbp.prior_from_p(float(g["p"])),
                                                 max_iter=int(g["max_iter"]), precision=jbp.FP64,
                                                 pp=True)
    np.testing.assert_array_equal(it, g["iterations"][:count])
    np.testing.assert_array_equal(ok, g["syn_match"][:count])
    np.testing.assert_array_equal(np.array([digest(e) for e in est]), g["est_sha1"][:count])
    # where the decode succeeded the flags are the reference's
    succ = g["success"][:count] == 1
    np.testing.assert_array_equal(flags[succ], g["flags"][:count][succ])


@pytest.mark.gpu
@pytest.mark.parametrize("max_iter", [8, 50])
def test_decode_llr_equals_reference(main_decoder, max_iter):
    """|LLR| of the last iteration (what OSD orders by), bit for bit with the
    reference's abs_llr_x / abs_llr_z, together with its iterations and
    estimate."""
    g = golden_npz("llr_decodes.npz")
    sel = np.nonzero(g["max_iter"] == max_iter)[0]
    trials = g["trial"][sel]
    print("decode_llr", max_iter, flush=True)
    err, sx, sz = main_decoder.sample(int(g["seed"]), int(trials[0]), len(trials), float(g["p"]))
    assert list(trials) == list(range(int(trials[0]), int(trials[0]) + len(trials)))
    it, ok, est, llr = main_decoder.decode_llr(sx, sz, jbp.prior_from_p(float(g["p"])),
                                               max_iter=max_iter)
    np.testing.assert_array_equal(it, g["iterations"][sel])
    np.testing.assert_array_equal(ok, g["syn_match"][sel])
    np.testing.assert_array_equal(np.array([digest(e) for e in est]), g["est_sha1"][sel])
    np.testing.assert_array_equal(llr[:, 0, :8], g["llr_head"][sel])
    got = np.array([hashlib.sha1(np.ascontiguousarray(l).tobytes()).hexdigest().encode() for l in llr])
    np.testing.assert_array_equal(got, g["llr_sha1"][sel])


@pytest.mark.gpu
def test_pp_interleaved_p_equals_single_p(pp_decoder):
    """With a p list, trial idx uses p[idx % len]: its post-processed outcome
    equals the same index of a single-p run at that p (channel prior of the
    stage's FP64 decode included)."""
    ps = [0.03, 0.04]
    n = 192
    print("pp interleaved", flush=True)
    _, mixed = pp_decoder.simulate(5, n, ps, max_iter=8, precision=jbp.FP64, pp=True, want_est=True)
    for k, p in enumerate(ps):
        _, single = pp_decoder.simulate(5, n, p, max_iter=8, precision=jbp.FP64, pp=True,
                                        want_est=True)
        idx = np.arange(n)
        sel = ((idx + 1) % len(ps)) == k  # trial index = first_trial + i = i + 1
        for key in ("iterations", "syn_match", "success", "exact", "pp_flags"):
            np.testing.assert_array_equal(mixed[key][sel], single[key][sel], err_msg=key)
        np.testing.assert_array_equal(mixed["est"][sel], single["est"][sel])


@pytest.mark.gpu
def test_pp_after_hbm_engine_first_pass(pp_decoder):
    """The HBM-resident FP32 engine as the first pass: the trials it leaves
    go through the same FP64 stage (outcomes agree with the LDS engine's
    first pass except where the two FP32 engines round differently)."""
    g = golden_npz("pp_trials_p0.04_it8.npz")
    print("pp hbm engine", flush=True)
    s_h, h = run(pp_decoder, g, jbp.FP32, engine=jbp.ENGINE_HBM, slots=128)
    s_l, l = run(pp_decoder, g, jbp.FP32, engine=jbp.ENGINE_LDS)
    assert np.mean((h["success"] == l["success"]) & (h["pp_flags"] == l["pp_flags"])) >= 0.97
    assert s_h["pp_trials"] > 0


@pytest.mark.gpu
@pytest.mark.parametrize("p,max_iter", [(0.04, 8), (0.08, 103)], ids=["it8", "it103"])
def test_pp_many_batches_equal_single_batch_calls(pp_decoder, p, max_iter):
    """One call whose stage runs three batches of up to 1,024 trials (two
    buffer sets and streams, the next batch queued right after the first
    device phase at max-iter 8, after the last step at max-iter 103) gives,
    trial for trial, what calls small enough for one batch give on the same
    trial indices.  The one-batch path is the one the reference fixtures pin."""
    n, part = 2600, 650
    kw = dict(max_iter=max_iter, precision=jbp.FP64, pp=True, want_est=True)
    print("pp batches", p, max_iter, flush=True)
    s_all, big = pp_decoder.simulate(41, n, p, first_trial=1, **kw)
    assert s_all["pp_trials"] > 2048  # three batches
    pieces, stats = [], []
    for k in range(0, n, part):
        s, per = pp_decoder.simulate(41, part, p, first_trial=1 + k, **kw)
        assert s["pp_trials"] < 1024
        pieces.append(per)
        stats.append(s)
    for key in ("iterations", "syn_match", "success", "exact", "pp_flags"):
        np.testing.assert_array_equal(big[key], np.concatenate([q[key] for q in pieces]), err_msg=key)
    np.testing.assert_array_equal(big["est"], np.concatenate([q["est"] for q in pieces]))
    for key in COUNTERS + ["failures", "bp_failures", "pp_trials"]:
        assert int(s_all[key]) == sum(int(s[key]) for s in stats), key

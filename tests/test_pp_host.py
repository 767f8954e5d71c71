"""Post-processing on the host (no device): table loading and the three
repairs (ETS lookup, flip-set solve, OSD) of libjbp against the reference's
own routines (jointbp_ets.cpp:3199-3258, 3525-3668, 3726-3880, 4078-4386),
whose outputs on constructed estimates are the fixtures written by
tests/golden/make_pp_golden.py."""
import os

import numpy as np
import pytest

from conftest import PARAMS, golden_json, golden_npz, unpack2

jbp = pytest.importorskip("joint_bp_plus_pp_amd")


@pytest.fixture(scope="module")
def code():
    return jbp.Code.from_params(PARAMS)


@pytest.fixture(scope="module")
def tables(code):
    return jbp.PPTables.from_dir(code, jbp.default_ets_dir(PARAMS))


def test_tables_load_like_the_reference(tables):
    want = golden_json("pp_ets_counts.json")["counts"]
    got = {label: [cx, cz] for label, cx, cz in tables.loaded}
    assert got == want


def test_missing_table_files_are_reported(code):
    t = jbp.PPTables(code)
    assert t.add(6, "6", "/nonexistent/ets_6_2.txt", None) == (-1, -1)
    with pytest.raises(jbp.JbpError):
        t.add(40, "40", PARAMS, PARAMS)  # more than 32 variables per set


@pytest.mark.parametrize("method", [jbp.PP_ETS, jbp.PP_FLIP, jbp.PP_OSD], ids=["ets", "flip", "osd"])
def test_repairs_match_reference(code, tables, method):
    g = golden_npz("pp_repairs.npz")
    n = code.n
    sel = np.nonzero(g["method"] == method)[0]
    assert sel.size >= 10
    applied_any = 0
    for i in sel:
        side = int(g["side"][i])
        m = code.mx if side == 0 else code.mz
        est = unpack2(g["est_in"][i])[:n]
        syn = np.unpackbits(g["syn"][i])[:m].astype(np.int32)
        kw = {}
        if method == jbp.PP_ETS:
            kw["tables"] = tables
        elif method == jbp.PP_FLIP:
            kw["flip_vars"] = g["vars"][i][:int(g["n_vars"][i])]
        else:
            kw["abs_llr"] = g["llr_q"][i].astype(np.float64) / 64.0
        out, applied = jbp.pp_repair(code, method, side, syn, est, **kw)
        assert applied == int(g["applied"][i]), "case %d" % i
        np.testing.assert_array_equal(out, unpack2(g["est_out"][i])[:n], err_msg="case %d" % i)
        applied_any += applied != 0
    # the cases include repairs that apply and ones that do not
    assert 0 < applied_any < sel.size


def test_repair_argument_errors(code):
    est = np.zeros(code.n, np.int32)
    syn = np.zeros(code.mx, np.int32)
    with pytest.raises(jbp.JbpError):
        jbp.pp_repair(code, jbp.PP_OSD, 0, syn, est)  # OSD needs |LLR|
    with pytest.raises(jbp.JbpError):
        jbp.pp_repair(code, 7, 0, syn, est)
    bad = est.copy()
    bad[3] = 5
    with pytest.raises(jbp.JbpError):
        jbp.pp_repair(code, jbp.PP_FLIP, 0, syn, bad, flip_vars=[1, 2])


def test_osd_closed_forms(code):
    """OSD facts the mathematics fixes, independent of the fixtures: a single
    wrong qubit with the lowest |LLR| is found (its column alone spans the
    difference), and a zero difference is never 'repaired'."""
    rng = np.random.default_rng(3)
    for side in (0, 1):
        for _ in range(5):
            err = np.zeros(code.n, np.int32)
            idx = rng.choice(code.n, 20, replace=False)
            err[idx] = rng.integers(1, 4, 20)
            H = code.checks(side)
            bit = 1 if side == 0 else 2
            syn = np.array([np.bitwise_xor.reduce((err[r[r >= 0]] & bit) != 0) for r in H], np.int32)
            v = int(rng.integers(code.n))
            est = err.copy()
            est[v] ^= bit
            llr = rng.random(code.n) + 1.0
            llr[v] = 0.0
            out, applied = jbp.pp_repair(code, jbp.PP_OSD, side, syn, est, abs_llr=llr)
            assert applied == 1
            np.testing.assert_array_equal(out, err)
            out, applied = jbp.pp_repair(code, jbp.PP_OSD, side, syn, err, abs_llr=llr)
            assert applied == 0

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


def _abs_llr_libm(tot, side):
    """|LLR| as the reference forms it (jointbp_ets.cpp:2483-2493), with the C
    library's log (math.log calls it)."""
    import math
    out = np.empty(len(tot))
    for v, t in enumerate(tot.tolist()):
        p1 = t[1] + t[3] if side == 0 else t[2] + t[3]
        p0 = t[0] + t[2] if side == 0 else t[0] + t[1]
        out[v] = abs(math.log(max(p1, 1e-300)) - math.log(max(p0, 1e-300)))
    return out


def _tie_pair(side):
    """Two marginals whose pair-sum ratios are equal while their rounded
    |LLR| differ (A below B): ordering by the ratio ties them, the
    reference's values do not."""
    import itertools
    import math
    base = np.array([0.3, 0.2, 0.3, 0.2]) if side == 0 else np.array([0.3, 0.3, 0.2, 0.2])
    seen = {}
    for steps in itertools.product(range(-2, 3), repeat=4):
        t = base.copy()
        for j, k in enumerate(steps):
            for _ in range(abs(k)):
                t[j] = np.nextafter(t[j], 1.0 if k > 0 else 0.0)
        p1 = t[1] + t[3] if side == 0 else t[2] + t[3]
        p0 = t[0] + t[2] if side == 0 else t[0] + t[1]
        key = max(p1, p0) / min(p1, p0)
        L = abs(math.log(p1) - math.log(p0))
        for L2, t2 in seen.get(key, []):
            if L2 != L:
                return (t, t2) if L < L2 else (t2, t)
        seen.setdefault(key, []).append((L, t))
    raise AssertionError("no tie pair")


def test_osd_from_marginals_equals_osd_from_values(code):
    """OSD given the marginals computes the exact |LLR| only around the prefix
    it walks (candidates picked by the ratio of the two pair sums).  Its
    result equals OSD given every exact |LLR|, on marginals built to make
    that hard: few distinct levels (large groups of exact ties, broken by
    index); a race between the wrong qubits and a dependent column set (the
    qubits of one check of the other side: H_X H_Z^T = 0) whose outcome
    depends on the order the rounded logs give, so that the ratio order
    alone gets it wrong; zero pair sums (clamped); and unnormalised values
    beyond the ratio bound's range."""
    rng = np.random.default_rng(11)
    n = code.n
    applied_seen = set()
    race_differs = 0
    for case in range(48):
        side = case % 2
        kind = (case // 2) % 6
        H = code.checks(side)
        Hother = code.checks(1 - side)
        bit = 1 if side == 0 else 2
        err = np.zeros(n, np.int32)
        idx = rng.choice(n, 12, replace=False)
        err[idx] = rng.integers(1, 4, 12)
        syn = np.array([np.bitwise_xor.reduce((err[r[r >= 0]] & bit) != 0) for r in H], np.int32)
        est = err.copy()
        wrong = rng.choice(n, int(rng.integers(1, 5)), replace=False)
        if kind == 0:      # few levels: ties everywhere
            tot = rng.integers(1, 5, (n, 4)).astype(np.float64) / 8.0
        elif kind <= 3:
            # race at the front of the order: the dependent set and two wrong
            # qubits share a ratio; the rounded logs put one group first,
            # index order the other
            tot = np.tile([0.4, 0.1, 0.4, 0.1], (n, 1)) * (1 + rng.random((n, 1)))
            if side == 1:
                tot = tot[:, [0, 2, 1, 3]]
            tA, tB = _tie_pair(side)
            row = Hother[int(rng.integers(len(Hother)))]
            dep = np.sort(row[row >= 0])
            rest = np.setdiff1d(np.arange(n), dep)
            if kind % 2:   # wrong first by value, last by index: repaired
                wrong = rng.choice(rest[rest > dep[-1]], 2, replace=False)
                tot[wrong], tot[dep] = tA, tB
            else:          # dependent set first by value, last by index: not
                wrong = rng.choice(rest[rest < dep[0]], 2, replace=False)
                tot[wrong], tot[dep] = tB, tA
            if kind == 3:  # with a crowd of equal values ahead of both
                crowd = rng.choice(np.setdiff1d(rest, wrong), 400, replace=False)
                tot[crowd] = [0.25, 0.25, 0.25, 0.25]
        elif kind == 4:    # zero pair sums, clamped to 1e-300
            tot = rng.random((n, 4))
            tot /= tot.sum(1, keepdims=True)
            z = rng.choice(n, 300, replace=False)
            tot[z[:150], 1] = tot[z[:150], 3] = 0.0
            tot[z[150:], 0] = tot[z[150:], 2] = 0.0
            tot[wrong] = 0.25
        else:              # beyond the bound's range: every value computed
            tot = rng.random((n, 4)) * 1e305
            tot[wrong] = 1.0
        est[wrong] ^= bit
        tot = np.ascontiguousarray(tot)
        llr = _abs_llr_libm(tot, side)
        want = jbp.pp_repair(code, jbp.PP_OSD, side, syn, est, abs_llr=llr)
        got = jbp.pp_repair(code, jbp.PP_OSD_MARGINALS, side, syn, est, abs_llr=tot.ravel())
        assert got[1] == want[1], "case %d" % case
        np.testing.assert_array_equal(got[0], want[0], err_msg="case %d" % case)
        applied_seen.add(want[1])
        if 1 <= kind <= 3:
            # what ordering by the ratio alone would have given
            p1 = np.maximum(tot[:, 1] + tot[:, 3] if side == 0 else tot[:, 2] + tot[:, 3], 1e-300)
            p0 = np.maximum(tot[:, 0] + tot[:, 2] if side == 0 else tot[:, 0] + tot[:, 1], 1e-300)
            ratio = np.maximum(p1, p0) / np.minimum(p1, p0)
            rank = np.empty(n, np.int64)
            rank[np.lexsort((np.arange(n), ratio))] = np.arange(n)
            alt = jbp.pp_repair(code, jbp.PP_OSD, side, syn, est, abs_llr=rank.astype(np.float64))
            race_differs += alt[1] != want[1]
    assert applied_seen == {0, 1}
    assert race_differs > 0

"""CPU-only tests of libjbp's host side: the C ABI loads and exports every
declared symbol, code construction matches the reference's, diagnostics
reproduce the README's code summary, and errors are reported (no compute
calls: there is no device here)."""
import ctypes as C
import hashlib
import os

import numpy as np
import pytest

import joint_bp_plus_pp_amd as jbp
from conftest import PARAMS, golden_json


def test_library_exports_every_declared_symbol():
    L = jbp.lib()
    names = jbp.header_symbols()
    assert "jbp_create" in names and "jbp_simulate" in names and "jbp_code_hgp" in names
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_params_code_matches_reference_construction(main_code_oracle):
    code = jbp.Code.from_params(PARAMS)
    summary = golden_json("code_summary.json")
    assert (code.n, code.mx, code.mz, code.x_edges, code.z_edges) == (9216, 2304, 2304, 27648, 27648)
    assert (code.P, code.J, code.L) == (768, 3, 12)
    for side, key in ((0, "checks_x"), (1, "checks_z")):
        ch = code.checks(side)
        assert ch.shape == (2304, 12)
        assert hashlib.sha256(ch.astype(np.int32).tobytes()).hexdigest() == summary[key + "_sha256"]
        off, ev = code.csr(side)
        ooff, oev = main_code_oracle.csr(side)
        assert np.array_equal(off, ooff) and np.array_equal(ev, oev)
    # fi / gi echo (README.md:118-119)
    assert code.params()[0].tolist() == [763, 679, 397, 61, 697, 373]


def test_var_adjacency_ascending_check_order():
    code = jbp.Code.from_params(PARAMS)
    g = code.graph()
    for side in (0, 1):
        voff = np.ctypeslib.as_array(g.x_var_offsets if side == 0 else g.z_var_offsets, shape=(code.n + 1,))
        vedge = np.ctypeslib.as_array(g.x_var_edges if side == 0 else g.z_var_edges, shape=(27648,))
        echk = np.ctypeslib.as_array(g.x_edge_check if side == 0 else g.z_edge_check, shape=(27648,))
        evar = np.ctypeslib.as_array(g.x_edge_var if side == 0 else g.z_edge_var, shape=(27648,))
        assert np.all(np.diff(voff) == 3)
        for v in range(0, code.n, 97):
            es = vedge[voff[v]:voff[v + 1]]
            assert np.all(evar[es] == v)
            assert np.all(np.diff(echk[es]) > 0)


def test_code_summary_known_answers():
    # README.md:124-132: rank 2302/2302, k = 4612, degrees 3/12, girth 8/8,
    # orthogonal with no odd pairs.
    d = jbp.Code.from_params(PARAMS).diagnostics(girth=True)
    summary = golden_json("code_summary.json")
    assert (d["rank_x"], d["rank_z"]) == (2302, 2302)
    assert 9216 - d["rank_x"] - d["rank_z"] == 4612
    assert d["odd_pairs"] == 0 == summary["odd_pairs"]
    assert (d["girth_x"], d["girth_z"]) == (8, 8) == (summary["girth_x"], summary["girth_z"])
    for s in "xz":
        assert d[s + "_var_deg_min"] == d[s + "_var_deg_max"] == 3
        assert d[s + "_check_deg_min"] == d[s + "_check_deg_max"] == 12


def test_girth_of_small_cycle_codes():
    # a 6-cycle: checks {0,1}, {1,2}, {2,0} -> Tanner girth 6; a repeated pair -> 4
    tri = np.array([[0, 1], [1, 2], [2, 0]], np.int32)
    d = jbp.Code.from_checks(3, tri, tri).diagnostics()
    assert d["girth_x"] == 6
    four = np.array([[0, 1], [0, 1]], np.int32)
    assert jbp.Code.from_checks(2, four, four).diagnostics()["girth_x"] == 4


def test_hgp_code_structure():
    code = jbp.Code.hgp(150, 200, 3, 4, seed=1)
    assert (code.n, code.mx, code.mz) == (62500, 30000, 30000)
    assert code.x_edges == code.z_edges == 210000
    d = code.diagnostics(girth=False)
    assert d["odd_pairs"] == 0  # Hx Hz^T = 0
    assert d["x_check_deg_min"] == d["x_check_deg_max"] == 7
    assert (d["x_var_deg_min"], d["x_var_deg_max"]) == (3, 4)
    # seeded: the same seed gives the same code, another seed another one
    again = jbp.Code.hgp(150, 200, 3, 4, seed=1)
    other = jbp.Code.hgp(150, 200, 3, 4, seed=2)
    assert np.array_equal(code.checks(0), again.checks(0))
    assert not np.array_equal(code.checks(0), other.checks(0))


def test_hgp_small_dense_orthogonality():
    code = jbp.Code.hgp(6, 8, 3, 4, seed=3)
    n = code.n
    H = [np.zeros((code.mx, n), np.int64), np.zeros((code.mz, n), np.int64)]
    for s in (0, 1):
        for c, row in enumerate(code.checks(s)):
            H[s][c, row] = 1
    assert not np.any((H[0] @ H[1].T) & 1)
    assert n == 8 * 8 + 6 * 6


def test_bad_params_file_reports_error(tmp_path):
    p = tmp_path / "bad.txt"
    p.write_text("768 3 11\n1\n")
    with pytest.raises(jbp.JbpError) as ei:
        jbp.Code.from_params(str(p))
    assert ei.value.code == 1
    with pytest.raises(jbp.JbpError):
        jbp.Code.from_params(str(tmp_path / "missing.txt"))


def test_create_without_device_fails_cleanly(gpu_available):
    if gpu_available:
        pytest.skip("a device is present")
    code = jbp.Code.from_params(PARAMS)
    with pytest.raises(jbp.JbpError) as ei:
        jbp.Decoder(code, 0)
    assert ei.value.code == 2


def test_fault_injection_create(monkeypatch):
    monkeypatch.setenv("JBP_FAULT_INJECT", "create")
    code = jbp.Code.from_params(PARAMS)
    with pytest.raises(jbp.JbpError) as ei:
        jbp.Decoder(code, 0)
    assert "fault injected" in str(ei.value)


def test_destroy_null_is_noop():
    jbp.lib().jbp_destroy(None)


def test_pack_bits_roundtrip():
    rng = np.random.default_rng(0)
    a = rng.integers(0, 2, (3, 77))
    b = jbp.pack_bits(a)
    assert b.shape == (3, 3)
    back = ((b[:, :, None] >> np.arange(32, dtype=np.uint32)) & 1).reshape(3, -1)[:, :77]
    assert np.array_equal(back, a)


def test_simulate_mixed_rejects_bad_arguments():
    # checked before any device call (1: JBP_ERR_ARG)
    L = jbp.lib()
    e = C.create_string_buffer(512)
    cfgs = (jbp.SimCfg * 3)()
    ctxs = (C.c_void_p * 3)()
    stats = (jbp.RunStats * 3)()
    assert L.jbp_simulate_mixed(ctxs, cfgs, 0, None, stats, e, 512) == 1
    assert L.jbp_simulate_mixed(ctxs, cfgs, 3, None, stats, e, 512) == 1
    assert b"codes" in e.value
    assert L.jbp_simulate_mixed(ctxs, cfgs, 1, None, stats, e, 512) == 1
    assert b"code 0" in e.value

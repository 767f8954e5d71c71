"""The jointbp_ets command line: the reference's flags, summary lines and
exit codes (jointbp_ets.cpp:4822-6470), driven as a subprocess."""
import os
import re
import subprocess

import numpy as np
import pytest

from conftest import PARAMS, ROOT, fnv1a, golden_json

CLI = os.path.join(ROOT, "joint_bp_plus_pp_amd", "jointbp_ets")


def run(args, **kw):
    if not os.path.exists(CLI):
        pytest.skip("jointbp_ets not built")
    return subprocess.run([CLI, "--params", PARAMS] + args, capture_output=True, text=True,
                          timeout=300, env=dict(os.environ, NO_COLOR="1"), **kw)


def summary_fields(text):
    line = [ln for ln in text.splitlines() if ln.startswith("trials=")][-1]
    return dict(re.findall(r"(\w+)=([^\s]+)", line))


def test_code_summary_lines_match_reference():
    out = run(["--simulate", "--trials", "1", "--seed", "1", "--max-iter", "3"])
    ref = golden_json("aggregates.json")["cli_simulate"]["lines"]
    for line in ref[:4]:
        assert line in out.stdout, line


def test_ets_load_table_matches_reference(tmp_path):
    """The trapping-set tables are found and reported as the reference
    reports them (jointbp_ets.cpp:5336-5387), also when the program runs
    elsewhere (the tables next to the params file), and as missing when
    there are none."""
    want = golden_json("pp_cli.json")["ets_load"]
    for cwd in (ROOT, str(tmp_path)):
        out = run(["--simulate", "--trials", "1", "--max-iter", "3", "--no-pp"], cwd=cwd)
        lines = out.stdout.splitlines()
        assert [ln for ln in lines if ln.startswith("ETS")] == want
    lonely = tmp_path / "lonely.txt"
    lonely.write_text(open(PARAMS).read())
    out = subprocess.run([CLI, "--params", str(lonely), "--simulate", "--trials", "1", "--no-pp"],
                         capture_output=True, text=True, timeout=300)
    rows = [ln for ln in out.stdout.splitlines() if ln.startswith("ETS")]
    assert len(rows) == 10 and all(r.endswith("X=- Z=- status=missing") for r in rows)


def test_argument_errors():
    assert run(["--simulate", "--no-pp", "--err", "x.txt"]).returncode == 1
    assert run(["--no-pp", "--trials", "3"]).returncode == 1
    assert run(["--simulate", "--no-pp", "--trial-index", "4", "--trials", "2"]).returncode == 1
    assert run(["--bogus"]).returncode == 1
    # --freeze-syn runs on the FP64 decoder with post-processing off; other
    # modes refuse it (never a silent drop)
    for mode in (["--simulate", "--no-pp", "--trials", "2"], ["--simulate", "--trials", "2"],
                 ["--simulate", "--trials", "2", "--precision", "64"]):
        out = run(mode + ["--freeze-syn"])
        assert out.returncode == 1 and "--freeze-syn" in out.stderr, mode
    bad = subprocess.run([CLI, "--params", "/nonexistent"], capture_output=True, text=True)
    assert bad.returncode == 1 and "Failed to read params" in bad.stderr


@pytest.mark.gpu
def test_simulate_fp64_matches_reference_run_summary():
    ref = golden_json("aggregates.json")["cli_simulate"]
    want = summary_fields("\n".join(ref["lines"]))
    out = run(["--simulate", "--p", "0.05", "--trials", "200", "--max-iter", "50", "--no-pp",
               "--seed", "106", "--report-every", "0", "--cuda", "--precision", "64"])
    assert out.returncode == 0, out.stderr
    got = summary_fields(out.stdout)
    for key in ("trials", "failures", "bp_fail", "stab_success", "fer", "ci95", "iters",
                "avg_iter", "exact_rate"):
        assert got[key] == want[key], key


@pytest.mark.gpu
def test_err_decode_matches_reference(tmp_path):
    from _oracle import Oracle
    ref = golden_json("aggregates.json")["cli_err_trial7"]
    err = Oracle().code_from_params(PARAMS).sample(106, 7, 0.05)
    assert fnv1a(err) == ref["err_hash"]
    ep, es = tmp_path / "err.txt", tmp_path / "est.txt"
    np.savetxt(ep, err, fmt="%d")
    out = run(["--err", str(ep), "--p", "0.05", "--max-iter", "50", "--no-pp", "--cuda",
               "--precision", "64", "--est", str(es)])
    assert out.returncode == 0, out.stderr
    for line in ref["lines"][4:]:
        assert line in out.stdout, line
    assert fnv1a(np.loadtxt(es, dtype=np.int64)) == ref["est_hash"]


@pytest.mark.gpu
def test_min_fails_and_hist(tmp_path):
    h = tmp_path / "hist.txt"
    out = run(["--simulate", "--p", "0.05", "--trials", "256", "--max-iter", "50", "--no-pp",
               "--seed", "106", "--report-every", "0", "--min-fails", "2", "--precision", "64",
               "--iter-hist-out", str(h)])
    assert out.returncode == 0, out.stderr
    got = summary_fields(out.stdout)
    assert got["failures"] == "2"
    lines = h.read_text().splitlines()
    assert lines[0].startswith("# costs") and lines[1] == "# iter fraction"
    assert abs(sum(float(x.split()[1]) for x in lines[2:]) - 1.0) < 1e-6


@pytest.mark.gpu
@pytest.mark.parametrize("graph", [False, True])
def test_syndrome_file_decode_matches_reference(tmp_path, graph):
    # --sx/--sz (one int per line, jointbp_ets.cpp:3081-3090) of the reference's
    # trial 7: the FP64 estimate must equal the reference's estimate of that
    # trial, with or without the captured graph (--cuda-graph)
    from _oracle import Oracle
    ref = golden_json("aggregates.json")["cli_err_trial7"]
    oc = Oracle().code_from_params(PARAMS)
    err = oc.sample(106, 7, 0.05)
    sx, sz = oc.syndrome(err)
    px, pz, es = tmp_path / "sx.txt", tmp_path / "sz.txt", tmp_path / "est.txt"
    np.savetxt(px, sx, fmt="%d")
    np.savetxt(pz, sz, fmt="%d")
    args = ["--sx", str(px), "--sz", str(pz), "--p", "0.05", "--max-iter", "50", "--no-pp",
            "--cuda", "--precision", "64", "--est", str(es)]
    out = run(args + (["--cuda-graph"] if graph else []))
    assert out.returncode == 0, out.stderr
    assert fnv1a(np.loadtxt(es, dtype=np.int64)) == ref["est_hash"]


@pytest.mark.gpu
def test_simulate_with_pp_matches_reference_run_summary():
    ref = golden_json("pp_cli.json")
    want = summary_fields(ref["run_summary"])
    out = run(ref["simulate_args"] + ["--report-every", "0", "--precision", "64"])
    assert out.returncode == 0, out.stderr
    got = summary_fields(out.stdout)
    for key in ("trials", "failures", "bp_fail", "pp_success", "flip_pp_success", "stab_success",
                "ets_used", "ets_saved", "ets6_used", "ets6_saved", "ets12_used", "ets12_saved",
                "fer", "ci95", "iters", "avg_iter", "exact_rate"):
        assert got[key] == want[key], key


@pytest.mark.gpu
def test_err_decode_with_pp_matches_reference(tmp_path):
    import hashlib
    from _oracle import Oracle
    ref = golden_json("pp_cli.json")
    err = Oracle().code_from_params(PARAMS).sample(ref["err_seed"], ref["err_trial"], 0.04)
    ep, es = tmp_path / "err.txt", tmp_path / "est.txt"
    np.savetxt(ep, err, fmt="%d")
    out = run(["--err", str(ep)] + ref["err_args"] + ["--precision", "64", "--est", str(es)])
    assert out.returncode == ref["err_exit"], out.stderr
    for line in ref["decode_result"]:
        assert line in out.stdout.splitlines(), line
    est = np.loadtxt(es, dtype=np.int64)
    assert hashlib.sha1(est.astype(np.uint8).tobytes()).hexdigest() == ref["est_sha1"]


REF_HOST = os.path.join(ROOT, "oracle", "_ref", "jointbp_ets_jbp")


def run_ref_host(tmp_path, extra):
    """The reference's own jointbp_ets.cpp, built with -DUSE_CUDA against
    joint_bp_plus_pp_amd/refhost/jointbp_cuda_jbp.cpp (its cuda_bp_* calls,
    jointbp_cuda.h:48-63, answered by libjbp), decoding the reference's
    trial 7 from an --err file."""
    if not os.path.exists(REF_HOST):
        pytest.skip("reference host on libjbp not built (needs the reference source)")
    from _oracle import Oracle
    err = Oracle().code_from_params(PARAMS).sample(106, 7, 0.05)
    ep, es = tmp_path / "err.txt", tmp_path / "est.txt"
    np.savetxt(ep, err, fmt="%d")
    out = subprocess.run([REF_HOST, "--params", PARAMS, "--err", str(ep), "--p", "0.05", "--max-iter",
                          "50", "--no-pp", "--est", str(es)] + extra, capture_output=True, text=True,
                         timeout=300, env=dict(os.environ, NO_COLOR="1"))
    return out, es


def test_reference_host_without_device_falls_back_to_its_cpu_decoder(tmp_path):
    # no device: cuda_bp_create returns null with libjbp's message, the
    # reference prints it and decodes on its own CPU decoder
    from conftest import hip_device_available
    if hip_device_available():
        pytest.skip("a device is present")
    ref = golden_json("aggregates.json")["cli_err_trial7"]
    out, es = run_ref_host(tmp_path, ["--cuda"])
    assert out.returncode == 0, out.stderr
    assert "CUDA init failed" in out.stderr and "falling back to CPU" in out.stderr
    for line in ref["lines"][4:]:
        assert line in out.stdout, line
    assert fnv1a(np.loadtxt(es, dtype=np.int64)) == ref["est_hash"]


@pytest.mark.gpu
def test_reference_host_on_libjbp_matches_reference(tmp_path):
    # the reference program's --err decode through its CUDA path, run by
    # libjbp's FP64 decoder: the same Decode Result lines and estimate as the
    # reference's CPU run, and no fallback
    ref = golden_json("aggregates.json")["cli_err_trial7"]
    out, es = run_ref_host(tmp_path, ["--cuda"])
    assert out.returncode == 0, out.stderr
    assert "falling back" not in out.stderr, out.stderr
    for line in ref["lines"][4:]:
        assert line in out.stdout, line
    assert fnv1a(np.loadtxt(es, dtype=np.int64)) == ref["est_hash"]
    # --freeze-syn goes to the device too (FP64 runs it)
    out, _ = run_ref_host(tmp_path, ["--cuda", "--freeze-syn"])
    assert out.returncode == 0 and "falling back" not in out.stderr, out.stderr

"""Write the golden fixtures under tests/golden/ from the reference decoder.

Runs only where the reference sources are present: it loads
oracle/_ref/libref_harness.so (built by `make -C oracle ref` from the
reference's jointbp_ets.cpp) and oracle/_ref/jointbp_ets (the reference
program itself).  The fixtures are data only: per-trial outputs of the
reference for seeded inputs, plus its code-summary known answers.

    python tests/golden/make_golden.py
"""
import ctypes as C
import hashlib
import json
import os
import re
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF_DIR = "/root/reference"
PARAMS = os.path.join(REF_DIR, "H_P768_J3_L12_dmax3_nc0-3_1-2_seed11579811919164041.txt")
HARNESS = os.path.join(ROOT, "oracle", "_ref", "libref_harness.so")
REF_BIN = os.path.join(ROOT, "oracle", "_ref", "jointbp_ets")


def fnv1a(arr):
    h = 1469598103934665603
    for v in np.asarray(arr, dtype=np.int64).tolist():
        h ^= v & 0xFFFFFFFFFFFFFFFF
        h = (h * 1099511628211) & 0xFFFFFFFFFFFFFFFF
    return h


def pack2(err):
    """Pack qubit states 0..3 four to a byte, qubit i in bits 2*(i%4)."""
    e = np.asarray(err, dtype=np.uint8).reshape(-1, 4)
    return (e[:, 0] | (e[:, 1] << 2) | (e[:, 2] << 4) | (e[:, 3] << 6)).astype(np.uint8)


def load():
    lib = C.CDLL(HARNESS)
    lib.ref_load_params.argtypes = [C.c_char_p]
    lib.ref_code_info.argtypes = [C.c_void_p, C.c_int, C.c_void_p]
    lib.ref_get_checks.argtypes = [C.c_int, C.c_void_p]
    lib.ref_trial_seed.restype = C.c_uint64
    lib.ref_trial_seed.argtypes = [C.c_uint64, C.c_uint64]
    lib.ref_sample.argtypes = [C.c_uint64, C.c_longlong, C.c_double, C.c_void_p]
    lib.ref_syndrome.argtypes = [C.c_void_p] * 3
    lib.ref_decode.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_double, C.c_int,
                               C.c_void_p, C.c_void_p]
    lib.ref_prior.argtypes = [C.c_double, C.c_void_p]
    lib.ref_stab_equiv.argtypes = [C.c_void_p, C.c_void_p]
    lib.ref_simulate.restype = C.c_double
    lib.ref_simulate.argtypes = [C.c_uint64, C.c_longlong, C.c_longlong, C.c_double, C.c_int,
                                 C.c_double, C.c_int] + [C.c_void_p] * 4
    assert lib.ref_load_params(PARAMS.encode()) == 0
    return lib


def code_summary(lib):
    info = np.zeros(9, np.int64)
    girth = np.zeros(2, np.int32)
    lib.ref_code_info(info.ctypes.data, 1, girth.ctypes.data)
    out = dict(zip(["n", "mx", "mz", "rank_x", "rank_z", "edges_x", "edges_z", "orthogonal",
                    "odd_pairs"], [int(v) for v in info]))
    out["girth_x"], out["girth_z"] = int(girth[0]), int(girth[1])
    for side, name in ((0, "x"), (1, "z")):
        flat = np.zeros(out["mx"] * 12, np.int32)
        lib.ref_get_checks(side, flat.ctypes.data)
        out["checks_%s_sha256" % name] = hashlib.sha256(flat.tobytes()).hexdigest()
        out["checks_%s_first_rows" % name] = flat[:36].tolist()
    return out


def per_trial(lib, base, first, count, p, max_iter, damping=0.0, keep_vectors=0):
    """Per-trial reference outputs.  The decode runs through ref_decode (one
    trial at a time) so that the estimate itself can be hashed."""
    n = 9216
    err = np.zeros(n, np.int32)
    sx = np.zeros(2304, np.int32)
    sz = np.zeros(2304, np.int32)
    est = np.zeros(n, np.int32)
    prior = np.zeros(4, np.float64)
    lib.ref_prior(p, prior.ctypes.data)
    rec = {k: [] for k in ("seed", "err_hash", "err_weight", "sx_hash", "sz_hash", "iterations",
                           "syn_ok", "success", "exact", "est_hash")}
    vecs = {"err": [], "est": []}
    for t in range(first, first + count):
        lib.ref_sample(base, t, p, err.ctypes.data)
        lib.ref_syndrome(err.ctypes.data, sx.ctypes.data, sz.ctypes.data)
        ok = C.c_int(0)
        it = lib.ref_decode(sx.ctypes.data, sz.ctypes.data, prior.ctypes.data, max_iter, damping, 0,
                            est.ctypes.data, C.byref(ok))
        succ = bool(ok.value) and bool(lib.ref_stab_equiv(err.ctypes.data, est.ctypes.data))
        rec["seed"].append(int(lib.ref_trial_seed(base, t)))
        rec["err_hash"].append(fnv1a(err))
        rec["err_weight"].append(int(np.count_nonzero(err)))
        rec["sx_hash"].append(fnv1a(sx))
        rec["sz_hash"].append(fnv1a(sz))
        rec["iterations"].append(int(it))
        rec["syn_ok"].append(int(ok.value))
        rec["success"].append(int(succ))
        rec["exact"].append(int(np.array_equal(est, err)))
        rec["est_hash"].append(fnv1a(est))
        if t - first < keep_vectors:
            vecs["err"].append(pack2(err))
            vecs["est"].append(pack2(est))
    arrays = {k: np.array(v, dtype=np.uint64 if k.endswith("hash") or k == "seed" else np.int32)
              for k, v in rec.items()}
    for k, v in vecs.items():
        if v:
            arrays[k + "_packed"] = np.stack(v)
    arrays["meta"] = np.array([base, first, count, max_iter], dtype=np.int64)
    arrays["p"] = np.array([p, damping], dtype=np.float64)
    arrays["prior"] = prior.copy()
    return arrays


def aggregate(lib, base, first, count, p, max_iter, threads=8):
    it = np.zeros(count, np.int32)
    ok = np.zeros(count, np.uint8)
    su = np.zeros(count, np.uint8)
    ex = np.zeros(count, np.uint8)
    lib.ref_simulate(base, first, count, p, max_iter, 0.0, threads, it.ctypes.data, ok.ctypes.data,
                     su.ctypes.data, ex.ctypes.data)
    return dict(base=base, first=first, count=count, p=p, max_iter=max_iter,
                failures=int(count - su.sum()), bp_fail=int(count - ok.sum()),
                iters=int(it.sum()), exact=int(ex.sum()),
                iter_hist={int(k): int(v) for k, v in zip(*np.unique(it, return_counts=True))})


def ref_cli_summary(args):
    """Run the reference program and parse its Run Summary / Decode Result."""
    out = subprocess.run([REF_BIN, "--params", PARAMS] + args, capture_output=True, text=True,
                         env=dict(os.environ, NO_COLOR="1"))
    text = out.stdout
    keep = [ln for ln in text.splitlines()
            if ln.startswith(("trials=", "iterations=", "decode_success=", "exact_match=",
                              "stabilizer_equiv=", "n=", "P=", "X: ", "Z: ", "hx_hz_"))]
    return {"args": args, "returncode": out.returncode, "lines": keep}


def xonly_decodes(lib, base, count, p, max_iter, damping):
    """Decodes of X-only errors (each sampled error with its Z components
    dropped: Y -> X, Z -> I) with --freeze-syn on and off.  The Z syndrome is
    then zero and matches at the first iteration, so the Z side freezes while
    the X side still decodes: the branch --freeze-syn changes."""
    n = 9216
    err = np.zeros(n, np.int32)
    sx = np.zeros(2304, np.int32)
    sz = np.zeros(2304, np.int32)
    est = np.zeros(n, np.int32)
    prior = np.zeros(4, np.float64)
    lib.ref_prior(p, prior.ctypes.data)
    rec = {k: [] for k in ("err_hash", "iterations", "syn_ok", "est_hash", "iterations_nofreeze",
                           "est_hash_nofreeze")}
    for t in range(1, count + 1):
        lib.ref_sample(base, t, p, err.ctypes.data)
        err &= 1
        lib.ref_syndrome(err.ctypes.data, sx.ctypes.data, sz.ctypes.data)
        for freeze, sfx in ((1, ""), (0, "_nofreeze")):
            ok = C.c_int(0)
            it = lib.ref_decode(sx.ctypes.data, sz.ctypes.data, prior.ctypes.data, max_iter, damping,
                                freeze, est.ctypes.data, C.byref(ok))
            rec["iterations" + sfx].append(int(it))
            rec["est_hash" + sfx].append(fnv1a(est))
            if freeze:
                rec["syn_ok"].append(int(ok.value))
        rec["err_hash"].append(fnv1a(err))
    arrays = {k: np.array(v, dtype=np.uint64 if "hash" in k else np.int32) for k, v in rec.items()}
    arrays["meta"] = np.array([base, 1, count, max_iter], dtype=np.int64)
    arrays["p"] = np.array([p, damping], dtype=np.float64)
    arrays["prior"] = prior.copy()
    return arrays


def freeze_fixtures(lib):
    """--freeze-syn (jointbp_ets.cpp:1884-1901, 2526-2535) on X-only errors."""
    np.savez_compressed(os.path.join(HERE, "decodes_xonly_p0.04_it60_freeze.npz"),
                        **xonly_decodes(lib, 106, 48, 0.04, 60, 0.0))


def main():
    if not os.path.exists(HARNESS):
        sys.exit("build the reference checker first: make -C oracle ref")
    lib = load()
    if sys.argv[1:] == ["freeze"]:
        freeze_fixtures(lib)
        print("freeze fixtures written to", HERE)
        return
    with open(os.path.join(HERE, "code_summary.json"), "w") as f:
        json.dump(code_summary(lib), f, indent=1)
    np.savez_compressed(os.path.join(HERE, "trials_p0.05_it50.npz"),
                        **per_trial(lib, 106, 1, 256, 0.05, 50, keep_vectors=8))
    np.savez_compressed(os.path.join(HERE, "trials_p0.04_it1000.npz"),
                        **per_trial(lib, 106, 1, 256, 0.04, 1000, keep_vectors=4))
    np.savez_compressed(os.path.join(HERE, "trials_p0.025_it200.npz"),
                        **per_trial(lib, 3, 1, 64, 0.025, 200))
    np.savez_compressed(os.path.join(HERE, "trials_p0.05_it50_damp0.3.npz"),
                        **per_trial(lib, 106, 1, 64, 0.05, 50, damping=0.3))
    freeze_fixtures(lib)
    agg = {"cfg1_p0.05_it50_seed106_1000": aggregate(lib, 106, 1, 1000, 0.05, 50),
           "p0.04_it1000_seed106_2000": aggregate(lib, 106, 1, 2000, 0.04, 1000)}
    # The reference program's own Run Summary on the first workload, and a
    # single --err decode with its printed result lines.
    agg["cli_simulate"] = ref_cli_summary(["--simulate", "--p", "0.05", "--trials", "200",
                                           "--max-iter", "50", "--no-pp", "--seed", "106",
                                           "--report-every", "0"])
    with tempfile.TemporaryDirectory() as td:
        err = np.zeros(9216, np.int32)
        lib.ref_sample(106, 7, 0.05, err.ctypes.data)
        ep = os.path.join(td, "err.txt")
        np.savetxt(ep, err, fmt="%d")
        agg["cli_err_trial7"] = ref_cli_summary(["--err", ep, "--p", "0.05", "--max-iter", "50",
                                                 "--no-pp", "--est", os.path.join(td, "est.txt")])
        est = np.loadtxt(os.path.join(td, "est.txt"), dtype=np.int64)
        agg["cli_err_trial7"]["est_hash"] = fnv1a(est)
        agg["cli_err_trial7"]["err_hash"] = fnv1a(err)
    with open(os.path.join(HERE, "aggregates.json"), "w") as f:
        json.dump(agg, f, indent=1)
    print("fixtures written to", HERE)


if __name__ == "__main__":
    main()

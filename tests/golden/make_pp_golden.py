"""Write the post-processing fixtures under tests/golden/ from the reference.

Runs only where the reference sources are present: it loads
oracle/_ref/libref_harness.so (built by `make -C oracle ref` from the
reference's jointbp_ets.cpp), whose entry points call the reference's own
read_ets_file, attempt_ets_fix_side_multi, attempt_flip_pp_side,
attempt_osd_pp_side and joint_bp_decode with PP on.  The fixtures are data:

  pp_ets_counts.json   entries per table and side, as the reference loads them
  pp_repairs.npz       single repairs (ETS, flip set, OSD) on constructed
                       estimates: inputs and the reference's outputs
  pp_trials_*.npz      per-trial outcome of the simulate loop with PP on

    python tests/golden/make_pp_golden.py
"""
import ctypes as C
import hashlib
import json
import os
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF_DIR = "/root/reference"
STEM = "H_P768_J3_L12_dmax3_nc0-3_1-2_seed11579811919164041"
PARAMS = os.path.join(REF_DIR, STEM + ".txt")
ETS_DIR = os.path.join(REF_DIR, STEM)
HARNESS = os.path.join(ROOT, "oracle", "_ref", "libref_harness.so")
REF_BIN = os.path.join(ROOT, "oracle", "_ref", "jointbp_ets")
P = 768
# (size, label, file) in the reference's load order (jointbp_ets.cpp:5309-5327)
GROUPS = [(6, "6", "ets_6_2.txt"), (8, "8", "ets_8_2.txt"), (8, "8p4", "ets_8_2_path4.txt"),
          (10, "10", "ets_10_2.txt"), (12, "12", "ets_12_2.txt"), (14, "14", "ets_14_2.txt"),
          (16, "16", "ets_16_2.txt"), (18, "18", "ets_18_2.txt"), (20, "20", "ets_20_2.txt"),
          (22, "22", "ets_22_2.txt")]
# per-trial workloads: (name, p, max_iter, trials, seed, flip_hist, damping)
WORKLOADS = [("p0.04_it8", 0.04, 8, 256, 7, 5, 0.0),
             ("p0.03_it5", 0.03, 5, 256, 7, 5, 0.0),
             ("p0.04_it8_hist0", 0.04, 8, 128, 21, 0, 0.0),
             ("p0.02_it1", 0.02, 1, 128, 5, 5, 0.0),
             ("p0.05_it115", 0.05, 115, 512, 11, 5, 0.0),
             ("p0.04_it10_damp0.3", 0.04, 10, 128, 13, 5, 0.3)]
# small hypergraph-product codes (built by libjbp's seeded generator), no
# trapping-set tables: name, (m, n, col_w, row_w, seed), p, max_iter, trials,
# seed.  HGP(21x28) has 588 checks per side: 19 words per syndrome, an odd
# count (the stage then copies syndromes trial by trial).
HGPS = [("hgp15x20", (15, 20, 3, 4, 5), 0.05, 6, 256, 3),
        ("hgp21x28", (21, 28, 3, 4, 2), 0.04, 9, 256, 4)]


def pack2(states):
    e = np.asarray(states, dtype=np.uint8).reshape(-1, 4)
    return (e[:, 0] | (e[:, 1] << 2) | (e[:, 2] << 4) | (e[:, 3] << 6)).astype(np.uint8)


def est_digest(est):
    return hashlib.sha1(np.asarray(est, dtype=np.uint8).tobytes()).hexdigest()


def load():
    h = C.CDLL(HARNESS)
    h.ref_load_params.argtypes = [C.c_char_p]
    h.ref_get_checks.argtypes = [C.c_int, C.c_void_p]
    h.ref_code_info.argtypes = [C.c_void_p, C.c_int, C.c_void_p]
    h.ref_sample.argtypes = [C.c_uint64, C.c_longlong, C.c_double, C.c_void_p]
    h.ref_syndrome.argtypes = [C.c_void_p] * 3
    h.ref_ets_add.argtypes = [C.c_int, C.c_char_p, C.c_char_p, C.c_char_p, C.c_int, C.c_void_p]
    h.ref_pp_repair.argtypes = [C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int,
                                C.c_void_p]
    h.ref_decode_llr.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_double,
                                 C.c_void_p, C.c_void_p, C.c_void_p]
    h.ref_prior.argtypes = [C.c_double, C.c_void_p]
    h.ref_simulate_pp.restype = C.c_double
    h.ref_simulate_pp.argtypes = [C.c_uint64, C.c_longlong, C.c_longlong, C.c_double, C.c_int,
                                  C.c_double, C.c_int, C.c_int] + [C.c_void_p] * 7
    assert h.ref_load_params(PARAMS.encode()) == 0
    return h


def load_tables(h):
    counts = {}
    c = np.zeros(2, np.int32)
    for size, label, name in GROUPS:
        path = os.path.join(ETS_DIR, name).encode()
        h.ref_ets_add(size, label.encode(), path, path, P, c.ctypes.data)
        counts[label] = [int(c[0]), int(c[1])]
    return counts


def read_entries(name, side):
    """(vars, unsat) pairs of one table for one side, parsed from the text
    (only to construct inputs; the outputs come from the reference)."""
    out = []
    for line in open(os.path.join(ETS_DIR, name)):
        t = line.split()
        if not t or t[0].upper() != side:
            continue
        rest = line.strip()[1:]
        v = rest.split("vars=")[1].split("unsat=")[0].strip().split(",")
        vars_ = [int(a[1:].split(":")[0]) * P + int(a.split(":")[1]) for a in v]
        out.append(vars_)
    return out


def hgp_fixtures(h):
    """Per-trial PP outcomes of the reference on the small HGP codes."""
    import sys
    sys.path.insert(0, ROOT)
    import joint_bp_plus_pp_amd as jbp
    h.ref_set_checks.argtypes = [C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_int, C.c_int, C.c_void_p]
    for name, (hm, hn, cw, rw, hseed), p, max_iter, trials, seed in HGPS:
        code = jbp.Code.hgp(hm, hn, cw, rw, hseed)
        xs, zs = code.checks(0), code.checks(1)
        h.ref_set_checks(code.n, xs.shape[0], xs.shape[1], np.ascontiguousarray(xs).ctypes.data,
                         zs.shape[0], zs.shape[1], np.ascontiguousarray(zs).ctypes.data)
        h.ref_ets_clear()
        it = np.zeros(trials, np.int32)
        ok, su, ex, fl = (np.zeros(trials, np.uint8) for _ in range(4))
        est = np.zeros((trials, code.n), np.int32)
        ctr = np.zeros(10, np.int64)
        sec = h.ref_simulate_pp(seed, 1, trials, p, max_iter, 0.0, 5, 8, it.ctypes.data, ok.ctypes.data,
                                su.ctypes.data, ex.ctypes.data, fl.ctypes.data, est.ctypes.data,
                                ctr.ctypes.data)
        digests = np.array([est_digest(e) for e in est], dtype="S40")
        np.savez_compressed(os.path.join(HERE, "pp_trials_%s.npz" % name),
                            hgp=np.array([hm, hn, cw, rw, hseed]), p=p, max_iter=max_iter, seed=seed,
                            flip_hist=5, damping=0.0, first_trial=1, iterations=it, syn_match=ok,
                            success=su, exact=ex, flags=fl, est_sha1=digests, counters=ctr)
        print(name, "%.1fs" % sec, "counters", ctr.tolist(), "bp failures",
              int(np.count_nonzero((fl & 1) == 0)))


def main():
    h = load()
    info = np.zeros(9, np.int64)
    h.ref_code_info(info.ctypes.data, 0, None)
    n, mx, mz = int(info[0]), int(info[1]), int(info[2])
    counts = load_tables(h)
    with open(os.path.join(HERE, "pp_ets_counts.json"), "w") as f:
        json.dump({"params": STEM + ".txt", "counts": counts}, f, indent=1)

    def syndromes(states):
        sx = np.zeros(mx, np.int32)
        sz = np.zeros(mz, np.int32)
        st = np.ascontiguousarray(states, dtype=np.int32)
        h.ref_syndrome(st.ctypes.data, sx.ctypes.data, sz.ctypes.data)
        return sx, sz

    rng = np.random.default_rng(20251203)
    cases = {k: [] for k in ("method", "side", "est_in", "syn", "vars", "n_vars", "llr_q", "applied",
                             "est_out")}
    max_vars = 2400

    def add(method, side, est, syn, vars_=None, llr_q=None):
        est = np.ascontiguousarray(est, dtype=np.int32)
        out = est.copy()
        vv = np.full(max_vars, -1, np.int32)
        nv = 0
        if vars_ is not None:
            nv = len(vars_)
            vv[:nv] = vars_
        llr = None if llr_q is None else np.ascontiguousarray(llr_q, dtype=np.float64) / 64.0
        applied = h.ref_pp_repair(method, side, np.ascontiguousarray(syn, np.int32).ctypes.data,
                                  None if llr is None else llr.ctypes.data,
                                  vv.ctypes.data, nv, out.ctypes.data)
        cases["method"].append(method)
        cases["side"].append(side)
        cases["est_in"].append(pack2(est))
        cases["syn"].append(np.packbits(np.asarray(syn, np.uint8)))
        cases["vars"].append(vv)
        cases["n_vars"].append(nv)
        cases["llr_q"].append(np.zeros(n, np.uint16) if llr_q is None else np.asarray(llr_q, np.uint16))
        cases["applied"].append(applied)
        cases["est_out"].append(pack2(out))

    def random_error(weight):
        e = np.zeros(n, np.int32)
        idx = rng.choice(n, weight, replace=False)
        e[idx] = rng.integers(1, 4, weight)
        return e

    bit = (1, 2)
    # ETS: estimates that differ from a syndrome-consistent error by one
    # table entry (both unsatisfied checks of that entry), and by random pairs
    for side, label_files in ((0, ["ets_6_2.txt", "ets_8_2_path4.txt", "ets_12_2.txt",
                                   "ets_16_2.txt", "ets_18_2.txt"]),
                              (1, ["ets_6_2.txt", "ets_8_2.txt", "ets_20_2.txt"])):
        for name in label_files:
            entries = read_entries(name, "XZ"[side])
            for k in rng.choice(len(entries), min(3, len(entries)), replace=False):
                err = random_error(30)
                est = err.copy()
                for v in entries[k]:
                    est[v] ^= bit[side]
                syn = syndromes(err)[side]
                add(0, side, est, syn)
        for _ in range(3):
            err = random_error(30)
            est = err.copy()
            for v in rng.choice(n, 2, replace=False):
                est[v] ^= bit[side]
            add(0, side, est, syndromes(err)[side])
    # flip sets: a candidate set V, the estimate wrong on a subset of V
    for side in (0, 1):
        for size, wrong in ((5, 2), (12, 5), (40, 9), (120, 20), (600, 30)):
            err = random_error(40)
            V = rng.choice(n, size, replace=False)
            S = rng.choice(V, wrong, replace=False)
            est = err.copy()
            for v in S:
                est[v] ^= bit[side]
            add(1, side, est, syndromes(err)[side], vars_=V)
        # a wrong qubit outside V (unsolvable), more columns than checks,
        # a set with repeats and the zero-difference case
        err = random_error(40)
        V = rng.choice(n, 30, replace=False)
        est = err.copy()
        est[int(np.setdiff1d(np.arange(n), V)[0])] ^= bit[side]
        add(1, side, est, syndromes(err)[side], vars_=V)
        V = rng.choice(n, 2350, replace=False)
        est = err.copy()
        est[int(V[0])] ^= bit[side]
        add(1, side, est, syndromes(err)[side], vars_=V)
        V = np.concatenate([V[:20], V[:10]])
        est = err.copy()
        est[int(V[3])] ^= bit[side]
        add(1, side, est, syndromes(err)[side], vars_=V)
        add(1, side, err, syndromes(err)[side], vars_=V[:15])
    # OSD: |LLR| on a 1/64 grid (ties included), wrong qubits given low |LLR|
    for side in (0, 1):
        for wrong, spread in ((1, 40), (4, 40), (10, 60), (25, 80), (35, 80), (8, 1)):
            err = random_error(40)
            S = rng.choice(n, wrong, replace=False)
            est = err.copy()
            for v in S:
                est[v] ^= bit[side]
            llr_q = rng.integers(spread, 320, n).astype(np.uint16)
            llr_q[S] = rng.integers(0, max(spread // 2, 1), wrong)
            add(2, side, est, syndromes(err)[side], llr_q=llr_q)
        err = random_error(40)
        est = err.copy()
        for v in rng.choice(n, 6, replace=False):
            est[v] ^= bit[side]
        add(2, side, est, syndromes(err)[side], llr_q=rng.integers(0, 320, n).astype(np.uint16))
        add(2, side, est, syndromes(err)[side], llr_q=np.full(n, 64, np.uint16))
    np.savez_compressed(os.path.join(HERE, "pp_repairs.npz"),
                        **{k: np.asarray(v) for k, v in cases.items()})
    print("repairs:", len(cases["method"]), "applied:", int(np.count_nonzero(cases["applied"])))

    for name, p, max_iter, trials, seed, hist, damping in WORKLOADS:
        it = np.zeros(trials, np.int32)
        ok, su, ex, fl = (np.zeros(trials, np.uint8) for _ in range(4))
        est = np.zeros((trials, n), np.int32)
        ctr = np.zeros(10, np.int64)
        sec = h.ref_simulate_pp(seed, 1, trials, p, max_iter, damping, hist, 8, it.ctypes.data, ok.ctypes.data,
                                su.ctypes.data, ex.ctypes.data, fl.ctypes.data, est.ctypes.data,
                                ctr.ctypes.data)
        digests = np.array([est_digest(e) for e in est], dtype="S40")
        np.savez_compressed(os.path.join(HERE, "pp_trials_%s.npz" % name), p=p, max_iter=max_iter,
                            seed=seed, flip_hist=hist, damping=damping, first_trial=1, iterations=it, syn_match=ok,
                            success=su, exact=ex, flags=fl, est_sha1=digests, counters=ctr)
        print(name, "%.1fs" % sec, "counters", ctr.tolist())

    # |LLR| of the last iteration of plain decodes (what OSD orders by)
    llr_rec = {k: [] for k in ("max_iter", "trial", "iterations", "syn_match", "est_sha1",
                               "llr_sha1", "llr_head")}
    prior = np.zeros(4)
    h.ref_prior(0.04, prior.ctypes.data)
    for max_iter in (8, 50):
        for trial in range(1, 7):
            err = np.zeros(n, np.int32)
            h.ref_sample(7, trial, 0.04, err.ctypes.data)
            sx, sz = syndromes(err)
            est = np.zeros(n, np.int32)
            ok = C.c_int(0)
            llr = np.zeros(2 * n)
            it = h.ref_decode_llr(sx.ctypes.data, sz.ctypes.data, prior.ctypes.data, max_iter, 0.0,
                                  est.ctypes.data, C.byref(ok), llr.ctypes.data)
            llr_rec["max_iter"].append(max_iter)
            llr_rec["trial"].append(trial)
            llr_rec["iterations"].append(it)
            llr_rec["syn_match"].append(ok.value)
            llr_rec["est_sha1"].append(est_digest(est))
            llr_rec["llr_sha1"].append(hashlib.sha1(llr.tobytes()).hexdigest())
            llr_rec["llr_head"].append(llr[:8].copy())
    np.savez_compressed(os.path.join(HERE, "llr_decodes.npz"), p=0.04, seed=7,
                        **{k: np.asarray(v, dtype="S40" if k.endswith("sha1") else None)
                           for k, v in llr_rec.items()})
    print("llr:", list(zip(llr_rec["max_iter"], llr_rec["iterations"], llr_rec["syn_match"])))

    hgp_fixtures(h)

    # the reference program itself, PP on: its ETS Load table, a Run Summary
    # and a single --err decode that PP repairs.  It runs in a scratch
    # directory holding links to the params file and its table directory
    # (it looks for the tables under the working directory).
    g = np.load(os.path.join(HERE, "pp_trials_p0.04_it8.npz"))
    saved = np.nonzero((g["flags"] & 2) != 0)[0]
    trial = int(saved[0]) + 1
    h.ref_load_params(PARAMS.encode())
    err = np.zeros(n, np.int32)
    h.ref_sample(7, trial, 0.04, err.ctypes.data)
    with tempfile.TemporaryDirectory() as tmp:
        os.symlink(PARAMS, os.path.join(tmp, STEM + ".txt"))
        os.symlink(ETS_DIR, os.path.join(tmp, STEM))
        np.savetxt(os.path.join(tmp, "err.txt"), err, fmt="%d")
        sim = subprocess.run([REF_BIN, "--params", STEM + ".txt", "--simulate", "--p", "0.04",
                              "--trials", "256", "--max-iter", "8", "--seed", "7", "--report-every",
                              "0"], cwd=tmp, capture_output=True, text=True, timeout=600)
        one = subprocess.run([REF_BIN, "--params", STEM + ".txt", "--err", "err.txt", "--p", "0.04",
                              "--max-iter", "8", "--est", "est.txt"], cwd=tmp, capture_output=True,
                             text=True, timeout=600)
        est = np.loadtxt(os.path.join(tmp, "est.txt"), dtype=np.int64)
    lines = sim.stdout.splitlines()
    start = lines.index(next(ln for ln in lines if "ETS Load" in ln))
    ets_lines = [ln for ln in lines[start + 2:] if ln.startswith("ETS")]
    summary = [ln for ln in lines if ln.startswith("trials=")][-1]
    one_lines = one.stdout.splitlines()
    k = one_lines.index(next(ln for ln in one_lines if "Decode Result" in ln))
    result = [ln for ln in one_lines[k + 2:] if "=" in ln and not ln.startswith("avg_latency")]
    with open(os.path.join(HERE, "pp_cli.json"), "w") as f:
        json.dump({"ets_load": ets_lines, "simulate_args": ["--simulate", "--p", "0.04", "--trials", "256",
                                                           "--max-iter", "8", "--seed", "7"],
                   "run_summary": summary, "err_trial": trial, "err_seed": 7,
                   "err_args": ["--p", "0.04", "--max-iter", "8"], "err_exit": one.returncode,
                   "decode_result": result,
                   "est_sha1": hashlib.sha1(est.astype(np.uint8).tobytes()).hexdigest()}, f, indent=1)
    print("cli:", summary)
    print("err trial", trial, result)


if __name__ == "__main__":
    import sys
    if sys.argv[1:] == ["hgp"]:  # only the HGP fixtures
        hgp_fixtures(load())
    else:
        main()

"""MI355X joint belief-propagation decoder: Python host over libjbp's C ABI.

The library (libjbp.so, built from csrc/ into this directory) holds the HIP
kernels for gfx950 and the C entry points declared in include/jbp.h and
include/jbp_code.h.  This module is a thin ctypes layer with the reference
GPU boundary's vocabulary (jointbp_cuda.h): a graph, a context created on a
device, single decodes, and the batched Monte-Carlo call.

There is no CPU fallback here: if libjbp.so is missing, importing the
bindings raises.
"""
import atexit
import ctypes as C
import os
import re
import weakref

import numpy as np

__all__ = ["Code", "Decoder", "LearnedModel", "PPTables", "JbpError", "lib", "LIB_PATH", "FP32",
           "FP64", "prior_from_p", "pp_repair", "default_ets_dir"]

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG_DIR, "libjbp.so")
INCLUDE_DIR = os.path.join(os.path.dirname(PKG_DIR), "include")

FP32 = 32
FP64 = 64
STATUS = {0: "ok", 1: "invalid argument", 2: "device error", 3: "unsupported", 4: "out of memory"}
STATUS_UNSUPPORTED = 3
ENGINE_AUTO, ENGINE_LDS, ENGINE_HBM = 0, 1, 2
LEARNED_SUM_PRODUCT, LEARNED_MIN_SUM, LEARNED_LAYERED_MIN_SUM = 0, 1, 2


class JbpError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("%s (status %d: %s)" % (msg, code, STATUS.get(code, "?")))
        self.code = code


class GraphDesc(C.Structure):
    _fields_ = [(n, C.c_int) for n in ("nvars", "mX", "mZ", "x_edges", "z_edges")] + [
        (n, C.POINTER(C.c_int)) for n in (
            "x_check_offsets", "x_check_edges", "x_edge_var", "x_edge_check", "x_edge_pos",
            "x_var_offsets", "x_var_edges", "z_check_offsets", "z_check_edges", "z_edge_var",
            "z_edge_check", "z_edge_pos", "z_var_offsets", "z_var_edges")]


class DecodeResult(C.Structure):
    _fields_ = [("iterations", C.c_int), ("check_count", C.c_int), ("syndrome_match", C.c_int)] + [
        (n, C.c_double) for n in ("kernel_ms", "memcpy_ms", "check_kernel_ms", "check_memcpy_ms",
                                  "init_kernel_ms", "host_ms")]


class SimCfg(C.Structure):
    _fields_ = [("base_seed", C.c_uint64), ("first_trial", C.c_int64), ("n_trials", C.c_int64),
                ("p", C.c_double), ("max_iter", C.c_int), ("damping", C.c_double),
                ("precision", C.c_int), ("check_warmup", C.c_int), ("check_interval", C.c_int),
                ("slots", C.c_int), ("block", C.c_int), ("min_fails", C.c_int64),
                ("p_list", C.POINTER(C.c_double)), ("n_p", C.c_int), ("engine", C.c_int),
                ("schedule", C.c_int), ("graph", C.c_int), ("pp", C.c_int), ("flip_hist", C.c_int),
                ("chunk", C.c_int), ("freeze_syn", C.c_int)]


class TrialOut(C.Structure):
    _fields_ = [("iterations", C.POINTER(C.c_int)), ("syn_match", C.POINTER(C.c_uint8)),
                ("success", C.POINTER(C.c_uint8)), ("exact", C.POINTER(C.c_uint8)),
                ("err", C.POINTER(C.c_int)), ("est", C.POINTER(C.c_int)),
                ("pp_flags", C.POINTER(C.c_uint8))]


class RunStats(C.Structure):
    _fields_ = [(n, C.c_int64) for n in ("done", "failures", "bp_failures", "exact", "total_iters")] + [
        ("device_ms", C.c_double), ("elapsed_ms", C.c_double)] + [
        (n, C.c_int64) for n in ("pp_success", "flip_pp_success", "osd_pp_success", "stab_success",
                                 "ets_used", "ets_saved", "ets6_used", "ets6_saved", "ets12_used",
                                 "ets12_saved", "pp_trials")] + [("pp_ms", C.c_double)]


# per-trial post-processing flags (jbp_trial_out.pp_flags)
PPF_BP_OK, PPF_PP_SUCCESS, PPF_ETS_SAVED, PPF_FLIP_SAVED = 1, 2, 4, 8
PPF_OSD_SAVED, PPF_ETS_USED, PPF_ETS6, PPF_ETS12 = 16, 32, 64, 128
# jbp_pp_repair methods
PP_ETS, PP_FLIP, PP_OSD = 0, 1, 2
PP_OSD_MARGINALS = 3  # OSD given the normalised marginals [n][4] instead of |LLR|
# the reference's table groups in the order it loads and tries them
# (size, label, file name), jointbp_ets.cpp:5229-5327
ETS_GROUPS = [(6, "6", "ets_6_2.txt"), (8, "8", "ets_8_2.txt"), (8, "8p4", "ets_8_2_path4.txt"),
              (10, "10", "ets_10_2.txt"), (12, "12", "ets_12_2.txt"), (14, "14", "ets_14_2.txt"),
              (16, "16", "ets_16_2.txt"), (18, "18", "ets_18_2.txt"), (20, "20", "ets_20_2.txt"),
              (22, "22", "ets_22_2.txt")]


_LIB = None


def lib():
    """Load libjbp.so (once).  Raises when it has not been built."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise ImportError("libjbp.so not built (%s); run __graft_entry__.build() or make -C "
                          "joint_bp_plus_pp_amd/csrc" % LIB_PATH)
    L = C.CDLL(LIB_PATH)
    vp, cp, sz, i32, i64, u64, dbl = C.c_void_p, C.c_char_p, C.c_size_t, C.c_int, C.c_int64, C.c_uint64, C.c_double
    sig = {
        "jbp_create": (i32, [C.POINTER(GraphDesc), i32, C.POINTER(vp), cp, sz]),
        "jbp_destroy": (None, [vp]),
        "jbp_decode": (i32, [vp, vp, vp, vp, i32, i32, i32, i32, i32, i32, dbl, i32,
                             C.POINTER(DecodeResult), vp, cp, sz]),
        "jbp_decode_batch": (i32, [vp, i32, vp, vp, vp, i32, dbl, i32, vp, vp, vp, cp, sz]),
        "jbp_decode_batch_noalloc": (i32, [vp, i32, vp, vp, vp, i32, vp, vp, vp, vp, vp, cp, sz]),
        "jbp_simulate": (i32, [vp, C.POINTER(SimCfg), C.POINTER(TrialOut), C.POINTER(RunStats), cp, sz]),
        "jbp_simulate_mixed": (i32, [C.POINTER(vp), C.POINTER(SimCfg), i32, C.POINTER(TrialOut),
                                     C.POINTER(RunStats), cp, sz]),
        "jbp_sample": (i32, [vp, u64, i64, i32, dbl, vp, vp, vp, cp, sz]),
        "jbp_in_row_space": (i32, [vp, i32, i32, vp, vp, cp, sz]),
        "jbp_last_kernel_ms": (dbl, [vp, i32, C.POINTER(i64)]),
        "jbp_info": (i32, [vp, vp]),
        "jbp_learned_create": (i32, [vp, i32, C.POINTER(vp), cp, sz]),
        "jbp_learned_create_kind": (i32, [vp, i32, i32, C.POINTER(vp), cp, sz]),
        "jbp_learned_kind": (i32, [vp]),
        "jbp_learned_destroy": (None, [vp]),
        "jbp_learned_param_count": (i32, [vp]),
        "jbp_learned_set_params": (i32, [vp, vp, i32, cp, sz]),
        "jbp_learned_get_params": (i32, [vp, vp]),
        "jbp_learned_loss_grad": (i32, [vp, i32, i32, vp, vp, vp, vp, vp, cp, sz]),
        "jbp_learned_train_step": (i32, [vp, u64, i64, i32, dbl, dbl, dbl, vp, cp, sz]),
        "jbp_simulate_learned": (i32, [vp, vp, C.POINTER(SimCfg), C.POINTER(TrialOut), C.POINTER(RunStats), cp, sz]),
        "jbp_code_from_params": (i32, [cp, C.POINTER(vp), cp, sz]),
        "jbp_code_from_checks": (i32, [i32, i32, i32, vp, i32, i32, vp, C.POINTER(vp), cp, sz]),
        "jbp_code_hgp": (i32, [i32, i32, i32, i32, u64, C.POINTER(vp), cp, sz]),
        "jbp_code_free": (None, [vp]),
        "jbp_code_dims": (i32, [vp, vp]),
        "jbp_code_graph": (i32, [vp, C.POINTER(GraphDesc)]),
        "jbp_code_params": (i32, [vp, vp]),
        "jbp_code_checks": (i32, [vp, i32, vp]),
        "jbp_code_diagnostics": (i32, [vp, i32, vp]),
        "jbp_code_logicals": (i32, [vp, i32, C.POINTER(i64), C.POINTER(i64), vp, vp]),
        "jbp_set_logicals": (i32, [vp, i32, i64, vp, vp, cp, sz]),
        "jbp_tune": (i32, [vp, C.POINTER(SimCfg), vp, vp, C.POINTER(i32), cp, sz]),
        "jbp_pp_tables_create": (i32, [C.POINTER(vp)]),
        "jbp_pp_tables_free": (None, [vp]),
        "jbp_pp_tables_add": (i32, [vp, C.POINTER(GraphDesc), i64, i32, cp, cp, cp, vp, cp, sz]),
        "jbp_set_pp_tables": (i32, [vp, vp, cp, sz]),
        "jbp_pp_repair": (i32, [C.POINTER(GraphDesc), vp, i32, i32, vp, vp, vp, i32, vp, vp, cp, sz]),
        "jbp_decode_batch_pp": (i32, [vp, i32, vp, vp, vp, i32, dbl, i32, i32, vp, vp, vp, vp, cp, sz]),
        "jbp_decode_llr": (i32, [vp, i32, vp, vp, vp, i32, dbl, vp, vp, vp, vp, cp, sz]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _LIB = L
    return L


def header_symbols():
    """Every function name declared in include/*.h."""
    names = []
    for h in sorted(os.listdir(INCLUDE_DIR)):
        if not h.endswith(".h"):
            continue
        text = open(os.path.join(INCLUDE_DIR, h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names += re.findall(r"\b(jbp_[a-z0-9_]+)\s*\(", text)
    return sorted(set(names))


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _errbuf():
    return C.create_string_buffer(512)


def _check(rc, buf):
    if rc != 0:
        raise JbpError(rc, buf.value.decode(errors="replace") or "libjbp call failed")


def prior_from_p(p):
    """normalize_msg({1-p, p/3, p/3, p/3}) (jointbp_ets.cpp:5484-5485)."""
    m = [1.0 - p, p / 3.0, p / 3.0, p / 3.0]
    s = ((m[0] + m[1]) + m[2]) + m[3]
    return np.array([0.25] * 4 if s <= 0 else [v / s for v in m], dtype=np.float64)


class Code:
    """A CSS code's two Tanner graphs (host side)."""

    def __init__(self, handle):
        self._h = C.c_void_p(handle)
        dims = np.zeros(8, np.int64)
        lib().jbp_code_dims(self._h, _ptr(dims))
        self.n, self.mx, self.mz, self.x_edges, self.z_edges, self.P, self.J, self.L = map(int, dims)

    @classmethod
    def from_params(cls, path):
        h, e = C.c_void_p(), _errbuf()
        _check(lib().jbp_code_from_params(os.fspath(path).encode(), C.byref(h), e, 512), e)
        return cls(h.value)

    @classmethod
    def from_checks(cls, n, x_checks, z_checks):
        xs = np.ascontiguousarray(x_checks, dtype=np.int32)
        zs = np.ascontiguousarray(z_checks, dtype=np.int32)
        h, e = C.c_void_p(), _errbuf()
        _check(lib().jbp_code_from_checks(n, xs.shape[0], xs.shape[1], _ptr(xs), zs.shape[0],
                                          zs.shape[1], _ptr(zs), C.byref(h), e, 512), e)
        return cls(h.value)

    @classmethod
    def hgp(cls, m=150, n=200, col_w=3, row_w=4, seed=1):
        h, e = C.c_void_p(), _errbuf()
        _check(lib().jbp_code_hgp(m, n, col_w, row_w, seed, C.byref(h), e, 512), e)
        return cls(h.value)

    def close(self):
        if self._h and self._h.value:
            lib().jbp_code_free(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def graph(self):
        g = GraphDesc()
        lib().jbp_code_graph(self._h, C.byref(g))
        return g

    def checks(self, side):
        m = self.mx if side == 0 else self.mz
        deg = lib().jbp_code_checks(self._h, side, None)
        out = np.zeros((m, max(deg, 0)), np.int32)
        lib().jbp_code_checks(self._h, side, _ptr(out))
        return out

    def params(self):
        out = np.zeros(2 * self.L, np.int64)
        lib().jbp_code_params(self._h, _ptr(out))
        return out.reshape(4, -1)

    def logicals(self, side):
        """Logical-operator representatives (row_off int64, qubits int32) of
        a hypergraph-product code, or None for other codes: side 0 spans
        ker(Hz) / rowspace(Hx), side 1 spans ker(Hx) / rowspace(Hz)."""
        rows, ent = C.c_int64(), C.c_int64()
        rc = lib().jbp_code_logicals(self._h, side, C.byref(rows), C.byref(ent), None, None)
        if rc == STATUS_UNSUPPORTED:
            return None
        msg = C.create_string_buffer(b"jbp_code_logicals failed")
        _check(rc, msg)
        off = np.zeros(rows.value + 1, np.int64)
        q = np.zeros(max(ent.value, 1), np.int32)
        _check(lib().jbp_code_logicals(self._h, side, C.byref(rows), C.byref(ent), _ptr(off), _ptr(q)),
               msg)
        return off, q[:ent.value]

    def diagnostics(self, girth=True):
        out = np.zeros(15, np.int64)
        lib().jbp_code_diagnostics(self._h, 1 if girth else 0, _ptr(out))
        keys = ["rank_x", "rank_z", "x_var_deg_min", "x_var_deg_max", "x_check_deg_min",
                "x_check_deg_max", "z_var_deg_min", "z_var_deg_max", "z_check_deg_min",
                "z_check_deg_max", "odd_pairs", "girth_x", "girth_z"]
        return dict(zip(keys, map(int, out[:13])))

    def csr(self, side):
        """(check_offsets, edge_var) of one side as numpy arrays."""
        g = self.graph()
        m = self.mx if side == 0 else self.mz
        E = self.x_edges if side == 0 else self.z_edges
        off = g.x_check_offsets if side == 0 else g.z_check_offsets
        ev = g.x_edge_var if side == 0 else g.z_edge_var
        return (np.ctypeslib.as_array(off, shape=(m + 1,)).copy(),
                np.ctypeslib.as_array(ev, shape=(E,)).copy())


class PPTables:
    """Trapping-set tables for post-processing (jbp_pp_tables_*): host
    memory, no device needed."""

    def __init__(self, code):
        self.code = code
        self._graph = code.graph()
        h = C.c_void_p()
        rc = lib().jbp_pp_tables_create(C.byref(h))
        _check(rc, C.create_string_buffer(b"jbp_pp_tables_create failed"))
        self._h = h
        self.loaded = []  # (label, count_x, count_z)

    def add(self, size, label, path_x, path_z):
        """One group from the reference's text files; returns (count_x,
        count_z): entries, 0 for an empty file, -1 for a missing one."""
        counts = np.zeros(2, np.int32)
        e = _errbuf()
        enc = lambda x: None if x is None else os.fspath(x).encode()
        _check(lib().jbp_pp_tables_add(self._h, C.byref(self._graph), self.code.P, size,
                                       label.encode(), enc(path_x), enc(path_z), _ptr(counts),
                                       e, 512), e)
        self.loaded.append((label, int(counts[0]), int(counts[1])))
        return int(counts[0]), int(counts[1])

    @classmethod
    def from_dir(cls, code, directory):
        """The reference's default set: every group of ETS_GROUPS read from
        directory/<file> for both sides."""
        t = cls(code)
        for size, label, name in ETS_GROUPS:
            path = os.path.join(os.fspath(directory), name)
            t.add(size, label, path, path)
        return t

    def close(self):
        if self._h and self._h.value:
            lib().jbp_pp_tables_free(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def default_ets_dir(params_path):
    """Where the reference looks for a params file's tables: the directory
    named after the file without its extension (jointbp_ets.cpp:5031)."""
    return os.path.splitext(os.fspath(params_path))[0]


def pp_repair(code, method, side, syndrome, est, tables=None, abs_llr=None, flip_vars=None):
    """One host-side repair (jbp_pp_repair): returns (est_after, applied)."""
    g = code.graph()
    syn = np.ascontiguousarray(syndrome, dtype=np.int32)
    out = np.ascontiguousarray(est, dtype=np.int32).copy()
    llr = None if abs_llr is None else np.ascontiguousarray(abs_llr, dtype=np.float64)
    fv = None if flip_vars is None else np.ascontiguousarray(flip_vars, dtype=np.int32)
    applied = C.c_int(0)
    e = _errbuf()
    _check(lib().jbp_pp_repair(C.byref(g), tables._h if tables is not None else None, method, side,
                               _ptr(syn), _ptr(llr), _ptr(fv), 0 if fv is None else int(fv.size),
                               _ptr(out), C.byref(applied), e, 512), e)
    return out, applied.value


# Decoders still open when the interpreter exits are destroyed (with their
# learned models) before module teardown, while the library handle and this
# module's names are still usable, instead of leaving device objects to the
# runtime's own exit.
_LIVE = weakref.WeakSet()


@atexit.register
def _destroy_live_decoders():
    for d in list(_LIVE):
        try:
            d.destroy()
        except Exception:
            pass


class Decoder:
    """A decoder context on one device: cuda_bp_create / cuda_bp_decode /
    cuda_bp_destroy (jointbp_cuda.h:48-63) plus the batched calls."""

    def __init__(self, code, device=0):
        self.code = code
        self._graph = code.graph()
        h, e = C.c_void_p(), _errbuf()
        _check(lib().jbp_create(C.byref(self._graph), device, C.byref(h), e, 512), e)
        self._h = h
        self._models = weakref.WeakSet()  # learned models bound to this context
        _LIVE.add(self)
        # hypergraph-product codes carry their logical operators: the
        # success test uses them (the dense row basis is too large to build)
        lx, lz = code.logicals(0), code.logicals(1)
        if lx is not None and lz is not None:
            self.set_logicals(lx, lz)

    def set_logicals(self, lx, lz):
        """Judge success by logical operators (row_off, qubits) per side
        (jbp_set_logicals); None for both restores the row-basis test."""
        e = _errbuf()
        if lx is None and lz is None:
            _check(lib().jbp_set_logicals(self._h, -1, 0, None, None, e, 512), e)
            return
        for side, (off, q) in enumerate((lx, lz)):
            off = np.ascontiguousarray(off, np.int64)
            q = np.ascontiguousarray(q, np.int32)
            _check(lib().jbp_set_logicals(self._h, side, off.size - 1, _ptr(off), _ptr(q), e, 512), e)

    def set_pp_tables(self, tables):
        """Install trapping-set tables for simulate(pp=True) (None: none)."""
        e = _errbuf()
        _check(lib().jbp_set_pp_tables(self._h, tables._h if tables is not None else None, e, 512), e)

    def destroy(self):
        """Free the context; learned models bound to it are freed first."""
        for m in list(getattr(self, "_models", ())):
            m.destroy()
        if self._h and self._h.value:
            lib().jbp_destroy(self._h)
            self._h = C.c_void_p()

    close = destroy

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.destroy()

    def info(self):
        out = np.zeros(8, np.int64)
        lib().jbp_info(self._h, _ptr(out))
        return dict(zip(["qc", "P", "J", "L", "device", "cus", "block", "basis"], map(int, out)))

    def last_kernel_ms(self, which):
        n = C.c_int64(0)
        ms = lib().jbp_last_kernel_ms(self._h, which, C.byref(n))
        return ms, n.value

    def decode(self, sx, sz, prior, max_iter=50, check_warmup=0, check_interval=1, damping=0.0,
               precision=FP32, freeze_syn=False, use_graph=False):
        """One syndrome pair -> (est, iterations, syndrome_match, result)."""
        sx = np.ascontiguousarray(sx, dtype=np.int32)
        sz = np.ascontiguousarray(sz, dtype=np.int32)
        pr = np.ascontiguousarray(prior, dtype=np.float64)
        est = np.zeros(self.code.n, np.int32)
        res, e = DecodeResult(), _errbuf()
        _check(lib().jbp_decode(self._h, _ptr(sx), _ptr(sz), _ptr(pr), max_iter, check_warmup,
                                check_interval, 0, int(use_graph), int(freeze_syn), damping,
                                precision, C.byref(res), _ptr(est), e, 512), e)
        return est, res.iterations, bool(res.syndrome_match), res

    def decode_batch(self, sx, sz, prior, max_iter=50, damping=0.0, precision=FP32, want_est=True,
                     pp=False, flip_hist=5):
        """Batched decode of int syndrome arrays (count x m).  pp=True adds
        the reference's post-processing (jbp_decode_batch_pp) and returns a
        fourth array of PPF_* flags ("saved" = the repair produced the
        syndrome match)."""
        sx = np.atleast_2d(np.asarray(sx))
        sz = np.atleast_2d(np.asarray(sz))
        count = sx.shape[0]
        bx, bz = pack_bits(sx), pack_bits(sz)
        pr = np.ascontiguousarray(prior, dtype=np.float64)
        it = np.zeros(count, np.int32)
        ok = np.zeros(count, np.uint8)
        est = np.zeros((count, self.code.n), np.int32) if want_est else None
        e = _errbuf()
        if pp:
            flags = np.zeros(count, np.uint8)
            _check(lib().jbp_decode_batch_pp(self._h, count, _ptr(bx), _ptr(bz), _ptr(pr), max_iter,
                                             damping, precision, flip_hist, _ptr(it), _ptr(ok),
                                             _ptr(flags), _ptr(est), e, 512), e)
            return it, ok, est, flags
        _check(lib().jbp_decode_batch(self._h, count, _ptr(bx), _ptr(bz), _ptr(pr), max_iter,
                                      damping, precision, _ptr(it), _ptr(ok), _ptr(est), e, 512), e)
        return it, ok, est

    def decode_llr(self, sx, sz, prior, max_iter=50, damping=0.0):
        """FP64 decodes (the reference's, per decode) with the |LLR| of the
        last iteration: returns (iterations, syn_match, est, llr) with llr of
        shape (count, 2, n), X then Z (jbp_decode_llr)."""
        sx = np.atleast_2d(np.asarray(sx))
        sz = np.atleast_2d(np.asarray(sz))
        count = sx.shape[0]
        bx, bz = pack_bits(sx), pack_bits(sz)
        pr = np.ascontiguousarray(prior, dtype=np.float64)
        it = np.zeros(count, np.int32)
        ok = np.zeros(count, np.uint8)
        est = np.zeros((count, self.code.n), np.int32)
        llr = np.zeros((count, 2, self.code.n), np.float64)
        e = _errbuf()
        _check(lib().jbp_decode_llr(self._h, count, _ptr(bx), _ptr(bz), _ptr(pr), max_iter, damping,
                                    _ptr(it), _ptr(ok), _ptr(est), _ptr(llr), e, 512), e)
        return it, ok, est, llr

    def sample(self, base_seed, first_trial, count, p):
        """Device-sampled errors and syndromes of trials first..first+count-1."""
        err = np.zeros((count, self.code.n), np.int32)
        sx = np.zeros((count, self.code.mx), np.int32)
        sz = np.zeros((count, self.code.mz), np.int32)
        e = _errbuf()
        _check(lib().jbp_sample(self._h, base_seed, first_trial, count, p, _ptr(err), _ptr(sx),
                                _ptr(sz), e, 512), e)
        return err, sx, sz

    def in_row_space(self, side, bits):
        bits = np.ascontiguousarray(np.atleast_2d(bits), dtype=np.int32)
        out = np.zeros(bits.shape[0], np.uint8)
        e = _errbuf()
        _check(lib().jbp_in_row_space(self._h, side, bits.shape[0], _ptr(bits), _ptr(out), e, 512), e)
        return out.astype(bool)

    def tune(self, base_seed, n_trials, p, max_iter=50, precision=FP32, engine=0, first_trial=1,
             check_warmup=0, check_interval=1):
        """jbp_tune on this workload: returns (slots, block, {(slots, block): ms})
        and keeps the fastest for later calls that pass slots = block = 0."""
        plist = None
        if np.ndim(p) > 0:
            plist = np.ascontiguousarray(p, dtype=np.float64)
            p = float(plist[0])
        cfg = SimCfg(base_seed, first_trial, n_trials, p, max_iter, 0.0, precision, check_warmup,
                     check_interval, 0, 0, 0,
                     plist.ctypes.data_as(C.POINTER(C.c_double)) if plist is not None else None,
                     int(plist.size) if plist is not None else 0, engine, 0, 0)
        best = np.zeros(2, np.int32)
        ms = np.zeros(64, np.float64)
        n = C.c_int(64)
        e = _errbuf()
        _check(lib().jbp_tune(self._h, C.byref(cfg), _ptr(best), _ptr(ms), C.byref(n), e, 512), e)
        return int(best[0]), int(best[1]), ms[:n.value].tolist()

    def simulate_learned(self, model, base_seed, n_trials, p, max_iter=1000, first_trial=1,
                         check_warmup=0, check_interval=1, min_fails=0, per_trial=True,
                         want_err=False, want_est=False, chunk=0):
        """jbp_simulate with the learned decoder for model.T iterations, then
        plain FP32 BP from the prior on trials that did not converge."""
        return self.simulate(base_seed, n_trials, p, max_iter=max_iter, first_trial=first_trial,
                             precision=FP32, check_warmup=check_warmup,
                             check_interval=check_interval, min_fails=min_fails,
                             per_trial=per_trial, want_err=want_err, want_est=want_est,
                             chunk=chunk, _model=model)

    def simulate(self, base_seed, n_trials, p, max_iter=50, first_trial=1, damping=0.0,
                 precision=FP32, check_warmup=0, check_interval=1, min_fails=0, slots=0, block=0,
                 per_trial=True, want_err=False, want_est=False, engine=0, schedule=0,
                 graph=False, pp=False, flip_hist=5, chunk=0, freeze_syn=False, _model=None):
        """The --simulate inner loop on the device (post-processing off unless
        pp=True, i.e. the reference's --no-pp by default).  Returns (stats,
        per-trial dict or None).  p may be a sequence: trial index idx then
        uses p[idx % len(p)] (interleaved mixed-noise run).  engine selects
        the FP32 decoder (ENGINE_AUTO, ENGINE_LDS, ENGINE_HBM); slots is the
        number of resident trials of the HBM engine (FP64: of the parity
        decoder).  schedule=1 decodes in index order instead of heaviest
        syndrome first (results are the same); graph=True replays the HBM
        engine's step batches from a captured graph.  pp=True runs the
        reference's post-processing with the tables of set_pp_tables
        (flip_hist: --flip-hist); per-trial "pp_flags" then holds PPF_* bits.
        chunk: trials per persistent launch of the LDS engine when a run
        spans several launches (0 = 8192)."""
        plist = None
        if np.ndim(p) > 0:
            plist = np.ascontiguousarray(p, dtype=np.float64)
            p = float(plist[0]) if plist.size else 0.0
        cfg = SimCfg(base_seed, first_trial, n_trials, p, max_iter, damping, precision, check_warmup,
                     check_interval, slots, block, min_fails,
                     plist.ctypes.data_as(C.POINTER(C.c_double)) if plist is not None else None,
                     int(plist.size) if plist is not None else 0, engine, schedule, int(graph),
                     int(pp), flip_hist, chunk, int(freeze_syn))
        out = None
        arrays = {}
        if per_trial:
            arrays = {"iterations": np.zeros(n_trials, np.int32),
                      "syn_match": np.zeros(n_trials, np.uint8),
                      "success": np.zeros(n_trials, np.uint8),
                      "exact": np.zeros(n_trials, np.uint8),
                      "pp_flags": np.zeros(n_trials, np.uint8)}
            if want_err:
                arrays["err"] = np.zeros((n_trials, self.code.n), np.int32)
            if want_est:
                arrays["est"] = np.zeros((n_trials, self.code.n), np.int32)
            out = TrialOut()
            for k, a in arrays.items():
                setattr(out, k, a.ctypes.data_as(C.POINTER(C.c_int if a.dtype == np.int32 else C.c_uint8)))
        st, e = RunStats(), _errbuf()
        if _model is None:
            rc = lib().jbp_simulate(self._h, C.byref(cfg), C.byref(out) if out else None,
                                    C.byref(st), e, 512)
        else:
            rc = lib().jbp_simulate_learned(self._h, _model._h, C.byref(cfg),
                                            C.byref(out) if out else None, C.byref(st), e, 512)
        _check(rc, e)
        stats = {k: getattr(st, k) for k, _ in RunStats._fields_}
        if per_trial:
            d = int(stats["done"])
            arrays = {k: v[:d] for k, v in arrays.items()}
        return stats, (arrays if per_trial else None)


class LearnedModel:
    """Learned-BP parameters on a decoder's device (include/jbp.h,
    jbp_learned_*).  Layout: wc[2][nb][T] | wv[2][nb][T] | mlp[65].  kind:
    LEARNED_SUM_PRODUCT (the reference's joint BP with weights),
    LEARNED_MIN_SUM (normalised min-sum) or LEARNED_LAYERED_MIN_SUM (the
    min-sum model on a layered schedule: fewer iterations; no damping)."""

    MLP_PARAMS = 65

    def __init__(self, decoder, T=25, kind=0):
        self.decoder = decoder
        self.T = T
        self.kind = int(kind)
        h, e = C.c_void_p(), _errbuf()
        _check(lib().jbp_learned_create_kind(decoder._h, T, self.kind, C.byref(h), e, 512), e)
        self._h = h
        decoder._models.add(self)
        self.count = lib().jbp_learned_param_count(h)
        self.nb = (self.count - self.MLP_PARAMS) // (4 * T)
        self.use_mlp = self.kind != LEARNED_LAYERED_MIN_SUM

    def destroy(self):
        if self._h and self._h.value:
            lib().jbp_learned_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass

    def get_params(self):
        out = np.zeros(self.count, np.float64)
        lib().jbp_learned_get_params(self._h, _ptr(out))
        return out

    def set_params(self, params, use_mlp=True):
        params = np.ascontiguousarray(params, dtype=np.float64)
        assert params.size == self.count
        e = _errbuf()
        _check(lib().jbp_learned_set_params(self._h, _ptr(params), int(use_mlp), e, 512), e)
        self.use_mlp = bool(use_mlp)

    def split(self, params):
        k = 2 * self.nb * self.T
        return (params[:k].reshape(2, self.nb, self.T), params[k:2 * k].reshape(2, self.nb, self.T),
                params[2 * k:])

    def loss_grad(self, err, p, precision=FP64, grad=True):
        err = np.ascontiguousarray(np.atleast_2d(err), dtype=np.int32)
        batch = err.shape[0]
        p = np.ascontiguousarray(np.broadcast_to(p, (batch,)), dtype=np.float64)
        loss = C.c_double(0.0)
        g = np.zeros(self.count, np.float64) if grad else None
        conv = np.zeros(batch, np.int32)
        e = _errbuf()
        _check(lib().jbp_learned_loss_grad(self._h, precision, batch, _ptr(err), _ptr(p),
                                           C.byref(loss), _ptr(g), _ptr(conv), e, 512), e)
        return loss.value, g, conv

    def train_step(self, base_seed, first_trial, batch, p_lo=0.03, p_hi=0.05, lr=1e-2):
        loss = C.c_double(0.0)
        e = _errbuf()
        _check(lib().jbp_learned_train_step(self._h, base_seed, first_trial, batch, p_lo, p_hi, lr,
                                            C.byref(loss), e, 512), e)
        return loss.value


def simulate_mixed(runs, per_trial=True):
    """jbp_simulate_mixed: up to two (decoder, settings) pairs, each a
    different code's Decoder on the same device, decoded together by the
    HBM-resident FP32 engine (every engine step's kernels cover both codes'
    trials).  settings: base_seed, n_trials, p (a number or a list: trial
    index idx uses p[idx % len(p)]), and optionally max_iter (50),
    first_trial (1), damping, check_warmup, check_interval, slots,
    schedule.  Per trial the outcome equals that decoder's simulate() with
    engine=ENGINE_HBM.  Returns [(stats, per-trial dict or None), ...]."""
    n = len(runs)
    cfgs = (SimCfg * n)()
    ctxs = (C.c_void_p * n)()
    outs = (TrialOut * n)()
    stats = (RunStats * n)()
    keep, arrays = [], []
    for i, (dec, kw) in enumerate(runs):
        p = kw["p"]
        plist = None
        if np.ndim(p) > 0:
            plist = np.ascontiguousarray(p, dtype=np.float64)
            keep.append(plist)
            p = float(plist[0]) if plist.size else 0.0
        nt = int(kw["n_trials"])
        cfgs[i] = SimCfg(kw["base_seed"], kw.get("first_trial", 1), nt, p, kw.get("max_iter", 50),
                         kw.get("damping", 0.0), FP32, kw.get("check_warmup", 0),
                         kw.get("check_interval", 1), kw.get("slots", 0), 0, 0,
                         plist.ctypes.data_as(C.POINTER(C.c_double)) if plist is not None else None,
                         int(plist.size) if plist is not None else 0, ENGINE_HBM,
                         kw.get("schedule", 0), 0, 0, 5, 0, 0)
        ctxs[i] = dec._h
        a = {}
        if per_trial:
            a = {"iterations": np.zeros(nt, np.int32), "syn_match": np.zeros(nt, np.uint8),
                 "success": np.zeros(nt, np.uint8), "exact": np.zeros(nt, np.uint8),
                 "pp_flags": np.zeros(nt, np.uint8)}
            for k, v in a.items():
                setattr(outs[i], k, v.ctypes.data_as(C.POINTER(C.c_int if v.dtype == np.int32 else C.c_uint8)))
        arrays.append(a)
    e = _errbuf()
    _check(lib().jbp_simulate_mixed(ctxs, cfgs, n, outs if per_trial else None, stats, e, 512), e)
    res = []
    for i in range(n):
        st = {k: getattr(stats[i], k) for k, _ in RunStats._fields_}
        res.append((st, arrays[i] if per_trial else None))
    return res


def pack_bits(a):
    """int arrays (count x len, nonzero = 1) -> uint32 bit planes (count x words)."""
    a = np.atleast_2d(np.asarray(a)) != 0
    count, length = a.shape
    words = max(1, (length + 31) // 32)
    pad = np.zeros((count, words * 32), dtype=bool)
    pad[:, :length] = a
    b = pad.reshape(count, words, 32).astype(np.uint64)
    return np.ascontiguousarray((b << np.arange(32, dtype=np.uint64)).sum(axis=2).astype(np.uint32))

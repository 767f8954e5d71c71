"""ctypes view of oracle/liboracle.so (the CPU checker).  Test-only."""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


class OracleCode:
    def __init__(self, lib, handle):
        self.lib = lib
        self.h = C.c_void_p(handle)
        info = np.zeros(7, np.int64)
        lib.orc_code_info(self.h, _p(info))
        self.n, self.mx, self.mz, self.rank_x, self.rank_z, self.ex, self.ez = map(int, info)

    def __del__(self):
        try:
            self.lib.orc_code_free(self.h)
        except Exception:
            pass

    def csr(self, side):
        m = self.mx if side == 0 else self.mz
        E = self.ex if side == 0 else self.ez
        off = np.zeros(m + 1, np.int32)
        ev = np.zeros(E, np.int32)
        self.lib.orc_code_csr(self.h, side, _p(off), _p(ev))
        return off, ev

    def sample(self, base, idx, p):
        err = np.zeros(self.n, np.int32)
        self.lib.orc_sample(self.h, base, idx, p, _p(err))
        return err

    def syndrome(self, err):
        err = np.ascontiguousarray(err, dtype=np.int32)
        sx = np.zeros(self.mx, np.int32)
        sz = np.zeros(self.mz, np.int32)
        self.lib.orc_syndrome(self.h, _p(err), _p(sx), _p(sz))
        return sx, sz

    def decode(self, sx, sz, prior, max_iter, damping=0.0, freeze_syn=False):
        sx = np.ascontiguousarray(sx, dtype=np.int32)
        sz = np.ascontiguousarray(sz, dtype=np.int32)
        pr = np.ascontiguousarray(prior, dtype=np.float64)
        est = np.zeros(self.n, np.int32)
        ok = C.c_int(0)
        it = self.lib.orc_decode(self.h, _p(sx), _p(sz), _p(pr), max_iter, damping,
                                 1 if freeze_syn else 0, _p(est), C.byref(ok))
        return est, it, bool(ok.value)

    def stab_equiv(self, err, est):
        err = np.ascontiguousarray(err, dtype=np.int32)
        est = np.ascontiguousarray(est, dtype=np.int32)
        return bool(self.lib.orc_stab_equiv(self.h, _p(err), _p(est)))

    def in_row_space(self, side, bits):
        bits = np.ascontiguousarray(bits, dtype=np.int32)
        return bool(self.lib.orc_in_row_space(self.h, side, _p(bits)))

    def simulate(self, base, first, count, p, max_iter, damping=0.0, threads=8):
        it = np.zeros(count, np.int32)
        ok = np.zeros(count, np.uint8)
        su = np.zeros(count, np.uint8)
        ex = np.zeros(count, np.uint8)
        eh = np.zeros(count, np.uint64)
        secs = self.lib.orc_simulate(self.h, base, first, count, p, max_iter, damping, threads,
                                     _p(it), _p(ok), _p(su), _p(ex), _p(eh))
        return {"iterations": it, "syn_ok": ok, "success": su, "exact": ex, "est_hash": eh,
                "seconds": secs}


class Oracle:
    def __init__(self, path=ORACLE_SO):
        if not os.path.exists(path):
            raise RuntimeError("oracle not built: make -C oracle oracle")
        L = C.CDLL(path)
        vp = C.c_void_p
        L.orc_code_from_params.restype = vp
        L.orc_code_from_params.argtypes = [C.c_char_p]
        L.orc_code_from_checks.restype = vp
        L.orc_code_from_checks.argtypes = [C.c_int, C.c_int, C.c_int, vp, C.c_int, C.c_int, vp]
        L.orc_code_free.argtypes = [vp]
        L.orc_code_info.argtypes = [vp, vp]
        L.orc_code_csr.argtypes = [vp, C.c_int, vp, vp]
        L.orc_splitmix64.restype = C.c_uint64
        L.orc_splitmix64.argtypes = [C.c_uint64]
        L.orc_trial_seed.restype = C.c_uint64
        L.orc_trial_seed.argtypes = [C.c_uint64, C.c_uint64]
        L.orc_mt_outputs.argtypes = [C.c_uint64, C.c_longlong, vp]
        L.orc_sample.argtypes = [vp, C.c_uint64, C.c_uint64, C.c_double, vp]
        L.orc_syndrome.argtypes = [vp, vp, vp, vp]
        L.orc_prior.argtypes = [C.c_double, vp]
        L.orc_decode.restype = C.c_int
        L.orc_decode.argtypes = [vp, vp, vp, vp, C.c_int, C.c_double, C.c_int, vp, vp]
        L.orc_stab_equiv.argtypes = [vp, vp, vp]
        L.orc_in_row_space.argtypes = [vp, C.c_int, vp]
        L.orc_simulate.restype = C.c_double
        L.orc_simulate.argtypes = [vp, C.c_uint64, C.c_longlong, C.c_longlong, C.c_double, C.c_int,
                                   C.c_double, C.c_int, vp, vp, vp, vp, vp]
        self.lib = L

    def code_from_params(self, path):
        h = self.lib.orc_code_from_params(os.fspath(path).encode())
        if not h:
            raise RuntimeError("oracle could not read " + path)
        return OracleCode(self.lib, h)

    def code_from_checks(self, n, x_checks, z_checks):
        xs = np.ascontiguousarray(x_checks, dtype=np.int32)
        zs = np.ascontiguousarray(z_checks, dtype=np.int32)
        h = self.lib.orc_code_from_checks(n, xs.shape[0], xs.shape[1], _p(xs), zs.shape[0],
                                          zs.shape[1], _p(zs))
        return OracleCode(self.lib, h)

    def splitmix64(self, x):
        return int(self.lib.orc_splitmix64(x))

    def trial_seed(self, base, idx):
        return int(self.lib.orc_trial_seed(base, idx))

    def mt_outputs(self, seed, count):
        out = np.zeros(count, np.uint64)
        self.lib.orc_mt_outputs(seed, count, _p(out))
        return out

    def prior(self, p):
        out = np.zeros(4, np.float64)
        self.lib.orc_prior(p, _p(out))
        return out


class LearnedOracle:
    """FP64 learned-BP forward / loss / adjoint (oracle/learned_oracle.cpp).
    kind 0: sum-product, 1: min-sum, 2: layered min-sum.  Weight blocks: edge
    at position i of check r*P + x is block r*L + i, or explicit per-edge
    blocks (blk, nb).  Layers (kind 2): check r*P + x is in layer r, or
    explicit per-check layers (lay, nlay)."""

    def __init__(self, code, P=768, L=12, path=ORACLE_SO, blk=None, nb=None, lay=None, nlay=None):
        L_ = C.CDLL(path)
        vp = C.c_void_p
        L_.orl_loss.restype = C.c_double
        L_.orl_loss.argtypes = [C.c_int] * 3 + [vp] * 6 + [C.c_int, C.c_int, vp, vp, vp, C.c_int,
                                                           C.c_int, vp, vp, vp, vp, vp, vp, vp]
        L_.orl_loss2.restype = C.c_double
        L_.orl_loss2.argtypes = [C.c_int] * 4 + [vp] * 6 + [C.c_int, C.c_int, vp, vp, vp, C.c_int,
                                                            C.c_int, vp, vp, vp, vp, vp, vp, vp,
                                                            vp, vp]
        L_.orl_loss3.restype = C.c_double
        L_.orl_loss3.argtypes = ([C.c_int] * 4 + [vp] * 6 + [C.c_int, vp, vp, C.c_int, C.c_int] +
                                 [vp] * 3 + [C.c_int, C.c_int] + [vp] * 9)
        self.lib = L_
        self.code = code
        self.xoff, self.xvar = code.csr(0)
        self.zoff, self.zvar = code.csr(1)
        if lay is not None:
            self.lay = [np.ascontiguousarray(x, dtype=np.int32) for x in lay]
            self.nlay = int(nlay)
        else:
            self.lay = [np.arange(m, dtype=np.int32) // P for m in (code.mx, code.mz)]
            self.nlay = max(1, code.mx // P)
        if blk is not None:
            self.blk = [np.ascontiguousarray(b, dtype=np.int32) for b in blk]
            self.nb = int(nb)
            return
        self.nb = (code.mx // P) * L
        self.blk = []
        for off in (self.xoff, self.zoff):
            b = np.zeros(off[-1], np.int32)
            for c in range(len(off) - 1):
                r = c // P
                b[off[c]:off[c + 1]] = r * L + np.arange(off[c + 1] - off[c])
            self.blk.append(b)

    def loss(self, T, wc, wv, mlp, use_mlp, err, p, grad=False, kind=0, extra=False):
        """(loss, [grads], conv, est) and, with extra=True, a dict with the
        forward pass's signature hash and the last marginals (min-sum)."""
        err = np.ascontiguousarray(err, dtype=np.int32)
        p = np.ascontiguousarray(np.broadcast_to(p, (err.shape[0],)), dtype=np.float64)
        wc = np.ascontiguousarray(wc, dtype=np.float64)
        wv = np.ascontiguousarray(wv, dtype=np.float64)
        mlp = np.ascontiguousarray(mlp, dtype=np.float64)
        g = [np.zeros_like(wc), np.zeros_like(wv), np.zeros_like(mlp)] if grad else [None] * 3
        conv = np.zeros(err.shape[0], np.int32)
        est = np.zeros_like(err)
        sig = np.zeros(1, np.uint64)
        marg = np.zeros((err.shape[0], 2 * err.shape[1]), np.float64)
        c = self.code
        val = self.lib.orl_loss3(int(kind), c.n, c.mx, c.mz, _p(self.xoff), _p(self.xvar),
                                 _p(self.zoff), _p(self.zvar), _p(self.blk[0]), _p(self.blk[1]),
                                 self.nb, _p(self.lay[0]), _p(self.lay[1]), self.nlay, T, _p(wc),
                                 _p(wv), _p(mlp), int(use_mlp), err.shape[0], _p(err), _p(p),
                                 _p(g[0]), _p(g[1]), _p(g[2]), _p(conv), _p(est), _p(sig), _p(marg))
        out = (val, g, conv, est) if grad else (val, conv, est)
        if extra:
            out = out + ({"signature": int(sig[0]), "marginals": marg},)
        return out

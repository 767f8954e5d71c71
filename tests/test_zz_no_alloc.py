"""GPU: the allocation-free decode (jbp_decode_batch_noalloc) against the
allocating batched decode (jbp_decode_batch) on the same syndromes, per
decode (iterations, syndrome match, estimate), on two calls in a row, the
second smaller (the scratch array keeps the first call's contents)."""
import ctypes as C

import numpy as np
import pytest

import joint_bp_plus_pp_amd as jbp

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


class DeviceBuffers:
    """Device buffers from the HIP runtime libjbp itself links."""

    def __init__(self):
        self.hip = C.CDLL("libamdhip64.so.7")
        self.hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.hip.hipFree.argtypes = [C.c_void_p]
        self.hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.hip.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
        self.ptrs = []

    def alloc(self, nbytes, fill=None):
        p = C.c_void_p()
        assert self.hip.hipMalloc(C.byref(p), C.c_size_t(max(nbytes, 4))) == 0
        self.ptrs.append(p)
        if fill is not None:
            assert self.hip.hipMemset(p, fill, C.c_size_t(nbytes)) == 0
        return p

    def upload(self, arr):
        arr = np.ascontiguousarray(arr)
        p = self.alloc(arr.nbytes)
        assert self.hip.hipMemcpy(p, arr.ctypes.data_as(C.c_void_p), arr.nbytes, H2D) == 0
        return p

    def download(self, p, shape, dtype):
        out = np.empty(shape, dtype)
        assert self.hip.hipDeviceSynchronize() == 0
        assert self.hip.hipMemcpy(out.ctypes.data_as(C.c_void_p), p, out.nbytes, D2H) == 0
        return out

    def free(self):
        for p in self.ptrs:
            self.hip.hipFree(p)
        self.ptrs = []


def unpack_planes(bx, bz, n):
    bits = lambda b: ((b[:, :, None] >> np.arange(32, dtype=np.uint32)) & 1).reshape(b.shape[0], -1)[:, :n]
    return (bits(bx) | (bits(bz) << 1)).astype(np.int32)


def test_noalloc_equals_decode_batch_two_calls(main_decoder):
    dec = main_decoder
    L = jbp.lib()
    n = dec.code.n
    wn = (n + 31) // 32
    prior = np.ascontiguousarray(jbp.prior_from_p(0.05), dtype=np.float64)
    mem = DeviceBuffers()
    try:
        for count, base, max_iter in ((3000, 106, 60), (700, 107, 45)):
            print("noalloc call: %d trials, seed %d, max-iter %d" % (count, base, max_iter), flush=True)
            _, sx, sz = dec.sample(base, 1, count, 0.05)
            it, ok, est = dec.decode_batch(sx, sz, prior, max_iter=max_iter)
            d_sx = mem.upload(jbp.pack_bits(sx))
            d_sz = mem.upload(jbp.pack_bits(sz))
            # outputs start as garbage so that stale values cannot pass
            d_it = mem.alloc(4 * count, fill=0x7f)
            d_ok = mem.alloc(count, fill=0x7f)
            d_ex = mem.alloc(4 * count * wn, fill=0x5a)
            d_ez = mem.alloc(4 * count * wn, fill=0x5a)
            e = C.create_string_buffer(512)
            rc = L.jbp_decode_batch_noalloc(dec._h, count, d_sx, d_sz, prior.ctypes.data_as(C.c_void_p),
                                            max_iter, d_it, d_ok, d_ex, d_ez, None, e, 512)
            assert rc == 0, e.value
            g_it = mem.download(d_it, count, np.int32)
            g_ok = mem.download(d_ok, count, np.uint8)
            g_est = unpack_planes(mem.download(d_ex, (count, wn), np.uint32),
                                  mem.download(d_ez, (count, wn), np.uint32), n)
            assert np.array_equal(g_it, it)
            assert np.array_equal(g_ok, ok)
            assert np.array_equal(g_est, est)
            assert 0 < int(np.sum(ok)) < count  # converged and unconverged decodes both seen
            mem.free()
        # sizes beyond the scratch array are refused, nothing launched
        e = C.create_string_buffer(512)
        rc = L.jbp_decode_batch_noalloc(dec._h, 65537, None, None, prior.ctypes.data_as(C.c_void_p),
                                        10, None, None, None, None, None, e, 512)
        assert rc == 1 and b"65536" in e.value  # JBP_ERR_ARG
    finally:
        mem.free()

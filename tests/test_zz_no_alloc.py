"""mw3_cheby_spectrum_noalloc against mw3_cheby_spectrum through the C ABI.

The no-allocation entry point takes device buffers; the test allocates them
with the HIP runtime that libmw3 itself links (libamdhip64.so.7, through
ctypes), never with another runtime's allocator.  Two calls in a row, the
second smaller, must give the allocating entry point's spectrum bit for bit
(same kernels, same order of summation); a kmax beyond the library's array
is refused before any launch.
"""
import ctypes as ct

import numpy as np
import pytest

from madwave3_amd import lib as L

pytestmark = pytest.mark.gpu


class Hip:
    def __init__(self):
        L.load_library()                 # loads libamdhip64.so.7 as its dependency
        self.h = ct.CDLL("libamdhip64.so.7")
        self.h.hipMalloc.argtypes = [ct.POINTER(ct.c_void_p), ct.c_size_t]
        self.h.hipFree.argtypes = [ct.c_void_p]
        self.h.hipMemcpy.argtypes = [ct.c_void_p, ct.c_void_p, ct.c_size_t, ct.c_int]
        self.h.hipDeviceSynchronize.argtypes = []
        self.bufs = []

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        p = ct.c_void_p()
        assert self.h.hipMalloc(ct.byref(p), max(a.nbytes, 8)) == 0
        self.bufs.append(p)
        assert self.h.hipMemcpy(p, a.ctypes.data_as(ct.c_void_p), a.nbytes, 1) == 0   # H2D
        return p

    def empty(self, n):
        p = ct.c_void_p()
        assert self.h.hipMalloc(ct.byref(p), max(8 * n, 8)) == 0
        self.bufs.append(p)
        return p

    def download(self, p, n):
        out = np.empty(n)
        assert self.h.hipMemcpy(out.ctypes.data_as(ct.c_void_p), p, 8 * n, 2) == 0     # D2H
        return out

    def free(self):
        for p in self.bufs:
            self.h.hipFree(p)
        self.bufs = []


def _noalloc(lib, hip, autoc, e, emindlt, delta2, hbr, xkgam):
    da, de, ds = hip.upload(autoc), hip.upload(e), hip.empty(e.size)
    rc = lib.mw3_cheby_spectrum_noalloc(da, autoc.size - 1, ct.c_double(emindlt),
                                        ct.c_double(delta2), ct.c_double(hbr),
                                        ct.c_double(xkgam), de, e.size, ds, None)
    assert rc == 0, lib.mw3_last_error(None)
    assert hip.h.hipDeviceSynchronize() == 0
    return hip.download(ds, e.size)


def test_noalloc_spectrum_matches_allocating_entry_point():
    lib = L.load_library()
    hip = Hip()
    rng = np.random.default_rng(11)
    emindlt, delta2, hbr, xkgam = 0.37, 2.6, 0.0656, 1.5e-4
    try:
        # first call: 60001 coefficients (a 60,000-step run), 700 energies
        k1 = 60000
        a1 = np.concatenate([[1.0], rng.standard_normal(k1) * np.exp(-np.arange(1, k1 + 1) / 9e3)])
        e1 = np.linspace(-2.0, 3.0, 700)
        s1 = _noalloc(lib, hip, a1, e1, emindlt, delta2, hbr, xkgam)
        r1 = L.cheby_spectrum(a1, emindlt, delta2, hbr, xkgam, e1)
        print("call 1: max |d| %.3e of max %.3e" % (np.max(np.abs(s1 - r1)), np.max(np.abs(r1))),
              flush=True)
        assert np.array_equal(s1, r1)
        # second call, smaller: the array still holds the first call's tail
        k2 = 900
        a2 = np.concatenate([[1.0], rng.standard_normal(k2) * 0.3])
        e2 = np.linspace(-1.0, 2.0, 97)
        s2 = _noalloc(lib, hip, a2, e2, emindlt, delta2, hbr, xkgam)
        r2 = L.cheby_spectrum(a2, emindlt, delta2, hbr, xkgam, e2)
        assert np.array_equal(s2, r2)
        assert np.all(np.isfinite(s2)) and np.max(np.abs(r2)) > 0
        # above the library's array: refused with a message, nothing launched
        dummy = hip.empty(4)
        rc = lib.mw3_cheby_spectrum_noalloc(dummy, 1 << 17, ct.c_double(emindlt),
                                            ct.c_double(delta2), ct.c_double(hbr),
                                            ct.c_double(xkgam), dummy, 4, dummy, None)
        assert rc == -1
        assert b"exceeds" in lib.mw3_last_error(None)
    finally:
        hip.free()

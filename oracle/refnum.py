"""ctypes access to the reference's own numerics library -- test infrastructure.

oracle/_ref/libliboctdyn.so is SRC/liboctdyn.f of the reference, compiled in
place by oracle/Makefile.  Fortran calling convention: every argument by
reference.  Routines that overwrite an input (sinmom/fftmom overwrite box)
get copies.
"""
import ctypes as ct
import os

import numpy as np

PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "_ref", "libliboctdyn.so")
_L = None


def available():
    return os.path.exists(PATH)


def lib():
    global _L
    if _L is None:
        _L = ct.CDLL(PATH)
    return _L


def _i(v):
    return ct.byref(ct.c_int(int(v)))


def _d(v):
    return ct.byref(ct.c_double(float(v)))


def _p(a):
    return a.ctypes.data_as(ct.POINTER(ct.c_double))


def gauleg(n):
    w = np.zeros(n)
    x = np.zeros(n)
    lib().gauleg_(_p(w), _p(x), _i(n))
    return w, x


def dwigner(jmax2, m2, mp2, beta):
    ndim = max(3 * jmax2, jmax2)
    dj = np.zeros(ndim + 1)
    lib().dwigner_(_p(dj), _i(jmax2), _i(m2), _i(mp2), _d(beta), _i(ndim))
    return dj[: jmax2 + 1]


def noptfft(nin, ntot):
    nout = ct.c_int(0)
    lib().noptfft_(_i(nin), ct.byref(nout), _i(ntot))
    return nout.value


def sinmom(box, npun, npundim, xmred, hbr):
    pr = np.zeros(npundim)
    p2r = np.zeros(npundim)
    lib().sinmom_(_d(box), _i(npun), _i(npundim), _d(xmred), _d(hbr), _p(pr), _p(p2r))
    return pr, p2r


def fftmom(box, npun, npundim, xmred, hbr):
    pr = np.zeros(npundim)
    p2r = np.zeros(npundim)
    lib().fftmom_(_d(box), _i(npun), _i(npundim), _d(xmred), _d(hbr), _p(pr), _p(p2r))
    return pr, p2r


def l2mat(facmass, iom0, iom1, j, Jtot):
    n = iom1 - iom0 + 1
    m = np.zeros((n, n), order="F")
    lib().l2mat_(_p(m), _d(facmass), _i(iom0), _i(iom1), _i(j), _i(Jtot), _i(iom0), _i(iom1))
    return np.asarray(m)


def besph2(l, x):
    F, DF, G, DG = (ct.c_double() for _ in range(4))
    lib().besph2_(ct.byref(F), ct.byref(DF), ct.byref(G), ct.byref(DG),
                  _d(l * (l + 1)), _d(x), _i(-1), _i(0))
    return F.value, DF.value, G.value, DG.value


def rgauscolini(r, rcol, alpha, xk, factor, l=0):
    out = ct.c_double()
    lib().rgauscolini_(ct.byref(out), _d(r), _d(rcol), _d(alpha), _d(xk), _d(factor), _i(l))
    return out.value


def diagon(h):
    n = h.shape[0]
    a = np.array(h, order="F", dtype=float)
    t = np.zeros((n, n), order="F")
    e = np.zeros(n)
    lib().diagon_(_p(a), _i(n), _i(n), _p(t), _p(e))
    return e


def tresj(j1, j2, j3, m1, m2, m3):
    """tresj of the reference (3j symbol); fills its factorial table first."""
    L = lib()
    L.factorial_()
    out = ct.c_double()
    L.tresj_(_i(j1), _i(j2), _i(j3), _i(m1), _i(m2), _i(m3), ct.byref(out))
    return out.value

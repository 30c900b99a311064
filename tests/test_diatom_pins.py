"""The diatomic-function solver (madwave3_amd/diatom.py) against the
reference's own numerics library (SRC/liboctdyn.f compiled into oracle/_ref):
the reference's preparation program (radial_functions01_write,
SRC/mod_baseYfunciones_01y2.f:553-760) finds the bound states as the
eigenvalues of the finite-difference Hamiltonian (TQLI), refines each by
SCHR and carries it to the propagation grid with SPLSET/SPLINQQ.  Here the
same sequence of the reference's routines, called through ctypes, is the
pin; the restatement agrees with it to rounding.  Independent checks: the
Morse oscillator's exact levels."""
import ctypes as ct
import math

import numpy as np
import pytest

from madwave3_amd import diatom
from oracle import refnum

needs_ref = pytest.mark.skipif(not refnum.available(), reason="oracle/_ref not built")
pd = ct.POINTER(ct.c_double)


def _i(v):
    return ct.byref(ct.c_int(int(v)))


def _d(v):
    return ct.byref(ct.c_double(float(v)))


def morse(D=0.17, a=1.03, re=1.4):
    return lambda r: D * (np.exp(-2 * a * (r - re)) - 2 * np.exp(-a * (r - re)))


def reference_states(vdiat, rmin, rmax, xm, r_grid, j, nvmax, npunt=2000):
    """radial_functions01_write's steps with the reference's own routines."""
    L = refnum.lib()
    ah = (rmax - rmin) / (npunt - 1)
    x = np.array([rmin + (i) * ah for i in range(npunt)])
    v = vdiat(x)
    xz = 0.5 / xm
    vv = v + 0.5 * j * (j + 1) / xm / x / x
    alpha = -2.0 - 2.0 * xm * vv * ah * ah
    beta = np.ones(npunt)
    a = np.ascontiguousarray(alpha)
    L.tqli_(a.ctypes.data_as(pd), beta.ctypes.data_as(pd), _i(npunt), _i(npunt))
    ev = np.sort(-xz * a / ah / ah)
    out_e, out_f = [], []
    for iv in range(nvmax + 1):
        if not (v[-1] - ev[iv] > 1e-2):
            continue
        xz1 = 2.0 * xm
        vs = np.ascontiguousarray(vv * xz1)
        p = np.zeros(npunt)
        e2, kv, itry = ct.c_double(), ct.c_int(), ct.c_int()
        L.schr_(_d(ev[iv] * xz1), _d(rmin), _d(rmax), _i(npunt), _i(20), _d(1e-10), ct.byref(e2),
                ct.byref(kv), ct.byref(itry), vs.ctypes.data_as(pd), p.ctypes.data_as(pd),
                _i(npunt))
        F = np.zeros((2, npunt))
        F[0] = p
        X = x.copy()
        L.splset_(F.ctypes.data_as(pd), X.ctypes.data_as(pd), _i(npunt), _i(npunt))
        fd = np.zeros(len(r_grid))
        for i, r in enumerate(r_grid):
            if X[0] <= r <= X[-1]:
                out = ct.c_double()
                L.splinqq_(F.ctypes.data_as(pd), X.ctypes.data_as(pd), _i(2), _i(npunt), _d(r),
                           _i(npunt), ct.byref(out))
                fd[i] = out.value
        fd /= math.sqrt(np.sum(fd * fd))
        out_e.append(e2.value / xz1)
        out_f.append(fd)
    return np.array(out_e), np.array(out_f)


@needs_ref
@pytest.mark.parametrize("j", [0, 3, 12])
def test_radial_functions_match_reference_routines(j):
    xm = 1.0078 * 2.0141 / (1.0078 + 2.0141) * 1822.888486
    rmin, rmax = 0.189, 15.0
    r_grid = rmin + np.arange(128) * (rmax - rmin) / 127
    vd = morse()
    mine = diatom.radial_functions(vd, rmin, rmax, xm, r_grid, [j], 0, 3)[j]
    ref_e, ref_f = reference_states(vd, rmin, rmax, xm, r_grid, j, 3)
    assert len(mine[0]) == len(ref_e) > 0
    assert np.allclose(mine[0], ref_e, rtol=1e-11, atol=0)
    for a, b in zip(mine[1], ref_f):
        assert np.max(np.abs(a - b)) < 1e-9


def test_numerov_levels_of_the_morse_oscillator():
    """E_v = -D + w (v + 1/2) - w^2 (v + 1/2)^2 / (4 D), w = a sqrt(2 D / m)."""
    D, a, re, m = 0.17, 1.03, 1.4, 1200.0
    w = a * math.sqrt(2 * D / m)
    r = np.linspace(0.4, 8.0, 4001)
    h = r[1] - r[0]
    g = 2 * m * morse(D, a, re)(r)
    est = diatom.fd_eigenvalues(g, h)
    for v in range(4):
        e, p, nodes = diatom.cooley_numerov(est[v], g, h, r[-1])
        exact = -D + w * (v + 0.5) - (w * (v + 0.5)) ** 2 / (4 * D)
        assert abs(e / (2 * m) - exact) < 1e-8
        assert nodes == v
        assert abs(h * np.sum(p * p) - 1.0) < 1e-12

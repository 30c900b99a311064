"""Pins of the oracle and of the host set-up (CPU only).

Three kinds of evidence:
  * the reference's own numerics library (SRC/liboctdyn.f compiled from the
    reference tree into oracle/_ref; skipped where it is not built),
  * numbers printed in the reference's golden run log
    (EXAMPLES/HD+H_collision/J000-check/sal.000, fixtures in tests/golden),
  * closed forms and textbook identities.
"""
import json
import math
import os

import numpy as np
import pytest

from madwave3_amd import consts as C
from madwave3_amd import model, numerics
from madwave3_amd import namelist as NL
from oracle import oracle as ORC
from oracle import refnum

GOLD = os.path.join(os.path.dirname(__file__), "golden")
needs_ref = pytest.mark.skipif(not refnum.available(), reason="oracle/_ref not built")


def L():
    return ORC.lib()


# ------------------------------------------------------------------ liboctdyn
@needs_ref
@pytest.mark.parametrize("n", [2, 7, 8, 60, 96, 120, 140])
def test_gauleg_matches_reference_bitwise(n):
    w, x = numerics.gauleg(n)
    wr, xr = refnum.gauleg(n)
    assert np.array_equal(w, wr) and np.array_equal(x, xr)


@pytest.mark.parametrize("n", [8, 60, 140])
def test_gauleg_is_gauss_legendre(n):
    w, x = numerics.gauleg(n)
    xe, we = np.polynomial.legendre.leggauss(n)
    assert np.max(np.abs(x - xe)) < 2e-14 and np.max(np.abs(w - we)) < 2e-14


@needs_ref
@pytest.mark.parametrize("m2,mp2", [(0, 0), (2, 0), (4, 0), (12, 0), (2, 2), (0, 2), (-2, 0), (4, 2)])
@pytest.mark.parametrize("beta", [0.1, 1.3, 2.9])
def test_dwigner_matches_reference(m2, mp2, beta):
    # the recurrence is arranged differently from the reference's: agreement
    # to rounding (values are O(1), 30 recurrence steps)
    a = numerics.dwigner(60, m2, mp2, beta)
    b = refnum.dwigner(60, m2, mp2, beta)
    assert np.max(np.abs(a - b)) <= 2e-14


def test_dwigner_legendre_and_associated_closed_forms():
    from scipy.special import eval_legendre, lpmv
    for beta in (0.3, 1.1, 2.5):
        d = numerics.dwigner(40, 0, 0, beta)
        for j in range(21):
            assert abs(d[2 * j] - eval_legendre(j, math.cos(beta))) < 1e-13
        for m in (1, 2, 5):
            d = numerics.dwigner(40, 2 * m, 0, beta)
            for j in range(m, 21):
                # d^j_{m0}(b) = sqrt((j-m)!/(j+m)!) P_j^m(cos b), with the
                # Condon-Shortley phase inside scipy's lpmv (d^1_10 = -sin b/sqrt 2)
                ref = math.sqrt(math.factorial(j - m) / math.factorial(j + m)) \
                    * lpmv(m, j, math.cos(beta))
                assert abs(d[2 * j] - ref) < 1e-12


def test_dwigner_general_closed_forms_and_orthogonality():
    """Textbook values of d^{1/2} and d^1, the symmetry
    d^j_{m m'} = (-1)^(m-m') d^j_{m' m}, and orthogonality of the rows of
    each d^j matrix (sum over m' of d^j_{m m'} d^j_{m'' m'} = delta)."""
    for b in (0.4, 1.7, 2.8):
        c, s, ch, sh = math.cos(b), math.sin(b), math.cos(b / 2), math.sin(b / 2)
        assert abs(numerics.dwigner(1, 1, 1, b)[1] - ch) < 1e-15
        assert abs(numerics.dwigner(1, 1, -1, b)[1] + sh) < 1e-15
        assert abs(numerics.dwigner(1, -1, 1, b)[1] - sh) < 1e-15
        assert abs(numerics.dwigner(2, 2, 2, b)[2] - 0.5 * (1 + c)) < 1e-15
        assert abs(numerics.dwigner(2, 2, 0, b)[2] + s / math.sqrt(2)) < 1e-15
        assert abs(numerics.dwigner(2, 0, 2, b)[2] - s / math.sqrt(2)) < 1e-15
        assert abs(numerics.dwigner(2, 2, -2, b)[2] - 0.5 * (1 - c)) < 1e-15
        for j2 in (7, 10, 13):
            ms = range(-j2, j2 + 1, 2)
            d = np.array([[numerics.dwigner(j2, m, mp, b)[j2] for mp in ms] for m in ms])
            assert np.max(np.abs(d @ d.T - np.eye(len(ms)))) < 1e-13
            sgn = np.array([[(-1.0) ** ((m - mp) // 2) for mp in ms] for m in ms])
            assert np.max(np.abs(d - sgn * d.T)) < 1e-13


@needs_ref
def test_noptfft_matches_reference_and_brute_force():
    for ntot in (64, 128, 256, 257, 512):
        for nin in range(0, ntot + 3, 3):
            a = L().oref_noptfft(nin, ntot)
            assert a == refnum.noptfft(nin, ntot) == numerics.noptfft(nin, ntot)
            smooth = [i for i in range(max(nin, 1), nin + ntot)
                      if all(p in (2, 3, 5, 7, 11) for p in _pf(i))]
            expect = min(smooth[0], ntot) if smooth else 1
            assert a == expect


def _pf(n):
    f, d = [], 2
    while d * d <= n:
        while n % d == 0:
            f.append(d)
            n //= d
        d += 1
    if n > 1:
        f.append(n)
    return f


@needs_ref
@pytest.mark.parametrize("n", [14, 60, 231, 256, 512])
def test_sinmom_fftmom_match_reference(n):
    ah, xm = 0.0312, 0.6718
    p1 = np.zeros(n)
    p2 = np.zeros(n)
    L().oref_sinmom((n - 1) * ah, n, xm, C.hbr, p1.ctypes.data_as(ORC._pd), p2.ctypes.data_as(ORC._pd))
    pr, p2r = refnum.sinmom((n - 1) * ah, n, n, xm, C.hbr)
    assert np.max(np.abs(p1 - pr)) <= 1e-15 * np.max(np.abs(pr)) + 0
    assert np.max(np.abs(p2 - p2r)) <= 1e-15 * np.max(p2r)
    k = np.arange(1, n + 1)
    assert np.allclose(p1, np.pi * k / ((n + 1) * ah), rtol=1e-14)
    L().oref_fftmom((n - 1) * ah, n, xm, C.hbr, p1.ctypes.data_as(ORC._pd), p2.ctypes.data_as(ORC._pd))
    pr, p2r = refnum.fftmom((n - 1) * ah, n, n, xm, C.hbr)
    assert np.max(np.abs(p1 - pr)) <= 1e-14 * np.max(np.abs(pr))


@needs_ref
@pytest.mark.parametrize("Jtot,j,om0,om1", [(0, 3, 0, 0), (12, 5, 0, 5), (12, 9, 1, 6), (5, 7, 2, 5)])
def test_l2mat_matches_reference_and_eigen_extremes(Jtot, j, om0, om1):
    n = om1 - om0 + 1
    fac = 0.37
    m = np.zeros(n * n)
    L().oref_l2mat(m.ctypes.data_as(ORC._pd), fac, om0, om1, j, Jtot)
    m = m.reshape(n, n)
    r = refnum.l2mat(fac, om0, om1, j, Jtot)
    assert np.max(np.abs(m - r)) < 1e-14
    if Jtot == 0:
        assert abs(m[0, 0] - j * (j + 1) * fac) < 1e-14
    lo, hi = ORC._pd, None
    a, b = np.zeros(1), np.zeros(1)
    L().oref_sym_eig_minmax(np.ascontiguousarray(m).ctypes.data_as(ORC._pd), n,
                            a.ctypes.data_as(ORC._pd), b.ctypes.data_as(ORC._pd))
    ev = np.linalg.eigvalsh(m)
    er = np.sort(refnum.diagon(m))
    assert abs(a[0] - ev[0]) < 1e-12 * abs(ev).max() and abs(b[0] - ev[-1]) < 1e-12 * abs(ev).max()
    assert abs(er[0] - ev[0]) < 1e-10 * abs(ev).max()


@needs_ref
@pytest.mark.parametrize("l", [0, 1, 3, 8, 15])
def test_besph2_matches_reference(l):
    for x in (0.5, 2.0, 7.3, 25.0, 60.0):
        F, dF, G, dG = numerics.besph2(l, np.array(x))
        Fr, dFr, Gr, dGr = refnum.besph2(l, x)
        for a, b in ((F, Fr), (dF, dFr), (G, Gr), (dG, dGr)):
            assert abs(float(a) - b) <= 1e-9 * max(1.0, abs(b))
        assert abs(float(F * dG - dF * G) - 1.0) < 1e-9   # Wronskian


@needs_ref
def test_rgauscolini_matches_reference():
    for r in (4.0, 5.2, 6.1):
        a = numerics.rgauscolini(r, 5.2, 0.414, 14.97, 1.23)
        b = refnum.rgauscolini(r, 5.2, 0.414, 14.97, 1.23)
        assert abs(float(a) - b) <= 1e-15 * max(1, abs(b))


def test_dst1_closed_form_and_involution():
    for n in (1, 5, 14, 31):
        x = np.random.default_rng(n).standard_normal(n)
        y = np.zeros(n)
        z = np.zeros(n)
        L().oref_dst1(n, x.ctypes.data_as(ORC._pd), y.ctypes.data_as(ORC._pd))
        j = np.arange(n)
        ref = np.array([2 * np.sum(x * np.sin(np.pi * (j + 1) * (k + 1) / (n + 1))) for k in range(n)])
        assert np.max(np.abs(y - ref)) < 1e-12
        L().oref_dst1(n, y.ctypes.data_as(ORC._pd), z.ctypes.data_as(ORC._pd))
        assert np.max(np.abs(z - 2 * (n + 1) * x)) < 1e-11 * n


# --------------------------------------------------------- golden example log
def _example_case():
    """Set-up values of EXAMPLES/HD+H_collision (J=0 reference run)."""
    text = """
&INPUTGRIDBASE NPUN1=256, RMIS1=0.1d0, RFIN1=20.d0, NPUN1MIN=64, NPUN2=256,
 RMIS2=0.1d0, RFIN2=20.d0, NANGU=140, NELECMAX=1, JTOT=0, IPARITY=1, IOMMIN=0,
 IOMMAX=0, J0=0, INC=1, JINI=0, JMAX=39, NVINI=0, NVMAX=5, NVREF=0, JREF=0,
 IOMREF=0, IELECREF=1 /
&inputprocess iphoto=0 /
&inputabsorcion absalp1=0.5d0, absr1=7.0d0, n1expo=2, ninterv1=1,
 absalp2=0.5d0, absr2=7.d0, n2expo=2, ninterv2=1 /
&inputpotmass system='H+DH', xm0=2.014101779d0, xm1=1.007825035d0,
 xm2=1.007825035d0, VcutmaxeV=2.5d0, radcutmaxeV=2.5d0, rotcutmaxeV=5.d0 /
&inputcol Rcolini_ang=5.2d0, ecol_ev=0.62d0, deltaE_ev=0.2d0 /
&inputflux r1flux_ang=4.d0, netot=500, ekinmin_eV=0.1d0, ekinmax_eV=1.5d0, ncontfile=0 /
&inputtime ntimes=500, nloop=7, kminelastic=400 /
"""
    nl = NL.parse(text)
    c = model.RefCase()
    model._grid(c, nl)
    model._absorcion(c, nl)
    model._pot0(c, nl)
    model.colini_params(c, nl)
    model.flux_params(c, nl)
    return c


def _pins():
    return json.load(open(os.path.join(GOLD, "hd_h_j000_pins.json")))


def test_golden_angles():
    c = _example_case()
    ang = np.loadtxt(os.path.join(GOLD, "hd_h_j000_angles.txt"))
    deg = np.degrees(np.arccos(c.cgamma))
    assert len(ang) == 140
    assert np.max(np.abs(deg[ang[:, 0].astype(int) - 1] - ang[:, 1])) < 2e-13


def test_golden_absorption_and_flux_indices():
    c = _example_case()
    p = _pins()
    assert (c.ir1abs, c.ir2abs) == (p["ir1abs"], p["ir2abs"])
    ir4 = max(i for i in range(1, c.npun1 + 1) if c.rmis1 + (i - 1) * c.ah1 <= c.r1flux)
    assert ir4 == p["ir4flux"]


def test_golden_kinetic_ranges():
    """E_kin ranges printed by Tradial (SRC/mod_Hphi_01y2.f:601-608).  The R
    grid uses n = npun2; the r extremes come from the largest per-angle DST
    size, which the PES sets: both printed r values must be reproduced by one
    noptFFT-smooth size."""
    c = _example_case()
    p = _pins()
    to_ev = 1.0 / C.conve1 / 8065.5
    p2 = np.zeros(256)
    L().oref_sinmom((256 - 1) * c.ah2, 256, c.xm2red, C.hbr, None, p2.ctypes.data_as(ORC._pd))
    assert abs(p2.max() * to_ev - p["ekin2_max_eV"]) < 1e-13 * p["ekin2_max_eV"]
    assert abs(p2.min() * to_ev - p["ekin2_min_eV"]) < 1e-13 * p["ekin2_min_eV"]
    hits = []
    for n in range(c.npun1min, c.npun1 + 1):
        p1 = np.zeros(n)
        L().oref_sinmom((n - 1) * c.ah1, n, c.xm1red, C.hbr, None, p1.ctypes.data_as(ORC._pd))
        if (abs(p1.max() * to_ev - p["ekin1_max_eV"]) < 1e-12 * p["ekin1_max_eV"]
                and abs(p1.min() * to_ev - p["ekin1_min_eV"]) < 1e-12 * p["ekin1_min_eV"]):
            hits.append(n)
    assert len(hits) == 1
    assert L().oref_noptfft(hits[0], c.npun1) == hits[0]


def test_golden_packet_parameters_and_energy_norm():
    c = _example_case()
    p = _pins()
    assert abs(c.xk0 - p["xk0"]) < 1e-12 * p["xk0"]
    assert abs(c.alpha0 - p["alpha0"]) < 1e-13
    assert abs(c.norm_gauss_R2 - p["norm_gauss_R2"]) < 1e-14
    assert c.il0 == 0
    _, _, norm = model.smatini_norm(c)
    assert abs(norm - p["norm_energy"]) < 1e-13 * p["norm_energy"] * 1e3


# -------------------------------------------------------------- closed forms
def test_coef_echeby_closed_form():
    from tests.cases import case
    c = case("j0")
    o = ORC.Oracle(c)
    E = np.linspace(c.emindlt - 0.9 * c.delta2, c.emindlt + 0.9 * c.delta2, 7)
    for k in (0, 1, 2, 17):
        z = o.coef_echeby(k, E)
        Es = (E - c.emindlt) / c.delta2
        d = 1.0 if k == 0 else 2.0
        th = np.arccos(Es)
        assert np.allclose(z.real, d * C.hbr / c.delta2 * np.cos(k * th) / np.sin(th), rtol=1e-12)
        assert np.allclose(z.imag, -d * C.hbr / c.delta2 * np.sin(k * th) / np.sin(th), rtol=1e-12,
                           atol=1e-300)


def test_zar1flux_is_half_the_fourier_derivative():
    """zAR1flux (SRC/mod_flux_01y2.f:133-144) applied to a band-limited
    periodic function gives half its derivative at r(ir4flux)."""
    from tests.cases import case
    c = case("mid")
    t = ORC.Oracle(c).tables()
    za, ir4 = t["zAR1flux"], t["ir4flux"]
    n, ah = c.npun1, c.ah1
    r = c.rmis1 + np.arange(n) * ah
    Lbox = n * ah
    for m0 in (1, 3, 7):
        f = np.sin(2 * np.pi * m0 * (r - 0.37) / Lbox)
        df = 2 * np.pi * m0 / Lbox * np.cos(2 * np.pi * m0 * (r[ir4 - 1] - 0.37) / Lbox)
        s = np.sum(za * f)
        assert abs(s.real - 0.5 * df) < 1e-10 * abs(df) + 1e-12 and abs(s.imag) < 1e-10 * abs(df) + 1e-12


def test_cheby_spectrum_closed_form():
    """cheby-spectra (SRC/photo/cheby-spectra.f:71-90) of auto(k) = cos(k t0)
    (one eigenstate at Es0 = cos t0): with r = e^-xkgam and kmax large enough
    for r^kmax to vanish, spec * sqrt(1-Es^2) delta2 / hbr is the sum
    sum_k (2 - d_k0) r^k cos(k t) cos(k t0) = [P(t - t0) + P(t + t0)] / 2 with the
    Poisson kernel P(a) = (1 - r^2) / (1 - 2 r cos a + r^2)."""
    from oracle.oracle import cheby_spectrum

    emindlt, delta2, hbr, xkgam = 0.7, 2.3, 0.9, 0.01
    kmax = 6000
    t0 = 1.1
    autoc = np.cos(np.arange(kmax + 1) * t0)
    E = emindlt + delta2 * np.linspace(-0.99, 0.99, 397)
    spec = cheby_spectrum(autoc, emindlt, delta2, hbr, xkgam, E)
    Es = (E - emindlt) / delta2
    t = np.arccos(Es)
    r = np.exp(-xkgam)
    P = lambda a: (1 - r * r) / (1 - 2 * r * np.cos(a) + r * r)
    want = 0.5 * (P(t - t0) + P(t + t0)) * hbr / (np.sqrt(1 - Es * Es) * delta2)
    assert np.allclose(spec, want, rtol=1e-9, atol=1e-12 * np.max(np.abs(want)))
    # |Es| >= 1 gives NaN in the sum, reported as 0
    out = cheby_spectrum(autoc[:10], emindlt, delta2, hbr, xkgam, np.array([emindlt + 1.5 * delta2]))
    assert out[0] == 0.0

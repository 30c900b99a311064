"""Pins of the product-state projection (iprod = 2): the host tables of
madwave3_amd.prodtrans (ini_transcoor / changeini / bfreac2prod,
SRC/mod_coortrans01_02.f:31-690) and the oracle's prodpaq / prodcvj
(SRC/mod_coortrans01_02.f:694-822).

What fixes each piece:
  spline (SPLSET/SPLINQQ)  the reference's own build of SRC/liboctdyn.f, and
                           exactness on cubics
  geometry                 the rp range the reference prints for the HD+H
                           J=0 example (J000-check/sal.000:907) and Cartesian
                           positions of the three atoms
  transRgjac               exact sine-basis interpolation of band-limited
                           functions on a grid with rmis1 = ah1
  dtrans                   unitarity of the parity-adapted rotation over
                           Omega' = 0..J
  prodwf                   the product-state norm printed by changeini (0.9994
                           for v'=0 in J000-check/sal.000:909): close to 1
  prodpaq / prodcvj        a dense numpy restatement on the same state, and the
                           Chebyshev sum of the per-step amplitudes
"""
import json
import math
import os
import types

import numpy as np
import pytest

from madwave3_amd import consts as C
from madwave3_amd import leps, model, prodtrans
from madwave3_amd import namelist as NL
from madwave3_amd.numerics import dwigner, dwigner_vec, gauleg, splinqq, splset
from oracle import refnum
from oracle.oracle import Oracle
from tests.cases import case

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(not refnum.available(), reason="oracle/_ref not built")
def test_spline_matches_reference_build():
    import ctypes as ct
    L = refnum.lib()
    nx = 120
    x = 0.1 + np.arange(nx) * 0.05
    f = np.exp(-(x - 3.0) ** 2) * np.cos(2.0 * x)
    ndim = nx + 3
    F = np.zeros((2, ndim))
    F[0, :nx] = f
    X = np.zeros(ndim)
    X[:nx] = x
    pd = ct.POINTER(ct.c_double)
    L.splset_(F.ctypes.data_as(pd), X.ctypes.data_as(pd), ct.byref(ct.c_int(nx)), ct.byref(ct.c_int(ndim)))
    d = splset(f, x)
    # the reference build may contract products into FMAs: agreement to rounding
    assert np.allclose(d, F[1, :nx], rtol=1e-13, atol=0)
    rs = np.concatenate([np.linspace(x[0], x[-1], 37)[:-1] + 0.013, [x[0], x[-1], x[-1] + 0.4]])
    mine = splinqq(f, d, x, rs)
    for r, m in zip(rs, mine):
        out = ct.c_double()
        L.splinqq_(F.ctypes.data_as(pd), X.ctypes.data_as(pd), ct.byref(ct.c_int(2)),
                   ct.byref(ct.c_int(nx)), ct.byref(ct.c_double(r)), ct.byref(ct.c_int(ndim)),
                   ct.byref(out))
        assert m == pytest.approx(out.value, rel=1e-13, abs=1e-300)


def test_spline_reproduces_cubics():
    """The end slopes are 4-point Lagrange derivatives, so the clamped spline
    of a cubic is the cubic itself."""
    x = np.sort(np.random.default_rng(3).uniform(0.0, 5.0, 40))
    f = 0.3 * x ** 3 - x ** 2 + 2.0 * x - 1.0
    d = splset(f, x)
    assert np.allclose(d, 0.9 * x ** 2 - 2.0 * x + 2.0, rtol=0, atol=1e-10)
    r = np.linspace(x[0], x[-1] - 1e-9, 200)
    assert np.allclose(splinqq(f, d, x, r), 0.3 * r ** 3 - r ** 2 + 2.0 * r - 1.0, atol=1e-11)


def test_dwigner_vec_equals_scalar():
    b = np.random.default_rng(5).uniform(0.0, math.pi, 30)
    for m2, mp2 in [(0, 0), (2, 0), (0, 4), (4, -2), (-6, 2), (2, 2)]:
        v = dwigner_vec(30, m2, mp2, b)
        s = np.array([dwigner(30, m2, mp2, x) for x in b]).T
        assert np.allclose(v, s, rtol=0, atol=1e-14)


def _hdh_geometry_case():
    """Grid, masses and product surface of EXAMPLES/HD+H_collision/J000
    (input echoed in J000-check/sal.000:19-58, 296-298)."""
    w, x = gauleg(140)
    return types.SimpleNamespace(
        xm0=2.014101779, xm1=1.007825035, xm2=1.007825035, Rbalinprod=6.0,
        rmis2=0.1, ah2=(20.0 - 0.1) / 255.0, n2prod0=1, n2prod1=256,
        cgamma=x, nangproj0=1, nangproj1=140)


def test_product_geometry_matches_reference_rp_range():
    """changeini prints the range of the product bond length rpb sampled on
    the projection grid (SRC/mod_coortrans01_02.f:378-407)."""
    pins = json.load(open(os.path.join(ROOT, "tests", "golden", "hd_h_j000_pins.json")))
    c = _hdh_geometry_case()
    R = c.rmis2 + np.arange(c.n2prod0 - 1, c.n2prod1) * c.ah2
    g = prodtrans.product_geometry(c, R[None, :], np.asarray(c.cgamma)[:, None])
    rpb = g["rpb"][g["ok_sqr"]]
    assert abs(rpb.min() - pins["rpbmin"]) <= 1e-13 * pins["rpbmin"]
    assert abs(rpb.max() - pins["rpbmax"]) <= 1e-13 * pins["rpbmax"]


def test_product_geometry_cartesian():
    """At the returned r, atom 1 is R' = Rbalinprod from the 0-2 centre of
    mass, rpb = |x2 - x0|, angjac = cos(r', R'), xjaco = dr/dR', and the
    frame angle of bfreac2prod is the angle between the reactant R (0-1 centre
    of mass -> 2) and the product R' (0-2 centre of mass -> 1)."""
    c = _hdh_geometry_case()
    rng = np.random.default_rng(11)
    R = rng.uniform(0.5, 12.0, 200)
    cg = rng.uniform(-1.0, 1.0, 200)
    g = prodtrans.product_geometry(c, R, cg)
    ok = g["ok_sqr"]
    m01, m02 = c.xm0 + c.xm1, c.xm0 + c.xm2
    r = g["rpa"]
    u = np.stack([np.ones_like(cg), np.zeros_like(cg)], -1)                 # r direction
    v = np.stack([cg, np.sqrt(1.0 - cg * cg)], -1)                          # R direction
    x0 = -(c.xm1 / m01) * r[:, None] * u
    x1 = (c.xm0 / m01) * r[:, None] * u
    x2 = R[:, None] * v
    com02 = (c.xm0 * x0 + c.xm2 * x2) / m02
    Rp = x1 - com02
    rp = x2 - x0
    assert np.allclose(np.linalg.norm(Rp, axis=1)[ok], c.Rbalinprod, rtol=1e-12)
    assert np.allclose(np.linalg.norm(rp, axis=1)[ok], g["rpb"][ok], rtol=1e-12)
    cos_p = np.sum(rp * Rp, 1) / np.linalg.norm(rp, axis=1) / np.linalg.norm(Rp, axis=1)
    assert np.allclose(cos_p[ok], g["angjac"][ok], atol=1e-12)
    h = 1e-5
    cp, cm = types.SimpleNamespace(**vars(c)), types.SimpleNamespace(**vars(c))
    cp.Rbalinprod += h
    cm.Rbalinprod -= h
    drdR = (prodtrans.product_geometry(cp, R, cg)["rpa"] - prodtrans.product_geometry(cm, R, cg)["rpa"]) / (2 * h)
    assert np.allclose(g["xjaco"][ok], drdR[ok], rtol=1e-7)
    # bfreac2prod's frame angle
    B = g["B"]
    ang = np.arccos(np.clip(g["angjac"], -1, 1))
    deno = (B * g["rpb"]) ** 2 + c.Rbalinprod ** 2 - 2.0 * B * g["rpb"] * c.Rbalinprod * np.cos(ang)
    ctrans = (B * g["rpb"] * np.cos(ang) - c.Rbalinprod) / np.sqrt(deno)
    cos_rr = np.sum(v * Rp, 1) / np.linalg.norm(Rp, axis=1)
    assert np.allclose(ctrans[ok], cos_rr[ok], atol=1e-12)


def _full_grid_case(n1=24, n2=20, na=6):
    """Every (r, R) point kept at every angle, one channel, r grid with
    rmis1 = ah1 so that sin(x1p_m r1) is the orthogonal sine basis."""
    ah1 = 0.2
    w, x = gauleg(na)
    c = types.SimpleNamespace(
        npun1=n1, rmis1=ah1, ah1=ah1, rfin1=ah1 * n1, npun2=n2, rmis2=0.3, ah2=0.5,
        nangu=na, cgamma=x, xm0=2.0141, xm1=1.00783, xm2=1.00783, Rbalinprod=1.5,
        nangproj0=1, nangproj1=na, ncan=1)
    c.indangreal = np.arange(na, 0, -1, dtype=np.int32)
    c.npunreal = np.full(na, n1 * n2, dtype=np.int32)
    i2, i1 = np.meshgrid(np.arange(1, n2 + 1), np.arange(1, n1 + 1), indexing="ij")
    c.indr1 = np.tile(i1.ravel(), (na, 1)).astype(np.int32)
    c.indr2 = np.tile(i2.ravel(), (na, 1)).astype(np.int32)
    c.off = (np.arange(na) * n1 * n2).reshape(na, 1)
    c.ntot = na * n1 * n2
    return c


def test_transrgjac_interpolates_sine_modes_exactly():
    """sum_r transRgjac psi for psi(r) = sin(x1p_p r) is sin(x1p_p r1inter)
    sqrt(dr/dR') / sqrt(ah1) wherever r1inter lies inside the r box."""
    c = _full_grid_case()
    T, xt1 = prodtrans.transrgjac(c)
    assert xt1 == pytest.approx(float(np.sum(T * T)))
    xk = math.pi * np.arange(1, c.npun1 + 1) / ((c.npun1 + 1) * c.ah1)
    R = c.rmis2 + np.arange(c.npun2) * c.ah2
    checked = 0
    for s in range(c.nangu):
        ia = int(c.indangreal[s])
        o = int(c.off[s, 0])
        r1 = c.rmis1 + (c.indr1[ia - 1] - 1) * c.ah1
        ir2 = c.indr2[ia - 1] - 1
        g = prodtrans.product_geometry(c, R, c.cgamma[ia - 1])
        inside = g["ok_sqr"] & (g["rpa"] > c.rmis1) & (g["rpa"] < c.rfin1)
        for p in (1, 3, 7):
            psi = np.sin(xk[p - 1] * r1)
            fun = np.bincount(ir2, weights=T[o:o + c.npun1 * c.npun2] * psi, minlength=c.npun2)
            want = np.sin(xk[p - 1] * g["rpa"]) * np.sqrt(g["xjaco"]) / math.sqrt(c.ah1)
            assert np.allclose(fun[inside], want[inside], rtol=0, atol=1e-12)
            assert np.all(fun[~inside] == 0.0)
            checked += int(inside.sum())
    assert checked > 20


def test_dtrans_unitary_over_all_product_helicities():
    """With Omega' = 0..J every reactant channel's parity-adapted rotation is a
    unit vector (bfreac2prod, SRC/mod_coortrans01_02.f:655-676)."""
    c = case("prodJ")
    assert c.iomminprod == 0 and c.iommaxprod == c.Jtot
    d = c.dtrans                                     # [ncan][nomb][nap][n2p]
    s = np.sum(d * d, axis=1)
    used = np.abs(d).sum(axis=1) > 0
    assert used.sum() > 50
    assert np.allclose(s[used], 1.0, atol=1e-13)


def test_product_state_norms_near_one():
    """changeini's check: the product states mapped to (R, gamma) keep their
    norm (the reference prints 0.9994 for v'=0, j'=0 of HD+H); configuration
    1's HD(v'=0) states lie inside the grid and keep it within 2%."""
    p = leps.config(1)
    text, pot, fdat = leps.make_inputs(p)
    c = model.build_case(NL.parse(text), pot, fdat)
    for j in (0, 5, 10):
        assert 0.98 < c.prod_norms[(0, j, 0)] < 1.0 + 1e-9


@pytest.mark.parametrize("name", ["prod", "prodJ", "prodrange"])
def test_oracle_prodpaq_matches_dense_restatement(name):
    """Oracle Cvjprod after each step against numpy on the dense grid:
    funprod = sum_r T psi, times dtrans, contracted with prodwf over (R,
    angle) per channel and summed; zS2prod = sum_k C_k(E) Cvjprod_k."""
    c = case(name)
    o = Oracle(c)
    o.begin(c.rpaq0)
    sl, ch, ir, ir1, ir2 = model.point_index(c)
    ia = c.indangreal[sl] - 1

    def dense(v):
        d = np.zeros((c.ncan, c.nangu, c.npun1, c.npun2))
        d[ch, ia, ir1 - 1, ir2 - 1] = v
        return d

    T = dense(c.transRgjac)
    W = c.prodwf                                     # [omb][j][v][e][a][r]
    z = np.zeros(W.shape[:3] + (c.netot,), complex)
    for _ in range(5):
        o.steps(1)
        _, psi, k = o.state()
        fun = np.einsum("caxr,caxr->car", T, dense(psi))
        fun = fun[:, c.nangproj0 - 1:c.nangproj1, c.n2prod0 - 1:c.n2prod1]
        cv = np.zeros(W.shape[:3])
        for ic in range(c.ncan):
            aux2 = fun[ic][None] * c.dtrans[ic]       # [omb][a][r]
            cv += np.einsum("ojvar,oar->ojv", W[:, :, :, c.nelebas[ic] - 1], aux2)
        ref = o.cvjprod()
        assert np.max(np.abs(ref - cv)) <= 1e-12 * np.max(np.abs(cv))
        z += cv[..., None] * o.coef_echeby(k, c.etotS2)[None, None, None, :]
    assert np.max(np.abs(o.zs2prod() - z)) <= 1e-12 * np.max(np.abs(z))


def test_prod_tables_keep_reference_layout():
    c = case("prodJ")
    nomb = c.iommaxprod - c.iomminprod + 1
    nj, nv = c.jmaxprod - c.jiniprod + 1, c.nvmaxprod - c.nviniprod + 1
    nap, n2p = c.nangproj1 - c.nangproj0 + 1, c.n2prod1 - c.n2prod0 + 1
    assert c.prodwf.shape == (nomb, nj, nv, c.nelecmax, nap, n2p)
    assert c.dtrans.shape == (c.ncan, nomb, nap, n2p)
    assert c.transRgjac.shape == (c.ntot,)
    # states the reference skips (j' < Omega') carry no function
    for io in range(nomb):
        for j in range(c.jiniprod, min(io, c.jmaxprod + 1)):
            assert not np.any(c.prodwf[io, j - c.jiniprod])
    assert C.hbr > 0

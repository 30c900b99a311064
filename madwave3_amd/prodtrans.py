"""Host set-up of the product-state projection (iprod = 2) at nproc = 1.

Restates the set-up half of module mod_coortrans01_02, which builds the tables
the per-step projection prodpaq/prodcvj reads:

  ini_transcoor   SRC/mod_coortrans01_02.f:31-209  S2prodfac and the tables
  changeini       SRC/mod_coortrans01_02.f:213-482 transRgjac, prodwf, norms
  prodwftrans     SRC/mod_coortrans01_02.f:484-588 one product state in
                                                    (R_reac, gamma_reac)
  bfreac2prod     SRC/mod_coortrans01_02.f:592-690 dtrans

Geometry (atoms 0, 1, 2; reactant diatom 0-1, product diatom 0-2; m01 =
xm0 + xm1, m02 = xm0 + xm2): the reactant Jacobi point (r, R, gamma) whose
product Jacobi distance is R' = Rbalinprod has
    A r = R cos(gamma) + sqrt(R^2 (cos^2 gamma - 1) + (m02 R'/xm2)^2),
    A = m02/xm2 - xm1/m01,
the product bond rpb = |x2 - x0| and the product Jacobi angle
cos gamma' = (R'^2 + B^2 rpb^2 - (m01 R/xm1)^2) / (2 B rpb R'),
B = m01/xm1 - xm2/m02.  transRgjac interpolates psi(r, R, gamma) in r to that
point with the sine basis of the r grid, times sqrt(dr/dR'); prodwf is the
product state evaluated there with the volume-element factor R r/(rpb R').

Arrays keep the reference's Fortran memory order (a Fortran A(n1,n2) is the
numpy array [n2][n1]), so they go to the C ABI unchanged:
  transRgjac  [ntot]                                  A0 layout
  prodwf      [nomb][nj][nv][nelec][nangp][n2p]       prodwf(n2p,nangp,nelec,nv,nj,nomb)
  dtrans      [ncan][nomb][nangp][n2p]                dtrans(n2p,nangp,nomb,ncan)
  S2prodfac   [nj][nv][netot]                         S2prodfac(netot,nv,nj)
"""
import math

import numpy as np

from . import consts as C
from .numerics import dwigner_vec, splinqq, splset

__all__ = ["ini_transcoor", "transrgjac", "product_geometry", "prodwf_tables", "bfreac2prod",
           "s2prodfac"]


def _grid2(c, lo, hi):
    return c.rmis2 + np.arange(lo - 1, hi) * c.ah2


def product_geometry(c, R, cg):
    """Product-coordinate quantities of prodwftrans / bfreac2prod at reactant
    (R, cos gamma) on the dividing surface R' = Rbalinprod (arrays broadcast):
    returns dict with rpa (r), rpb, angjac (cos gamma'), xjaco (dr/dR'),
    ok_sqr (the square root is real)."""
    xm0, xm1, xm2 = c.xm0, c.xm1, c.xm2
    Rgb = c.Rbalinprod
    A = (xm0 + xm2) / xm2 - xm1 / (xm0 + xm1)
    B = (xm0 + xm1) / xm1 - xm2 / (xm0 + xm2)
    R = np.asarray(R, dtype=float)
    cg = np.asarray(cg, dtype=float)
    sqr = R * R * (cg * cg - 1.0) + (Rgb * (xm0 + xm2) / xm2) ** 2
    ok = sqr > 0.0
    ssq = np.sqrt(np.where(ok, sqr, 1.0))
    rpa = (R * cg + ssq) / A
    rpb = (rpa * xm1 / (xm0 + xm1)) ** 2 + R * R + 2.0 * rpa * R * cg * xm1 / (xm0 + xm1)
    rpb = np.sqrt(np.where(ok, rpb, 1.0))
    angjac = Rgb * Rgb + B * B * rpb * rpb - (R * (xm0 + xm1) / xm1) ** 2
    angjac = angjac * 0.5 / (B * rpb * Rgb)
    xjac1 = (R ** 2) * (cg ** 2 - 1.0) + (Rgb * (xm0 + xm2) / xm2) ** 2
    xjaco = Rgb * (((xm0 + xm2) / xm2) ** 2) / (A * np.sqrt(np.where(ok, xjac1, 1.0)))
    return dict(rpa=rpa, rpb=rpb, angjac=angjac, xjaco=xjaco, ok_sqr=ok, A=A, B=B)


def _sin_momenta(c):
    """x1p of sinmom(box1, npun1, ...) (SRC/liboctdyn.f:1004-1031) with box1 =
    (npun1 - 1) ah1 (SRC/mod_coortrans01_02.f:248-250)."""
    n = c.npun1
    box1 = float(n - 1) * c.ah1
    ah = box1 / float(n - 1)
    box = 2.0 * (float(n) + 1.0) * ah
    dpi = 2.0 * math.acos(-1.0)
    return dpi * np.arange(1, n + 1, dtype=float) / box


def transrgjac(c):
    """transRgjac (SRC/mod_coortrans01_02.f:303-370): for the angles
    nangproj0..nangproj1 every kept point (ir1, ir2) of every channel gets
    raux(ir2, ir1) = (2/(npun1+1)) sum_m sin(x1p_m r1) sin(x1p_m r1inter)
    sqrt(dr/dR') / sqrt(ah1), where r1inter(R, gamma) is the r of the dividing
    surface R' = Rbalinprod (0 when it falls outside (rmis1, rfin1)).
    Returns (transRgjac [ntot], xt1 = sum of squares)."""
    xm0, xm1, xm2 = c.xm0, c.xm1, c.xm2
    A = (xm0 + xm2) / xm2 - xm1 / (xm0 + xm1)
    Rgprod = c.Rbalinprod
    xk = _sin_momenta(c)
    r1 = c.rmis1 + np.arange(c.npun1) * c.ah1
    S1 = np.sin(np.outer(r1, xk))                       # [ir1][m]
    r2 = c.rmis2 + np.arange(c.npun2) * c.ah2
    out = np.zeros(c.ntot)
    for s in range(c.nangu):
        iang = int(c.indangreal[s])
        if iang < c.nangproj0 or iang > c.nangproj1:
            continue
        cg = float(c.cgamma[iang - 1])
        xxx1 = ((xm0 + xm2) * Rgprod / xm2) ** 2
        xxx2 = r2 * r2 * (cg * cg - 1.0)
        pos = (xxx1 + xxx2) > 0.0
        r1inter = (r2 * cg + np.sqrt(np.where(pos, xxx1 + xxx2, 0.0))) / A
        use = pos & (r1inter > c.rmis1) & (r1inter < c.rfin1)
        raux = np.zeros((c.npun2, c.npun1))               # raux(ir2, ir1) as [ir2][ir1]
        if use.any():
            S2 = np.sin(np.outer(xk, r1inter[use]))       # [m][ir2 used]
            ssum = (S1 @ S2).T * 2.0 / float(c.npun1 + 1)  # [ir2 used][ir1]
            x1 = np.sqrt(r2[use] * r2[use] * (cg * cg - 1.0) + (Rgprod * (xm0 + xm2) / xm2) ** 2)
            xj = Rgprod * (((xm0 + xm2) / xm2) ** 2)
            xj = np.sqrt(xj / (A * x1))
            raux[use] = ssum * xj[:, None] / math.sqrt(c.ah1)
        n = int(c.npunreal[iang - 1])
        i1 = c.indr1[iang - 1, :n] - 1
        i2 = c.indr2[iang - 1, :n] - 1
        vals = raux[i2, i1]
        for ic in range(c.ncan):
            o = int(c.off[s, ic])
            out[o:o + n] = vals
    return out, float(np.sum(out * out))


def prodwf_tables(c):
    """prodwf (SRC/mod_coortrans01_02.f:398-423 with prodwftrans 484-588):
    for every Omega' = iomminprod..iommaxprod, j' = max(jiniprod, Omega')..
    jmaxprod, v' = nviniprod..noBCprod(j') and electronic state e:
      prodwf(R, gamma) = fdprod_spline(rpb) d^{j'}_{m,m'}(gamma')
                         sqrt((2j'+1)/2) sqrt(w_gamma) facnorm R r sqrt(dr/dR')
                         / (rpb R'),
    m = Omega' - iomatom(e), m' = iomdiat(e); 0 where the geometry has no
    solution or |cos gamma'| > 1.  Returns (prodwf, norms) with
    norms[(iom, j, v)] the sum of squares the reference prints, and the range
    (rpbmin, rpbmax) of rpb it prints."""
    n2p = c.n2prod1 - c.n2prod0 + 1
    nap = c.nangproj1 - c.nangproj0 + 1
    nomb = c.iommaxprod - c.iomminprod + 1
    nj = c.jmaxprod - c.jiniprod + 1
    nv = c.nvmaxprod - c.nviniprod + 1
    lmax = (c.nangu - 1) * c.inc
    if c.jmaxprod > lmax:
        raise ValueError("jmaxprod = %d > (nangu-1)*inc = %d" % (c.jmaxprod, lmax))
    Rga = _grid2(c, c.n2prod0, c.n2prod1)[None, :]                     # [1][n2p]
    cgam = np.asarray(c.cgamma[c.nangproj0 - 1:c.nangproj1])[:, None]  # [nap][1]
    g = product_geometry(c, Rga, cgam)
    rpa, rpb, angjac, xjaco = g["rpa"], g["rpb"], g["angjac"], g["xjaco"]
    ok = g["ok_sqr"]
    rpbmin = float(rpb[ok].min()) if ok.any() else 1e10
    rpbmax = float(rpb[ok].max()) if ok.any() else 0.0
    inrange = ok & (rpb >= c.rmis2) & (rpb <= c.rfin2)
    angok = ok & (np.abs(angjac) <= 1.0)
    angb = np.arccos(np.where(angok, angjac, 1.0))
    facnorm = 1.0 / math.sqrt(2.0) if c.inc == 2 else 1.0
    wsq = np.sqrt(np.asarray(c.weiprod[c.nangproj0 - 1:c.nangproj1]))[:, None]
    xx = c.rmis2 + np.arange(c.n2prod1) * c.ah2
    rb = rpb[inrange]
    prodwf = np.zeros((nomb, nj, nv, c.nelecmax, nap, n2p))
    norms = {}
    nelec = c.nelecmax
    for iom in range(c.iomminprod, c.iommaxprod + 1):
        # angular factor per electronic state and j'
        ang = {}
        for ie in range(1, nelec + 1):
            m2 = 2 * (iom - int(c.iomatom[ie - 1]))
            mp2 = 2 * int(c.iomdiat[ie - 1])
            d = dwigner_vec(2 * c.jmaxprod, m2, mp2, angb[angok])
            ang[ie] = d
        jinip = max(c.jiniprod, iom)
        for j in range(jinip, c.jmaxprod + 1):
            xn = math.sqrt((2.0 * j + 1.0) * 0.5)
            for iv in range(c.nviniprod, int(c.noBCprod[j - c.jiniprod]) + 1):
                tot = 0.0
                for ie in range(1, nelec + 1):
                    f = c.fdprod[ie - 1, j - c.jiniprod, iv - c.nviniprod, :c.n2prod1]
                    dsl = splset(f, xx)
                    fdi = np.zeros_like(rpb)
                    fdi[inrange] = splinqq(f, dsl, xx, rb)
                    angpart = np.zeros_like(rpb)
                    angpart[angok] = ang[ie][2 * j] * xn
                    angpart = angpart * wsq * facnorm
                    xxx = fdi * angpart * Rga * rpa * np.sqrt(xjaco) / (rpb * c.Rbalinprod)
                    xxx = np.where(angok, xxx, 0.0)
                    prodwf[iom - c.iomminprod, j - c.jiniprod, iv - c.nviniprod, ie - 1] = xxx
                    tot += float(np.sum(xxx * xxx))
                norms[(iom, j, iv)] = tot
    return prodwf, norms, (rpbmin, rpbmax)


def bfreac2prod(c):
    """dtrans (SRC/mod_coortrans01_02.f:592-690): body-fixed frame rotation
    from reactant Omega (channel) to product Omega' at each (R, gamma) of the
    dividing surface.  1 at Jtot = 0; otherwise, with beta the angle between
    the reactant and product R vectors,
      d^J_{Omega,Omega'}(beta), parity-combined with
      iparity (-1)^{J+Omega} d^J_{-Omega,Omega'}(beta) over sqrt(2) when
      Omega > 0, times sqrt(2) when Omega' > 0."""
    if c.iommaxprod > c.Jtot:
        raise ValueError("iommaxprod = %d > Jtot = %d" % (c.iommaxprod, c.Jtot))
    n2p = c.n2prod1 - c.n2prod0 + 1
    nap = c.nangproj1 - c.nangproj0 + 1
    nomb = c.iommaxprod - c.iomminprod + 1
    dtrans = np.zeros((c.ncan, nomb, nap, n2p))
    if c.Jtot == 0:
        dtrans[:] = 1.0
        return dtrans
    Rga = _grid2(c, c.n2prod0, c.n2prod1)[None, :]
    cgam = np.asarray(c.cgamma[c.nangproj0 - 1:c.nangproj1])[:, None]
    g = product_geometry(c, Rga, cgam)
    rpb, angjac, B = g["rpb"], g["angjac"], g["B"]
    Rgb = c.Rbalinprod
    ok = g["ok_sqr"] & (np.abs(angjac) <= 1.0)
    angb = np.arccos(np.where(ok, angjac, 1.0))
    deno = (B * rpb) ** 2 + Rgb * Rgb - 2.0 * B * rpb * Rgb * np.cos(angb)
    ctrans = (B * rpb * np.cos(angb) - Rgb) / np.sqrt(np.where(ok, deno, 1.0))
    ok = ok & (np.abs(ctrans) <= 1.0)
    beta = np.arccos(np.where(ok, ctrans, 1.0))[ok]
    jt2 = 2 * c.Jtot
    for ican in range(c.ncan):
        iomreac = int(c.iombas[ican])
        for iomprod in range(c.iomminprod, c.iommaxprod + 1):
            m2, mp2 = 2 * iomreac, 2 * iomprod
            x1 = dwigner_vec(jt2, m2, mp2, beta)[jt2]
            if iomreac != 0:
                ss = float(c.iparity) * ((-1.0) ** abs(c.Jtot + iomreac))
                x1 = x1 + ss * dwigner_vec(jt2, -m2, mp2, beta)[jt2]
                x1 = x1 / math.sqrt(2.0)
            if iomprod != 0:
                x1 = x1 * math.sqrt(2.0)
            blk = np.zeros((nap, n2p))
            blk[ok] = x1
            dtrans[ican, iomprod - c.iomminprod] = blk
    return dtrans


def s2prodfac(c):
    """S2prodfac(netot, nv', j') (SRC/mod_coortrans01_02.f:119-165):
    iphoto = 0: hbr^2 pfin pini / (2 paqini) / xmasa0 / xm2prod (0 where
    paqini < 1e-10); iphoto > 0: 4 pfin / (pi xm2prod); pfin from the product
    kinetic energy E - ediatprod(v', j')."""
    nj = c.jmaxprod - c.jiniprod + 1
    nv = c.nvmaxprod - c.nviniprod + 1
    fac = np.zeros((nj, nv, c.netot))
    for ie in range(c.netot):
        if c.iphoto == 0:
            ekinini = c.etotS2[ie]
            pini = math.sqrt(ekinini * c.xmasa0 * 2.0 / C.hbr / C.hbr) if ekinini > 0 else 0.0
            paqini = c.apaqini[ie]
        for j in range(c.jiniprod, c.jmaxprod + 1):
            for iv in range(c.nviniprod, int(c.noBCprod[j - c.jiniprod]) + 1):
                ekinfin = c.etotS2[ie] - c.ediatprod[j - c.jiniprod, iv - c.nviniprod]
                pfin = math.sqrt(ekinfin * c.xm2prod * 2.0 / C.hbr / C.hbr) if ekinfin > 0 else 0.0
                if c.iphoto == 0:
                    v = 0.0 if paqini < 1e-10 else (
                        C.hbr * C.hbr * pfin * pini / (2.0 * paqini) / c.xmasa0 / c.xm2prod)
                else:
                    v = 4.0 * pfin / (C.pi * c.xm2prod)
                fac[j - c.jiniprod, iv - c.nviniprod, ie] = v
    return fac


def ini_transcoor(c):
    """All product-projection tables of a case (ini_transcoor + changeini +
    bfreac2prod), stored on the case: transRgjac, prodwf, dtrans, S2prodfac,
    and the diagnostics the reference prints (prod_xt1, prod_rpb_range,
    prod_norms)."""
    c.S2prodfac = s2prodfac(c)
    c.transRgjac, c.prod_xt1 = transrgjac(c)
    c.prodwf, c.prod_norms, c.prod_rpb_range = prodwf_tables(c)
    c.dtrans = bfreac2prod(c)

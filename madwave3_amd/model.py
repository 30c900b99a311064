"""Host-side set-up of a MADWAVE3 propagation at nproc = 1.

Builds, from input.dat namelists and the pot/ and func/ data, the module state
the reference's set-up routines leave behind for the hot path, with the
reference's names:

  input_grid            SRC/mod_gridYpara_01y2.f:85-324
  ini_absorcion         SRC/mod_absorcion_01y2.f:28-107
  pot0                  SRC/mod_pot_01y2.f:56-144
  basis                 SRC/mod_baseYfunciones_01y2.f:50-239
  paralelizacion        SRC/mod_baseYfunciones_01y2.f:242-452 (nproc = 1)
  angular_functions     SRC/mod_baseYfunciones_01y2.f:454-551
  radial_functions01_read  SRC/mod_baseYfunciones_01y2.f:837-923
  pot2                  SRC/mod_pot_01y2.f:344-533
  set_colini, colini_reacJacobi  SRC/mod_colini_01y2.f:25-177
  ini_flux, Smatini, Smatini_col, Smatini_photo
                        SRC/mod_flux_01y2.f:54-581
  input_PsiE            SRC/mod_PsiE_01y2.f:19-88

Arrays keep the reference's Fortran memory order so they can be handed to the
C ABI of libmw3 unchanged (a Fortran A(n1,n2) is the numpy C array [n2][n1]).
"""
import math
import os

import numpy as np

from . import consts as C
from . import namelist as NL
from . import refio
from .numerics import besph2, dwigner, gauleg, rgauscolini

__all__ = ["RefCase", "build_case", "load_case", "write_case", "InputError"]


class InputError(ValueError):
    """Inconsistent input; the reference writes a message and stops."""


class RefCase:
    """Module state of one propagation (attribute names follow the reference)."""

    def __repr__(self):
        return "<RefCase %dx%dx%d ncan=%d ntot=%d>" % (
            self.npun1, self.npun2, self.nangu, self.ncan, self.ntot)


def _grid(case, nl):
    g = NL.group(nl, "inputgridbase",
                  {"npun1min": 32, "iomref": 0, "ielecref": 1, "iommin": 0})
    for k in ("npun1", "rmis1", "rfin1", "npun2", "rmis2", "rfin2", "nangu",
              "jtot", "iparity", "inc", "nelecmax", "iommax", "j0", "jini",
              "jmax", "nvini", "nvmax", "nvref", "jref"):
        if k not in g:
            raise InputError("inputgridbase: %s missing" % k)
    c = case
    c.npun1, c.npun1min, c.npun2 = int(g["npun1"]), int(g["npun1min"]), int(g["npun2"])
    c.rmis1, c.rfin1 = float(g["rmis1"]), float(g["rfin1"])
    c.rmis2, c.rfin2 = float(g["rmis2"]), float(g["rfin2"])
    c.nangu, c.Jtot, c.iparity, c.inc = int(g["nangu"]), int(g["jtot"]), int(g["iparity"]), int(g["inc"])
    c.nelecmax, c.iommin, c.iommax, c.j0 = int(g["nelecmax"]), int(g["iommin"]), int(g["iommax"]), int(g["j0"])
    c.jini, c.jmax, c.nvini, c.nvmax = int(g["jini"]), int(g["jmax"]), int(g["nvini"]), int(g["nvmax"])
    c.nvref, c.jref, c.iomref, c.ielecref = int(g["nvref"]), int(g["jref"]), int(g["iomref"]), int(g["ielecref"])
    c.nangu2 = c.nangu * c.inc

    # products data (SRC/mod_gridYpara_01y2.f:131-158)
    pr = NL.group(nl, "inputprod", {"iomminprod": 0, "iommaxprod": 0,
                                     "n2prod1": c.npun2, "n2prod0": 1,
                                     "nangproj0": 1, "nangproj1": c.nangu,
                                     "iprod": 0, "emincut_prod_ev": 0.0,
                                     "nviniprod": 0, "nvmaxprod": 0,
                                     "jiniprod": 0, "jmaxprod": 0,
                                     "rbalinprod": 0.0}, required=False)
    c.iprod = int(pr["iprod"])
    c.nviniprod, c.nvmaxprod = int(pr["nviniprod"]), int(pr["nvmaxprod"])
    c.jiniprod, c.jmaxprod = int(pr["jiniprod"]), int(pr["jmaxprod"])
    c.iomminprod, c.iommaxprod = int(pr["iomminprod"]), int(pr["iommaxprod"])
    if c.iommaxprod == 0:
        c.iommaxprod = min(c.Jtot, c.jmaxprod)
    c.Rbalinprod = float(pr["rbalinprod"])
    c.n2prod0, c.n2prod1 = int(pr["n2prod0"]), min(int(pr["n2prod1"]), c.npun2)
    c.nangproj0, c.nangproj1 = int(pr["nangproj0"]), int(pr["nangproj1"])
    pc = NL.group(nl, "inputprocess", {"nelec_bnd": 1})
    c.iphoto = int(pc["iphoto"])
    c.nelec_bnd = int(pc["nelec_bnd"])

    # radial steps (SRC/mod_gridYpara_01y2.f:213-225)
    c.ah2 = (c.rfin2 - c.rmis2) / float(c.npun2 - 1) if c.npun2 > 1 else 0.0
    c.ah1 = (c.rfin1 - c.rmis1) / float(c.npun1 - 1) if c.npun1 > 1 else 0.0
    # Gauss-Legendre angular grid (SRC/mod_gridYpara_01y2.f:237-281)
    if c.nangu == 1:
        c.wreal = np.ones(1)
        c.cgamma = np.ones(1)
        c.cgamprod = np.ones(1)
        c.weiprod = np.ones(1)
    else:
        if c.inc == 1 and c.j0 != 0:
            raise InputError("inc=1 --> j0=0, here j0=%d" % c.j0)
        if c.inc == 2 and c.j0 > 1:
            raise InputError("inc=2 --> j0=0 or 1, here j0=%d" % c.j0)
        if c.inc not in (1, 2):
            raise InputError("inc=1 or 2, and it is inc=%d" % c.inc)
        w, x = gauleg(c.nangu2)
        c.cgamprod = x.copy()
        c.weiprod = w.copy()
        c.cgamma = x[: c.nangu].copy()
        c.wreal = w[: c.nangu] * float(c.inc)
    if abs(c.iparity) != 1:
        raise InputError("|iparity| must be 1")
    if c.jmax > (c.nangu2 - 1) * c.inc and c.npun2 > 1:
        raise InputError("jmax=%d > (nangu2-1)*inc" % c.jmax)


def _absorcion(c, nl):
    g = NL.group(nl, "inputabsorcion", {"ninterv1": 1, "ninterv2": 1,
                                         "n1expo": 2, "n2expo": 2})
    c.absr1, c.absalp1 = float(g["absr1"]), float(g["absalp1"])
    c.absr2, c.absalp2 = float(g["absr2"]), float(g["absalp2"])
    c.n1expo, c.n2expo = int(g["n1expo"]), int(g["n2expo"])
    c.ninterv1, c.ninterv2 = int(g["ninterv1"]), int(g["ninterv2"])
    c.absfr1 = np.ones(c.npun1)
    c.absfr2 = np.ones(c.npun2)
    c.ir1abs, c.ir2abs = c.npun1, c.npun2
    for ir1 in range(1, c.npun1 + 1):
        r1 = (ir1 - 1) * c.ah1 + c.rmis1
        if r1 >= c.absr1:
            c.ir1abs = min(c.ir1abs, ir1)
            delta = (r1 - c.absr1) / float(c.ninterv1)
            c.absfr1[ir1 - 1] = math.exp(-(c.absalp1 * delta ** c.n1expo))
    for ir2 in range(1, c.npun2 + 1):
        r2 = (ir2 - 1) * c.ah2 + c.rmis2
        if r2 >= c.absr2:
            c.ir2abs = min(c.ir2abs, ir2)
            delta = (r2 - c.absr2) / float(c.ninterv2)
            c.absfr2[ir2 - 1] = math.exp(-(c.absalp2 * delta ** c.n2expo))


def _pot0(c, nl):
    g = NL.group(nl, "inputpotmass", {"radau": 0, "r1inf_radial_functions": 100.0,
                                       "r2inf_radial_functions": 100.0,
                                       "no_ref_energy": 1, "diabatic_prod_pot": 0,
                                       "proj_j2_up": 0, "system": ""})
    c.system = g["system"]
    c.xm0, c.xm1, c.xm2 = float(g["xm0"]), float(g["xm1"]), float(g["xm2"])
    c.radau, c.No_ref_energy = int(g["radau"]), int(g["no_ref_energy"])
    c.Proj_j2_Up = int(g["proj_j2_up"])
    c.vcutmax = C.ev_to_zot(float(g["vcutmaxev"]))
    c.radcutmax = C.ev_to_zot(float(g["radcutmaxev"]))
    c.rotcutmax = C.ev_to_zot(float(g["rotcutmaxev"]))
    xm0, xm1, xm2 = c.xm0, c.xm1, c.xm2
    c.xmtot = xm0 + xm1 + xm2
    if c.radau == 0:
        c.xm1reac = (xm0 * xm1) / (xm0 + xm1)
        c.xm2reac = (xm2 * (xm0 + xm1)) / c.xmtot if xm2 > 1e-3 else c.xm1reac
        c.xm1prod = (xm0 * xm2) / (xm0 + xm2)
        c.xm2prod = (xm1 * (xm0 + xm2)) / c.xmtot
        c.xm1red = (xm0 * xm1) / (xm0 + xm1)
        c.xm2red = (xm2 * (xm0 + xm1)) / c.xmtot if xm2 > 1e-3 else c.xm1red
    else:
        c.xm1red, c.xm2red = xm1, xm2
        c.xm1reac, c.xm2reac = xm1, xm2
        c.xm1prod = c.xm2prod = xm1


def _basis(c, elec):
    """basis (SRC/mod_baseYfunciones_01y2.f:50-239) for xm2 > 1e-3."""
    c.nelec = len(elec)
    c.iomdiat = np.array([e[0] for e in elec], dtype=np.int32)
    c.iomatom = np.array([e[1] for e in elec], dtype=np.int32)
    c.sigdiat = np.array([e[2] for e in elec], dtype=float)
    c.sigatom = np.array([e[3] for e in elec], dtype=float)
    if c.iommax < c.iommin or c.iommax > c.Jtot:
        raise InputError("iommin/iommax inconsistent with Jtot")
    if c.xm2 < 1e-3:
        raise InputError("diatomic (xm2 < 1e-3) propagation is not on this path")
    iparbc = (-1) ** c.j0
    iombas, nelebas, j00, jbas, nojbas = [], [], [], [], []
    iommin_real = c.iommax
    for iom in range(c.iommin, c.iommax + 1):
        for ielec in range(1, c.nelec + 1):
            iomdi, iomat = int(c.iomdiat[ielec - 1]), int(c.iomatom[ielec - 1])
            isign = -1 if c.Jtot % 2 != 0 else 1
            ipar = c.sigdiat[ielec - 1] * c.sigatom[ielec - 1] * isign
            if iom == 0 and iomdi == 0 and iomat == 0 and c.iparity != ipar:
                continue
            jmin = c.j0
            if c.inc == 2:
                jmin = 0 if c.sigdiat[ielec - 1] == iparbc else 1
            while jmin < abs(iom - iomat) or jmin < abs(iomdi):
                jmin += c.inc
            iombas.append(iom)
            iommin_real = min(iommin_real, iom)
            nelebas.append(ielec)
            j00.append(jmin)
            jl = [jmin + (j - 1) * c.inc for j in range(1, c.nangu + 1)]
            jl = [jj for jj in jl if jj <= c.inc * (c.nangu - 1)]
            jbas.append(jl)
            nojbas.append(len(jl))
    if iommin_real > c.iommin:
        raise InputError("iommin must be set to iommin_real=%d" % iommin_real)
    if not iombas:
        raise InputError("NCAN=0")
    c.ncan = len(iombas)
    c.iombas = np.array(iombas, dtype=np.int32)
    c.nelebas = np.array(nelebas, dtype=np.int32)
    c.j00 = np.array(j00, dtype=np.int32)
    c.nojbas = np.array(nojbas, dtype=np.int32)
    c.jbas = np.zeros((c.ncan, c.nangu), dtype=np.int32)
    for i, jl in enumerate(jbas):
        c.jbas[i, : len(jl)] = jl
    # paralelizacion at nproc = 1: one processor holds every Omega and angle;
    # local angle slots run over decreasing global angle
    # (SRC/mod_baseYfunciones_01y2.f:365-376).
    c.indangreal = np.arange(c.nangu, 0, -1, dtype=np.int32)


def _angular_functions(c):
    """Djmmp(iang, j, ielec, iom) (SRC/mod_baseYfunciones_01y2.f:454-516)."""
    lmax = (c.nangu - 1) * c.inc
    jmax2 = lmax * 2
    c.ndj = c.nangu2 + 1
    D = np.zeros((c.Jtot + 1, c.nelec, c.ndj, c.nangu))
    for ielec in range(1, c.nelec + 1):
        for iom in range(0, c.Jtot + 1):
            m2 = 2 * (iom - int(c.iomatom[ielec - 1]))
            mp2 = 2 * int(c.iomdiat[ielec - 1])
            mmax = max(m2, mp2) // 2
            jprin = max(mmax, c.jini)
            if c.nangu == 1:
                D[iom, ielec - 1, jprin:lmax + 1, 0] = 1.0
                continue
            for iang in range(1, c.nangu + 1):
                gamma = math.acos(c.cgamprod[iang - 1])
                pm = dwigner(jmax2, m2, mp2, gamma)
                for j in range(jprin, lmax + 1):
                    xn = math.sqrt((2.0 * j + 1.0) * 0.5)
                    D[iom, ielec - 1, j, iang - 1] = pm[2 * j] * xn * math.sqrt(c.wreal[iang - 1])
    c.Djmmp = D


def _radial_functions(c, fdat):
    """radial_functions01_read (SRC/mod_baseYfunciones_01y2.f:837-923)."""
    nj = c.jmax - c.jini + 1
    nv = c.nvmax - c.nvini + 1
    c.noBCstates = np.zeros((c.nelecmax, nj), dtype=np.int32)
    c.ediat = np.zeros((c.nelecmax, nj, nv))
    c.fd = np.zeros((c.nelecmax, nj, nv, c.npun1))
    for ie in range(1, c.nelecmax + 1):
        for j in range(c.jini, c.jmax + 1):
            nb = int(fdat.nobc[ie - 1][j - c.jini])
            c.noBCstates[ie - 1, j - c.jini] = nb
            for iv in range(c.nvini, nb + 1):
                if iv > c.nvmax:
                    raise InputError("bcwf holds v=%d > nvmax" % iv)
                c.ediat[ie - 1, j - c.jini, iv - c.nvini] = fdat.ediat[(ie, j, iv)]
                c.fd[ie - 1, j - c.jini, iv - c.nvini, :] = fdat.fd[(ie, j, iv)]
    c.ediatref = fdat.eref
    for ie in range(1, c.nelec + 1):
        for j in range(c.jini, c.jmax + 1):
            for iv in range(c.nvini, int(c.noBCstates[ie - 1, j - c.jini]) + 1):
                c.ediat[ie - 1, j - c.jini, iv - c.nvini] -= c.ediatref


def _product_radial(c, pdat):
    """product_radialf_read (SRC/mod_baseYfunciones_01y2.f:1310-1443):
    fdprod(ir2, v, j, e) on the first n2prod1 R points, ediatprod(v, j) minus
    the reference energy of func/eref, noBCprod(j)."""
    if pdat is None:
        raise InputError("iprod > 0 needs func/prodwf")
    nj = c.jmaxprod - c.jiniprod + 1
    nv = c.nvmaxprod - c.nviniprod + 1
    c.noBCprod = np.array(pdat.nobc, dtype=np.int32)
    c.fdprod = np.zeros((c.nelecmax, nj, nv, c.n2prod1))
    c.ediatprod = np.zeros((nj, nv))
    c.elecprodmax = np.zeros((nj, nv), dtype=np.int32)
    c.vibelecprod = np.zeros((nj, nv))
    # v range per (j, state) printed by printprob (SRC/mod_baseYfunciones_01y2.f:1331-1332, 1419-1432)
    c.minvibprod = np.full((nj, c.nelecmax), c.nviniprod, dtype=np.int32)
    c.maxvibprod = np.full((nj, c.nelecmax), -1, dtype=np.int32)
    for j in range(c.jiniprod, c.jmaxprod + 1):
        for iv in range(c.nviniprod, int(c.noBCprod[j - c.jiniprod]) + 1):
            c.ediatprod[j - c.jiniprod, iv - c.nviniprod] = pdat.ediat[(j, iv)] - c.ediatref
            c.elecprodmax[j - c.jiniprod, iv - c.nviniprod] = pdat.elec[(j, iv)]
            c.vibelecprod[j - c.jiniprod, iv - c.nviniprod] = pdat.vibel[(j, iv)]
            ie = pdat.elec[(j, iv)]
            if pdat.vibel[(j, iv)] > c.maxvibprod[j - c.jiniprod, ie - 1]:
                c.maxvibprod[j - c.jiniprod, ie - 1] = int(pdat.vibel[(j, iv)])
            f = pdat.fd[(j, iv)]
            for ie in range(pdat.nelec):
                c.fdprod[ie, j - c.jiniprod, iv - c.nviniprod, :] = f[:, ie]


def _pot2(c, pot):
    """pot2 (SRC/mod_pot_01y2.f:344-533) at nproc = 1."""
    if (c.npun1 != pot.grid1[2] or c.npun2 != pot.grid2[2] or c.nangu != pot.nangu
            or c.inc != pot.inc):
        raise InputError("Problem in input no. of points and incj values")
    if (abs(c.rmis1 - pot.grid1[0]) > 1e-5 or abs(c.rfin1 - pot.grid1[1]) > 1e-5
            or abs(c.rmis2 - pot.grid2[0]) > 1e-5 or abs(c.rfin2 - pot.grid2[1]) > 1e-5):
        raise InputError("Problem in input in radial intervals")
    c.npunreal = pot.npunreal.astype(np.int32)
    c.npuntot = int(c.npunreal.max())
    na, np_ = c.nangu, c.npuntot
    c.indr1 = np.zeros((na, np_), dtype=np.int32)
    c.indr2 = np.zeros((na, np_), dtype=np.int32)
    for ia in range(na):
        n = c.npunreal[ia]
        c.indr1[ia, :n] = pot.indr1[ia]
        c.indr2[ia, :n] = pot.indr2[ia]
    ne = c.nelecmax
    c.VVV = np.zeros((ne, ne, na, np_))
    for sl in range(na):
        ia = int(c.indangreal[sl]) - 1
        n = c.npunreal[ia]
        v = pot.vmat[ia]  # [n][je][ie]
        for je in range(ne):
            for ie in range(ne):
                c.VVV[je, ie, sl, :n] = v[:, je, ie]
    # A0 layout offsets (SRC/mod_pot_01y2.f:435-455)
    c.off = np.zeros((na, c.ncan), dtype=np.int64)
    t = 0
    for sl in range(na):
        ia = int(c.indangreal[sl]) - 1
        for ic in range(c.ncan):
            c.off[sl, ic] = t
            t += int(c.npunreal[ia])
    c.ntot = t
    # spectral range (SRC/mod_pot_01y2.f:500-505)
    c.vmin = pot.vmin
    c.vmaxtot = pot.vmaxtot
    c.vmintot = pot.vmin
    c.emaxtot = c.vmaxtot + c.radcutmax + 3.0 * c.rotcutmax
    c.delta2 = 0.5 * (c.emaxtot - c.vmintot)
    c.emindlt = c.vmin + c.delta2


def point_index(c):
    """Per A0 element: (slot, channel, ir, ir1, ir2), the decode of indiproc
    (SRC/mod_pot_01y2.f:628-654), as int arrays."""
    sl = np.empty(c.ntot, dtype=np.int32)
    ch = np.empty(c.ntot, dtype=np.int32)
    ir = np.empty(c.ntot, dtype=np.int32)
    for s in range(c.nangu):
        ia = int(c.indangreal[s]) - 1
        n = int(c.npunreal[ia])
        for ic in range(c.ncan):
            o = int(c.off[s, ic])
            sl[o:o + n] = s
            ch[o:o + n] = ic
            ir[o:o + n] = np.arange(n)
    ia = c.indangreal[sl] - 1
    return sl, ch, ir, c.indr1[ia, ir], c.indr2[ia, ir]


def colini_params(c, nl):
    """Packet parameters of set_colini (SRC/mod_colini_01y2.f:36-84): k0,
    alpha, normalisation factor, closest l, and the norm of the R packet."""
    g = NL.group(nl, "inputcol")
    c.Rcolini = float(g["rcolini_ang"])
    c.ecol = C.ev_to_zot(float(g["ecol_ev"]))
    c.deltaE = C.ev_to_zot(float(g["deltae_ev"]))
    c.xmasa0 = c.xm2prod if c.iprod == 1 else c.xm2reac
    c.xk0 = math.sqrt(2.0 * c.xmasa0 * c.ecol) / C.hbr
    c.alpha0 = math.sqrt(2.0 * c.ecol / c.xmasa0) * C.hbr / c.deltaE
    c.factor0 = (2.0 / C.pi / c.alpha0 / c.alpha0) ** 0.25
    det = float(c.Jtot * (c.Jtot + 1) + c.jref * (c.jref + 1) - 2 * c.iomref * c.iomref)
    c.il0 = int(0.5 * (-1.0 + math.sqrt(1.0 + 4.0 * det)) + 0.5)
    r2 = c.rmis2 + np.arange(c.npun2) * c.ah2
    xr2 = rgauscolini(r2, c.Rcolini, c.alpha0, c.xk0, c.factor0)
    c.norm_gauss_R2 = float(np.sum(xr2 * xr2) * c.ah2)
    return xr2


def _colini(c, nl):
    """set_colini + colini_reacJacobi (SRC/mod_colini_01y2.f:25-177)."""
    c.xr2colini = colini_params(c, nl)
    if c.iprod == 1:
        raise InputError("iprod=1 (product Jacobi propagation) is not on this path")
    c.rpaq0 = colini_packet(c)
    c.norm_initial = float(np.sum(c.rpaq0 * c.rpaq0))


def colini_packet(c, jref=None):
    """The collision packet of colini_reacJacobi (SRC/mod_colini_01y2.f:86-177)
    for the reference state (nvref, jref, iomref) of the case, or for another
    rotational state jref of the same v and Omega choice."""
    xr2 = c.xr2colini
    jref = c.jref if jref is None else int(jref)
    sl, ch, ir, ir1, ir2 = point_index(c)
    iang = c.indangreal[sl]
    ielec = c.nelebas[ch]
    iom = c.iombas[ch]
    angpart = np.zeros(c.ntot)
    if c.iomref >= 0:
        m = (ielec == c.ielecref) & (iom == c.iomref)
        angpart[m] = c.Djmmp[iom[m], ielec[m] - 1, jref, iang[m] - 1]
    else:
        m = (ielec == c.ielecref) & (iom <= jref)
        angpart[m] = ((-1.0) ** iom[m]) * c.Djmmp[iom[m], ielec[m] - 1, jref, iang[m] - 1]
    if c.npun1 == 1:
        fdi = np.ones(c.ntot)
    else:
        fdi = c.fd[c.ielecref - 1, jref - c.jini, c.nvref - c.nvini, ir1 - 1]
    xr = xr2[ir2 - 1]
    return np.where(ielec == c.ielecref, -xr * angpart * 1.0 * fdi * math.sqrt(c.ah2), 0.0)


def flux_params(c, nl):
    """inputflux / inputtime values and the energy grid of Smatini."""
    g = NL.group(nl, "inputflux", {"ncontfile": 0})
    c.r1flux = float(g.get("r1flux_ang", c.absr1 - 1.0))
    c.netot = int(g["netot"])
    c.ekinmin = C.ev_to_zot(float(g["ekinmin_ev"]))
    c.ekinmax = C.ev_to_zot(float(g["ekinmax_ev"]))
    c.ncontfile = int(g["ncontfile"])
    tm = NL.group(nl, "inputtime", {"kminelastic": 0})
    c.ntimes, c.nloop = int(tm["ntimes"]), int(tm["nloop"])
    c.kminelastic = int(tm["kminelastic"])
    estep = (c.ekinmax - c.ekinmin) / float(c.netot - 1)
    c.etotS2 = c.ekinmin + np.arange(c.netot) * estep
    c.estep = estep


def smatini_norm(c):
    """Energy distribution of the incident packet in Smatini_col
    (SRC/mod_flux_01y2.f:427-503): per energy |<packet|l-wave>|^2 and its
    integral ("Norm ini. wvp in energy"); returns (paqini, pini, norm)."""
    npunt = 1000
    ahg = (c.rfin2 - c.rmis2) / float(npunt - 1)
    rg = c.rmis2 + np.arange(npunt) * ahg
    rgaussr = rgauscolini(rg, c.Rcolini, c.alpha0, c.xk0, c.factor0) / math.sqrt(2.0)
    pepe = float(c.il0 * (c.il0 + 1))
    r = c.Rcolini + 1.0
    erot = C.hbr * C.hbr * pepe * 0.5 / (r * r * c.xmasa0)
    paq = np.zeros(c.netot)
    pin = np.zeros(c.netot)
    for ie in range(c.netot):
        ekinini = c.etotS2[ie]
        if ekinini > 0.0 and ekinini > erot:
            pini = math.sqrt(ekinini * c.xmasa0 * 2.0 / C.hbr / C.hbr)
            F, _, G, _ = besph2(c.il0, rg * pini)
            keep = np.abs(F) > 1e-20
            zr = np.sum(np.where(keep, rgaussr * (-G), 0.0)) * ahg / 2.0 / C.pi
            zi = np.sum(np.where(keep, rgaussr * F, 0.0)) * ahg / 2.0 / C.pi
            paq[ie] = zr * zr + zi * zi
            pin[ie] = pini
    return paq, pin, float(np.sum(paq) * c.estep)


def _flux(c, nl):
    """ini_flux inputs and Smatini (SRC/mod_flux_01y2.f:54-334, 338-581)."""
    flux_params(c, nl)
    if c.iprod == 0 or c.npun1 == 1:  # SRC/mod_flux_01y2.f:258-265
        c.nviniprod, c.nvmaxprod = c.nvini, c.nvmax
        c.jiniprod, c.jmaxprod = c.jini, c.jmax
        c.iomminprod, c.iommaxprod = c.iommin, c.iommax
    # reference channel (SRC/mod_flux_01y2.f:348-370)
    c.icanref = 0
    if c.No_ref_energy == 1:
        for ic in range(c.ncan):
            if c.nelebas[ic] == c.ielecref and c.iombas[ic] == abs(c.iomref):
                c.icanref = ic + 1
        if c.icanref == 0:
            raise InputError("icanref can not be set")
    nj = c.jmax - c.jini + 1
    nv = c.nvmax - c.nvini + 1
    c.S2factor = np.zeros((c.nelecmax, nj, nv, c.netot))
    c.reacfct = np.zeros(c.netot)
    c.apaqini = np.zeros(c.netot)
    if c.iphoto == 0:
        c.photonorm = 1.0
        paq, pin, c.norm_energy = smatini_norm(c)
        c.apaqini = paq
        for ie in range(c.netot):
            paqini, pini = paq[ie], pin[ie]
            for iele in range(c.nelec):
                for j in range(c.jini, c.jmax + 1):
                    for iv in range(c.nvini, c.nvmax + 1):
                        ekinfin = c.etotS2[ie] - c.ediat[iele, j - c.jini, iv - c.nvini]
                        pfin = math.sqrt(ekinfin * c.xm2reac * 2.0 / C.hbr / C.hbr) if ekinfin > 0 else 0.0
                        if paqini >= 1e-10:
                            c.S2factor[iele, j - c.jini, iv - c.nvini, ie] = (
                                C.hbr * C.hbr * pfin * pini / (2.0 * paqini) / c.xm2reac / c.xmasa0)
            if paqini >= 1e-10:
                c.reacfct[ie] = C.hbr * C.hbr * pini / paqini / c.xmasa0 / c.xm1reac / 4.0 / C.pi / C.pi
    else:
        c.photonorm = 1.0
        for ie in range(c.netot):
            c.reacfct[ie] = 0.25 / C.pi / c.xm1reac
            for iele in range(c.nelec):
                for j in range(c.jini, c.jmax + 1):
                    for iv in range(c.nvini, c.nvmax + 1):
                        ekinfin = c.etotS2[ie] - c.ediat[iele, j - c.jini, iv - c.nvini]
                        pfin = math.sqrt(ekinfin * c.xm2reac * 2.0 / C.hbr / C.hbr) if ekinfin > 0 else 0.0
                        c.S2factor[iele, j - c.jini, iv - c.nvini, ie] = 4.0 * pfin / (C.pi * c.xm2reac)
    # psi(E) energies (SRC/mod_PsiE_01y2.f:19-88)
    c.nofe = 0
    c.efe = np.zeros(0)
    if c.iphoto > 0:
        pe = nl.get("inputpsie")
        if pe and int(pe.get("nofe", 0)) > 0:
            c.nofe = int(pe["nofe"])
            ev = pe.get("resonancese_ev")
            if ev is None:
                fn = pe["file_resonancese"]
                ev = np.loadtxt(os.path.join(c.base_dir, fn))[: c.nofe]
            c.efe = C.ev_to_zot(np.atleast_1d(np.asarray(ev, dtype=float)))


def build_case(nl, pot, fdat, base_dir="."):
    """Module state from namelist groups (namelist.parse output), PotData and
    FuncData (with its product states for iprod > 0), in the reference's
    initialisation order (SRC/main_madwave3.f:127-198)."""
    c = RefCase()
    c.base_dir = base_dir
    c.nl = nl
    _grid(c, nl)
    _absorcion(c, nl)
    _pot0(c, nl)
    if pot.nelec != c.nelecmax:
        raise InputError("cont.pot nelec=%d != nelecmax=%d" % (pot.nelec, c.nelecmax))
    _basis(c, pot.elec)
    _angular_functions(c)
    _radial_functions(c, fdat)
    if c.iprod > 0 and c.npun1 != 1:
        _product_radial(c, fdat.prod)
    _pot2(c, pot)
    if c.Proj_j2_Up != 0:
        raise InputError("Proj_j2_Up != 0 is not built")
    if c.iphoto == 0:
        _colini(c, nl)
    else:
        c.rpaq0 = None
        c.xmasa0 = c.xm2reac
    _flux(c, nl)
    if c.iprod == 2 and c.npun1 > 1:
        from . import prodtrans
        prodtrans.ini_transcoor(c)
    return c


def load_case(case_dir):
    """Read input.dat, pot/ and func/ like the reference propagator does."""
    with open(os.path.join(case_dir, "input.dat")) as f:
        nl = NL.parse(f.read())
    pot = refio.read_pot(refio.find_dir(case_dir, "pot"))
    g = NL.group(nl, "inputgridbase", {"npun1min": 32})
    fdir = (os.path.join(case_dir, "..", "func")
            if os.path.isdir(os.path.join(case_dir, "..", "func"))
            else os.path.join(case_dir, "func"))
    fdat = refio.read_func(fdir, int(g["nelecmax"]), int(g["jini"]), int(g["jmax"]),
                           int(g["nvini"]), int(g["npun1"]))
    pr = NL.group(nl, "inputprod", {"iprod": 0, "nviniprod": 0, "nvmaxprod": 0,
                                     "jiniprod": 0, "jmaxprod": 0,
                                     "n2prod1": int(g["npun2"])}, required=False)
    if int(pr["iprod"]) > 0 and int(g["npun1"]) != 1:
        fdat.prod = refio.read_prodwf(fdir, int(pr["jiniprod"]), int(pr["jmaxprod"]),
                                      int(pr["nviniprod"]), int(pr["nvmaxprod"]),
                                      min(int(pr["n2prod1"]), int(g["npun2"])))
    c = build_case(nl, pot, fdat, base_dir=case_dir)
    if c.iphoto in (1, 2):
        # read_trans_dipole + dip_bnd (SRC/main_madwave3.f:177-181): ../dip and
        # ../bndJxx{p,m} beside the case directory
        from . import photo
        b = NL.group(nl, "inputbnd", {"nprocbnd": 1, "maxbnddim": None, "igridbnd": 1})
        if int(b["igridbnd"]) != 1:
            raise InputError("only igridbnd = 1 (bound state on the grid) is built")
        up = os.path.join(case_dir, "..")
        dipol = photo.read_dip(os.path.join(up, "dip"), c, c.nelec_bnd) if c.iphoto == 1 else None
        bs = photo.read_bnd(up, c, int(b["jtotini"]), int(b["iparini"]), int(b["nvbound"]),
                            int(b["nprocbnd"]), b["maxbnddim"] and int(b["maxbnddim"]))
        photo.dip_bnd(c, bs, dipol, c.iphoto, int(b["jtotini"]), int(b["iparini"]))
    elif c.iphoto == 3:
        raise InputError("iphoto = 3 (V_elec times the bound state) is not built")
    return c


def write_case(case_dir, nl_text, pot, fdat):
    """Write input.dat, pot/ and func/ in the reference's formats."""
    os.makedirs(case_dir, exist_ok=True)
    with open(os.path.join(case_dir, "input.dat"), "w") as f:
        f.write(nl_text)
    refio.write_pot(os.path.join(case_dir, "pot"), pot)
    refio.write_func(os.path.join(case_dir, "func"), fdat)

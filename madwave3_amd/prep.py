"""The reference's preparation program for any single-state surface: the
pot/ and func/ content that main_potini writes (SRC/main_potini.f; the
potential grid of pot1, SRC/mod_pot_01y2.f:147-341, and the reactant
functions of radial_functions01_write, SRC/mod_baseYfunciones_01y2.f:553-835)
from a `potelebond` function of the reference's contract: distances r01, r02
in bohr and the cosine of the angle 1-0-2, energy in hartree.

leps.py plays the same part for its analytic surface; this module serves any
other one (the HD+H example's network surface in the golden-example test).
"""
import math

import numpy as np

from . import consts as C
from . import diatom
from . import namelist as NL
from .numerics import gauleg, splinqq, splset
from .refio import FuncData, PotData, ProdFuncData

__all__ = ["grid_potential", "reactant_functions", "product_functions", "product_functions_dvr",
           "prepare"]

H2Z = 1.0 / C.conve * C.conve1  # hartree -> zot


def grid_potential(potelebond, npun1, rmis1, rfin1, npun2, rmis2, rfin2, nangu, inc, xm0, xm1,
                   eref, vcut):
    """pot1 for one electronic state (radau = 0): at every angle of the
    Gauss-Legendre grid and every (r, R), the surface at r01 = r,
    r02 = |R + gamma r| (gamma = xm1/(xm0 + xm1)) and the angle at atom 0,
    in zots minus eref, capped at vcut; the points below vcut are kept, R
    outer and r inner.  Lengths in angstrom, eref and vcut in zots."""
    ah1 = (rfin1 - rmis1) / (npun1 - 1)
    ah2 = (rfin2 - rmis2) / (npun2 - 1)
    _, cg = gauleg(nangu * inc)
    cg = cg[:nangu]
    r = (rmis1 + np.arange(npun1) * ah1) / C.convl
    R = (rmis2 + np.arange(npun2) * ah2) / C.convl
    rr, RR = np.meshgrid(r, R)               # [ir2][ir1]
    gam = xm1 / (xm0 + xm1)
    indr1, indr2, vmat = [], [], []
    vmin, vmaxtot = 1e10, vcut
    for ia in range(nangu):
        c = cg[ia]
        x2 = np.sqrt(RR * RR + gam * gam * rr * rr + 2.0 * gam * RR * rr * c)
        ca = np.clip((RR * c + gam * rr) / x2, -1.0, 1.0)
        v = potelebond(rr, x2, ca) / C.conve * C.conve1 - eref
        v = np.where(v > vcut, vcut, v)
        vmin = min(vmin, float(v.min()))
        i2, i1 = np.nonzero(v < vcut)
        indr1.append((i1 + 1).astype(np.int32))
        indr2.append((i2 + 1).astype(np.int32))
        vmat.append(v[i2, i1].reshape(-1, 1, 1))
    return PotData(1, [(0, 0, 1.0, 1.0)], vmin, vmaxtot, (rmis1, rfin1, npun1),
                   (rmis2, rfin2, npun2), nangu, inc, indr1, indr2, vmat)


def reactant_functions(potelebond, npun1, rmis1, rfin1, xm0, xm1, jini, jmax, nvini, nvmax,
                       nvref, jref, r2inf=100.0):
    """radial_functions01_write for one electronic state: the 0-1 diatom with
    the third atom at r2inf angstrom on the line (cos = 1), its bound states
    for j = jini..jmax, v = nvini..nvmax on the r grid (diatom.py), energies
    in zots; eref = the (nvref, jref) energy."""
    ah1 = (rfin1 - rmis1) / (npun1 - 1)
    rmin = rmis1 / C.convl
    rmax = (rmis1 + (npun1 - 1) * ah1) / C.convl
    rg = r2inf / C.convl
    r_grid = (rmis1 + np.arange(npun1) * ah1) / C.convl
    mass = xm0 * xm1 / (xm0 + xm1) * C.convm
    states = diatom.radial_functions(lambda x: potelebond(x, rg, 1.0), rmin, rmax, mass, r_grid,
                                     range(jini, jmax + 1), nvini, nvmax)
    nobc = [[nvini - 1] * (jmax - jini + 1)]
    ediat, fd = {}, {}
    for j, (en, f) in states.items():
        for k in range(len(en)):
            iv = nvini + k
            ediat[(1, j, iv)] = float(en[k]) / C.conve * C.conve1
            fd[(1, j, iv)] = f[k]
            nobc[0][j - jini] = iv
    eref = ediat[(1, jref, nvref)]
    return FuncData(1, jini, jmax, nvini, npun1, nobc, ediat, fd, eref)


def product_functions(potelebond, npun2, rmis2, rfin2, n2prod1, xm0, xm2, jini, jmax, nvini, nvmax,
                      r1inf=100.0, n_fine=2000):
    """product_radialf_write for one electronic state
    (SRC/mod_baseYfunciones_01y2.f:1065-1308, with bndbc1ele,
    SRC/liboctdyn.f:1815-1913): the 0-2 diatom with atom 1 at r1inf angstrom
    (cos = 1) on a fine grid over the R range, its levels by the
    finite-difference estimates and Cooley's refinement (diatom.py), the
    functions splined onto the first n2prod1 R points and normalised to
    sum fd^2 = 1 there.

    The levels are counted from the lowest: eigenvalue number i = 0, 1, ..
    (at most nvmax - nvini + 1 of them, while bound by more than 0.01
    hartree) is stored as v' = nvini + i, its vibelecprod being i.
    Energies in zots as written (the reference energy is subtracted when
    the run reads them); lengths in angstrom."""
    rmin, rmax = rmis2 / C.convl, rfin2 / C.convl
    x = rmin + np.arange(n_fine) * ((rmax - rmin) / (n_fine - 1))
    h = (rmax - rmin) / (n_fine - 1)
    mass = xm0 * xm2 / (xm0 + xm2) * C.convm
    ah2 = (rfin2 - rmis2) / (npun2 - 1)
    r_grid = (rmis2 + np.arange(n2prod1) * ah2) / C.convl
    inside = (r_grid >= x[0]) & (r_grid <= x[-1])
    v = np.asarray(potelebond(r1inf / C.convl, x, 1.0), dtype=float)
    two_m = 2.0 * mass
    nobc, ediat, elec, vibel, fd = [], {}, {}, {}, {}
    for j in range(jini, jmax + 1):
        vv = v + 0.5 * j * (j + 1) / mass / (x * x)
        est = diatom.fd_eigenvalues(two_m * vv, h) / two_m
        top = nvini - 1
        for i in range(0, min(n_fine, nvmax - nvini) + 1):
            if not (v[-1] - est[i] > 1e-2):
                continue
            e, p, _ = diatom.cooley_numerov(est[i] * two_m, two_m * vv, h, rmax)
            top += 1
            f = np.zeros(n2prod1)
            f[inside] = splinqq(p, splset(p, x), x, r_grid[inside])
            s = float(np.sum(f * f))
            if s * ah2 > 1e-8:
                f /= math.sqrt(s)
            ediat[(j, top)] = float(e / two_m) * H2Z
            elec[(j, top)] = 1
            vibel[(j, top)] = float(i)
            fd[(j, top)] = f.reshape(n2prod1, 1)
        nobc.append(min(top, nvmax))
    return ProdFuncData(jini, jmax, nvini, n2prod1, 1, nobc, ediat, elec, vibel, fd)


def product_functions_dvr(potelebond, npun2, rmis2, rfin2, n2prod1, xm0, xm2, jini, jmax, nvini,
                          nvmax, r1inf=100.0):
    """The diabatic_prod_pot > 0 branch of product_radialf_write for one
    electronic state (SRC/mod_baseYfunciones_01y2.f:1083-1215, with
    bndbcele, SRC/liboctdyn.f:1687-1813): a sine-type DVR on the npun2
    points of the R grid itself, kinetic matrix
        T(i, i') = (-1)^(i - i') [1/(i - i')^2 + 1/(i + i')^2] / (mu h^2),
        T(i, i)  = [pi^2/6 + 1/(2i)^2] / (mu h^2),
    plus V(R) + j(j+1)/(2 mu R^2) on the diagonal.  Its lowest
    nvmax - nvini + 1 eigenvectors are kept for every j as v' = nvini..nvmax
    (sum fd^2 = 1 over the first n2prod1 points, vibelecprod = v'),
    whether or not they lie below the asymptote.  Energies in zots as
    written; lengths in angstrom."""
    rmin, rmax = rmis2 / C.convl, rfin2 / C.convl
    h = (rmax - rmin) / (npun2 - 1)
    r = rmin + np.arange(npun2) * h
    mass = xm0 * xm2 / (xm0 + xm2) * C.convm
    v = np.asarray(potelebond(r1inf / C.convl, r, 1.0), dtype=float)
    i = np.arange(1, npun2 + 1, dtype=float)
    d = i[:, None] - i[None, :]
    s = i[:, None] + i[None, :]
    sign = np.where(np.abs(d) % 2 == 1, -1.0, 1.0)
    np.fill_diagonal(d, 1.0)
    t = sign * (1.0 / (d * d) + 1.0 / (s * s))
    np.fill_diagonal(t, np.pi ** 2 / 6.0 + 1.0 / (2.0 * i) ** 2)
    t /= mass * h * h
    nv = nvmax - nvini + 1
    nobc, ediat, elec, vibel, fd = [], {}, {}, {}, {}
    for j in range(jini, jmax + 1):
        hmat = t + np.diag(v + 0.5 * j * (j + 1) / mass / (r * r))
        e, vec = np.linalg.eigh(hmat)
        for k in range(nv):
            iv = nvini + k
            ediat[(j, iv)] = float(e[k]) * H2Z
            elec[(j, iv)] = 1
            vibel[(j, iv)] = float(iv)
            fd[(j, iv)] = vec[:n2prod1, k].reshape(n2prod1, 1).copy()
        nobc.append(nvmax)
    return ProdFuncData(jini, jmax, nvini, n2prod1, 1, nobc, ediat, elec, vibel, fd)


def prepare(input_text, potelebond):
    """(PotData, FuncData) for the run of input_text (the input.dat
    namelists, as text or as namelist.parse groups) on the surface
    potelebond, as the preparation program writes them: functions first (the
    reference energy comes from them), then the potential grid; with
    iprod > 0 the product functions too (FuncData.prod), by the DVR when
    diabatic_prod_pot > 0."""
    nl = NL.parse(input_text) if isinstance(input_text, str) else input_text
    g = NL.group(nl, "inputgridbase", {"npun1min": 32, "inc": 1})
    pm = NL.group(nl, "inputpotmass")
    fdat = reactant_functions(potelebond, int(g["npun1"]), float(g["rmis1"]), float(g["rfin1"]),
                              float(pm["xm0"]), float(pm["xm1"]), int(g["jini"]), int(g["jmax"]),
                              int(g["nvini"]), int(g["nvmax"]), int(g["nvref"]), int(g["jref"]))
    pr = NL.group(nl, "inputprod", {"iprod": 0, "nviniprod": 0, "jiniprod": 0,
                                    "n2prod1": int(g["npun2"])})
    if int(pr["iprod"]) > 0:
        dvr = int(pm.get("diabatic_prod_pot", 0)) > 0
        fdat.prod = (product_functions_dvr if dvr else product_functions)(
            potelebond, int(g["npun2"]), float(g["rmis2"]), float(g["rfin2"]),
            min(int(pr["n2prod1"]), int(g["npun2"])), float(pm["xm0"]), float(pm["xm2"]),
            int(pr["jiniprod"]), int(pr["jmaxprod"]), int(pr["nviniprod"]), int(pr["nvmaxprod"]),
            float(pm.get("r1inf_radial_functions", 100.0)))
    vcut = C.ev_to_zot(float(pm["vcutmaxev"]))
    pot = grid_potential(potelebond, int(g["npun1"]), float(g["rmis1"]), float(g["rfin1"]),
                         int(g["npun2"]), float(g["rmis2"]), float(g["rfin2"]), int(g["nangu"]),
                         int(g["inc"]), float(pm["xm0"]), float(pm["xm1"]), fdat.eref, vcut)
    return pot, fdat

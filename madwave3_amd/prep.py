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
from .numerics import gauleg
from .refio import FuncData, PotData

__all__ = ["grid_potential", "reactant_functions", "prepare"]

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


def prepare(input_text, potelebond):
    """(PotData, FuncData) for the run of input_text (the input.dat
    namelists) on the surface potelebond, as the preparation program
    writes them: functions first (the reference energy comes from them),
    then the potential grid."""
    nl = NL.parse(input_text)
    g = NL.group(nl, "inputgridbase", {"npun1min": 32, "inc": 1})
    pm = NL.group(nl, "inputpotmass")
    fdat = reactant_functions(potelebond, int(g["npun1"]), float(g["rmis1"]), float(g["rfin1"]),
                              float(pm["xm0"]), float(pm["xm1"]), int(g["jini"]), int(g["jmax"]),
                              int(g["nvini"]), int(g["nvmax"]), int(g["nvref"]), int(g["jref"]))
    vcut = C.ev_to_zot(float(pm["vcutmaxev"]))
    pot = grid_potential(potelebond, int(g["npun1"]), float(g["rmis1"]), float(g["rfin1"]),
                         int(g["npun2"]), float(g["rmis2"]), float(g["rfin2"]), int(g["nangu"]),
                         int(g["inc"]), float(pm["xm0"]), float(pm["xm1"]), fdat.eref, vcut)
    return pot, fdat

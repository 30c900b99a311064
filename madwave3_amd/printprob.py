"""The probability tables of printprob (SRC/main_madwave3.f:555-844) for the
collision runs this path propagates (iprod = 0 or 2, triatomic, iphoto = 0 or
1), from the results of a propagation: P(E) (S2prodtot), the per-state flux
(S2prodelectot), the reactant amplitudes zS2 and, for iprod = 2, the product
amplitudes zS2prod.

Files, as the reference names and formats them:
  S2prod.vNN.JNNN.kNNNNN  iprod = 0: E (eV), P, P + S2reac, reacfct
                          iprod = 2: E (eV), P, P + S2reac, S2no, vibprod(v')
                                                                (501(1x,e17.7e3))
  S2prodelec              E (eV), P_elec for every state          (501(1x,e17.7e3))
  distriS2prod.elecNN     iprod = 2: E (eV), vibprod(e, v')       (501(1x,e17.7e3))
  with iwrt_reac_distri = 1:
  distriS2reac.elecNN     E (eV), S2reac of the state             (1000(1x,e15.7))
  distriS2reac.vNN.eN     E (eV), S2(v, j) for j = jini..jmax      (1000(1x,e15.7))
"""
import math
import os

import numpy as np

from . import consts as C


def fortran_e(x, w, d, e=2):
    """Fortran Ew.dEe edit descriptor (mantissa in [0.1, 1), e exponent
    digits), right-justified in w columns."""
    x = float(x)
    if x == 0.0 or not math.isfinite(x):
        mant, ex = 0.0, 0
        if not math.isfinite(x):
            return ("%" + str(w) + "s") % ("NaN" if x != x else ("Infinity" if x > 0 else "-Infinity"))
    else:
        ex = int(math.floor(math.log10(abs(x)))) + 1
        mant = abs(x) / 10.0 ** ex
        s = ("%." + str(d) + "f") % mant
        if s.startswith("1."):  # rounding carried into the next decade
            ex += 1
            mant = abs(x) / 10.0 ** ex
    digits = ("%." + str(d) + "f") % mant
    sign = "-" if x < 0 else ""
    es = "+" if ex >= 0 else "-"
    txt = "%s%sE%s%0*d" % (sign, digits, es, e, abs(ex))
    return txt.rjust(w)


def _line(vals, w, d, e):
    return "".join(" " + fortran_e(v, w, d, e) for v in vals) + "\n"


def reactant_probabilities(case, zs2):
    """S2(iv, j, ican)(E) = |zS2/2pi|^2 S2factor and their sum S2reac(E)
    (SRC/main_madwave3.f:621-633); zs2[ican][j - jini][v - nvini][ie]."""
    nj = case.jmax - case.jini + 1
    nv = case.nvmax - case.nvini + 1
    S2 = np.zeros((case.ncan, nj, nv, case.netot))
    for ican in range(case.ncan):
        iele = int(case.nelebas[ican])
        for j in range(int(case.j00[ican]), case.jmax + 1, case.inc):
            for iv in range(case.nvini, int(case.noBCstates[iele - 1, j - case.jini]) + 1):
                z = zs2[ican, j - case.jini, iv - case.nvini] / 2.0 / C.pi
                S2[ican, j - case.jini, iv - case.nvini] = (
                    (z * np.conj(z)).real * case.S2factor[iele - 1, j - case.jini, iv - case.nvini])
    return S2, S2.sum(axis=(0, 1, 2))


def product_probabilities(case, zs2prod):
    """Product-state probabilities of printprob for iprod = 2
    (SRC/main_madwave3.f:669-697): S2pro(v', j', Omega')(E) =
    |zS2prod|^2 / (4 pi^2) S2prodfac(E, v', j'), their sum S2no(E) and the
    vibrational distribution vibprod(e, v')(E) (state 1 when nelecmax = 1,
    else the state and v of largest weight, elecprodmax / vibelecprod).
    zs2prod[Omega' - iomminprod][j' - jiniprod][v' - nviniprod][ie]."""
    nomb = case.iommaxprod - case.iomminprod + 1
    nj = case.jmaxprod - case.jiniprod + 1
    nv = case.nvmaxprod - case.nviniprod + 1
    S2pro = np.zeros((nomb, nj, nv, case.netot))
    vib = np.zeros((case.nelecmax, nv, case.netot))
    for io in range(nomb):
        for j in range(case.jiniprod, case.jmaxprod + 1):
            for iv in range(case.nviniprod, int(case.noBCprod[j - case.jiniprod]) + 1):
                z = zs2prod[io, j - case.jiniprod, iv - case.nviniprod]
                av = (z * np.conj(z)).real * 0.25 / (C.pi * C.pi)
                s = av * case.S2prodfac[j - case.jiniprod, iv - case.nviniprod]
                S2pro[io, j - case.jiniprod, iv - case.nviniprod] = s
                if case.nelecmax == 1:
                    vib[0, iv - case.nviniprod] += s
                else:
                    ivbc = int(case.vibelecprod[j - case.jiniprod, iv - case.nviniprod])
                    ie = int(case.elecprodmax[j - case.jiniprod, iv - case.nviniprod])
                    if case.nviniprod <= ivbc <= case.nvmaxprod:
                        vib[ie - 1, ivbc - case.nviniprod] += s
    return S2pro, S2pro.sum(axis=(0, 1, 2)), vib


def write_tables(case, outdir, iloop, P, Pelec, zs2, iwrt_reac_distri=0, photonorm=None,
                 zs2prod=None):
    """Writes the printprob files of one analysis line into outdir; returns the
    list of paths.  zs2prod is required for iprod = 2."""
    iprod = getattr(case, "iprod", 0)
    if iprod not in (0, 2):
        raise ValueError("printprob tables are written for iprod = 0 or 2")
    if iprod == 2 and zs2prod is None:
        raise ValueError("iprod = 2 tables need zS2prod")
    pn = case.photonorm if photonorm is None else photonorm
    E = np.asarray(case.etotS2) / C.conve1 / C.eV2cm
    S2, S2reac = reactant_probabilities(case, zs2)
    Pt = np.where(np.asarray(P) < 1e-90, 0.0, np.asarray(P))
    Sr = np.where(S2reac < 1e-90, 0.0, S2reac)
    os.makedirs(outdir, exist_ok=True)
    paths = []
    name = os.path.join(outdir, "S2prod.v%02d.J%03d.k%05d" % (case.nvref, case.Jtot, iloop))
    if iprod == 2:
        _, S2no, vib = product_probabilities(case, zs2prod)
        vib = np.where(vib < 1e-90, 0.0, vib)
        ranges = [(int(case.minvibprod[0, e]), int(case.maxvibprod[0, e]))
                  for e in range(case.nelec)]
        with open(name, "w") as f:
            for ie in range(case.netot):
                cols = [E[ie], Pt[ie] * pn, (Pt[ie] + Sr[ie]) * pn, S2no[ie] * pn]
                for e, (v0, v1) in enumerate(ranges):
                    cols += [vib[e, v - case.nviniprod, ie] * pn for v in range(v0, v1 + 1)]
                f.write(_line(cols, 17, 7, 3))
        paths.append(name)
        for e, (v0, v1) in enumerate(ranges):
            fe = os.path.join(outdir, "distriS2prod.elec%02d" % (e + 1))
            with open(fe, "w") as f:
                for ie in range(case.netot):
                    f.write(_line([E[ie]] + [vib[e, v - case.nviniprod, ie] * pn
                                             for v in range(v0, v1 + 1)], 17, 7, 3))
            paths.append(fe)
    else:
        with open(name, "w") as f:
            for ie in range(case.netot):
                f.write(_line([E[ie], Pt[ie] * pn, (Pt[ie] + Sr[ie]) * pn, case.reacfct[ie]],
                              17, 7, 3))
        paths.append(name)
    name = os.path.join(outdir, "S2prodelec")
    with open(name, "w") as f:
        for ie in range(case.netot):
            f.write(_line([E[ie]] + [Pelec[e][ie] * pn for e in range(case.nelecmax)], 17, 7, 3))
    paths.append(name)
    if iwrt_reac_distri == 1:
        for ielec in range(1, case.nelec + 1):
            fe = os.path.join(outdir, "distriS2reac.elec%02d" % ielec)
            tot = np.zeros(case.netot)
            for iv in range(case.nvini, int(case.noBCstates[ielec - 1, 0]) + 1):
                rot = np.zeros((case.jmax - case.jini + 1, case.netot))
                for ican in range(case.ncan):
                    if int(case.nelebas[ican]) == ielec:
                        for j in range(int(case.j00[ican]), case.jmax + 1, case.inc):
                            rot[j - case.jini] += S2[ican, j - case.jini, iv - case.nvini]
                tot += rot.sum(axis=0)
                fv = os.path.join(outdir, "distriS2reac.v%02d.e%1d" % (iv, ielec))
                with open(fv, "w") as f:
                    for ie in range(case.netot):
                        f.write(_line([E[ie]] + list(rot[:, ie] * pn), 15, 7, 2))
                paths.append(fv)
            with open(fe, "w") as f:
                for ie in range(case.netot):
                    f.write(_line([E[ie], tot[ie] * pn], 15, 7, 2))
            paths.append(fe)
    return paths


def read_table(path):
    """Columns of a printprob table as a float array [rows][cols]."""
    return np.loadtxt(path, ndmin=2)

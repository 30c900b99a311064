"""Photodissociation spectrum from the autocorrelation function: the program
cheby-spectra of the reference (SRC/photo/cheby-spectra.f) on the GPU.

The autocorrelation file ("auto", written by check, SRC/main_madwave3.f:
444-455) holds a first line `xnorm ebndini emindlt delta2` and then lines
`k auto(k)`; auto(0) = 1.  The Chebyshev sum itself runs in libmw3
(mw3_cheby_spectrum); the unit conversions of the output follow
SRC/photo/cheby-spectra.f:28-97.
"""
import math

import numpy as np

from . import consts as C
from . import lib
from . import namelist as NL

# constants of cheby-spectra (SRC/photo/cheby-spectra.f:32-49)
_CLUZ_AU = 137.0
_EPS0_AU = 0.07957747
CS_CONSTANT_AU = 1.0 / (_CLUZ_AU * _EPS0_AU)
_ZOT2AU = (1.0 / C.conve1) * C.conve


def read_auto(path, kmax=None):
    """(xnorm, ebndini, emindlt, delta2, auto[0..kmaxreal]) from an auto file."""
    with open(path) as f:
        head = f.readline().split()
        xnorm, ebndini, emindlt, delta2 = (float(x.replace("D", "E").replace("d", "e"))
                                           for x in head[:4])
        vals = {0: 1.0}
        for ln in f:
            t = ln.split()
            if len(t) < 2:
                continue
            k = int(t[0])
            if kmax is not None and k > kmax:
                break
            vals[k] = float(t[1].replace("D", "E").replace("d", "e"))
    kreal = max(vals)
    auto = np.zeros(kreal + 1)
    for k, v in vals.items():
        auto[k] = v
    return xnorm, ebndini, emindlt, delta2, auto


def write_auto(path, xnorm, ebndini, emindlt, delta2, autocorr, k0=1):
    """Writes an auto file: the header line and `k auto(k)` for k = k0, k0+1, ..."""
    with open(path, "w") as f:
        f.write(" %.15e %.15e %.15e %.15e\n" % (xnorm, ebndini, emindlt, delta2))
        for i, v in enumerate(autocorr):
            f.write(" %d %.15e\n" % (k0 + i, v))


def spectrum_rows(emin_ev, emax_ev, ne, xkgam, auto, xnorm, ebndini, emindlt, delta2, eref):
    """Rows of the cheby-spectra output (SRC/photo/cheby-spectra.f:71-95):
    energy (eV), spectrum*xnorm, cross section (au), photon energy (au),
    spectrum (1/energy, au), CS constant; and the integral printed last."""
    estep = (emax_ev - emin_ev) / (ne - 1)
    ee = emin_ev + np.arange(ne) * estep
    e_zot = ee * C.eV2cm * C.conve1
    spec = lib.cheby_spectrum(auto, emindlt, delta2, C.hbr, xkgam, e_zot)
    eshift = eref / (C.conve1 * C.eV2cm)
    ephoton_au = (ee + eshift - ebndini) / C.au2eV
    spec_au = spec * xnorm / (C.pi * C.hbr * _ZOT2AU)
    xsec_au = CS_CONSTANT_AU * ephoton_au * spec_au
    rows = np.stack([ee, spec * xnorm, xsec_au, ephoton_au, spec_au,
                     np.full(ne, CS_CONSTANT_AU)], axis=1)
    integral = float(np.sum(spec_au / xnorm)) * estep / C.au2eV
    return rows, integral


def run(input_text, auto_path, eref):
    """cheby-spectra: the namelist inputEspectro of input.dat (Emin, Emax, ne,
    xkgam, kmax), the auto file and the reference energy (../func/eref)."""
    g = NL.group(NL.parse(input_text), "inputespectro")
    kmax = int(g["kmax"])
    xnorm, ebndini, emindlt, delta2, auto = read_auto(auto_path, kmax)
    return spectrum_rows(float(g["emin"]), float(g["emax"]), int(g["ne"]), float(g["xkgam"]),
                         auto, xnorm, ebndini, emindlt, delta2, eref)


def format_rows(rows, integral):
    """The text cheby-spectra writes to unit 6: (20(1x,e15.7)) per row."""
    out = []
    for r in rows:
        out.append("".join(" %15.7E" % v for v in r))
    out.append("  integral= %.15g" % integral)
    return "\n".join(out) + "\n"

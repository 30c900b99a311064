"""Physical constants and unit conversions of MADWAVE3.

Values follow SRC/mod_gridYpara_01y2.f:13-31: lengths in angstrom, masses in
amu, energies in "zots" (1 zot = 1/conve1 cm^-1).
"""
import math

pi = math.acos(-1.0)
conve1 = 1.197e-4          # zot per cm^-1
convl = 0.52917726         # angstrom per bohr
convm = 0.182288853e4      # electron masses per amu
conve = 4.55633538e-6      # hartree per cm^-1
eV2cm = 8065.5             # cm^-1 per eV
hbr = convl / math.sqrt(convm * conve / conve1)
au2eV = 27.21113957
zot2au = 0.0380646230576441
zot2eV = 1.03578177070099
hartree2cm = 219474.625


def ev_to_zot(e_ev):
    """eV -> zots as the reference converts input energies (e.g. mod_pot_01y2.f:131)."""
    return e_ev * eV2cm * conve1


def zot_to_ev(e_zot):
    """zots -> eV as the reference prints them (e.g. main_madwave3.f:728)."""
    return e_zot / conve1 / eV2cm


def bohr_to_ang(x):
    return x * convl

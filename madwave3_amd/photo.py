"""Photodissociation initial state (iphoto = 1, 2): the transition dipole and
bound-state files of the reference and the packet built from them.

The reference prepares this state in module mod_photoini_01y2 before the
propagation:
  * read_trans_dipole (SRC/mod_photoini_01y2.f:145-187) reads the spherical
    components dip(q = -1..1, ielec, ielec_bnd) of the transition dipole on
    the kept grid points from ../dip/dip.NNN.dat, one file per angle;
  * dip_bndgrid (SRC/mod_photoini_01y2.f:509-718) reads the bound state of
    total angular momentum Jtotini from ../bndJxx{p,m}/bnd.ivNNN.ipNNN and its
    point list from grid.ipNNN, and forms
        iphoto = 1: psi0(i) = sum_{ielec_bnd, Omega_ini, |q| <= 1}
                        bnd(r1, R, gamma, Omega_ini, ielec_bnd)
                        dip(q, ielec(i), ielec_bnd)(point) factresj(Omega_ini, q, Omega(i))
        iphoto = 2: psi0(i) = bnd(r1, R, gamma, Omega(i), ielec(i));
  * dip_bnd (SRC/mod_photoini_01y2.f:192-423) holds the rotational factors
    factresj (3j symbols of the one-photon transition, linear polarisation),
    normalises psi0 to 1 and writes the norm before normalisation, the
    bound-state energy, emindlt and delta2 as the header of the `auto` file.

The propagation then runs on the GPU like any other (mw3_cheby_begin with
this psi0, iphoto > 0 keeps <psi0|psi_k>), and spectra.py turns the
autocorrelation into the spectrum.

Files here are written and read in the reference's list-directed layouts.
"""
import math
import os
from fractions import Fraction

import numpy as np

from . import model

__all__ = ["tresj", "factresj_table", "write_dip", "read_dip", "write_bnd", "read_bnd",
           "BoundState", "dip_bndgrid", "dip_bnd", "MAX_OMEGA"]

# dip_bnd tabulates the rotational factors for Omega, Omega_ini = 0..5 only
# (max_Omega, SRC/mod_photoini_01y2.f:198)
MAX_OMEGA = 5


def tresj(j1, j2, j3, m1, m2, m3):
    """Wigner 3j symbol (j1 j2 j3; m1 m2 m3) for integer arguments, the value
    tresj returns (SRC/liboctdyn.f:1154-1194).

    Racah's single-sum formula evaluated in exact rational arithmetic: with
    T = (j1+j2-j3)! (j1-j2+j3)! (-j1+j2+j3)!/(j1+j2+j3+1)! and
    F = prod (j_i + m_i)! (j_i - m_i)!,
        3j = (-1)^(j1-j2-m3) sqrt(T F) sum_k (-1)^k / [k! (j3-j2+m1+k)!
             (j3-j1-m2+k)! (j1+j2-j3-k)! (j1-m1-k)! (j2+m2-k)!],
    the square root taken once in floating point.  Zero unless m1+m2+m3 = 0,
    |m_i| <= j_i and the triangle condition holds.
    """
    if m1 + m2 + m3 != 0:
        return 0.0
    if any(abs(m) > j for j, m in ((j1, m1), (j2, m2), (j3, m3))):
        return 0.0
    if j3 < abs(j1 - j2) or j3 > j1 + j2:
        return 0.0
    f = math.factorial
    tri = Fraction(f(j1 + j2 - j3) * f(j1 - j2 + j3) * f(-j1 + j2 + j3), f(j1 + j2 + j3 + 1))
    fm = f(j1 + m1) * f(j1 - m1) * f(j2 + m2) * f(j2 - m2) * f(j3 + m3) * f(j3 - m3)
    a1, a2 = j3 - j2 + m1, j3 - j1 - m2
    b1, b2, b3 = j1 + j2 - j3, j1 - m1, j2 + m2
    acc = Fraction(0)
    for k in range(max(0, -a1, -a2), min(b1, b2, b3) + 1):
        den = f(k) * f(a1 + k) * f(a2 + k) * f(b1 - k) * f(b2 - k) * f(b3 - k)
        acc += Fraction(-1 if k % 2 else 1, den)
    if acc == 0:
        return 0.0
    mag = math.sqrt(float(tri * fm * acc * acc))
    sign = 1.0 if acc > 0 else -1.0
    if (j1 - j2 - m3) % 2:
        sign = -sign
    return sign * mag


def factresj_table(Jtot, iparity, Jtotini, iparini, max_omega=MAX_OMEGA):
    """Rotational factors of the one-photon transition Jtotini -> Jtot with
    linearly polarised light, factresj(Omega_ini, q, Omega) of dip_bnd
    (SRC/mod_photoini_01y2.f:229-266), as an array [Omega_ini][q+1][Omega]
    for Omega, Omega_ini = 0..max_omega:
        (-1)^(Omega+q) n(Omega) n(Omega_ini) [ (Jini 1 J; Om_ini q -Om)
            + p_ini (-1)^Jini (Jini 1 J; -Om_ini q -Om) ]
    with n(0) = 1/sqrt(2), n(>0) = 1, and zero when iparity iparini >= 0 (the
    transition changes parity), Omega_ini > Jtotini or Omega > Jtot.  Raises
    ValueError when every factor vanishes (the reference stops: forbidden
    transition)."""
    tab = np.zeros((max_omega + 1, 3, max_omega + 1))
    for om in range(max_omega + 1):
        for omi in range(max_omega + 1):
            for q in (-1, 0, 1):
                if omi > Jtotini or om > Jtot:
                    continue
                norm = 1.0
                if om == 0:
                    norm /= math.sqrt(2.0)
                if omi == 0:
                    norm /= math.sqrt(2.0)
                if iparity * iparini >= 0:
                    norm = 0.0
                sym = tresj(Jtotini, 1, Jtot, omi, q, -om) + \
                    tresj(Jtotini, 1, Jtot, -omi, q, -om) * float(iparini * (-1) ** Jtotini)
                tab[omi, q + 1, om] = (-1.0) ** (om + q) * norm * sym
    if not np.any(np.abs(tab) > 1e-10):
        raise ValueError("forbidden transition for iphoto=1 (check Jtot, Jtotini and parities)")
    return tab


# ---------------------------------------------------------------- dip files
def write_dip(dip_dir, c, dipol):
    """Writes dip/dip.NNN.dat for every angle iang = 1..nangu: per kept point
    of the angle (indr1/indr2 order) a line `ir1 ir2` followed by
    dip(q, ielec, ielec_bnd) with q = -1..1 outermost, then ielec, ielec_bnd
    innermost (SRC/mod_photoini_01y2.f:71-133).  dipol is a list over
    iang - 1 of arrays [npunreal(iang)][3 (q+1)][nelecmax][nelec_bnd]."""
    os.makedirs(dip_dir, exist_ok=True)
    for ia in range(c.nangu):
        n = int(c.npunreal[ia])
        d = np.asarray(dipol[ia], dtype=float).reshape(n, -1)
        with open(os.path.join(dip_dir, "dip.%03d.dat" % (ia + 1)), "w") as f:
            for ir in range(n):
                f.write(" %d %d " % (c.indr1[ia, ir], c.indr2[ia, ir]))
                f.write(" ".join("%.17g" % v for v in d[ir]))
                f.write("\n")


def read_dip(dip_dir, c, nelec_bnd):
    """read_trans_dipole (SRC/mod_photoini_01y2.f:145-187): dipol as a list
    over iang - 1 of arrays [npunreal(iang)][3][nelecmax][nelec_bnd]; the
    point order of each file must be that of the potential files."""
    out = []
    per = 3 * c.nelecmax * nelec_bnd
    for ia in range(c.nangu):
        n = int(c.npunreal[ia])
        path = os.path.join(dip_dir, "dip.%03d.dat" % (ia + 1))
        vals = np.loadtxt(path, ndmin=2)
        if vals.shape != (n, 2 + per):
            raise ValueError("%s: expected %d lines of %d values" % (path, n, 2 + per))
        if not (np.array_equal(vals[:, 0].astype(int), c.indr1[ia, :n]) and
                np.array_equal(vals[:, 1].astype(int), c.indr2[ia, :n])):
            raise ValueError("%s: grid points differ from the potential files" % path)
        out.append(vals[:, 2:].reshape(n, 3, c.nelecmax, nelec_bnd))
    return out


# ---------------------------------------------------------------- bnd files
class BoundState:
    """One bound state of total angular momentum Jtotini on a list of grid
    points: ir1, ir2, iang (1-based), Omega and the electronic state
    (1-based) of each point, its amplitude, the vibrational index and the
    energy in eV, and the grid it was computed on (rmis1, rfin1, npun1,
    rmis2, rfin2, npun2, nangu, inc)."""

    def __init__(self, ir1, ir2, iang, omega, ielec, value, iv, energy_ev, grid):
        self.ir1 = np.asarray(ir1, dtype=np.int64)
        self.ir2 = np.asarray(ir2, dtype=np.int64)
        self.iang = np.asarray(iang, dtype=np.int64)
        self.omega = np.asarray(omega, dtype=np.int64)
        self.ielec = np.asarray(ielec, dtype=np.int64)
        self.value = np.asarray(value, dtype=float)
        self.iv = int(iv)
        self.energy_ev = float(energy_ev)
        self.grid = tuple(grid)


def bnd_dirname(Jtotini, iparini):
    if iparini not in (1, -1):
        raise ValueError("iparini must be +1 or -1")
    return "bndJ%02d%s" % (Jtotini, "p" if iparini == 1 else "m")


def write_bnd(base_dir, Jtotini, iparini, bs, nparts=1):
    """Writes the bound state as the bound-state program leaves it for
    dip_bndgrid, split over nparts `processor` files: bndJxx{p,m}/bnd.ivNNN.ipNNN
    (a line `iv energy_eV`, then one amplitude per line) and grid.ipNNN
    (nbnddim; rmis1 rfin1 npun1; rmis2 rfin2 npun2; nangu inc; then per value
    `i ican ielec Omega iangp ir ir1 ir2 iang`), SRC/mod_photoini_01y2.f:537-645."""
    d = os.path.join(base_dir, bnd_dirname(Jtotini, iparini))
    os.makedirs(d, exist_ok=True)
    parts = np.array_split(np.arange(bs.value.size), nparts)
    rmis1, rfin1, npun1, rmis2, rfin2, npun2, nangu, inc = bs.grid
    for ip, idx in enumerate(parts):
        with open(os.path.join(d, "bnd.iv%03d.ip%03d" % (bs.iv, ip)), "w") as f:
            f.write(" %d %.17g\n" % (bs.iv, bs.energy_ev))
            for v in bs.value[idx]:
                f.write(" %.17g\n" % v)
        with open(os.path.join(d, "grid.ip%03d" % ip), "w") as f:
            f.write(" %d\n" % idx.size)
            f.write(" %.17g %.17g %d\n" % (rmis1, rfin1, npun1))
            f.write(" %.17g %.17g %d\n" % (rmis2, rfin2, npun2))
            f.write(" %d %d\n" % (nangu, inc))
            for n, i in enumerate(idx):
                f.write(" %d 1 %d %d %d %d %d %d %d\n" % (
                    n + 1, bs.ielec[i], bs.omega[i], bs.iang[i], n + 1, bs.ir1[i], bs.ir2[i],
                    bs.iang[i]))


def read_bnd(base_dir, c, Jtotini, iparini, nvbound, nprocbnd, maxbnddim=None):
    """Reads the nprocbnd pieces of bound state nvbound as dip_bndgrid does
    (SRC/mod_photoini_01y2.f:537-645), checking that it was computed on the
    propagation's grid (r and R origin and step within 1e-4, same angles).
    Returns a BoundState holding every piece."""
    d = os.path.join(base_dir, bnd_dirname(Jtotini, iparini))
    cols = [[] for _ in range(6)]
    iv = energy = None
    for ip in range(nprocbnd):
        with open(os.path.join(d, "grid.ip%03d" % ip)) as f:
            lines = f.read().split("\n")
        nb = int(lines[0].split()[0])
        if maxbnddim is not None and nb > maxbnddim:
            raise ValueError("nbnddim=%d > maxbnddim=%d" % (nb, maxbnddim))
        rm1, rf1, n1 = lines[1].split()
        rm2, rf2, n2 = lines[2].split()
        na, inc = (int(x) for x in lines[3].split()[:2])
        ah1 = (float(rf1) - float(rm1)) / (int(n1) - 1)
        ah2 = (float(rf2) - float(rm2)) / (int(n2) - 1)
        if abs(c.rmis1 - float(rm1)) > 1e-4 or abs(c.ah1 - ah1) > 1e-4:
            raise ValueError("grid in r1 for bnd non equal")
        if abs(c.rmis2 - float(rm2)) > 1e-4 or abs(c.ah2 - ah2) > 1e-4:
            raise ValueError("grid in r2 for bnd non equal")
        if na != c.nangu or inc != c.inc:
            raise ValueError("grid in angle for bnd non equal")
        pts = np.array([ln.split() for ln in lines[4:4 + nb]], dtype=np.int64).reshape(nb, 9)
        with open(os.path.join(d, "bnd.iv%03d.ip%03d" % (nvbound, ip))) as f:
            head = f.readline().split()
            iv, energy = int(head[0]), float(head[1].replace("D", "E").replace("d", "e"))
            vals = np.array(f.read().replace("D", "E").replace("d", "e").split(), dtype=float)
        if vals.size < nb:
            raise ValueError("bnd file holds fewer values than its grid file")
        cols[0].append(pts[:, 6])
        cols[1].append(pts[:, 7])
        cols[2].append(pts[:, 8])
        cols[3].append(pts[:, 3])
        cols[4].append(pts[:, 2])
        cols[5].append(vals[:nb])
    grid = (c.rmis1, c.rmis1 + (c.npun1 - 1) * c.ah1, c.npun1, c.rmis2,
            c.rmis2 + (c.npun2 - 1) * c.ah2, c.npun2, c.nangu, c.inc)
    return BoundState(*(np.concatenate(x) for x in cols), iv, energy, grid)


# ---------------------------------------------------------------- the packet
def dip_bndgrid(c, bs, dipol, iphoto, fres=None):
    """Unnormalised initial packet in A0 layout (SRC/mod_photoini_01y2.f:646-703):
    for every state-point i inside the bound state's grid box (ir1, ir2 and
    iang not above the largest the state holds):
      iphoto = 1: sum over ielec_bnd and Omega_ini = 0..max(Omega of the
                  state) with |Omega(i) - Omega_ini| <= 1 of
                  bnd(ir1, ir2, iang, Omega_ini, ielec_bnd)
                  dip(q, ielec(i), ielec_bnd) factresj(Omega_ini, q, Omega(i));
      iphoto = 2: bnd(ir1, ir2, iang, Omega(i), ielec(i)) when Omega(i) and
                  ielec(i) are within the state's.
    dipol: list per angle of [npunreal][3][nelecmax][nelec_bnd] (read_dip);
    fres: factresj_table (iphoto = 1)."""
    if iphoto not in (1, 2):
        raise ValueError("dip_bndgrid: iphoto must be 1 or 2")
    sl, ch, ir, ir1, ir2 = model.point_index(c)
    iang = c.indangreal[sl].astype(np.int64)
    ielec = c.nelebas[ch].astype(np.int64)
    iom = c.iombas[ch].astype(np.int64)
    nr1, nr2, nan = int(bs.ir1.max()), int(bs.ir2.max()), int(bs.iang.max())
    ommax = int(max(bs.omega.max(), 0))
    nelb = int(bs.ielec.max())
    box = np.zeros((nelb, ommax + 1, nan, nr2, nr1))
    box[bs.ielec - 1, bs.omega, bs.iang - 1, bs.ir2 - 1, bs.ir1 - 1] = bs.value
    inside = (ir1 <= nr1) & (ir2 <= nr2) & (iang <= nan)
    psi = np.zeros(c.ntot)
    i1, i2, ia = ir1 - 1, ir2 - 1, iang - 1
    if iphoto == 2:
        ok = inside & (iom <= ommax) & (ielec <= nelb)
        psi[ok] = box[ielec[ok] - 1, iom[ok], ia[ok], i2[ok], i1[ok]]
        return psi
    if fres is None:
        raise ValueError("dip_bndgrid: iphoto = 1 needs the factresj table")
    # the dipole of each A0 element: per angle, rows in point order
    dip = np.zeros((c.ntot, 3, nelb))
    for s in range(c.nangu):
        a = int(c.indangreal[s]) - 1
        n = int(c.npunreal[a])
        for icn in range(c.ncan):
            o = int(c.off[s, icn])
            e = int(c.nelebas[icn]) - 1
            dip[o:o + n] = dipol[a][:, :, e, :nelb]
    for eb in range(nelb):
        for omi in range(ommax + 1):
            q = iom - omi
            ok = inside & (np.abs(q) <= 1) & (iom <= MAX_OMEGA) & (omi <= MAX_OMEGA)
            if not ok.any():
                continue
            fac = fres[omi, q[ok] + 1, iom[ok]]
            psi[ok] += box[eb, omi, ia[ok], i2[ok], i1[ok]] * dip[ok, q[ok] + 1, eb] * fac
    return psi


def dip_bnd(c, bs, dipol, iphoto, Jtotini, iparini):
    """The initial packet of a photodissociation (SRC/mod_photoini_01y2.f:192-423):
    dip_bndgrid, then normalisation to 1.  Sets c.rpaq0 and returns
    (xnorm_ini, energy_ini_eV): the norm before normalisation and the bound
    state's energy, the first two values of the auto file's header."""
    fres = factresj_table(c.Jtot, c.iparity, Jtotini, iparini) if iphoto == 1 else None
    if iphoto == 2 and (c.Jtot != Jtotini or c.iparity != iparini):
        raise ValueError("iphoto=2: Jtot, iparity must equal Jtotini, iparini")
    psi = dip_bndgrid(c, bs, dipol, iphoto, fres)
    xnorm = float(np.sum(psi * psi))
    if xnorm < 1e-15:
        raise ValueError("norm of initial wave packet too low: %g" % xnorm)
    c.rpaq0 = psi / math.sqrt(xnorm)
    c.photonorm = xnorm
    c.energy_ini_ev = bs.energy_ev
    return xnorm, bs.energy_ev

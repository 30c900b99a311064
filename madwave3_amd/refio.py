"""Readers and writers for MADWAVE3's on-disk inputs.

Formats (written by the reference's preparation program, read by the
propagator):
  pot/cont.pot      SRC/mod_pot_01y2.f:292-303 (write), 361-378 (read)
  pot/pot.NNN.dat   SRC/mod_pot_01y2.f:275-276 (write), 472-494 (read):
                    list-directed lines "ir1 ir2 V(1,1) V(2,1) ... V(ne,ne)"
                    for the kept points of angle NNN, R outer / r inner
  func/bcwf         SRC/mod_baseYfunciones_01y2.f:797-809 (write), 871-883 (read)
  func/eref         SRC/mod_baseYfunciones_01y2.f:788-790 (write), 889-891 (read)
Files are Fortran list-directed text; numbers may carry D exponents.
"""
import os

import numpy as np

__all__ = [
    "PotData", "FuncData", "write_pot", "read_pot", "write_func", "read_func",
    "find_dir",
]


def _numbers(text):
    if "D" in text or "d" in text:
        text = text.replace("D", "E").replace("d", "E")
    return np.array(text.split(), dtype=float)


def _fmt_real(x):
    return "%25.16E" % x


class PotData:
    """Content of pot/: electronic-state labels, cut statistics and, per global
    angle (0-based list index), the kept points and their V matrices."""

    def __init__(self, nelec, elec, vmin, vmaxtot, grid1, grid2, nangu, inc,
                 indr1, indr2, vmat):
        self.nelec = nelec
        self.elec = elec            # list of (iomdiat, iomatom, sigdiat, sigatom)
        self.vmin = vmin
        self.vmaxtot = vmaxtot
        self.grid1 = grid1          # (rmis1, rfin1, npun1)
        self.grid2 = grid2          # (rmis2, rfin2, npun2)
        self.nangu = nangu
        self.inc = inc
        self.indr1 = indr1          # list over angles of int arrays (1-based)
        self.indr2 = indr2
        self.vmat = vmat            # list over angles of [npts, ne(je), ne(ie)]

    @property
    def npunreal(self):
        return np.array([len(a) for a in self.indr1], dtype=np.int32)

    @property
    def maxpoint(self):
        return int(self.npunreal.sum())


class FuncData:
    """Content of func/: noBCstates(j, ielec), ediat(iv, j, ielec) in zots as
    written (before the reference energy is subtracted), fd(ir1, iv, j, ielec)
    and the reference energy."""

    def __init__(self, nelec, jini, jmax, nvini, npun1, nobc, ediat, fd, eref):
        self.nelec = nelec
        self.jini, self.jmax, self.nvini, self.npun1 = jini, jmax, nvini, npun1
        self.nobc = nobc            # [nelec][jmax-jini+1] ints
        self.ediat = ediat          # dict (ielec, j, iv) -> E
        self.fd = fd                # dict (ielec, j, iv) -> array[npun1]
        self.eref = eref


def find_dir(base, name):
    """The reference looks for ../pot and ../func before ./pot
    (SRC/mod_pot_01y2.f:360-365, 475-478)."""
    up = os.path.join(base, "..", name)
    if os.path.isdir(up):
        return up
    return os.path.join(base, name)


def write_pot(potdir, p):
    os.makedirs(potdir, exist_ok=True)
    with open(os.path.join(potdir, "cont.pot"), "w") as f:
        f.write("%12d\n" % p.nelec)
        for e in p.elec:
            f.write("%12d%12d%s%s\n" % (e[0], e[1], _fmt_real(e[2]), _fmt_real(e[3])))
        f.write("%s%12d%s\n" % (_fmt_real(p.vmin), p.maxpoint, _fmt_real(p.vmaxtot)))
        f.write("%s%s%12d\n" % (_fmt_real(p.grid1[0]), _fmt_real(p.grid1[1]), p.grid1[2]))
        f.write("%s%s%12d\n" % (_fmt_real(p.grid2[0]), _fmt_real(p.grid2[1]), p.grid2[2]))
        f.write("%12d%12d\n" % (p.nangu, p.inc))
        f.write("".join(" %12d" % n for n in p.npunreal) + "\n")
    ne = p.nelec
    for ia in range(p.nangu):
        i1, i2, v = p.indr1[ia], p.indr2[ia], p.vmat[ia]
        # columns: ir1 ir2 then V(ie,je) with ie fastest
        cols = [i1.astype(np.int64), i2.astype(np.int64)]
        vv = v.reshape(len(i1), ne * ne)
        with open(os.path.join(potdir, "pot.%03d.dat" % (ia + 1)), "w") as f:
            fmt = "%12d%12d" + "%25.16E" * (ne * ne) + "\n"
            for k in range(len(i1)):
                f.write(fmt % ((cols[0][k], cols[1][k]) + tuple(vv[k])))


def read_pot(potdir):
    with open(os.path.join(potdir, "cont.pot")) as f:
        toks = _numbers(f.read())
    pos = 0
    nelec = int(toks[pos]); pos += 1
    elec = []
    for _ in range(nelec):
        elec.append((int(toks[pos]), int(toks[pos + 1]), float(toks[pos + 2]),
                     float(toks[pos + 3])))
        pos += 4
    vmin, maxpoint, vmaxtot = float(toks[pos]), int(toks[pos + 1]), float(toks[pos + 2])
    pos += 3
    grid1 = (float(toks[pos]), float(toks[pos + 1]), int(toks[pos + 2])); pos += 3
    grid2 = (float(toks[pos]), float(toks[pos + 1]), int(toks[pos + 2])); pos += 3
    nangu, inc = int(toks[pos]), int(toks[pos + 1]); pos += 2
    npunreal = toks[pos:pos + nangu].astype(np.int64)
    indr1, indr2, vmat = [], [], []
    ncol = 2 + nelec * nelec
    for ia in range(nangu):
        n = int(npunreal[ia])
        with open(os.path.join(potdir, "pot.%03d.dat" % (ia + 1))) as f:
            arr = _numbers(f.read())
        if arr.size < n * ncol:
            raise ValueError("pot.%03d.dat has fewer than npunreal=%d lines" % (ia + 1, n))
        arr = arr[: n * ncol].reshape(n, ncol)
        indr1.append(arr[:, 0].astype(np.int32))
        indr2.append(arr[:, 1].astype(np.int32))
        vmat.append(arr[:, 2:].reshape(n, nelec, nelec).copy())
    return PotData(nelec, elec, vmin, vmaxtot, grid1, grid2, nangu, inc,
                   indr1, indr2, vmat)


def write_func(funcdir, fdat):
    os.makedirs(funcdir, exist_ok=True)
    with open(os.path.join(funcdir, "eref"), "w") as f:
        f.write(_fmt_real(fdat.eref) + "\n")
    with open(os.path.join(funcdir, "bcwf"), "w") as f:
        for ie in range(1, fdat.nelec + 1):
            for j in range(fdat.jini, fdat.jmax + 1):
                nb = fdat.nobc[ie - 1][j - fdat.jini]
                f.write("%12d%12d%12d\n" % (ie, j, nb))
                for iv in range(fdat.nvini, nb + 1):
                    f.write("%12d%s\n" % (iv, _fmt_real(fdat.ediat[(ie, j, iv)])))
                    f.write("".join(_fmt_real(x) + "\n" for x in fdat.fd[(ie, j, iv)]))


def read_func(funcdir, nelec, jini, jmax, nvini, npun1):
    with open(os.path.join(funcdir, "eref")) as f:
        eref = float(_numbers(f.read())[0])
    with open(os.path.join(funcdir, "bcwf")) as f:
        toks = _numbers(f.read())
    pos = 0
    nobc = [[0] * (jmax - jini + 1) for _ in range(nelec)]
    ediat, fd = {}, {}
    for ie in range(1, nelec + 1):
        for j in range(jini, jmax + 1):
            pos += 2
            nb = int(toks[pos]); pos += 1
            nobc[ie - 1][j - jini] = nb
            for iv in range(nvini, nb + 1):
                pos += 1
                ediat[(ie, j, iv)] = float(toks[pos]); pos += 1
                fd[(ie, j, iv)] = toks[pos:pos + npun1].copy(); pos += npun1
    return FuncData(nelec, jini, jmax, nvini, npun1, nobc, ediat, fd, eref)

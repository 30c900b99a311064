"""Readers and writers for MADWAVE3's on-disk inputs.

Formats (written by the reference's preparation program, read by the
propagator):
  pot/cont.pot      SRC/mod_pot_01y2.f:292-303 (write), 361-378 (read)
  pot/pot.NNN.dat   SRC/mod_pot_01y2.f:275-276 (write), 472-494 (read):
                    list-directed lines "ir1 ir2 V(1,1) V(2,1) ... V(ne,ne)"
                    for the kept points of angle NNN, R outer / r inner
  func/bcwf         SRC/mod_baseYfunciones_01y2.f:797-809 (write), 871-883 (read)
  func/eref         SRC/mod_baseYfunciones_01y2.f:788-790 (write), 889-891 (read)
  func/prodwf       SRC/mod_baseYfunciones_01y2.f:1268-1279 (write), 1353-1385 (read):
                    per j "j noBCprod n2prod1 nelecmax", per v "v E ielec v_e",
                    then n2prod1 lines of fdprod(ir2, v, j, 1..nelecmax)
Files are Fortran list-directed text; numbers may carry D exponents.
"""
import os

import numpy as np

__all__ = [
    "PotData", "FuncData", "ProdFuncData", "write_pot", "read_pot", "write_func",
    "read_func", "write_prodwf", "read_prodwf", "find_dir",
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
    written (before the reference energy is subtracted), fd(ir1, iv, j, ielec),
    the reference energy and, when the run projects on products, the
    ProdFuncData of func/prodwf (prod; None otherwise)."""

    def __init__(self, nelec, jini, jmax, nvini, npun1, nobc, ediat, fd, eref, prod=None):
        self.nelec = nelec
        self.jini, self.jmax, self.nvini, self.npun1 = jini, jmax, nvini, npun1
        self.nobc = nobc            # [nelec][jmax-jini+1] ints
        self.ediat = ediat          # dict (ielec, j, iv) -> E
        self.fd = fd                # dict (ielec, j, iv) -> array[npun1]
        self.eref = eref
        self.prod = prod


class ProdFuncData:
    """Content of func/prodwf: product (0-2 diatom) states on the R grid.
    nobc[j - jini] = noBCprod(j); per (j, v): energy in zots as written,
    elecprodmax, vibelecprod and fdprod [n2prod1][nelec]."""

    def __init__(self, jini, jmax, nvini, n2prod1, nelec, nobc, ediat, elec, vibel, fd):
        self.jini, self.jmax, self.nvini = jini, jmax, nvini
        self.n2prod1, self.nelec = n2prod1, nelec
        self.nobc = nobc      # [jmax-jini+1] ints
        self.ediat = ediat    # dict (j, iv) -> E
        self.elec = elec      # dict (j, iv) -> electronic state of largest weight
        self.vibel = vibel    # dict (j, iv) -> vibrational number in that state
        self.fd = fd          # dict (j, iv) -> array[n2prod1][nelec]


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
    if fdat.prod is not None:
        write_prodwf(funcdir, fdat.prod)


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


def write_prodwf(funcdir, pd):
    os.makedirs(funcdir, exist_ok=True)
    with open(os.path.join(funcdir, "prodwf"), "w") as f:
        for j in range(pd.jini, pd.jmax + 1):
            nb = pd.nobc[j - pd.jini]
            f.write("%12d%12d%12d%12d\n" % (j, nb, pd.n2prod1, pd.nelec))
            for iv in range(pd.nvini, nb + 1):
                f.write("%12d%s%12d%s\n" % (iv, _fmt_real(pd.ediat[(j, iv)]), pd.elec[(j, iv)],
                                           _fmt_real(pd.vibel[(j, iv)])))
                fd = np.asarray(pd.fd[(j, iv)]).reshape(pd.n2prod1, pd.nelec)
                f.write("".join("".join(_fmt_real(x) for x in row) + "\n" for row in fd))


def read_prodwf(funcdir, jini, jmax, nvini, nvmax, n2prod1):
    """product_radialf_read (SRC/mod_baseYfunciones_01y2.f:1353-1385): stops
    (ValueError) when a j holds more states than nvmax or another n2prod1."""
    with open(os.path.join(funcdir, "prodwf")) as f:
        toks = _numbers(f.read())
    pos = 0
    nobc = [0] * (jmax - jini + 1)
    ediat, elec, vibel, fd = {}, {}, {}, {}
    nelec = 1
    for j in range(jini, jmax + 1):
        nb, nn2, nelec = int(toks[pos + 1]), int(toks[pos + 2]), int(toks[pos + 3])
        pos += 4
        if nvmax < nb:
            raise ValueError("nvmaxprod = %d < noBCprod(%d) = %d in func/prodwf" % (nvmax, j, nb))
        if nn2 != n2prod1:
            raise ValueError("func/prodwf holds n2prod1 = %d, input.dat %d" % (nn2, n2prod1))
        nobc[j - jini] = nb
        for iv in range(nvini, nb + 1):
            ediat[(j, iv)] = float(toks[pos + 1])
            elec[(j, iv)] = int(toks[pos + 2])
            vibel[(j, iv)] = float(toks[pos + 3])
            pos += 4
            fd[(j, iv)] = toks[pos:pos + nn2 * nelec].reshape(nn2, nelec).copy()
            pos += nn2 * nelec
    return ProdFuncData(jini, jmax, nvini, n2prod1, nelec, nobc, ediat, elec, vibel, fd)

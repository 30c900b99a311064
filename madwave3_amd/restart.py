"""Continuation files of a propagation: the ones `wvpcont` writes at the end
of every time loop when ncontfile = 1 (SRC/main_madwave3.f:494-552) and the
run reads back at start-up (main_madwave3.f:223-261; mod_flux_01y2.f:298-322;
mod_coortrans01_02.f:180-196):

  cont.data        list-directed: "iit0 indt iloop" and "abstot nflux ncan"
  contNN.paq       unformatted: rflanz0proc (psi_{k-1}), rpaqproc (psi_k)
  contNN.S2        unformatted: zS2(netot, v, j, channel), zCR(netot, flux
                   point), zCdR(netot, flux point)
  contNN.S2prod    unformatted (iprod > 1): zS2prod(netot, v', j', Omega')

and the per-step records of every time loop (main_madwave3.f:311-322,
353-365, 388-394), written by rank 0:

  Cvj.LLLL         unformatted, one record "it, Cvj" per step (iphoto >= 1,
                   iprod = 1 or iwrt_reac_distri = 2, and xm2 > 1e-3)
  Cvjprod.LLLL     unformatted, one record "it, Cvjprod" per step (iprod = 2)

with LLLL the loop number iloop.
NN is the rank; one process holds the whole grid here, so the files are
those of rank 0 of a one-rank run (the state in the A0 layout of
nproc = 1, the flux points in its order).  Unformatted sequential files
follow the usual Fortran record framing: a 4-byte little-endian byte count
before and after each record.  wvpcont writes the second line of cont.data
from a variable it never sets (abstot); 0 is written here.

write_cont(dir, prop, iloop) and resume(prop, dir) move a Propagator's state
and accumulators through these files; read_cont returns their contents.
time_loop(prop, dir, iloop, ntimes) runs one loop of the reference's main
program on a Propagator: ntimes steps, the loop's record files and the
continuation files."""
import os
import struct

import numpy as np

__all__ = ["write_records", "read_records", "write_cont", "read_cont", "resume",
           "write_step_records", "read_step_records", "record_kinds", "time_loop"]

_MAXREC = 2 ** 31 - 1


def write_records(path, arrays):
    """One unformatted sequential record per array (Fortran order of the
    array's dimensions given in C order reversed; complex as (re, im))."""
    with open(path, "wb") as f:
        for a in arrays:
            a = np.asarray(a)
            if np.iscomplexobj(a):
                a = np.ascontiguousarray(a, dtype=np.complex128)
            else:
                a = np.ascontiguousarray(a, dtype=np.float64)
            b = a.tobytes()
            if len(b) > _MAXREC:
                raise ValueError("record of %d bytes exceeds one Fortran record" % len(b))
            mark = struct.pack("<i", len(b))
            f.write(mark)
            f.write(b)
            f.write(mark)


def read_records(path):
    """The records of an unformatted sequential file, as bytes."""
    out = []
    with open(path, "rb") as f:
        data = f.read()
    pos = 0
    while pos < len(data):
        (n,) = struct.unpack_from("<i", data, pos)
        if n < 0:
            raise ValueError("%s: continued records are not read" % path)
        rec = data[pos + 4:pos + 4 + n]
        (m,) = struct.unpack_from("<i", data, pos + 4 + n)
        if m != n or len(rec) != n:
            raise ValueError("%s: broken record framing at byte %d" % (path, pos))
        out.append(rec)
        pos += 8 + n
    return out


def _fmt_int(v):
    return "%12d" % v


def write_cont(dirpath, prop, iloop, indt=0, idproc=0):
    """wvpcont for one rank: the state, zS2, zCR, zCdR (and zS2prod with
    iprod > 1) of Propagator prop after its current step, and cont.data with
    iit0 = k, indt + 1 and iloop."""
    case = prop.case
    os.makedirs(dirpath, exist_ok=True)
    psim1, psi, k = prop.state()
    write_records(os.path.join(dirpath, "cont%02d.paq" % idproc), [psim1, psi])
    zcr, zcdr = prop.flux_accum()
    write_records(os.path.join(dirpath, "cont%02d.S2" % idproc), [prop.zs2(), zcr, zcdr])
    if getattr(case, "iprod", 0) > 1:
        write_records(os.path.join(dirpath, "cont%02d.S2prod" % idproc), [prop.zs2prod()])
    with open(os.path.join(dirpath, "cont.data"), "w") as f:
        f.write(" %s%s%s\n" % (_fmt_int(k), _fmt_int(indt + 1), _fmt_int(iloop)))
        f.write("   %.16E %s%s\n" % (0.0, _fmt_int(prop.nflux), _fmt_int(case.ncan)))
    return k


def read_cont(dirpath, case, nflux, idproc=0):
    """Contents of the continuation files for a case with nflux flux points:
    dict with it0, indt, iloop, psim1, psi, zs2, zcr, zcdr and zs2prod (or
    None), shaped as the Propagator getters return them."""
    with open(os.path.join(dirpath, "cont.data")) as f:
        first = f.readline().split()
        second = f.readline().split()
    it0, indt, iloop = (int(x) for x in first[:3])
    if second and (int(second[1]) != nflux or int(second[2]) != case.ncan):
        raise ValueError("cont.data: nflux %s ncan %s, this case has %d, %d"
                         % (second[1], second[2], nflux, case.ncan))
    paq = read_records(os.path.join(dirpath, "cont%02d.paq" % idproc))
    if len(paq) != 2:
        raise ValueError("cont%02d.paq: expected 2 records, found %d" % (idproc, len(paq)))
    psim1, psi = (np.frombuffer(r, dtype=np.float64).copy() for r in paq)
    if psi.size != case.ntot or psim1.size != case.ntot:
        raise ValueError("cont%02d.paq: %d values, the case has %d" % (idproc, psi.size, case.ntot))
    s2 = read_records(os.path.join(dirpath, "cont%02d.S2" % idproc))
    if len(s2) != 3:
        raise ValueError("cont%02d.S2: expected 3 records, found %d" % (idproc, len(s2)))
    nv, nj = case.nvmax - case.nvini + 1, case.jmax - case.jini + 1
    zs2 = np.frombuffer(s2[0], dtype=np.complex128).reshape(case.ncan, nj, nv, case.netot).copy()
    zcr = np.frombuffer(s2[1], dtype=np.complex128).reshape(nflux, case.netot).copy()
    zcdr = np.frombuffer(s2[2], dtype=np.complex128).reshape(nflux, case.netot).copy()
    zs2prod = None
    if getattr(case, "iprod", 0) > 1:
        rec = read_records(os.path.join(dirpath, "cont%02d.S2prod" % idproc))
        shape = (case.iommaxprod - case.iomminprod + 1, case.jmaxprod - case.jiniprod + 1,
                 case.nvmaxprod - case.nviniprod + 1, case.netot)
        zs2prod = np.frombuffer(rec[0], dtype=np.complex128).reshape(shape).copy()
    return {"it0": it0, "indt": indt, "iloop": iloop, "psim1": psim1, "psi": psi, "zs2": zs2,
            "zcr": zcr, "zcdr": zcdr, "zs2prod": zs2prod}


def resume(prop, dirpath, idproc=0):
    """Start-up with it0 > 0 and ncontfile = 1: puts the state (k = iit0)
    and the accumulators of the files into prop; the initial packet of the
    case stays the autocorrelation's reference.  Returns the cont.data
    counters (it0, indt, iloop)."""
    case = prop.case
    d = read_cont(dirpath, case, prop.nflux, idproc)
    prop.put_state(d["psim1"], d["psi"], d["it0"], psi0=case.rpaq0)
    prop.put_accum(d["zcr"], d["zcdr"], d["zs2"])
    if d["zs2prod"] is not None:
        prop.put_zs2prod(d["zs2prod"])
    return d["it0"], d["indt"], d["iloop"]


def write_step_records(path, its, values):
    """A Cvj.LLLL / Cvjprod.LLLL file: per step i one unformatted record
    holding the 4-byte integer its[i] and then values[i] (float64, its
    dimensions in C order reversed: Fortran order).  Opened as the reference
    opens it (status='new'): an existing file is an error."""
    its = np.asarray(its, dtype=np.int32)
    values = np.ascontiguousarray(values, dtype=np.float64).reshape(len(its), -1)
    n = 4 + 8 * values.shape[1]
    if n > _MAXREC:
        raise ValueError("record of %d bytes exceeds one Fortran record" % n)
    mark = struct.pack("<i", n)
    with open(path, "xb") as f:
        for i in range(len(its)):
            f.write(mark)
            f.write(struct.pack("<i", int(its[i])))
            f.write(values[i].tobytes())
            f.write(mark)


def read_step_records(path, shape):
    """(its, values[n, *shape]) of a file written as write_step_records
    writes it."""
    m = int(np.prod(shape))
    its, vals = [], []
    for rec in read_records(path):
        if len(rec) != 4 + 8 * m:
            raise ValueError("%s: record of %d bytes, expected %d" % (path, len(rec), 4 + 8 * m))
        its.append(struct.unpack_from("<i", rec, 0)[0])
        vals.append(np.frombuffer(rec, dtype=np.float64, offset=4))
    v = np.array(vals).reshape((len(its),) + tuple(shape))
    return np.array(its, dtype=np.int32), v


def record_kinds(case):
    """(Cvj, Cvjprod): whether the main program writes each record file for
    this case (main_madwave3.f:311-322)."""
    cvj = ((getattr(case, "iphoto", 0) >= 1 or getattr(case, "iprod", 0) == 1
            or getattr(case, "iwrt_reac_distri", 0) == 2) and getattr(case, "xm2", 1.0) > 1e-3)
    return cvj, getattr(case, "iprod", 0) == 2


def time_loop(prop, dirpath, iloop, ntimes, indt=0, idproc=0):
    """One pass of the main program's time loop (main_madwave3.f:308-394)
    on Propagator prop, begun or resumed: ntimes Chebyshev steps with their
    per-step records into Cvj.LLLL / Cvjprod.LLLL (as the case asks for
    them), then the continuation files.  Returns the steps' norms."""
    case = prop.case
    want_cvj, want_prod = record_kinds(case)
    want_cvj = want_cvj and case.netot > 0
    os.makedirs(dirpath, exist_ok=True)
    names = [os.path.join(dirpath, "%s.%04d" % (w, iloop))
             for w, on in (("Cvj", want_cvj), ("Cvjprod", want_prod)) if on]
    for n in names:
        if os.path.exists(n):
            raise FileExistsError(n)
    prop.records(want_cvj or want_prod)
    try:
        norms = prop.steps(ntimes)[0]
        its, cvj, cvjprod = prop.take_records() if (want_cvj or want_prod) else (None,) * 3
    finally:
        prop.records(False)
    if want_prod and cvjprod is None:
        raise ValueError("iprod = 2: the context has no product tables attached")
    if want_cvj:
        write_step_records(os.path.join(dirpath, "Cvj.%04d" % iloop), its, cvj)
    if want_prod:
        write_step_records(os.path.join(dirpath, "Cvjprod.%04d" % iloop), its, cvjprod)
    write_cont(dirpath, prop, iloop, indt, idproc)
    return norms

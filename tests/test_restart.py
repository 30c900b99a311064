"""Continuation files (madwave3_amd/restart.py; wvpcont and the start-up
reads, SRC/main_madwave3.f:223-261, 494-552): round trip through the
files, the record framing, and a Fortran program (tests/fortran/read_cont.f90,
built here with amdflang) that reads them with the reference's whole-array
unformatted reads and list-directed cont.data read.  The per-step record
files Cvj.LLLL / Cvjprod.LLLL (main_madwave3.f:311-365) the same way, with
tests/fortran/read_cvj.f90 reading them as "read(15) it, Cvj"."""
import os
import shutil
import struct
import subprocess

import numpy as np
import pytest

from madwave3_amd import restart
from tests.cases import case

HERE = os.path.dirname(os.path.abspath(__file__))
FLANG = os.environ.get("FLANG", "/opt/rocm/llvm/bin/amdflang")


class FakeProp:
    """The getters and putters of a Propagator over synthetic arrays."""

    def __init__(self, c, nflux, seed):
        rng = np.random.default_rng(seed)
        self.case, self.nflux = c, nflux
        nv, nj = c.nvmax - c.nvini + 1, c.jmax - c.jini + 1

        def z(*shape):
            return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
        self.s = (rng.standard_normal(c.ntot), rng.standard_normal(c.ntot), 137)
        self.a = (z(nflux, c.netot), z(nflux, c.netot))
        self.z2 = z(c.ncan, nj, nv, c.netot)
        self.zp = None
        if getattr(c, "iprod", 0) > 1:
            self.zp = z(c.iommaxprod - c.iomminprod + 1, c.jmaxprod - c.jiniprod + 1,
                        c.nvmaxprod - c.nviniprod + 1, c.netot)
        self.got = {}

    def state(self):
        return self.s

    def flux_accum(self):
        return self.a

    def zs2(self):
        return self.z2

    def zs2prod(self):
        return self.zp

    def put_state(self, psim1, psi, k, psi0=None):
        self.got["state"] = (psim1, psi, k, psi0)

    def put_accum(self, zcr, zcdr, zs2):
        self.got["accum"] = (zcr, zcdr, zs2)

    def put_zs2prod(self, z):
        self.got["zs2prod"] = z

    # the recording interface of a Propagator over synthetic records
    def records(self, on=True):
        self.got.setdefault("records", []).append(bool(on))

    def steps(self, n):
        c = self.case
        rng = np.random.default_rng(n)
        nv, nj = c.nvmax - c.nvini + 1, c.jmax - c.jini + 1
        k0 = self.s[2]
        self.its = np.arange(k0 + 1, k0 + 1 + n, dtype=np.int32)
        self.cvj = rng.standard_normal((n, c.ncan, nj, nv))
        self.cvjp = None if self.zp is None else rng.standard_normal((n,) + self.zp.shape[:3])
        self.s = (self.s[0], self.s[1], k0 + n)
        return np.ones(n), None

    def take_records(self):
        return self.its, self.cvj, self.cvjp


@pytest.mark.parametrize("name", ["j0", "prod"])
def test_continuation_files_round_trip(tmp_path, name):
    c = case(name)
    p = FakeProp(c, 37, 5)
    k = restart.write_cont(str(tmp_path), p, iloop=3, indt=1)
    assert k == 137
    d = restart.read_cont(str(tmp_path), c, 37)
    assert (d["it0"], d["indt"], d["iloop"]) == (137, 2, 3)
    assert np.array_equal(d["psim1"], p.s[0]) and np.array_equal(d["psi"], p.s[1])
    assert np.array_equal(d["zcr"], p.a[0]) and np.array_equal(d["zcdr"], p.a[1])
    assert np.array_equal(d["zs2"], p.z2)
    if p.zp is not None:
        assert np.array_equal(d["zs2prod"], p.zp)
    q = FakeProp(c, 37, 6)
    assert restart.resume(q, str(tmp_path)) == (137, 2, 3)
    pm, ps, kk, p0 = q.got["state"]
    assert kk == 137 and np.array_equal(ps, p.s[1]) and p0 is c.rpaq0
    assert np.array_equal(q.got["accum"][2], p.z2)
    # framing: 4-byte count, data, count
    raw = open(os.path.join(str(tmp_path), "cont00.paq"), "rb").read()
    assert struct.unpack_from("<i", raw, 0)[0] == 8 * c.ntot
    assert struct.unpack_from("<i", raw, 4 + 8 * c.ntot)[0] == 8 * c.ntot
    assert len(raw) == 2 * (8 + 8 * c.ntot)
    with pytest.raises(ValueError):
        restart.read_cont(str(tmp_path), c, 36)


@pytest.mark.skipif(not os.path.exists(FLANG), reason="amdflang not installed")
def test_fortran_reads_the_continuation_files(tmp_path):
    c = case("prod")
    p = FakeProp(c, 23, 11)
    d = str(tmp_path)
    restart.write_cont(d, p, iloop=4)
    nv, nj = c.nvmax - c.nvini + 1, c.jmax - c.jini + 1
    ns2 = c.ncan * nj * nv
    nprod = p.zp.size // c.netot
    with open(os.path.join(d, "dims.txt"), "w") as f:
        f.write("%d %d %d %d %d\n" % (c.ntot, c.netot, ns2, 23, nprod))
    exe = os.path.join(d, "read_cont")
    subprocess.run([FLANG, "-O1", "-o", exe, os.path.join(HERE, "fortran", "read_cont.f90")],
                   check=True, cwd=d, capture_output=True)
    out = subprocess.run([exe], check=True, cwd=d, capture_output=True, text=True).stdout.split()
    ints = [int(x) for x in out[:5]]
    assert ints == [137, 1, 4, 23, c.ncan]
    v = [float(x) for x in out[5:]]
    want = [p.s[0].sum(), p.s[1].sum(), p.z2.real.sum(), p.z2.imag.sum(), p.a[0].real.sum(),
            p.a[1].imag.sum(), p.z2.reshape(-1, c.netot)[-1, -1].real, p.a[0][-1, 0].imag,
            p.zp.real.sum(), p.zp.imag.sum()]
    for a, b in zip(v, want):
        assert a == pytest.approx(b, rel=1e-12, abs=1e-12)


def test_step_records_round_trip_and_framing(tmp_path):
    rng = np.random.default_rng(3)
    its = np.arange(5, 12, dtype=np.int32)
    v = rng.standard_normal((7, 3, 4, 5))
    path = os.path.join(str(tmp_path), "Cvj.0002")
    restart.write_step_records(path, its, v)
    i2, v2 = restart.read_step_records(path, (3, 4, 5))
    assert np.array_equal(i2, its) and np.array_equal(v2, v)
    raw = open(path, "rb").read()
    n = 4 + 8 * 60
    assert len(raw) == 7 * (8 + n)
    assert struct.unpack_from("<ii", raw, 0) == (n, 5)
    assert struct.unpack_from("<d", raw, 8)[0] == v[0, 0, 0, 0]
    assert struct.unpack_from("<i", raw, 4 + n)[0] == n
    with pytest.raises(FileExistsError):       # status='new'
        restart.write_step_records(path, its, v)
    with pytest.raises(ValueError):
        restart.read_step_records(path, (3, 4, 4))


def test_record_kinds_follow_the_main_program():
    assert restart.record_kinds(case("photo")) == (True, False)
    assert restart.record_kinds(case("prod")) == (False, True)
    c = case("odd")
    assert restart.record_kinds(c) == (False, False)


@pytest.mark.parametrize("name", ["prod", "photo"])
def test_time_loop_writes_records_and_continuation(tmp_path, name):
    c = case(name)
    p = FakeProp(c, 11, 7)
    d = str(tmp_path)
    norms = restart.time_loop(p, d, iloop=3, ntimes=9)
    assert len(norms) == 9 and p.got["records"] == [True, False]
    nv, nj = c.nvmax - c.nvini + 1, c.jmax - c.jini + 1
    if name == "photo":
        its, v = restart.read_step_records(os.path.join(d, "Cvj.0003"), (c.ncan, nj, nv))
        assert np.array_equal(its, np.arange(138, 147)) and np.array_equal(v, p.cvj)
        assert not os.path.exists(os.path.join(d, "Cvjprod.0003"))
    else:
        its, v = restart.read_step_records(os.path.join(d, "Cvjprod.0003"), p.cvjp.shape[1:])
        assert np.array_equal(its, np.arange(138, 147)) and np.array_equal(v, p.cvjp)
        assert not os.path.exists(os.path.join(d, "Cvj.0003"))
    assert restart.read_cont(d, c, 11)["it0"] == 146
    with pytest.raises(FileExistsError):       # the loop's files exist
        restart.time_loop(p, d, iloop=3, ntimes=2)


@pytest.mark.skipif(not os.path.exists(FLANG), reason="amdflang not installed")
def test_fortran_reads_the_step_records(tmp_path):
    rng = np.random.default_rng(9)
    nch, nj, nv = 3, 5, 4
    its = np.array([41, 42, 43], dtype=np.int32)
    v = rng.standard_normal((3, nch, nj, nv))       # C order of Cvj(nv, nj, nch)
    d = str(tmp_path)
    restart.write_step_records(os.path.join(d, "Cvj.0001"), its, v)
    with open(os.path.join(d, "dims.txt"), "w") as f:
        f.write("%d %d %d Cvj.0001\n" % (nv, nj, nch))
    exe = os.path.join(d, "read_cvj")
    subprocess.run([FLANG, "-O1", "-o", exe, os.path.join(HERE, "fortran", "read_cvj.f90")],
                   check=True, cwd=d, capture_output=True)
    out = subprocess.run([exe], check=True, cwd=d, capture_output=True, text=True).stdout
    rows = [l.split() for l in out.strip().splitlines()]
    assert [int(r[0]) for r in rows] == [41, 42, 43]
    for r, x in zip(rows, v):
        got = [float(t) for t in r[1:]]
        want = [x.sum(), x[0, 0, 0], x[0, 0, 1], x[0, 1, 0], x[-1, -1, -1]]
        for a, b in zip(got, want):
            assert a == pytest.approx(b, rel=1e-12, abs=1e-12)

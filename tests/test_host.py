"""Host logic and the C ABI boundary (CPU only; no device calls)."""
import ctypes as ct
import os
import re
import subprocess
import tempfile

import numpy as np
import pytest

from madwave3_amd import leps, model, refio
from madwave3_amd import namelist as NL
from madwave3_amd.lib import LIB_PATH, Mw3Setup, load_library
from tests.cases import case

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "mw3.h")


def _declared():
    txt = open(HEADER).read()
    return re.findall(r"MW3_API[^;]*?\b(mw3_\w+)\s*\(", txt)


def test_library_exports_every_declared_symbol():
    lib = load_library()
    names = _declared()
    assert len(names) >= 20
    for n in names:
        assert hasattr(lib, n), n
    assert lib.mw3_abi_version() == 1
    assert lib.mw3_kernel_name(0).decode() == "angular_fwd"
    assert isinstance(lib.mw3_last_error(None), bytes)


def test_setup_struct_layout_matches_header():
    fields = [f[0] for f in Mw3Setup._fields_]
    src = "#include <stddef.h>\n#include <stdio.h>\n#include \"%s\"\nint main(void){\n" % HEADER
    for f in fields:
        src += 'printf("%s %%zu\\n", offsetof(mw3_setup, %s));\n' % (f, f)
    src += 'printf("sizeof %zu\\n", sizeof(mw3_setup));\nreturn 0;}\n'
    with tempfile.TemporaryDirectory() as d:
        cfile = os.path.join(d, "layout.c")
        exe = os.path.join(d, "layout")
        open(cfile, "w").write(src)
        subprocess.run(["gcc", "-o", exe, cfile], check=True)
        out = subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split("\n")
    got = dict(l.split() for l in out if l)
    for f in fields:
        assert int(got[f]) == getattr(Mw3Setup, f).offset, f
    assert int(got["sizeof"]) == ct.sizeof(Mw3Setup)


def test_create_rejects_bad_setup_without_device_work():
    """Argument checks run before any device call; with no device in the
    container mw3_create fails cleanly either way."""
    lib = load_library()
    h = ct.c_void_p()
    rc = lib.mw3_create(None, 0, ct.byref(h))
    assert rc < 0 and not h.value
    assert lib.mw3_last_error(None)


def test_namelist_parse_and_roundtrip():
    txt = """
&inputflux
!----> data for total flux
r1flux_ang=4.d0
netot=500
ekinmin_eV=0.1d0,ekinmax_eV=1.5d0      ! comment
/
&inputprod iprod=2, jiniprod=0, jmaxprod=20  iomminprod=0!  , iommaxprod=0
 Rbalinprod=6.d0 /
&inputpotmass system='H+DH' xm0=2.014101779d0 /
"""
    g = NL.parse(txt)
    assert g["inputflux"] == {"r1flux_ang": 4.0, "netot": 500, "ekinmin_ev": 0.1, "ekinmax_ev": 1.5}
    assert g["inputprod"]["iprod"] == 2 and g["inputprod"]["rbalinprod"] == 6.0
    assert "iommaxprod" not in g["inputprod"]
    assert g["inputpotmass"]["system"] == "H+DH"
    back = NL.parse(NL.dump_group("inputflux", g["inputflux"]))
    assert back["inputflux"] == g["inputflux"]
    assert NL.group(g, "inputflux", {"ncontfile": 0})["ncontfile"] == 0


def test_reference_formats_roundtrip(tmp_path):
    p = leps.config(5, scale=12)
    p.update(iphoto=0)   # the collision packet (the photo files: tests/test_photo.py)
    text, pot, fdat = leps.make_inputs(p)
    model.write_case(str(tmp_path), text, pot, fdat)
    c1 = model.build_case(NL.parse(text), pot, fdat)
    c2 = model.load_case(str(tmp_path))
    for a in ("VVV", "indr1", "indr2", "npunreal", "Djmmp", "fd", "rpaq0", "etotS2", "reacfct"):
        x, y = getattr(c1, a), getattr(c2, a)
        assert np.array_equal(x, y) or np.max(np.abs(x - y)) <= 1e-15 * max(np.max(np.abs(x)), 1), a


def test_a0_layout_and_point_decode():
    c = case("coriol")
    sl, ch, ir, i1, i2 = model.point_index(c)
    assert len(sl) == c.ntot
    # indiproc order: slots outer, channels, points (SRC/mod_pot_01y2.f:435-455)
    assert np.all(np.diff(sl) >= 0)
    assert c.indangreal[0] == c.nangu and c.indangreal[-1] == 1
    for s in range(c.nangu):
        ia = c.indangreal[s] - 1
        for q in range(c.ncan):
            o = c.off[s, q]
            n = c.npunreal[ia]
            assert np.all(ch[o:o + n] == q) and np.all(ir[o:o + n] == np.arange(n))
            # R outer / r inner within an angle
            key = (i2[o:o + n].astype(np.int64) << 20) + i1[o:o + n]
            assert np.all(np.diff(key) > 0)


def test_leps_surface_limits_and_diatomic_states():
    pes = leps.LEPS()
    r = np.linspace(0.8, 3.0, 9)
    far = pes(r, np.full_like(r, 60.0), np.full_like(r, 61.0))
    assert np.max(np.abs(far - pes.morse(r))) < 1e-12
    assert abs(pes.morse(np.array([pes.re]))[0] + pes.D) < 1e-15
    a, b, cc = 1.3, 2.1, 2.9
    assert abs(pes(np.array(a), np.array(b), np.array(cc)) - pes(np.array(cc), np.array(a), np.array(b))) < 1e-15
    c = case("j0")
    f = c.fd[0, 0, :, :]
    nb = c.noBCstates[0, 0] - c.nvini + 1
    G = f[:nb] @ f[:nb].T
    assert np.max(np.abs(G - np.eye(nb))) < 1e-12
    assert abs(c.ediat[0, c.jref - c.jini, c.nvref - c.nvini]) < 1e-12   # eref subtracted


def test_initial_packet_matches_colini_form():
    """rpaq0 = -g(R) D(gamma, jref) fd(r) sqrt(ah2) on the kept points
    (SRC/mod_colini_01y2.f:143-170)."""
    c = case("j0")
    sl, ch, ir, i1, i2 = model.point_index(c)
    from madwave3_amd.numerics import rgauscolini
    R = c.rmis2 + (i2 - 1) * c.ah2
    g = rgauscolini(R, c.Rcolini, c.alpha0, c.xk0, c.factor0)
    D = c.Djmmp[0, 0, c.jref, c.indangreal[sl] - 1]
    fd = c.fd[0, c.jref - c.jini, c.nvref - c.nvini, i1 - 1]
    assert np.max(np.abs(c.rpaq0 - (-g * D * fd * np.sqrt(c.ah2)))) < 1e-15


FLANG = "/opt/rocm/llvm/bin/amdflang"
FMODDIR = os.path.join(ROOT, "madwave3_amd", "_build", "fortran")


@pytest.mark.skipif(not os.path.exists(os.path.join(FMODDIR, "mw3_capi.mod")),
                    reason="Fortran module not built (run __graft_entry__.build())")
def test_fortran_setup_type_layout_matches_header():
    """type(mw3_setup), bind(C) of module mw3_capi has the C struct's member
    offsets and size (the ctypes mirror is checked against gcc above)."""
    fields = [f[0] for f in Mw3Setup._fields_]
    src = ["program layout", "  use, intrinsic :: iso_c_binding", "  use mw3_capi",
           "  implicit none", "  type(mw3_setup), target :: s", "  integer(c_intptr_t) :: b",
           "  b = transfer(c_loc(s), b)"]
    for f in fields:
        src.append("  print '(a,1x,i0)', '%s', transfer(c_loc(s%%%s), b) - b" % (f, f))
    src.append("  print '(a,1x,i0)', 'sizeof', c_sizeof(s)")
    src.append("end program layout")
    with tempfile.TemporaryDirectory() as d:
        f90 = os.path.join(d, "layout.f90")
        exe = os.path.join(d, "layout")
        open(f90, "w").write("\n".join(src) + "\n")
        # the program uses only the derived type: no library symbols to link
        subprocess.run([FLANG, "-I", FMODDIR, "-o", exe, f90], check=True)
        out = subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split("\n")
    got = dict(l.split() for l in out if l.strip())
    for f in fields:
        assert int(got[f]) == getattr(Mw3Setup, f).offset, f
    assert int(got["sizeof"]) == ct.sizeof(Mw3Setup)


@pytest.mark.skipif(not os.path.exists(os.path.join(ROOT, "madwave3_amd", "fortran", "mw3_fdrive")),
                    reason="Fortran host not built (run __graft_entry__.build())")
def test_fortran_host_reads_case_file_and_reports_library_status(tmp_path):
    """The Fortran host reads a case file written by lib.write_case_file and
    calls mw3_create; without a device it stops with the library's message
    (the reference's write-and-stop convention)."""
    from madwave3_amd.lib import MW3_FP64, write_case_file

    c = case("j0")
    cf = str(tmp_path / "case.bin")
    rf = str(tmp_path / "res.bin")
    write_case_file(cf, c, c.rpaq0, 3, MW3_FP64)
    r = subprocess.run([os.path.join(ROOT, "madwave3_amd", "fortran", "mw3_fdrive"), cf, rf],
                       capture_output=True, text=True, timeout=120)
    if r.returncode != 0:
        assert "mw3_create: libmw3 status -3" in r.stdout, r.stdout + r.stderr
        assert "no HIP device" in r.stdout
    else:  # a device is visible: the run completed
        assert os.path.exists(rf)


@pytest.mark.skipif(not os.path.exists("/root/reference/SRC/main_madwave3.f"),
                    reason="reference tree not present")
def test_fortran_dropin_compiles_with_the_reference_main():
    """The replacement mod_Hphi_01y2 and mod_mw3_cheby compile in the reference's
    module chain, and the reference's unmodified main_madwave3.f and the modules
    that use mod_Hphi_01y2 compile against them (compile only; oracle/Makefile
    target shimcheck)."""
    r = subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "shimcheck"],
                       capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "main_madwave3.f compiled" in r.stdout


def test_auto_file_roundtrip(tmp_path):
    """auto file of check / cheby-spectra: header `xnorm ebndini emindlt delta2`,
    then `k auto(k)`; auto(0) = 1 and unwritten k read as 0."""
    from madwave3_amd import spectra

    a = np.array([0.5, -0.25, 0.125])
    p = str(tmp_path / "auto")
    spectra.write_auto(p, 2.0, -0.1, 0.7, 1.9, a, k0=2)
    xn, eb, em, de, auto = spectra.read_auto(p)
    assert (xn, eb, em, de) == (2.0, -0.1, 0.7, 1.9)
    assert np.array_equal(auto, [1.0, 0.0, 0.5, -0.25, 0.125])
    _, _, _, _, cut = spectra.read_auto(p, kmax=3)
    assert np.array_equal(cut, [1.0, 0.0, 0.5, -0.25])


def test_printprob_number_format_matches_golden_table():
    """The reference's S2prod table (EXAMPLES/HD+H_collision/J000-check) is in
    (1x,e15.7) fields: re-formatting its numbers gives its text back."""
    from madwave3_amd.printprob import fortran_e

    lines = open(os.path.join(ROOT, "tests", "golden", "hd_h_j000_S2prod.txt")).read().splitlines()
    for ln in lines[:50]:
        vals = [float(x) for x in ln.split()]
        assert "".join(" " + fortran_e(v, 15, 7, 2) for v in vals) == ln.rstrip("\n")


@pytest.mark.parametrize("name", ["states", "j0"])
def test_printprob_tables_roundtrip(tmp_path, name):
    """S2prod / S2prodelec columns: E(eV), P, P + S2reac, then reacfct
    (iprod = 0) or S2no and the product vibrational distribution (iprod = 2,
    also in distriS2prod.elec01); values below 1e-90 written as 0
    (SRC/main_madwave3.f:669-760)."""
    from madwave3_amd import printprob

    c = case(name)
    rng = np.random.default_rng(1)
    P = rng.random(c.netot) * 0.5
    P[0] = 1e-95
    Pe = [P.copy() for _ in range(c.nelecmax)]
    nj, nv = c.jmax - c.jini + 1, c.nvmax - c.nvini + 1
    zs2 = rng.standard_normal((c.ncan, nj, nv, c.netot)) + 1j * rng.standard_normal((c.ncan, nj, nv, c.netot))
    zp = None
    if c.iprod == 2:
        sh = (c.iommaxprod - c.iomminprod + 1, c.jmaxprod - c.jiniprod + 1,
              c.nvmaxprod - c.nviniprod + 1, c.netot)
        zp = rng.standard_normal(sh) + 1j * rng.standard_normal(sh)
    paths = printprob.write_tables(c, str(tmp_path), 3, P, Pe, zs2, iwrt_reac_distri=1,
                                   zs2prod=zp)
    t = printprob.read_table(paths[0])
    _, S2reac = printprob.reactant_probabilities(c, zs2)
    assert os.path.basename(paths[0]) == "S2prod.v%02d.J%03d.k00003" % (c.nvref, c.Jtot)
    assert np.allclose(t[:, 0], np.asarray(c.etotS2) / 1.197e-4 / 8065.5, rtol=1e-6)
    assert t[0, 1] == 0.0
    assert np.allclose(t[1:, 1], P[1:], rtol=1e-6)
    assert np.allclose(t[1:, 2], P[1:] + S2reac[1:], rtol=1e-6)
    if c.iprod == 2:
        S2pro, S2no, vib = printprob.product_probabilities(c, zp)
        assert t.shape[1] == 4 + int(c.maxvibprod[0, 0]) - int(c.minvibprod[0, 0]) + 1
        assert np.allclose(t[:, 3], S2no, rtol=1e-6, atol=1e-300)
        assert np.allclose(t[:, 4:].sum(axis=1), S2no, rtol=1e-5)
        d = printprob.read_table([p for p in paths if p.endswith("distriS2prod.elec01")][0])
        assert np.allclose(d[:, 1:], t[:, 4:], rtol=1e-6, atol=1e-300)
        # S2pro = |zS2prod|^2 / (4 pi^2) S2prodfac, zero for states beyond noBCprod
        z = zp[0, 0, 0]
        assert np.allclose(S2pro[0, 0, 0], np.abs(z) ** 2 / (4 * np.pi ** 2) * c.S2prodfac[0, 0])
    else:
        assert np.allclose(t[:, 3], c.reacfct, rtol=1e-6, atol=1e-300)
    tot = printprob.read_table([p for p in paths if p.endswith("distriS2reac.elec01")][0])
    assert np.allclose(tot[:, 1], S2reac if c.nelecmax == 1 else tot[:, 1], rtol=1e-6, atol=1e-12)

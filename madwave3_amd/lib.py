"""ctypes binding of libmw3 (include/mw3.h).

The Python host hands the module state built by :mod:`madwave3_amd.model` to
the C ABI exactly as a Fortran caller would: arrays in the reference's Fortran
memory order, 1-based index values.  Every failing call raises Mw3Error with
the library's message; a missing library raises at load time (there is no
fallback path).
"""
import ctypes as ct
import os

import numpy as np

__all__ = ["Mw3Setup", "Mw3ProdSetup", "make_setup", "make_prod_setup", "load_library",
           "Propagator", "Mw3Error",
           "MW3_FP64", "MW3_MIXED", "MW3_OPT_GRAPHS", "MW3_OPT_RADIAL", "MW3_OPT_ANGULAR",
           "MW3_OPT_SIDE_STREAM", "MW3_OPT_RADIAL_GRID", "MW3_OPT_ANGULAR_GRID",
           "MW3_OPT_FP16_PRODUCTS", "OPTION_NAMES",
           "LIB_PATH"]

MW3_ABI_VERSION = 1
MW3_NKERNELS = 11  # include/mw3.h
MW3_FP64 = 0
MW3_MIXED = 1
# include/mw3.h execution options
MW3_OPT_GRAPHS, MW3_OPT_RADIAL, MW3_OPT_ANGULAR, MW3_OPT_SIDE_STREAM = 0, 1, 2, 3
MW3_OPT_RADIAL_GRID = 4
MW3_OPT_ANGULAR_GRID = 5
MW3_OPT_FP16_PRODUCTS = 6
MW3_NOPTIONS = 7
OPTION_NAMES = ("graphs", "radial", "angular", "side_stream", "radial_grid", "angular_grid",
                "fp16_products")
LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libmw3.so")

_pi = ct.POINTER(ct.c_int32)
_pd = ct.POINTER(ct.c_double)


class Mw3Setup(ct.Structure):
    """Mirror of struct mw3_setup (include/mw3.h)."""
    _fields_ = [
        ("abi_version", ct.c_int32),
        ("npun1", ct.c_int32), ("npun1min", ct.c_int32), ("npun2", ct.c_int32),
        ("nangu", ct.c_int32), ("inc", ct.c_int32),
        ("Jtot", ct.c_int32), ("iparity", ct.c_int32), ("iommin", ct.c_int32),
        ("iommax", ct.c_int32), ("j0", ct.c_int32),
        ("nelecmax", ct.c_int32),
        ("rmis1", ct.c_double), ("ah1", ct.c_double),
        ("rmis2", ct.c_double), ("ah2", ct.c_double),
        ("hbr", ct.c_double),
        ("xm1red", ct.c_double), ("xm2red", ct.c_double),
        ("radcutmax", ct.c_double), ("rotcutmax", ct.c_double),
        ("emindlt", ct.c_double), ("delta2", ct.c_double),
        ("proj_j2_up", ct.c_int32),
        ("iomdiat", _pi), ("iomatom", _pi),
        ("sigdiat", _pd), ("sigatom", _pd),
        ("npuntot", ct.c_int32),
        ("indangreal", _pi), ("npunreal", _pi), ("indr1", _pi), ("indr2", _pi),
        ("VVV", _pd),
        ("ncan", ct.c_int32),
        ("iombas", _pi), ("nelebas", _pi), ("nojbas", _pi), ("j00", _pi),
        ("jbas", _pi),
        ("ndj", ct.c_int32),
        ("Djmmp", _pd),
        ("absfr1", _pd), ("absfr2", _pd),
        ("ir2abs", ct.c_int32),
        ("netot", ct.c_int32),
        ("etots2", _pd), ("reacfct", _pd),
        ("r1flux", ct.c_double),
        ("nvini", ct.c_int32), ("nvmax", ct.c_int32), ("jini", ct.c_int32),
        ("jmax", ct.c_int32),
        ("noBCstates", _pi),
        ("fd", _pd),
        ("iphoto", ct.c_int32), ("iprod", ct.c_int32), ("kminelastic", ct.c_int32),
        ("icanref", ct.c_int32), ("nvref", ct.c_int32), ("jref", ct.c_int32),
        ("iomref", ct.c_int32),
        ("nofe", ct.c_int32),
        ("efe", _pd),
    ]


class Mw3ProdSetup(ct.Structure):
    """Mirror of struct mw3_prod_setup (include/mw3.h)."""
    _fields_ = [
        ("n2prod0", ct.c_int32), ("n2prod1", ct.c_int32),
        ("nangproj0", ct.c_int32), ("nangproj1", ct.c_int32),
        ("iomminprod", ct.c_int32), ("iommaxprod", ct.c_int32),
        ("nviniprod", ct.c_int32), ("nvmaxprod", ct.c_int32),
        ("jiniprod", ct.c_int32), ("jmaxprod", ct.c_int32),
        ("noBCprod", _pi),
        ("transRgjac", _pd), ("prodwf", _pd), ("dtrans", _pd),
    ]


def _arr(keep, a, dt):
    a = np.ascontiguousarray(a, dtype=dt)
    keep.append(a)
    return a.ctypes.data_as(_pi if dt == np.int32 else _pd)


def make_setup(c):
    """mw3_setup for a model.RefCase; returns (struct, keep-alive list)."""
    keep = []
    s = Mw3Setup()
    s.abi_version = MW3_ABI_VERSION
    for k in ("npun1", "npun1min", "npun2", "nangu", "inc", "Jtot", "iparity",
              "iommin", "iommax", "j0", "nelecmax", "npuntot", "ncan", "ndj",
              "ir2abs", "netot", "nvini", "nvmax", "jini", "jmax", "iphoto",
              "iprod", "kminelastic", "icanref", "nvref", "jref", "iomref", "nofe"):
        setattr(s, k, int(getattr(c, k)))
    from . import consts
    s.rmis1, s.ah1, s.rmis2, s.ah2 = c.rmis1, c.ah1, c.rmis2, c.ah2
    s.hbr = consts.hbr
    s.xm1red, s.xm2red = c.xm1red, c.xm2red
    s.radcutmax, s.rotcutmax = c.radcutmax, c.rotcutmax
    s.emindlt, s.delta2 = c.emindlt, c.delta2
    s.proj_j2_up = int(c.Proj_j2_Up)
    s.r1flux = c.r1flux
    s.iomdiat = _arr(keep, c.iomdiat, np.int32)
    s.iomatom = _arr(keep, c.iomatom, np.int32)
    s.sigdiat = _arr(keep, c.sigdiat, np.float64)
    s.sigatom = _arr(keep, c.sigatom, np.float64)
    s.indangreal = _arr(keep, c.indangreal, np.int32)
    s.npunreal = _arr(keep, c.npunreal, np.int32)
    s.indr1 = _arr(keep, c.indr1, np.int32)
    s.indr2 = _arr(keep, c.indr2, np.int32)
    s.VVV = _arr(keep, c.VVV, np.float64)
    s.iombas = _arr(keep, c.iombas, np.int32)
    s.nelebas = _arr(keep, c.nelebas, np.int32)
    s.nojbas = _arr(keep, c.nojbas, np.int32)
    s.j00 = _arr(keep, c.j00, np.int32)
    s.jbas = _arr(keep, c.jbas, np.int32)
    s.Djmmp = _arr(keep, c.Djmmp, np.float64)
    s.absfr1 = _arr(keep, c.absfr1, np.float64)
    s.absfr2 = _arr(keep, c.absfr2, np.float64)
    s.etots2 = _arr(keep, c.etotS2, np.float64)
    s.reacfct = _arr(keep, c.reacfct, np.float64)
    s.noBCstates = _arr(keep, c.noBCstates, np.int32)
    s.fd = _arr(keep, c.fd, np.float64)
    s.efe = _arr(keep, c.efe if len(c.efe) else np.zeros(1), np.float64)
    return s, keep


def make_prod_setup(c):
    """mw3_prod_setup from the product tables of a case built with iprod = 2
    (madwave3_amd.prodtrans); returns (struct, keep-alive list)."""
    keep = []
    p = Mw3ProdSetup()
    for k in ("n2prod0", "n2prod1", "nangproj0", "nangproj1", "iomminprod", "iommaxprod",
              "nviniprod", "nvmaxprod", "jiniprod", "jmaxprod"):
        setattr(p, k, int(getattr(c, k)))
    p.noBCprod = _arr(keep, c.noBCprod, np.int32)
    p.transRgjac = _arr(keep, c.transRgjac, np.float64)
    p.prodwf = _arr(keep, c.prodwf, np.float64)
    p.dtrans = _arr(keep, c.dtrans, np.float64)
    return p, keep


def has_prod_tables(c):
    """The case projects on product states (iprod = 2, npun1 > 1)."""
    return getattr(c, "iprod", 0) == 2 and hasattr(c, "transRgjac")


def prod_shape(c):
    """(nomb, nj', nv') of the product-state arrays of a case."""
    return (c.iommaxprod - c.iomminprod + 1, c.jmaxprod - c.jiniprod + 1,
            c.nvmaxprod - c.nviniprod + 1)


def write_case_file(path, case, psi0, nsteps, precision):
    """Write the case file read by the Fortran host madwave3_amd/fortran/mw3_fdrive
    (format in its header): the mw3_setup members in declaration order, pointer
    members as an int64 count and their elements, then psi0, nsteps and the
    precision, then the product tables (mw3_prod_setup) when the case has them."""
    def struct_bytes(st, keep, fields):
        by_addr = {a.ctypes.data: a for a in keep}
        out = []
        for name, typ in fields:
            v = getattr(st, name)
            if typ is ct.c_int32:
                out.append(np.int32(v).tobytes())
            elif typ is ct.c_double:
                out.append(np.float64(v).tobytes())
            else:
                addr = ct.cast(v, ct.c_void_p).value
                a = by_addr.get(addr) if addr else None
                if a is None:
                    out.append(np.int64(0).tobytes())
                else:
                    a = np.ascontiguousarray(a).ravel()
                    out.append(np.int64(a.size).tobytes())
                    out.append(a.astype(np.int32 if typ is _pi else np.float64).tobytes())
        return out

    s, keep = make_setup(case)
    out = [b"MW3D", np.int32(2).tobytes()] + struct_bytes(s, keep, Mw3Setup._fields_)
    psi0 = np.ascontiguousarray(psi0, dtype=np.float64)
    out.append(np.int64(psi0.size).tobytes())
    out.append(psi0.tobytes())
    out.append(np.int32(nsteps).tobytes())
    out.append(np.int32(precision).tobytes())
    if has_prod_tables(case):
        ps, pkeep = make_prod_setup(case)
        out.append(np.int32(1).tobytes())
        out += struct_bytes(ps, pkeep, Mw3ProdSetup._fields_)
    else:
        out.append(np.int32(0).tobytes())
    with open(path, "wb") as f:
        f.write(b"".join(out))


def read_fdrive_result(path):
    """(norms, P, zS2 interleaved, zS2prod interleaved or empty) from a
    mw3_fdrive result file."""
    b = open(path, "rb").read()
    o = 0
    ns = int(np.frombuffer(b, np.int32, 1, o)[0]); o += 4
    norms = np.frombuffer(b, np.float64, ns, o).copy(); o += 8 * ns
    ne = int(np.frombuffer(b, np.int32, 1, o)[0]); o += 4
    P = np.frombuffer(b, np.float64, ne, o).copy(); o += 8 * ne
    m = int(np.frombuffer(b, np.int64, 1, o)[0]); o += 8
    z = np.frombuffer(b, np.float64, m, o).copy(); o += 8 * m
    mp = int(np.frombuffer(b, np.int64, 1, o)[0]); o += 8
    zp = np.frombuffer(b, np.float64, mp, o).copy()
    return norms, P, z, zp


class Mw3Error(RuntimeError):
    pass


def cheby_spectrum(autoc, emindlt, delta2, hbr, xkgam, e_zot):
    """mw3_cheby_spectrum: the Chebyshev-sum spectrum of cheby-spectra on the GPU
    (autoc[0..kmax], energies in zots)."""
    lib = load_library()
    a = np.ascontiguousarray(autoc, dtype=np.float64)
    e = np.ascontiguousarray(e_zot, dtype=np.float64)
    out = np.zeros(e.size)
    rc = lib.mw3_cheby_spectrum(a.ctypes.data_as(ct.c_void_p), a.size - 1, emindlt, delta2, hbr,
                                xkgam, e.ctypes.data_as(ct.c_void_p), e.size,
                                out.ctypes.data_as(ct.c_void_p))
    if rc != 0:
        raise Mw3Error("mw3_cheby_spectrum: %s" % lib.mw3_last_error(None).decode())
    return out


def steps_batch(props, n):
    """mw3_cheby_steps_batch over Propagators (one GPU): n steps of each,
    interleaved; returns norms [len(props)][n]."""
    lib = props[0].lib
    hs = [p.h.value if isinstance(p.h, ct.c_void_p) else p.h for p in props]
    arr = (ct.c_void_p * len(props))(*hs)
    norms = np.zeros((len(props), n))
    rc = lib.mw3_cheby_steps_batch(arr, len(props), n, norms.ctypes.data_as(ct.c_void_p))
    if rc != 0:
        raise Mw3Error("mw3_cheby_steps_batch: %s" % lib.mw3_last_error(None).decode())
    return norms


_lib = None


def load_library(path=None):
    """Load libmw3.so; raises OSError when it is not built."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise OSError("libmw3.so is not built (%s); run __graft_entry__.build()" % p)
    lib = ct.CDLL(p)
    lib.mw3_abi_version.restype = ct.c_int32
    lib.mw3_last_error.restype = ct.c_char_p
    lib.mw3_last_error.argtypes = [ct.c_void_p]
    lib.mw3_create.argtypes = [ct.POINTER(Mw3Setup), ct.c_int32, ct.POINTER(ct.c_void_p)]
    lib.mw3_destroy.argtypes = [ct.c_void_p]
    lib.mw3_ntot.restype = ct.c_int64
    lib.mw3_ntot.argtypes = [ct.c_void_p]
    lib.mw3_nflux.restype = ct.c_int64
    lib.mw3_nflux.argtypes = [ct.c_void_p]
    lib.mw3_hphi.argtypes = [ct.c_void_p, _pd, _pd]
    lib.mw3_cheby_begin.argtypes = [ct.c_void_p, _pd, _pd]
    lib.mw3_cheby_steps.argtypes = [ct.c_void_p, ct.c_int32, _pd, _pd]
    lib.mw3_cheby_k.argtypes = [ct.c_void_p]
    lib.mw3_flux_result.argtypes = [ct.c_void_p, _pd, _pd]
    lib.mw3_zs2_get.argtypes = [ct.c_void_p, _pd]
    lib.mw3_flux_energy_test.argtypes = [ct.c_void_p, ct.c_int32]
    lib.mw3_records.argtypes = [ct.c_void_p, ct.c_int32]
    lib.mw3_records_get.argtypes = [ct.c_void_p, ct.c_int32, _pi, _pd, _pd, _pi]
    lib.mw3_flux_accum_get.argtypes = [ct.c_void_p, _pd, _pd]
    lib.mw3_flux_accum_put.argtypes = [ct.c_void_p, _pd, _pd, _pd]
    lib.mw3_psie_get.argtypes = [ct.c_void_p, _pd]
    lib.mw3_state_get.argtypes = [ct.c_void_p, _pd, _pd, _pi]
    lib.mw3_state_put.argtypes = [ct.c_void_p, _pd, _pd, ct.c_int32, _pd]
    lib.mw3_timing_enable.argtypes = [ct.c_void_p, ct.c_int32]
    lib.mw3_timing_get.argtypes = [ct.c_void_p, ct.c_int32, _pd, ct.POINTER(ct.c_int64)]
    lib.mw3_cheby_spectrum.restype = ct.c_int32
    lib.mw3_cheby_spectrum.argtypes = [ct.c_void_p, ct.c_int32, ct.c_double, ct.c_double,
                                       ct.c_double, ct.c_double, ct.c_void_p, ct.c_int32,
                                       ct.c_void_p]
    lib.mw3_autotune.argtypes = [ct.c_void_p, ct.c_int32, _pi, _pd]
    lib.mw3_set_option.argtypes = [ct.c_void_p, ct.c_int32, ct.c_int32]
    lib.mw3_get_option.argtypes = [ct.c_void_p, ct.c_int32, _pi]
    lib.mw3_cheby_spectrum_noalloc.restype = ct.c_int32
    lib.mw3_cheby_spectrum_noalloc.argtypes = [ct.c_void_p, ct.c_int32, ct.c_double, ct.c_double,
                                               ct.c_double, ct.c_double, ct.c_void_p, ct.c_int32,
                                               ct.c_void_p, ct.c_void_p]
    lib.mw3_cheby_steps_batch.restype = ct.c_int32
    lib.mw3_cheby_steps_batch.argtypes = [ct.POINTER(ct.c_void_p), ct.c_int32, ct.c_int32,
                                          ct.c_void_p]
    lib.mw3_prod_attach.argtypes = [ct.c_void_p, ct.POINTER(Mw3ProdSetup)]
    lib.mw3_zs2prod_get.argtypes = [ct.c_void_p, _pd]
    lib.mw3_zs2prod_put.argtypes = [ct.c_void_p, _pd]
    lib.mw3_cvjprod_get.argtypes = [ct.c_void_p, _pd]
    lib.mw3_kernel_name.restype = ct.c_char_p
    lib.mw3_kernel_name.argtypes = [ct.c_int32]
    if lib.mw3_abi_version() != MW3_ABI_VERSION:
        raise OSError("libmw3 ABI %d != %d" % (lib.mw3_abi_version(), MW3_ABI_VERSION))
    if path is None:
        _lib = lib
    return lib


def _dp(a):
    return a.ctypes.data_as(_pd) if a is not None else None


class Propagator:
    """One libmw3 context: the drop-in for mod_Hphi_01y2 + the recursion."""

    def __init__(self, case, precision=MW3_FP64, lib=None, prod=None):
        """prod: attach the case's product tables (default: when the case was
        built with iprod = 2, as the reference calls ini_transcoor)."""
        self.lib = lib or load_library()
        self.case = case
        self._setup, self._keep = make_setup(case)
        h = ct.c_void_p()
        rc = self.lib.mw3_create(ct.byref(self._setup), precision, ct.byref(h))
        if rc != 0:
            raise Mw3Error("mw3_create: %s" % self.lib.mw3_last_error(None).decode())
        self.h = h
        self.ntot = int(self.lib.mw3_ntot(h))
        self.nflux = int(self.lib.mw3_nflux(h))
        if prod is None:
            prod = has_prod_tables(case)
        if prod:
            self.prod_attach()

    def _check(self, rc, what):
        if rc != 0:
            raise Mw3Error("%s: %s" % (what, self.lib.mw3_last_error(self.h).decode()))

    def close(self):
        if getattr(self, "h", None):
            self.lib.mw3_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def hphi(self, psi):
        psi = np.ascontiguousarray(psi, dtype=np.float64)
        out = np.empty(self.ntot)
        self._check(self.lib.mw3_hphi(self.h, _dp(psi), _dp(out)), "mw3_hphi")
        return out

    def begin(self, psi0):
        psi0 = np.ascontiguousarray(psi0, dtype=np.float64)
        st = np.zeros(3)
        self._check(self.lib.mw3_cheby_begin(self.h, _dp(psi0), _dp(st)), "mw3_cheby_begin")
        return st

    def steps(self, n, want_norms=True, want_auto=False):
        norms = np.zeros(n) if want_norms else None
        au = np.zeros(n) if want_auto else None
        self._check(self.lib.mw3_cheby_steps(self.h, n, _dp(norms), _dp(au)), "mw3_cheby_steps")
        return norms, au

    @property
    def k(self):
        return int(self.lib.mw3_cheby_k(self.h))

    def flux_result(self):
        c = self.case
        tot = np.zeros(c.netot)
        el = np.zeros((c.nelecmax, c.netot))
        self._check(self.lib.mw3_flux_result(self.h, _dp(tot), _dp(el)), "mw3_flux_result")
        return tot, el

    def flux_energy_test(self, on):
        """Keep (1, default) or drop (0) the E > V(flux point) test of the
        flux sums (mw3_flux_energy_test)."""
        self._check(self.lib.mw3_flux_energy_test(self.h, int(on)), "mw3_flux_energy_test")

    def records(self, on=True):
        """Start (on) or stop recording the per-step Cvj / Cvjprod records
        (mw3_records); stopping drops what is held."""
        self._check(self.lib.mw3_records(self.h, int(bool(on))), "mw3_records")

    def take_records(self, maxrec=None):
        """The held records, oldest first, moved out of the context:
        (it[n], cvj[n, ncan, nj, nv], cvjprod[n, nomb, nj', nv'] or None)."""
        c = self.case
        big = 1 << 30 if maxrec is None else int(maxrec)
        shp = (c.ncan, c.jmax - c.jini + 1, c.nvmax - c.nvini + 1) if c.netot > 0 else (0,)
        pshape = getattr(self, "_prod_shape", None)
        its, cvs, cps = [], [], []
        while big > 0:
            m = min(big, 4096)
            it = np.zeros(m, dtype=np.int32)
            cv = np.zeros((m,) + shp)
            cp = np.zeros((m,) + pshape) if pshape else None
            n = ct.c_int32(0)
            self._check(self.lib.mw3_records_get(self.h, m, it.ctypes.data_as(_pi), _dp(cv),
                                                 _dp(cp), ct.byref(n)), "mw3_records_get")
            n = n.value
            its.append(it[:n])
            cvs.append(cv[:n])
            if cp is not None:
                cps.append(cp[:n])
            big -= n
            if n < m:
                break
        return (np.concatenate(its), np.concatenate(cvs),
                np.concatenate(cps) if pshape else None)

    def zs2(self):
        c = self.case
        z = np.zeros((c.ncan, c.jmax - c.jini + 1, c.nvmax - c.nvini + 1, c.netot, 2))
        self._check(self.lib.mw3_zs2_get(self.h, _dp(z)), "mw3_zs2_get")
        return z[..., 0] + 1j * z[..., 1]

    def flux_accum(self):
        a = np.zeros((self.nflux, self.case.netot, 2))
        b = np.zeros((self.nflux, self.case.netot, 2))
        self._check(self.lib.mw3_flux_accum_get(self.h, _dp(a), _dp(b)), "mw3_flux_accum_get")
        return a[..., 0] + 1j * a[..., 1], b[..., 0] + 1j * b[..., 1]

    def put_accum(self, zcr, zcdr, zs2):
        def inter(z):
            return None if z is None else np.ascontiguousarray(
                np.stack([np.real(z), np.imag(z)], -1), dtype=np.float64)
        a, b, s = inter(zcr), inter(zcdr), inter(zs2)
        self._check(self.lib.mw3_flux_accum_put(self.h, _dp(a), _dp(b), _dp(s)),
                    "mw3_flux_accum_put")

    def prod_attach(self):
        """Hand the case's product tables (iprod = 2) to the context."""
        ps, keep = make_prod_setup(self.case)
        self._check(self.lib.mw3_prod_attach(self.h, ct.byref(ps)), "mw3_prod_attach")
        self._prod_shape = prod_shape(self.case)

    def zs2prod(self):
        z = np.zeros(self._prod_shape + (self.case.netot, 2))
        self._check(self.lib.mw3_zs2prod_get(self.h, _dp(z)), "mw3_zs2prod_get")
        return z[..., 0] + 1j * z[..., 1]

    def put_zs2prod(self, z):
        a = np.ascontiguousarray(np.stack([np.real(z), np.imag(z)], -1), dtype=np.float64)
        self._check(self.lib.mw3_zs2prod_put(self.h, _dp(a)), "mw3_zs2prod_put")

    def cvjprod(self):
        v = np.zeros(self._prod_shape)
        self._check(self.lib.mw3_cvjprod_get(self.h, _dp(v)), "mw3_cvjprod_get")
        return v

    def psie(self):
        z = np.zeros((self.case.nofe, self.ntot, 2))
        self._check(self.lib.mw3_psie_get(self.h, _dp(z)), "mw3_psie_get")
        return z[..., 0] + 1j * z[..., 1]

    def state(self):
        a = np.zeros(self.ntot)
        b = np.zeros(self.ntot)
        k = ct.c_int32()
        self._check(self.lib.mw3_state_get(self.h, _dp(a), _dp(b), ct.byref(k)), "mw3_state_get")
        return a, b, k.value

    def put_state(self, psim1, psi, k, psi0=None):
        psim1 = np.ascontiguousarray(psim1, dtype=np.float64)
        psi = np.ascontiguousarray(psi, dtype=np.float64)
        p0 = None if psi0 is None else np.ascontiguousarray(psi0, dtype=np.float64)
        self._check(self.lib.mw3_state_put(self.h, _dp(psim1), _dp(psi), k, _dp(p0)),
                    "mw3_state_put")

    def set_option(self, option, value):
        self._check(self.lib.mw3_set_option(self.h, int(option), int(value)), "mw3_set_option")

    def options(self):
        """{name: value} of the execution options in effect."""
        return {n: self.get_option(i) for i, n in enumerate(OPTION_NAMES)}

    def autotune(self, steps=16):
        """mw3_autotune: times every option combination over `steps` trial
        steps and keeps the fastest; returns ({name: value}, ms per step)."""
        ch = (ct.c_int32 * MW3_NOPTIONS)()
        ms = ct.c_double()
        self._check(self.lib.mw3_autotune(self.h, int(steps), ch, ct.byref(ms)), "mw3_autotune")
        return dict(zip(OPTION_NAMES, list(ch))), ms.value

    def get_option(self, option):
        v = ct.c_int32()
        self._check(self.lib.mw3_get_option(self.h, int(option), ct.byref(v)), "mw3_get_option")
        return v.value

    def timing(self, on=True, every=1):
        """HIP-event timing of the kernels; every > 1 times every N-th step only."""
        n = (max(int(every), 1) if on else 0)
        self._check(self.lib.mw3_timing_enable(self.h, n), "mw3_timing_enable")

    def kernel_times(self):
        out = {}
        for i in range(MW3_NKERNELS):
            t = ct.c_double()
            n = ct.c_int64()
            self._check(self.lib.mw3_timing_get(self.h, i, ct.byref(t), ct.byref(n)), "mw3_timing_get")
            if n.value:
                out[self.lib.mw3_kernel_name(i).decode()] = (t.value, n.value)
        return out

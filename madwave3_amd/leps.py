"""Synthetic MADWAVE3 inputs: a LEPS triatomic surface and its diatomic states.

Plays the part of the reference's preparation program (SRC/main_potini.f:1-108
with pot1, SRC/mod_pot_01y2.f:147-341, and radial_functions01_write,
SRC/mod_baseYfunciones_01y2.f:553-835) for an analytic surface, and produces
the same pot/ and func/ content (see refio).

Surface: London-Eyring-Polanyi-Sato form with the Sato parameter s, built from
a Morse singlet E1 = D (e^{-2a x} - 2 e^{-a x}) and an anti-Morse triplet
E3 = D/2 (e^{-2a x} + 2 e^{-a x}) (x = r - re) for each of the three pairs:
    Q_i = D/4 [(3 + s) e^{-2 a x} - (2 + 6 s) e^{-a x}]
    J_i = D/4 [(1 + 3 s) e^{-2 a x} - (6 + 2 s) e^{-a x}]
    V = [Q_1 + Q_2 + Q_3 - sqrt(J_1^2 + J_2^2 + J_3^2 - J_1 J_2 - J_2 J_3
         - J_3 J_1)] / (1 + s)
which tends to the Morse curve -D at re when the third atom is far away.

Diatomic states: eigenvectors of the grid Hamiltonian of the 0-1 diatom on the
r grid with the sine-transform kinetic operator of the propagator
(p_k = pi k/((npun1+1) ah1)), normalised to sum fd^2 = 1 as the reference
normalises them (SRC/mod_baseYfunciones_01y2.f:733-747).
"""
import math

import numpy as np

from . import consts as C
from . import namelist as NL
from .refio import FuncData, PotData, ProdFuncData

__all__ = ["LEPS", "diatomic_states", "product_states", "make_inputs", "config"]


class LEPS:
    """LEPS surface of three Morse pairs (hartree, bohr): the 0-1 and 0-2 pairs
    share De, a, re; the 1-2 pair has the same a, re and its own well depth
    (De12_eV, default De_eV: the symmetric surface).  A shallow 1-2 well
    closes the 1-2 + 0 arrangement."""

    def __init__(self, De_eV=4.75, a_invA=1.94, re_A=0.742, sato=0.17, De12_eV=None):
        self.D = De_eV / C.au2eV
        self.D12 = (De_eV if De12_eV is None else De12_eV) / C.au2eV
        self.a = a_invA * C.convl          # 1/bohr
        self.re = re_A / C.convl           # bohr
        self.s = sato

    def morse(self, r):
        e = np.exp(-self.a * (r - self.re))
        return self.D * (e * e - 2.0 * e)

    def _qj(self, r, D):
        e = np.exp(-self.a * (r - self.re))
        s = self.s
        q = 0.25 * D * ((3.0 + s) * e * e - (2.0 + 6.0 * s) * e)
        j = 0.25 * D * ((1.0 + 3.0 * s) * e * e - (6.0 + 2.0 * s) * e)
        return q, j

    def __call__(self, r01, r02, r12):
        q1, j1 = self._qj(r01, self.D)
        q2, j2 = self._qj(r02, self.D)
        q3, j3 = self._qj(r12, self.D12)
        x = j1 * j1 + j2 * j2 + j3 * j3 - j1 * j2 - j2 * j3 - j3 * j1
        return (q1 + q2 + q3 - np.sqrt(np.maximum(x, 0.0))) / (1.0 + self.s)


def jacobi_distances(r_bohr, R_bohr, cosg, xm0, xm1):
    """Atom-atom distances for reactant Jacobi (r = 0-1 bond, R from the 0-1
    centre of mass to atom 2), as in pot1 (SRC/mod_pot_01y2.f:196-209)."""
    gam = xm1 / (xm0 + xm1)
    r02 = np.sqrt(R_bohr ** 2 + gam ** 2 * r_bohr ** 2 + 2.0 * gam * R_bohr * r_bohr * cosg)
    g1 = 1.0 - gam
    r12 = np.sqrt(R_bohr ** 2 + g1 ** 2 * r_bohr ** 2 - 2.0 * g1 * R_bohr * r_bohr * cosg)
    return r_bohr, r02, r12


def diatomic_states(vdiat_zot, npun1, ah1, rmis1, mu, j, nstates):
    """Lowest nstates eigenpairs (zots, grid vectors) of
    -hbar^2/2mu d2/dr2 + V(r) + hbar^2 j(j+1)/(2 mu r^2) on the r grid."""
    n = npun1
    k = np.arange(1, n + 1)
    U = math.sqrt(2.0 / (n + 1)) * np.sin(np.pi * np.outer(k, k) / (n + 1))
    p = np.pi * k / ((n + 1) * ah1)
    T = U @ np.diag(0.5 * C.hbr * C.hbr * p * p / mu) @ U
    r = rmis1 + np.arange(n) * ah1
    H = T + np.diag(vdiat_zot + 0.5 * C.hbr * C.hbr * j * (j + 1) / (mu * r * r))
    e, v = np.linalg.eigh(H)
    e = e[:nstates]
    v = v[:, :nstates].copy()
    for i in range(v.shape[1]):
        if v[np.argmax(np.abs(v[:, i])), i] < 0:
            v[:, i] = -v[:, i]
        v[:, i] /= math.sqrt(np.sum(v[:, i] ** 2))
    return e, v


def product_states(p, pes):
    """func/prodwf content for the 0-2 diatom (one electronic state): its
    bound states below the dissociation limit on the R grid (n2prod1 =
    npun2 points), v = 0..nvmaxprod, j = 0..jmaxprod, as the reference's
    preparation program writes them for iprod > 0
    (SRC/mod_baseYfunciones_01y2.f:1083-1279)."""
    if p["nelec"] != 1:
        raise ValueError("product states are generated for one electronic state")
    conv = C.convl
    npun2 = p["npun2"]
    rmis2, rfin2 = p["rmis2_bohr"] * conv, p["rfin2_bohr"] * conv
    ah2 = (rfin2 - rmis2) / (npun2 - 1)
    mu02 = p["xm0"] * p["xm2"] / (p["xm0"] + p["xm2"])
    h2z = 1.0 / C.conve * C.conve1
    n2p1 = int(p.get("n2prod1", npun2))
    R = rmis2 + np.arange(n2p1) * ah2
    vd = pes.morse(R / conv) * h2z
    nobc, ediat, elec, vibel, fd = [], {}, {}, {}, {}
    for j in range(0, p["jmaxprod"] + 1):
        ev, v = diatomic_states(vd, n2p1, ah2, rmis2, mu02, j, p["nvmaxprod"] + 1)
        nb = -1
        for iv in range(0, p["nvmaxprod"] + 1):
            if ev[iv] < 0.0:
                ediat[(j, iv)] = ev[iv]
                elec[(j, iv)] = 1
                vibel[(j, iv)] = float(iv)
                fd[(j, iv)] = v[:, iv].reshape(n2p1, 1)
                nb = iv
        nobc.append(nb)
    return ProdFuncData(0, p["jmaxprod"], 0, n2p1, 1, nobc, ediat, elec, vibel, fd)


def config(cfg_id, scale=1.0, J=0):
    """Parameters of the benchmark configurations (grid sizes may be divided
    by `scale` for small test cases).  Configurations 1 and 2 resolve the
    products (iprod = 2: HD(v' <= 3, j' <= 15) at R' = 4 bohr).
    Configuration 4 is one total angular momentum J of the batched J = 0..23
    workload."""
    base = dict(
        xm0=2.01410, xm1=1.00783, xm2=1.00783,   # D, H, H (amu)
        De_eV=4.75, a_invA=1.94, re_A=0.742, sato=0.17,
        vcut_eV=2.5, radcut_eV=2.5, rotcut_eV=5.0,
        R0_bohr=9.0, width_bohr=0.3, ecol_eV=0.8,
        absorb_bohr=3.0, absalp=0.5, nexpo=3,
        ekinmin_eV=0.3, ekinmax_eV=1.5, netot=500,
        nvini=0, nvmax=3, jini=0, jmax=10, nvref=0, jref=0, iomref=0,
        Jtot=0, iparity=1, iommin=0, iommax=0, nelec=1,
        npun1min=32, kminelastic=100, rmis1_bohr=0.1, rmis2_bohr=0.1,
        iprod=0, nvmaxprod=3, jmaxprod=15, Rbalinprod_bohr=4.0,
    )
    base["Jtot"] = J
    if cfg_id == 1:
        base.update(npun1=128, rfin1_bohr=8.0, npun2=256, rfin2_bohr=15.0,
                    nangu=60, ntimes=500, nloop=39, r1flux_bohr=4.0, iprod=2)
    elif cfg_id == 2:
        base.update(npun1=256, rfin1_bohr=8.0, npun2=512, rfin2_bohr=15.0,
                    nangu=120, ntimes=500, nloop=79, r1flux_bohr=4.0, iprod=2)
    elif cfg_id == 5:
        # photodissociation (iphoto = 1): a seeded bound state of J = 1,
        # parity -1 on the ground surface times a linear transition dipole
        # to the three coupled states (photo_packet)
        base.update(npun1=256, rfin1_bohr=8.0, npun2=384, rfin2_bohr=15.0,
                    nangu=96, ntimes=500, nloop=119, r1flux_bohr=4.0, nelec=3,
                    state_offsets_eV=[0.0, 0.3, 0.6], coupling_eV=0.05,
                    coupling_r_bohr=1.4, coupling_width_bohr=0.3,
                    iphoto=1, Jtotini=1, iparini=-1, nvbound=0, bnd_R_bohr=4.5,
                    bnd_width_bohr=0.4, bnd_energy_eV=-0.5, bnd_jmax=2,
                    dip_strength=[1.0, 0.7, 0.4], dip_slope_r=0.5, dip_slope_Rcos=0.2,
                    seed=20261202)
    elif cfg_id == 3:
        # initial packet: a seeded superposition of the v = 0 packets of
        # j = 0, 1, 2 (every Omega <= j, as set_colini builds them with iomref < 0)
        base.update(npun1=192, rfin1_bohr=8.0, npun2=384, rfin2_bohr=15.0,
                    nangu=96, ntimes=500, nloop=59, r1flux_bohr=4.0,
                    Jtot=12, iparity=1, iommin=0, iommax=6, iomref=-1,
                    jmix=(0, 1, 2), seed=20261202)
    elif cfg_id == 4:
        # one total J of the batched workload J = 0..23 (Jtot set by the caller;
        # Omega = 0..min(J, 6); parity +1 has no Omega = 0 block at odd J)
        J = int(base.get("Jtot", 0))
        base.update(npun1=160, rfin1_bohr=8.0, npun2=320, rfin2_bohr=15.0,
                    nangu=80, ntimes=500, nloop=59, r1flux_bohr=4.0,
                    Jtot=J, iparity=1, iommin=J % 2, iommax=min(J, 6),
                    iomref=-1 if J else 0)
    else:
        raise ValueError("configuration %r has no generator yet" % (cfg_id,))
    if scale != 1.0:
        for k in ("npun1", "npun2", "nangu"):
            base[k] = max(8, int(round(base[k] / scale)))
        base["npun1min"] = min(base["npun1min"], base["npun1"] // 2)
        base["jmax"] = min(base["jmax"], base["nangu"] - 1)
        base["jmaxprod"] = min(base["jmaxprod"], base["nangu"] - 1)
    return base


def jmix_weights(p):
    """Seeded amplitudes of the initial rotational states p["jmix"]: standard
    normal draws from numpy's default generator seeded with p["seed"],
    normalised to unit sum of squares."""
    rng = np.random.default_rng(int(p["seed"]))
    w = rng.standard_normal(len(p["jmix"]))
    return dict(zip(p["jmix"], w / math.sqrt(float(np.sum(w * w)))))


def photo_packet(p, c):
    """Synthetic inputs of the photodissociation run (configuration 5): the
    bound state that dip_bndgrid reads from ../bndJxx{p,m} and the transition
    dipole of ../dip (see photo.py), as (BoundState, dipol).

    Bound state (J = Jtotini, Omega = 0, one electronic state): on the kept
    points of the propagation grid,
        fd_{v=0,j=0}(r) exp(-(R - R_b)^2 / (2 w^2)) sum_j a_j P~_j(cos gamma)
    with P~_j the normalised Legendre functions times sqrt(GL weight), the
    a_j (j = 0..bnd_jmax) seeded standard normal draws, normalised to 1.
    Transition dipole, linear in the coordinates (atomic units, along the
    body-fixed z axis, so only its q = 0 component is non-zero):
        d_e(r, R, gamma) = s_e (1 + a (r - r_e) + b R cos(gamma)),  r, R in bohr,
    s_e = dip_strength[e], a = dip_slope_r, b = dip_slope_Rcos."""
    from . import photo
    from .numerics import dwigner_vec

    rng = np.random.default_rng(int(p["seed"]) + 1)
    nj = int(p["bnd_jmax"]) + 1
    a = rng.standard_normal(nj)
    conv = C.convl
    gam = np.arccos(c.cgamma)
    leg = dwigner_vec(2 * (nj - 1), 0, 0, gam)[0::2]            # [j][iang]
    ang = np.zeros(c.nangu)
    for j in range(nj):
        ang += a[j] * leg[j] * math.sqrt(j + 0.5) * np.sqrt(c.wreal)
    fd0 = c.fd[0, 0 - c.jini, 0 - c.nvini]                      # v = 0, j = 0, state 1
    ir1s, ir2s, ias, vals, dipol = [], [], [], [], []
    re_b = p["re_A"] / conv
    for ia in range(c.nangu):
        n = int(c.npunreal[ia])
        i1 = c.indr1[ia, :n].astype(np.int64)
        i2 = c.indr2[ia, :n].astype(np.int64)
        r_b = (c.rmis1 + (i1 - 1) * c.ah1) / conv
        R_b = (c.rmis2 + (i2 - 1) * c.ah2) / conv
        v = fd0[i1 - 1] * np.exp(-((R_b - p["bnd_R_bohr"]) ** 2) / (2.0 * p["bnd_width_bohr"] ** 2))
        v = v * ang[ia]
        keep = np.abs(v) > 0.0
        ir1s.append(i1[keep])
        ir2s.append(i2[keep])
        ias.append(np.full(int(keep.sum()), ia + 1))
        vals.append(v[keep])
        lin = 1.0 + p["dip_slope_r"] * (r_b - re_b) + p["dip_slope_Rcos"] * R_b * c.cgamma[ia]
        d = np.zeros((n, 3, c.nelecmax, 1))
        for e in range(c.nelecmax):
            d[:, 1, e, 0] = p["dip_strength"][e] * lin
        dipol.append(d)
    val = np.concatenate(vals)
    val /= math.sqrt(float(np.sum(val * val)))
    npt = val.size
    grid = (c.rmis1, c.rfin1, c.npun1, c.rmis2, c.rfin2, c.npun2, c.nangu, c.inc)
    bs = photo.BoundState(np.concatenate(ir1s), np.concatenate(ir2s), np.concatenate(ias),
                          np.zeros(npt, dtype=np.int64), np.ones(npt, dtype=np.int64), val,
                          p["nvbound"], p["bnd_energy_eV"], grid)
    return bs, dipol


def build(p):
    """Case of parameter dict p (see config) with its initial packet: the
    collision packet of set_colini, a seeded superposition of them over
    p["jmix"], or the photodissociation packet of dip_bnd (iphoto > 0)."""
    from . import model
    from . import photo

    text, pot, fdat = make_inputs(p)
    c = model.build_case(NL.parse(text), pot, fdat)
    if int(p.get("iphoto", 0)) > 0:
        bs, dipol = photo_packet(p, c)
        photo.dip_bnd(c, bs, dipol, int(p["iphoto"]), int(p["Jtotini"]), int(p["iparini"]))
    elif p.get("jmix"):
        w = jmix_weights(p)
        c.rpaq0 = sum(wj * model.colini_packet(c, jref=j) for j, wj in w.items())
        c.norm_initial = float(np.sum(c.rpaq0 * c.rpaq0))
        c.jmix_weights = w
    return c


def make_inputs(p):
    """(input.dat text, PotData, FuncData) for parameter dict p (see config)."""
    conv = C.convl
    npun1, npun2, nangu = p["npun1"], p["npun2"], p["nangu"]
    rmis1, rfin1 = p["rmis1_bohr"] * conv, p["rfin1_bohr"] * conv
    rmis2, rfin2 = p["rmis2_bohr"] * conv, p["rfin2_bohr"] * conv
    ah1 = (rfin1 - rmis1) / (npun1 - 1)
    ah2 = (rfin2 - rmis2) / (npun2 - 1)
    xm0, xm1, xm2 = p["xm0"], p["xm1"], p["xm2"]
    mu1 = xm0 * xm1 / (xm0 + xm1)
    mu2 = xm2 * (xm0 + xm1) / (xm0 + xm1 + xm2)
    pes = LEPS(p["De_eV"], p["a_invA"], p["re_A"], p["sato"], p.get("De12_eV"))
    h2z = 1.0 / C.conve * C.conve1          # hartree -> zot
    vcut = C.ev_to_zot(p["vcut_eV"])

    # diatomic functions and reference energy
    r = rmis1 + np.arange(npun1) * ah1
    vd = pes.morse(r / conv) * h2z
    nj = p["jmax"] - p["jini"] + 1
    nel = p["nelec"]
    offs_e = [C.ev_to_zot(x) for x in p.get("state_offsets_eV", [0.0] * nel)]
    nobc = [[0] * nj for _ in range(nel)]
    ediat, fd = {}, {}
    for e in range(1, nel + 1):
        for j in range(p["jini"], p["jmax"] + 1):
            ev, v = diatomic_states(vd + offs_e[e - 1], npun1, ah1, rmis1, mu1, j, p["nvmax"] + 1)
            nb = -1
            for iv in range(p["nvini"], p["nvmax"] + 1):
                if ev[iv] < offs_e[e - 1]:    # bound below the state's dissociation limit
                    ediat[(e, j, iv)] = ev[iv]
                    fd[(e, j, iv)] = v[:, iv]
                    nb = iv
            nobc[e - 1][j - p["jini"]] = nb
    eref = ediat[(1, p["jref"], p["nvref"])]
    fdat = FuncData(nel, p["jini"], p["jmax"], p["nvini"], npun1, nobc, ediat, fd, eref)
    if p.get("iprod", 0) == 2:
        fdat.prod = product_states(p, pes)

    # potential on the grid, compressed as pot1 does (SRC/mod_pot_01y2.f:175-279)
    from .numerics import gauleg
    ne = p["nelec"]
    offs = [C.ev_to_zot(x) for x in p.get("state_offsets_eV", [0.0] * ne)]
    vc = C.ev_to_zot(p.get("coupling_eV", 0.0))
    rc = p.get("coupling_r_bohr", 1.4) * conv
    wc = p.get("coupling_width_bohr", 0.3) * conv
    _, cg = gauleg(nangu)
    R = rmis2 + np.arange(npun2) * ah2
    indr1, indr2, vmat = [], [], []
    vmin = 1e10
    vmaxtot = vcut
    for ia in range(nangu):
        rr, RR = np.meshgrid(r / conv, R / conv)            # [ir2][ir1]
        d01, d02, d12 = jacobi_distances(rr, RR, cg[ia], xm0, xm1)
        v0 = pes(d01, d02, d12) * h2z
        if ne == 1:
            v = v0 - eref
            v = np.where(v > vcut, vcut, v)
            vmin = min(vmin, float(v.min()))
            keep = v < vcut
            i2, i1 = np.nonzero(keep)                      # R outer, r inner
            vm = v[i2, i1].reshape(-1, 1, 1)
        else:
            # V[ir2, ir1, je, ie]: diabatic matrix, reference energy on the diagonal
            V = np.zeros(v0.shape + (ne, ne))
            cpl = vc * np.exp(-((rr * conv - rc) / wc) ** 2)
            for e in range(ne):
                V[..., e, e] = v0 + offs[e] - eref
                for f in range(ne):
                    if f != e:
                        V[..., e, f] = cpl
            # cut as pot1: a diagonal above vcut loses its couplings and is capped
            for e in range(ne):
                hi = V[..., e, e] > vcut
                for f in range(ne):
                    V[..., e, f] = np.where(hi, 0.0, V[..., e, f])
                    V[..., f, e] = np.where(hi, 0.0, V[..., f, e])
                V[..., e, e] = np.where(hi, vcut, V[..., e, e])
            ev = np.linalg.eigvalsh(V)
            vmin = min(vmin, float(ev.min()))
            vmaxtot = max(vmaxtot, float(ev.max()))
            keep = (ev < vcut).any(axis=-1)
            i2, i1 = np.nonzero(keep)
            vm = V[i2, i1]                                 # [n][je][ie] (symmetric)
        indr1.append((i1 + 1).astype(np.int32))
        indr2.append((i2 + 1).astype(np.int32))
        vmat.append(vm)
    pot = PotData(ne, [(0, 0, 1.0, 1.0)] * ne, vmin, vmaxtot, (rmis1, rfin1, npun1),
                  (rmis2, rfin2, npun2), nangu, 1, indr1, indr2, vmat)

    # namelists
    ecol = p["ecol_eV"]
    alpha = p["width_bohr"] * conv
    ecol_z = C.ev_to_zot(ecol)
    deltaE = math.sqrt(2.0 * ecol_z / mu2) * C.hbr / alpha / (C.eV2cm * C.conve1)
    absr1 = rfin1 - p["absorb_bohr"] * conv
    absr2 = rfin2 - p["absorb_bohr"] * conv
    text = ""
    text += NL.dump_group("INPUTGRIDBASE", dict(
        NPUN1=npun1, RMIS1=rmis1, RFIN1=rfin1, NPUN1MIN=p["npun1min"],
        NPUN2=npun2, RMIS2=rmis2, RFIN2=rfin2, NANGU=nangu,
        NELECMAX=p["nelec"], JTOT=p["Jtot"], IPARITY=p["iparity"],
        IOMMIN=p["iommin"], IOMMAX=p["iommax"], J0=0, INC=1,
        JINI=p["jini"], JMAX=p["jmax"], NVINI=p["nvini"], NVMAX=p["nvmax"],
        NVREF=p["nvref"], JREF=p["jref"], IOMREF=p["iomref"], IELECREF=1))
    text += NL.dump_group("inputprocess", dict(iphoto=int(p.get("iphoto", 0))))
    if int(p.get("iphoto", 0)) > 0:
        text += NL.dump_group("inputbnd", dict(
            Jtotini=p["Jtotini"], iparini=p["iparini"], nvbound=p["nvbound"],
            nprocbnd=int(p.get("nprocbnd", 1)),
            maxbnddim=npun1 * npun2 * nangu, igridbnd=1))
    text += NL.dump_group("inputabsorcion", dict(
        absalp1=p["absalp"], absr1=absr1, n1expo=p["nexpo"], ninterv1=1,
        absalp2=p["absalp"], absr2=absr2, n2expo=p["nexpo"], ninterv2=1))
    if p.get("iprod", 0) == 2:
        prod = dict(iprod=2, nviniprod=0, nvmaxprod=p["nvmaxprod"], jiniprod=0,
                    jmaxprod=p["jmaxprod"], iomminprod=0,
                    Rbalinprod=p["Rbalinprod_bohr"] * conv)
        for k in ("n2prod0", "n2prod1", "nangproj0", "nangproj1", "iommaxprod"):
            if k in p:
                prod[k] = p[k]
        text += NL.dump_group("inputprod", prod)
    else:
        text += NL.dump_group("inputprod", dict(iprod=0))
    text += NL.dump_group("inputtime", dict(ntimes=p["ntimes"], nloop=p["nloop"],
                                            kminelastic=p["kminelastic"]))
    text += NL.dump_group("inputpotmass", dict(
        system="LEPS", xm0=xm0, xm1=xm1, xm2=xm2,
        VcutmaxeV=p["vcut_eV"], radcutmaxeV=p["radcut_eV"], rotcutmaxeV=p["rotcut_eV"]))
    text += NL.dump_group("inputcol", dict(
        Rcolini_ang=p["R0_bohr"] * conv, ecol_ev=ecol, deltaE_ev=deltaE))
    text += NL.dump_group("inputflux", dict(
        r1flux_ang=p["r1flux_bohr"] * conv, netot=p["netot"],
        ekinmin_eV=p["ekinmin_eV"], ekinmax_eV=p["ekinmax_eV"], ncontfile=0))
    return text, pot, fdat

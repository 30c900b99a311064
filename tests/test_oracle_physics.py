"""Closed-form and brute-force checks of the oracle's operator and recursion
(CPU only).  Each check is chosen so that a plausible mistake -- a transposed
angular table, a wrong DST size or momentum, a dropped cap, a wrong sign or
index in the recursion -- fails it."""
import numpy as np
import pytest

from madwave3_amd import consts as C
from madwave3_amd import model
from oracle.oracle import Oracle
from tests.cases import box_case, case


def _dense(c, psi):
    """A0 vector -> dense [slot][ir1][ir2] array (channel 0)."""
    sl, ch, ir, i1, i2 = model.point_index(c)
    out = np.zeros((c.ncan, c.nangu, c.npun1, c.npun2))
    out[ch, sl, i1 - 1, i2 - 1] = psi
    return out


def _from_dense(c, d):
    sl, ch, ir, i1, i2 = model.point_index(c)
    return d[ch, sl, i1 - 1, i2 - 1]


def test_kinetic_sine_modes_are_eigenvectors():
    """Full grid, V = 0, J = 0: psi = D(g, j=0) sin(pi k1 i1/(n1+1)) sin(pi k2 i2/(n2+1))
    satisfies H psi = min(p_r^2/2mu1 + p_R^2/2mu2, radcut) psi (the j = 0
    angular terms vanish at J = 0)."""
    c = box_case()
    o = Oracle(c)
    n1, n2 = c.npun1, c.npun2
    assert (o.tables()["nopr1"] == n1).all()
    D0 = c.Djmmp[0, 0, 0, :]                      # D(iang, j=0)
    for k1, k2 in [(1, 1), (3, 5), (n1, n2), (n1 // 2, 2)]:
        g = D0[c.indangreal - 1][:, None, None]
        s1 = np.sin(np.pi * k1 * np.arange(1, n1 + 1) / (n1 + 1))[None, :, None]
        s2 = np.sin(np.pi * k2 * np.arange(1, n2 + 1) / (n2 + 1))[None, None, :]
        d = (g * s1 * s2)[None]
        psi = _from_dense(c, d)
        e1 = (np.pi * k1 / ((n1 + 1) * c.ah1)) ** 2 * C.hbr ** 2 / (2 * c.xm1red)
        e2 = (np.pi * k2 / ((n2 + 1) * c.ah2)) ** 2 * C.hbr ** 2 / (2 * c.xm2red)
        lam = min(e1 + e2, c.radcutmax)
        h = o.hphi(psi)
        assert np.max(np.abs(h - lam * psi)) < 1e-10 * lam * np.max(np.abs(psi))


def test_angular_tables_diagonal_in_fbr_at_J0():
    """xj2(r1) and xl2p(R, q=0) are D diag D^T: each Legendre column D(., j)
    is an eigenvector with eigenvalue min(hb^2 j(j+1)/(2 mu1 r^2), rotcut) and
    min(hb^2 j(j+1)/(2 mu2 R^2), rotcut) respectively."""
    c = box_case(na=8)
    t = Oracle(c).tables()
    xj2, xl2p = t["xj2"], t["xl2p"]
    slots = c.indangreal - 1
    for j in range(c.nangu):
        Dj = c.Djmmp[0, 0, j, :]
        for ir1 in (0, 3, c.npun1 - 1):
            r1 = c.rmis1 + ir1 * c.ah1
            K = xj2[0, :, :, ir1]          # [source slot][target iang]
            v = K.T @ Dj[slots]
            lam = min(C.hbr ** 2 * j * (j + 1) / (2 * c.xm1red * r1 * r1), c.rotcutmax)
            assert np.max(np.abs(v - lam * Dj)) < 1e-10 * max(lam, 1e-300) + 1e-12
        for ir2 in (0, 5, c.npun2 - 1):
            R = c.rmis2 + ir2 * c.ah2
            K = xl2p[1, 0, :, :, ir2]
            v = K.T @ Dj[slots]
            lam = min(C.hbr ** 2 * j * (j + 1) / (2 * c.xm2red * R * R), c.rotcutmax)
            assert np.max(np.abs(v - lam * Dj)) < 1e-10 * max(lam, 1e-300) + 1e-12


def test_coriolis_tables_map_columns_to_neighbour_columns():
    """J = 2, Omega = 0..2: xl2p(R, q) of source Omega_s maps D_s(., j) onto a
    multiple of D_t(., j), Omega_t = Omega_s - q, for every j (the Coriolis
    term is diagonal in j; a transposed or Omega-swapped table is not)."""
    c = box_case(na=8, Jtot=2, iommax=2)
    t = Oracle(c).tables()
    xl2p = t["xl2p"]
    slots = c.indangreal - 1
    found = 0
    for s in range(c.ncan):
        for q in (-1, 1):
            om_t = c.iombas[s] - q
            if om_t < 0 or om_t > 2:
                continue
            for j in range(max(c.iombas[s], om_t), c.nangu):
                Ds = c.Djmmp[c.iombas[s], 0, j, :]
                Dt = c.Djmmp[om_t, 0, j, :]
                K = xl2p[q + 1, s, :, :, c.npun2 // 2]
                v = K.T @ Ds[slots]
                a = v @ Dt / (Dt @ Dt)
                assert np.max(np.abs(v - a * Dt)) < 1e-10 * max(np.max(np.abs(v)), 1e-300)
                if abs(a) > 1e-12:
                    found += 1
    assert found > 0


def _hmatrix(o, n):
    H = np.zeros((n, n))
    for i in range(n):
        e = np.zeros(n)
        e[i] = 1.0
        H[:, i] = o.hphi(e)
    return H


def test_undamped_recursion_is_chebyshev_polynomial():
    """With no absorption, psi_k = T_k(H_norm) psi_0 (SRC/main_madwave3.f:
    231-339) for the spectrally scaled H_norm = (H - emindlt)/delta2."""
    c = box_case(n1=8, n2=10, na=4, damp=False, vzero=False)
    o = Oracle(c)
    n = c.ntot
    H = _hmatrix(o, n)
    assert np.max(np.abs(H - H.T)) < 1e-12 * np.max(np.abs(H))
    Hn = (H - c.emindlt * np.eye(n)) / c.delta2
    lam, V = np.linalg.eigh(Hn)
    assert lam.min() > -1.0 and lam.max() < 1.0
    psi0 = np.random.default_rng(3).standard_normal(n)
    o.begin(psi0)
    for k_target in (5, 20, 60):
        o.steps(k_target - o.state()[2])
        pm, p, k = o.state()
        th = np.arccos(lam)
        ref = V @ (np.cos(k * th) * (V.T @ psi0))
        refm = V @ (np.cos((k - 1) * th) * (V.T @ psi0))
        assert np.max(np.abs(p - ref)) < 1e-10 * np.max(np.abs(ref))
        assert np.max(np.abs(pm - refm)) < 1e-10 * np.max(np.abs(refm))


def test_radial_kinetic_past_the_sine_box_is_the_identity():
    """Rows ir1 > nopr1 of an angle lie outside its 2-D sine transform: raux
    keeps psi there and rHpaqrec adds it unchanged (SRC/mod_Hphi_01y2.f:768-831).
    With the kept rows 5..14 of 16, nopr1 = noptFFT(10) = 10.  At J = 0 and
    V = 0 the column of a point past row 10 has no entry off its (r1, R) line,
    and the line block is I + A with A the angular block, whose j = 0
    eigenvalue is 0, so its smallest eigenvalue is exactly 1.  Returning zero
    there, or the transform's value, or taking row nopr1 itself out of the
    transform fails it."""
    c = box_case(n1=16, n2=12, na=2, damp=False, vzero=True, r1range=(5, 14))
    o = Oracle(c)
    n = c.ntot
    H = _hmatrix(o, n)
    _, _, _, i1, i2 = model.point_index(c)
    nn1 = int(o.tables()["nopr1"][0])
    assert nn1 == 10 and i1.min() == 5 and i1.max() == 14
    for r1 in (nn1, nn1 + 1, 14):
        for r2 in (1, 6, 12):
            line = np.where((i1 == r1) & (i2 == r2))[0]
            off = np.setdiff1d(np.arange(n), line)
            coupling = np.max(np.abs(H[np.ix_(off, line)]))
            lo = np.linalg.eigvalsh(H[np.ix_(line, line)]).min()
            if r1 <= nn1:        # inside the transform: the kinetic term couples rows
                assert coupling > 1e-3 and abs(lo - 1.0) > 1e-2
            else:
                assert coupling == 0.0
                assert abs(lo - 1.0) < 1e-12


def test_autocorrelation_and_psie_are_chebyshev_sums():
    """Photodissociation outputs with no absorption: the autocorrelation
    <psi_0|psi_k> (SRC/main_madwave3.f:444-455) is sum_i c_i^2 cos(k theta_i)
    in the eigenbasis of H_norm (c = V^T psi_0, theta = arccos lambda), and
    psi(E) (funE, SRC/mod_PsiE_01y2.f:92-123, called for it = 2, 3, ... at
    SRC/main_madwave3.f:350-353) is V diag(sum_k C_k(E) cos(k theta)) c with the
    coefficients C_k(E) of CoefEcheby (pinned by its closed form).  Pairing a
    coefficient with the wrong step, a missing step or a wrong real/imaginary
    part fails it."""
    c = box_case(n1=8, n2=10, na=4, damp=False, vzero=False)
    c.iphoto = 1
    c.nofe = 3
    c.efe = c.emindlt + c.delta2 * np.array([-0.55, 0.1, 0.7])  # |E_s| < 1
    o = Oracle(c)
    n = c.ntot
    H = _hmatrix(o, n)
    Hn = (H - c.emindlt * np.eye(n)) / c.delta2
    lam, V = np.linalg.eigh(Hn)
    th = np.arccos(lam)
    psi0 = np.random.default_rng(5).standard_normal(n)
    cf = V.T @ psi0
    o.begin(psi0)
    nst = 40
    _, au, _ = o.steps(nst)
    ks = np.arange(2, 2 + nst)          # the steps' Chebyshev orders
    auto_ref = np.cos(np.outer(ks, th)) @ (cf * cf)
    assert np.max(np.abs(au - auto_ref)) < 1e-11 * np.max(np.abs(auto_ref))
    ze = o.psie()
    for ie, E in enumerate(c.efe):
        ck = np.array([o.coef_echeby(int(k), np.array([E]))[0] for k in ks])
        ref = V @ ((np.cos(np.outer(th, ks)) @ ck) * cf)
        assert np.max(np.abs(ze[ie] - ref)) < 1e-10 * np.max(np.abs(ref))


def test_damped_recursion_matches_matrix_iteration():
    """Damped update psi_{k+1} = f (2 H_norm psi_k - f psi_{k-1}) with
    f = absfr1(r) absfr2(R) decoded per point (SRC/main_madwave3.f:332-339)."""
    c = case("j0")
    o = Oracle(c)
    n = c.ntot
    H = _hmatrix(o, n)
    Hn = (H - c.emindlt * np.eye(n)) / c.delta2
    _, _, _, i1, i2 = model.point_index(c)
    f = c.absfr1[i1 - 1] * c.absfr2[i2 - 1]
    assert f.min() < 0.99            # the case does absorb
    psi0 = c.rpaq0
    pm, p = psi0, Hn @ psi0
    o.begin(psi0)
    norms, _, _ = o.steps(40)
    for s in range(40):
        pm, p = p, f * (2.0 * (Hn @ p) - f * pm)
        assert abs(norms[s] - p @ p) < 1e-11 * max(p @ p, 1e-30)
    assert np.max(np.abs(o.state()[1] - p)) < 1e-11 * np.max(np.abs(p))


def test_energy_projection_is_linear_in_the_steps():
    """zS2(E) = sum_k C_k(E) Cvj_k over the projected steps, with the
    reference state's k < kminelastic terms left out (Cvjflux_k)."""
    c = case("j0")
    o = Oracle(c)
    o.begin(c.rpaq0)
    nst = 12
    _, _, cvj = o.steps(nst, want_cvj=True)
    z = o.zs2()
    ks = np.arange(2, 2 + nst)
    ref = np.zeros_like(z)
    for s, k in enumerate(ks):
        ck = o.coef_echeby(int(k), c.etotS2)
        m = np.ones(cvj.shape[1:])
        if k < c.kminelastic:
            m[c.icanref - 1, c.jref - c.jini, c.nvref - c.nvini] = 0.0
        ref += (cvj[s] * m)[..., None] * ck[None, None, None, :]
    assert np.max(np.abs(z - ref)) < 1e-12 * max(np.max(np.abs(ref)), 1e-300)

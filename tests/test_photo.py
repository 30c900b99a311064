"""Photodissociation initial state (iphoto = 1, 2; madwave3_amd/photo.py)
against the reference's own 3j routine, closed forms of the rotational factors
and of the packet, and round trips through the reference's file formats
(SRC/mod_photoini_01y2.f:145-718)."""
import math
import os

import numpy as np
import pytest

from madwave3_amd import leps, model, photo
from madwave3_amd import namelist as NL
from oracle import refnum
from tests.cases import case


@pytest.mark.skipif(not refnum.available(), reason="oracle/_ref not built")
def test_tresj_matches_reference_build():
    for args in [(1, 1, 0, 0, 0, 0), (1, 1, 1, 1, -1, 0), (2, 1, 3, -1, 1, 0), (3, 1, 2, 2, -1, -1),
                 (5, 1, 4, 0, 1, -1), (6, 1, 6, 3, 0, -3), (12, 1, 11, 5, 1, -6),
                 (4, 2, 3, 1, 1, -2), (7, 1, 8, -2, 0, 2)]:
        a = photo.tresj(*args)
        b = refnum.tresj(*args)
        assert a == pytest.approx(b, rel=1e-12, abs=1e-15), args


def test_tresj_closed_forms_and_orthogonality():
    # (1 1 0; m -m 0) = (-1)^(1-m)/sqrt(3)
    for m in (-1, 0, 1):
        assert photo.tresj(1, 1, 0, m, -m, 0) == pytest.approx((-1) ** (1 - m) / math.sqrt(3.0))
    # (j j 1; m -m 0) = (-1)^(j-m) m / sqrt(j (j+1) (2j+1)) (Edmonds, Table 2);
    # exchanging the last two columns multiplies by (-1)^(2j+1) = -1
    for j in (1, 3, 6):
        for m in range(-j, j + 1):
            ref = -((-1) ** (j - m)) * m / math.sqrt(j * (j + 1) * (2 * j + 1))
            assert photo.tresj(j, 1, j, m, 0, -m) == pytest.approx(ref, abs=1e-15)
    # sum_{m1,m2} (2 j3 + 1) (j1 j2 j3; m1 m2 m3)(j1 j2 j3'; m1 m2 m3') = delta
    j1, j2 = 3, 2
    for j3 in range(1, 6):
        for j3p in range(1, 6):
            for m3 in range(-min(j3, j3p), min(j3, j3p) + 1):
                s = sum(photo.tresj(j1, j2, j3, m1, m2, m3) * photo.tresj(j1, j2, j3p, m1, m2, m3)
                        for m1 in range(-j1, j1 + 1) for m2 in range(-j2, j2 + 1))
                assert (2 * j3 + 1) * s == pytest.approx(1.0 if j3 == j3p else 0.0, abs=1e-13)
    assert photo.tresj(1, 1, 3, 0, 0, 0) == 0.0 and photo.tresj(2, 1, 2, 1, 1, 0) == 0.0


def test_factresj_parity_and_selection():
    # J = 0 from Jini = 1 (p_ini = -1): (Omega_ini, q, Omega) = (0, 0, 0) gives
    # (1/2) (1 - p_ini) (1 1 0; 0 0 0) = -1/sqrt(3); (1, -1, 0) gives
    # (-1)^(0-1) (1/sqrt 2) (1 1 0; 1 -1 0) = -1/sqrt(6); (1, +1, 0) the same
    # through the parity partner (1 1 0; -1 1 0) p_ini (-1)^Jini; nothing else
    t = photo.factresj_table(0, 1, 1, -1)
    assert t[0, 1, 0] == pytest.approx(-1.0 / math.sqrt(3.0), rel=1e-14)
    assert t[1, 0, 0] == pytest.approx(-1.0 / math.sqrt(6.0), rel=1e-14)
    assert t[1, 2, 0] == pytest.approx(-1.0 / math.sqrt(6.0), rel=1e-14)
    t[0, 1, 0] = t[1, 0, 0] = t[1, 2, 0] = 0.0
    assert np.max(np.abs(t)) == 0.0
    # the same parity in both states: every factor vanishes (forbidden)
    with pytest.raises(ValueError):
        photo.factresj_table(0, 1, 1, 1)
    with pytest.raises(ValueError):
        photo.factresj_table(0, 1, 0, -1)   # J = 0 <- Jini = 0


def test_photo_packet_is_dipole_times_bound_state():
    """With one q = 0 term (J = 0 from Jini = 1) the unnormalised packet is
    factresj(0,0,0) d_e(point) bnd(point) on every state-point."""
    c = case("photo")
    p = leps.config(5, scale=12)
    p.update(nvmax=2, jmax=4, netot=60)
    bs, dipol = leps.photo_packet(p, c)
    psi = photo.dip_bndgrid(c, bs, dipol, 1, photo.factresj_table(0, 1, 1, -1))
    sl, ch, ir, ir1, ir2 = model.point_index(c)
    ia = c.indangreal[sl] - 1
    bmap = {(a, b, g): v for a, b, g, v in zip(bs.ir1, bs.ir2, bs.iang, bs.value)}
    f0 = -1.0 / math.sqrt(3.0)
    for i in range(0, c.ntot, 7):
        b = bmap.get((ir1[i], ir2[i], ia[i] + 1), 0.0)
        e = int(c.nelebas[ch[i]]) - 1
        d = dipol[ia[i]][ir[i], 1, e, 0]
        assert psi[i] == pytest.approx(f0 * d * b, rel=1e-14, abs=1e-300)
    assert np.sum(c.rpaq0 ** 2) == pytest.approx(1.0, rel=1e-13)
    # iphoto = 2 at J = Jini returns the bound state itself
    c2 = case("photo")
    bs2 = photo.BoundState(bs.ir1, bs.ir2, bs.iang, bs.omega, bs.ielec, bs.value, 0, -0.5,
                           bs.grid)
    psi2 = photo.dip_bndgrid(c2, bs2, None, 2)
    assert np.sum(psi2 ** 2) == pytest.approx(np.sum(bs.value ** 2), rel=1e-12)


def test_photo_files_round_trip_through_load_case(tmp_path):
    """A run directory with input.dat, pot/, func/ and, beside it, dip/ and
    bndJ01m/ (split over two bound-state files) gives the packet that the
    in-memory generator builds."""
    p = leps.config(5, scale=12)
    p.update(nvmax=2, jmax=4, netot=60, nprocbnd=2)
    text, pot, fdat = leps.make_inputs(p)
    ref = leps.build(dict(p))
    bs, dipol = leps.photo_packet(p, ref)
    run = tmp_path / "J000"
    model.write_case(str(run), text, pot, fdat)
    photo.write_dip(str(tmp_path / "dip"), ref, dipol)
    photo.write_bnd(str(tmp_path), 1, -1, bs, nparts=2)
    back = photo.read_dip(str(tmp_path / "dip"), ref, 1)
    assert all(np.array_equal(a, b) for a, b in zip(back, dipol))
    bs2 = photo.read_bnd(str(tmp_path), ref, 1, -1, 0, 2)
    assert np.array_equal(bs2.value, bs.value) and np.array_equal(bs2.ir1, bs.ir1)
    assert bs2.energy_ev == bs.energy_ev
    c = model.load_case(str(run))
    assert np.array_equal(c.rpaq0, ref.rpaq0)
    assert c.photonorm == ref.photonorm
    # a bound state computed on another grid is refused
    bs3 = photo.BoundState(bs.ir1, bs.ir2, bs.iang, bs.omega, bs.ielec, bs.value, 0, -0.5,
                           (bs.grid[0] + 0.01,) + bs.grid[1:])
    photo.write_bnd(str(tmp_path / "other"), 1, -1, bs3)
    with pytest.raises(ValueError):
        photo.read_bnd(str(tmp_path / "other"), ref, 1, -1, 0, 1)


def test_jmix_packet_is_the_seeded_superposition():
    c = case("jmix")
    w = c.jmix_weights
    assert sorted(w) == [0, 1, 2]
    assert sum(v * v for v in w.values()) == pytest.approx(1.0)
    parts = {j: model.colini_packet(c, jref=j) for j in w}
    assert np.allclose(c.rpaq0, sum(w[j] * parts[j] for j in w), rtol=0, atol=1e-15)
    # the components are orthogonal (different j in the angular functions)
    assert abs(np.dot(parts[0], parts[2])) < 1e-12 * np.dot(parts[2], parts[2])

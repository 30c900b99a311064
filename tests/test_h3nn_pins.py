"""The restated H3 surface of the HD+H example (tests/h3nn.py) against the
example's own Fortran compiled into oracle/_ref/libh3nn.so."""
import ctypes as ct
import os

import numpy as np
import pytest

from tests import h3nn

LIB = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle", "_ref",
                   "libh3nn.so")


@pytest.mark.skipif(not os.path.exists(LIB), reason="oracle/_ref/libh3nn.so not built")
def test_potelebond_matches_the_example_build():
    L = ct.CDLL(LIB)
    pd = ct.POINTER(ct.c_double)
    rng = np.random.default_rng(9)
    r1 = rng.uniform(0.1, 8.0, 300)
    r2 = rng.uniform(0.1, 12.0, 300)
    c = rng.uniform(-1.0, 1.0, 300)
    mine = h3nn.potelebond(r1, r2, c)
    for a, b, cc, m in zip(r1, r2, c, mine):
        pot = np.zeros(1)
        L.potelebond_(ct.byref(ct.c_double(a)), ct.byref(ct.c_double(b)),
                      ct.byref(ct.c_double(cc)), pot.ctypes.data_as(pd),
                      ct.byref(ct.c_int(1)), ct.byref(ct.c_int(1)))
        assert m == pytest.approx(pot[0], rel=1e-12, abs=1e-14)


def test_h3_surface_is_symmetric_and_dissociates():
    """Permutation invariance of the three distances; with the third atom far,
    the diatom's well sits at about 1.4 bohr at the zero of energy (the
    reference point is H + H2 at 1.4 bohr) and dissociates about 4.7 eV up."""
    d = np.array([1.3, 2.7, 3.9])
    e = h3nn.nn_energy(*d)
    for perm in ([1, 0, 2], [2, 1, 0], [0, 2, 1]):
        assert h3nn.nn_energy(*d[perm]) == pytest.approx(e, rel=1e-14)
    r = np.linspace(0.8, 3.0, 200)
    v = h3nn.potelebond(r, 60.0, 1.0) * 27.211386
    assert 1.3 < r[np.argmin(v)] < 1.5 and abs(v.min()) < 0.01
    assert 4.4 < h3nn.potelebond(10.0, 60.0, 1.0) * 27.211386 < 5.0

"""Small propagation cases for the tests, built with the synthetic LEPS
generator at reduced grid sizes (cached per process)."""
import functools

from madwave3_amd import leps, model
from madwave3_amd import namelist as NL


def _build(p):
    text, pot, fdat = leps.make_inputs(p)
    return model.build_case(NL.parse(text), pot, fdat)


@functools.lru_cache(maxsize=None)
def case(name):
    """Named small cases:
      j0      J = 0, one Omega block, one state (configuration 1 shape / 8)
      coriol  J = 2, Omega = 0..2 coupled by Coriolis (configuration 3 shape)
      states  J = 0, three coupled diabatic states (configuration 5 shape)
      odd     J = 0 on odd, non-smooth grid sizes (per-angle DST sizes vary)
    """
    if name == "j0":
        p = leps.config(1, scale=8)
    elif name == "coriol":
        p = leps.config(3, scale=12)
        p.update(Jtot=2, iommax=2, iomref=-1, jref=1, nvmax=2, jmax=4)
    elif name == "states":
        p = leps.config(5, scale=12)
        p.update(nvmax=2, jmax=4)
    elif name == "odd":
        p = leps.config(1, scale=1)
        p.update(npun1=23, npun2=37, nangu=11, npun1min=8, jmax=5, netot=40,
                 rfin1_bohr=6.0, rfin2_bohr=12.0, R0_bohr=7.0, width_bohr=0.6)
    elif name == "mid":
        p = leps.config(1, scale=2)
        p.update(netot=100)
    else:
        raise KeyError(name)
    p.update(kminelastic=5)
    return _build(p)

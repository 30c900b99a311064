"""Autocorrelation, psi(E) and the spectrum from the autocorrelation on the
GPU against the oracle (photodissociation mode, iphoto = 1)."""
import numpy as np
import pytest

from madwave3_amd import consts as C
from madwave3_amd import spectra
from madwave3_amd.lib import MW3_FP64, MW3_MIXED, Propagator, cheby_spectrum
from oracle.oracle import Oracle
from oracle.oracle import cheby_spectrum as oracle_spectrum
from tests.cases import photo_case


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["j0", "states"])
def test_autocorrelation_and_psie_match_oracle(name):
    c = photo_case(name)
    n = 90
    print("photo case %s: %d steps" % (name, n), flush=True)
    g = Propagator(c, MW3_FP64)
    o = Oracle(c)
    g.begin(c.rpaq0)
    o.begin(c.rpaq0)
    ng, ag = g.steps(n, want_auto=True)
    no, ao, _ = o.steps(n)
    zg, zo = g.psie(), o.psie()
    g.close()
    assert np.max(np.abs(ag - ao)) <= 1e-11 * np.max(np.abs(ao))
    assert np.max(np.abs(ng - no)) <= 1e-11 * np.max(no)
    assert np.max(np.abs(zg - zo)) <= 1e-10 * np.max(np.abs(zo))


@pytest.mark.gpu
def test_spectrum_kernel_matches_oracle():
    rng = np.random.default_rng(20261202)
    kmax = 3000
    autoc = rng.standard_normal(kmax + 1) * np.exp(-np.arange(kmax + 1) / 800.0)
    autoc[0] = 1.0
    emindlt, delta2, xkgam = 1.3, 2.1, 1e-4
    e = emindlt + delta2 * np.linspace(-1.05, 1.05, 777)  # a few |Es| > 1 (NaN -> 0)
    sg = cheby_spectrum(autoc, emindlt, delta2, C.hbr, xkgam, e)
    so = oracle_spectrum(autoc, emindlt, delta2, C.hbr, xkgam, e)
    assert np.array_equal(sg == 0.0, so == 0.0)
    assert np.max(np.abs(sg - so)) <= 1e-10 * np.max(np.abs(so))


@pytest.mark.gpu
def test_spectrum_from_a_propagation_through_the_auto_file(tmp_path):
    """Mixed-precision propagation -> auto file -> cheby-spectra driver, against
    the oracle's autocorrelation through the oracle's sum."""
    c = photo_case("states", nofe=0)
    n = 300
    g = Propagator(c, MW3_MIXED)
    g.begin(c.rpaq0)
    _, ag = g.steps(n, want_auto=True)
    g.close()
    o = Oracle(c)
    o.begin(c.rpaq0)
    _, ao, _ = o.steps(n)
    path = str(tmp_path / "auto")
    spectra.write_auto(path, 1.0, 0.0, c.emindlt, c.delta2, ag, k0=2)
    text = "&inputEspectro emin=0.2, emax=1.6, ne=300, xkgam=0.002, kmax=%d /\n" % (n + 1)
    rows, integ = spectra.run(text, path, eref=0.0)
    # same energies through the oracle: auto(1) is not written by the run
    # (steps start at k = 2), as in a reference auto file read from k = 1
    a_full = np.zeros(n + 2)
    a_full[0] = 1.0
    a_full[2:] = ao
    e_zot = rows[:, 0] * C.eV2cm * C.conve1
    so = oracle_spectrum(a_full, c.emindlt, c.delta2, C.hbr, 0.002, e_zot)
    assert np.max(np.abs(rows[:, 1] - so)) <= 1e-5 * np.max(np.abs(so))
    assert np.isfinite(integ)

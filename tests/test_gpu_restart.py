"""A propagation stopped, written to the continuation files (wvpcont,
SRC/main_madwave3.f:494-552), read back into a new context (start-up with
it0 > 0 and ncontfile = 1, main_madwave3.f:223-261, mod_flux_01y2.f:298-322,
mod_coortrans01_02.f:180-196) and continued gives the run that was not
stopped: norms to 1e-13, P(E), zS2 and zS2prod to 1e-12 relative.

The per-step records (mw3_records; the Cvj.LLLL / Cvjprod.LLLL files of
main_madwave3.f:311-365 through restart.time_loop) against the oracle's
per-step projections, 1e-11 relative in FP64."""
import os

import numpy as np
import pytest

from madwave3_amd import restart
from madwave3_amd.lib import MW3_FP64, Propagator
from oracle.oracle import Oracle
from tests.cases import case

pytestmark = pytest.mark.gpu


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


@pytest.mark.parametrize("name", ["prod", "photo"])
def test_continue_from_files(tmp_path, name):
    c = case(name)
    print("case", name, c, flush=True)
    g1 = Propagator(c, MW3_FP64)
    g1.begin(c.rpaq0)
    g1.steps(300)
    restart.write_cont(str(tmp_path), g1, iloop=1)
    n1, a1 = g1.steps(200, want_auto=c.iphoto > 0)
    g2 = Propagator(c, MW3_FP64)
    assert restart.resume(g2, str(tmp_path)) == (301, 1, 1)
    n2, a2 = g2.steps(200, want_auto=c.iphoto > 0)
    assert g2.k == g1.k == 501
    print("norms %.1e P %.1e" % (rel(n2, n1), rel(g2.flux_result()[0], g1.flux_result()[0])),
          flush=True)
    assert rel(n2, n1) < 1e-13
    if c.iphoto > 0:
        assert rel(a2, a1) < 1e-13
    assert rel(g2.flux_result()[0], g1.flux_result()[0]) < 1e-12
    assert rel(g2.zs2(), g1.zs2()) < 1e-12
    if c.iprod == 2:
        assert rel(g2.zs2prod(), g1.zs2prod()) < 1e-12
    g1.close()
    g2.close()


def test_records_across_calls_match_the_oracle():
    """Cvj of every step over two calls (40 + 260 steps: the second flushes
    its ring at 256 steps and at its end), taken in two parts."""
    c = case("prod")
    g = Propagator(c, MW3_FP64)
    g.begin(c.rpaq0)
    g.records(True)
    g.steps(40)
    g.steps(260)
    i1, v1, p1 = g.take_records(100)
    i2, v2, p2 = g.take_records()
    assert len(i1) == 100 and len(i2) == 200 and len(g.take_records()[0]) == 0
    its = np.concatenate([i1, i2])
    assert np.array_equal(its, np.arange(2, 302))
    o = Oracle(c)
    o.begin(c.rpaq0)
    _, _, cv_o = o.steps(300, want_cvj=True)
    e = rel(np.concatenate([v1, v2]), cv_o)
    print("Cvj records: rel err %.1e" % e, flush=True)
    assert e < 1e-11
    assert p1.shape == (100,) + g.cvjprod().shape
    assert np.array_equal(p2[-1], g.cvjprod())      # the last record is the last step's
    g.records(False)
    g.steps(3)
    assert len(g.take_records()[0]) == 0
    g.close()


@pytest.mark.parametrize("name", ["photo", "prod"])
def test_time_loop_record_files_match_the_oracle(tmp_path, name):
    """One time loop of 300 steps: Cvj.0001 (photodissociation, iphoto = 1)
    or Cvjprod.0001 (iprod = 2) against the oracle stepped one step at a
    time, and the loop's continuation files."""
    c = case(name)
    print("case", name, c, flush=True)
    g = Propagator(c, MW3_FP64)
    g.begin(c.rpaq0)
    d = str(tmp_path)
    restart.time_loop(g, d, iloop=1, ntimes=300)
    o = Oracle(c)
    o.begin(c.rpaq0)
    nv, nj = c.nvmax - c.nvini + 1, c.jmax - c.jini + 1
    if name == "photo":
        its, v = restart.read_step_records(os.path.join(d, "Cvj.0001"), (c.ncan, nj, nv))
        want = o.steps(300, want_cvj=True)[2]
    else:
        shp = g.cvjprod().shape
        its, v = restart.read_step_records(os.path.join(d, "Cvjprod.0001"), shp)
        want = []
        for _ in range(300):
            o.steps(1)
            want.append(o.cvjprod())
        want = np.array(want)
    assert np.array_equal(its, np.arange(2, 302))
    e = rel(v, want)
    print("%s records: rel err %.1e" % (name, e), flush=True)
    assert e < 1e-11
    assert restart.read_cont(d, c, g.nflux)["it0"] == 301
    g.close()

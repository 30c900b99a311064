"""A propagation stopped, written to the continuation files (wvpcont,
SRC/main_madwave3.f:494-552), read back into a new context (start-up with
it0 > 0 and ncontfile = 1, main_madwave3.f:223-261, mod_flux_01y2.f:298-322,
mod_coortrans01_02.f:180-196) and continued gives the run that was not
stopped: norms to 1e-13, P(E), zS2 and zS2prod to 1e-12 relative."""
import numpy as np
import pytest

from madwave3_amd import restart
from madwave3_amd.lib import MW3_FP64, Propagator
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

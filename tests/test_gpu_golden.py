"""The reference's own HD+H example (EXAMPLES/HD+H_collision, J = 0) run on
the GPU from restated inputs, against the results the reference's run log
and table hold (tests/golden/hd_h_j000_norms.txt: |psi_k|^2 for k = 1..4001
printed by check; tests/golden/hd_h_j000_S2prod.txt: the printprob table
after the 4000 steps, columns E, P(E), P(E) + sum of reactant-state
probabilities).

Inputs: the example's input.dat (tests/golden/hd_h_input.dat), its neural-
network surface restated in tests/h3nn.py (pinned against its own build),
the potential grid and the HD functions by madwave3_amd.prep / diatom
(pinned against the reference's routines); the set-up reproduces the
log's Vmin, compressed-point count and initial norm (checked here too).
The product-state analysis (iprod = 2) does not enter these columns and is
left off.  Bars: the norms to 1e-9 of |psi_1|^2 in FP64 (the reference ran
on 7 ranks, so its sums are ordered differently) and 1e-5 in mixed
precision; P and P + S2reac to the table's 7 printed digits plus 1e-9 (FP64)
or 1e-4 absolute (mixed)."""
import json
import os

import numpy as np
import pytest

from madwave3_amd import consts as C
from madwave3_amd import model, prep, printprob
from madwave3_amd import namelist as NL
from madwave3_amd.lib import MW3_FP64, MW3_MIXED, Propagator
from tests import h3nn

HERE = os.path.dirname(os.path.abspath(__file__))
G = os.path.join(HERE, "golden")

pytestmark = pytest.mark.gpu

_CASE = {}


def golden_case():
    if "c" not in _CASE:
        text = open(os.path.join(G, "hd_h_input.dat")).read()
        pot, fdat = prep.prepare(text, h3nn.potelebond)
        nl = NL.parse(text)
        nl["inputprod"]["iprod"] = 0
        _CASE["c"] = model.build_case(nl, pot, fdat)
        _CASE["vmin"] = pot.vmin
    return _CASE["c"], _CASE["vmin"]


def test_golden_setup_matches_the_log():
    c, vmin = golden_case()
    pins = json.load(open(os.path.join(G, "hd_h_j000_pins.json")))
    assert c.ntot == 704278
    assert abs(C.zot_to_ev(vmin) - pins["emin_eV"]) < 1e-12
    assert abs(c.norm_initial - pins["norm_initial"]) < 1e-12


@pytest.mark.parametrize("precision,tol_n,tol_p", [(MW3_FP64, 1e-9, 1e-9), (MW3_MIXED, 1e-5, 1e-4)])
def test_golden_norms_and_probabilities(precision, tol_n, tol_p):
    c, _ = golden_case()
    ref_n = np.loadtxt(os.path.join(G, "hd_h_j000_norms.txt"))[:, 1]
    table = np.loadtxt(os.path.join(G, "hd_h_j000_S2prod.txt"))
    nsteps = int(c.ntimes) * (int(c.nloop) + 1)
    assert nsteps == 4000 and len(ref_n) == nsteps + 1
    print("golden example:", c, "precision", precision, flush=True)
    g = Propagator(c, precision)
    g.begin(c.rpaq0)
    p1 = g.state()[1]
    norms = [float(p1 @ p1)]
    for _ in range(8):
        norms.extend(g.steps(500)[0])
    norms = np.array(norms)
    P = g.flux_result()[0]
    _, S2reac = printprob.reactant_probabilities(c, g.zs2())
    g.close()
    dn = float(np.max(np.abs(norms - ref_n)) / ref_n[0])
    E = C.zot_to_ev(np.asarray(c.etotS2))
    assert np.allclose(E, table[:, 0], rtol=5e-7, atol=0)
    print7 = 5e-8 * np.abs(table[:, 1:3])       # the table's 7 significant digits
    dP = np.abs(P - table[:, 1]) - print7[:, 0]
    dT = np.abs(P + S2reac - table[:, 2]) - print7[:, 1]
    print("max |d norm|/|psi_1|^2 %.2e; max |dP| beyond print %.2e; max |d(P + S2reac)| %.2e;"
          " P max %.4f" % (dn, dP.max(), dT.max(), table[:, 1].max()), flush=True)
    assert dn < tol_n
    assert dP.max() < tol_p and dT.max() < tol_p

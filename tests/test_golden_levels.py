"""Diatomic levels of the HD+H example (EXAMPLES/HD+H_collision) from the
restated preparation (madwave3_amd.prep, diatom) against the levels the
reference's run log prints after subtracting the reference energy
(J000-check/sal.000; tests/golden/hd_h_j000_reac_levels.txt and
hd_h_j000_prod_levels.txt, written by tests/golden/make_fixtures.py).

The log comes from an earlier version of the reference than its current
source.  Two differences show in these levels and are checked here as
facts about the fixture:
  - its reactant list also holds levels within 0.01 hartree of the
    asymptote, which the current source's test
    v(r_max) - E > 1e-2 (SRC/mod_baseYfunciones_01y2.f:702) drops; every
    level both lists hold agrees;
  - its product levels are those of the DVR on the R grid (bndbcele,
    SRC/liboctdyn.f:1687-1813, the diabatic_prod_pot > 0 branch of the
    current source), not of the finite-difference + Cooley route of
    bndbc1ele.
The Cooley route for the products is pinned by symmetry instead: with
xm1 = xm2 and a permutation-invariant surface, the 0-2 diatom is the 0-1
diatom, so its levels equal the reactant ones."""
import os

import numpy as np
import pytest

from madwave3_amd import consts as C
from madwave3_amd import namelist as NL
from madwave3_amd import prep
from tests import h3nn

G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
H2EV = C.zot_to_ev(1.0 / C.conve * C.conve1)


@pytest.fixture(scope="module")
def example():
    nl = NL.parse(open(os.path.join(G, "hd_h_input.dat")).read())
    g = NL.group(nl, "inputgridbase")
    pm = NL.group(nl, "inputpotmass")
    pr = NL.group(nl, "inputprod")
    reac = prep.reactant_functions(
        h3nn.potelebond, int(g["npun1"]), float(g["rmis1"]), float(g["rfin1"]), float(pm["xm0"]),
        float(pm["xm1"]), int(g["jini"]), int(g["jmax"]), int(g["nvini"]), int(g["nvmax"]),
        int(g["nvref"]), int(g["jref"]))
    args = (h3nn.potelebond, int(g["npun2"]), float(g["rmis2"]), float(g["rfin2"]),
            int(pr["n2prod1"]), float(pm["xm0"]), float(pm["xm2"]), 0, int(pr["jmaxprod"]),
            int(pr["nviniprod"]), int(pr["nvmaxprod"]))
    return reac, prep.product_functions_dvr(*args), prep.product_functions(*args)


def test_reactant_levels_match_the_log(example):
    reac = example[0]
    rows = np.loadtxt(os.path.join(G, "hd_h_j000_reac_levels.txt"))
    eref_ev = C.zot_to_ev(reac.eref)
    asym_ev = float(h3nn.potelebond(37.8, 100.0 / C.convl, 1.0)) * H2EV
    both, extra = 0, 0
    for ie, j, v, e in rows:
        k = (int(ie), int(j), int(v))
        if k in reac.ediat:
            assert abs(C.zot_to_ev(reac.ediat[k]) - eref_ev - e) < 1e-12, k
            both += 1
        elif e == -eref_ev or abs(e + eref_ev) < 1e-12:
            extra += 1           # a placeholder row (no level: zero energy)
        else:
            # a level within 0.01 hartree of the asymptote
            assert 0.0 < asym_ev - (e + eref_ev) < 0.01 * H2EV, k
            extra += 1
    assert both == 187 and both + extra == len(rows)


def test_product_dvr_levels_match_the_log(example):
    dvr = example[1]
    rows = np.loadtxt(os.path.join(G, "hd_h_j000_prod_levels.txt"))
    eref_ev = C.zot_to_ev(example[0].eref)
    assert len(rows) == 231
    for j, v, e in rows:
        assert abs(C.zot_to_ev(dvr.ediat[(int(j), int(v))]) - eref_ev - e) < 1e-11, (j, v)
    for j in range(21):
        f = dvr.fd[(j, 0)][:, 0]
        assert abs(float(f @ f) - 1.0) < 1e-12


def test_product_cooley_levels_equal_the_reactant_levels(example):
    reac, _, cool = example
    n = 0
    for (j, v), e in cool.ediat.items():
        if (1, j, v) in reac.ediat:
            assert abs(C.zot_to_ev(e - reac.ediat[(1, j, v)])) < 1e-9, (j, v)
            n += 1
    assert n == 21 * 6            # j = 0..20, v = 0..5
    assert cool.vibel[(0, 3)] == 3.0 and cool.nobc[0] == 10

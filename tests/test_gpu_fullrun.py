"""Full-length propagation of configuration 1 (128 r x 256 R x 60 gamma, J=0,
LEPS, 20000 Chebyshev steps) with both GPU propagators against the FP64 CPU
oracle's run of the same workload (tests/golden/cfg1_oracle_full.json, made by
tests/golden/make_cfg1_oracle.py).  Tolerances: FP64 max |dP(E)| <= 1e-8;
mixed (FP32 operator terms) max |dP(E)| <= 1e-4, the bound the mixed
propagator is held to; norms |psi_k|^2 every 100 steps to 1e-9 / 1e-5 of
the initial norm."""
import json
import os

import numpy as np
import pytest

from madwave3_amd import leps, model
from madwave3_amd import namelist as NL
from madwave3_amd.lib import MW3_FP64, MW3_MIXED, Propagator

HERE = os.path.dirname(os.path.abspath(__file__))
FIX = os.path.join(HERE, "golden", "cfg1_oracle_full.json")


def _case():
    p = leps.config(1)
    text, pot, fdat = leps.make_inputs(p)
    return model.build_case(NL.parse(text), pot, fdat), int(p["ntimes"]) * (int(p["nloop"]) + 1)


@pytest.mark.gpu
@pytest.mark.skipif(not os.path.exists(FIX), reason="full-run oracle fixture not generated")
@pytest.mark.parametrize("precision,tol_p,tol_n", [(MW3_FP64, 1e-8, 1e-9), (MW3_MIXED, 1e-4, 1e-5)])
def test_full_run_matches_oracle(precision, tol_p, tol_n):
    ref = json.load(open(FIX))
    case, nsteps = _case()
    assert nsteps == ref["steps"]
    print("full run: configuration 1, %d steps, precision %d" % (nsteps, precision), flush=True)
    g = Propagator(case, precision)
    g.begin(case.rpaq0)
    norms = np.concatenate([g.steps(min(4000, nsteps - i))[0] for i in range(0, nsteps, 4000)])
    P = g.flux_result()[0]
    z = g.zs2()
    g.close()
    Pr = np.asarray(ref["P"])
    dP = float(np.max(np.abs(P - Pr)))
    # norm error against the initial norm (absorption takes |psi|^2 from 1 to 0.005)
    dn = float(np.max(np.abs(norms[99::100] - ref["norm_every_100"])) / ref["norm_every_100"][0])
    zr = np.asarray(ref["zs2_v0_abs2"])
    dz = float(np.max(np.abs(np.abs(z[0, :, 0, :]) ** 2 - zr)))
    print("max |dP| %.3e  max d|psi|^2 / |psi_1|^2 %.3e  max d|zS2|^2 %.3e  (P max %.3f)"
          % (dP, dn, dz, float(Pr.max())), flush=True)
    assert dP <= tol_p
    assert dn <= tol_n
    assert dz <= tol_p * max(1.0, float(zr.max()))

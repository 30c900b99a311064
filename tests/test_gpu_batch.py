"""mw3_cheby_steps_batch: contexts of different shapes advanced together give
exactly the results of advancing each alone (same kernels, same order within
each context's stream)."""
import numpy as np
import pytest

from madwave3_amd.lib import MW3_FP64, MW3_MIXED, Propagator, steps_batch
from tests.cases import case

NAMES = ["j0", "coriol", "states"]


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [MW3_FP64, MW3_MIXED])
def test_batch_equals_separate_runs(precision):
    nsteps = 300  # crosses a 256-step ring boundary
    print("batch of %s, %d steps, precision %d" % (NAMES, nsteps, precision), flush=True)
    alone = []
    for n in NAMES:
        c = case(n)
        g = Propagator(c, precision)
        g.begin(c.rpaq0)
        norms, _ = g.steps(nsteps)
        alone.append((norms, g.flux_result()[0], g.zs2()))
        g.close()
    props = []
    for n in NAMES:
        c = case(n)
        g = Propagator(c, precision)
        g.begin(c.rpaq0)
        props.append(g)
    norms_b = steps_batch(props, nsteps)
    for i, g in enumerate(props):
        assert np.array_equal(norms_b[i], alone[i][0])
        assert np.array_equal(g.flux_result()[0], alone[i][1])
        assert np.array_equal(g.zs2(), alone[i][2])
        g.close()

"""Chebyshev steps as captured HIP graphs (MW3_OPT_GRAPHS = 1, the default)
against the same steps enqueued kernel by kernel: the kernels and the order
of every sum are the same, so every result must agree bit for bit.  Each run
crosses ring flushes (256 steps), replays the two captured graphs (one per
psi buffer) many times with host work and other calls in between, and ends
with calls of one and of an odd number of steps."""
import numpy as np
import pytest

from madwave3_amd.lib import MW3_FP64, MW3_MIXED, MW3_OPT_GRAPHS, Propagator, steps_batch
from tests.cases import case

pytestmark = pytest.mark.gpu


def run(c, precision, graphs):
    g = Propagator(c, precision)
    g.set_option(MW3_OPT_GRAPHS, graphs)
    assert g.get_option(MW3_OPT_GRAPHS) == graphs
    g.begin(c.rpaq0)
    out = {}
    n1, a1 = g.steps(300, want_auto=True)
    pm, p, k = g.state()
    n2, a2 = g.steps(1, want_auto=True)
    n3, a3 = g.steps(37, want_auto=True)
    out["norms"] = np.concatenate([n1, n2, n3])
    out["auto"] = np.concatenate([a1, a2, a3])
    out["mid_state"] = np.concatenate([pm, p])
    pm, p, k = g.state()
    out["state"] = np.concatenate([pm, p])
    out["k"] = k
    out["P"] = g.flux_result()[0]
    out["zs2"] = g.zs2()
    za, zb = g.flux_accum()
    out["zcr"], out["zcdr"] = za, zb
    if c.iprod == 2:
        out["zs2prod"] = g.zs2prod()
    g.close()
    return out


@pytest.mark.parametrize("name", ["j0", "coriol", "states", "prod", "prodJ", "photo"])
@pytest.mark.parametrize("precision", [MW3_FP64, MW3_MIXED])
def test_graph_steps_equal_kernel_steps(name, precision):
    c = case(name)
    print("case", name, c, "precision", precision, flush=True)
    a = run(c, precision, 0)
    b = run(c, precision, 1)
    assert a["k"] == b["k"] == 1 + 338
    for key in a:
        assert np.array_equal(np.asarray(a[key]), np.asarray(b[key])), key


def test_graph_steps_in_a_batch():
    names = ["j0", "coriol", "states"]
    cs = [case(n) for n in names]
    res = {}
    for graphs in (0, 1):
        props = [Propagator(c, MW3_MIXED) for c in cs]
        for g, c in zip(props, cs):
            g.set_option(MW3_OPT_GRAPHS, graphs)
            g.begin(c.rpaq0)
        n = np.concatenate([steps_batch(props, 260), steps_batch(props, 3)], axis=1)
        res[graphs] = (n, [g.flux_result()[0] for g in props])
        for g in props:
            g.close()
    assert np.array_equal(res[0][0], res[1][0])
    for x, y in zip(res[0][1], res[1][1]):
        assert np.array_equal(x, y)

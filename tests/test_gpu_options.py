"""Execution options (MW3_OPT_*) and the autotuner.

Every combination of the radial variant (persistent low-mode kernels or
GEMMs), the angular variant (folded one-pass kernel, unfolded one-pass
kernel, GEMMs) and the stream layout must give the oracle's H psi and
recursion to the usual bars (FP64 1e-11 / 1e-10 relative, mixed 2e-5 on
H psi); the autotuner must leave the propagation state and every
accumulator as it found them and keep a combination the context can run."""
import itertools

import numpy as np
import pytest

from madwave3_amd.lib import (MW3_FP64, MW3_MIXED, MW3_OPT_ANGULAR, MW3_OPT_ANGULAR_GRID,
                              MW3_OPT_FP16_PRODUCTS, MW3_OPT_RADIAL, MW3_OPT_RADIAL_GRID,
                              MW3_OPT_SIDE_STREAM, Propagator)
from oracle.oracle import Oracle
from tests.cases import case

pytestmark = pytest.mark.gpu

# radial, angular, side stream, radial grid divisor, angular grid divisor
COMBOS = [k + (1, 1) for k in itertools.product([0, 1], [0, 1, 2], [0, 1])] + \
    [(1, 0, 1, 2, 1), (1, 0, 0, 4, 1), (1, 0, 1, 2, 2), (0, 0, 0, 1, 2)]


def rel(a, b):
    return float(np.max(np.abs(np.asarray(a) - np.asarray(b))) / max(np.max(np.abs(b)), 1e-300))


@pytest.mark.parametrize("name", ["j0", "coriol", "states", "r512", "photo"])
def test_every_option_combination_matches_oracle(name):
    c = case(name)
    print("case", name, c, flush=True)
    o = Oracle(c)
    rng = np.random.default_rng(3)
    psi = rng.standard_normal(c.ntot)
    h_ref = o.hphi(psi)
    o.begin(c.rpaq0)
    n_ref, _, _ = o.steps(60)
    P_ref = o.flux_result()[0]
    seen = set()
    for radial, angular, side, grid, agrid in COMBOS:
        for prec in (MW3_FP64, MW3_MIXED):
            g = Propagator(c, prec)
            g.set_option(MW3_OPT_RADIAL, radial)
            g.set_option(MW3_OPT_ANGULAR, angular)
            g.set_option(MW3_OPT_SIDE_STREAM, side)
            g.set_option(MW3_OPT_RADIAL_GRID, grid)
            g.set_option(MW3_OPT_ANGULAR_GRID, agrid)
            eff = tuple(g.get_option(k) for k in (MW3_OPT_RADIAL, MW3_OPT_ANGULAR,
                                                  MW3_OPT_SIDE_STREAM, MW3_OPT_RADIAL_GRID,
                                                  MW3_OPT_ANGULAR_GRID))
            seen.add(eff)
            label = "%s prec %d options %s (in effect %s)" % (
                name, prec, (radial, angular, side, grid, agrid), eff)
            e_h = rel(g.hphi(psi), h_ref)
            if prec == MW3_FP64:
                g.begin(c.rpaq0)
                n, _ = g.steps(60)
                e_n = rel(n, n_ref)
                e_p = rel(g.flux_result()[0], P_ref)
                print(label, "H psi %.1e norms %.1e P %.1e" % (e_h, e_n, e_p), flush=True)
                assert e_h < 1e-11 and e_n < 1e-10 and e_p < 1e-9, label
            else:
                print(label, "H psi %.1e" % e_h, flush=True)
                assert e_h < 2e-5, label
            g.close()
    print(name, "variants in effect:", sorted(seen), flush=True)


@pytest.mark.parametrize("prec", [MW3_FP64, MW3_MIXED])
def test_autotune_keeps_state_and_accumulators(prec):
    c = case("prod")
    g = Propagator(c, prec)
    g.begin(c.rpaq0)
    g.steps(70)
    pm, p, k = g.state()
    za, zb = g.flux_accum()
    zs = g.zs2()
    zp = g.zs2prod()
    chosen, ms = g.autotune(10)
    print("precision", prec, "chosen", chosen, "%.4f ms/step" % ms, flush=True)
    assert ms > 0 and chosen["radial"] in (0, 1) and chosen["angular"] in (0, 1, 2)
    assert chosen == g.options()
    pm2, p2, k2 = g.state()
    assert k2 == k and np.array_equal(pm2, pm) and np.array_equal(p2, p)
    za2, zb2 = g.flux_accum()
    assert np.array_equal(za2, za) and np.array_equal(zb2, zb)
    assert np.array_equal(g.zs2(), zs) and np.array_equal(g.zs2prod(), zp)
    # stepping on agrees with an untuned context to rounding
    n1, _ = g.steps(40)
    h = Propagator(c, prec)
    h.begin(c.rpaq0)
    h.steps(70)
    n2, _ = h.steps(40)
    assert rel(n1, n2) < (1e-11 if prec == MW3_FP64 else 1e-5)
    assert rel(g.flux_result()[0], h.flux_result()[0]) < (1e-9 if prec == MW3_FP64 else 1e-4)


@pytest.mark.parametrize("name", ["j0", "states", "r512", "r384"])
def test_fp16_products_option(name):
    """MW3_OPT_FP16_PRODUCTS: mixed H psi with the inverse R pass and the
    folded angular kernel multiplying as compensated FP16 pairs (default)
    and in FP32, both within the mixed bar of the oracle; the two differ
    (other kernels ran), the FP64 context reports 0, and the autotuner keeps
    the value it was given."""
    c = case(name)
    rng = np.random.default_rng(5)
    psi = rng.standard_normal(c.ntot)
    h_ref = Oracle(c).hphi(psi)
    out = {}
    for fp16 in (1, 0):
        g = Propagator(c, MW3_MIXED)
        g.set_option(MW3_OPT_SIDE_STREAM, 0)
        g.set_option(MW3_OPT_FP16_PRODUCTS, fp16)
        assert g.get_option(MW3_OPT_FP16_PRODUCTS) == fp16
        out[fp16] = g.hphi(psi)
        e = rel(out[fp16], h_ref)
        print(name, "fp16 products", fp16, "H psi rel err %.2e" % e, flush=True)
        assert e < 2e-5
        g.begin(c.rpaq0)
        chosen, _ = g.autotune(4)
        assert chosen["fp16_products"] == fp16
        g.close()
    assert not np.array_equal(out[0], out[1])
    g = Propagator(c, MW3_FP64)
    assert g.get_option(MW3_OPT_FP16_PRODUCTS) == 0
    g.close()

"""The benchmark configurations at their full grid sizes: the FP64 GPU
propagator against the FP64 CPU oracle over the first steps of each run.

This is the check that makes the FP64 GPU path the long-run reference of the
full-length mixed-precision runs (`bench.py --accuracy`): after K steps the
state psi_K agrees with the oracle's to ||dpsi||/||psi|| <= 1e-12 (1e-11
with the parity-folded angular kernel, see below), the norms (sums over the
whole grid, added in another order) to 1e-11, and the energy-resolved
accumulators (whose terms cancel) to 1e-9 relative to their largest entry
plus 1e-13.  The step counts keep the oracle's share to seconds on the
host.

  configuration 2  J = 0, 256 r x 512 R x 120 gamma, product states (iprod = 2)
  configuration 3  J = 12, Omega = 0..6 (7 Coriolis-coupled blocks), 192 x 384 x 96,
                   seeded j = 0, 1, 2 initial packet
  configuration 5  three coupled states, 256 x 384 x 96, photodissociation packet
                   (bound state x linear transition dipole, iphoto = 1)
  configuration 4  J = 12 (7 blocks of 160 x 320 x 80) advanced inside the batch
                   of all 24 J values (mw3_cheby_steps_batch)
"""
import functools

import numpy as np
import pytest

from madwave3_amd import leps
from madwave3_amd.lib import MW3_FP64, MW3_MIXED, MW3_OPT_ANGULAR, Propagator, steps_batch
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu

K_STEPS = {2: 12, 3: 6, 5: 12}


def rel(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def close(a, b, rtol=1e-9, atol=1e-13):
    """max |a - b| within rtol of the largest |b| plus atol: the energy
    accumulators of a run whose packet has not reached the analysis line
    (the photodissociation packet at configuration 5) hold rounding-level
    values only."""
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.max(np.abs(a - b))) <= rtol * float(np.max(np.abs(b))) + atol


def state_rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@functools.lru_cache(maxsize=None)
def full_case(cfg, J=0):
    return leps.build(leps.config(cfg, J=J))


def oracle_run(c, nsteps):
    o = Oracle(c)
    r = {"stats": o.begin(c.rpaq0)}
    r["norms"], r["auto"], _ = o.steps(nsteps)
    r["psim1"], r["psi"], r["k"] = o.state()
    r["zcr"], r["zcdr"] = o.flux_accum()
    r["zs2"] = o.zs2()
    r["P"] = o.flux_result()[0]
    if c.iprod == 2:
        r["zs2prod"] = o.zs2prod()
    return r


def compare(g, r, c, nsteps, label, tol):
    """GPU run against the oracle's results r; tol bounds the state
    (||dpsi||/||psi||); the norms and autocorrelations, sums of 5-17 million
    terms that the GPU and the oracle add in different orders, get 1e-11; the
    energy-resolved accumulators 1e-9."""
    st_g = g.begin(c.rpaq0)
    assert rel(st_g, r["stats"]) < 1e-11, label
    n_g, a_g = g.steps(nsteps, want_auto=True)
    e_n = rel(n_g, r["norms"])
    pm_g, p_g, k_g = g.state()
    e_p, e_pm = state_rel(p_g, r["psi"]), state_rel(pm_g, r["psim1"])
    print(label, "norm rel err %.2e, state rel err %.2e / %.2e (bar %.0e)" % (e_n, e_p, e_pm, tol),
          flush=True)
    assert e_n < 1e-11, label
    if c.iphoto > 0:
        assert rel(a_g, r["auto"]) < 1e-11, label
    assert k_g == r["k"] == 1 + nsteps
    assert e_p < tol and e_pm < tol, label
    za_g, zb_g = g.flux_accum()
    zs = g.zs2()
    print(label, "accumulators: zS2 max %.2e err %.2e" % (np.max(np.abs(r["zs2"])),
                                                         np.max(np.abs(zs - r["zs2"]))), flush=True)
    assert close(za_g, r["zcr"]) and close(zb_g, r["zcdr"]), label
    assert close(zs, r["zs2"]), label
    assert close(g.flux_result()[0], r["P"]), label
    if c.iprod == 2:
        assert close(g.zs2prod(), r["zs2prod"]), label


@pytest.mark.parametrize("cfg", [2, 3, 5])
def test_full_size_fp64_matches_oracle(cfg):
    """The unfolded angular kernel (or the GEMMs where Coriolis coupling
    needs them) to 1e-12; the default parity-folded angular kernel to 1e-11:
    it uses D(n_gamma-1-gamma, j) = +-D(gamma, j), which the Gauss-Legendre
    tables hold only to their own rounding (4e-14 of the largest entry at
    configuration 2), and that difference adds up over the steps."""
    c = full_case(cfg)
    label = "configuration %d %s" % (cfg, c)
    print(label, flush=True)
    r = oracle_run(c, K_STEPS[cfg])
    for angular, tol in ((1, 1e-12), (0, 1e-11)):
        g = Propagator(c, MW3_FP64)
        g.set_option(MW3_OPT_ANGULAR, angular)
        compare(g, r, c, K_STEPS[cfg], "%s angular variant %d" % (label, g.get_option(MW3_OPT_ANGULAR)),
                tol)
        g.close()


@pytest.mark.parametrize("cfg", [2, 3, 5])
def test_full_size_hphi_mixed_matches_oracle(cfg):
    """One H psi in mixed precision at the full grid size, so that the FP32
    operator kernels run at their production shapes (the folded angular
    kernel with 60 and 48 angle pairs, the R passes at 512 and 384 points,
    seven Coriolis-coupled blocks), against the oracle on the
    configuration's initial packet and on a random vector: within 2e-5 of the
    largest entry, the bound of the small cases, for each angular variant."""
    c = full_case(cfg)
    o = Oracle(c)
    rng = np.random.default_rng(cfg)
    for name, psi in (("packet", c.rpaq0), ("random", rng.standard_normal(c.ntot))):
        h_ref = o.hphi(psi)
        for angular in (0, 1):
            g = Propagator(c, MW3_MIXED)
            g.set_option(MW3_OPT_ANGULAR, angular)
            e = rel(g.hphi(psi), h_ref)
            label = "configuration %d, %s, angular variant %d" % (cfg, name,
                                                                  g.get_option(MW3_OPT_ANGULAR))
            g.close()
            print(label, "H psi rel err %.2e" % e, flush=True)
            assert e < 2e-5, label


def test_config4_batch_member_matches_oracle():
    """J = 12 of configuration 4, stepped by mw3_cheby_steps_batch together
    with the other 23 J values, against the oracle run of J = 12 alone; the
    other members must stay finite and keep the norms a lone run gives."""
    nsteps = 4
    cases = [full_case(4, J) for J in range(24)]
    print("configuration 4:", sum(c.ncan for c in cases), "Omega blocks", flush=True)
    props = [Propagator(c, MW3_FP64) for c in cases]
    for g, c in zip(props, cases):
        g.begin(c.rpaq0)
    norms = steps_batch(props, nsteps)
    assert np.all(np.isfinite(norms))
    c = cases[12]
    o = Oracle(c)
    o.begin(c.rpaq0)
    n_o, _, _ = o.steps(nsteps)
    assert rel(norms[12], n_o) < 1e-11
    pm_o, p_o, _ = o.state()
    pm_g, p_g, _ = props[12].state()
    e = state_rel(p_g, p_o)
    print("configuration 4 J=12 in the batch: state rel err %.2e" % e, flush=True)
    assert e < 1e-11 and state_rel(pm_g, pm_o) < 1e-11
    assert close(props[12].flux_result()[0], o.flux_result()[0])
    # a member alone gives its batch norms bit for bit (J = 5: 5 blocks)
    alone = Propagator(cases[5], MW3_FP64)
    alone.begin(cases[5].rpaq0)
    n_alone, _ = alone.steps(nsteps)
    assert np.array_equal(n_alone, norms[5])
    for g in props + [alone]:
        g.close()

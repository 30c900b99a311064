"""Chebyshev H.psi propagation benchmark on one MI355X.

Workload: the J = 0 collision of configuration 2 -- nr = 256 r points on
[0.1, 8] bohr, nR = 512 R points on [0.1, 15] bohr, 120 Gauss-Legendre gamma
points (15.7M grid values), LEPS surface, v=0 j=0 packet at R = 9 bohr,
reactant states v <= 3, j <= 10 and product states HD(v' <= 3, j' <= 15) at
R' = 4 bohr (iprod = 2) -- built in memory by the synthetic input generator.
A "step" is one full Chebyshev iteration (H psi, damped update, per-step
projections on the flux line, the reactant states and the product states;
the energy sums run once per 256 steps and are included).

Prints one JSON line: value = steps/s of the mixed-precision propagator (FP32
operator terms, FP64 Chebyshev state and accumulators); the FP64 propagator's
steps/s and max |dP(E)| between the two over the timed run are reported
beside it, with the roofline of the dominant kernel group and the CPU oracle
timed on the host.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
TIMING_EVERY = 8
FP64_PEAK_TF = 78.6
FP32_PEAK_TF = 157.3


def build_case(cfg):
    from madwave3_amd import leps

    return leps.build(leps.config(cfg))


# kernel launches per step in each group (mixed and FP64 alike)
GROUP_LAUNCHES = {"radial_dst": 3, "angular": 1, "update_fused": 1, "update_tail": 2,
                  "projections": 1, "energy_sums": 2, "coriolis": 1, "products": 1}


def kernel_groups(times, steps):
    """Per-step milliseconds of each kernel group from mw3_timing_get totals:
    radial_dst (the radial correction passes: k_rad_fwd, k_rad_mid, k_rad_inv
    at configuration 2), angular (the one-pass angular kernel; in mixed
    precision on a third stream beside the radial passes), update_fused
    (the update kernel), update_tail (flux columns and norm of the step,
    second stream),
    projections (reactant-state sums; second stream, overlapped with the next
    step), products (the per-step product-state reduction, same stream as
    projections), energy_sums (once per ring flush, with the product-state
    GEMM of the ring: its total time over the timed region divided by the
    `steps` timed steps, since every flush is timed) and coriolis (the FBR
    coefficient pass with the Omega +- 1 stencil, run instead of the fused
    forward epilogue when channels are Coriolis-coupled)."""
    acc = {}
    for name, (ms, cnt) in times.items():
        if name.startswith("dst"):
            g = "radial_dst"
        elif name.startswith("angular"):
            g = "angular"
        elif name == "update_fused":
            g = "update_fused"
        elif name == "update_tail":
            g = "update_tail"
        elif name == "energy_accum":
            g = "energy_sums"
        elif name == "coriolis":
            g = "coriolis"
        elif name == "prod_projection":
            g = "products"
        else:
            g = "projections"
        tot, n = acc.get(g, (0.0, 0))
        acc[g] = (tot + ms, max(n, cnt))
    per = {}
    for g, (ms, n) in acc.items():
        if n:
            per[g] = ms / steps if g == "energy_sums" else ms / n
    return per


def per_kernel(times, steps):
    """Milliseconds per step of each timed kernel id: per timed launch,
    except the energy accumulation (once per ring flush, every flush timed),
    whose total is divided by the timed steps."""
    return {k: (ms / steps if k == "energy_accum" else ms / max(cnt, 1))
            for k, (ms, cnt) in times.items()}


def angular_flops(case, variant, N):
    """Algorithmic flops of the angular term per step: 4 n_gamma N for the
    unfolded contraction (two gamma sums of 2 n_gamma flops per grid value,
    n_j = n_gamma).  The parity-folded kernel (variant 0) sums over half the
    angles in each class, 2 n_gamma flops per value, and computes only the
    16-column groups of the (r, R) plane that hold a compressed-grid point
    at some angle; it is credited with those columns only."""
    if variant != 0:
        return 4.0 * case.nangu * N
    n1, n2 = case.npun1, case.npun2
    live = np.zeros(n1 * n2, dtype=bool)
    for ia in range(case.nangu):
        n = int(case.npunreal[ia])
        live[(case.indr1[ia, :n].astype(np.int64) - 1) * n2 + case.indr2[ia, :n] - 1] = True
    ngroups = int(np.count_nonzero(np.pad(live, (0, (-live.size) % 16)).reshape(-1, 16).any(1)))
    cols = min(16 * ngroups, n1 * n2)
    return 2.0 * case.nangu * case.nangu * cols * case.ncan


def run_gpu(case, precision, steps, warmup, timing, tune=0, options=None):
    from madwave3_amd.lib import OPTION_NAMES, Propagator

    g = Propagator(case, precision)
    g.begin(case.rpaq0)
    for name, value in (options or {}).items():
        g.set_option(OPTION_NAMES.index(name), value)
    tuned = None
    if tune:
        # the autotuner picks the kernel variants and stream/graph layout from
        # timed trial steps; the state is left as it was
        opts, ms = g.autotune(tune)
        tuned = {"options": opts, "trial_ms_per_step": ms}
    if warmup:
        g.steps(warmup)
    if timing:
        # HIP events around every kernel group on every TIMING_EVERY-th step of
        # the timed region (the event records of every step would add ~15% to it)
        g.timing(True, every=TIMING_EVERY)
    t0 = time.perf_counter()
    g.steps(steps)
    dt = time.perf_counter() - t0
    times = g.kernel_times() if timing else {}
    P = g.flux_result()[0]
    nflux = g.nflux
    if tuned is None:
        tuned = {"options": g.options()}
    g.close()
    return dt, times, P, nflux, tuned


def build_case_j(J):
    from madwave3_amd import leps

    return leps.build(leps.config(4, J=J))


def run_batch(cases, precision, steps, warmup):
    """All contexts advanced together with mw3_cheby_steps_batch; seconds for
    `steps` batched steps."""
    from madwave3_amd.lib import Propagator, steps_batch

    props = [Propagator(c, precision) for c in cases]
    for g, c in zip(props, cases):
        g.begin(c.rpaq0)
    if warmup:
        steps_batch(props, warmup)
    t0 = time.perf_counter()
    steps_batch(props, steps)
    dt = time.perf_counter() - t0
    for g in props:
        g.close()
    return dt


def run_sequential(cases, precision, steps, warmup):
    """Each context alone, one after the other; summed seconds for `steps`
    steps of every context."""
    from madwave3_amd.lib import Propagator

    tot = 0.0
    for c in cases:
        g = Propagator(c, precision)
        g.begin(c.rpaq0)
        if warmup:
            g.steps(warmup)
        t0 = time.perf_counter()
        g.steps(steps)
        tot += time.perf_counter() - t0
        g.close()
    return tot


def main_batched(a):
    """Configuration 4: J = 0..23, Omega = 0..min(J, 6) (odd J from 1 at
    parity +1), 160 r x 320 R x 80 gamma per Omega block, all in one batch."""
    from madwave3_amd.lib import MW3_FP64, MW3_MIXED

    t0 = time.perf_counter()
    cases = [build_case_j(J) for J in range(24)]
    nblk = sum(c.ncan for c in cases)
    N = sum(c.ncan * c.nangu * c.npun1 * c.npun2 for c in cases)
    print("# 24 J values, %d Omega blocks built in %.1fs" % (nblk, time.perf_counter() - t0),
          file=sys.stderr, flush=True)
    dmx = run_batch(cases, MW3_MIXED, a.steps, a.warmup)
    print("# batched mixed: %.3f ms/step" % (1e3 * dmx / a.steps), file=sys.stderr, flush=True)
    d64 = run_batch(cases, MW3_FP64, a.steps, a.warmup)
    print("# batched fp64: %.3f ms/step" % (1e3 * d64 / a.steps), file=sys.stderr, flush=True)
    dseq = run_sequential(cases, MW3_FP64, a.steps, a.warmup)
    print("# sequential fp64: %.3f ms/step" % (1e3 * dseq / a.steps), file=sys.stderr, flush=True)
    line = {
        "metric": "aggregate Chebyshev H.psi steps/s of the J=0..23 batch (every Omega block "
                  "advanced one step)",
        "value": a.steps / dmx, "unit": "steps/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": 1e3 * dmx / a.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 operator terms (the inverse R pass and the folded angular kernel multiply on "
                 "the fp16 matrix cores as compensated hi/lo pairs, 22-bit operands, f32 sums), "
                 "f64 Chebyshev state/update/accumulators",
        "data": "synthetic (LEPS surface, generated in memory)",
        "config": {"workload": "configuration 4: J = 0..23 batched, %d Omega blocks of "
                               "160 r x 320 R x 80 gamma (%.0fM grid values)" % (nblk, N / 1e6),
                   "grid_values": N, "parallelism": "single GPU, %d contexts" % len(cases)},
        "fp64_batched_steps_per_s": a.steps / d64,
        "fp64_sequential_steps_per_s": a.steps / dseq,
        "batched_fp64_vs_sequential_fp64": dseq / d64,
        "batched_mixed_vs_sequential_fp64": dseq / dmx,
    }
    print(json.dumps(line))


def run_full(case, precision, nsteps, label, chunk=5120):
    """One propagation of nsteps Chebyshev steps (the configuration's full
    length), in chunks with a progress line each; returns the results the
    reference prints at the analysis line."""
    from madwave3_amd.lib import Propagator

    g = Propagator(case, precision)
    g.begin(case.rpaq0)
    norms, autoc = [], []
    t0 = time.perf_counter()
    done = 0
    while done < nsteps:
        n = min(chunk, nsteps - done)
        nn, au = g.steps(n, want_auto=case.iphoto > 0)
        norms.append(nn)
        if au is not None:
            autoc.append(au)
        done += n
        print("# %s: %d / %d steps, %.1fs" % (label, done, nsteps, time.perf_counter() - t0),
              file=sys.stderr, flush=True)
    out = {"seconds": time.perf_counter() - t0, "norms": np.concatenate(norms),
           "P": g.flux_result()[0], "zs2": g.zs2()}
    if autoc:
        out["auto"] = np.concatenate(autoc)
    if getattr(case, "iprod", 0) == 2:
        out["zs2prod"] = g.zs2prod()
    g.close()
    return out


def spectrum_of(case, autoc, ne=2000):
    """cheby-spectra's sum over the autocorrelation (auto[0] = 1) on ne
    energies across the propagator's spectral range (98% of it, where the
    Chebyshev weight 1/sqrt(1 - Es^2) stays finite), in the reference's units
    before the cross-section factor (SRC/photo/cheby-spectra.f:58-97)."""
    from madwave3_amd import consts as C
    from madwave3_amd.lib import cheby_spectrum

    a = np.concatenate([[1.0], autoc])
    e = case.emindlt + case.delta2 * np.linspace(-0.98, 0.98, ne)
    return cheby_spectrum(a, case.emindlt, case.delta2, C.hbr, 0.0, e)


def main_accuracy(a):
    """Full-length mixed against FP64 on the GPU for configuration a.config:
    the FP64 propagator is the reference of the long runs (it matches the CPU
    oracle over the first steps at full size, tests/test_gpu_configs.py).
    Prints one JSON line with the largest differences of what the run
    reports: P(E), the reactant state-resolved probabilities |S|^2, the
    product v'j'-resolved probabilities (iprod = 2), the norm series and,
    for photodissociation, the autocorrelation and the spectrum."""
    from madwave3_amd import printprob
    from madwave3_amd.lib import MW3_FP64, MW3_MIXED

    case = build_case(a.config)
    nsteps = a.full_steps or int(case.ntimes) * (int(case.nloop) + 1)
    print("# accuracy: configuration %d %s, %d steps" % (a.config, case, nsteps), file=sys.stderr,
          flush=True)
    r64 = run_full(case, MW3_FP64, nsteps, "fp64")
    rmx = run_full(case, MW3_MIXED, nsteps, "mixed")
    res = {"config": a.config, "workload": workload_name(a.config, case,
                                                         case.ncan * case.nangu * case.npun1 * case.npun2),
           "steps": nsteps, "fp64_seconds": r64["seconds"], "mixed_seconds": rmx["seconds"],
           "max_abs_dP": float(np.max(np.abs(rmx["P"] - r64["P"]))),
           "max_P": float(np.max(np.abs(r64["P"]))),
           "max_rel_dnorm": float(np.max(np.abs(rmx["norms"] - r64["norms"]) / r64["norms"])),
           "final_norm": float(r64["norms"][-1])}
    s64, t64 = printprob.reactant_probabilities(case, r64["zs2"])
    smx, tmx = printprob.reactant_probabilities(case, rmx["zs2"])
    res["max_abs_dS2_reactant"] = float(np.max(np.abs(smx - s64)))
    res["max_S2_reactant"] = float(np.max(np.abs(s64)))
    if "zs2prod" in r64:
        p64, q64, _ = printprob.product_probabilities(case, r64["zs2prod"])
        pmx, qmx, _ = printprob.product_probabilities(case, rmx["zs2prod"])
        res["max_abs_dS2_product"] = float(np.max(np.abs(pmx - p64)))
        res["max_S2_product"] = float(np.max(np.abs(p64)))
        res["max_abs_dS2no"] = float(np.max(np.abs(qmx - q64)))
    if "auto" in r64:
        res["max_abs_dauto"] = float(np.max(np.abs(rmx["auto"] - r64["auto"])))
        sp64 = spectrum_of(case, r64["auto"])
        spmx = spectrum_of(case, rmx["auto"])
        res["max_rel_dspectrum"] = float(np.max(np.abs(spmx - sp64)) / np.max(np.abs(sp64)))
    print(json.dumps(res))


def workload_name(cfg, case, N):
    kind = {1: "J=0 collision", 2: "J=0 collision", 3: "J=%d collision, Omega=0..%d coupled by "
            "Coriolis" % (case.Jtot, case.iommax), 5: "%d coupled electronic states"
            % case.nelecmax}.get(cfg, "collision")
    prod = ""
    if getattr(case, "iprod", 0) == 2:
        prod = ", product states v'<=%d j'<=%d" % (case.nvmaxprod, case.jmaxprod)
    return ("configuration %d: %s, %d r x %d R x %d gamma, %d channel(s) (%.1fM grid values)%s"
            % (cfg, kind, case.npun1, case.npun2, case.nangu, case.ncan, N / 1e6, prod))


def copy_calibration(mib):
    """A plain device copy of known size (torch.neg over float32, reads and
    writes mib MiB each) run three times, so that hardware byte counters
    collected over the same process can be scaled against known traffic."""
    import torch

    n = mib * (1 << 20) // 4
    x = torch.ones(n, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    for _ in range(3):
        torch.neg(x, out=y)
    torch.cuda.synchronize()
    del x, y


def load_traffic(kernel):
    """HBM bytes per step of a kernel group from the committed byte-counter
    summary (profiles/traffic_cfg2.json), or None."""
    path = os.path.join(ROOT, "profiles", "traffic_cfg2.json")
    try:
        with open(path) as f:
            return json.load(f)["mixed"][kernel]["hbm_bytes_per_step"]
    except (OSError, KeyError, ValueError):
        return None


def load_accuracy(cfg):
    """Full-length mixed vs FP64 differences of configuration cfg from the
    committed summary of `bench.py --accuracy` runs, or None."""
    path = os.path.join(ROOT, "profiles", "accuracy.json")
    try:
        with open(path) as f:
            return json.load(f).get(str(cfg))
    except (OSError, ValueError):
        return None


def cpu_baseline(case, nsteps):
    from oracle.oracle import Oracle, set_threads

    nthr = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or os.cpu_count() or 1
    nthr = set_threads(nthr)
    o = Oracle(case)
    o.begin(case.rpaq0)
    t0 = time.perf_counter()
    o.steps(nsteps)
    dt = time.perf_counter() - t0
    return {"value": nsteps / dt, "unit": "steps/s", "cores": nthr, "kind": "port",
            "sample": "%d FP64 Chebyshev steps of the same workload by the C/OpenMP "
                      "oracle (dense sine-matrix DST-I, DVR angular sums)" % nsteps}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--config", type=int, default=2)
    ap.add_argument("--cpu-steps", type=int, default=60)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-timing", action="store_true")
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="run a copy of this many MiB first (byte-counter calibration)")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default execution options instead of autotuning them")
    ap.add_argument("--option", action="append", default=[], metavar="NAME=VALUE",
                    help="set an execution option (graphs, radial, angular, side_stream, "
                         "radial_grid, angular_grid) in both propagators instead of autotuning; "
                         "repeatable")
    ap.add_argument("--accuracy", action="store_true",
                    help="full-length mixed vs FP64 run of --config instead of the timing")
    ap.add_argument("--full-steps", type=int, default=0,
                    help="with --accuracy: steps instead of the configuration's ntimes (nloop+1)")
    a = ap.parse_args()
    if a.gpus != 1:
        print("bench.py measures one GPU; this build has no multi-GPU path (--gpus %d)" % a.gpus,
              file=sys.stderr)
        sys.exit(2)
    from madwave3_amd.lib import MW3_FP64, MW3_MIXED

    if a.accuracy:
        main_accuracy(a)
        return
    if a.config == 4:
        main_batched(a)
        return
    if a.copy_mib:
        copy_calibration(a.copy_mib)
    t0 = time.perf_counter()
    case = build_case(a.config)
    print("# case %s built in %.1fs" % (case, time.perf_counter() - t0), file=sys.stderr, flush=True)
    N = case.ncan * case.nangu * case.npun1 * case.npun2
    opts = dict((kv.split("=")[0], int(kv.split("=")[1])) for kv in a.option)
    tune = 0 if (a.no_autotune or opts) else 48
    dt64, tim64, P64, nflux, tuned64 = run_gpu(case, MW3_FP64, a.steps, a.warmup,
                                               not a.no_timing, tune, opts)
    print("# fp64: %.3f ms/step %s" % (1e3 * dt64 / a.steps, tuned64), file=sys.stderr, flush=True)
    dtmx, timmx, Pmx, _, tunedmx = run_gpu(case, MW3_MIXED, a.steps, a.warmup,
                                           not a.no_timing, tune, opts)
    print("# mixed: %.3f ms/step %s" % (1e3 * dtmx / a.steps, tunedmx), file=sys.stderr,
          flush=True)
    dP = float(np.max(np.abs(Pmx - P64)))
    acc = load_accuracy(a.config)

    # roofline of the dominant kernel group of the headline (mixed) run: the
    # largest per-step time among the groups on the library's main stream
    per = kernel_groups(timmx, a.steps)
    roof = None
    kern = {}
    if per:
        s_op = 4  # FP32 operator I/O
        # the update visits only the compressed grid (the reference's loop over
        # ntotproc): N_c = case.ntot points over all channels
        n_c = float(case.ntot) / case.nelecmax
        ne = case.nelecmax
        # product projection: transRgjac read by the update (operator precision)
        tprod = s_op * n_c * ne if getattr(case, "iprod", 0) == 2 else 0.0
        nprod_l = ((case.nangproj1 - case.nangproj0 + 1) * (case.n2prod1 - case.n2prod0 + 1)
                   if tprod else 0)
        work = {
            # (a) radial kinetic: 2 s N bytes per step
            "radial_dst": ("hbm", 2.0 * s_op * N, "GB/s"),
            # (b) angular: 4 n_gamma N flops per step; the parity-folded
            # kernel does half of them, on the columns it computes
            "angular": ("mfma", angular_flops(case, tunedmx["options"].get("angular", 1), N),
                        "TFLOP/s"),
            # (c) fused V + absorption + update: 8 + 8 + s + (V) + 8 bytes per
            # state-point on the N_c compressed points, V being the unique
            # entries of the symmetric n_e x n_e matrix per spatial point
            # (s_V (n_e + 1)/2 per state-point: 4 B at one state)
            "update_fused": ("hbm", n_c * ne * (8 + 8 + s_op + 4.0 * (ne + 1) / 2 + 8) + tprod,
                             "GB/s"),
            # (e) energy projection: 8 netot n_flux flops per step (time only)
            "energy_sums": ("hbm", 16.0 * nflux, "GB/s"),
            "projections": ("hbm", 8.0 * case.nangu * case.npun1, "GB/s"),
            # flux columns and norm of the step (second stream): the flux
            # line values written to the ring (time only, like energy_sums)
            "update_tail": ("hbm", 16.0 * nflux, "GB/s"),
            # (d) Coriolis stencil in the FBR: 2 s N bytes over the Omega-coupled blocks
            "coriolis": ("hbm", 2.0 * s_op * N, "GB/s"),
            # product reduction: 64 partials per update item, dtrans, ring row
            "products": ("hbm", 8.0 * 3 * max(nprod_l, 1), "GB/s"),
        }
        for k, ms in per.items():
            bound, w, unit = work[k]
            sec = ms / 1e3
            ach = w / sec / (1e9 if unit == "GB/s" else 1e12)
            peak = HBM_PEAK_GBS if unit == "GB/s" else FP32_PEAK_TF
            kern[k] = {"ms_per_step": ms, "achieved": ach, "unit": unit, "peak": peak,
                       "frac": ach / peak}
            if k == "angular" and tunedmx["options"].get("angular", 1) == 0:
                # the folded kernel multiplies on the fp16 matrix cores (three
                # fp16 products per f32 one); its flops are still priced at the
                # f32 peak, and its HBM rate is the psi read plus the result
                # written over the columns it computes
                cols = angular_flops(case, 0, N) / (2.0 * case.nangu * case.nangu)
                kern[k]["products"] = "fp16 x3 (compensated)"
                kern[k]["hbm_GBps"] = 2.0 * s_op * case.nangu * cols / sec / 1e9
            tr = load_traffic(k) if a.config == 2 else None
            if tr and unit == "GB/s":
                # HBM bytes the counters saw per step over the algorithmic bytes
                kern[k]["traffic"] = tr
                kern[k]["traffic_vs_algorithmic"] = tr / w
        main = [k for k in per if k not in ("projections", "energy_sums", "products",
                                            "update_tail")]
        dom = max(main, key=lambda k: per[k])
        d = kern[dom]
        roof = {"bound": work[dom][0], "achieved": d["achieved"], "peak": d["peak"],
                "unit": d["unit"], "frac": d["frac"], "traffic": load_traffic(dom) if a.config == 2 else None,
                "kernel": dom, "launches_per_unit": GROUP_LAUNCHES[dom],
                "unit_of_work": "one Chebyshev step of the group (its per-step work formula, "
                                "as in `kernels`); traffic = HBM bytes per step from the "
                                "committed byte-counter summary"}
    cpu = None
    if not a.no_cpu:
        cpu = cpu_baseline(case, a.cpu_steps)
    steps_s = a.steps / dtmx
    line = {
        "metric": "Chebyshev H.psi steps/s (mixed vs FP64 GPU) at <=1e-4 P(E) err; HBM GB/s",
        "value": steps_s,
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": 1e3 * dtmx / a.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 operator terms (the inverse R pass and the folded angular kernel multiply on "
                 "the fp16 matrix cores as compensated hi/lo pairs, 22-bit operands, f32 sums), "
                 "f64 Chebyshev state/update/accumulators",
        "data": "synthetic (LEPS surface, generated in memory)",
        "config": {"workload": workload_name(a.config, case, N),
                   "grid_values": N, "active_fraction": float(case.ntot) / N,
                   "parallelism": "single GPU"},
        "fp64_steps_per_s": a.steps / dt64,
        "fp64_ms_per_step": 1e3 * dt64 / a.steps,
        # the metric's error bound, measured over the configuration's full
        # length (bench.py --accuracy; committed in profiles/accuracy.json)
        "mixed_vs_fp64_max_abs_dP": acc.get("max_abs_dP") if acc else None,
        "accuracy_full_run": acc,
        # the same difference after this run's warmup + timed steps only
        "short_run_max_abs_dP": dP,
        "short_run_steps": a.warmup + a.steps,
        "autotune": {"mixed": tunedmx, "fp64": tuned64},
        "roofline": roof,
        "kernels": kern,
        "kernels_fp64": {k: {"ms_per_step": v} for k, v in kernel_groups(tim64, a.steps).items()},
        "kernel_ms_per_step": per_kernel(timmx, a.steps),
        "fp64_kernel_ms_per_step": per_kernel(tim64, a.steps),
        "energy_accum_ms_per_flush": {k: ms / max(cnt, 1) for k, (ms, cnt) in
                                      (("mixed", timmx.get("energy_accum", (0.0, 0))),
                                       ("fp64", tim64.get("energy_accum", (0.0, 0))))},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()

"""Full-length FP64 oracle run of configuration 1 (128 r x 256 R x 60 gamma,
J = 0, LEPS surface, 20000 Chebyshev steps k = 2 .. 20001, products resolved
with iprod = 2), written as the fixture cfg1_oracle_full.json: P(E) on the
500 energies, the per-state sums, |psi_k|^2 every 100 steps,
|zS2(E, v=0, j)|^2, and the product side: S2no(E), the product vibrational
distribution vibprod(v', E) and |zS2prod(E, v'=0, j')|^2.  The GPU tests
compare both GPU propagators with it.  Runs on the CPU (about 17 minutes on 8
cores):

    python tests/golden/make_cfg1_oracle.py
"""
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from madwave3_amd import leps, model, printprob  # noqa: E402
from madwave3_amd import namelist as NL  # noqa: E402
from oracle.oracle import Oracle, set_threads  # noqa: E402


def main():
    p = leps.config(1)
    nsteps = int(p["ntimes"]) * (int(p["nloop"]) + 1)
    text, pot, fdat = leps.make_inputs(p)
    case = model.build_case(NL.parse(text), pot, fdat)
    nthr = set_threads(os.cpu_count() or 1)
    o = Oracle(case)
    o.begin(case.rpaq0)
    t0 = time.perf_counter()
    norms = np.zeros(nsteps)
    done = 0
    while done < nsteps:
        n = min(1000, nsteps - done)
        out = o.steps(n)
        norms[done:done + n] = out[0] if isinstance(out, tuple) else out
        done += n
        print("%d / %d steps, %.0f s" % (done, nsteps, time.perf_counter() - t0), flush=True)
    P, Pe = o.flux_result()
    res = {
        "workload": "configuration 1: 128 r x 256 R x 60 gamma, J=0, LEPS, %d steps" % nsteps,
        "steps": nsteps, "threads": nthr,
        "etot_zot": [float(x) for x in np.asarray(case.etotS2)],
        "P": [float(x) for x in P],
        "P_elec": np.asarray(Pe).tolist(),
        "norm_every_100": [float(x) for x in norms[99::100]],
        # |zS2(E, v=0, j)|^2 of channel 1, j = jini..jmax
        "zs2_v0_abs2": (np.abs(o.zs2()[0, :, 0, :]) ** 2).tolist(),
    }
    zp = o.zs2prod()
    _, S2no, vib = printprob.product_probabilities(case, zp)
    res.update({
        "S2no": [float(x) for x in S2no],
        "vibprod": vib[0].tolist(),
        # |zS2prod(E, v'=0, j')|^2 of Omega' = 0, j' = jiniprod..jmaxprod
        "zs2prod_v0_abs2": (np.abs(zp[0, :, 0, :]) ** 2).tolist(),
    })
    with open(os.path.join(HERE, "cfg1_oracle_full.json"), "w") as f:
        json.dump(res, f)
    print("wrote cfg1_oracle_full.json", flush=True)


if __name__ == "__main__":
    main()

"""The H3 neural-network surface of the reference's HD+H example
(EXAMPLES/HD+H_collision/pes: FFmodules.f90 NN_VH3 / generate_input,
FFroutines.f step, pothead.f potelebond), restated for the golden-example
parity test.  Its parameters are the data of tests/golden/h3_nn_params.json
(made by tests/golden/make_h3_nn_params.py).

Surface, distances in bohr, energy in hartree:
  x_i = exp(-0.7 d_i) for the three atom-atom distances d_i;
  invariants p = (sum x, sum x^2, sum x^3), standardised (p - mean)/std;
  two sigmoid layers of 20 units and a linear output y;
  V = (y + 1)/2 (y_max - y_min) + y_min,
and potelebond(r01, r02, cos) = V(r01, r02, r12) - V(100, 100, 1.4), with r12
from r01, r02 and the angle at atom 0 and then each distance clamped below at
0.25 bohr.  The example is compiled with
default-real (single precision) literals, so the parameters and the
constant 0.7 are their single-precision values.
"""
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_P = None


def _sp(x):
    return np.float64(np.float32(x))


def params():
    global _P
    if _P is None:
        raw = json.load(open(os.path.join(HERE, "golden", "h3_nn_params.json")))
        _P = {k: np.vectorize(lambda s: _sp(float(s)))(np.array(v, dtype=object)).astype(float)
              for k, v in raw.items()}
    return _P


def nn_energy(d1, d2, d3):
    """NN_VH3: hartree, distances in bohr (arrays broadcast together)."""
    P = params()
    alpha = _sp(0.7)
    x = np.exp(-alpha * np.stack(np.broadcast_arrays(d1, d2, d3)).astype(float))
    inv = np.stack([x.sum(0), (x * x).sum(0), (x * x * x).sum(0)])
    inv = (inv - P["xstandard"][:, 0].reshape(3, *([1] * (inv.ndim - 1)))) / \
        P["xstandard"][:, 1].reshape(3, *([1] * (inv.ndim - 1)))
    z = np.moveaxis(inv, 0, -1)
    h1 = 1.0 / (1.0 + np.exp(-(P["b"] + z @ P["w"])))
    h2 = 1.0 / (1.0 + np.exp(-(P["b1"] + h1 @ P["w1"])))
    y = P["b2"][0] + h2 @ P["w2"][:, 0]
    lin = P["ylinear"]
    return 0.5 * (y + 1.0) * (lin[0] - lin[1]) + lin[1]


def potelebond(r01, r02, costet):
    """The example's potelebond: V(r01, r02, angle at atom 0) - V(asymptote)."""
    r01 = np.asarray(r01, dtype=float)
    r02 = np.asarray(r02, dtype=float)
    # the 1-2 distance from the distances as given; then each is clamped
    r12 = np.sqrt(np.maximum(r01 * r01 + r02 * r02 - 2.0 * r01 * r02 * costet, 0.0))
    r01, r02, r12 = (np.maximum(x, 0.25) for x in (r01, r02, r12))
    return nn_energy(r12, r02, r01) - nn_energy(100.0, 100.0, 1.4)

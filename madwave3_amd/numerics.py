"""Numerical routines of MADWAVE3's set-up, restated for the host side.

Each function computes what a routine of SRC/liboctdyn.f computes (cited per
function), written from the mathematics of that quantity, so the host builds the
same grids, angular functions and packets as the reference.
"""
import math

import numpy as np

__all__ = ["gauleg", "dwigner", "dwigner_vec", "rgauscolini", "besph2", "noptfft",
           "splset", "splinqq", "lagrange_slope"]


def gauleg(n):
    """Gauss-Legendre nodes and weights on [-1, 1], nodes ascending.

    Same quadrature as GAULEG (SRC/liboctdyn.f:174-210).  The roots of P_n are
    found by Newton's method from the asymptotic guesses
    cos(pi (k - 1/4)/(n + 1/2)), k = 1..ceil(n/2), all roots at once; P_n and
    P_{n-1} come from Bonnet's recurrence
        m P_m(z) = (2m - 1) z P_{m-1}(z) - (m - 1) P_{m-2}(z),
    P_n'(z) = n (z P_n - P_{n-1})/(z^2 - 1), and a root is final once its
    Newton step is at most 1e-15.  The weight of a root z is
    2/((1 - z^2) P_n'(z')^2), with P_n' taken at the iterate z' before the last
    step.  Written with the recurrence in this operation order, the nodes and
    weights are bitwise those of the reference build.
    Returns (weights, nodes).
    """
    half = (n + 1) // 2
    root = np.array([math.cos(math.pi * (k - 0.25) / (n + 0.5)) for k in range(1, half + 1)])
    slope = np.zeros(half)
    todo = np.arange(half)
    while todo.size:
        z = root[todo]
        p_lo, p_hi = np.zeros_like(z), np.ones_like(z)
        for m in range(1, n + 1):
            p_lo, p_hi = p_hi, ((2.0 * m - 1.0) * z * p_hi - (m - 1.0) * p_lo) / m
        dpn = n * (z * p_hi - p_lo) / (z * z - 1.0)
        z_new = z - p_hi / dpn
        root[todo] = z_new
        slope[todo] = dpn
        todo = todo[np.abs(z_new - z) > 1e-15]
    w_half = 2.0 / ((1.0 - root * root) * slope * slope)
    nodes = np.empty(n)
    weights = np.empty(n)
    nodes[:half] = -root          # the guesses run from the largest root down
    nodes[n - half:] = root[::-1]
    weights[:half] = w_half
    weights[n - half:] = w_half[::-1]
    return weights, nodes


def _dsmall_first(j2, m2, mp2, cb, sb):
    """d^j_{m,m'}(beta) at the lowest j = max(|m|, |m'|) (all arguments
    doubled: j2 = 2j, ...), with cb = cos(beta/2), sb = sin(beta/2):
        d^j_{m,m'} = s sqrt(C(2j, j + m')) cb^(j+m'') sb^(j-m'')
    in the Condon-Shortley convention, where for m = +-j the binomial and the
    powers take m' (and for m' = +-j they take m), and the sign s is
    (-1)^(j - m') when m = j, (-1)^(j + m) when m' = -j, +1 otherwise."""
    if abs(m2) == j2:
        other = mp2
        if m2 == j2:
            pc, ps = (j2 + other) // 2, (j2 - other) // 2
            sign = -1.0 if ((j2 - other) // 2) % 2 else 1.0
        else:
            pc, ps = (j2 - other) // 2, (j2 + other) // 2
            sign = 1.0
    else:
        other = m2
        if mp2 == j2:
            pc, ps = (j2 + other) // 2, (j2 - other) // 2
            sign = 1.0
        else:
            pc, ps = (j2 - other) // 2, (j2 + other) // 2
            sign = -1.0 if ((j2 + other) // 2) % 2 else 1.0
    amp = math.sqrt(float(math.comb(j2, (j2 + other) // 2)))
    return sign * amp * cb ** pc * sb ** ps


def dwigner_vec(jmax2, m2, mp2, beta):
    """Wigner small-d functions d^j_{m,m'}(beta) for 2j = 0..jmax2 over an
    array of angles: d[2j, i] = d^j_{m,m'}(beta[i]) (m2 = 2m, mp2 = 2m').

    The same functions as dwigner (SRC/liboctdyn.f:3-171), evaluated here from
    the closed form at j = max(|m|, |m'|) and the standard three-term
    recurrence in j (e.g. Varshalovich et al., Quantum Theory of Angular
    Momentum, 4.8.2):
        j sqrt(((j+1)^2 - m^2)((j+1)^2 - m'^2)) d^{j+1}
          = (2j + 1)(j (j+1) cos(beta) - m m') d^j
            - (j + 1) sqrt((j^2 - m^2)(j^2 - m'^2)) d^{j-1};
    at m = m' = 0 it starts from P_0 = 1, P_1 = cos(beta) (Legendre).  Rows
    with 2j below max(|m2|, |mp2|), or of the other parity, are 0.
    """
    beta = np.asarray(beta, dtype=float)
    out = np.zeros((jmax2 + 1, beta.size))
    j2min = max(abs(m2), abs(mp2))
    if (j2min - m2) % 2 or (j2min - mp2) % 2:
        raise ValueError("dwigner: 2j, 2m and 2m' must have one parity")
    if j2min > jmax2:
        return out
    c = np.cos(beta)
    m, mp = 0.5 * m2, 0.5 * mp2
    if j2min == 0:
        out[0] = 1.0
        if jmax2 >= 2:
            out[2] = c
        j_from = 1.0
    else:
        out[j2min] = _dsmall_first(j2min, m2, mp2, np.cos(0.5 * beta), np.sin(0.5 * beta))
        j_from = 0.5 * j2min
    lo = int(round(2 * j_from))
    for jj2 in range(lo, jmax2 - 1, 2):
        j = 0.5 * jj2
        below = out[jj2 - 2] if jj2 >= 2 else 0.0
        upper = (2.0 * j + 1.0) * (j * (j + 1.0) * c - m * mp) * out[jj2]
        lower = (j + 1.0) * math.sqrt(max((j * j - m * m) * (j * j - mp * mp), 0.0)) * below
        scale = j * math.sqrt(((j + 1.0) ** 2 - m * m) * ((j + 1.0) ** 2 - mp * mp))
        out[jj2 + 2] = (upper - lower) / scale
    return out


def dwigner(jmax2, m2, mp2, beta):
    """d^j_{m,m'}(beta) at one angle for 2j = 0..jmax2, indexed by 2j (see
    dwigner_vec; SRC/liboctdyn.f:3-171)."""
    return dwigner_vec(jmax2, m2, mp2, np.array([float(beta)]))[:, 0]


def lagrange_slope(x, y, at):
    """Derivative at the node x[at] of the polynomial interpolating (x, y),
    the quantity DLAGRA returns (SRC/liboctdyn.f:960-983).  Barycentric form:
    with w_i = 1/prod_{k != i}(x_i - x_k),
        p'(x_a) = sum_{i != a} (w_i/w_a) y_i/(x_a - x_i)
                  + y_a sum_{i != a} 1/(x_a - x_i)."""
    x = np.asarray(x, dtype=float)
    y = np.asarray(y, dtype=float)
    diff = x[:, None] - x[None, :]
    np.fill_diagonal(diff, 1.0)
    w = 1.0 / np.prod(diff, axis=1)
    others = np.arange(len(x)) != at
    gap = x[at] - x[others]
    return float(np.sum(w[others] / w[at] * y[others] / gap) + y[at] * np.sum(1.0 / gap))


def splset(f, x):
    """Knot slopes of the cubic spline through (x, f) with clamped ends, the
    F(:,2) column that SPLSET fills (SRC/liboctdyn.f:342-403).

    End slopes: the derivative of the cubic through the first (last) four
    knots at the first (last) knot.  Interior slopes s_i: C^2 continuity of
    the piecewise Hermite cubic,
        h_i s_{i-1} + 2 (h_{i-1} + h_i) s_i + h_{i-1} s_{i+1}
          = 3 (h_i D_{i-1} + h_{i-1} D_i),
    h_i = x_{i+1} - x_i, D_i = (f_{i+1} - f_i)/h_i, solved as a banded system.
    """
    from scipy.linalg import solve_banded

    f = np.asarray(f, dtype=float)
    x = np.asarray(x, dtype=float)
    n = len(x)
    h = np.diff(x)
    dd = np.diff(f) / h
    s = np.zeros(n)
    s[0] = lagrange_slope(x[:4], f[:4], 0)
    s[-1] = lagrange_slope(x[-4:], f[-4:], 3)
    m = n - 2
    if m <= 0:
        return s
    rhs = 3.0 * (h[1:] * dd[:-1] + h[:-1] * dd[1:])
    rhs[0] -= h[1] * s[0]
    rhs[-1] -= h[-2] * s[-1]
    ab = np.zeros((3, m))
    ab[0, 1:] = h[:-2]             # superdiagonal: h_{i-1} s_{i+1}
    ab[1, :] = 2.0 * (h[:-1] + h[1:])
    ab[2, :-1] = h[2:]             # subdiagonal: h_i s_{i-1}
    s[1:-1] = solve_banded((1, 1), ab, rhs)
    return s


def splinqq(f, d, x, r):
    """Value at the points r of the spline with knot values f and slopes d,
    as SPLINQQ evaluates it (SRC/liboctdyn.f:298-341).

    Inside the knots (and before the first): the cubic Hermite interpolant
    on the interval [x_{i-1}, x_i) that holds r (the first interval below
    x_0), t = (r - x_{i-1})/h,
        f_{i-1} H00(t) + f_i H01(t) + h (d_{i-1} H10(t) + d_i H11(t)).
    At or beyond the last knot x_e: the tail A (x_e/r)^6 + B (x_e/r)^8 whose
    value and slope match f_e and d_e at x_e, B = -(3 f_e + x_e d_e/2),
    A = f_e - B.
    """
    f = np.asarray(f, dtype=float)
    d = np.asarray(d, dtype=float)
    x = np.asarray(x, dtype=float)
    r = np.asarray(r, dtype=float)
    out = np.empty_like(r)
    tail = r >= x[-1]
    ri = r[~tail]
    hi_k = np.clip(np.searchsorted(x, ri, side="right"), 1, len(x) - 1)
    lo_k = hi_k - 1
    h = x[hi_k] - x[lo_k]
    t = (ri - x[lo_k]) / h
    t2 = t * t
    h00 = (2.0 * t - 3.0) * t2 + 1.0
    h01 = 1.0 - h00
    h10 = t * (t - 1.0) ** 2
    h11 = t2 * (t - 1.0)
    out[~tail] = f[lo_k] * h00 + f[hi_k] * h01 + h * (d[lo_k] * h10 + d[hi_k] * h11)
    if tail.any():
        xe, fe, de = x[-1], f[-1], d[-1]
        b = -(3.0 * fe + 0.5 * xe * de)
        a = fe - b
        q = (xe / r[tail]) ** 2
        out[tail] = (a + b * q) * q ** 3
    return out


def rgauscolini(r, rcol, alpha, xk, factor):
    """Incident packet sqrt(2) cos(k (R - R0)) factor exp(-(R - R0)^2/alpha^2).

    SRC/liboctdyn.f:2051-2066 (rgauscolini); vectorised over r.
    """
    r = np.asarray(r, dtype=float)
    arg = (r - rcol) * xk
    g = np.cos(arg) * math.sqrt(2.0)
    return g * factor * np.exp(-(r - rcol) * (r - rcol) / alpha / alpha)


def besph2(l, x):
    """Riccati-Bessel pair for an open channel, x > 0.

    Same functions as SRC/liboctdyn.f:743-843 (BESPH2 with ARG > 0):
    F = x j_l(x) ~ sin(x - l pi/2), G = x y_l(x) ~ -cos(x - l pi/2), and their
    derivatives, from scipy's spherical Bessel functions.  Returns
    (F, dF, G, dG); vectorised over x.
    """
    from scipy.special import spherical_jn, spherical_yn

    x = np.asarray(x, dtype=float)
    if np.any(x <= 0.0):
        raise ValueError("besph2: open-channel argument must be positive")
    if l == 0:
        return np.sin(x), np.cos(x), -np.cos(x), np.sin(x)
    j = spherical_jn(l, x)
    y = spherical_yn(l, x)
    dj = spherical_jn(l, x, derivative=True)
    dy = spherical_yn(l, x, derivative=True)
    return x * j, j + x * dj, x * y, y + x * dy


def noptfft(nin, ntot):
    """noptFFT (SRC/liboctdyn.f:1033-1117): smallest 2^a 3^b 5^c 7^d 11^e at
    or above nin (exponents bounded by ntot, distance below ntot), capped at
    ntot; 1 when no candidate exists."""
    exps = []
    for p in (2, 3, 5, 7, 11):
        v, n = 1, 0
        for n in range(0, 101):
            v *= p
            if v > ntot:
                break
        exps.append(n)
    nout, imin = 1, ntot
    for a11 in range(exps[4] + 1):
        for a7 in range(exps[3] + 1):
            for a5 in range(exps[2] + 1):
                for a3 in range(exps[1] + 1):
                    for a2 in range(exps[0] + 1):
                        i = 2 ** a2 * 3 ** a3 * 5 ** a5 * 7 ** a7 * 11 ** a11
                        idis = i - nin
                        if 0 <= idis < imin:
                            nout, imin = i, idis
    return min(nout, ntot)

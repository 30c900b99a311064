"""Numerical routines of MADWAVE3's set-up, restated for the host side.

Each function follows a routine of SRC/liboctdyn.f (cited per function) so the
host builds the same grids, angular functions and packets as the reference.
"""
import math

import numpy as np

__all__ = ["gauleg", "dwigner", "dwigner_vec", "rgauscolini", "besph2", "noptfft",
           "splset", "splinqq"]


def gauleg(n):
    """Gauss-Legendre nodes and weights on [-1, 1], ascending nodes.

    SRC/liboctdyn.f:174-210 (GAULEG): Newton iteration on P_n from the guess
    cos(pi (i - 1/4)/(n + 1/2)) to |dz| <= 1e-15, symmetric fill.
    Returns (weights, nodes).
    """
    w = np.zeros(n)
    x = np.zeros(n)
    x1, x2 = -1.0, 1.0
    yn = float(n)
    eps = 1.0e-15
    m = (n + 1) // 2
    xm = 0.5 * (x1 + x2)
    xl = 0.5 * (x2 - x1)
    pi = math.acos(-1.0)
    for i in range(1, m + 1):
        z = math.cos(pi * (i - 0.25) / (yn + 0.5))
        while True:
            p1, p2 = 1.0, 0.0
            for j in range(1, n + 1):
                p3 = p2
                p2 = p1
                yj = float(j)
                p1 = ((2.0 * yj - 1.0) * z * p2 - (yj - 1.0) * p3) / yj
            pp = yn * (z * p1 - p2) / (z * z - 1.0)
            z1 = z
            z = z1 - p1 / pp
            if abs(z - z1) <= eps:
                break
        x[i - 1] = xm - xl * z
        x[n - i] = xm + xl * z
        w[i - 1] = 2.0 * xl / ((1.0 - z * z) * pp * pp)
        w[n - i] = w[i - 1]
    return w, x


def dwigner(jmax2, m2, mp2, beta):
    """Wigner d^J_{M,M'}(beta) for 2J = 0..jmax2, returned indexed by 2J.

    SRC/liboctdyn.f:3-171 (dwigner): closed form at J = max(|M|,|M'|), then
    the three-term recurrence in J.  m2, mp2 are 2M and 2M'.  Entries with
    2J below max(|m2|,|mp2|) are 0.
    """
    dj = np.zeros(jmax2 + 1)
    jmin2 = max(abs(m2), abs(mp2))
    y = math.cos(beta)
    if jmin2 == 0:
        dj[0] = 1.0
        if jmax2 >= 2:
            dj[2] = y
        for ind in range(4, jmax2 + 1, 2):
            xj = ind * 0.5
            dj[ind] = (y * (2.0 * xj - 1.0) * dj[ind - 2] - (xj - 1.0) * dj[ind - 4]) / xj
        return dj
    facnum = sum(math.log(i) for i in range(1, jmin2 + 1))
    mreal2 = mp2 if jmin2 == abs(m2) else m2
    facden1 = 0.0
    if jmin2 + mreal2 > 0:
        for i in range(2, jmin2 + mreal2 + 1, 2):
            facden1 += math.log(i * 0.5)
    facden2 = 0.0
    if jmin2 - mreal2 > 0:
        for i in range(2, jmin2 - mreal2 + 1, 2):
            facden2 += math.log(i * 0.5)
    fact = math.exp(0.5 * (facnum - facden1 - facden2))
    sign = 1.0
    xsin = math.sin(beta * 0.5)
    xcos = math.cos(beta * 0.5)
    a = b = 0.0
    if jmin2 == abs(m2):
        if jmin2 == m2:
            if (jmin2 - mp2) % 2 != 0:
                raise ValueError("dwigner: complex phase")
            sign = (-1.0) ** int((jmin2 - mp2) * 0.5 + 0.5)
            a = xsin ** ((jmin2 - mp2) * 0.5)
            b = xcos ** ((jmin2 + mp2) * 0.5)
        elif jmin2 == -m2:
            a = xsin ** ((jmin2 + mp2) * 0.5)
            b = xcos ** ((jmin2 - mp2) * 0.5)
    elif jmin2 == abs(mp2):
        if jmin2 == mp2:
            a = xsin ** ((jmin2 - m2) * 0.5)
            b = xcos ** ((jmin2 + m2) * 0.5)
        elif jmin2 == -mp2:
            if (jmin2 + m2) % 2 != 0:
                raise ValueError("dwigner: complex phase")
            sign = (-1.0) ** int((jmin2 + m2) * 0.5 + 0.5)
            a = xsin ** ((jmin2 + m2) * 0.5)
            b = xcos ** ((jmin2 - m2) * 0.5)
    if jmin2 <= jmax2:
        dj[jmin2] = fact * sign * a * b
    xj = jmin2 * 0.5
    xjp = xj + 1.0
    xm = m2 * 0.5
    xmp = mp2 * 0.5
    facnum = xjp * (2.0 * xj + 1.0)
    factot = facnum / math.sqrt((xjp * xjp - xm * xm) * (xjp * xjp - xmp * xmp))
    facj = y - xm * xmp / (xj * (xj + 1.0))
    if jmin2 + 2 <= jmax2:
        dj[jmin2 + 2] = factot * facj * dj[jmin2]
        for indp in range(jmin2 + 4, jmax2 + 1, 2):
            ind = indp - 2
            indm = indp - 4
            xjp = indp * 0.5
            xj = xjp - 1.0
            facnum = xjp * (2.0 * xj + 1.0)
            factot = facnum / math.sqrt((xjp * xjp - xm * xm) * (xjp * xjp - xmp * xmp))
            facj = y - xm * xmp / (xj * (xj + 1.0))
            facjm = math.sqrt((xj * xj - xm * xm) * (xj * xj - xmp * xmp)) / (xj * (2.0 * xj + 1.0))
            dj[indp] = factot * (facj * dj[ind] - facjm * dj[indm])
    return dj


def dwigner_vec(jmax2, m2, mp2, beta):
    """dwigner over an array of angles: d[2J, i] = d^J_{M,M'}(beta[i]).

    The same closed form at J = max(|M|,|M'|) and the same recurrence in J as
    dwigner (SRC/liboctdyn.f:3-171), evaluated elementwise with numpy.
    """
    beta = np.asarray(beta, dtype=float)
    nb = beta.size
    dj = np.zeros((jmax2 + 1, nb))
    jmin2 = max(abs(m2), abs(mp2))
    y = np.cos(beta)
    if jmin2 == 0:
        dj[0] = 1.0
        if jmax2 >= 2:
            dj[2] = y
        for ind in range(4, jmax2 + 1, 2):
            xj = ind * 0.5
            dj[ind] = (y * (2.0 * xj - 1.0) * dj[ind - 2] - (xj - 1.0) * dj[ind - 4]) / xj
        return dj
    facnum = sum(math.log(i) for i in range(1, jmin2 + 1))
    mreal2 = mp2 if jmin2 == abs(m2) else m2
    facden1 = 0.0
    if jmin2 + mreal2 > 0:
        for i in range(2, jmin2 + mreal2 + 1, 2):
            facden1 += math.log(i * 0.5)
    facden2 = 0.0
    if jmin2 - mreal2 > 0:
        for i in range(2, jmin2 - mreal2 + 1, 2):
            facden2 += math.log(i * 0.5)
    fact = math.exp(0.5 * (facnum - facden1 - facden2))
    sign = 1.0
    xsin = np.sin(beta * 0.5)
    xcos = np.cos(beta * 0.5)
    a = b = np.zeros(nb)
    if jmin2 == abs(m2):
        if jmin2 == m2:
            if (jmin2 - mp2) % 2 != 0:
                raise ValueError("dwigner: complex phase")
            sign = (-1.0) ** int((jmin2 - mp2) * 0.5 + 0.5)
            a = xsin ** ((jmin2 - mp2) * 0.5)
            b = xcos ** ((jmin2 + mp2) * 0.5)
        elif jmin2 == -m2:
            a = xsin ** ((jmin2 + mp2) * 0.5)
            b = xcos ** ((jmin2 - mp2) * 0.5)
    elif jmin2 == abs(mp2):
        if jmin2 == mp2:
            a = xsin ** ((jmin2 - m2) * 0.5)
            b = xcos ** ((jmin2 + m2) * 0.5)
        elif jmin2 == -mp2:
            if (jmin2 + m2) % 2 != 0:
                raise ValueError("dwigner: complex phase")
            sign = (-1.0) ** int((jmin2 + m2) * 0.5 + 0.5)
            a = xsin ** ((jmin2 + m2) * 0.5)
            b = xcos ** ((jmin2 - m2) * 0.5)
    if jmin2 <= jmax2:
        dj[jmin2] = fact * sign * a * b
    xj = jmin2 * 0.5
    xjp = xj + 1.0
    xm = m2 * 0.5
    xmp = mp2 * 0.5
    facnum = xjp * (2.0 * xj + 1.0)
    factot = facnum / math.sqrt((xjp * xjp - xm * xm) * (xjp * xjp - xmp * xmp))
    facj = y - xm * xmp / (xj * (xj + 1.0))
    if jmin2 + 2 <= jmax2:
        dj[jmin2 + 2] = factot * facj * dj[jmin2]
        for indp in range(jmin2 + 4, jmax2 + 1, 2):
            ind = indp - 2
            indm = indp - 4
            xjp = indp * 0.5
            xj = xjp - 1.0
            facnum = xjp * (2.0 * xj + 1.0)
            factot = facnum / math.sqrt((xjp * xjp - xm * xm) * (xjp * xjp - xmp * xmp))
            facj = y - xm * xmp / (xj * (xj + 1.0))
            facjm = math.sqrt((xj * xj - xm * xm) * (xj * xj - xmp * xmp)) / (xj * (2.0 * xj + 1.0))
            dj[indp] = factot * (facj * dj[ind] - facjm * dj[indm])
    return dj


def _dlagra(x, y, ip):
    """Derivative at x[ip] of the Lagrange polynomial through (x, y)
    (SRC/liboctdyn.f:960-983, DLAGRA; ip 0-based here)."""
    n = len(x)
    d = 0.0
    for i in range(n):
        if i == ip:
            continue
        yp = y[i]
        for j in range(n):
            if j == ip or j == i:
                continue
            yp = yp * (x[ip] - x[j])
        for j in range(n):
            if j == i:
                continue
            yp = yp / (x[i] - x[j])
        d = d + yp
    for i in range(n):
        if i == ip:
            continue
        d = d + y[ip] / (x[ip] - x[i])
    return d


def splset(f, x):
    """Knot derivatives of the clamped cubic spline through (x, f).

    SRC/liboctdyn.f:342-393 (SPLSET): the end slopes are the derivatives of the
    4-point Lagrange polynomials at the first and last knot (DLAGRA); the
    interior slopes m solve lambda_i m_{i-1} + 2 m_i + mu_i m_{i+1} =
    3 lambda_i (f_i - f_{i-1})/h_{i-1} + 3 mu_i (f_{i+1} - f_i)/h_i,
    lambda_i = h_i/(h_{i-1} + h_i), mu_i = 1 - lambda_i, by the pivoted
    forward sweep and back substitution of JORDAN (SRC/liboctdyn.f:926-958).
    Returns the slopes (the F(:,2) column of the reference).
    """
    f = np.asarray(f, dtype=float)
    x = np.asarray(x, dtype=float)
    nx = len(x)
    nx2 = nx - 2
    hx = [x[i] - x[i - 1] for i in range(1, nx)]
    rlx = [hx[i + 1] / (hx[i] + hx[i + 1]) for i in range(nx2)]
    rmux = [1.0 - v for v in rlx]
    p0 = _dlagra(x[:4], f[:4], 0)
    pn = _dlagra(x[nx - 4:], f[nx - 4:], 3)
    b = [3.0 * rlx[i] / hx[i] * (f[i + 1] - f[i]) + 3.0 * rmux[i] / hx[i + 1] * (f[i + 2] - f[i + 1])
         for i in range(nx2)]
    b[0] = b[0] - rlx[0] * p0
    b[nx2 - 1] = b[nx2 - 1] - rmux[nx2 - 1] * pn
    piv = [0.0] * nx2
    piv[0] = 2.0
    for i in range(1, nx2):
        piv[i] = 2.0 - rlx[i] * rmux[i - 1] / piv[i - 1]
        b[i] = b[i] - rlx[i] / piv[i - 1] * b[i - 1]
    xi = [0.0] * nx2
    xi[nx2 - 1] = b[nx2 - 1] / piv[nx2 - 1]
    for i in range(nx2 - 2, -1, -1):
        xi[i] = (b[i] - xi[i + 1] * rmux[i]) / piv[i]
    d = np.zeros(nx)
    d[0] = p0
    d[nx - 1] = pn
    d[1:nx - 1] = xi
    return d


def splinqq(f, d, x, r):
    """Spline value at r (array) from knots x, values f and slopes d.

    SRC/liboctdyn.f:298-341 (SPLINQQ): cubic Hermite form on the interval
    [x_{i-1}, x_i) holding r (searched from the second knot); at or beyond the
    last knot the extension C6/r^6 + C8/r^8 matched to the last value and slope.
    """
    f = np.asarray(f, dtype=float)
    d = np.asarray(d, dtype=float)
    x = np.asarray(x, dtype=float)
    r = np.asarray(r, dtype=float)
    nx = len(x)
    out = np.empty_like(r)
    beyond = r >= x[nx - 1]
    inner = ~beyond
    ri = r[inner]
    idol = np.maximum(np.searchsorted(x, ri, side="right"), 1)  # 0-based index of X(IDOL)
    hi = x[idol] - x[idol - 1]
    xr = (ri - x[idol - 1]) / hi
    u1 = xr * xr * (-2.0 * xr + 3.0)
    u3 = hi * xr * xr * (xr - 1.0)
    u2 = 1.0 - u1
    u4 = hi * xr * ((xr - 2.0) * xr + 1.0)
    out[inner] = u1 * f[idol] + u2 * f[idol - 1] + u3 * d[idol] + u4 * d[idol - 1]
    if beyond.any():
        ro = x[nx - 1]
        yo = f[nx - 1]
        yp = d[nx - 1]
        ain = yo + yp * ro / 6.0
        c8 = -ain * 3.0 * ro ** 8
        c6 = yo * ro ** 6 - c8 / ro / ro
        rb = r[beyond]
        out[beyond] = c6 / rb ** 6 + c8 / rb ** 8
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

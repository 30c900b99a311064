"""Bound states of the reactant diatom on the propagation grid: the func/bcwf
content that the reference's preparation program writes
(radial_functions01_write, SRC/mod_baseYfunciones_01y2.f:553-835), for a
diatomic potential given as a function.

The method is the reference's, written here from its mathematics:
  1. On a fine uniform grid of n_fine points spanning the r range of the
     propagation grid, the three-point finite-difference Hamiltonian
     -1/(2m) d^2/dr^2 + V(r) + j(j+1)/(2 m r^2) has tridiagonal eigenvalues
     (the reference diagonalises it with TQLI, SRC/liboctdyn.f:233-296); they
     are the first estimates, lowest first.
  2. Each estimate below the asymptote by more than 0.01 hartree x 2m/2m
     (the reference's test v(rmax) - e > 1e-2 in its units) is refined by
     Cooley's method (J. W. Cooley, Math. Comp. 15 (1961) 363), the SCHR
     routine of the reference (SRC/liboctdyn.f:405-548): Numerov integration
     of u'' = g(r) u, g = 2m (V - E), inward from r_max and outward from
     r_min, matched at the maximum of the inward branch, and the energy
     correction
         dE = [ (2 Y_m - Y_out - Y_in)/h^2 + g_m ] / sum_i P_i^2,
     Y = (1 - h^2 g/12) P, repeated until it is below eps; the returned
     function is normalised to h sum P^2 = 1.
  3. The function is carried to the propagation grid by the clamped cubic
     spline (numerics.splset / splinqq) and normalised to sum fd^2 = 1 there.
Energies in the units of V; lengths as given (the reference works in bohr
and hartree here).
"""
import math

import numpy as np

from .numerics import splinqq, splset

__all__ = ["cooley_numerov", "fd_eigenvalues", "radial_functions"]


def fd_eigenvalues(g_diag, h):
    """Eigenvalues (ascending) of the tridiagonal finite-difference operator
    -u'' + g u on a uniform grid of step h with u = 0 beyond both ends, for
    g_diag = 2 m (V + centrifugal): the reference's TQLI step, divided by 2m
    by the caller."""
    from scipy.linalg import eigvalsh_tridiagonal

    n = len(g_diag)
    d = 2.0 / (h * h) + np.asarray(g_diag, dtype=float)
    e = np.full(n - 1, -1.0 / (h * h))
    return eigvalsh_tridiagonal(d, e)


def cooley_numerov(e0, g, h, rmax, maxit=20, eps=1e-10):
    """Cooley's eigenvalue refinement of u'' = (g - E) u on a uniform grid
    ending at rmax (g and E both in the scaled units 2m x energy), from the
    estimate e0.  Returns (E, P, nodes): the refined eigenvalue, the
    eigenfunction on the grid normalised to h sum P^2 = 1, and its node
    count.  The function is the one integrated at the energy before the
    last correction, as the reference's SCHR returns it."""
    g = np.asarray(g, dtype=float)
    n = len(g)
    h2 = h * h
    w = h2 / 12.0
    e = float(e0)
    de = 0.0
    test = -1.0
    big = 1e34
    p = np.zeros(n)
    for _ in range(maxit):
        p[:] = 0.0
        q = g - e
        # inward: start in the classically forbidden tail (WKB decay)
        if q[n - 2] < 0.0:
            e = g[n - 3]
            q = g - e
        p[n - 1] = 1e-30
        # WKB decay between the last two points, at their absolute positions
        decay = (rmax - h) * math.sqrt(max(q[n - 2], 0.0)) - rmax * math.sqrt(max(q[n - 1], 0.0))
        p[n - 2] = p[n - 1] * math.exp(-min(decay, 50.0))
        y_prev = (1.0 - w * q[n - 1]) * p[n - 1]
        y_cur = (1.0 - w * q[n - 2]) * p[n - 2]
        m = n - 3
        while True:
            y_next = 2.0 * y_cur - y_prev + h2 * q[m + 1] * p[m + 1]
            p[m] = y_next / (1.0 - w * q[m])
            if abs(p[m]) >= big:          # rescale the branch and redo this point
                s = p[m + 1]
                p[m + 1:] /= s
                y_prev /= s
                y_cur /= s
                continue
            if abs(p[m]) <= abs(p[m + 1]) or m <= 1:
                break
            y_prev, y_cur = y_cur, y_next
            m -= 1
        pm = p[m]
        y_in = y_cur / pm
        p[m:] /= pm
        # outward from r_min
        p[0] = 1.0
        y_prev = 0.0
        y_cur = (1.0 - w * q[0]) * p[0]
        i = 1
        while i <= m:
            y_next = 2.0 * y_cur - y_prev + h2 * q[i - 1] * p[i - 1]
            p[i] = y_next / (1.0 - w * q[i])
            if abs(p[i]) >= big:
                s = p[i - 1]
                p[:i] /= s
                y_prev /= s
                y_cur /= s
                continue
            y_prev, y_cur = y_cur, y_next
            i += 1
        pm = p[m]
        d_old = de
        if pm != 0.0:
            y_out = y_prev / pm
            y_m = y_cur / pm
            p[:m + 1] /= pm
            norm2 = float(np.sum(p * p))
            f = (2.0 * y_m - y_out - y_in) / h2 + q[m]
            if abs(f) < 1e37:
                de = f / norm2
            else:
                norm2 = 9.99999e29
                de = abs(1e-4 * e)
        else:
            norm2 = 9.99999e29
            de = abs(1e-4 * e)
        e_old = e
        e = e + de
        test = max(abs(d_old) - abs(de), test)
        if test >= 0.0 and abs(e - e_old) < eps:
            break
    nodes = int(np.sum((p[2:n - 2] * p[1:n - 3] < 0.0)))
    p = p / math.sqrt(h * float(np.sum(p * p)))
    return e, p, nodes


def radial_functions(vdiat, rmin, rmax, mass, r_grid, jvals, nvini, nvmax, n_fine=2000,
                     bound_margin=1e-2):
    """Bound states of the diatom on r_grid for every j in jvals.

    vdiat(r): the diatomic potential (energy units E) on the fine grid
    positions r (length units L); mass: reduced mass in units with
    hbar = 1, so that the kinetic operator is -1/(2 mass) d^2/dr^2.
    Returns {j: (energies [nv], fd [nv][len(r_grid)])} with the states
    v = nvini.. whose estimate lies more than bound_margin below the
    potential at rmax (each state's fd normalised to sum fd^2 = 1 on
    r_grid), as radial_functions01_write keeps them.  As there, once a j
    keeps at most its v = nvini state while a higher v of the range is not
    bound, the j above it get no state."""
    x = rmin + np.arange(n_fine) * ((rmax - rmin) / (n_fine - 1))
    h = (rmax - rmin) / (n_fine - 1)
    v = np.asarray(vdiat(x), dtype=float)
    two_m = 2.0 * mass
    out = {}
    stopped = False
    for j in jvals:
        if stopped:
            out[j] = (np.zeros(0), np.zeros((0, len(r_grid))))
            continue
        vv = v + 0.5 * j * (j + 1) / mass / (x * x)
        est = fd_eigenvalues(two_m * vv, h) / two_m
        gs = two_m * vv
        ens, fds = [], []
        last_failed = False
        for iv in range(nvini, nvmax + 1):
            last_failed = not (v[-1] - est[iv] > bound_margin)
            if last_failed:
                continue
            e, p, _ = cooley_numerov(est[iv] * two_m, gs, h, rmax)
            f = p.copy()
            d = splset(f, x)
            inside = (r_grid >= x[0]) & (r_grid <= x[-1])
            fd = np.zeros(len(r_grid))
            fd[inside] = splinqq(f, d, x, r_grid[inside])
            fd /= math.sqrt(float(np.sum(fd * fd)))
            ens.append(e / two_m)
            fds.append(fd)
        out[j] = (np.array(ens), np.array(fds).reshape(len(ens), len(r_grid)))
        if len(ens) <= 1 and last_failed:
            stopped = True
    return out

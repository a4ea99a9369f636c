"""Occupation optimiser and the orbital-optimiser dispatch (minimization.jl:100-139,
242-630), plus the host bookkeeping the reference runs between them (gamma guesses,
utils.jl:539-616 / guess.jl:184-213).

Every energy, gradient, Lagrangian, rotation, the occupation optimiser's CG iteration and
the re-ordering of the pairs run in libdonofhip on the device; the host-driven CG
(conjugate_gradient over device calcocce/calcoccg calls) is kept as the iteration the
device optimiser is checked against.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from . import api

# ---- gamma guesses and conversions (guess.jl:184-213, utils.jl:539-584) ---------------


def guess_gamma_softmax(ndoc, ncwo):
    """guess.jl:200-213."""
    return np.full(ndoc * ncwo, math.log(max(0.001 / ncwo, 1e-16)) - math.log(0.999))


def guess_gamma_trigonometric(ndoc, ncwo):
    """guess.jl:184-198."""
    g = np.zeros(ndoc * ncwo)
    for i in range(1, ndoc + 1):
        g[i - 1] = math.acos(math.sqrt(2.0 * 0.999 - 1.0))
        for j in range(1, ncwo):
            g[ndoc + (ndoc - i) + ndoc * (j - 1)] = math.asin(math.sqrt(1.0 / (ncwo - j + 1)))
    return g


def _weak(p, l):
    return p.no1 + p.ndns + (p.ndoc - l) + p.ndoc * np.arange(p.ncwo)


def pair_weak(p, l):
    """0-based columns of the weak orbitals of strong orbital l (1-based):
    ll:ndoc:ul with ll = no1 + ndns + (ndoc - l) + 1 (pnof.jl:509-510)."""
    return _weak(p, l)


def n_to_gammas_softmax(n, p):
    """utils.jl:539-561 (reads n[i] as the reference does)."""
    g = np.zeros(p.nv)
    for i in range(1, p.ndoc + 1):
        gi = (p.ndoc - i) + p.ndoc * np.arange(p.ncwo)
        g[gi] = np.log(np.maximum(n[_weak(p, i)], 1e-16)) - math.log(n[i - 1])
    return g


def n_to_gammas_trigonometric(n, p):
    """utils.jl:563-584."""
    g = np.zeros(p.nv)
    for i in range(1, p.ndoc + 1):
        idx = p.no1 + i - 1
        g[i - 1] = math.acos(math.sqrt(2.0 * n[idx] - 1.0))
        pref = max(1 - n[idx], 1e-14)
        w = _weak(p, i)
        for j in range(p.ncwo - 1):
            gj = math.asin(math.sqrt(n[w[j]] / pref))
            g[p.ndoc + (p.ndoc - i) + p.ndoc * j] = gj
            pref = pref * math.cos(gj) ** 2
    return g


def order_occupations_softmax(old_C, old_gamma, p, I):
    """Sort each pair's orbitals by gamma, strongest first, gammas shifted by their
    maximum (utils.jl:586-616), on the device of the System I."""
    return api.order_occupations_softmax(old_C, old_gamma, p, I)


# ---- nonlinear conjugate gradient ------------------------------------------------------


@dataclass
class CGResult:
    x: np.ndarray
    minimum: float
    iterations: int
    g_converged: bool


def _approx_wolfe(f, d, f0, d0, c1, c2, eps=1e-6):
    """Hager-Zhang approximate Wolfe conditions: the slope has risen to c2 d0 (and is
    below (2 c1 - 1) d0) while f stays within eps |f0| of f0.  They keep the search
    going when energy differences fall below round-off."""
    return (2 * c1 - 1) * d0 >= d >= c2 * d0 and f <= f0 + eps * abs(f0)


def _line_search(phi, f0, d0, a1, c1=0.1, c2=0.9, amax=1e6, maxiter=40):
    """Strong-Wolfe line search with bracketing and cubic/bisection zoom (Nocedal &
    Wright, Numerical Optimization, Alg. 3.5/3.6), also accepting the approximate Wolfe
    conditions of Hager & Zhang.  phi(a) -> (f, g, slope)."""
    a_prev, f_prev, d_prev, g_prev = 0.0, f0, d0, None
    a = a1
    for i in range(maxiter):
        f, g, d = phi(a)
        if np.isfinite(f) and _approx_wolfe(f, d, f0, d0, c1, c2):
            return a, f, g
        if not np.isfinite(f) or f > f0 + c1 * a * d0 or (i > 0 and f >= f_prev):
            return _zoom(phi, f0, d0, a_prev, f_prev, d_prev, a, f, d, c1, c2)
        if abs(d) <= -c2 * d0:
            return a, f, g
        if d >= 0:
            return _zoom(phi, f0, d0, a, f, d, a_prev, f_prev, d_prev, c1, c2)
        a_prev, f_prev, d_prev, g_prev = a, f, d, g
        a = min(2.0 * a, amax)
    return a_prev, f_prev, g_prev  # the last point evaluated


def _zoom(phi, f0, d0, alo, flo, dlo, ahi, fhi, dhi, c1, c2, maxiter=40):
    best = None
    for _ in range(maxiter):
        # cubic interpolation between the bracket ends, safeguarded to the interior
        a = None
        if np.isfinite(fhi) and np.isfinite(dhi) and ahi != alo:
            d1 = dlo + dhi - 3 * (flo - fhi) / (alo - ahi)
            rad = d1 * d1 - dlo * dhi
            if rad >= 0:
                d2 = math.copysign(math.sqrt(rad), ahi - alo)
                den = dhi - dlo + 2 * d2
                if den != 0:
                    a = ahi - (ahi - alo) * (dhi + d2 - d1) / den
        lo, hi = min(alo, ahi), max(alo, ahi)
        if a is None or not (lo + 0.1 * (hi - lo) <= a <= hi - 0.1 * (hi - lo)):
            a = 0.5 * (alo + ahi)
        f, g, d = phi(a)
        if best is None or f < best[1]:
            best = (a, f, g)
        if np.isfinite(f) and _approx_wolfe(f, d, f0, d0, c1, c2):
            return a, f, g
        if not np.isfinite(f) or f > f0 + c1 * a * d0 or f >= flo:
            ahi, fhi, dhi = a, f, d
        else:
            if abs(d) <= -c2 * d0:
                return a, f, g
            if d * (ahi - alo) >= 0:
                ahi, fhi, dhi = alo, flo, dlo
            alo, flo, dlo = a, f, d
        if abs(ahi - alo) < 1e-16 * max(1.0, abs(alo)):
            break
    return best


def conjugate_gradient(fun, grad, x0, g_abstol=1e-8, maxiter=1000, eta=0.4, psi0=0.01):
    """Nonlinear conjugate gradient with the Hager-Zhang direction update (Hager & Zhang,
    SIAM J. Optim. 16 (2005) 170-192; eta-truncated beta) and a strong-Wolfe line
    search.  Stops when max|g| <= g_abstol, the criterion of Optim.Options(g_abstol)."""
    x = np.array(x0, dtype=np.float64, copy=True)
    f = fun(x)
    g = grad(x)
    if np.max(np.abs(g)) <= g_abstol:
        return CGResult(x, f, 0, True)
    d = -g
    xn = np.max(np.abs(x))
    a = psi0 * xn / np.max(np.abs(g)) if xn > 0 else psi0 * max(abs(f), 1.0) / np.dot(g, g)
    converged = False
    it = 0
    for it in range(1, maxiter + 1):
        d0 = float(np.dot(g, d))
        if d0 >= 0:
            d = -g
            d0 = float(np.dot(g, d))

        def phi(t, x=x, d=d):
            xt = x + t * d
            ft = fun(xt)
            gt = grad(xt)
            return ft, gt, float(np.dot(gt, d))

        a, fn, gn = _line_search(phi, f, d0, a)
        if not (np.isfinite(fn) and fn <= f + 1e-6 * abs(f)) or a == 0.0:
            break  # no acceptable step: stagnation at the round-off floor
        x = x + a * d
        if np.max(np.abs(gn)) <= g_abstol:
            f, g = fn, gn
            converged = True
            break
        yk = gn - g
        dy = float(np.dot(d, yk))
        if dy == 0 or not np.isfinite(dy):
            beta = 0.0
        else:
            beta = float(np.dot(yk - 2.0 * d * float(np.dot(yk, yk)) / dy, gn)) / dy
            etak = -1.0 / (np.linalg.norm(d) * min(eta, np.linalg.norm(g)))
            beta = max(beta, etak)
        d = -gn + beta * d
        f, g = fn, gn
    return CGResult(x, float(f), it, converged)


# ---- minimization.jl:100-139 -----------------------------------------------------------


def occoptr_host_cg(gamma, p, I, g_abstol=None):
    """The same CG driven from the host (one device call per energy or gradient): the
    reference iteration the device optimiser is checked against.  A trial point whose
    energy or gradient is not finite reads as inf / nan, as the device CG treats it."""
    from ._lib import DN_ERR_NONFINITE, DnError

    def fun(g):
        try:
            return api.calcocce(g, None, None, None, p, I)
        except DnError as e:
            if e.code != DN_ERR_NONFINITE:
                raise
            return math.inf

    def grad(g):
        try:
            return api.calcoccg(g, None, None, None, p, I)
        except DnError as e:
            if e.code != DN_ERR_NONFINITE:
                raise
            return np.full(p.nv, math.nan)

    return conjugate_gradient(fun, grad, gamma,
                              g_abstol=p.threshen if g_abstol is None else g_abstol)


def occoptr(gamma, C, H, I, freeze_occ, p):
    """Occupation optimisation with J/K/H_core frozen at C (minimization.jl:100-139).
    One J/K build on the device, then nonlinear CG over gamma on the device (dn_occoptr:
    iterate, line search and every calcocce/calcoccg evaluation stay in HBM) against the
    resident J/K/H_core."""
    res = None
    if p.ndoc > 0 and not freeze_occ:
        api.computeJKH_MO(C, H, I, p)
        x, fmin, nit, conv = api.occ_conjugate_gradient(gamma, None, None, None, p, I,
                                                         p.threshen)
        res = CGResult(x, fmin, nit, conv)
        gamma = res.x
        if res.iterations < 1:
            p.maxloop = p.maxloop + 10
    n, _ = api.ocupacion(gamma, p, I)
    cj12, ck12 = api.PNOFi_selector(n, p, I)
    if res is not None:
        return res.minimum, res.iterations, res.g_converged, gamma, n, cj12, ck12
    return -1, 0, True, gamma, n, cj12, ck12


ORBOPT = {"ADAM": api.orbopt_adam, "ADABelief": api.orbopt_adabelief,
          "YOGI": api.orbopt_yogi, "Demon": api.orbopt_demon}

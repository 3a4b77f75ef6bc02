"""Exact solution of the Riemann problem for the 1D Euler equations (gamma law).

Textbook construction (Toro, "Riemann Solvers and Numerical Methods for Fluid
Dynamics", ch. 4): the star pressure solves f_L(p) + f_R(p) + (u_R - u_L) = 0
with the shock (Rankine-Hugoniot) or rarefaction (isentrope + Riemann
invariant) branch on each side; the solution is sampled at x/t.  Written
independently of the reference's iteration so that it can pin it.  Pressure is
solved by safeguarded Newton to machine precision.
"""
import math

import numpy as np


def _f(p, rho, pk, c, g):
    """Velocity jump across one wave and its derivative in p."""
    if p > pk:                                     # shock
        A = 2.0 / ((g + 1.0) * rho)
        B = (g - 1.0) / (g + 1.0) * pk
        s = math.sqrt(A / (p + B))
        return (p - pk) * s, s * (1.0 - 0.5 * (p - pk) / (p + B))
    # rarefaction
    r = (p / pk) ** ((g - 1.0) / (2.0 * g))
    return 2.0 * c / (g - 1.0) * (r - 1.0), 1.0 / (rho * c) * (p / pk) ** (-(g + 1.0) / (2.0 * g))


def star_state(rhoL, uL, pL, rhoR, uR, pR, g=1.4, gR=None):
    """Star pressure and velocity; gR (default g) is the right state's gamma."""
    gL, gR = g, (g if gR is None else gR)
    cL, cR = math.sqrt(gL * pL / rhoL), math.sqrt(gR * pR / rhoR)
    du = uR - uL
    if 2.0 * cL / (gL - 1.0) + 2.0 * cR / (gR - 1.0) <= du:
        raise ValueError("vacuum")
    lo, hi = 0.0, max(pL, pR)
    while _f(hi, rhoL, pL, cL, gL)[0] + _f(hi, rhoR, pR, cR, gR)[0] + du < 0:
        hi *= 2.0
    p = 0.5 * (pL + pR)
    for _ in range(200):
        fl, dl = _f(p, rhoL, pL, cL, gL)
        fr, dr = _f(p, rhoR, pR, cR, gR)
        F = fl + fr + du
        if F > 0:
            hi = min(hi, p)
        else:
            lo = max(lo, p)
        pn = p - F / (dl + dr)
        if not (lo < pn < hi):
            pn = 0.5 * (lo + hi)
        if abs(pn - p) <= 1e-16 * p:
            p = pn
            break
        p = pn
    fl = _f(p, rhoL, pL, cL, gL)[0]
    fr = _f(p, rhoR, pR, cR, gR)[0]
    u = 0.5 * (uL + uR) + 0.5 * (fr - fl)
    return p, u


def sample(rhoL, uL, pL, rhoR, uR, pR, xi, g=1.4, gR=None):
    """(rho, u, p) at similarity coordinates xi = x / t (array); gR: the right gamma."""
    gL, gR = g, (g if gR is None else gR)
    ps, us = star_state(rhoL, uL, pL, rhoR, uR, pR, gL, gR)
    cL, cR = math.sqrt(gL * pL / rhoL), math.sqrt(gR * pR / rhoR)
    xi = np.asarray(xi, dtype=np.float64)
    rho = np.empty_like(xi)
    u = np.empty_like(xi)
    p = np.empty_like(xi)
    for k, s in enumerate(xi.flat):
        if s <= us:                                            # left of contact
            g = gL
            gm, gp = g - 1.0, g + 1.0
            if ps > pL:
                SL = uL - cL * math.sqrt(gp / (2 * g) * ps / pL + gm / (2 * g))
                if s <= SL:
                    st = (rhoL, uL, pL)
                else:
                    st = (rhoL * (ps / pL + gm / gp) / (gm / gp * ps / pL + 1.0), us, ps)
            else:
                cs = cL * (ps / pL) ** (gm / (2 * g))
                if s <= uL - cL:
                    st = (rhoL, uL, pL)
                elif s >= us - cs:
                    st = (rhoL * (ps / pL) ** (1.0 / g), us, ps)
                else:
                    c = 2.0 / gp * (cL + gm / 2.0 * (uL - s))
                    uu = 2.0 / gp * (cL + gm / 2.0 * uL + s)
                    st = (rhoL * (c / cL) ** (2.0 / gm), uu, pL * (c / cL) ** (2.0 * g / gm))
        else:
            g = gR
            gm, gp = g - 1.0, g + 1.0
            if ps > pR:
                SR = uR + cR * math.sqrt(gp / (2 * g) * ps / pR + gm / (2 * g))
                if s >= SR:
                    st = (rhoR, uR, pR)
                else:
                    st = (rhoR * (ps / pR + gm / gp) / (gm / gp * ps / pR + 1.0), us, ps)
            else:
                cs = cR * (ps / pR) ** (gm / (2 * g))
                if s >= uR + cR:
                    st = (rhoR, uR, pR)
                elif s <= us + cs:
                    st = (rhoR * (ps / pR) ** (1.0 / g), us, ps)
                else:
                    c = 2.0 / gp * (cR - gm / 2.0 * (uR - s))
                    uu = 2.0 / gp * (-cR + gm / 2.0 * uR + s)
                    st = (rhoR * (c / cR) ** (2.0 / gm), uu, pR * (c / cR) ** (2.0 * g / gm))
        rho.flat[k], u.flat[k], p.flat[k] = st
    return rho, u, p


def sod_l1(rho, u, p, t=0.2, x0=0.5):
    """Cell-average-free L1 errors (midpoint sampling) of a Sod-1000-type state."""
    m = len(rho)
    x = (np.arange(m) + 0.5) / m
    er, eu, ep = sample(1.0, 0.0, 1.0, 0.125, 0.0, 0.1, (x - x0) / t)
    h = 1.0 / m
    return np.array([np.sum(np.abs(rho - er)) * h, np.sum(np.abs(u - eu)) * h, np.sum(np.abs(p - ep)) * h])


def star_pressure_mp(rhoL, uL, pL, rhoR, uR, pR, g=1.4, gR=None, p0=None, dps=40):
    """The star pressure to `dps` digits (mpmath, Toro's f functions; bracketed
    from the double-precision root or p0), and p |f'(p)| there: the factor that
    turns a velocity error into a relative pressure error at the root."""
    import mpmath as mp
    gL, gR = g, (g if gR is None else gR)
    with mp.workdps(dps):
        rL, rR, PL, PR = mp.mpf(rhoL), mp.mpf(rhoR), mp.mpf(pL), mp.mpf(pR)
        cL, cR = mp.sqrt(gL * PL / rL), mp.sqrt(gR * PR / rR)
        du = mp.mpf(uR) - mp.mpf(uL)

        def f(p, rho, pk, c, gg):
            if p > pk:
                A = 2 / ((gg + 1) * rho)
                B = mp.mpf(gg - 1) / (gg + 1) * pk
                return (p - pk) * mp.sqrt(A / (p + B))
            return 2 * c / (gg - 1) * ((p / pk) ** (mp.mpf(gg - 1) / (2 * gg)) - 1)

        F = lambda p: f(p, rL, PL, cL, gL) + f(p, rR, PR, cR, gR) + du
        if p0 is None:
            p0 = star_state(rhoL, uL, pL, rhoR, uR, pR, g, gR)[0]
        lo, hi = mp.mpf(p0) * (1 - mp.mpf("1e-6")), mp.mpf(p0) * (1 + mp.mpf("1e-6"))
        while F(lo) > 0:
            lo /= 2
        while F(hi) < 0:
            hi *= 2
        p = mp.findroot(F, (lo, hi), solver="anderson")
        slope = mp.diff(F, p)
        return float(p), float(p * abs(slope))

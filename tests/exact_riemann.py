"""Exact solution of the 1-D Euler Riemann problem for an ideal gas.

Textbook two-rarefaction / two-shock solver (E. F. Toro, "Riemann Solvers
and Numerical Methods for Fluid Dynamics", ch. 4): Newton iteration on the
star pressure, then sampling of the self-similar solution.  Used only as an
independent check of the discrete scheme's convergence.
"""
from __future__ import annotations

import math

import numpy as np


def _f(p, rho, pk, c, g):
    if p > pk:
        A = 2.0 / ((g + 1) * rho)
        B = (g - 1) / (g + 1) * pk
        q = math.sqrt(A / (p + B))
        return (p - pk) * q, q * (1.0 - 0.5 * (p - pk) / (B + p))
    pr = p / pk
    f = 2.0 * c / (g - 1) * (pr ** ((g - 1) / (2 * g)) - 1.0)
    df = 1.0 / (rho * c) * pr ** (-(g + 1) / (2 * g))
    return f, df


def star_state(wl, wr, g=1.4):
    rl, ul, pl = wl
    rr, ur, pr = wr
    cl, cr = math.sqrt(g * pl / rl), math.sqrt(g * pr / rr)
    p = max(1e-8, 0.5 * (pl + pr))
    for _ in range(100):
        fl, dfl = _f(p, rl, pl, cl, g)
        fr, dfr = _f(p, rr, pr, cr, g)
        dp = (fl + fr + ur - ul) / (dfl + dfr)
        p_new = max(1e-12, p - dp)
        if abs(p_new - p) < 1e-14 * (p_new + p):
            p = p_new
            break
        p = p_new
    fl, _ = _f(p, rl, pl, cl, g)
    fr, _ = _f(p, rr, pr, cr, g)
    u = 0.5 * (ul + ur) + 0.5 * (fr - fl)
    return p, u


def sample(x, t, x0, wl, wr, g=1.4):
    """Primitive (rho, u, p) at positions x and time t > 0."""
    rl, ul, pl = wl
    rr, ur, pr = wr
    cl, cr = math.sqrt(g * pl / rl), math.sqrt(g * pr / rr)
    ps, us = star_state(wl, wr, g)
    out = np.zeros((len(x), 3))
    for i, xi in enumerate(x):
        s = (xi - x0) / t
        if s <= us:  # left of contact
            if ps > pl:  # left shock
                rs = rl * ((ps / pl + (g - 1) / (g + 1)) / ((g - 1) / (g + 1) * ps / pl + 1))
                sh = ul - cl * math.sqrt((g + 1) / (2 * g) * ps / pl + (g - 1) / (2 * g))
                out[i] = (rl, ul, pl) if s <= sh else (rs, us, ps)
            else:  # left rarefaction
                rs = rl * (ps / pl) ** (1 / g)
                cs = cl * (ps / pl) ** ((g - 1) / (2 * g))
                if s <= ul - cl:
                    out[i] = (rl, ul, pl)
                elif s >= us - cs:
                    out[i] = (rs, us, ps)
                else:
                    u = 2 / (g + 1) * (cl + (g - 1) / 2 * ul + s)
                    c = 2 / (g + 1) * (cl + (g - 1) / 2 * (ul - s))
                    r = rl * (c / cl) ** (2 / (g - 1))
                    out[i] = (r, u, pl * (c / cl) ** (2 * g / (g - 1)))
        else:
            if ps > pr:  # right shock
                rs = rr * ((ps / pr + (g - 1) / (g + 1)) / ((g - 1) / (g + 1) * ps / pr + 1))
                sh = ur + cr * math.sqrt((g + 1) / (2 * g) * ps / pr + (g - 1) / (2 * g))
                out[i] = (rr, ur, pr) if s >= sh else (rs, us, ps)
            else:
                rs = rr * (ps / pr) ** (1 / g)
                cs = cr * (ps / pr) ** ((g - 1) / (2 * g))
                if s >= ur + cr:
                    out[i] = (rr, ur, pr)
                elif s <= us + cs:
                    out[i] = (rs, us, ps)
                else:
                    u = 2 / (g + 1) * (-cr + (g - 1) / 2 * ur + s)
                    c = 2 / (g + 1) * (cr - (g - 1) / 2 * (ur - s))
                    r = rr * (c / cr) ** (2 / (g - 1))
                    out[i] = (r, u, pr * (c / cr) ** (2 * g / (g - 1)))
    return out

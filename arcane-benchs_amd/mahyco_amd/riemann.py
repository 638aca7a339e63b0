"""Exact solution of the 1-D Riemann problem for an ideal gas.

Standard two-rarefaction / two-shock pressure iteration (Toro, "Riemann
Solvers and Numerical Methods for Fluid Dynamics", 3rd ed., ch. 4: the
pressure function f(p) = f_L(p) + f_R(p) + u_R - u_L solved by Newton, then
sampling of the self-similar solution).  Used by the tests to pin the whole
method against the Sod shock tube.
"""
import numpy as np


def _f_and_df(p, rho, pk, ck, g):
    if p > pk:  # shock
        A = 2.0 / ((g + 1.0) * rho)
        B = (g - 1.0) / (g + 1.0) * pk
        s = np.sqrt(A / (p + B))
        f = (p - pk) * s
        df = s * (1.0 - 0.5 * (p - pk) / (B + p))
    else:  # rarefaction
        pr = p / pk
        f = 2.0 * ck / (g - 1.0) * (pr ** ((g - 1.0) / (2.0 * g)) - 1.0)
        df = 1.0 / (rho * ck) * pr ** (-(g + 1.0) / (2.0 * g))
    return f, df


def star_state(rl, ul, pl, rr, ur, pr, g):
    cl, cr = np.sqrt(g * pl / rl), np.sqrt(g * pr / rr)
    p = max(1e-12, 0.5 * (pl + pr) - 0.125 * (ur - ul) * (rl + rr) * (cl + cr))
    for _ in range(100):
        fl, dfl = _f_and_df(p, rl, pl, cl, g)
        fr, dfr = _f_and_df(p, rr, pr, cr, g)
        dp = (fl + fr + ur - ul) / (dfl + dfr)
        pn = max(1e-12, p - dp)
        if abs(pn - p) <= 1e-14 * (pn + p):
            p = pn
            break
        p = pn
    fl, _ = _f_and_df(p, rl, pl, cl, g)
    fr, _ = _f_and_df(p, rr, pr, cr, g)
    u = 0.5 * (ul + ur) + 0.5 * (fr - fl)
    return p, u


def sample(x, t, x0, left, right, g):
    """Density, velocity, pressure at positions x and time t.  left/right = (rho, u, p)."""
    rl, ul, pl = left
    rr, ur, pr = right
    ps, us = star_state(rl, ul, pl, rr, ur, pr, g)
    cl, cr = np.sqrt(g * pl / rl), np.sqrt(g * pr / rr)
    xi = (np.asarray(x, dtype=float) - x0) / t
    rho = np.empty_like(xi)
    u = np.empty_like(xi)
    p = np.empty_like(xi)
    gm = (g - 1.0) / (g + 1.0)
    for i, s in enumerate(xi):
        if s <= us:  # left of contact
            if ps > pl:  # left shock
                rs = rl * ((ps / pl + gm) / (gm * ps / pl + 1.0))
                sl = ul - cl * np.sqrt((g + 1.0) / (2 * g) * ps / pl + (g - 1.0) / (2 * g))
                rho[i], u[i], p[i] = (rl, ul, pl) if s <= sl else (rs, us, ps)
            else:  # left rarefaction
                rs = rl * (ps / pl) ** (1.0 / g)
                cs = cl * (ps / pl) ** ((g - 1.0) / (2 * g))
                shl, stl = ul - cl, us - cs
                if s <= shl:
                    rho[i], u[i], p[i] = rl, ul, pl
                elif s >= stl:
                    rho[i], u[i], p[i] = rs, us, ps
                else:
                    uu = 2.0 / (g + 1.0) * (cl + (g - 1.0) / 2.0 * ul + s)
                    cc = 2.0 / (g + 1.0) * (cl + (g - 1.0) / 2.0 * (ul - s))
                    rho[i] = rl * (cc / cl) ** (2.0 / (g - 1.0))
                    u[i] = uu
                    p[i] = pl * (cc / cl) ** (2.0 * g / (g - 1.0))
        else:  # right of contact
            if ps > pr:  # right shock
                rs = rr * ((ps / pr + gm) / (gm * ps / pr + 1.0))
                sr = ur + cr * np.sqrt((g + 1.0) / (2 * g) * ps / pr + (g - 1.0) / (2 * g))
                rho[i], u[i], p[i] = (rr, ur, pr) if s >= sr else (rs, us, ps)
            else:  # right rarefaction
                rs = rr * (ps / pr) ** (1.0 / g)
                cs = cr * (ps / pr) ** ((g - 1.0) / (2 * g))
                shr, str_ = ur + cr, us + cs
                if s >= shr:
                    rho[i], u[i], p[i] = rr, ur, pr
                elif s <= str_:
                    rho[i], u[i], p[i] = rs, us, ps
                else:
                    uu = 2.0 / (g + 1.0) * (-cr + (g - 1.0) / 2.0 * ur + s)
                    cc = 2.0 / (g + 1.0) * (cr - (g - 1.0) / 2.0 * (ur - s))
                    rho[i] = rr * (cc / cr) ** (2.0 / (g - 1.0))
                    u[i] = uu
                    p[i] = pr * (cc / cr) ** (2.0 * g / (g - 1.0))
    return rho, u, p

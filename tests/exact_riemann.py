"""Exact solution of the 1D Riemann problem for an ideal gas (Toro,
"Riemann Solvers and Numerical Methods for Fluid Dynamics", ch. 4: pressure
function f_K, Newton iteration for p*, sampling of the self-similar
solution).  Used to check the oracle's Sod shock tube against mathematics."""
import numpy as np


def _fk(p, dk, pk, ck, g):
    if p > pk:  # shock
        a = 2.0 / ((g + 1.0) * dk)
        b = (g - 1.0) / (g + 1.0) * pk
        q = np.sqrt(a / (p + b))
        return (p - pk) * q, q * (1.0 - 0.5 * (p - pk) / (b + p))
    # rarefaction
    r = (p / pk) ** ((g - 1.0) / (2.0 * g))
    return 2.0 * ck / (g - 1.0) * (r - 1.0), 1.0 / (dk * ck) * (p / pk) ** (-(g + 1.0) / (2.0 * g))


def star_state(dl, ul, pl, dr, ur, pr, g):
    cl, cr = np.sqrt(g * pl / dl), np.sqrt(g * pr / dr)
    p = max(1e-12, 0.5 * (pl + pr) - 0.125 * (ur - ul) * (dl + dr) * (cl + cr))
    for _ in range(200):
        fl, dfl = _fk(p, dl, pl, cl, g)
        fr, dfr = _fk(p, dr, pr, cr, g)
        dp = (fl + fr + ur - ul) / (dfl + dfr)
        pn = max(1e-12, p - dp)
        if abs(pn - p) < 1e-14 * (pn + p):
            p = pn
            break
        p = pn
    fl, _ = _fk(p, dl, pl, cl, g)
    fr, _ = _fk(p, dr, pr, cr, g)
    u = 0.5 * (ul + ur) + 0.5 * (fr - fl)
    return p, u


def sample(x, t, x0, left, right, g):
    """Density, velocity, pressure at positions x (array) and time t > 0;
    left/right = (density, velocity, pressure)."""
    dl, ul, pl = left
    dr, ur, pr = right
    ps, us = star_state(dl, ul, pl, dr, ur, pr, g)
    cl, cr = np.sqrt(g * pl / dl), np.sqrt(g * pr / dr)
    s = (np.asarray(x, dtype=float) - x0) / t
    d = np.empty_like(s)
    u = np.empty_like(s)
    p = np.empty_like(s)
    gm = (g - 1.0) / (g + 1.0)
    for i, si in enumerate(s):
        if si <= us:  # left of contact
            if ps > pl:  # left shock
                sl = ul - cl * np.sqrt((g + 1) / (2 * g) * ps / pl + (g - 1) / (2 * g))
                if si <= sl:
                    d[i], u[i], p[i] = dl, ul, pl
                else:
                    d[i] = dl * (ps / pl + gm) / (gm * ps / pl + 1)
                    u[i], p[i] = us, ps
            else:  # left rarefaction
                csl = cl * (ps / pl) ** ((g - 1) / (2 * g))
                if si <= ul - cl:
                    d[i], u[i], p[i] = dl, ul, pl
                elif si >= us - csl:
                    d[i], u[i], p[i] = dl * (ps / pl) ** (1 / g), us, ps
                else:
                    c = 2 / (g + 1) * (cl + (g - 1) / 2 * (ul - si))
                    u[i] = 2 / (g + 1) * (cl + (g - 1) / 2 * ul + si)
                    d[i] = dl * (c / cl) ** (2 / (g - 1))
                    p[i] = pl * (c / cl) ** (2 * g / (g - 1))
        else:  # right of contact
            if ps > pr:  # right shock
                sr = ur + cr * np.sqrt((g + 1) / (2 * g) * ps / pr + (g - 1) / (2 * g))
                if si >= sr:
                    d[i], u[i], p[i] = dr, ur, pr
                else:
                    d[i] = dr * (ps / pr + gm) / (gm * ps / pr + 1)
                    u[i], p[i] = us, ps
            else:  # right rarefaction
                csr = cr * (ps / pr) ** ((g - 1) / (2 * g))
                if si >= ur + cr:
                    d[i], u[i], p[i] = dr, ur, pr
                elif si <= us + csr:
                    d[i], u[i], p[i] = dr * (ps / pr) ** (1 / g), us, ps
                else:
                    c = 2 / (g + 1) * (cr - (g - 1) / 2 * (ur - si))
                    u[i] = 2 / (g + 1) * (-cr + (g - 1) / 2 * ur + si)
                    d[i] = dr * (c / cr) ** (2 / (g - 1))
                    p[i] = pr * (c / cr) ** (2 * g / (g - 1))
    return d, u, p

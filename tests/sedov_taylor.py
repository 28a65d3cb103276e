"""The Sedov-Taylor self-similar blast wave in closed form (test
infrastructure): Sedov's solution for a point explosion in a uniform gas
(L. I. Sedov, Similarity and Dimensional Methods in Mechanics, 1959), in the
parametrisation of J. R. Kamm and F. X. Timmes, "On efficient generation of
numerically robust Sedov solutions" (LA-UR-07-2849, 2007), standard case,
uniform ambient density (omega = 0).

j = 1, 2, 3: planar, cylindrical, spherical.  The similarity variable V runs
from V0 (the centre) to V2 (the shock); along it, `profiles` gives
r / R_shock and the ratios of velocity, density and pressure to their
values just behind the shock."""
import numpy as np


def _constants(g, j, w=0.0):
    a = (j + 2 - w) * (g + 1) / 4.0
    b = (g + 1) / (g - 1)
    c = (j + 2 - w) * g / 2.0
    d = (j + 2 - w) * (g + 1) / ((j + 2 - w) * (g + 1) - 2.0 * (2 + j * (g - 1)))
    e = (2 + j * (g - 1)) / 2.0
    a0 = 2.0 / (j + 2 - w)
    a2 = -(g - 1) / (2 * (g - 1) + j - g * w)
    a1 = (j + 2 - w) * g / (2 + j * (g - 1)) * (2 * (j * (2 - g) - w) / (g * (j + 2 - w) ** 2) - a2)
    a3 = (j - w) / (2 * (g - 1) + j - g * w)
    a4 = a1 * (j + 2 - w) / (2 - g)
    a5 = (w * (1 + g) - 2 * j) / (j * (2 - g) - w)
    return a, b, c, d, e, a0, a1, a2, a3, a4, a5


def v_range(g, j, w=0.0):
    """(V0, V2): the centre and the shock."""
    return 2.0 / ((j + 2 - w) * g), 4.0 / ((j + 2 - w) * (g + 1))


def profiles(V, g, j, w=0.0):
    """(lam, f, gd, h) at similarity variable(s) V: r / R_shock and the
    velocity, density and pressure over their post-shock values."""
    a, b, c, d, e, a0, a1, a2, a3, a4, a5 = _constants(g, j, w)
    V = np.asarray(V, dtype=float)
    x1 = a * V
    x2 = b * (c * V - 1)
    x3 = d * (1 - e * V)
    x4 = b * (1 - c * V / g)
    lam = x1 ** (-a0) * x2 ** (-a2) * x3 ** (-a1)
    f = x1 * lam
    gd = x1 ** (a0 * w) * x2 ** (a3 + a2 * w) * x3 ** (a4 + a1 * w) * x4 ** a5
    h = x1 ** (a0 * j) * x3 ** (a4 + a1 * (w - 2)) * x4 ** (1 + a5)
    return lam, f, gd, h


def tables(g, j, n=4000, offset=None):
    """Profiles at n values of V spaced evenly from V0 + offset (default
    1e-9 of the range) to the shock, by increasing r."""
    V0, V2 = v_range(g, j)
    V = np.linspace(V0 + (1e-9 * (V2 - V0) if offset is None else offset), V2, n)
    return profiles(V, g, j)


def energy_constant(g, j, n=200001):
    """alpha of E = alpha rho0 R^(j+2) / t^2 (per unit length or area for
    j < 3), from the energy inside the shock with the post-shock values of
    a strong shock of speed D = 2 R / ((j + 2) t).  The trapezoid rule runs
    over r on a grid of V clustered at both ends; inside its first point
    (r < 0.02 R) the pressure is the centre's and the kinetic energy nil."""
    V0, V2 = v_range(g, j)
    t = np.linspace(0.0, 1.0, n)
    V = V0 + (V2 - V0) * (1e-12 + (1 - 1e-12) * (0.5 - 0.5 * np.cos(np.pi * t)))
    lam, f, gd, h = profiles(V, g, j)
    u2 = 2.0 / (g + 1)         # post-shock values for rho0 = D = R = 1
    rho2 = (g + 1) / (g - 1)
    p2 = 2.0 / (g + 1)
    dens = 0.5 * rho2 * gd * (u2 * f) ** 2 + p2 * h / (g - 1)
    geo = {1: 2.0, 2: 2.0 * np.pi, 3: 4.0 * np.pi}[j]
    integrand = geo * lam ** (j - 1) * dens
    E = np.sum(0.5 * (integrand[1:] + integrand[:-1]) * np.diff(lam))
    E += geo * p2 * h[0] / (g - 1) * lam[0] ** j / j
    return E * (2.0 / (j + 2)) ** 2

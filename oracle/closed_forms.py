"""Closed-form solutions used to pin the oracle (our own code).

* stoker_wet_dambreak - SWASHES wet dam break without friction, as the case
  tests/runtime/1D_dambreak/dambreak.py:12-70 evaluates it: c_m solves
  -8 c_r^2 c_m^2 (c_l - c_m)^2 + (c_m^2 - c_r^2)^2 (c_m^2 + c_r^2) = 0.
* berthon_steady - steady hydro of the Berthon constant-flux Exner case
  (tests/runtime/2D_const_flux/2D_build_num_sol.py:8-33,
  analyze_convergence.py:9-25): u = ((alpha x + beta)/A)^(1/3), h = q/u,
  z_b = -(u^3 + 2 g q)/(2 g u) + gamma - alpha t.
* mpm_qb - Meyer-Peter-Mueller bedload magnitude
  q_b = 8 (theta - theta_c)^1.5 sqrt((s-1) g d^3) / (1 - p).
"""
import numpy as np
from scipy.optimize import brentq

from .constants import g


def stoker_cm(hl, hr):
    cl, cr = np.sqrt(g * hl), np.sqrt(g * hr)
    f = lambda cm: -8 * cr ** 2 * cm ** 2 * (cl - cm) ** 2 + (cm ** 2 - cr ** 2) ** 2 * (cm ** 2 + cr ** 2)  # noqa
    return brentq(f, cr * (1 + 1e-12), cl)


def stoker_wet_dambreak(x, t, hl, hr, x0):
    cm = stoker_cm(hl, hr)
    cl = np.sqrt(g * hl)
    xa = x0 - t * cl
    xb = x0 + t * (2 * cl - 3 * cm)
    xc = x0 + t * (2 * cm ** 2 * (cl - cm)) / (cm ** 2 - g * hr)
    h = np.full_like(x, hr, dtype=np.float64)
    u = np.zeros_like(x, dtype=np.float64)
    h = np.where(x <= xa, hl, h)
    fan = (x > xa) & (x <= xb)
    h = np.where(fan, (4 / (9 * g)) * (cl - (x - x0) / (2 * t)) ** 2, h)
    u = np.where(fan, (2 / 3) * ((x - x0) / t + cl), u)
    mid = (x > xb) & (x <= xc)
    h = np.where(mid, cm ** 2 / g, h)
    u = np.where(mid, 2 * (cl - cm), u)
    return h, u


def berthon_steady(x, t, q=1.0, A=0.005, alpha=0.005, beta=0.005, gamma=1.0):
    u = ((alpha * x + beta) / A) ** (1.0 / 3.0)
    h = q / u
    zb = -(u ** 3 + 2 * g * q) / (2 * g * u) + gamma - alpha * t
    return h, u, zb


def mpm_qb(h, speed, d50, density, porosity, theta_c=0.047):
    s = density / 1000.0
    n = d50 ** (1 / 6) / 21.1
    theta = n ** 2 * speed ** 2 / ((s - 1) * d50 * np.cbrt(h))
    dth = np.maximum(theta - theta_c, 0.0)
    return 8 * dth ** 1.5 * np.sqrt((s - 1) * g * d50 ** 3) / (1 - porosity)

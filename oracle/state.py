"""Plain structure-of-arrays states for the oracle.

Field sets follow state/roe_exner_state.py:8-48 (h, hu, hv, z, z_b, n, n_b, G,
seds) and state/roe_state.py:8-63 (h, hu, hv, z, n).  Arrays are (Ny, Nx),
row-major, x along axis 1, +y towards decreasing row index.
"""
from dataclasses import dataclass, replace, fields

import numpy as np


@dataclass
class ExnerState:
    h: np.ndarray
    hu: np.ndarray
    hv: np.ndarray
    z: np.ndarray
    z_b: np.ndarray
    n: np.ndarray
    n_b: np.ndarray
    G: np.ndarray
    seds: np.ndarray  # (7, nsed): frac, d50, density, k_d, k_e, theta_c, porosity

    def replace(self, **kw):
        return replace(self, **kw)

    def copy(self):
        return ExnerState(**{f.name: np.array(getattr(self, f.name), copy=True) for f in fields(self)})

    def astype(self, dtype):
        return ExnerState(**{f.name: np.asarray(getattr(self, f.name), dtype=dtype) for f in fields(self)})


@dataclass
class RoeState:
    h: np.ndarray
    hu: np.ndarray
    hv: np.ndarray
    z: np.ndarray
    n: np.ndarray

    def replace(self, **kw):
        return replace(self, **kw)

    def copy(self):
        return RoeState(**{f.name: np.array(getattr(self, f.name), copy=True) for f in fields(self)})

    def astype(self, dtype):
        return RoeState(**{f.name: np.asarray(getattr(self, f.name), dtype=dtype) for f in fields(self)})


EXNER_FIELDS = ("h", "hu", "hv", "z", "z_b", "n", "n_b", "G")
ROE_FIELDS = ("h", "hu", "hv", "z", "n")

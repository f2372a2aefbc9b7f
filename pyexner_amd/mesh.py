"""Uniform structured mesh metadata (domain/mesh.py:10-24)."""
from dataclasses import dataclass
from typing import Tuple

import numpy as np


@dataclass
class Mesh2D:
    global_Ny: int
    global_Nx: int
    local_Ny: int
    local_Nx: int
    x_offset: int
    y_offset: int
    local_X: np.ndarray
    local_Y: np.ndarray
    dh: float
    local_shape: Tuple
    global_shape: Tuple


def mesh_from_coords(x, y, dh=None):
    """Single-rank mesh from file coordinates (io/pnetcdf_reader.py:56-109):
    cell centres are (x + dh/2, y - dh/2); ``dh`` defaults to round(x[1]-x[0], 5)."""
    x = np.asarray(x, dtype=np.float64)
    y = np.asarray(y, dtype=np.float64)
    if dh is None:
        dh = round(float(x[1] - x[0]), 5)
    X, Y = np.meshgrid(x + dh / 2, y - dh / 2, indexing="xy")
    ny, nx = len(y), len(x)
    return Mesh2D(global_Ny=ny, global_Nx=nx, local_Ny=ny, local_Nx=nx, x_offset=0, y_offset=0,
                  local_X=X, local_Y=Y, dh=dh, local_shape=(ny, nx), global_shape=(ny, nx))


def uniform_mesh(ny, nx, dh, x0=0.0, y0=0.0):
    """Mesh whose file coordinates start at x0 and descend in y from y0 + (ny-1)*dh."""
    x = x0 + dh * np.arange(nx)
    y = y0 + dh * np.arange(ny - 1, -1, -1)
    return mesh_from_coords(x, y, dh)

"""Host-side state construction from YAML parameters and input arrays.

Mirrors the parts of io/pnetcdf_reader.py:111-220 that build the state:
the sediment table ``seds`` (rows frac, d50, density, deposition_flux,
erosion_flux, theta_c = 0.047, bulk_porosity; pnetcdf_reader.py:136-158), the
initial G (zeros in MPM mode, grass_factor everywhere otherwise;
pnetcdf_reader.py:126-131) and the cell-centre mesh.  NetCDF-3 input files
(what the reference cases write, e.g. tests/runtime/1D_dambreak/
build_dambreak_input.py:42) are read with scipy.io.netcdf_file.
"""
import numpy as np

from .mesh import mesh_from_coords

THETA_C_MPM = 0.047


def seds_from_params(params):
    """(7, nsed) sediment table, or zeros((1, 1)) for a constant grass factor."""
    ero = params.get("erosion", {}) or {}
    grass = ero.get("grass_factor", 0.0)
    if not isinstance(grass, str):
        return np.zeros((1, 1))
    rows = []
    for key, sed in (ero.get("sediments") or {}).items():
        rows.append([sed["fraction"], sed["diameter"], sed["density"], sed["deposition_flux"],
                     sed["erosion_flux"], THETA_C_MPM, ero["bulk_porosity"]])
    if not rows:
        raise KeyError("grass_factor 'MPM' needs at least one entry under erosion.sediments")
    return np.asarray(rows, dtype=np.float64).T


def initial_G(params, shape):
    ero = params.get("erosion", {}) or {}
    grass = ero.get("grass_factor", 0.0)
    if isinstance(grass, str):
        return np.zeros(shape)
    return float(grass) * np.ones(shape)


def read_netcdf(path, names=("h", "hu", "hv", "z", "z_b", "n")):
    """Read (y, x) fields and coordinates from a NetCDF-3 file (classic or 64-bit offset)."""
    from scipy.io import netcdf_file
    with netcdf_file(path, "r", mmap=False) as f:
        x = np.array(f.variables["x"][:], dtype=np.float64)
        y = np.array(f.variables["y"][:], dtype=np.float64)
        arrays = {}
        for nm in names:
            if nm in f.variables:
                arrays[nm] = np.array(f.variables[nm][:], dtype=np.float64)
    return x, y, arrays


def mesh_and_arrays_from_netcdf(path, params):
    x, y, arrays = read_netcdf(path)
    mesh = mesh_from_coords(x, y)
    shape = mesh.local_shape
    for nm in ("h", "hu", "hv", "z", "n"):
        if nm not in arrays:
            raise KeyError(f"Variable '{nm}' not found in NetCDF file.")
    if params.get("flux_scheme") == "Roe Exner":
        if "z_b" not in arrays:
            raise KeyError("Variable 'z_b' not found in NetCDF file.")
        arrays["G"] = initial_G(params, shape)
    return mesh, arrays

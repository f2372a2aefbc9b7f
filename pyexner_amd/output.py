"""Snapshot output: the write side of io/pnetcdf_reader.py (PnetCDFStateIO).

`NetCDFWriter(path, mesh)` creates a NetCDF-3 file (64-bit offset, what
scipy.io.netcdf_file writes) with dimensions (t unlimited, y, x), the file
coordinates x and y (float64; pnetcdf_reader.py:180-212) and one float32
(t, y, x) variable per state field (seds excluded).  `write_state(state, mesh,
mask)` appends one snapshot with NaN wherever `mask` (the nodata mask) is set
(pnetcdf_reader.py:222-256).  The integrators call `write_state` at every
output window, as the reference's do.

The file holds a single domain (one GPU; the reference's parallel offsets are
always zero here).  Snapshots are kept in host memory until `close()` (or the
end of a `with` block), which writes the file.
"""
from dataclasses import fields

import numpy as np


class NetCDFWriter:
    def __init__(self, path, mesh, x=None, y=None):
        """x, y: the file coordinates; by default recovered from the mesh's
        cell centres (centre = (x + dh/2, y - dh/2), pnetcdf_reader.py:90-91)."""
        self.path = path
        self.x = np.asarray(mesh.local_X[0, :] - mesh.dh / 2 if x is None else x, dtype=np.float64)
        self.y = np.asarray(mesh.local_Y[:, 0] + mesh.dh / 2 if y is None else y, dtype=np.float64)
        self.shape = (len(self.y), len(self.x))
        self.snapshots = []
        self.names = None

    def write_state(self, state, mesh, mask):
        """Append one snapshot of every field of state (torch tensors or
        numpy arrays), float32, NaN where mask."""
        host = state.to_host() if hasattr(state, "to_host") else state
        m = mask.detach().cpu().numpy() if hasattr(mask, "detach") else np.asarray(mask)
        m = m.astype(bool)
        snap = {}
        for f in fields(host):
            if f.name in ("seds", "z_b_lo"):
                continue
            arr = getattr(host, f.name)
            if arr is None:
                continue
            arr = np.asarray(arr)
            if arr.shape != self.shape:
                raise ValueError(f"field {f.name} has shape {arr.shape}, the file {self.shape}")
            snap[f.name] = np.where(m, np.nan, arr).astype(np.float32)
        if self.names is None:
            self.names = list(snap)
        elif list(snap) != self.names:
            raise ValueError("every snapshot must hold the same fields")
        self.snapshots.append(snap)

    def close(self):
        from scipy.io import netcdf_file
        with netcdf_file(self.path, "w", version=2) as f:
            f.createDimension("t", None)
            f.createDimension("y", self.shape[0])
            f.createDimension("x", self.shape[1])
            vx = f.createVariable("x", "d", ("x",))
            vy = f.createVariable("y", "d", ("y",))
            vx[:] = self.x
            vy[:] = self.y
            for name in self.names or []:
                v = f.createVariable(name, "f", ("t", "y", "x"))
                for k, snap in enumerate(self.snapshots):
                    v[k] = snap[name]

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
        return False


def read_snapshots(path):
    """(x, y, {name: (t, y, x) float32 array}) of a file NetCDFWriter wrote."""
    from scipy.io import netcdf_file
    with netcdf_file(path, "r", mmap=False) as f:
        # NetCDF stores big-endian; return native arrays
        x = np.array(f.variables["x"][:], dtype=np.float64)
        y = np.array(f.variables["y"][:], dtype=np.float64)
        out = {k: np.array(v[:], dtype=np.float32) for k, v in f.variables.items() if k not in ("x", "y")}
    return x, y, out

"""Ghost-cell boundary handlers, restated from domain/boundary_registry.py and
domain/boundaries/*.py.

Handlers are applied in YAML order (boundary_registry.py:183-187); each one
gathers from the state as it was before that handler and scatters to its
boundary cells (jax ``.at[].set``).

Periodic has no counterpart in the reference (README.md:163 advertises it,
domain/boundaries/__init__.py registers none).  This oracle defines it: a
boundary cell takes every field from the cell ``N-2`` cells away against the
outward normal, i.e. the last owned column/row on the opposite side.
"""
import numpy as np


def point_in_polygon(points, polygon):
    """boundary_registry.py:19-44, vectorised over points (N, 2)."""
    x = points[:, 0][:, None]
    y = points[:, 1][:, None]
    xi, yi = polygon[:, 0][None, :], polygon[:, 1][None, :]
    xj, yj = np.roll(polygon[:, 0], 1)[None, :], np.roll(polygon[:, 1], 1)[None, :]
    crosses_y = (yi > y) != (yj > y)
    denom = yj - yi
    safe = np.where(crosses_y, denom, 1.0)
    test_x = xi + (xj - xi) * (y - yi) / safe
    intersect = crosses_y & (x <= test_x + 1e-8)
    return np.sum(intersect, axis=1) % 2 == 1


def compute_label_mask(X, Y, polygon):
    pts = np.stack([X.ravel(), Y.ravel()], axis=-1)
    return point_in_polygon(pts, np.asarray(polygon, dtype=np.float64)).reshape(X.shape)


def compute_reflective_indices(mask, normal):
    """boundary_registry.py:61-82."""
    dx = -int(np.round(normal[0]))
    dy = int(np.round(normal[1]))
    b = np.argwhere(mask)
    Ny, Nx = mask.shape
    i = b + np.array([dy, dx])
    i[:, 0] = np.clip(i[:, 0], 0, Ny - 1)
    i[:, 1] = np.clip(i[:, 1], 0, Nx - 1)
    return (b[:, 0], b[:, 1]), (i[:, 0], i[:, 1])


def compute_periodic_indices(mask, normal):
    """Periodic source: N-2 cells against the outward normal (new, no reference)."""
    Ny, Nx = mask.shape
    dx = -int(np.round(normal[0])) * (Nx - 2)
    dy = int(np.round(normal[1])) * (Ny - 2)
    b = np.argwhere(mask)
    i = b + np.array([dy, dx])
    i[:, 0] = np.clip(i[:, 0], 0, Ny - 1)
    i[:, 1] = np.clip(i[:, 1], 0, Nx - 1)
    return (b[:, 0], b[:, 1]), (i[:, 0], i[:, 1])


class Handler:
    def __init__(self, kind, scheme, mask, normal):
        self.kind = kind
        self.scheme = scheme
        self.mask = mask
        self.normal = np.asarray(normal, dtype=np.float64)
        if kind == "Periodic":
            self.b, self.i = compute_periodic_indices(mask, normal)
        else:
            self.b, self.i = compute_reflective_indices(mask, normal)

    def apply(self, s):
        by, bx = self.b
        iy, ix = self.i
        dt = s.h.dtype
        nx, ny = (np.asarray(c, dtype=dt) for c in self.normal)
        kw = {}

        def copy(name):
            a = getattr(s, name).copy()
            a[by, bx] = getattr(s, name)[iy, ix]
            kw[name] = a

        if self.scheme == "Roe Exner":
            if self.kind == "Reflective":  # roeexner_reflective.py:31-63
                hu_int = s.hu[iy, ix]
                hv_int = s.hv[iy, ix]
                mom_n = hu_int * nx + hv_int * ny
                mom_t = -hu_int * ny + hv_int * nx
                mom_n_reflected = -mom_n
                hu_ref = mom_n_reflected * nx - mom_t * ny
                hv_ref = mom_n_reflected * ny + mom_t * nx
                for name in ("h", "z_b", "n", "G"):
                    copy(name)
                hu = s.hu.copy(); hu[by, bx] = hu_ref
                hv = s.hv.copy(); hv[by, bx] = hv_ref
                kw["hu"], kw["hv"] = hu, hv
            elif self.kind == "Transmissive":  # roeexner_transmissive.py:17-31
                for name in ("h", "hu", "hv"):
                    copy(name)
            elif self.kind == "TransmissiveBed":  # roeexner_transmissivebed.py:17-27
                copy("z_b")
            elif self.kind == "Periodic":
                for name in ("h", "hu", "hv", "z", "z_b", "n", "n_b", "G"):
                    copy(name)
            else:
                raise KeyError(f"Roe Exner {self.kind}")
        elif self.scheme == "Roe":
            if self.kind == "Reflective":  # roe_reflective.py:18-44
                for name in ("h", "z", "n"):
                    copy(name)
                hu = s.hu.copy(); hu[by, bx] = s.hu[iy, ix] * (1 - 2 * nx * nx)
                hv = s.hv.copy(); hv[by, bx] = s.hv[iy, ix] * (1 - 2 * ny * ny)
                kw["hu"], kw["hv"] = hu, hv
            elif self.kind in ("Transmissive", "Periodic"):  # roe_transmissive.py:17-43
                for name in ("h", "z", "n", "hu", "hv"):
                    copy(name)
            else:
                raise KeyError(f"Roe {self.kind}")
        return s.replace(**kw)


class Boundaries:
    """BoundaryManager (boundary_registry.py:84-187) for one rank."""

    def __init__(self, scheme, specs, X, Y):
        self.handlers = []
        masks = []
        for name, spec in specs.items():
            m = compute_label_mask(X, Y, spec["polygon"])
            masks.append(m)
            self.handlers.append(Handler(spec["type"], scheme, m, spec["normal"]))
        self.boundary_mask = np.any(np.stack(masks), axis=0) if masks else np.zeros(X.shape, bool)

    def apply(self, s):
        for hd in self.handlers:
            s = hd.apply(s)
        return s

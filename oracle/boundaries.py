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


G_ACC = 9.81


def _berthon(values):
    """u, h = q/u, zb0 and alpha of the Berthon et al. steady state from the
    float32 values [q, A, alpha, beta, gamma, x] (roeexner_berthon.py:35-49,
    roeexner_normalflowdepth.py:24-40); float32 arithmetic as the reference's
    float32 value array gives."""
    q, A, alpha, beta, gamma, x = values[:6]
    u = np.power((alpha * x + beta) / A, np.float32(1.0 / 3.0))
    h = q / u
    u3 = u * (u * u)  # lax.integer_pow(u, 3)
    zb0 = -(u3 + np.float32(2.0 * G_ACC) * q) / (np.float32(2.0 * G_ACC) * np.maximum(u, np.float32(1e-12))) + gamma
    return u, h, zb0, alpha


class Handler:
    def __init__(self, kind, scheme, mask, normal, values=()):
        self.kind = kind
        self.scheme = scheme
        self.mask = mask
        self.normal = np.asarray(normal, dtype=np.float64)
        # boundary_registry.py:103-107: the YAML values as a float32 array
        self.values = np.array([np.nan if str(v).lower() == "nan" else float(v) for v in values], dtype=np.float32)
        if kind == "Periodic":
            self.b, self.i = compute_periodic_indices(mask, normal)
        else:
            self.b, self.i = compute_reflective_indices(mask, normal)

    def apply(self, s, time=0.0):
        by, bx = self.b
        iy, ix = self.i
        dt = s.h.dtype
        nx, ny = (np.asarray(c, dtype=dt) for c in self.normal)
        kw = {}
        tt = dt.type(time)

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
            elif self.kind == "ZeroMomentum":  # roeexner_zeromomentum.py:17-35
                v1 = self.values[1].astype(dt)
                hu_val, hv_val = v1 * nx, v1 * ny
                m = self.mask
                kw["hu"] = np.where(m, s.hu if np.isnan(hu_val) else hu_val, s.hu)
                kw["hv"] = np.where(m, s.hv if np.isnan(hv_val) else hv_val, s.hv)
            elif self.kind == "ConstantInflux":  # roeexner_constantinflux.py:17-47
                q, vx, vy = self.values[:3]
                hu = s.hu.copy(); hu[by, bx] = (q * vx).astype(dt)
                hv = s.hv.copy(); hv[by, bx] = (q * vy).astype(dt)
                kw["hu"], kw["hv"] = hu, hv
            elif self.kind == "ConstantOutflux":  # roeexner_constantoutflux.py:20-71
                q, vx, vy = self.values[:3]
                qd, nxd, nyd = (np.asarray(v, dtype=dt) for v in (q, vx, vy))
                fl = dt.type(1e-8)
                h_i = np.maximum(s.h[iy, ix], fl)
                u_i = s.hu[iy, ix] / h_i
                v_i = s.hv[iy, ix] / h_i
                un_i = u_i * nxd + v_i * nyd
                c_i = np.sqrt(dt.type(G_ACC) * h_i)
                Rp = un_i + dt.type(2.0) * c_i
                hb = h_i
                for _ in range(5):
                    hb = np.maximum(hb, fl)
                    c = np.sqrt(dt.type(G_ACC) * hb)
                    f = qd / hb + dt.type(2.0) * c - Rp
                    df = -qd / (hb * hb) + dt.type(G_ACC) / np.maximum(c, fl)
                    hb = np.maximum(hb - f / df, fl)
                h = s.h.copy(); h[by, bx] = hb
                hu = s.hu.copy(); hu[by, bx] = (q * vx).astype(dt)
                hv = s.hv.copy(); hv[by, bx] = (q * vy).astype(dt)
                kw["h"], kw["hu"], kw["hv"] = h, hu, hv
            elif self.kind in ("Berthon", "NormalFlowDepth"):  # roeexner_berthon.py, roeexner_normalflowdepth.py
                u, hb, zb0, alpha = _berthon(self.values)
                zb = (-alpha).astype(dt) * tt + zb0.astype(dt)
                z_b = s.z_b.copy(); z_b[by, bx] = zb
                kw["z_b"] = z_b
                if self.kind == "NormalFlowDepth":
                    h = s.h.copy(); h[by, bx] = hb.astype(dt)
                    kw["h"] = h
                    copy("hu")
                    copy("hv")
            elif self.kind == "SteepFall":  # roeexner_steepfall.py:17-26
                z_b = s.z_b.copy(); z_b[by, bx] = s.z_b[iy, ix] - 1
                kw["z_b"] = z_b
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
            self.handlers.append(Handler(spec["type"], scheme, m, spec["normal"], spec.get("values", ())))
        self.boundary_mask = np.any(np.stack(masks), axis=0) if masks else np.zeros(X.shape, bool)

    def apply(self, s, time=0.0):
        for hd in self.handlers:
            s = hd.apply(s, time)
        return s

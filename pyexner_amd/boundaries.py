"""Ghost-cell boundaries: the handler registry and BoundaryManager.

Mirrors domain/boundary_registry.py:11-187.  Handlers are registered under
``f"{flux_scheme} {type}"`` and constructed from a YAML polygon and normal.  The
polygons are rasterised on the host (point_in_polygon, boundary_registry.py:19-44)
and the handler list is resolved, in YAML order, into one composed program: for
every target cell and field group, the source cell in the state before the
fill and the chain of momentum transforms applied on the way
(boundary_registry.py:183-187 applies the handlers one after another, each
gathering from the state its predecessors left).  The GPU applies the program
in one launch (pxh_bc_apply).

Periodic is not in the reference (README.md:163 advertises it; nothing is
registered in domain/boundaries/__init__.py).  Here a Periodic boundary cell takes
every field from the cell N-2 cells against its outward normal: the last owned
column or row on the opposite side of the domain.
"""
from dataclasses import dataclass

import numpy as np
import torch

from . import _hip

BOUNDARY_REGISTRY = {}


def register_boundary(name):
    def decorator(cls):
        BOUNDARY_REGISTRY[name] = cls
        return cls
    return decorator


# ---------------------------------------------------------------------------
# geometry (host, set-up only)
# ---------------------------------------------------------------------------
def point_in_polygon(points, polygon):
    """Even-odd ray test with the reference's half-open crossing rule and its
    1e-8 tolerance (boundary_registry.py:19-44), vectorised over points (N, 2)."""
    polygon = np.asarray(polygon, dtype=np.float64)
    x = points[:, 0][:, None]
    y = points[:, 1][:, None]
    xi, yi = polygon[None, :, 0], polygon[None, :, 1]
    xj, yj = np.roll(polygon[:, 0], 1)[None, :], np.roll(polygon[:, 1], 1)[None, :]
    inside = np.zeros(points.shape[0], dtype=np.int64)
    crosses = (yi > y) != (yj > y)
    denom = np.where(crosses, yj - yi, 1.0)
    test_x = xi + (xj - xi) * (y - yi) / denom
    inside = np.sum(crosses & (x <= test_x + 1e-8), axis=1)
    return inside % 2 == 1


def compute_label_mask(X, Y, polygon):
    """Cells whose centre is inside the polygon.  Only points inside the
    polygon's bounding box (widened by the 1e-8 tolerance) are tested: a point
    outside it crosses no edge (above, below or right of the box) or every
    edge its row crosses (left of the box), an even count either way."""
    X = np.asarray(X, np.float64)
    Y = np.asarray(Y, np.float64)
    poly = np.asarray(polygon, dtype=np.float64)
    lo, hi = poly.min(axis=0), poly.max(axis=0)
    cand = np.nonzero(((X >= lo[0] - 1e-8) & (X <= hi[0] + 1e-8) & (Y >= lo[1]) & (Y <= hi[1])).ravel())[0]
    out = np.zeros(X.size, dtype=bool)
    pts = np.stack([X.ravel()[cand], Y.ravel()[cand]], axis=-1)
    step = 1 << 20
    for s in range(0, len(cand), step):
        out[cand[s:s + step]] = point_in_polygon(pts[s:s + step], poly)
    return out.reshape(X.shape)


def _clip_pairs(b, shift, shape):
    ny, nx = shape
    i = b + np.asarray(shift, dtype=np.int64)
    i[:, 0] = np.clip(i[:, 0], 0, ny - 1)
    i[:, 1] = np.clip(i[:, 1], 0, nx - 1)
    return i


def compute_reflective_indices(mask, normal):
    """boundary_registry.py:61-82: interior = boundary + (round(ny), -round(nx)), clipped."""
    b = np.argwhere(mask)
    return b, _clip_pairs(b, (int(np.round(normal[1])), -int(np.round(normal[0]))), mask.shape)


def compute_periodic_indices(mask, normal):
    ny, nx = mask.shape
    b = np.argwhere(mask)
    shift = (int(np.round(normal[1])) * (ny - 2), -int(np.round(normal[0])) * (nx - 2))
    return b, _clip_pairs(b, shift, mask.shape)


# ---------------------------------------------------------------------------
# handlers: what each writes, and the momentum transform it applies
# ---------------------------------------------------------------------------
G = _hip


@dataclass
class Handler:
    mask: np.ndarray
    normal: np.ndarray
    boundary_indices: np.ndarray = None
    interior_indices: np.ndarray = None
    values: np.ndarray = None

    groups = ()
    kind = _hip.BC_COPY
    periodic = False

    def __post_init__(self):
        self.normal = np.asarray(self.normal, dtype=np.float64)
        if self.boundary_indices is None:
            fn = compute_periodic_indices if self.periodic else compute_reflective_indices
            self.boundary_indices, self.interior_indices = fn(self.mask, self.normal)

    def value_spec(self, grp):
        """(kind, a, b, c) of the value this handler writes into group grp
        (pxh_bc_value), or None when it copies the group or leaves it."""
        return None


def _f32(v):
    """A YAML boundary value as the reference holds it: a float32 array
    element (boundary_registry.py:103-107)."""
    return np.float32(v)


def _berthon_profile(vals):
    """u, h = q / u and zb0 of the Berthon et al. steady state at the boundary's
    x (roeexner_berthon.py:35-49, roeexner_normalflowdepth.py:31-40), in
    float32 like the reference's float32 values: values = [q, A, alpha, beta,
    gamma, x]."""
    q, A, alpha, beta, gamma, x = (_f32(v) for v in vals[:6])
    u = np.power(np.float32((alpha * x + beta) / A), np.float32(1.0 / 3.0))
    h = np.float32(q / u)
    two_g = np.float32(2.0 * 9.81)
    zb0 = np.float32(-(u * (u * u) + two_g * q) / (two_g * np.maximum(u, np.float32(1e-12))) + gamma)
    return u, h, zb0, alpha


@register_boundary("Roe Exner Reflective")
class RoeExner_ReflectiveBoundary(Handler):
    """roeexner_reflective.py:17-63: h, z_b, n, G copied; normal momentum negated."""
    groups = (G.G_H, G.G_HUV, G.G_ZB, G.G_N, G.G_G)
    kind = _hip.BC_REFLECT_RX


@register_boundary("Roe Exner Transmissive")
class RoeExner_TransmissiveBoundary(Handler):
    """roeexner_transmissive.py:17-31: h, hu, hv copied."""
    groups = (G.G_H, G.G_HUV)


@register_boundary("Roe Exner TransmissiveBed")
class RoeExner_TransmissiveBedBoundary(Handler):
    """roeexner_transmissivebed.py:17-27: z_b copied."""
    groups = (G.G_ZB,)


@register_boundary("Roe Exner Periodic")
class RoeExner_PeriodicBoundary(Handler):
    groups = (G.G_H, G.G_HUV, G.G_Z, G.G_ZB, G.G_N, G.G_NB, G.G_G)
    periodic = True


@register_boundary("Roe Exner ZeroMomentum")
class RoeExner_ZeroMomentumBoundary(Handler):
    """roeexner_zeromomentum.py:17-35: hu = values[1] * normal[0], hv =
    values[1] * normal[1] on every boundary cell (a NaN keeps the cell's)."""
    def value_spec(self, grp):
        if grp != G.G_HUV:
            return None
        v1 = float(_f32(self.values[1]))
        return (_hip.BV_CONST, v1 * float(self.normal[0]), v1 * float(self.normal[1]), 0.0)


@register_boundary("Roe Exner ConstantInflux")
class RoeExner_ConstantInfluxBoundary(Handler):
    """roeexner_constantinflux.py:17-47: (hu, hv) = q (values[1], values[2])
    with q = values[0]."""
    def value_spec(self, grp):
        if grp != G.G_HUV:
            return None
        q, nx, ny = (_f32(v) for v in self.values[:3])
        return (_hip.BV_CONST, float(np.float32(q * nx)), float(np.float32(q * ny)), 0.0)


@register_boundary("Roe Exner ConstantOutflux")
class RoeExner_ConstantOutfluxBoundary(Handler):
    """roeexner_constantoutflux.py:20-71: h from the interior cell's outgoing
    Riemann invariant at discharge q = values[0] through (values[1],
    values[2]); (hu, hv) = q (values[1], values[2])."""
    def value_spec(self, grp):
        q, nx, ny = (_f32(v) for v in self.values[:3])
        if grp == G.G_H:
            return (_hip.BV_OUTFLUX, float(q), float(nx), float(ny))
        if grp == G.G_HUV:
            return (_hip.BV_CONST, float(np.float32(q * nx)), float(np.float32(q * ny)), 0.0)
        return None


@register_boundary("Roe Exner NormalFlowDepth")
class RoeExner_NormalFlowDepthBoundary(Handler):
    """roeexner_normalflowdepth.py:20-47: h = q / u of the Berthon profile,
    (hu, hv) copied from the interior cell, z_b = zb0 - alpha t."""
    groups = (G.G_HUV,)

    def value_spec(self, grp):
        u, h, zb0, alpha = _berthon_profile(self.values)
        if grp == G.G_H:
            return (_hip.BV_CONST, float(h), 0.0, 0.0)
        if grp == G.G_ZB:
            return (_hip.BV_AFFINE_T, float(zb0), -float(alpha), 0.0)
        return None


@register_boundary("Roe Exner Berthon")
class RoeExner_BerthonBoundary(Handler):
    """roeexner_berthon.py:19-53: z_b = zb0 - alpha t of the Berthon profile."""
    def value_spec(self, grp):
        if grp != G.G_ZB:
            return None
        u, h, zb0, alpha = _berthon_profile(self.values)
        return (_hip.BV_AFFINE_T, float(zb0), -float(alpha), 0.0)


@register_boundary("Roe Exner SteepFall")
class RoeExner_SteepFall(Handler):
    """roeexner_steepfall.py:17-26: z_b = z_b of the interior cell - 1."""
    def value_spec(self, grp):
        return (_hip.BV_OFFSET, -1.0, 0.0, 0.0) if grp == G.G_ZB else None


@register_boundary("Roe Reflective")
class Roe_ReflectiveBoundary(Handler):
    """roe_reflective.py:18-44: h, z, n copied; hu*(1-2nx^2), hv*(1-2ny^2)."""
    groups = (G.G_H, G.G_Z, G.G_N, G.G_HUV)
    kind = _hip.BC_REFLECT_ROE


@register_boundary("Roe Transmissive")
class Roe_TransmissiveBoundary(Handler):
    """roe_transmissive.py:17-43: h, z, n, hu, hv copied."""
    groups = (G.G_H, G.G_Z, G.G_N, G.G_HUV)


@register_boundary("Roe Periodic")
class Roe_PeriodicBoundary(Handler):
    groups = (G.G_H, G.G_HUV, G.G_Z, G.G_N)
    periodic = True


STATIC_GROUPS = (G.G_Z, G.G_N, G.G_NB)


# ---------------------------------------------------------------------------
# composed program
# ---------------------------------------------------------------------------
# pxh_bc_entry (16 bytes)
ENTRY_DTYPE = np.dtype([("target", "<i4"), ("src", "<i4"), ("group", "u1"), ("nchain", "u1"),
                        ("chain", "u1", (_hip.MAX_CHAIN,))])


def _reflect(hd, hu, hv):
    """The momentum transform of handler hd (as bc_transform applies it) on
    host floats."""
    nx, ny = float(hd.normal[0]), float(hd.normal[1])
    if hd.kind == _hip.BC_REFLECT_RX:
        mn = hu * nx + hv * ny
        mt = -hu * ny + hv * nx
        mr = -mn
        return mr * nx - mt * ny, mr * ny + mt * nx
    if hd.kind == _hip.BC_REFLECT_ROE:
        return hu * (1.0 - 2.0 * nx * nx), hv * (1.0 - 2.0 * ny * ny)
    return hu, hv


def _spec_key(spec):
    """Identity of a value spec bit for bit (-0.0 differs from 0.0, NaN equals NaN)."""
    return (spec[0],) + tuple(float(v).hex() for v in spec[1:])


def compose(handlers, shape):
    """Resolve the sequential handlers into per-target entries.

    Returns (entries: structured numpy array of pxh_bc_entry, idempotent: bool,
    values: list of (kind, a, b, c) pxh_bc_value specs the value entries
    index, level_start: the entries' dependency levels, entries[level_start[L]:
    level_start[L + 1]] being level L; the entries are ordered by level, then
    group).  Each (group, target) ends up as one entry: a copy of a source
    cell through a chain of momentum transforms, or a value (a constant, an
    affine function of time, an offset of a source cell, or the outflux depth
    of a source cell).  A program is idempotent when no entry reads a cell
    the program writes in a group the entry reads: then one launch reads only
    values the launch never writes, and applying it twice equals applying it
    once."""
    ny, nx = shape
    rows = []
    idempotent = True
    values, value_keys = [], []
    # cells written by handlers before handler k, in any group (a value
    # handler reading such a cell is not supported)
    written_before = []
    acc = set()
    for hd in handlers:
        written_before.append(set(acc))
        acc |= set((hd.boundary_indices[:, 0] * nx + hd.boundary_indices[:, 1]).tolist())
    targets_of = {}
    reads = []  # (src, groups read) of the entries
    for grp in range(7):
        src_map, chain_map, val_map = {}, {}, {}
        for k, hd in enumerate(handlers):
            spec = hd.value_spec(grp)
            copies = grp in hd.groups
            if spec is None and not copies:
                continue
            b = hd.boundary_indices[:, 0] * nx + hd.boundary_indices[:, 1]
            i = hd.interior_indices[:, 0] * nx + hd.interior_indices[:, 1]
            transform = grp == G.G_HUV and hd.kind != _hip.BC_COPY
            new = []
            for bb, ii in zip(b.tolist(), i.tolist()):
                if spec is not None:
                    if spec[0] == _hip.BV_OFFSET:
                        # an offset of a cell an earlier handler copied into reads that copy's source
                        if ii in val_map or chain_map.get(ii):
                            raise NotImplementedError("an offset boundary reading another boundary's value")
                        new.append((bb, ("val", spec, src_map.get(ii, ii))))
                    elif spec[0] == _hip.BV_OUTFLUX:
                        if ii in written_before[k]:
                            raise NotImplementedError("an outflux boundary reading a cell an earlier handler writes")
                        new.append((bb, ("val", spec, ii)))
                    else:
                        new.append((bb, ("val", spec, bb)))
                elif ii in val_map:
                    vs, vsrc = val_map[ii]
                    if transform:
                        # a reflected constant momentum is a constant
                        if vs[0] != _hip.BV_CONST or vs[1] != vs[1] or vs[2] != vs[2]:
                            raise NotImplementedError("a reflection of a value boundary's momentum")
                        vs = (_hip.BV_CONST,) + _reflect(hd, vs[1], vs[2]) + (0.0,)
                    new.append((bb, ("val", vs, vsrc if vs[0] in (_hip.BV_OFFSET, _hip.BV_OUTFLUX) else bb)))
                else:
                    ch = chain_map.get(ii, ())
                    new.append((bb, ("copy", src_map.get(ii, ii), ch + (k,) if transform else ch)))
            for bb, x in new:
                if x[0] == "val":
                    val_map[bb] = (x[1], x[2])
                    src_map.pop(bb, None)
                    chain_map.pop(bb, None)
                else:
                    src_map[bb] = x[1]
                    chain_map[bb] = x[2]
                    val_map.pop(bb, None)
        targets = set(src_map) | set(val_map)
        targets_of[grp] = targets
        for t, s in src_map.items():
            ch = chain_map[t]
            if s == t and not ch:
                continue
            if s in targets:
                idempotent = False
            if len(ch) > _hip.MAX_CHAIN:
                raise NotImplementedError(f"boundary chain of {len(ch)} reflections exceeds {_hip.MAX_CHAIN}")
            rows.append((t, s, grp, ch, None))
        for t, (spec, src) in val_map.items():
            key = _spec_key(spec)
            if key not in value_keys:
                value_keys.append(key)
                values.append(spec)
            if spec[0] == _hip.BV_OFFSET:
                reads.append((src, (grp,)))
            elif spec[0] == _hip.BV_OUTFLUX:
                reads.append((src, (G.G_H, G.G_HUV)))
            rows.append((t, src, grp, (), value_keys.index(key)))
    for src, grps in reads:
        if any(src in targets_of.get(g_, ()) for g_ in grps):
            idempotent = False
    if len(values) > _hip.MAX_VALUES:
        raise NotImplementedError(f"more than {_hip.MAX_VALUES} boundary values")
    level = dependency_levels(rows, values)
    if level and max(level) + 1 > _hip.MAX_LEVELS:
        raise NotImplementedError(f"a ring fill of more than {_hip.MAX_LEVELS} dependency levels")
    order = sorted(range(len(rows)), key=lambda q: (level[q], rows[q][2]))
    rows = [rows[q] for q in order]
    level = [level[q] for q in order]
    ent = np.zeros(len(rows), dtype=ENTRY_DTYPE)
    for q, (t, s, grp, ch, vi) in enumerate(rows):
        ent[q]["target"], ent[q]["src"], ent[q]["group"] = t, s, grp
        if vi is None:
            ent[q]["nchain"] = len(ch)
            ent[q]["chain"][:len(ch)] = ch
        else:
            ent[q]["nchain"] = _hip.BC_VALUE
            ent[q]["chain"][0] = vi
    assert ent.itemsize == 16
    n_levels = (max(level) + 1) if level else 1
    level_start = [int(v) for v in np.searchsorted(np.asarray(level, dtype=np.int64), np.arange(n_levels + 1))]
    return ent, n_levels == 1, values, level_start


def dependency_levels(rows, values):
    """The level of each composed entry (t, src, group, chain, value index):
    an entry that reads a (cell, group) which another entry writes runs in an
    earlier level than that writer, so that every entry reads its source as
    it was before the program, as the sequential handlers of
    boundary_registry.py:183-187 read theirs (each handler gathers its sources
    before it writes).  Copies read their source's group, an offset value its
    source's group, an outflux depth its source's h, hu and hv."""
    writer = {(r[0], r[2]): q for q, r in enumerate(rows)}
    reads = []
    for t, src, grp, ch, vi in rows:
        if vi is None:
            reads.append(((src, grp),))
        elif values[vi][0] == _hip.BV_OFFSET:
            reads.append(((src, grp),))
        elif values[vi][0] == _hip.BV_OUTFLUX:
            reads.append(((src, G.G_H), (src, G.G_HUV)))
        else:
            reads.append(())
    level = [0] * len(rows)
    for _ in range(len(rows) + 1):
        changed = False
        for q, rd in enumerate(reads):
            for key in rd:
                w = writer.get(key)
                if w is not None and w != q and level[w] < level[q] + 1:
                    level[w] = level[q] + 1
                    changed = True
        if not changed:
            return level
    raise NotImplementedError("a ring fill whose entries read each other's targets in a cycle")


def frame_local(entries, shape):
    """True when every entry's target lies on the outer ring of cells and its
    source within two cells of the grid edge: the ring fill then reads and
    writes only the two-cell frame, so the bed update of every cell off the
    frame can run before the fill (the fused stage)."""
    ny, nx = shape
    if len(entries) == 0:
        return True
    t = entries["target"].astype(np.int64)
    s = entries["src"].astype(np.int64)
    tr, tc = t // nx, t % nx
    sr, sc = s // nx, s % nx
    ring = (tr == 0) | (tr == ny - 1) | (tc == 0) | (tc == nx - 1)
    frame = (sr <= 1) | (sr >= ny - 2) | (sc <= 1) | (sc >= nx - 2)
    return bool(np.all(ring) and np.all(frame))


def set_values(bc, values, time=0.0):
    """Fill pxh_bc's value table from (kind, a, b, c) specs."""
    bc.n_values = len(values)
    bc.time = float(time)
    for k, (kind, a, b, c) in enumerate(values):
        bc.values[k].kind, bc.values[k].a, bc.values[k].b, bc.values[k].c = kind, a, b, c


class BoundaryManager:
    """BoundaryManager(params, X, Y) (boundary_registry.py:84-187)."""

    def __init__(self, params, X, Y, device="cuda"):
        self.X, self.Y = np.asarray(X), np.asarray(Y)
        self.flux_scheme = params["flux_scheme"]
        self.boundary_specs = params.get("boundaries") or {}
        self.boundary_handlers = []
        self.names = []
        shape = self.X.shape
        masks = []
        for key, spec in self.boundary_specs.items():
            self.names.append(key)
            btype = f"{self.flux_scheme} {spec['type']}"
            if btype not in BOUNDARY_REGISTRY:
                raise KeyError(f"Unknown boundary '{btype}'. Available options: {list(BOUNDARY_REGISTRY.keys())}")
            m = compute_label_mask(self.X, self.Y, spec["polygon"])
            masks.append(m)
            vals = np.array([np.nan if str(v).lower() == "nan" else float(v) for v in spec.get("values", [])],
                            dtype=np.float32).astype(np.float64)
            self.boundary_handlers.append(BOUNDARY_REGISTRY[btype](mask=m, normal=spec["normal"], values=vals))
        self.boundary_mask = np.any(np.stack(masks), axis=0) if masks else np.zeros(shape, dtype=bool)
        if len(self.boundary_handlers) > _hip.MAX_HANDLERS:
            raise NotImplementedError(f"more than {_hip.MAX_HANDLERS} boundary handlers")
        entries, self.idempotent, self.values, self.level_start = compose(self.boundary_handlers, shape)
        self.n_levels = len(self.level_start) - 1
        self.time_dependent = any(v[0] == _hip.BV_AFFINE_T for v in self.values)
        self.entries = entries
        self.n_entries = len(entries)
        self.frame_local = frame_local(entries, shape)
        self.device = device
        self._entries_dev = torch.from_numpy(entries.view(np.uint8).copy()).to(device) if len(entries) else None
        self.groups_present = sorted({int(g) for g in entries["group"]}) if len(entries) else []
        grp = entries["group"].astype(np.int64)
        if self.n_levels == 1:
            assert np.all(np.diff(grp) >= 0), "compose() emits one level's entries group by group"
            self.group_start = [int(v) for v in np.searchsorted(grp, np.arange(_hip.N_GROUPS + 1))]
        else:
            self.group_start = None

    def program(self, groups=None):
        """pxh_bc struct restricted to the given groups (default: all)."""
        bc = _hip.BC()
        bc.entries = self._entries_dev.data_ptr() if self._entries_dev is not None else None
        bc.n_entries = self.n_entries
        bc.n_handlers = len(self.boundary_handlers)
        mask = 0
        for g_ in (range(7) if groups is None else groups):
            mask |= 1 << g_
        bc.group_mask = mask
        bc.flags = _hip.BC_FRAME if self.frame_local else 0
        if self.n_levels == 1:
            bc.flags |= _hip.BC_GROUPED
            for g_ in range(_hip.N_GROUPS + 1):
                bc.group_start[g_] = self.group_start[g_]
        else:
            bc.n_levels = self.n_levels
            for k, v in enumerate(self.level_start):
                bc.level_start[k] = v
        for k, hd in enumerate(self.boundary_handlers):
            bc.handlers[k].kind = hd.kind
            bc.handlers[k].nx = float(hd.normal[0])
            bc.handlers[k].ny = float(hd.normal[1])
        set_values(bc, self.values)
        return bc

    def adjoint_program(self, groups=None):
        """The program ordered for the adjoint fill (pxh_adj_*): entries
        sorted by (group, src), stably, so that the targets filled from one
        source are adjacent and their adjoints are summed in a fixed order
        (PXH_BC_SRC_SORTED)."""
        if self.n_levels > 1:
            raise NotImplementedError("the adjoint ring fill needs a program of one dependency level")
        if not hasattr(self, "_adj_entries_dev"):
            ent = self.entries
            order = np.lexsort((ent["src"], ent["group"])) if len(ent) else np.zeros(0, dtype=np.int64)
            self._adj_entries = ent[order]
            self._adj_entries_dev = (torch.from_numpy(self._adj_entries.view(np.uint8).copy()).to(self.device)
                                     if len(ent) else None)
        bc = self.program(groups)
        bc.entries = self._adj_entries_dev.data_ptr() if self._adj_entries_dev is not None else None
        bc.flags = _hip.BC_GROUPED | _hip.BC_SRC_SORTED
        return bc

    def apply_inplace(self, state, groups=None, time=0.0):
        from .state import fields_struct, dtype_code
        if self.n_entries == 0:
            return state
        grid = _hip.Grid(state.h.shape[0], state.h.shape[1], dtype_code(state.h.dtype), 0, 0.0)
        fs = fields_struct(state)
        bc = self.program(groups)
        bc.time = float(time)
        stream = torch.cuda.current_stream().cuda_stream
        _hip.call("pxh_bc_apply", grid, fs, bc, stream)
        return state

    def apply(self, state, time=0.0):
        """Functional ring fill: returns a new state (boundary_registry.py:183-187)."""
        return self.apply_inplace(state.clone(), time=time)

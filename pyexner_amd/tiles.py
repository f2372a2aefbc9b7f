"""A parNy x parNx tile decomposition of one Roe-Exner domain on one device.

The reference decomposes a domain over MPI ranks (`parallel/mpi_handler.py`),
gives every rank a tile with a one-cell ghost ring, and swaps the ring with
`make_halo_exchange` (solvers/kernels/roe_exner.py:719-803) inside the step
(solvers/roe_exner_solver.py:92-110).  Here each tile keeps its owned block
plus two or three ghost rings on every side it shares with another tile
(none on the sides of the physical boundary).  Tiles on one device swap
ghost strips by batched device copies straight from the neighbours' owned
strips (`pxh_copy_batch`, one launch per exchange phase); with
direct=False the strips are packed and unpacked per side with
`pxh_halo_pack` / `pxh_halo_unpack`, the buffers a multi-GPU run would
send and receive over RCCL.

An owned cell's stage result reads the stage input up to three cells away:
its bed update reads its neighbours' updated hydro, whose momentum
corrections read the updated depth one cell further, whose hydro update
reads one more.  (With the reference's single ring that depth is computed
from an incomplete stencil.)  Two paths make the tiled run equal the
single-domain run cell for cell:

* fused (frame-local rings, the default): three ghost rings, so every owned
  cell's stage is complete from the tile's own data; each tile runs the
  fused stage (stage sweep -> fill (h, hu, hv) -> frame sweep -> fill
  (z_b, G)) and the tiles swap every evolving field once per stage, then the
  dt from every tile's own minimum after the swap, reduced over tiles:

      b  = S(a) per tile;                      exchange(b)
      a  = 0.5 a + 0.5 S(b) per tile;          exchange(a)
      dt = min over tiles of compute_dt(a)     (the MIN allreduce)

* two-sweep: two ghost rings, and an exchange after each sweep:

      dt = min over tiles of compute_dt(a)
      b  = hydro(a); fill(b: h, hu, hv); exchange(b: h, hu, hv)
      b  = bed(b, a); fill(b: z_b, G); exchange(b: z_b, G)
      c  = hydro(b); fill(c: h, hu, hv); exchange(c: h, hu, hv)
      a  = 0.5 a + 0.5 bed(c, b); fill(a: z_b, G); exchange(a: z_b, G)
"""
import numpy as np
import torch

from . import _hip
from .boundaries import ENTRY_DTYPE, set_values
from .solvers.roe_exner_solver import workspace_bytes
from .state import fields_struct

HW = 2  # ghost rings on a shared side, two-sweep path
_ALL_FIELDS = ("h", "hu", "hv", "z", "z_b", "n", "n_b", "G", "z_b_lo")
HW_FUSED = 3  # ghost rings on a shared side, fused path
_SIDES = (_hip.HALO_W, _hip.HALO_N, _hip.HALO_E, _hip.HALO_S)
_OPPOSITE = {_hip.HALO_W: _hip.HALO_E, _hip.HALO_E: _hip.HALO_W, _hip.HALO_N: _hip.HALO_S,
             _hip.HALO_S: _hip.HALO_N}
_STEP = (1 << _hip.G_H) | (1 << _hip.G_HUV) | (1 << _hip.G_ZB) | (1 << _hip.G_G)
_HYDRO = (1 << _hip.G_H) | (1 << _hip.G_HUV)
_BED = (1 << _hip.G_ZB) | (1 << _hip.G_G)
# the fields of each group in the order pxh_halo_pack takes them
_GROUP_FIELDS = ((_hip.G_H, ("h",)), (_hip.G_HUV, ("hu", "hv")), (_hip.G_Z, ("z",)), (_hip.G_ZB, ("z_b", "z_b_lo")),
                 (_hip.G_N, ("n",)), (_hip.G_NB, ("n_b",)), (_hip.G_G, ("G",)))


def split(n, parts):
    """Owned extents [start, stop) of `parts` near-equal bands of n cells."""
    edges = np.linspace(0, n, parts + 1).round().astype(int)
    return [(int(edges[k]), int(edges[k + 1])) for k in range(parts)]


def _like(state, names):
    kw = {n: torch.empty_like(getattr(state, n)) for n in names}
    return state.replace(**kw)


def _tile_program(entries, shape, owned, ext):
    """The global ring-fill program restricted to the fills whose target this
    tile owns, in the tile's local indices (sources must lie in the tile)."""
    ny, nx = shape
    r0, r1, c0, c1 = owned
    e0, _, f0, _ = ext
    tny, tnx = ext[1] - ext[0], ext[3] - ext[2]
    t = entries["target"].astype(np.int64)
    s = entries["src"].astype(np.int64)
    tr, tc, sr, sc = t // nx, t % nx, s // nx, s % nx
    keep = (tr >= r0) & (tr < r1) & (tc >= c0) & (tc < c1)
    sub = entries[keep].copy()
    srl, scl = sr[keep] - e0, sc[keep] - f0
    if np.any((srl < 0) | (srl >= tny) | (scl < 0) | (scl >= tnx)):
        raise NotImplementedError("a ring fill reads a cell outside its tile (periodic rings do not tile)")
    sub["target"] = ((tr[keep] - e0) * tnx + (tc[keep] - f0)).astype(np.int32)
    sub["src"] = (srl * tnx + scl).astype(np.int32)
    grp = sub["group"].astype(np.int64)
    start = [int(v) for v in np.searchsorted(grp, np.arange(_hip.N_GROUPS + 1))]
    return sub, start


class Tile:
    """One tile: its extent in the global arrays, three buffer sets, grid,
    mask, ring-fill program and dt workspace."""

    def __init__(self, backend, state, mask, entries, handlers, pos, par, rows, cols, device, values=(), hw=HW):
        ty, tx = pos
        pny, pnx = par
        r0, r1 = rows[ty]
        c0, c1 = cols[tx]
        self.pos = pos
        self.hw = hw
        self.ring = {_hip.HALO_N: hw if ty > 0 else 0, _hip.HALO_S: hw if ty < pny - 1 else 0,
                     _hip.HALO_W: hw if tx > 0 else 0, _hip.HALO_E: hw if tx < pnx - 1 else 0}
        self.owned = (r0, r1, c0, c1)
        self.ext = (r0 - self.ring[_hip.HALO_N], r1 + self.ring[_hip.HALO_S],
                    c0 - self.ring[_hip.HALO_W], c1 + self.ring[_hip.HALO_E])
        e0, e1, f0, f1 = self.ext
        if min(r1 - r0, c1 - c0) < 2 * hw + 1:
            raise ValueError(f"tile {pos} owns {r1 - r0}x{c1 - c0} cells: too small for {hw} ghost rings")
        names = [n for n in ("h", "hu", "hv", "z", "z_b", "n", "n_b", "G", "z_b_lo") if getattr(state, n) is not None]
        self.a = state.replace(**{n: getattr(state, n)[e0:e1, f0:f1].contiguous() for n in names})
        evolving = ("h", "hu", "hv", "z_b", "G") + (("z_b_lo",) if state.z_b_lo is not None else ())
        self.b = _like(self.a, evolving)
        self.c = _like(self.a, evolving)
        self.fa, self.fb, self.fc = (fields_struct(x) for x in (self.a, self.b, self.c))
        self.mask = mask[:, e0:e1, f0:f1].contiguous()
        self.m0 = self.mask[0].data_ptr()
        g = backend.grid
        self.grid = _hip.Grid(e1 - e0, f1 - f0, g.dtype, g.flags, g.dx)
        sub, start = _tile_program(entries, state.h.shape, self.owned, self.ext)
        self.sub = sub  # local indices; TiledExner folds every tile's into one arena program
        self.n_entries = len(sub)
        self._entries = torch.from_numpy(sub.view(np.uint8).copy()).to(device) if len(sub) else None
        self.programs = {}
        for key, groups in (("hy", _HYDRO), ("bed1", (1 << _hip.G_ZB) | (0 if backend.sed.mpm else 1 << _hip.G_G)),
                            ("bed", _BED)):
            bc = _hip.BC()
            bc.entries = self._entries.data_ptr() if self._entries is not None else None
            bc.n_entries = self.n_entries
            bc.n_handlers = len(handlers)
            bc.group_mask = groups
            bc.flags = _hip.BC_GROUPED
            for k in range(_hip.N_GROUPS + 1):
                bc.group_start[k] = start[k]
            for k, (kind, nxn, nyn) in enumerate(handlers):
                bc.handlers[k].kind, bc.handlers[k].nx, bc.handlers[k].ny = kind, nxn, nyn
            set_values(bc, values)
            self.programs[key] = bc
        ws = workspace_bytes(self.grid)  # sized for any strip height: it never moves
        if ws < 0:
            raise ValueError("invalid tile geometry")
        self.workspace = torch.empty(int(ws), dtype=torch.uint8, device=device)
        self.dt = torch.zeros((), dtype=state.h.dtype, device=device)
        lib = _hip.load()
        self.send = {}
        for side in _SIDES:
            if self.ring[side]:
                nbytes = lib.pxh_halo_bytes(self.grid, self.fa, hw, _STEP, side)
                if nbytes < 0:
                    raise ValueError("invalid halo geometry")
                self.send[side] = torch.empty(int(nbytes), dtype=torch.uint8, device=device)

    def fields(self, which):
        return {"a": self.fa, "b": self.fb, "c": self.fc}[which]


class TiledExner:
    """Roe-Exner SSPRK2 on a par = (parNy, parNx) decomposition of an
    initialised single-domain state (init_fn applied), every tile on this device."""

    def __init__(self, backend, state, mask, boundaries, par, cfl, fused=None, direct=True, multi=True,
                 static=True):
        """fused: the fused-stage path (three ghost rings); default when the
        ring fill is frame-local.  direct: exchange ghost strips by batched
        device copies (default) rather than pack/unpack per side.  multi: run
        every tile's fused sweeps in one launch each (default) rather than
        one launch per tile.  static: check once whether no tile has a
        nodata cell and each tile's z holds one value, and if so let the
        fused stage sweeps skip those streams (PXH_NO_NODATA /
        PXH_Z_UNIFORM; results unchanged)."""
        self.be = backend
        self.cfl = float(cfl)
        self.par = tuple(par)
        pny, pnx = self.par
        ny, nx = state.h.shape
        self.rows, self.cols = split(ny, pny), split(nx, pnx)
        entries = boundaries.entries if boundaries is not None else np.zeros(0, dtype=ENTRY_DTYPE)
        values = boundaries.values if boundaries is not None else []
        if boundaries is not None and boundaries.time_dependent:
            raise NotImplementedError("tiles run at time 0: boundaries whose values move with time are not tiled")
        if boundaries is not None and boundaries.n_levels > 1:
            raise NotImplementedError("tiles take ring fills of one dependency level")
        handlers = [(hd.kind, float(hd.normal[0]), float(hd.normal[1])) for hd in
                    (boundaries.boundary_handlers if boundaries is not None else [])]
        device = state.h.device
        frame_ok = boundaries is None or boundaries.n_entries == 0 or boundaries.frame_local
        self.fused = frame_ok if fused is None else bool(fused)
        if self.fused and not frame_ok:
            raise NotImplementedError("the fused tile path needs a frame-local ring fill")
        self.hw = HW_FUSED if self.fused else HW
        self.tiles = {(ty, tx): Tile(backend, state, mask, entries, handlers, (ty, tx), self.par, self.rows,
                                     self.cols, device, values, self.hw) for ty in range(pny) for tx in range(pnx)}
        ts = list(self.tiles.values())
        self.no_nodata = not any(bool(t.mask[0].any()) for t in ts)
        self.z_uniform = self.no_nodata and all(bool((t.a.z == t.a.z.reshape(-1)[0]).all()) for t in ts)
        if static and self.no_nodata:
            extra = _hip.NO_NODATA | (_hip.Z_UNIFORM if self.z_uniform else 0)
            for t in ts:
                t.grid.flags = int(t.grid.flags) | extra
        self.dt = torch.zeros((), dtype=state.h.dtype, device=device)
        self.dt_ready = False
        self.direct = bool(direct)
        self.multi = bool(multi) and self.fused
        self._jobs = {}
        self._dt = None
        self._sweep_tabs = {}
        self._arena_fill(handlers, values, device)

    def _arena_fill(self, handlers, values, device):
        """Every tile's buffer sets in one arena per field and set (each tile
        a contiguous view of it), and the tiles' ring-fill programs folded
        into one program over the arenas: a fill of every tile is one launch
        (per dependency group set) instead of one per tile.  A tile's fill
        reads and writes only its own cells, so the folded program is the
        per-tile programs side by side."""
        tiles = list(self.tiles.values())
        off, total = [], 0
        for t in tiles:
            off.append(total)
            total += t.grid.ny * t.grid.nx
        total += total & 1  # the arena grid below is 2 x total/2
        self._arena_cells = total
        static = ("z", "n", "n_b")
        names = [n for n in _ALL_FIELDS if getattr(tiles[0].a, n) is not None]
        self._arena = {}
        for which in ("a", "b", "c"):
            kw = {}
            for name in names:
                if which != "a" and name in static:
                    continue
                src0 = getattr(getattr(tiles[0], which), name)
                arena = torch.empty(total, dtype=src0.dtype, device=device)
                for t, o in zip(tiles, off):
                    cur = getattr(getattr(t, which), name)
                    view = arena[o:o + cur.numel()].view(cur.shape)
                    view.copy_(cur)
                    kw.setdefault(id(t), {})[name] = view
                self._arena[(which, name)] = arena
            for t in tiles:
                setattr(t, which, getattr(t, which).replace(**kw[id(t)]))
        for t in tiles:
            # the stage buffer sets share the tile's static fields
            stat = {n: getattr(t.a, n) for n in static}
            t.b, t.c = t.b.replace(**stat), t.c.replace(**stat)
            t.fa, t.fb, t.fc = (fields_struct(x) for x in (t.a, t.b, t.c))
        self._arena_fields = {}
        for which in ("a", "b", "c"):
            f = _hip.Fields()
            for name in names:
                key = ("a", name) if name in static else (which, name)
                setattr(f, name, self._arena[key].data_ptr())
            self._arena_fields[which] = f
        g = self.be.grid
        self._arena_grid = _hip.Grid(2, total // 2, g.dtype, 0, g.dx)
        parts = []
        for t, o in zip(tiles, off):
            if t.n_entries:
                e = t.sub.copy()
                e["target"] = (e["target"].astype(np.int64) + o).astype(np.int32)
                e["src"] = (e["src"].astype(np.int64) + o).astype(np.int32)
                parts.append(e)
        self.n_fill_entries = int(sum(len(p) for p in parts))
        self._fill_progs = {}
        if not self.n_fill_entries:
            return
        allp = np.concatenate(parts)
        allp = allp[np.argsort(allp["group"], kind="stable")]
        grp = allp["group"].astype(np.int64)
        start = [int(v) for v in np.searchsorted(grp, np.arange(_hip.N_GROUPS + 1))]
        self._fill_entries = torch.from_numpy(allp.view(np.uint8).copy()).to(device)
        for key, groups in (("hy", _HYDRO), ("bed1", (1 << _hip.G_ZB) | (0 if self.be.sed.mpm else 1 << _hip.G_G)),
                            ("bed", _BED)):
            bc = _hip.BC()
            bc.entries = self._fill_entries.data_ptr()
            bc.n_entries = len(allp)
            bc.n_handlers = len(handlers)
            bc.group_mask = groups
            bc.flags = _hip.BC_GROUPED
            for k in range(_hip.N_GROUPS + 1):
                bc.group_start[k] = start[k]
            for k, (kind, nxn, nyn) in enumerate(handlers):
                bc.handlers[k].kind, bc.handlers[k].nx, bc.handlers[k].ny = kind, nxn, nyn
            set_values(bc, values)
            self._fill_progs[key] = bc

    def _fill(self, which, key):
        """The ring fill `key` ("hy", "bed1", "bed") of buffer set `which` of
        every tile, in one launch over the arenas."""
        if self.n_fill_entries:
            _hip.call("pxh_bc_apply", self._arena_grid, self._arena_fields[which], self._fill_progs[key],
                      torch.cuda.current_stream().cuda_stream)

    def set_strip_rows(self, rows):
        """Owned rows per strip of every tile's fused stage sweep (8..128):
        a tile is a small grid, and shorter strips give its sweep enough
        wavefronts to fill the GPU.  Results do not depend on it.
        A hipGraph captured earlier keeps the launch shape it was captured
        with: its sweep tables (keyed by strip height) and workspaces stay
        allocated."""
        lib = _hip.load()
        rows = int(rows)
        if not 8 <= rows <= 128:
            raise ValueError(f"strip height {rows} is outside 8..128")
        for t in self.tiles.values():
            t.grid.flags = (int(t.grid.flags) & ~(0xFF << _hip.STRIP_ROWS_SHIFT)) | (rows << _hip.STRIP_ROWS_SHIFT)
            ws = lib.pxh_dt_workspace_size(t.grid)
            if ws < 0 or ws > t.workspace.numel():
                raise ValueError(f"strip height {rows}: dt workspace of {ws} bytes exceeds the allocation")

    def _neighbour(self, pos, side):
        ty, tx = pos
        d = {_hip.HALO_W: (0, -1), _hip.HALO_E: (0, 1), _hip.HALO_N: (-1, 0), _hip.HALO_S: (1, 0)}[side]
        return self.tiles[(ty + d[0], tx + d[1])]

    def _copy_jobs(self, which, groups, phase):
        """The strip copies of one exchange phase as a device table of
        pxh_copy_job (built once per buffer set, groups and phase): every
        ghost strip of the phase from its neighbour's owned strip, in the
        order and with the depths pxh_halo_pack/unpack use."""
        key = (which, groups, phase)
        if key in self._jobs:
            return self._jobs[key]
        names = [n for bit, ns in _GROUP_FIELDS if groups & (1 << bit) for n in ns]
        esz = 8 if self.be.dtype == torch.float64 else 4
        jobs = []
        for t in self.tiles.values():
            ft = getattr(t, which)
            ny, nx = ft.h.shape
            hw = t.hw
            for side in phase:
                if not t.ring[side]:
                    continue
                n = self._neighbour(t.pos, side)
                fn = getattr(n, which)
                my, mx = fn.h.shape
                # (dst row, dst col, src row, src col, rows, cols)
                rect = {_hip.HALO_W: (0, 0, 0, mx - 2 * hw, ny, hw),
                        _hip.HALO_E: (0, nx - hw, 0, hw, ny, hw),
                        _hip.HALO_N: (0, 0, my - 2 * hw, 0, hw, nx),
                        _hip.HALO_S: (ny - hw, 0, hw, 0, hw, nx)}[side]
                dr, dc, sr, sc, rows, cols = rect
                if (side in (_hip.HALO_W, _hip.HALO_E) and my != ny) or \
                        (side in (_hip.HALO_N, _hip.HALO_S) and mx != nx):
                    raise ValueError("neighbouring tiles disagree on the shared strip's length")
                for name in names:
                    a, b = getattr(ft, name), getattr(fn, name)
                    if a is None or b is None:
                        continue
                    jobs.append((b.data_ptr() + (sr * mx + sc) * esz, a.data_ptr() + (dr * nx + dc) * esz,
                                 mx, nx, rows, cols))
        table = (_hip.CopyJob * max(len(jobs), 1))()
        start = 0
        for k, (src, dst, sld, dld, rows, cols) in enumerate(jobs):
            table[k].src, table[k].dst, table[k].src_ld, table[k].dst_ld = src, dst, sld, dld
            table[k].start, table[k].rows, table[k].cols = start, rows, cols
            start += rows * cols
        raw = np.frombuffer(bytes(table), dtype=np.uint8).copy()
        entry = (torch.from_numpy(raw).to(self.tiles[(0, 0)].a.h.device), len(jobs), start)
        self._jobs[key] = entry
        return entry

    def exchange(self, which, groups):
        """Swap the ghost strips of buffer set `which` ("a", "b", "c") for the
        field groups `groups`: west/east first, then north/south (corners).
        direct (default): one batched copy per phase straight from the
        neighbours' owned strips; else a pack and an unpack per side, the
        form whose packed strips would travel between devices."""
        st = torch.cuda.current_stream().cuda_stream
        if self.direct:
            dtype = _hip.PXH_F64 if self.be.dtype == torch.float64 else _hip.PXH_F32
            for phase in ((_hip.HALO_W, _hip.HALO_E), (_hip.HALO_N, _hip.HALO_S)):
                table, n, total = self._copy_jobs(which, groups, phase)
                if n:
                    _hip.call("pxh_copy_batch", dtype, table.data_ptr(), n, total, st)
            return
        for phase in ((_hip.HALO_W, _hip.HALO_E), (_hip.HALO_N, _hip.HALO_S)):
            for t in self.tiles.values():
                for side in phase:
                    if t.ring[side]:
                        _hip.call("pxh_halo_pack", t.grid, t.fields(which), t.hw, groups, side,
                                  t.send[side].data_ptr(), st)
            for t in self.tiles.values():
                for side in phase:
                    if t.ring[side]:
                        src = self._neighbour(t.pos, side).send[_OPPOSITE[side]]
                        _hip.call("pxh_halo_unpack", t.grid, t.fields(which), t.hw, groups, side, src.data_ptr(),
                                  st)

    def _sweep_table(self, src, dst, old):
        """The tiles as a device table of pxh_sweep_domain for one stage
        (src -> dst, combined into old when given), the stage, frame and
        frame-dt block counts of the multi-domain sweeps, and the second
        stage's dt buffers: the per-block minima (stage blocks, then frame-dt
        blocks) and every tile's seam records (built once per stage kind and
        strip height)."""
        key = (src, dst, old, int(next(iter(self.tiles.values())).grid.flags))
        if key not in self._sweep_tabs:
            lib = _hip.load()
            tiles = list(self.tiles.values())
            table = (_hip.SweepDomain * len(tiles))()
            sb = fb = tb = xe = 0
            for k, t in enumerate(tiles):
                d = table[k]
                d.in_, d.out = t.fields(src), t.fields(dst)
                if old is not None:
                    d.old = t.fields(old)
                d.mask0, d.ny, d.nx, d.sblk0, d.fblk0, d.dx = t.m0, t.grid.ny, t.grid.nx, sb, fb, t.grid.dx
                d.xs0, d.tblk0 = xe, tb
                d.shared = sum(side for side in _SIDES if t.ring[side])
                ns, nf = lib.pxh_multi_stage_blocks(t.grid), lib.pxh_multi_frame_blocks(t.grid.ny, t.grid.nx)
                nt, ne = lib.pxh_multi_dt_frame_blocks(t.grid), lib.pxh_multi_seam_elems(t.grid.ny, t.grid.nx)
                if ns <= 0 or nf <= 0 or nt <= 0 or ne <= 0:
                    raise ValueError("a tile the fused stage cannot sweep")
                sb += ns
                fb += nf
                tb += nt
                xe += ne
            raw = np.frombuffer(bytes(table), dtype=np.uint8).copy()
            dev = self.dt.device
            part = xs = None
            if old is not None:
                part = torch.empty(sb + tb, dtype=self.dt.dtype, device=dev)
                xs = torch.empty(xe, dtype=self.dt.dtype, device=dev)
            self._sweep_tabs[key] = (torch.from_numpy(raw).to(dev), len(tiles), sb, fb, tb, part, xs)
        return self._sweep_tabs[key]

    def _dt_table(self):
        """The tiles' buffer set a as a device table of pxh_dt_domain (built
        once) and the partial minima buffer of pxh_dt_multi."""
        if self._dt is None:
            lib = _hip.load()
            tiles = list(self.tiles.values())
            table = (_hip.DtDomain * len(tiles))()
            blk = 0
            for k, t in enumerate(tiles):
                d = table[k]
                d.h, d.hu, d.hv, d.G = (getattr(t.a, f).data_ptr() for f in ("h", "hu", "hv", "G"))
                d.mask0, d.ny, d.nx, d.blk0, d.dx = t.m0, t.grid.ny, t.grid.nx, blk, t.grid.dx
                nb = lib.pxh_dt_multi_blocks(t.grid.ny, t.grid.nx)
                if nb <= 0:
                    raise ValueError("invalid tile geometry")
                blk += nb
            device = self.dt.device
            raw = np.frombuffer(bytes(table), dtype=np.uint8).copy()
            self._dt = (torch.from_numpy(raw).to(device), len(tiles), blk,
                        torch.empty(blk * 8, dtype=torch.uint8, device=device))
        return self._dt

    def compute_dt(self):
        """compute_dt_fn over every tile (the MIN allreduce): one sweep launch
        over all tiles and one reduction (pxh_dt_multi)."""
        st = torch.cuda.current_stream().cuda_stream
        table, n, nblocks, partial = self._dt_table()
        g = self.be.grid
        _hip.call("pxh_dt_multi", g.dtype, g.flags, table.data_ptr(), n, nblocks, self.cfl, self.dt.data_ptr(),
                  partial.data_ptr(), st)
        return self.dt

    def _sweeps(self, hyd_in, hyd_out, bed_in, old, fill_bed):
        st = torch.cuda.current_stream().cuda_stream
        dtp = self.dt.data_ptr()
        for t in self.tiles.values():
            _hip.call("pxh_roe_exner_hydro_sweep", t.grid, t.fields(hyd_in), t.fields(hyd_out), t.m0, dtp, st)
        self._fill(hyd_out, "hy")
        self.exchange(hyd_out, _HYDRO)
        out = old if old is not None else hyd_out
        for t in self.tiles.values():
            _hip.call("pxh_roe_exner_bed_sweep", t.grid, t.fields(hyd_out), t.fields(bed_in), t.fields(out),
                      t.fields(old) if old is not None else None, t.m0, self.be.sed, dtp, None, st)
        self._fill(out, fill_bed)
        self.exchange(out, _BED)

    def _fused_stage(self, src, dst, old):
        """One fused stage of every tile, src -> dst (old None) or old =
        0.5 old + 0.5 S(src) with dst holding S(src)'s hydro near the frame,
        then one exchange of every evolving field of the result."""
        st = torch.cuda.current_stream().cuda_stream
        dtp = self.dt.data_ptr()
        sed = self.be.sed
        skip = _hip.SWEEP_SKIP_G if old is None else 0
        if self.multi:
            # every tile's sweep in one launch each; the ring fills stay per tile
            tab, n, nsb, nfb, ntb, part, xs = self._sweep_table(src, dst, old)
            g = self.tiles[(0, 0)].grid
            lo = int(self.tiles[(0, 0)].a.z_b_lo is not None)
            second = int(old is not None)
            # the second stage reduces the next dt as it goes (its minima and
            # seam records, then the frame and seam edges after the frame sweep)
            pp, xp = (part.data_ptr(), xs.data_ptr()) if second else (None, None)
            _hip.call("pxh_roe_exner_stage_sweep_multi", g.dtype, g.flags, tab.data_ptr(), n, nsb, second, lo, sed,
                      dtp, skip, pp, xp, st)
            self._fill(dst, "hy")
            _hip.call("pxh_roe_exner_frame_sweep_multi", g.dtype, g.flags, tab.data_ptr(), n, nfb, second, sed, dtp,
                      skip, st)
            if second:
                _hip.call("pxh_dt_frame_multi", g.dtype, g.flags, tab.data_ptr(), n, ntb, nsb, self.cfl, pp, xp, dtp,
                          st)
            self._fill(old if old is not None else dst, "bed" if old is not None else "bed1")
            self.exchange(old if old is not None else dst, _STEP)
            return
        for t in self.tiles.values():
            fs, fd = t.fields(src), t.fields(dst)
            fo = t.fields(old) if old is not None else None
            _hip.call("pxh_roe_exner_stage_sweep", t.grid, fs, fd, fo, t.m0, sed, dtp, skip, None, st)
            if t.n_entries:
                _hip.call("pxh_bc_apply", t.grid, fd, t.programs["hy"], st)
            _hip.call("pxh_roe_exner_frame_sweep", t.grid, fd, fs, None if old is not None else fd, fo, t.m0, sed,
                      dtp, skip, self.cfl, None, st)
            if t.n_entries:
                _hip.call("pxh_bc_apply", t.grid, fo if old is not None else fd,
                          t.programs["bed" if old is not None else "bed1"], st)
        self.exchange(old if old is not None else dst, _STEP)

    def ssprk2_step(self):
        """One SSPRK2 step (integrators/SSPRK2.py:32-40) of every tile."""
        if not self.fused:
            self.compute_dt()
            self._sweeps("a", "b", "a", None, "bed1")
            self._sweeps("b", "c", "b", "a", "bed")
            return
        if not self.dt_ready:
            self.compute_dt()
        self._fused_stage("a", "b", None)
        self._fused_stage("b", "c", "a")  # with multi, leaves the next step's dt
        if not self.multi:
            self.compute_dt()  # the next step's dt, from the exchanged state
        self.dt_ready = True

    def halo_bytes_per_step(self):
        """Bytes the tiles send each other in one step (four exchanges)."""
        lib = _hip.load()
        total = 0
        for t in self.tiles.values():
            for side in _SIDES:
                if t.ring[side]:
                    if self.fused:
                        total += 2 * lib.pxh_halo_bytes(t.grid, t.fa, t.hw, _STEP, side)
                    else:
                        total += 2 * lib.pxh_halo_bytes(t.grid, t.fa, t.hw, _HYDRO, side)
                        total += 2 * lib.pxh_halo_bytes(t.grid, t.fa, t.hw, _BED, side)
        return int(total)

    def capture(self):
        """Record one step into a hipGraph (one warm-up step first)."""
        s = torch.cuda.Stream()
        s.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(s):
            self.ssprk2_step()
        torch.cuda.current_stream().wait_stream(s)
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            self.ssprk2_step()
        self.graph = g
        return g

    def replay(self):
        self.graph.replay()

    def gather(self, name):
        """The owned cells of every tile, as one global array."""
        parts = []
        for ty in range(self.par[0]):
            row = []
            for tx in range(self.par[1]):
                t = self.tiles[(ty, tx)]
                r0, r1, c0, c1 = t.owned
                e0, _, f0, _ = t.ext
                row.append(getattr(t.a, name)[r0 - e0:r1 - e0, c0 - f0:c1 - f0])
            parts.append(torch.cat(row, dim=1))
        return torch.cat(parts, dim=0)

"""Batched ensembles: many independent Roe-Exner domains advanced together.

Each member is an ordinary case (mesh, arrays, YAML params) of its own shape,
sediment and boundary layout; the members' fields are packed into one arena
per field, their ring-fill programs into one program over the arenas, and one
SSPRK2 step of every member is `pxh_ensemble_ssprk2`: one launch per sweep for
the whole ensemble, each member with its own CFL dt.  The reference runs one
domain per process; this is the batched form of the same step
(integrators/SSPRK2.py:32-40 with solvers/roe_exner_solver.py:80-122 per
member).
"""
import ctypes

import numpy as np
import torch

from . import _hip
from .boundaries import ENTRY_DTYPE
from .runtime.driver import setup_case
from .state import fields_struct

EVOLVING = ("h", "hu", "hv", "z_b", "G")


class Ensemble:
    """Members given as (mesh, arrays, params) triples, all Roe-Exner with
    frame-local rings, one dtype; `cfl` is shared (each member its own dt)."""

    def __init__(self, members, dtype=torch.float32, device="cuda", cfl=0.5, exact_f32=False, keep_members=False,
                 static=True, packed=True):
        """static: check once whether no member has a nodata cell and each
        member's z holds one value, and if so let the fused stage sweeps skip
        those streams (PXH_NO_NODATA / PXH_Z_UNIFORM; results unchanged).
        packed (float32): the first stage's output keeps h, hu, hv, z_b as one
        four-float record per cell in a 4-cell-wide b.h (PXH_PACKED_STAGE;
        results unchanged)."""
        self.n = len(members)
        # with keep_members, (backend, state, mask) of every member set up alone
        # (for comparisons against single-domain stepping)
        self.alone = [] if keep_members else None
        self.cfl = float(cfl)
        self.dtype = dtype
        states, masks, progs, handlers, meta = [], [], [], None, []
        for mesh, arrays, params in members:
            params = dict(params, backend=dict(params.get("backend") or {}, exact_f32=exact_f32))
            state, bm, solver, sc, _, _ = setup_case(mesh, arrays, params, dtype=dtype, device=device)
            mask = solver.mask_fn(state, sc.mpi_handler, bm.boundary_mask)
            state = solver.init_fn(state, mask, sc)
            if bm.n_entries and not bm.frame_local:
                raise NotImplementedError("ensemble members need frame-local rings (PXH_BC_FRAME)")
            if bm.values:
                raise NotImplementedError("ensemble members take copy and reflection boundaries only")
            if bm.n_levels > 1:
                raise NotImplementedError("ensemble members take ring fills of one dependency level")
            hd = [(h.kind, float(h.normal[0]), float(h.normal[1])) for h in bm.boundary_handlers]
            if handlers is None:
                handlers = hd
            elif hd != handlers:
                raise NotImplementedError("ensemble members must share one boundary-handler list")
            states.append(state)
            if keep_members:
                self.alone.append((sc.backend, state, mask))
            masks.append(mask[0].reshape(-1))
            progs.append(bm.entries)
            meta.append((state.h.shape, sc.backend.sed, float(sc.backend.dx), sc.backend.grid.flags))
        self.handlers = handlers or []
        self.shapes = [m[0] for m in meta]
        sizes = [int(s[0] * s[1]) for s in self.shapes]
        self.offsets = np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(np.int64)
        total = int(sum(sizes))
        if total >= 2 ** 31:
            raise ValueError("an ensemble spans at most 2^31 - 1 cells")
        names = [n for n in ("h", "hu", "hv", "z", "z_b", "n", "n_b", "G", "z_b_lo")
                 if getattr(states[0], n) is not None]
        arena = {n: torch.cat([getattr(s, n).reshape(-1) for s in states]) for n in names}
        self.a = states[0].replace(**arena)
        ev = EVOLVING + (("z_b_lo",) if "z_b_lo" in names else ())
        self.b = self.a.replace(**{n: torch.empty_like(arena[n]) for n in ev})
        self.packed = bool(packed) and dtype == torch.float32
        if self.packed:
            wide = torch.empty(4 * total, dtype=dtype, device=device)
            self.b = self.b.replace(h=wide[:total])
        self.c = self.a.replace(**{n: torch.empty_like(arena[n]) for n in ev})
        self.fa, self.fb, self.fc = (fields_struct(x) for x in (self.a, self.b, self.c))
        self.mask = torch.cat(masks).to(torch.uint8).contiguous()
        del arena
        # one ring-fill program over the arenas, ordered by field group
        ents = []
        for k, e in enumerate(progs):
            e = e.copy()
            e["target"] = (e["target"].astype(np.int64) + self.offsets[k]).astype(np.int32)
            e["src"] = (e["src"].astype(np.int64) + self.offsets[k]).astype(np.int32)
            ents.append(e)
        ent = np.concatenate(ents) if ents else np.zeros(0, dtype=ENTRY_DTYPE)
        ent = ent[np.argsort(ent["group"], kind="stable")]
        start = [int(v) for v in np.searchsorted(ent["group"].astype(np.int64), np.arange(_hip.N_GROUPS + 1))]
        self._entries = torch.from_numpy(ent.view(np.uint8).copy()).to(device) if len(ent) else None
        bc = _hip.BC()
        bc.entries = self._entries.data_ptr() if self._entries is not None else None
        bc.n_entries = len(ent)
        bc.n_handlers = len(self.handlers)
        bc.group_mask = (1 << _hip.N_GROUPS) - 1
        bc.flags = _hip.BC_FRAME | _hip.BC_GROUPED
        for k in range(_hip.N_GROUPS + 1):
            bc.group_start[k] = start[k]
        for k, (kind, nxn, nyn) in enumerate(self.handlers):
            bc.handlers[k].kind, bc.handlers[k].nx, bc.handlers[k].ny = kind, nxn, nyn
        self.bc = bc
        # the plan: member table and sediment folds, host and device copies
        self._mem = (_hip.Member * self.n)()
        for k, (shape, sed, dx, _) in enumerate(meta):
            self._mem[k].off, self._mem[k].ny, self._mem[k].nx = int(self.offsets[k]), shape[0], shape[1]
            self._mem[k].dx, self._mem[k].sed = dx, sed
        self._code = _hip.PXH_F32 if dtype == torch.float32 else _hip.PXH_F64
        self.no_nodata = not bool(self.mask.any())
        self.z_uniform = self.no_nodata and all(bool((s.z == s.z.reshape(-1)[0]).all()) for s in states)
        self._flags = int(meta[0][3]) & ~(_hip.NO_NODATA | _hip.Z_UNIFORM | _hip.PACKED_STAGE)
        if static and self.no_nodata:
            self._flags |= _hip.NO_NODATA | (_hip.Z_UNIFORM if self.z_uniform else 0)
        if self.packed:
            self._flags |= _hip.PACKED_STAGE
        self.device = device
        self.workspace = None
        self._build_plan()
        self.dt = torch.zeros(self.n, dtype=dtype, device=device)
        self.dt_ready = False
        self.graph = None

    def _build_plan(self):
        lib = _hip.load()
        size = lib.pxh_ensemble_plan(self._code, self._flags, self._mem, self.n, None, 0)
        if size < 0:
            raise ValueError(f"pxh_ensemble_plan: {lib.pxh_error_string(int(size)).decode()}")
        self._plan_host = (ctypes.c_uint8 * size)()
        got = lib.pxh_ensemble_plan(self._code, self._flags, self._mem, self.n, self._plan_host, size)
        if got != size:
            raise ValueError(f"pxh_ensemble_plan: {lib.pxh_error_string(int(got)).decode()}")
        self._plan_dev = torch.frombuffer(bytearray(self._plan_host), dtype=torch.uint8).to(self.device)
        ws = lib.pxh_ensemble_workspace_size(self._plan_host)
        if self.workspace is None or ws > self.workspace.numel():
            self.workspace = torch.empty(int(ws), dtype=torch.uint8, device=self.device)

    def set_strip_rows(self, rows):
        """Owned rows per strip of the fused stage sweep for every member
        (pxh_grid.flags PXH_STRIP_ROWS); rebuilds the plan.  A captured graph
        must be captured again."""
        self._flags = (int(self._flags) & ~(0xFF << _hip.STRIP_ROWS_SHIFT)) | (int(rows) << _hip.STRIP_ROWS_SHIFT)
        self._build_plan()
        self.graph = None

    def _stream(self):
        return torch.cuda.current_stream().cuda_stream

    def compute_dt(self):
        _hip.call("pxh_ensemble_dt", ctypes.addressof(self._plan_host), self._plan_dev.data_ptr(), self.fa,
                  self.mask.data_ptr(), self.cfl, self.dt.data_ptr(), self.workspace.data_ptr(), self._stream())
        self.dt_ready = True

    def ssprk2_step(self):
        """One SSPRK2 step of every member; afterwards self.dt holds each
        member's dt of its new state."""
        _hip.call("pxh_ensemble_ssprk2", ctypes.addressof(self._plan_host), self._plan_dev.data_ptr(), self.fa,
                  self.fb, self.fc, self.mask.data_ptr(), self.bc, self.cfl, self.dt.data_ptr(),
                  self.workspace.data_ptr(), int(self.dt_ready), self._stream())
        self.dt_ready = True

    def member(self, k, name):
        """Member k's field `name` as a (ny, nx) view of the arena."""
        ny, nx = self.shapes[k]
        o = int(self.offsets[k])
        return getattr(self.a, name)[o:o + ny * nx].view(ny, nx)

    def capture(self):
        """Record one step into a hipGraph (after one eager warm-up step)."""
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

    @property
    def cells(self):
        return int(sum(s[0] * s[1] for s in self.shapes))


def channel_members(n, seed=4, nx_range=(256, 2048), ny_range=(64, 512), dx=0.02, bumps=32):
    """The seeded channel scenarios of an uneven ensemble: per member nx, ny,
    inflow velocity U[0.2, 1.0] m/s, depth U[0.05, 0.3] m, sediment diameter
    log-uniform in [0.2, 4] mm, its own bed bumps; Transmissive in/out and
    Reflective side walls."""
    from . import cases
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n):
        nx = int(rng.integers(nx_range[0], nx_range[1] + 1))
        ny = int(rng.integers(ny_range[0], ny_range[1] + 1))
        u = float(rng.uniform(0.2, 1.0))
        h = float(rng.uniform(0.05, 0.3))
        d = float(np.exp(rng.uniform(np.log(2e-4), np.log(4e-3))))
        mesh, arr, params = cases.channel(nx=nx, ny=ny, dx=dx, n_bumps=bumps, seed=int(rng.integers(2 ** 31)), h0=h,
                                          u0=u)
        seds = {key: dict(v, diameter=d) for key, v in params["erosion"]["sediments"].items()}
        params = dict(params, erosion=dict(params["erosion"], sediments=seds))
        out.append((mesh, arr, params))
    return out

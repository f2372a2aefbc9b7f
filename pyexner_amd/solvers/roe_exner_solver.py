"""The "Roe Exner" solver bundle (solvers/roe_exner_solver.py:21-133) on the GPU.

Same names, signatures and functional semantics as the reference bundle:
``mask_fn(state, mpi_handler, b_mask)``, ``init_fn(state, mask, config)``,
``step_fn(state, time, dt, mask, config)``, ``compute_dt_fn(state, cfl, mask,
config)``.  Each returns new tensors; ``dt`` stays on the device (a 0-d tensor)
so a loop of steps never waits on the host.
"""
import numpy as np
import torch

from .. import _hip
from ..constants import g
from ..parallel import identity_halo
from ..state import dtype_code, fields_struct
from .registry import SolverBundle, SolverConfig, register_solver_bundle


def sediment_table(seds, constant, np_dtype):
    """pxh_sediments from seds (7, nsed) = [frac, d50, density, k_d, k_e, theta_c,
    porosity] (io/pnetcdf_reader.py:146-158), evaluated in the working dtype the
    way compute_G / compute_n do (roe_exner.py:64-130)."""
    sd = _hip.Sediments()
    if constant:
        sd.n_sed, sd.mpm, sd.one_minus_por, sd.nb_sed = 0, 0, 1.0, 0.0
        return sd
    s7 = np.asarray(seds, dtype=np_dtype)
    if s7.ndim != 2 or s7.shape[0] != 7:
        raise ValueError(f"seds must be (7, nsed), got {s7.shape}")
    nsed = s7.shape[1]
    if not 1 <= nsed <= _hip.MAX_SED:
        raise ValueError(f"1..{_hip.MAX_SED} sediment classes supported, got {nsed}")
    frac, d50, density, k_d, k_e, theta_c, porosity = s7
    s = density / np_dtype(1000.0)
    np_ = d50 ** np_dtype(1 / 6) / np_dtype(21.1)
    sqrt_g = np.sqrt(np_dtype(g))
    A = np_ * np_
    B = (s - np_dtype(1.0)) * d50
    Cc = np_ * (np_ * np_) * sqrt_g
    D = s - np_dtype(1.0)
    sd.n_sed, sd.mpm = nsed, 1
    sd.one_minus_por = float(np_dtype(1.0) - porosity.mean(dtype=np_dtype))
    d50m = np.mean(d50 * frac, dtype=np_dtype)
    sd.nb_sed = float(d50m ** np_dtype(1 / 6) / np_dtype(21.1))
    for k in range(nsed):
        sd.frac[k], sd.A[k], sd.B[k] = float(frac[k]), float(A[k]), float(B[k])
        sd.C[k], sd.D[k], sd.theta_c[k] = float(Cc[k]), float(D[k]), float(theta_c[k])
        sd.AB[k], sd.CD[k] = float(A[k]) / float(B[k]), float(Cc[k]) / float(D[k])
    return sd


def workspace_bytes(grid):
    """pxh_dt_workspace_size of the grid at the largest strip count any strip
    height gives (8 rows), or of its own height if that needs more."""
    lib = _hip.load()
    g8 = _hip.Grid(grid.ny, grid.nx, grid.dtype, (int(grid.flags) & ~(0xFF << _hip.STRIP_ROWS_SHIFT))
                   | (8 << _hip.STRIP_ROWS_SHIFT), grid.dx)
    a, b = lib.pxh_dt_workspace_size(grid), lib.pxh_dt_workspace_size(g8)
    return a if a < 0 else max(a, b)


class ExnerBackend:
    """Host-side description of one domain: grid, sediment table, ring-fill
    programs and the dt workspace.  Built once by config_fn."""

    STEP_GROUPS = (_hip.G_H, _hip.G_HUV, _hip.G_ZB, _hip.G_G)

    def __init__(self, shape, dtype, dx, seds, constant, boundaries, device="cuda", compensated_bed=None,
                 exact_f32=False, fused_stage=True):
        self.shape = tuple(shape)
        self.dtype = dtype
        self.np_dtype = np.float32 if dtype == torch.float32 else np.float64
        self.dx = float(dx)
        f32 = dtype == torch.float32
        # a ring fill of several dependency levels (a handler overwriting a
        # cell an earlier one read) runs the stage as the reference sequences
        # it, through the single-kernel entries (_seq_stage)
        self.sequential = boundaries is not None and boundaries.n_levels > 1
        # float32 keeps the bed as z_b + z_b_lo unless told otherwise (the
        # single-kernel Exner update of the sequential stage keeps a plain bed)
        self.compensated_bed = f32 if compensated_bed is None else (bool(compensated_bed) and f32)
        if self.sequential:
            self.compensated_bed = False
        self.exact_f32 = bool(exact_f32) and f32
        flags = _hip.EXACT_F32 if self.exact_f32 else 0
        self.grid = _hip.Grid(self.shape[0], self.shape[1], dtype_code(dtype), flags, self.dx)
        self.constant = constant
        self.sed = sediment_table(seds, constant, self.np_dtype)
        self.boundaries = boundaries
        self.bc_all = boundaries.program() if boundaries is not None else None
        self.bc_step = boundaries.program(self.STEP_GROUPS) if boundaries is not None else None
        # The step entries fuse the hydro and bed updates of a stage when the
        # program is frame-local (pxh_bc.flags); fused_stage=False keeps the
        # two sweeps with the ring fill between them.
        self.fused_stage = bool(fused_stage) and not self.sequential
        if self.bc_step is not None and not self.fused_stage:
            self.bc_step.flags &= ~_hip.BC_FRAME
        # the sequential stage's copy of the filled bed
        self._bed = torch.empty(self.shape, dtype=dtype, device=device) if self.sequential else None
        # The dt workspace is sized once for the shortest strips (8 rows: the
        # most strips and seams), so set_strip_rows never moves it and a
        # hipGraph captured before a retune still points at live memory.
        ws = workspace_bytes(self.grid)
        if ws < 0:
            raise ValueError("invalid grid")
        self.workspace = torch.empty(int(ws), dtype=torch.uint8, device=device)
        self.device = device

    def grid_for(self, no_nodata, z_uniform=False, packed=False):
        """The grid of a call: with no_nodata (the caller checked that mask0
        is zero in every cell) PXH_NO_NODATA is set, and the fused stage
        sweeps skip the mask loads; with z_uniform as well (the caller checked
        that z holds one value) PXH_Z_UNIFORM, and they read z once; with
        packed (b.h holds 4 ny nx elements) PXH_PACKED_STAGE."""
        extra = (_hip.NO_NODATA | (_hip.Z_UNIFORM if z_uniform else 0)) if no_nodata else 0
        extra |= _hip.PACKED_STAGE if packed else 0
        if not extra:
            return self.grid
        g = self.grid
        key = (int(g.flags), int(g.ny), int(g.nx), extra)
        cache = self.__dict__.setdefault("_static_grids", {})
        if key not in cache:
            cache[key] = _hip.Grid(g.ny, g.nx, g.dtype, int(g.flags) | extra, g.dx)
        return cache[key]

    def set_stage_waves(self, w1, w2):
        """The most wavefronts per SIMD of the first / second SSPRK2 stage's
        fused sweep (PXH_STAGE1_WAVES / PXH_STAGE2_WAVES; 0: no cap, registers
        set the occupancy).  Results do not depend on it.  A hipGraph
        captured earlier keeps the launch shape it was captured with."""
        w1, w2 = int(w1), int(w2)
        if not (0 <= w1 <= 8 and 0 <= w2 <= 8):
            raise ValueError(f"waves per SIMD ({w1}, {w2}) outside 0..8")
        mask = (0xF << _hip.STAGE1_WAVES_SHIFT) | (0xF << _hip.STAGE2_WAVES_SHIFT)
        self.grid.flags = (int(self.grid.flags) & ~mask) | (w1 << _hip.STAGE1_WAVES_SHIFT) \
            | (w2 << _hip.STAGE2_WAVES_SHIFT)

    def stage_waves(self):
        f = int(self.grid.flags)
        return (f >> _hip.STAGE1_WAVES_SHIFT) & 0xF, (f >> _hip.STAGE2_WAVES_SHIFT) & 0xF

    def set_strip_rows(self, rows):
        """Owned rows per strip of the fused stage sweep (pyexner_amd.autotune).
        A hipGraph captured earlier keeps the launch shape it was captured
        with; its buffers stay valid (the workspace does not move)."""
        rows = int(rows)
        if not 8 <= rows <= 128:
            raise ValueError(f"strip height {rows} is outside 8..128")
        flags = (int(self.grid.flags) & ~(0xFF << _hip.STRIP_ROWS_SHIFT)) | (rows << _hip.STRIP_ROWS_SHIFT)
        self.grid.flags = flags
        ws = _hip.load().pxh_dt_workspace_size(self.grid)
        if ws < 0 or ws > self.workspace.numel():
            raise ValueError(f"strip height {rows}: dt workspace of {ws} bytes exceeds the {self.workspace.numel()} "
                             "allocated")

    # raw entry points on preallocated buffers (no allocation) -----------------
    def init_into(self, fs, mask0_ptr, stream):
        _hip.call("pxh_roe_exner_init", self.grid, fs, mask0_ptr, self.sed, stream)
        if self.bc_all is not None and self.bc_all.n_entries:
            _hip.call("pxh_bc_apply", self.grid, fs, self.bc_all, stream)

    def step_into(self, fin, fout, mask0_ptr, dt_ptr, stream, time=0.0, states=None, no_nodata=False,
                  z_uniform=False):
        """One stage; time: the simulation time the ring fill sees (value
        boundaries whose bed moves with time).  states = (in, out) state
        objects, needed by the sequential stage.  no_nodata: mask0 is zero in
        every cell; z_uniform: z holds one value (grid_for)."""
        if self.sequential:
            self._seq_stage(states[0], states[1], mask0_ptr, dt_ptr, stream, time)
            return
        if self.bc_step is not None:
            self.bc_step.time = float(time)
        _hip.call("pxh_roe_exner_step", self.grid_for(no_nodata, z_uniform), fin, fout, mask0_ptr, self.sed,
                  self.bc_step, dt_ptr, stream)

    def _seq_stage(self, sin, sout, mask0_ptr, dt_ptr, stream, time):
        """step_fn (roe_exner_solver.py:80-116) kernel by kernel with the ring
        fill of every group after the hydro update and after the bed update,
        as the reference applies its handlers: for a program of several
        dependency levels two fills differ from one.  sout must not share
        evolving fields with sin."""
        for f in ("z_b", "G"):
            getattr(sout, f).copy_(getattr(sin, f))  # the stage state carries the input bed
        fin, fout = fields_struct(sin), fields_struct(sout)
        g = self.grid
        self.bc_all.time = float(time)
        _hip.call("pxh_roe_solve_2D", g, fin, fout, mask0_ptr, dt_ptr, stream)
        _hip.call("pxh_momentum_corrections", g, fin, fout, mask0_ptr, stream)
        _hip.call("pxh_bc_apply", g, fout, self.bc_all, stream)
        _hip.call("pxh_compute_G", g, fout, fout, mask0_ptr, self.sed, stream)
        self._bed.copy_(sout.z_b)  # exner_solve_2D reads the filled bed and writes sout's
        fb = fields_struct(sout)
        fb.z_b = self._bed.data_ptr()
        _hip.call("pxh_exner_solve_2D", g, fb, fout, mask0_ptr, dt_ptr, stream)
        _hip.call("pxh_bc_apply", g, fout, self.bc_all, stream)

    def dt_into(self, fs, mask0_ptr, cfl, dt_ptr, stream):
        _hip.call("pxh_roe_exner_dt", self.grid, fs, mask0_ptr, float(cfl), dt_ptr, self.workspace.data_ptr(),
                  stream)

    def combine_into(self, fa, fb, fo, stream):
        _hip.call("pxh_ssprk2_combine", self.grid, fa, fb, fo, stream)

    def ssprk2_into(self, fa, fb, fc, mask0_ptr, cfl, dt_ptr, stream, dt_ready=False, time=0.0, states=None,
                    no_nodata=False, z_uniform=False, packed=False):
        """dt (unless dt_ready), two stages and the combine; a is advanced in
        place and dt_ptr receives the dt of the new a.  Both stages' ring
        fills see `time`, the step's start time (SSPRK2.py:34-35).  states =
        (a, b, c), needed by the sequential stage.  packed: fb.h points at
        4 ny nx elements (PXH_PACKED_STAGE; b is then scratch whose hu, hv,
        z_b are not written)."""
        if self.sequential:
            a, b, c = states
            if not dt_ready:
                self.dt_into(fa, mask0_ptr, cfl, dt_ptr, stream)
            self._seq_stage(a, b, mask0_ptr, dt_ptr, stream, time)
            self._seq_stage(b, c, mask0_ptr, dt_ptr, stream, time)
            self.combine_into(fa, fc, fa, stream)
            self.dt_into(fa, mask0_ptr, cfl, dt_ptr, stream)
            return
        if self.bc_step is not None:
            self.bc_step.time = float(time)
        _hip.call("pxh_roe_exner_ssprk2", self.grid_for(no_nodata, z_uniform, packed), fa, fb, fc, mask0_ptr, self.sed,
                  self.bc_step, float(cfl), dt_ptr, self.workspace.data_ptr(), int(bool(dt_ready)), stream)


def _stream():
    return torch.cuda.current_stream().cuda_stream


def _dt_tensor(dt, dtype, device):
    if isinstance(dt, torch.Tensor):
        t = dt.to(device=device, dtype=dtype)
        return t.reshape(()).contiguous()
    return torch.full((), float(dt), dtype=dtype, device=device)


def get_mask(state, mpi_handler=None, b_mask=None):
    """roe_exner_solver.py:21-39 on one rank: nodata = isnan(z) plus the outer
    ring, minus boundary cells; returns bool (2, Ny, Nx) = [nodata, b_mask]."""
    z = state.z
    mask = torch.isnan(z)
    mask[0, :] = True
    mask[-1, :] = True
    mask[:, 0] = True
    mask[:, -1] = True
    if b_mask is None:
        b_mask = torch.zeros_like(mask)
    bm = torch.as_tensor(np.asarray(b_mask), device=z.device).to(torch.bool) if not isinstance(
        b_mask, torch.Tensor) else b_mask.to(device=z.device, dtype=torch.bool)
    mask = torch.where(bm, torch.zeros_like(mask), mask)
    return torch.stack([mask, bm], dim=0).contiguous()


def config_fn_roeexner(state, mpi_handler, boundaries, dx, params):
    """roe_exner_solver.py:41-56.  Backend options (not in the reference YAML):
    ``backend: {compensated_bed: bool, exact_f32: bool, fused_stage: bool,
    stepper_loop: bool, autotune: bool, check_finite: bool}``."""
    grass = params["erosion"]["grass_factor"]
    constant = not isinstance(grass, str)
    opts = params.get("backend", {}) or {}
    backend = ExnerBackend(state.h.shape, state.h.dtype, dx, state.seds, constant, boundaries,
                           device=state.h.device, compensated_bed=opts.get("compensated_bed"),
                           exact_f32=opts.get("exact_f32", False), fused_stage=opts.get("fused_stage", True))
    # the integrators' run_fn steps an ExnerStepper (preallocated buffers, the
    # fused SSPRK2 entry, hipGraph replay) unless stepper_loop is false
    backend.stepper_loop = bool(opts.get("stepper_loop", True))
    backend.autotune = opts.get("autotune")
    # the integrators' loops raise on a non-finite dt or state (per output window)
    backend.check_finite = bool(opts.get("check_finite", True))
    return SolverConfig(mpi_handler=mpi_handler, boundaries=boundaries, dx=dx, halo_exchange=identity_halo,
                        compute_G="constant" if constant else "MPM", compute_n="constant" if constant else "MPM",
                        backend=backend)


def _mask0_ptr(mask):
    m0 = mask[0]
    if m0.dtype not in (torch.bool, torch.uint8) or not m0.is_contiguous():
        raise TypeError("mask must be a contiguous bool/uint8 (2, Ny, Nx) tensor")
    return m0.data_ptr()


def init_fn_roeexner(state, mask, config):
    """roe_exner_solver.py:58-78: n_b, G, then the ring fill at t = 0."""
    s = state.clone()
    if config.backend.compensated_bed and s.z_b_lo is None:
        s.z_b_lo = torch.zeros_like(s.z_b)
    config.backend.init_into(fields_struct(s), _mask0_ptr(mask), _stream())
    return s


def _new_evolving(state):
    kw = dict(h=torch.empty_like(state.h), hu=torch.empty_like(state.hu), hv=torch.empty_like(state.hv),
              z_b=torch.empty_like(state.z_b), G=torch.empty_like(state.G))
    if state.z_b_lo is not None:
        kw["z_b_lo"] = torch.empty_like(state.z_b_lo)
    return state.replace(**kw)


def step_fn_roeexner(state, time, dt, mask, config):
    """roe_exner_solver.py:80-116: one Forward-Euler step, out of place."""
    be = config.backend
    out = _new_evolving(state)
    dtt = _dt_tensor(dt, state.h.dtype, state.h.device)
    be.step_into(fields_struct(state), fields_struct(out), _mask0_ptr(mask), dtt.data_ptr(), _stream(), time=time,
                 states=(state, out))
    return out


def compute_dt_roeexner(state, cfl, mask, config):
    """roe_exner_solver.py:118-122: max(cfl * min over active edges, 1e-6), on device."""
    dt = torch.empty((), dtype=state.h.dtype, device=state.h.device)
    config.backend.dt_into(fields_struct(state), _mask0_ptr(mask), cfl, dt.data_ptr(), _stream())
    return dt


@register_solver_bundle("Roe Exner")
def solver_roeexner():
    return SolverBundle(name="Roe Exner", config=config_fn_roeexner, mask_fn=get_mask, init_fn=init_fn_roeexner,
                        step_fn=step_fn_roeexner, compute_dt_fn=compute_dt_roeexner)

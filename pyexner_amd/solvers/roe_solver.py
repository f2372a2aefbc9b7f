"""The "Roe" solver bundle (solvers/roe_solver.py:14-71) on the GPU.

The reference bundle's init_fn/step_fn take no mask (roe_solver.py:30,40-41)
while the integrators pass one (integrators/forwardeuler.py:40,59); here the
bundle takes the Roe-Exner signatures so both integrators can drive it.
``init_fn`` is the identity (halo exchange only, roe_solver.py:30-38), so the
first step also fills the static fields z and n of the ring; later steps only
refill h, hu, hv (the static fill gives the same static fields every time).
"""
import torch

from .. import _hip
from ..parallel import identity_halo
from ..state import dtype_code, fields_struct
from .registry import SolverBundle, SolverConfig, register_solver_bundle


class RoeBackend:
    def __init__(self, shape, dtype, dx, boundaries, device="cuda"):
        self.shape = tuple(shape)
        self.dtype = dtype
        self.dx = float(dx)
        self.grid = _hip.Grid(self.shape[0], self.shape[1], dtype_code(dtype), 0, self.dx)
        self.boundaries = boundaries
        if boundaries is not None and boundaries.n_levels > 1:
            # pxh_roe_step fills once per step; a program whose handlers read
            # cells another one writes needs the reference's repeated fills
            raise NotImplementedError("plain-Roe stepping of a ring fill of several dependency levels")
        self.bc_all = boundaries.program() if boundaries is not None else None
        self.bc_step = boundaries.program((_hip.G_H, _hip.G_HUV)) if boundaries is not None else None
        self.bc_static = boundaries.program((_hip.G_Z, _hip.G_N)) if boundaries is not None else None
        ws = _hip.load().pxh_dt_workspace_size(self.grid)
        self.workspace = torch.empty(int(ws), dtype=torch.uint8, device=device)

    def step_into(self, fin, fout, dt_ptr, stream):
        _hip.call("pxh_roe_step", self.grid, fin, fout, self.bc_step, dt_ptr, stream)

    def static_fill_into(self, fs, stream):
        if self.bc_static is not None and self.bc_static.n_entries:
            _hip.call("pxh_bc_apply", self.grid, fs, self.bc_static, stream)

    def dt_into(self, fs, cfl, dt_ptr, stream):
        _hip.call("pxh_roe_dt", self.grid, fs, float(cfl), dt_ptr, self.workspace.data_ptr(), stream)


def _stream():
    return torch.cuda.current_stream().cuda_stream


def get_mask(state, *args, **kwargs):
    """roe_solver.py:14-17: ones on the interior (used by nothing: roe.py:26-27)."""
    mask = torch.zeros(state.h.shape, dtype=torch.bool, device=state.h.device)
    mask[1:-1, 1:-1] = True
    return torch.stack([~mask, torch.zeros_like(mask)], dim=0)


def config_fn_roe(state, mpi_handler, boundaries, dx, params=None):
    backend = RoeBackend(state.h.shape, state.h.dtype, dx, boundaries, device=state.h.device)
    backend.check_finite = bool(((params or {}).get("backend") or {}).get("check_finite", True))
    return SolverConfig(mpi_handler=mpi_handler, boundaries=boundaries, dx=dx, halo_exchange=identity_halo,
                        backend=backend)


def init_fn_roe(state, mask=None, config=None):
    return state


def step_fn_roe(state, time, dt, mask, config):
    """roe_solver.py:40-54: roe_solve_2D, then the ring fill."""
    be = config.backend
    static_done = getattr(state, "_static_filled", False)
    kw = dict(h=torch.empty_like(state.h), hu=torch.empty_like(state.hu), hv=torch.empty_like(state.hv))
    if not static_done:
        kw.update(z=state.z.clone(), n=state.n.clone())
    out = state.replace(**kw)
    if isinstance(dt, torch.Tensor):
        dtt = dt.to(device=state.h.device, dtype=state.h.dtype).reshape(()).contiguous()
    else:
        dtt = torch.full((), float(dt), dtype=state.h.dtype, device=state.h.device)
    fo = fields_struct(out)
    be.step_into(fields_struct(state), fo, dtt.data_ptr(), _stream())
    if not static_done:
        be.static_fill_into(fo, _stream())
    object.__setattr__(out, "_static_filled", True)
    return out


def compute_dt_roe(state, cfl, mask, config):
    """roe_solver.py:56-60: cfl * min over cells, no clamp."""
    dt = torch.empty((), dtype=state.h.dtype, device=state.h.device)
    config.backend.dt_into(fields_struct(state), cfl, dt.data_ptr(), _stream())
    return dt


@register_solver_bundle("Roe")
def solver_roe():
    return SolverBundle(name="Roe", config=config_fn_roe, mask_fn=get_mask, init_fn=init_fn_roe,
                        step_fn=step_fn_roe, compute_dt_fn=compute_dt_roe)

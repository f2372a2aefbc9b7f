"""pxh_roe_exner_ssprk2_noalloc (the dt workspace in a static device array)
against pxh_roe_exner_ssprk2 (the caller's workspace), through the C-ABI.

Bar: bit for bit (the same kernels on the same values), state and next dt,
on two calls in a row with the second on a smaller grid, so that the second
runs on the first's leftovers in the static array.  The reference entry's
workspace starts zeroed: a partial minimum read before it is written would
clamp its dt to 1e-6 and the two would differ."""
import pytest
import torch

from pyexner_amd import _hip, cases
from pyexner_amd.stepper import ExnerStepper
from tests.helpers import FIELDS_EXNER, gpu_case

pytestmark = pytest.mark.gpu


def _pair(mesh, arr, params, dtype):
    sg, mask, solver, sc = gpu_case(mesh, arr, params, dtype)
    ref = ExnerStepper(sc.backend, sg, mask, 0.5)
    ref.be.workspace.zero_()
    new = ExnerStepper(sc.backend, sg, mask, 0.5)
    return sc.backend, ref, new


def _noalloc_step(be, st, dt_ready):
    stream = torch.cuda.current_stream().cuda_stream
    be.bc_step.time = 0.0
    _hip.call("pxh_roe_exner_ssprk2_noalloc", be.grid, st.fa, st.fb, st.fc, st.m0, be.sed, be.bc_step, st.cfl,
              st.dt.data_ptr(), int(dt_ready), stream)


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_noalloc_equals_workspace_entry(cuda, dtype):
    shapes = [(150, 390), (41, 97)]  # the second call is on a smaller grid
    for k, (ny, nx) in enumerate(shapes):
        mesh, arr, params = cases.channel(nx=nx, ny=ny, n_bumps=8, seed=30 + k)
        be, ref, new = _pair(mesh, arr, params, dtype)
        for step in range(12):
            ref.ssprk2_step()
            _noalloc_step(be, new, dt_ready=step > 0)
        torch.cuda.synchronize()
        print(f"noalloc {dtype} {ny}x{nx}: dt {float(new.dt):.9e} (workspace entry {float(ref.dt):.9e})",
              flush=True)
        assert float(ref.dt) > 1e-6
        assert torch.equal(new.dt, ref.dt), (ny, nx)
        names = FIELDS_EXNER + (("z_b_lo",) if new.a.z_b_lo is not None else ())
        for f in names:
            assert torch.equal(getattr(new.a, f), getattr(ref.a, f)), (ny, nx, f)


def test_noalloc_refuses_grids_beyond_its_scratch(cuda):
    """A grid whose dt workspace exceeds PXH_NOALLOC_SCRATCH_BYTES is refused
    with PXH_E_RANGE before anything is launched (the field pointers are
    never read)."""
    lib = _hip.load()
    g = _hip.Grid(16384, 16384, _hip.PXH_F64, 0, 0.05)
    need = lib.pxh_dt_workspace_size(g)
    assert need > _hip.NOALLOC_SCRATCH_BYTES
    empty = _hip.Fields()
    st = lib.pxh_roe_exner_ssprk2_noalloc(g, empty, empty, empty, None, None, None, 0.5, None, 0,
                                          torch.cuda.current_stream().cuda_stream)
    assert st == _hip.PXH_E_RANGE
    g8 = _hip.Grid(8192, 8192, _hip.PXH_F32, 0, 0.05)
    assert lib.pxh_dt_workspace_size(g8) <= _hip.NOALLOC_SCRATCH_BYTES

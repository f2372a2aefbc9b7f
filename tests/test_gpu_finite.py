"""Failure detection in the output-window loops (integrators/loop.py
FiniteWatch, pxh_count_nonfinite): a non-finite dt raises at the step that
meets it, a NaN the dt never sees (the bed of a dry cell) is found by the
window's scan, and a clean run passes every window.  NaN inputs only: no
kernel indexes memory by a field value."""
import numpy as np
import pytest
import torch

from pyexner_amd import _hip, cases
from pyexner_amd.integrators.loop import NonFiniteState
from pyexner_amd.runtime.driver import run_case

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("stepper", [True, False])
@pytest.mark.parametrize("integrator", ["SSPRK2", "Forward Euler"])
def test_nan_momentum_raises_at_the_dt(cuda, integrator, stepper):
    mesh, arr, params = cases.channel(nx=80, ny=24, n_bumps=4, seed=3)
    hu = np.array(arr["hu"], dtype=np.float64)
    hu[12, 40] = np.nan
    arr = dict(arr, hu=hu)
    p = dict(params, integrator=integrator, end_time=0.2, out_freq=0.1, backend={"stepper_loop": stepper})
    print(f"nan hu, {integrator}, stepper={stepper}", flush=True)
    with pytest.raises(NonFiniteState, match="dt = nan"):
        run_case(mesh, arr, p, dtype=torch.float32)


@pytest.mark.parametrize("integrator", ["SSPRK2", "Forward Euler"])
def test_nan_in_a_dry_bed_found_by_the_window_scan(cuda, integrator):
    mesh, arr, params = cases.wetdry_obstacle(nx=90, ny=40)
    h = np.asarray(arr["h"])
    zb = np.array(arr["z_b"], dtype=np.float64)
    r, c = 20, 85  # dry and far from the front for the run's two windows
    assert h[r - 2:r + 3, c - 2:c + 3].max() == 0.0
    zb[r, c] = np.nan
    arr = dict(arr, z_b=zb)
    p = dict(params, integrator=integrator, end_time=0.1, out_freq=0.05)
    print(f"nan dry z_b, {integrator}", flush=True)
    with pytest.raises(NonFiniteState, match="1 cells hold NaN or infinite values .* output window 1"):
        run_case(mesh, arr, p, dtype=torch.float64)
    # the same run unwatched ends normally, with the NaN still in place
    p2 = dict(p, backend={"check_finite": False})
    s, _ = run_case(mesh, arr, p2, dtype=torch.float64)
    zb_out = s.z_b.cpu().numpy()
    assert np.isnan(zb_out[r, c]) and np.isfinite(np.delete(zb_out.ravel(), r * zb_out.shape[1] + c)).all()


def test_count_nonfinite_counts_cells(cuda):
    """Cells, not values: a cell with NaN in two fields counts once; inf
    counts; the float32 low bed part is scanned."""
    from pyexner_amd.state import RoeExnerState, fields_struct
    ny, nx = 300, 257
    rng = np.random.default_rng(3)
    host = {k: rng.uniform(0.1, 1.0, (ny, nx)) for k in ("h", "hu", "hv", "z", "z_b", "n", "n_b", "G")}
    host["h"][5, 7] = np.nan
    host["hu"][5, 7] = np.inf
    host["G"][299, 256] = -np.inf
    host["hv"][0, 0] = np.nan
    st = RoeExnerState.from_arrays(host, np.zeros((7, 1)), dtype=torch.float32, device="cuda")
    st.z_b_lo = torch.zeros_like(st.z_b)
    st.z_b_lo[100, 100] = float("nan")
    g = _hip.Grid(ny, nx, _hip.PXH_F32, 0, 0.1)
    cnt = torch.full((), 12345, dtype=torch.int64, device="cuda")
    _hip.call("pxh_count_nonfinite", g, fields_struct(st), cnt.data_ptr(), torch.cuda.current_stream().cuda_stream)
    assert int(cnt.item()) == 4

"""run_driver end to end from a YAML dict and a NetCDF-3 input, with the
output windows written by NetCDFWriter.

Bar (written here, checked below): one snapshot at t = 0 and one per output
window (SSPRK2 and Forward Euler, as the reference's run_fn write them), NaN
exactly under the nodata mask; with SSPRK2 the last snapshot equals the
returned state in float32; the run's final state equals the same case
stepped through run_case without output, bit for bit (writing snapshots does
not change the run).
"""
import numpy as np
import pytest
import torch

from pyexner_amd import cases
from pyexner_amd.output import read_snapshots
from pyexner_amd.runtime.driver import run_case, run_driver

pytestmark = pytest.mark.gpu


def _input_file(path, mesh, arr):
    from scipy.io import netcdf_file
    x = mesh.local_X[0] - mesh.dh / 2
    y = mesh.local_Y[:, 0] + mesh.dh / 2
    with netcdf_file(str(path), "w", version=2) as f:
        f.createDimension("x", len(x))
        f.createDimension("y", len(y))
        f.createVariable("x", "f8", ("x",))[:] = x
        f.createVariable("y", "f8", ("y",))[:] = y
        for k in ("h", "hu", "hv", "z", "z_b", "n"):
            f.createVariable(k, "f8", ("y", "x"))[:] = np.asarray(arr[k], dtype=np.float64)


@pytest.mark.parametrize("integrator", ["SSPRK2", "Forward Euler"])
def test_run_driver_writes_output_windows(cuda, tmp_path, integrator):
    mesh, arr, params = cases.channel(nx=80, ny=24, n_bumps=4, seed=3)
    _input_file(tmp_path / "in.nc", mesh, arr)
    p = dict(params, input_file="in.nc", output_file="out.nc", integrator=integrator, end_time=0.3, out_freq=0.1)
    (tmp_path / "case.yaml").write_text("")  # the YAML's directory anchors relative paths
    import yaml
    (tmp_path / "case.yaml").write_text(yaml.safe_dump(p))
    state, _ = run_driver(str(tmp_path / "case.yaml"), dtype=torch.float32)
    x, y, out = read_snapshots(str(tmp_path / "out.nc"))
    assert out["h"].shape == (4, 24, 80)  # t = 0, 0.1, 0.2, 0.3
    from pyexner_amd.solvers.roe_exner_solver import get_mask
    from pyexner_amd.boundaries import BoundaryManager
    bm = BoundaryManager(p, mesh.local_X, mesh.local_Y, device="cuda")
    m0 = get_mask(state, None, bm.boundary_mask)[0].cpu().numpy()
    if integrator == "SSPRK2":
        # the window-closing step is committed: the last snapshot is the final state
        for name in ("h", "hu", "hv", "z_b", "G"):
            want = np.where(m0, np.nan, getattr(state, name).cpu().numpy()).astype(np.float32)
            np.testing.assert_array_equal(out[name][-1], want)
    else:
        # Forward Euler writes a probe step to the window and commits a CFL
        # step instead (forwardeuler.py:68-117): the snapshot is the state at
        # the window time, close to (not equal to) the final state
        assert np.nanmax(np.abs(out["h"][-1] - state.h.cpu().numpy())) < 1e-3
    assert np.isnan(out["h"][0][m0]).all() and np.isfinite(out["h"][0][~m0]).all()
    alone, _ = run_case(mesh, {k: np.asarray(v, dtype=np.float64) for k, v in arr.items()}, p, dtype=torch.float32)
    for name in ("h", "hu", "hv", "z_b", "G"):
        assert torch.equal(getattr(alone, name), getattr(state, name)), name


class _Collect:
    """An io object that keeps host copies of the states it is given."""

    def __init__(self):
        self.states = []

    def write_state(self, state, mesh, mask0):
        self.states.append(state.to_host())


@pytest.mark.parametrize("integrator", ["SSPRK2", "Forward Euler"])
def test_run_fn_stepper_loop_matches_oracle_f64(cuda, integrator):
    """The integrators' run_fn on an ExnerStepper (preallocated buffers, the
    fused SSPRK2 entry with hipGraph replay, the dt read back one step behind)
    against the oracle's window loops (oracle.steps.run_ssprk2 /
    run_forward_euler), float64: the iteration count, the final time, every
    output window's state (SSPRK2: after the clipped closing step; Forward
    Euler: the probe step) and the final state within 1e-12 per field."""
    from oracle import steps as osteps
    from pyexner_amd.integrators import forwardeuler, ssprk2
    from tests.helpers import FIELDS_EXNER, compare, oracle_exner, rel_linf
    mesh, arr, params = cases.channel(nx=130, ny=40, n_bumps=6, seed=17)
    end, of = 0.25, 0.1
    p = dict(params, integrator=integrator, end_time=end, out_freq=of)
    so, prob = oracle_exner(mesh, arr, p)
    outs = []
    run = osteps.run_ssprk2 if integrator == "SSPRK2" else osteps.run_forward_euler
    so, t_o, it_o = run(prob, so, end, of, lambda t, s: outs.append((t, s)))
    io = _Collect()
    sg, _ = run_case(mesh, arr, p, dtype=torch.float64, io=io)
    mod = ssprk2 if integrator == "SSPRK2" else forwardeuler
    assert mod.last_run["path"] == "stepper"
    assert mod.last_run["iters"] == it_o, (mod.last_run, it_o)
    assert abs(mod.last_run["time"] - float(t_o)) <= 1e-12 * end
    assert len(io.states) == len(outs) + 1  # t = 0 and one per window
    for (t, s), got in zip(outs, io.states[1:]):
        errs = {f: rel_linf(getattr(got, f), getattr(s, f)) for f in ("h", "hu", "hv", "z_b")}
        assert max(errs.values()) <= 1e-12, (t, errs)
    errs = compare(sg, so, FIELDS_EXNER)
    print(f"run_fn {integrator}: {it_o} steps", " ".join(f"{k}={v:.2e}" for k, v in errs.items()), flush=True)
    assert max(errs.values()) <= 1e-12

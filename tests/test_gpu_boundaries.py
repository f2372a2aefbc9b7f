"""Value boundaries (ZeroMomentum, ConstantInflux, ConstantOutflux,
NormalFlowDepth, Berthon, SteepFall) against the oracle's restatement of the
reference handlers.

Bar (written here, checked below): in float64 the GPU ring fill equals the
handlers applied one after another bit for bit (at a nonzero time, for the
beds that move with time), and runs with these boundaries stay within 1e-12
relative L-inf of the float64 oracle per field, the boundaries seeing each
step's start time.  Float32 runs stay within the float32 bar of the float64
oracle (1e-5 relative on h, hu; hv against max|hu|; 1e-4 of max|dz_b| on the
bed).  The adjoint refuses an outflux program (PXH_E_UNSUPPORTED).
"""
import numpy as np
import pytest
import torch

from oracle import steps as osteps
from pyexner_amd import _hip, cases
from pyexner_amd.state import fields_struct
from tests.helpers import FIELDS_EXNER, compare, gpu_case, gpu_fe, gpu_ssprk2, oracle_exner

pytestmark = pytest.mark.gpu

BUILDERS = {"berthon": lambda: cases.berthon_channel(nx=90, ny=14),
            "outflux": lambda: cases.outflux_channel(nx=70, ny=22),
            "steepfall": lambda: cases.outflux_channel(nx=70, ny=22, steep_fall=True)}


@pytest.mark.parametrize("name", sorted(BUILDERS))
def test_value_fill_matches_sequential_handlers_f64(cuda, name):
    mesh, arr, params = BUILDERS[name]()
    sg, mask, solver, sc = gpu_case(mesh, arr, params, torch.float64)
    so, prob = oracle_exner(mesh, arr, params)
    rng = np.random.default_rng(3)
    host = sg.to_host()
    noise = {k: rng.uniform(-0.05, 0.05, host.h.shape) for k in FIELDS_EXNER}
    so = so.replace(**{k: getattr(so, k) + noise[k] for k in FIELDS_EXNER})
    sg = sg.replace(**{k: torch.as_tensor(getattr(so, k), device="cuda") for k in FIELDS_EXNER})
    t = 0.37
    got = sc.boundaries.apply(sg, t).to_host()
    want = prob.bcs.apply(so, t)
    for k in ("h", "hu", "hv", "z", "z_b", "n", "n_b", "G"):
        np.testing.assert_array_equal(getattr(got, k), getattr(want, k), err_msg=k)


@pytest.mark.parametrize("name,integrator", [("berthon", "Forward Euler"), ("outflux", "SSPRK2"),
                                             ("steepfall", "SSPRK2")])
def test_value_boundary_runs_f64(cuda, name, integrator):
    mesh, arr, params = BUILDERS[name]()
    so, prob = oracle_exner(mesh, arr, params)
    sg, mask, solver, sc = gpu_case(mesh, arr, params, torch.float64)
    step = gpu_fe if integrator == "Forward Euler" else gpu_ssprk2
    ostep = osteps.fe_step if integrator == "Forward Euler" else osteps.ssprk2_step
    t_o, t_g = 0.0, 0.0
    for _ in range(60):
        so, dto = ostep(prob, so, t_o)
        sg, dtg = step(sg, mask, solver, sc, prob.cfl, time=t_g)
        t_o = float(np.float64(t_o) + dto)
        t_g = t_g + float(dtg)
    errs = compare(sg, so, FIELDS_EXNER)
    print(name, integrator, "t", t_g, {k: float(f"{v:.3e}") for k, v in errs.items()})
    assert abs(t_g - t_o) <= 1e-12 * t_o
    assert max(errs.values()) <= 1e-12, errs
    assert np.isfinite(sg.to_host().h).all()


@pytest.mark.parametrize("name", ["outflux", "berthon"])
def test_value_boundary_f32_within_tolerance(cuda, name):
    mesh, arr, params = BUILDERS[name]()
    so, prob = oracle_exner(mesh, arr, params)
    z0 = so.z_b.copy()
    sg, mask, solver, sc = gpu_case(mesh, arr, params, torch.float32)
    t_o, t_g = 0.0, 0.0
    for _ in range(100):
        so, dto = osteps.ssprk2_step(prob, so, t_o)
        sg, dtg = gpu_ssprk2(sg, mask, solver, sc, prob.cfl, time=t_g)
        t_o += float(dto)
        t_g += float(dtg)
    g = sg.to_host()
    hu_scale = np.max(np.abs(so.hu))
    assert np.max(np.abs(g.h - so.h)) <= 1e-5 * np.max(np.abs(so.h))
    assert np.max(np.abs(g.hu - so.hu)) <= 1e-5 * hu_scale
    assert np.max(np.abs(g.hv - so.hv)) <= 1e-5 * hu_scale
    dz = np.max(np.abs(so.z_b - z0))
    assert np.max(np.abs(g.z_b - so.z_b)) <= 1e-4 * dz + 1e-7
    print(name, "max|dz_b|", dz)


def test_adjoint_refuses_an_outflux_program(cuda):
    mesh, arr, params = BUILDERS["outflux"]()
    sg, mask, solver, sc = gpu_case(mesh, arr, params, torch.float64)
    be = sc.backend
    bc = sc.boundaries.adjoint_program(be.STEP_GROUPS)
    ws = torch.empty(int(_hip.load().pxh_adj_workspace_size(be.grid, bc)), dtype=torch.uint8, device="cuda")
    lam = sg.clone()
    one = torch.ones(1, dtype=torch.float64, device="cuda")
    st = _hip.load().pxh_adj_roe_exner_ssprk2(be.grid, fields_struct(sg), fields_struct(sg), fields_struct(sg),
                                              mask[0].data_ptr(), be.sed, bc, 0.5, one.data_ptr(),
                                              fields_struct(lam), None, ws.data_ptr(), None)
    assert st == _hip.PXH_E_UNSUPPORTED


@pytest.mark.parametrize("integrator", ["Forward Euler", "SSPRK2"])
def test_nodata_island_f64_and_f32(cuda, integrator):
    """NaN bed cells inside the domain are nodata: the fluxes, the momentum
    corrections and the dt skip them.  float64 within 1e-12 of the oracle on
    every cell that carries data, float32 finite there and within its bar."""
    mesh, arr, params = cases.nodata_island()
    ostep = osteps.fe_step if integrator == "Forward Euler" else osteps.ssprk2_step
    gstep = gpu_fe if integrator == "Forward Euler" else gpu_ssprk2
    so, prob = oracle_exner(mesh, arr, params)
    data = ~np.isnan(np.asarray(arr["z"]))
    s64, mask, solver, sc = gpu_case(mesh, arr, params, torch.float64)
    s32, mask32, solver32, sc32 = gpu_case(mesh, arr, params, torch.float32)
    for _ in range(40):
        so, _ = ostep(prob, so)
        s64, _ = gstep(s64, mask, solver, sc, prob.cfl)
        s32, _ = gstep(s32, mask32, solver32, sc32, prob.cfl)
    g64, g32 = s64.to_host(), s32.to_host()
    for k in FIELDS_EXNER:
        o, a, b = getattr(so, k)[data], getattr(g64, k)[data], getattr(g32, k)[data]
        scale = max(np.max(np.abs(o)), 1e-300)
        assert np.max(np.abs(a - o)) <= 1e-12 * scale, k
        assert np.isfinite(b).all(), k
    assert np.max(np.abs(g32.h[data] - so.h[data])) <= 1e-5 * np.max(so.h[data])


def test_sequential_ring_fill_matches_handlers_f64(cuda):
    """The reference's Berthon convergence inputs (2D_const_flux/dh_0.05/
    input_0.05.yaml: bottom and top polygons two rows of cells thick) give a
    ring fill that reads cells it writes.  pxh_bc_apply runs it level by
    level and equals the handlers applied one after another
    (boundary_registry.py:183-187), bit for bit."""
    import torch
    from oracle import boundaries as ob
    from pyexner_amd.boundaries import BoundaryManager
    from pyexner_amd.state import RoeExnerState
    mesh, arr, params = cases.berthon_convergence(0.05)
    bm = BoundaryManager(params, mesh.local_X, mesh.local_Y)
    assert bm.n_levels > 1
    rng = np.random.default_rng(21)
    shape = mesh.local_shape
    host = {k: rng.uniform(0.1, 1.0, shape) for k in ("h", "hu", "hv", "z", "z_b", "n", "n_b", "G")}
    st = RoeExnerState.from_arrays(host, np.zeros((7, 1)), dtype=torch.float64, device="cuda")
    bm.apply_inplace(st, time=0.37)
    from oracle.state import ExnerState
    exp = ob.Boundaries("Roe Exner", params["boundaries"], mesh.local_X, mesh.local_Y).apply(
        ExnerState(seds=None, **host), 0.37)
    for k in ("h", "hu", "hv", "z", "z_b", "n", "n_b", "G"):
        np.testing.assert_array_equal(getattr(st, k).cpu().numpy(), getattr(exp, k), err_msg=k)


@pytest.mark.parametrize("integ", ["Forward Euler", "SSPRK2"])
def test_sequential_ring_fill_run_f64(cuda, integ):
    """60-odd steps of that case (Forward Euler is its own integrator)
    through run_case against the oracle's window loop: each stage runs the
    single-kernel sequence with both full fills; within 1e-12 per field."""
    import torch
    from oracle import steps as osteps
    from pyexner_amd.runtime.driver import run_case
    from tests.helpers import FIELDS_EXNER, compare, oracle_exner
    mesh, arr, params = cases.berthon_convergence(0.05, end_time=0.36, out_freq=0.12, integrator=integ)
    so, prob = oracle_exner(mesh, arr, params)
    run = osteps.run_forward_euler if integ == "Forward Euler" else osteps.run_ssprk2
    so, t_o, it_o = run(prob, so, 0.36, 0.12)
    sg, _ = run_case(mesh, arr, params, dtype=torch.float64)
    from pyexner_amd.integrators import forwardeuler, ssprk2
    last = forwardeuler.last_run if integ == "Forward Euler" else ssprk2.last_run
    assert last["iters"] == it_o and it_o >= 50
    errs = compare(sg, so, FIELDS_EXNER)
    print(f"berthon convergence {integ} x{it_o}", " ".join(f"{k}={v:.2e}" for k, v in errs.items()), flush=True)
    assert max(errs.values()) <= 1e-12


def test_sequential_ring_fill_refused_by_fused_entries(cuda):
    """The fused stage and SSPRK2 entries fill once per stage: they refuse a
    program of several dependency levels rather than fill it differently
    from the reference."""
    import torch
    from pyexner_amd import _hip
    from pyexner_amd.boundaries import BoundaryManager
    from pyexner_amd.solvers.roe_exner_solver import ExnerBackend
    from pyexner_amd.state import RoeExnerState, fields_struct
    mesh, arr, params = cases.berthon_convergence(0.05)
    bm = BoundaryManager(params, mesh.local_X, mesh.local_Y)
    be = ExnerBackend(mesh.local_shape, torch.float64, mesh.dh, None, True, bm)
    assert be.sequential and not be.fused_stage
    host = {k: np.ones(mesh.local_shape) for k in ("h", "hu", "hv", "z", "z_b", "n", "n_b", "G")}
    st = RoeExnerState.from_arrays(host, np.zeros((7, 1)), dtype=torch.float64, device="cuda")
    fa, fb, fc = (fields_struct(s) for s in (st, st.clone(), st.clone()))
    dt = torch.zeros((), dtype=torch.float64, device="cuda")
    m0 = torch.ones(mesh.local_shape, dtype=torch.uint8, device="cuda")
    rc = _hip.load().pxh_roe_exner_step(be.grid, fa, fb, m0.data_ptr(), be.sed, be.bc_all, dt.data_ptr(),
                                        torch.cuda.current_stream().cuda_stream)
    assert rc == _hip.PXH_E_UNSUPPORTED
    rc = _hip.load().pxh_roe_exner_ssprk2(be.grid, fa, fb, fc, m0.data_ptr(), be.sed, be.bc_all, 0.5, dt.data_ptr(),
                                          be.workspace.data_ptr(), 0, torch.cuda.current_stream().cuda_stream)
    assert rc == _hip.PXH_E_UNSUPPORTED

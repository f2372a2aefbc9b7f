"""GPU parity: the HIP path through the C-ABI against the float64 oracle on the
same seeded inputs.

Tolerances (written here, checked below):
* float64 build: relative L-inf per field <= 1e-12 (against max |oracle field|),
  after 1000 SSPRK2 steps on a small channel and after full runs of the other
  cases.
* float32 build (hardware arithmetic, CFL cubic in float32 with a float64
  fallback next to its double root, compensated bed):
  h, hu <= 1e-5 relative; hv <= 1e-5 of max|hu| (hv is a small cross-flow);
  z_b <= 1e-4 of max |z_b - z_b(0)|, all after 1000 SSPRK2 steps.
"""
import numpy as np
import pytest
import torch

from oracle import steps as osteps
from oracle.closed_forms import stoker_wet_dambreak
from pyexner_amd import _hip, cases
from tests.helpers import (FIELDS_EXNER, FIELDS_ROE, compare, gpu_case, gpu_fe, gpu_ssprk2, oracle_exner,
                           oracle_roe, rel_linf)

pytestmark = pytest.mark.gpu

F64_TOL = 1e-12


def _print(tag, errs):
    print(tag, " ".join(f"{k}={v:.3e}" for k, v in errs.items()), flush=True)


def test_init_and_dt_f64(cuda):
    mesh, arr, params = cases.channel(nx=96, ny=40, n_bumps=12, seed=5)
    so, prob = oracle_exner(mesh, arr, params)
    sg, mask, solver, sc = gpu_case(mesh, arr, params, torch.float64)
    errs = compare(sg, so, ("h", "hu", "hv", "z_b", "G", "n_b"))
    _print("init", errs)
    assert max(errs.values()) <= 1e-14
    dto = prob.dt(so)
    dtg = float(solver.compute_dt_fn(sg, prob.cfl, mask, sc))
    assert abs(dtg - float(dto)) <= 1e-14 * float(dto)


@pytest.mark.parametrize("at", [1e-4, 1e-2, 0.3, 1.0])
def test_dt_f32_fast_cubic_matches_f64(cuda, at):
    """The float32 path solves the CFL cubic in float32 (float64 next to the
    double root, |x| >= 0.999).  On uniform states every edge carries the same
    (ut, at, ct), so the domain dt is the one edge's: float32 dt within 5e-6 of
    the float64 dt across subcritical, critical and supercritical flow."""
    mesh, arr, params = cases.periodic_box(nx=24, ny=20)
    got = {}
    for dtype in (torch.float64, torch.float32):
        sg, mask, solver, sc = gpu_case(mesh, arr, params, dtype)
        for fr in (0.3, 0.7, 0.95, 1.0, 1.3, 2.0, 3.5):
            h = 1.0
            u = fr * np.sqrt(9.81 * h)
            G = at * h / (3.0 * u * u)  # at = G (3 u^2) / h on a uniform state
            sg.h.fill_(h)
            sg.hu.fill_(h * u)
            sg.hv.fill_(0.0)
            sg.G.fill_(G)
            got.setdefault(fr, {})[dtype] = float(solver.compute_dt_fn(sg, 0.5, mask, sc))
    for fr, d in got.items():
        e = abs(d[torch.float32] - d[torch.float64]) / d[torch.float64]
        print(f"at={at:g} Fr={fr:g} dt64={d[torch.float64]:.9e} rel(f32)={e:.2e}", flush=True)
        assert e <= 5e-6, (at, fr, e)


def test_bc_apply_matches_oracle(cuda):
    mesh, arr, params = cases.channel(nx=40, ny=24, n_bumps=4, seed=6)
    rng = np.random.default_rng(0)
    so, prob = oracle_exner(mesh, arr, params)
    for f in ("h", "hu", "hv", "z_b", "G", "n"):
        setattr(so, f, rng.uniform(-1, 1, so.h.shape))
    sg, mask, solver, sc = gpu_case(mesh, arr, params, torch.float64)
    for f in ("h", "hu", "hv", "z_b", "G", "n"):
        getattr(sg, f).copy_(torch.from_numpy(getattr(so, f)))
    out_g = sc.boundaries.apply(sg, 0.0).to_host()
    out_o = prob.bcs.apply(so)
    for f in ("h", "hu", "hv", "z_b", "G", "n"):
        assert np.array_equal(getattr(out_g, f), getattr(out_o, f)), f


@pytest.mark.parametrize("integrator", ["Forward Euler", "SSPRK2"])
def test_exner_one_step_f64(cuda, integrator):
    mesh, arr, params = cases.channel(nx=96, ny=40, n_bumps=12, seed=5)
    so, prob = oracle_exner(mesh, arr, params)
    sg, mask, solver, sc = gpu_case(mesh, arr, params, torch.float64)
    if integrator == "Forward Euler":
        so, _ = osteps.fe_step(prob, so)
        sg, _ = gpu_fe(sg, mask, solver, sc, prob.cfl)
    else:
        so, _ = osteps.ssprk2_step(prob, so)
        sg, _ = gpu_ssprk2(sg, mask, solver, sc, prob.cfl)
    errs = compare(sg, so, FIELDS_EXNER)
    _print(integrator, errs)
    assert max(errs.values()) <= 1e-14


def test_exner_ssprk2_1000_steps_f64(cuda):
    mesh, arr, params = cases.channel(nx=64, ny=32, n_bumps=10, seed=7)
    so, prob = oracle_exner(mesh, arr, params)
    sg, mask, solver, sc = gpu_case(mesh, arr, params, torch.float64)
    for _ in range(1000):
        so, _ = osteps.ssprk2_step(prob, so)
        sg, _ = gpu_ssprk2(sg, mask, solver, sc, prob.cfl)
    errs = compare(sg, so, FIELDS_EXNER)
    _print("ssprk2 x1000", errs)
    assert max(errs.values()) <= F64_TOL


@pytest.mark.parametrize("fused", [True, False])
def test_exner_ssprk2_fused_and_two_sweep_f64(cuda, fused):
    """The fused stage (one sweep per stage, the frame after the ring fill) and
    the two-sweep stage both match the oracle: 200 SSPRK2 steps of the channel
    (Transmissive + Reflective ring) and of the periodic box."""
    for mk in (lambda: cases.channel(nx=70, ny=36, n_bumps=8, seed=12),
               lambda: cases.periodic_box(nx=40, ny=34, seed=13)):
        mesh, arr, params = mk()
        params = dict(params, backend={"fused_stage": fused})
        so, prob = oracle_exner(mesh, arr, params)
        sg, mask, solver, sc = gpu_case(mesh, arr, params, torch.float64)
        assert (sc.backend.bc_step.flags & _hip.BC_FRAME) == (_hip.BC_FRAME if fused else 0)
        for _ in range(200):
            so, _ = osteps.ssprk2_step(prob, so)
            sg, _ = gpu_ssprk2(sg, mask, solver, sc, prob.cfl)
        errs = compare(sg, so, FIELDS_EXNER)
        _print(f"fused={fused} x200", errs)
        assert max(errs.values()) <= F64_TOL


def test_exner_radial_dambreak_fe_f64(cuda):
    mesh, arr, params = cases.radial_dambreak(nx=80, ny=72, dx=0.05, n_bumps=16, radius=1.0)
    so, prob = oracle_exner(mesh, arr, params)
    sg, mask, solver, sc = gpu_case(mesh, arr, params, torch.float64)
    for _ in range(300):
        so, _ = osteps.fe_step(prob, so)
        sg, _ = gpu_fe(sg, mask, solver, sc, prob.cfl)
    errs = compare(sg, so, FIELDS_EXNER)
    _print("radial fe x300", errs)
    assert max(errs.values()) <= F64_TOL


def _f32_run(params_backend, nsteps=1000, with_f32_oracle=False):
    mesh, arr, params = cases.channel(nx=96, ny=40, n_bumps=12, seed=8)
    params = dict(params, backend=params_backend)
    so, prob = oracle_exner(mesh, arr, params)
    s32 = p32 = None
    if with_f32_oracle:
        s32, p32 = oracle_exner(mesh, arr, params, dtype=np.float32, cubic_dtype=np.float64)
    zb0 = so.z_b.copy()
    sg, mask, solver, sc = gpu_case(mesh, arr, params, torch.float32)
    for _ in range(nsteps):
        so, _ = osteps.ssprk2_step(prob, so)
        if s32 is not None:
            s32, _ = osteps.ssprk2_step(p32, s32)
        sg, _ = gpu_ssprk2(sg, mask, solver, sc, prob.cfl)
    return sg.to_host(), so, s32, np.max(np.abs(so.z_b - zb0))


def test_exner_f32_within_tolerance(cuda):
    """float32 build (hardware arithmetic, compensated bed), 1000 SSPRK2 steps,
    against the float64 oracle: h, hu <= 1e-5 relative, hv <= 1e-5 of max|hu|,
    z_b <= 1e-4 of max |z_b - z_b(0)|."""
    hg, so, _, dzb = _f32_run({})
    e_h = rel_linf(hg.h, so.h)
    e_hu = rel_linf(hg.hu, so.hu)
    e_hv = np.max(np.abs(hg.hv - so.hv)) / np.max(np.abs(so.hu))
    e_zb = np.max(np.abs(hg.z_b - so.z_b)) / dzb
    print(f"f32 x1000: h={e_h:.3e} hu={e_hu:.3e} hv/|hu|={e_hv:.3e} zb/|dzb|={e_zb:.3e} (max dzb={dzb:.3e})",
          flush=True)
    assert e_h <= 1e-5 and e_hu <= 1e-5 and e_hv <= 1e-5
    assert e_zb <= 1e-4


def test_exner_f32_exact_matches_f32_oracle(cuda):
    """float32 with IEEE arithmetic and a plain float32 bed stays with the
    float32 restatement (the reference's own precision): h, hu within 1e-6;
    the bed within 1e-4 of max|dz_b| (float32 storage of the bed and the
    lambda_4 upwind switch make bed round-off the largest difference)."""
    hg, so, s32, dzb = _f32_run({"exact_f32": True, "compensated_bed": False}, nsteps=300, with_f32_oracle=True)
    e = {f: rel_linf(getattr(hg, f), getattr(s32, f)) for f in ("h", "hu")}
    e_zb = np.max(np.abs(hg.z_b - s32.z_b)) / dzb
    print(f"f32 exact vs f32 oracle x300: {e} zb/|dzb|={e_zb:.3e}", flush=True)
    assert max(e.values()) <= 1e-6
    assert e_zb <= 1e-4


def test_roe_dambreak_c1_f64(cuda):
    """Plain Roe, Forward Euler, fp64, 512 x 32, to t = 2 s: oracle parity and Stoker."""
    from pyexner_amd.runtime.driver import run_case
    from pyexner_amd.integrators import forwardeuler
    mesh, arr, params = cases.dambreak_1d()
    so, prob = oracle_roe(mesh, arr, params)
    so, t_o, it_o = osteps.run_forward_euler(prob, so, params["end_time"], params["out_freq"])
    sg, _ = run_case(mesh, arr, params, dtype=torch.float64)
    assert forwardeuler.last_run["iters"] == it_o
    errs = compare(sg, so, FIELDS_ROE)
    _print("roe c1", errs)
    assert max(errs.values()) <= F64_TOL
    # Stoker: the interior row against the closed form at the final time
    x = mesh.local_X[16, 1:-1]
    h_ex, u_ex = stoker_wet_dambreak(x, t_o, 1.0, 0.5, 25.6)
    hg = sg.to_host().h[16, 1:-1]
    l1 = np.mean(np.abs(hg - h_ex))
    print(f"stoker L1(h) = {l1:.4e} at t = {t_o:.4f}", flush=True)
    assert l1 < 5e-3
    # column-uniform data stays column-uniform
    hh = sg.to_host().h[1:-1, 1:-1]
    assert np.max(np.abs(hh - hh[0:1, :])) <= 1e-12


def test_roe_stepped_bed_f64(cuda):
    """Plain Roe over flat blocks with thin films (cases.roe_stepped_bed):
    step-face and wall thrusts and the source fix fire (roe.py:132-189);
    about 45 Forward-Euler steps through run_case against the oracle, within
    1e-12 per field, negative film depths included."""
    from pyexner_amd.runtime.driver import run_case
    from pyexner_amd.integrators import forwardeuler
    mesh, arr, params = cases.roe_stepped_bed()
    h, z = arr["h"], arr["z"]
    assert ((z[:, 1:] >= z[:, :-1]) & (h[:, :-1] + z[:, :-1] < z[:, 1:])).sum() > 50  # walls
    so, prob = oracle_roe(mesh, arr, params)
    so, t_o, it_o = osteps.run_forward_euler(prob, so, params["end_time"], params["out_freq"])
    assert np.isfinite(so.h).all() and so.h.min() < 0 and it_o >= 40
    sg, _ = run_case(mesh, arr, params, dtype=torch.float64)
    assert forwardeuler.last_run["iters"] == it_o
    errs = compare(sg, so, FIELDS_ROE)
    _print("roe stepped bed", errs)
    assert max(errs.values()) <= F64_TOL


def test_periodic_translation_invariance_f64(cuda):
    mesh, arr, params = cases.periodic_box(nx=40, ny=30)
    shift = 7

    def roll(a):
        inner = a[1:-1, 1:-1]
        out = a.copy()
        out[1:-1, 1:-1] = np.roll(inner, shift, axis=1)
        return out

    arr2 = {k: roll(v) for k, v in arr.items()}
    sa, ma, sol, sca = gpu_case(mesh, arr, params, torch.float64)
    sb, mb, _, scb = gpu_case(mesh, arr2, params, torch.float64)
    mass0 = sa.h[1:-1, 1:-1].sum().item()
    for _ in range(50):
        sa, _ = gpu_ssprk2(sa, ma, sol, sca, 0.5)
        sb, _ = gpu_ssprk2(sb, mb, sol, scb, 0.5)
    ha, hb = sa.to_host(), sb.to_host()
    for f in FIELDS_EXNER:
        assert np.array_equal(roll(getattr(ha, f))[1:-1, 1:-1], getattr(hb, f)[1:-1, 1:-1]), f
    mass1 = ha.h[1:-1, 1:-1].sum()
    assert abs(mass1 - mass0) <= 1e-12 * mass0


def test_reflective_box_mass_conservation_f64(cuda):
    mesh, arr, params = cases.radial_dambreak(nx=48, ny=48, dx=0.05, n_bumps=0, radius=0.6)
    sg, mask, solver, sc = gpu_case(mesh, arr, params, torch.float64)
    m0 = sg.h[1:-1, 1:-1].sum().item()
    for _ in range(100):
        sg, _ = gpu_fe(sg, mask, solver, sc, 0.5)
    m1 = sg.h[1:-1, 1:-1].sum().item()
    assert abs(m1 - m0) <= 1e-12 * m0


@pytest.mark.parametrize("integrator", ["Forward Euler", "SSPRK2"])
def test_wetdry_obstacle_f64(cuda, integrator):
    """Wet/dry fronts and an emergent block (momentum corrections fire, some
    through the out-of-window path), partial sweep strips in x and y."""
    mesh, arr, params = cases.wetdry_obstacle()
    so, prob = oracle_exner(mesh, arr, params)
    sg, mask, solver, sc = gpu_case(mesh, arr, params, torch.float64)
    for _ in range(120):
        if integrator == "Forward Euler":
            so, _ = osteps.fe_step(prob, so)
            sg, _ = gpu_fe(sg, mask, solver, sc, prob.cfl)
        else:
            so, _ = osteps.ssprk2_step(prob, so)
            sg, _ = gpu_ssprk2(sg, mask, solver, sc, prob.cfl)
    errs = compare(sg, so, FIELDS_EXNER)
    _print(f"wetdry {integrator}", errs)
    assert max(errs.values()) <= F64_TOL
    wet = (so.h > 1e-4).mean()
    assert 0.2 < wet < 0.95  # the front is still inside the domain


@pytest.mark.parametrize("shape", [(37, 200), (70, 150), (33, 63), (3, 125)])
def test_stepper_ssprk2_matches_oracle_f64(cuda, shape):
    """The fused SSPRK2 entry (dt, two stages, combine) on preallocated buffers."""
    from pyexner_amd.stepper import ExnerStepper
    ny, nx = shape
    mesh, arr, params = cases.channel(nx=nx, ny=ny, n_bumps=8, seed=10)
    so, prob = oracle_exner(mesh, arr, params)
    sg, mask, solver, sc = gpu_case(mesh, arr, params, torch.float64)
    st = ExnerStepper(sc.backend, sg, mask, prob.cfl)
    for _ in range(40):
        so, dto = osteps.ssprk2_step(prob, so)
        st.ssprk2_step()
    dnext = float(prob.dt(so))  # the step leaves the dt of the new state (fused reduction)
    assert abs(float(st.dt) - dnext) <= 1e-13 * dnext
    errs = compare(st.state, so, FIELDS_EXNER)
    _print(f"stepper {shape}", errs)
    assert max(errs.values()) <= F64_TOL


def test_stepper_graph_replay_matches_eager_f32(cuda):
    """hipGraph replays of the SSPRK2 step equal eager steps, replay after replay."""
    from pyexner_amd.stepper import ExnerStepper
    mesh, arr, params = cases.channel(nx=130, ny=66, n_bumps=8, seed=12)
    sg, mask, solver, sc = gpu_case(mesh, arr, params, torch.float32)
    eager = ExnerStepper(sc.backend, sg, mask, 0.5)
    graphed = ExnerStepper(sc.backend, sg, mask, 0.5)
    graphed.capture(1, "SSPRK2")  # capture warms up with one eager step
    eager.ssprk2_step()
    for k in range(5):
        graphed.replay()
        eager.ssprk2_step()
        torch.cuda.synchronize()
        for f in FIELDS_EXNER:
            assert torch.equal(getattr(graphed.state, f), getattr(eager.state, f)), (k, f)


@pytest.mark.parametrize("dtype,grass", [(torch.float32, "MPM"), (torch.float64, "MPM"), (torch.float32, 0.004)])
def test_stage_and_frame_sweeps_equal_fused_ssprk2(cuda, dtype, grass):
    """pxh_roe_exner_stage_sweep / pxh_roe_exner_frame_sweep with the ring fills
    in between, issued the way a caller that sequences the step itself issues
    them, give the state and next dt of pxh_roe_exner_ssprk2 bit for bit."""
    from pyexner_amd import _hip
    from pyexner_amd.stepper import ExnerStepper
    mesh, arr, params = cases.radial_dambreak(nx=133, ny=101, dx=0.05, n_bumps=6, radius=1.5)
    params = dict(params, erosion=dict(params["erosion"], grass_factor=grass))
    sg, mask, solver, sc = gpu_case(mesh, arr, params, dtype)
    assert sc.boundaries.frame_local
    ref = ExnerStepper(sc.backend, sg, mask, 0.5)
    seq = ExnerStepper(sc.backend, sg, mask, 0.5)
    be, m0, g, sz = sc.backend, seq.m0, sc.backend.grid, sc.backend.sed
    stream = torch.cuda.current_stream().cuda_stream
    part = be.workspace.data_ptr()
    bc_hy = sc.boundaries.program((_hip.G_H, _hip.G_HUV))
    bc_bed1 = sc.boundaries.program((_hip.G_ZB,) if sz.mpm else (_hip.G_ZB, _hip.G_G))
    bc_end = sc.boundaries.program((_hip.G_ZB, _hip.G_G))
    skip = _hip.SWEEP_SKIP_G
    dtp = seq.dt.data_ptr()
    ref.compute_dt()
    seq.compute_dt()
    for k in range(25):
        ref.ssprk2_step()
        fa, fb, fc = seq.fa, seq.fb, seq.fc
        _hip.call("pxh_roe_exner_stage_sweep", g, fa, fb, None, m0, sz, dtp, skip, None, stream)
        _hip.call("pxh_bc_apply", g, fb, bc_hy, stream)
        _hip.call("pxh_roe_exner_frame_sweep", g, fb, fa, fb, None, m0, sz, dtp, skip, 0.0, None, stream)
        _hip.call("pxh_bc_apply", g, fb, bc_bed1, stream)
        _hip.call("pxh_roe_exner_stage_sweep", g, fb, fc, fa, m0, sz, dtp, 0, part, stream)
        _hip.call("pxh_bc_apply", g, fc, bc_hy, stream)
        _hip.call("pxh_roe_exner_frame_sweep", g, fc, fb, None, fa, m0, sz, dtp, 0, 0.5, part, stream)
        _hip.call("pxh_bc_apply", g, fa, bc_end, stream)
        torch.cuda.synchronize()
        assert torch.equal(ref.dt, seq.dt), k
        for f in FIELDS_EXNER + (("z_b_lo",) if ref.state.z_b_lo is not None else ()):
            assert torch.equal(getattr(ref.state, f), getattr(seq.state, f)), (k, f)
    assert float(ref.dt) > 0


@pytest.mark.parametrize("fused", [True, False])
def test_next_dt_of_ssprk2_matches_oracle_f64(cuda, fused):
    """The dt an SSPRK2 step leaves for the next step equals compute_dt_fn of
    the new state (oracle), on the radial dam break whose fastest edges cross
    the sweeps' strip seams.  (The two-sweep combine once took the seam edges
    from `old` rows the strip below had already combined in place.)"""
    from pyexner_amd.stepper import ExnerStepper
    mesh, arr, params = cases.radial_dambreak(nx=60, ny=66, dx=0.05, n_bumps=6, radius=0.9)
    params = dict(params, backend={"fused_stage": fused})
    so, prob = oracle_exner(mesh, arr, params)
    sg, mask, solver, sc = gpu_case(mesh, arr, params, torch.float64)
    st = ExnerStepper(sc.backend, sg, mask, prob.cfl)
    for k in range(8):
        so, _ = osteps.ssprk2_step(prob, so)
        st.ssprk2_step()
        want = float(prob.dt(so))
        assert abs(float(st.dt) - want) <= 1e-13 * want, (k, float(st.dt), want)


# (fused, strip rows, pair of cells, direction): a pair of fast cells whose
# shared edge crosses a strip seam of the combine sweep that reduces the next
# dt.  Fused stage: 60-column strips of 16 rows here; two-sweep combine:
# 62-column strips of 16 rows.
SEAM_CASES = [
    (True, 16, ((31, 20), (32, 20)), "y"),
    (True, 16, ((20, 59), (20, 60)), "x"),
    (True, 24, ((47, 20), (48, 20)), "y"),
    (False, 0, ((31, 20), (32, 20)), "y"),
    (False, 0, ((20, 61), (20, 62)), "x"),
]


@pytest.mark.parametrize("fused", [True, False])
def test_float32_runs_are_reproducible(cuda, fused):
    """Two float32 runs of the same case give the same state and dt bit for
    bit: no sweep reads a value another block may or may not have written yet
    (the combine sweeps store over the step input in place).  A radial dam
    break on a grid of many strips in both directions (16-row fused strips)."""
    from pyexner_amd.stepper import ExnerStepper
    mesh, arr, params = cases.radial_dambreak(nx=700, ny=300, dx=0.05, n_bumps=24, radius=6.0)
    params = dict(params, backend={"fused_stage": fused})
    runs = []
    for _ in range(2):
        sg, mask, solver, sc = gpu_case(mesh, arr, params, torch.float32)
        sc.backend.set_strip_rows(16)
        st = ExnerStepper(sc.backend, sg, mask, params["cfl"])
        dts = []
        for _ in range(40):
            st.ssprk2_step()
            dts.append(st.dt.clone())
        torch.cuda.synchronize()
        runs.append((torch.stack(dts), {f: getattr(st.a, f).clone() for f in FIELDS_EXNER}))
    assert torch.equal(runs[0][0], runs[1][0])
    for f in FIELDS_EXNER:
        assert torch.equal(runs[0][1][f], runs[1][1][f]), f


@pytest.mark.parametrize("fused,rows,pair,axis", SEAM_CASES)
def test_next_dt_across_strip_seams_f64(cuda, fused, rows, pair, axis):
    """The next dt an SSPRK2 step leaves equals compute_dt_fn of the new state
    (oracle) when the fastest edge lies between two strips of the combine
    sweep.  A strip stores its combined cells over the step's input in place,
    so the edges between two strips are reduced after the sweep from the stored
    state; they were once left out (row seams of the fused stage) or taken from
    a neighbour's cell that might already hold its combined value."""
    from pyexner_amd.stepper import ExnerStepper
    mesh, arr, params = cases.periodic_box(nx=130, ny=70)
    arr = dict(arr)
    hu, hv = arr["hu"].copy(), arr["hv"].copy()
    for r, c in pair:
        if axis == "y":
            hv[r, c] = 2.0 * arr["h"][r, c]
        else:
            hu[r, c] = 2.0 * arr["h"][r, c]
    arr["hu"], arr["hv"] = hu, hv
    params = dict(params, backend={"fused_stage": fused})
    so, prob = oracle_exner(mesh, arr, params)
    sg, mask, solver, sc = gpu_case(mesh, arr, params, torch.float64)
    if rows:
        sc.backend.set_strip_rows(rows)
    st = ExnerStepper(sc.backend, sg, mask, prob.cfl)
    for k in range(4):
        so, _ = osteps.ssprk2_step(prob, so)
        st.ssprk2_step()
        want = float(prob.dt(so))
        assert abs(float(st.dt) - want) <= 1e-13 * want, (k, float(st.dt), want)

"""GPU parity of the throughput path: ExnerStepper / pxh_roe_exner_ssprk2 (the
fused SSPRK2 step that bench.py times) against the float64 oracle, plus the
single-kernel entries of the C-ABI.

Tolerances (written here, checked below):
* float64: relative L-inf per field <= 1e-12 (against max |oracle field|);
  the next dt the step leaves within 1e-13 relative of the oracle's.
* float32 FAST (hardware rcp/rsq/log2/exp2, float32 cubic, compensated bed),
  channel, 1000 SSPRK2 steps: h, hu <= 1e-5 relative, hv <= 1e-5 of max|hu|,
  z_b <= 1e-4 of max |z_b - z_b(0)|; each step's next dt within 2e-5
  relative of the oracle's compute_dt of the oracle state.  Radial dam break
  (160 x 144, 150 steps): h, hu, hv and dt as above; the bed within 3x of
  what float32 storage alone gives (the float64 oracle with h, hu, hv, G
  rounded to float32 after every step).
"""
import numpy as np
import pytest
import torch

from oracle import roe_exner as orx
from oracle import steps as osteps
from pyexner_amd import _hip, cases
from pyexner_amd.state import fields_struct
from pyexner_amd.stepper import ExnerStepper
from tests.helpers import FIELDS_EXNER, compare, gpu_case, oracle_exner, rel_linf

pytestmark = pytest.mark.gpu

F64_TOL = 1e-12


def _print(tag, errs):
    print(tag, " ".join(f"{k}={v:.3e}" for k, v in errs.items()), flush=True)


def _value_case(name):
    if name == "berthon":
        return cases.berthon_channel()
    if name == "outflux":
        return cases.outflux_channel()
    # SteepFall drains the channel through wet/dry cells: at 120 x 30 the
    # oracle itself turns a 1e-14 bed perturbation into 8e-8 within 60 steps,
    # so the run is checked at the size whose 60 steps stay well conditioned
    return cases.outflux_channel(nx=70, ny=22, steep_fall=True)


@pytest.mark.parametrize("name", ["berthon", "outflux", "steepfall"])
def test_stepper_value_boundaries_match_oracle_f64(cuda, name):
    """pxh_roe_exner_ssprk2 with value boundaries (a Berthon / NormalFlowDepth
    bed that moves with time, a SteepFall offset bed, ConstantInflux,
    ConstantOutflux, ZeroMomentum): the ring after a step is the average of
    the step's start ring and the second stage's filled ring
    (integrators/SSPRK2.py:40), like every other cell, and both stages fill at
    the step's start time (SSPRK2.py:34-35)."""
    mesh, arr, params = _value_case(name)
    so, prob = oracle_exner(mesh, arr, params)
    sg, mask, solver, sc = gpu_case(mesh, arr, params, torch.float64)
    st = ExnerStepper(sc.backend, sg, mask, prob.cfl)
    t = 0.0
    for k in range(60):
        so, dto = osteps.ssprk2_step(prob, so, t)
        st.ssprk2_step(time=t)
        t += float(dto)
    dnext = float(prob.dt(so))
    assert abs(float(st.dt) - dnext) <= 1e-13 * dnext
    errs = compare(st.state, so, FIELDS_EXNER)
    _print(f"stepper {name} t={t:.4f}", errs)
    assert max(errs.values()) <= F64_TOL


def test_stepper_capture_refuses_moving_bed(cuda):
    """A captured step would replay the ring fill at the time of capture."""
    mesh, arr, params = cases.berthon_channel()
    sg, mask, solver, sc = gpu_case(mesh, arr, params, torch.float64)
    st = ExnerStepper(sc.backend, sg, mask, 0.5)
    with pytest.raises(NotImplementedError):
        st.capture(1)


def _round_stored(s):
    """Float64 oracle state with h, hu, hv, G rounded to float32 (the fields a
    float32 build stores in float32; its bed is compensated)."""
    return s.replace(**{f: getattr(s, f).astype(np.float32).astype(np.float64) for f in ("h", "hu", "hv", "G")})


def _f32_stepper_run(mesh, arr, params, nsteps, dt_tol, storage_bed=False):
    so, prob = oracle_exner(mesh, arr, params)
    zb0 = so.z_b.copy()
    sr = _round_stored(so) if storage_bed else None
    sg, mask, solver, sc = gpu_case(mesh, arr, params, torch.float32)
    be = sc.backend
    assert be.compensated_bed and not be.exact_f32
    st = ExnerStepper(be, sg, mask, prob.cfl)
    worst_dt = 0.0
    for k in range(nsteps):
        so, _ = osteps.ssprk2_step(prob, so)
        if sr is not None:
            sr = _round_stored(osteps.ssprk2_step(prob, sr)[0])
        st.ssprk2_step()
        dnext = float(prob.dt(so))
        e = abs(float(st.dt) - dnext) / dnext
        worst_dt = max(worst_dt, e)
        assert e <= dt_tol, (k, float(st.dt), dnext)
    host = st.state.to_host()
    dzb = np.max(np.abs(so.z_b - zb0))
    errs = {"h": rel_linf(host.h, so.h), "hu": rel_linf(host.hu, so.hu),
            "hv/|hu|": np.max(np.abs(host.hv - so.hv)) / np.max(np.abs(so.hu)),
            "zb/|dzb|": np.max(np.abs(host.z_b + host.z_b_lo.astype(np.float64) - so.z_b)) / dzb,
            "dt": worst_dt}
    if sr is not None:
        errs["zb(storage)/|dzb|"] = np.max(np.abs(sr.z_b - so.z_b)) / dzb
    return errs, dzb


def test_stepper_f32_fast_channel_matches_oracle(cuda):
    """The kernel bench times (stage sweep MODE 1: combine + fused next dt,
    FAST float32, compensated bed) over 1000 SSPRK2 steps of a channel, every
    step's dt checked, against the float64 oracle."""
    mesh, arr, params = cases.channel(nx=96, ny=40, n_bumps=12, seed=8)
    errs, dzb = _f32_stepper_run(mesh, arr, params, 1000, 2e-5)
    _print(f"stepper f32 channel x1000 (max dzb {dzb:.3e})", errs)
    assert errs["h"] <= 1e-5 and errs["hu"] <= 1e-5 and errs["hv/|hu|"] <= 1e-5
    assert errs["zb/|dzb|"] <= 1e-4


def test_stepper_f32_fast_radial_matches_oracle(cuda):
    """The bench workload's physics (radial dam break over bumps, Reflective
    ring) at 160 x 144 for 150 SSPRK2 steps: the shock crosses the bumps.
    h, hu, hv at the float32 bar, every step's dt within 2e-5.  The erodible
    layer is lifted 5 mm so that no cell nears SED_TOL.  The bed still
    crosses the lambda_4 switch at |dz| = SED_TOL (roe_exner.py:653), where
    the upwind side can flip: the float64 oracle on float32 storage of
    h, hu, hv, G leaves the float64 oracle's bed by 2.7e-3 of max |dz_b|
    there, so the bed is held within 3x of that."""
    mesh, arr, params = cases.radial_dambreak(nx=160, ny=144, dx=0.05, n_bumps=16, seed=3, radius=2.0)
    arr = dict(arr, z_b=np.asarray(arr["z_b"], dtype=np.float64) + 0.005)
    errs, dzb = _f32_stepper_run(mesh, arr, params, 150, 2e-5, storage_bed=True)
    _print(f"stepper f32 radial x150 (max dzb {dzb:.3e})", errs)
    assert errs["h"] <= 1e-5 and errs["hu"] <= 1e-5 and errs["hv/|hu|"] <= 1e-5
    assert errs["zb/|dzb|"] <= 3 * errs["zb(storage)/|dzb|"] + 1e-6


def test_single_kernel_entries_match_oracle_f64(cuda):
    """pxh_roe_solve_2D -> pxh_momentum_corrections -> fill (h, hu, hv) ->
    pxh_compute_G -> pxh_exner_solve_2D -> fill (z_b, G): the four kernels of
    step_fn (roe_exner_solver.py:80-116) one call each, every intermediate
    against the oracle's, on the wet/dry obstacle (blocked momentum,
    dry cells) for 20 Forward-Euler steps."""
    mesh, arr, params = cases.wetdry_obstacle()
    so, prob = oracle_exner(mesh, arr, params)
    sg, mask, solver, sc = gpu_case(mesh, arr, params, torch.float64)
    be = sc.backend
    bm = sc.boundaries
    m0 = mask[0].data_ptr()
    stream = torch.cuda.current_stream().cuda_stream
    hy = bm.program((_hip.G_H, _hip.G_HUV))
    bed = bm.program((_hip.G_ZB, _hip.G_G))
    for k in range(20):
        dto = prob.dt(so)
        dt = torch.full((), float(dto), dtype=torch.float64, device=cuda)
        out = sg.replace(h=torch.empty_like(sg.h), hu=torch.empty_like(sg.hu), hv=torch.empty_like(sg.hv),
                         z_b=torch.empty_like(sg.z_b), G=torch.empty_like(sg.G))
        fi, fo = fields_struct(sg), fields_struct(out)
        _hip.call("pxh_roe_solve_2D", be.grid, fi, fo, m0, dt.data_ptr(), stream)
        o1 = orx.roe_solve_2D(so, dto, prob.dx, prob.mask)
        if k < 3:
            e = compare(out, o1, ("h", "hu", "hv"))
            assert max(e.values()) <= F64_TOL, ("roe_solve_2D", k, e)
        _hip.call("pxh_momentum_corrections", be.grid, fi, fo, m0, stream)
        o2 = orx.momentum_corrections(o1, prob.mask)
        if k < 3:
            e = compare(out, o2, ("h", "hu", "hv"))
            assert max(e.values()) <= F64_TOL, ("momentum_corrections", k, e)
        _hip.call("pxh_bc_apply", be.grid, fo, hy, stream)
        o3 = prob.bcs.apply(o2, 0.0)
        _hip.call("pxh_compute_G", be.grid, fi, fo, m0, be.sed, stream)
        Gs = orx.compute_G_mode(o3.h, o3.hu, o3.hv, o3.seds, prob.mask, o3.G, prob.constant)
        o4 = o3.replace(G=Gs)
        if k < 3:
            e = compare(out, o4, ("G",))
            assert max(e.values()) <= F64_TOL, ("compute_G", k, e)
        _hip.call("pxh_exner_solve_2D", be.grid, fi, fo, m0, dt.data_ptr(), stream)
        o5 = orx.exner_solve_2D(o4, dto, prob.dx, prob.mask)
        if k < 3:
            e = compare(out, o5, ("z_b",))
            assert max(e.values()) <= F64_TOL, ("exner_solve_2D", k, e)
        _hip.call("pxh_bc_apply", be.grid, fo, bed, stream)
        so = prob.bcs.apply(o5, 0.0)
        sg = out
    errs = compare(sg, so, FIELDS_EXNER)
    _print("single-kernel FE x20", errs)
    assert max(errs.values()) <= F64_TOL
    assert 0.2 < (so.h > 1e-4).mean() < 0.95


def test_two_sweep_entries_and_dt_finalize_equal_ssprk2(cuda):
    """pxh_roe_exner_hydro_sweep / pxh_roe_exner_bed_sweep (with the combine
    and the dt minima) / pxh_dt_finalize, sequenced by the caller with the ring
    fills between them, against the oracle's SSPRK2 step and next dt."""
    mesh, arr, params = cases.channel(nx=150, ny=70, n_bumps=8, seed=21)
    so, prob = oracle_exner(mesh, arr, params)
    sg, mask, solver, sc = gpu_case(mesh, arr, params, torch.float64)
    be, bm = sc.backend, sc.boundaries
    m0 = mask[0].data_ptr()
    stream = torch.cuda.current_stream().cuda_stream
    hy = bm.program((_hip.G_H, _hip.G_HUV))
    bedp = bm.program((_hip.G_ZB, _hip.G_G))
    ws = torch.empty(int(_hip.load().pxh_dt_workspace_size(be.grid)), dtype=torch.uint8, device=cuda)

    def fresh(s):
        return s.replace(h=torch.empty_like(s.h), hu=torch.empty_like(s.hu), hv=torch.empty_like(s.hv),
                         z_b=torch.empty_like(s.z_b), G=torch.empty_like(s.G))
    a = sg
    dt = torch.empty((), dtype=torch.float64, device=cuda)
    be.dt_into(fields_struct(a), m0, prob.cfl, dt.data_ptr(), stream)
    for k in range(30):
        so, dto = osteps.ssprk2_step(prob, so)
        assert abs(float(dt) - float(dto)) <= 1e-13 * float(dto)
        b, c = fresh(a), fresh(a)
        fa, fb, fc = fields_struct(a), fields_struct(b), fields_struct(c)
        _hip.call("pxh_roe_exner_hydro_sweep", be.grid, fa, fb, m0, dt.data_ptr(), stream)
        _hip.call("pxh_bc_apply", be.grid, fb, hy, stream)
        _hip.call("pxh_roe_exner_bed_sweep", be.grid, fb, fa, fb, None, m0, be.sed, dt.data_ptr(), None, stream)
        _hip.call("pxh_bc_apply", be.grid, fb, bedp, stream)
        _hip.call("pxh_roe_exner_hydro_sweep", be.grid, fb, fc, m0, dt.data_ptr(), stream)
        _hip.call("pxh_bc_apply", be.grid, fc, hy, stream)
        _hip.call("pxh_roe_exner_bed_sweep", be.grid, fc, fb, fa, fa, m0, be.sed, dt.data_ptr(), ws.data_ptr(),
                  stream)
        _hip.call("pxh_bc_apply", be.grid, fa, bedp, stream)
        _hip.call("pxh_dt_finalize", be.grid, ws.data_ptr(), prob.cfl, dt.data_ptr(), stream)
    dnext = float(prob.dt(so))
    assert abs(float(dt) - dnext) <= 1e-13 * dnext
    errs = compare(a, so, FIELDS_EXNER)
    _print("two-sweep entries x30", errs)
    assert max(errs.values()) <= F64_TOL


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_no_nodata_flag_changes_no_result(cuda, dtype):
    """PXH_NO_NODATA only skips loading a mask that is zero: 30 SSPRK2 steps
    of the radial dam break with and without it are equal bit for bit
    (float32 FAST with the compensated bed, and float64), the next dt too.
    A grid with nodata cells does not take the flag."""
    mesh, arr, params = cases.radial_dambreak(nx=300, ny=260, dx=0.05, n_bumps=8, radius=4.0)
    runs = []
    for nn in (True, False):
        sg, mask, solver, sc = gpu_case(mesh, arr, params, dtype)
        st = ExnerStepper(sc.backend, sg, mask, 0.5)
        assert st.no_nodata
        st.no_nodata = nn
        for _ in range(30):
            st.ssprk2_step()
        torch.cuda.synchronize()
        runs.append(st)
    a, b = runs
    for f in FIELDS_EXNER + (("z_b_lo",) if a.a.z_b_lo is not None else ()):
        assert torch.equal(getattr(a.a, f), getattr(b.a, f)), f
    assert torch.equal(a.dt, b.dt)
    mesh, arr, params = cases.nodata_island()
    sg, mask, solver, sc = gpu_case(mesh, arr, params, dtype)
    assert not ExnerStepper(sc.backend, sg, mask, 0.5).no_nodata


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_uniform_z_flag_changes_no_result(cuda, dtype):
    """PXH_Z_UNIFORM only skips streaming a fixed bed z that holds one value:
    30 SSPRK2 steps of the radial dam break raised onto z = 0.37 m, with and
    without it (both without the mask loads), are equal bit for bit, dt
    included, and equal to the same run with the mask loaded.  A varying z
    does not take the flag."""
    mesh, arr, params = cases.radial_dambreak(nx=300, ny=260, dx=0.05, n_bumps=8, radius=4.0)
    arr = dict(arr, z=np.full_like(np.asarray(arr["z"], dtype=np.float64), 0.37))
    runs = []
    for nn, zu in ((True, True), (True, False), (False, False)):
        sg, mask, solver, sc = gpu_case(mesh, arr, params, dtype)
        st = ExnerStepper(sc.backend, sg, mask, 0.5)
        assert st.no_nodata and st.z_uniform
        st.no_nodata, st.z_uniform = nn, zu
        for _ in range(30):
            st.ssprk2_step()
        st.fe_step()
        torch.cuda.synchronize()
        runs.append(st)
    for b in runs[1:]:
        for f in FIELDS_EXNER + (("z_b_lo",) if runs[0].a.z_b_lo is not None else ()):
            assert torch.equal(getattr(runs[0].a, f), getattr(b.a, f)), f
        assert torch.equal(runs[0].dt, b.dt)
    zb = dict(arr, z=np.asarray(arr["z"]) + 1e-3 * np.asarray(arr["z_b"]))
    sg, mask, solver, sc = gpu_case(mesh, zb, params, dtype)
    assert not ExnerStepper(sc.backend, sg, mask, 0.5).z_uniform


def _packed_case(name):
    if name == "radial":
        return cases.radial_dambreak(nx=300, ny=260, dx=0.05, n_bumps=8, radius=4.0)
    if name == "channel":
        return cases.channel(nx=256, ny=96, n_bumps=6)
    if name == "wetdry":
        return cases.wetdry_obstacle()
    if name == "periodic":
        return cases.periodic_box()
    if name == "nodata":
        return cases.nodata_island()
    return _value_case(name)


@pytest.mark.parametrize("name", ["radial", "channel", "wetdry", "periodic", "nodata", "berthon", "outflux",
                                  "steepfall"])
def test_packed_stage_changes_no_result(cuda, name):
    """PXH_PACKED_STAGE (float32 FAST) only changes how the first stage's
    output is laid out in b: 24 SSPRK2 steps, a Forward-Euler step (which
    swaps a and b, so b's wide h then serves as a) and 6 more SSPRK2 steps
    with and without it are equal bit for bit, dt included, over reflective,
    periodic, value (moving and offset bed, influx, outflux) rings and a
    grid with nodata cells."""
    mesh, arr, params = _packed_case(name)
    runs = []
    for pk in (True, False):
        sg, mask, solver, sc = gpu_case(mesh, arr, params, torch.float32)
        st = ExnerStepper(sc.backend, sg, mask, 0.5)
        assert st.packed and st.b.h.untyped_storage().nbytes() >= 4 * st.b.h.numel() * 4
        st.packed = pk
        for k in range(24):
            st.ssprk2_step(time=0.01 * k)
        st.fe_step()
        for k in range(6):
            st.ssprk2_step(time=0.25 + 0.01 * k)
        torch.cuda.synchronize()
        runs.append(st)
    a, b = runs
    for f in FIELDS_EXNER + (("z_b_lo",) if a.a.z_b_lo is not None else ()):
        assert torch.equal(getattr(a.a, f), getattr(b.a, f)), f
    assert torch.equal(a.dt, b.dt)


def test_packed_stage_graph_replays_equal_eager(cuda):
    """A hipGraph of 3 packed SSPRK2 steps, replayed 4 times with eager work
    and synchronisations between the replays, leaves the state and dt of 12
    eager unpacked steps bit for bit."""
    mesh, arr, params = cases.radial_dambreak(nx=300, ny=260, dx=0.05, n_bumps=8, radius=4.0)
    runs = []
    for pk in (True, False):
        sg, mask, solver, sc = gpu_case(mesh, arr, params, torch.float32)
        runs.append(ExnerStepper(sc.backend, sg, mask, 0.5))
        runs[-1].packed = pk
    g, e = runs
    g.capture(3)  # one eager warm-up step, then 3 captured steps
    e.ssprk2_step()
    for _ in range(4):
        g.replay()
        for _ in range(3):
            e.ssprk2_step()
        torch.cuda.synchronize()
        scratch = torch.empty(1 << 20, device="cuda").fill_(7.0)
        assert float(scratch.sum()) == 7.0 * (1 << 20)
        for f in FIELDS_EXNER + (("z_b_lo",) if g.a.z_b_lo is not None else ()):
            assert torch.equal(getattr(g.a, f), getattr(e.a, f)), f
        assert torch.equal(g.dt, e.dt)


def test_separate_entries_on_record_views_equal_packed_step(cuda):
    """The launches of the fused step issued one entry at a time
    (pxh_roe_exner_stage_sweep, pxh_bc_apply, pxh_roe_exner_frame_sweep) with
    the first stage's output as a record view (record_fields) leave the state
    and dt of pxh_roe_exner_ssprk2 with PXH_PACKED_STAGE bit for bit, 12 steps
    of the radial dam break (float32 FAST, compensated bed)."""
    from pyexner_amd.stepper import record_fields
    mesh, arr, params = cases.radial_dambreak(nx=300, ny=260, dx=0.05, n_bumps=8, radius=4.0)
    runs = []
    for _ in range(2):
        sg, mask, solver, sc = gpu_case(mesh, arr, params, torch.float32)
        runs.append((ExnerStepper(sc.backend, sg, mask, 0.5), sc))
    (ref, _), (st, sc) = runs
    assert ref.packed and st.packed and sc.boundaries.frame_local
    be, bm, m0 = sc.backend, sc.boundaries, st.m0
    stream = torch.cuda.current_stream().cuda_stream
    g = be.grid_for(st.no_nodata, st.z_uniform)
    G_H, G_HUV, G_ZB, G_G = _hip.G_H, _hip.G_HUV, _hip.G_ZB, _hip.G_G
    hy, bed1 = bm.program((G_H, G_HUV)), bm.program((G_ZB,) if be.sed.mpm else (G_ZB, G_G))
    end = bm.program((G_ZB, G_G))
    fa, fb, fc, dtp, part = st.fa, record_fields(st.b), st.fc, st.dt.data_ptr(), be.workspace.data_ptr()
    be.dt_into(fa, m0, 0.5, dtp, stream)
    for _ in range(12):
        ref.ssprk2_step()
        _hip.call("pxh_roe_exner_stage_sweep", g, fa, fb, None, m0, be.sed, dtp, _hip.SWEEP_SKIP_G, None, stream)
        _hip.call("pxh_bc_apply", g, fb, hy, stream)
        _hip.call("pxh_roe_exner_frame_sweep", g, fb, fa, fb, None, m0, be.sed, dtp, _hip.SWEEP_SKIP_G, 0.0, None,
                  stream)
        _hip.call("pxh_bc_apply", g, fb, bed1, stream)
        _hip.call("pxh_roe_exner_stage_sweep", g, fb, fc, fa, m0, be.sed, dtp, 0, part, stream)
        _hip.call("pxh_bc_apply", g, fc, hy, stream)
        _hip.call("pxh_roe_exner_frame_sweep", g, fc, fb, None, fa, m0, be.sed, dtp, 0, 0.5, part, stream)
        _hip.call("pxh_bc_apply", g, fa, end, stream)
    torch.cuda.synchronize()
    for f in FIELDS_EXNER + (("z_b_lo",) if st.a.z_b_lo is not None else ()):
        assert torch.equal(getattr(ref.a, f), getattr(st.a, f)), f
    assert torch.equal(ref.dt, st.dt)

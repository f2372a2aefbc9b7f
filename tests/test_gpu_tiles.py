"""Tile decomposition on one GPU: halo strips through the C-ABI and tiled runs
against the single-domain run.

Bars (written here, checked below): the strips are bit-exact copies; a tiled
SSPRK2 run (the fused path with three ghost rings, and the two-sweep path
with two) equals the single-domain run bit for bit in every owned cell and in
dt in float64, and within the float32 parity bar in float32."""
import numpy as np
import pytest
import torch

from pyexner_amd import _hip, cases
from pyexner_amd.state import RoeExnerState, dtype_code, fields_struct
from pyexner_amd.stepper import ExnerStepper
from pyexner_amd.tiles import TiledExner
from tests.helpers import FIELDS_EXNER, gpu_case

pytestmark = pytest.mark.gpu

GROUPS = ((1 << _hip.G_H) | (1 << _hip.G_HUV) | (1 << _hip.G_ZB) | (1 << _hip.G_G))
FIELD_ORDER = ("h", "hu", "hv", "z_b", "z_b_lo", "G")


def _state(ny, nx, dtype, lo, seed):
    g = torch.Generator().manual_seed(seed)
    mk = lambda: torch.rand((ny, nx), generator=g, dtype=torch.float64).to(dtype).cuda()
    kw = {n: mk() for n in ("h", "hu", "hv", "z", "z_b", "n", "n_b", "G")}
    kw["z_b_lo"] = mk() if lo else None
    kw["seds"] = torch.zeros((7, 1), dtype=dtype)
    return RoeExnerState(**kw)


def _cell(side, d, i, ny, nx, hw, recv):
    """Where element (side, depth d, position i) of a strip lives (the layout
    the header documents), written out independently of the kernel."""
    if side == _hip.HALO_W:
        return i, (hw - 1 - d) if recv else (hw + d)
    if side == _hip.HALO_E:
        return i, (nx - hw + d) if recv else (nx - hw - 1 - d)
    if side == _hip.HALO_N:
        return ((hw - 1 - d) if recv else (hw + d)), i
    return ((ny - hw + d) if recv else (ny - hw - 1 - d)), i


@pytest.mark.parametrize("dtype,lo", [(torch.float32, True), (torch.float64, False)])
@pytest.mark.parametrize("hw", [1, 2, 3])
def test_halo_strips_pack_and_unpack(cuda, dtype, lo, hw):
    ny, nx = 23, 31
    st = _state(ny, nx, dtype, lo, seed=5)
    grid = _hip.Grid(ny, nx, dtype_code(dtype), 0, 0.1)
    fs = fields_struct(st)
    names = [n for n in FIELD_ORDER if getattr(st, n) is not None]
    stream = torch.cuda.current_stream().cuda_stream
    lib = _hip.load()
    for sides in (_hip.HALO_W, _hip.HALO_N, _hip.HALO_E, _hip.HALO_S, _hip.HALO_W | _hip.HALO_E,
                  _hip.HALO_N | _hip.HALO_S):
        nbytes = lib.pxh_halo_bytes(grid, fs, hw, GROUPS, sides)
        sel = [s for s in (_hip.HALO_W, _hip.HALO_N, _hip.HALO_E, _hip.HALO_S) if sides & s]
        per = sum(hw * (ny if s in (_hip.HALO_W, _hip.HALO_E) else nx) for s in sel)
        assert nbytes == per * len(names) * st.h.element_size()
        buf = torch.empty(nbytes // st.h.element_size(), dtype=dtype, device="cuda")
        _hip.call("pxh_halo_pack", grid, fs, hw, GROUPS, sides, buf.data_ptr(), stream)
        want = []
        host = {n: getattr(st, n).cpu().numpy() for n in names}
        for n in names:
            for s in sel:
                L = ny if s in (_hip.HALO_W, _hip.HALO_E) else nx
                for d in range(hw):
                    want.extend(host[n][_cell(s, d, i, ny, nx, hw, False)] for i in range(L))
        assert np.array_equal(buf.cpu().numpy(), np.array(want, dtype=buf.cpu().numpy().dtype)), sides
        # unpack a distinct buffer: exactly the receive cells change
        src = torch.arange(buf.numel(), dtype=dtype, device="cuda") + 1000.0
        before = {n: getattr(st, n).clone() for n in names}
        _hip.call("pxh_halo_unpack", grid, fs, hw, GROUPS, sides, src.data_ptr(), stream)
        torch.cuda.synchronize()
        k = 0
        s_host = src.cpu().numpy()
        for n in names:
            exp = before[n].cpu().numpy().copy()
            for s in sel:
                L = ny if s in (_hip.HALO_W, _hip.HALO_E) else nx
                for d in range(hw):
                    for i in range(L):
                        exp[_cell(s, d, i, ny, nx, hw, True)] = s_host[k]
                        k += 1
            assert np.array_equal(getattr(st, n).cpu().numpy(), exp), (n, sides)
            getattr(st, n).copy_(before[n])


@pytest.mark.parametrize("fused", [True, False])
@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
@pytest.mark.parametrize("which,par", [("channel", (2, 2)), ("radial", (3, 2)), ("wetdry", (2, 3))])
def test_tiled_run_equals_single_domain(cuda, dtype, which, par, fused):
    """float64: every owned cell of a tiled SSPRK2 run, and every dt, equal the
    single-domain run of the same stage path bit for bit.  float32 (hardware
    math): the float32 parity bar, h, hu, G within 1e-5
    of max |field|, hv within 1e-5 of max |hu|, the compensated bed within 1e-4
    of max |z_b - z_b(0)|, dt within 1e-6."""
    if which == "channel":
        mesh, arr, params = cases.channel(nx=70, ny=40, n_bumps=8, seed=21)
    elif which == "radial":
        mesh, arr, params = cases.radial_dambreak(nx=60, ny=66, dx=0.05, n_bumps=6, radius=0.9)
    else:
        mesh, arr, params = cases.wetdry_obstacle(nx=75, ny=44)
    # the single-domain run of the same stage path: in float32 the fused and
    # two-sweep stages round differently, and a rounding difference at a
    # discontinuous switch of the reference (an SED_TOL threshold of the bed,
    # the Exner upwind side) can grow into a local bed difference beyond the
    # float32 bar between the two paths themselves
    params = dict(params, backend={"fused_stage": fused})
    sg, mask, solver, sc = gpu_case(mesh, arr, params, dtype)
    mono = ExnerStepper(sc.backend, sg, mask, 0.5)
    tiled = TiledExner(sc.backend, sg, mask, sc.boundaries, par, 0.5, fused=fused)
    assert tiled.hw == (3 if fused else 2)
    exact = dtype == torch.float64
    names = FIELDS_EXNER + (("z_b_lo",) if sg.z_b_lo is not None else ())
    zb0 = sg.z_b.double() + (sg.z_b_lo.double() if sg.z_b_lo is not None else 0)
    for k in range(60):
        if k == 0:
            mono.compute_dt()
        dt_used = mono.dt.clone()
        mono.ssprk2_step()
        tiled.ssprk2_step()
        torch.cuda.synchronize()
        # the fused path leaves the next step's dt (as the stepper does); the
        # two-sweep path the dt it used
        want_dt = mono.dt if fused else dt_used
        if exact:
            assert torch.equal(tiled.dt, want_dt), (k, float(tiled.dt), float(want_dt))
        else:
            assert abs(float(tiled.dt) - float(want_dt)) <= 1e-6 * float(want_dt), k
    if exact:
        for f in names:
            got, want = tiled.gather(f), getattr(mono.state, f)
            assert torch.equal(got, want), (f, float((got - want).abs().max()))
    else:
        for f in ("h", "hu", "hv", "G"):
            got, want = tiled.gather(f).double(), getattr(mono.state, f).double()
            ref = mono.state.hu if f == "hv" else getattr(mono.state, f)  # hv is a small cross-flow
            scale = max(float(ref.double().abs().max()), 1e-30)
            assert float((got - want).abs().max()) <= 1e-5 * scale, (f, float((got - want).abs().max()) / scale)
        bed = lambda z, lo: z.double() + (lo.double() if lo is not None else 0)  # noqa: E731
        got = bed(tiled.gather("z_b"), tiled.gather("z_b_lo") if sg.z_b_lo is not None else None)
        want = bed(mono.state.z_b, mono.state.z_b_lo)
        dz = max(float((want - zb0).abs().max()), 1e-30)
        assert float((got - want).abs().max()) <= 1e-4 * dz
    assert float(mono.state.h.sum()) > 0


@pytest.mark.parametrize("fused", [True, False])
def test_tiled_run_matches_oracle_f64(cuda, fused):
    """A 2 x 3 tiled float64 SSPRK2 run against the numpy oracle directly
    (not only against the single-domain HIP run): 40 steps of the channel,
    every field within 1e-12 of its largest magnitude."""
    from oracle import steps as osteps
    from tests.helpers import oracle_exner
    mesh, arr, params = cases.channel(nx=70, ny=40, n_bumps=8, seed=21)
    params = dict(params, backend={"fused_stage": fused})
    sg, mask, solver, sc = gpu_case(mesh, arr, params, torch.float64)
    tiled = TiledExner(sc.backend, sg, mask, sc.boundaries, (2, 3), 0.5, fused=fused)
    so, prob = oracle_exner(mesh, arr, params)
    for _ in range(40):
        tiled.ssprk2_step()
        so, _ = osteps.ssprk2_step(prob, so)
    torch.cuda.synchronize()
    for f in FIELDS_EXNER:
        got, want = tiled.gather(f).cpu().numpy(), getattr(so, f)
        scale = max(float(np.max(np.abs(want))), 1e-30)
        assert float(np.max(np.abs(got - want))) <= 1e-12 * scale, f


@pytest.mark.parametrize("fused", [True, False])
def test_tiled_graph_replays_equal_eager_steps(cuda, fused):
    """A captured tiled step replayed equals eager tiled steps, replay after replay."""
    mesh, arr, params = cases.channel(nx=90, ny=52, n_bumps=6, seed=3)
    params = dict(params, backend={"fused_stage": False})
    sg, mask, solver, sc = gpu_case(mesh, arr, params, torch.float32)
    eager = TiledExner(sc.backend, sg, mask, sc.boundaries, (2, 3), 0.5, fused=fused)
    graphed = TiledExner(sc.backend, sg, mask, sc.boundaries, (2, 3), 0.5, fused=fused)
    graphed.capture()  # one warm-up step before the capture
    eager.ssprk2_step()
    for k in range(4):
        graphed.replay()
        eager.ssprk2_step()
        torch.cuda.synchronize()
        assert torch.equal(graphed.dt, eager.dt), k
        for f in FIELDS_EXNER:
            assert torch.equal(graphed.gather(f), eager.gather(f)), (k, f)


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
@pytest.mark.parametrize("fused", [True, False])
def test_direct_exchange_equals_pack_and_unpack(cuda, dtype, fused):
    """The batched direct exchange (pxh_copy_batch, one launch per phase)
    leaves every buffer set of every tile exactly as the pack/unpack exchange
    does, for each group set the steps swap, on tiles of uneven size filled
    with seeded random values (ghost strips included)."""
    mesh, arr, params = cases.channel(nx=71, ny=45, n_bumps=8, seed=21)
    params = dict(params, backend={"fused_stage": fused})
    sg, mask, solver, sc = gpu_case(mesh, arr, params, dtype)
    made = [TiledExner(sc.backend, sg, mask, sc.boundaries, (3, 2), 0.5, fused=fused, direct=d)
            for d in (True, False)]
    g = torch.Generator().manual_seed(5)
    for which in ("a", "b", "c"):
        for groups in (GROUPS, (1 << _hip.G_H) | (1 << _hip.G_HUV), (1 << _hip.G_ZB) | (1 << _hip.G_G)):
            for pos, t in made[0].tiles.items():
                for f in FIELD_ORDER:
                    x = getattr(getattr(t, which), f)
                    if x is None:
                        continue
                    x.copy_(torch.rand(x.shape, generator=g, dtype=torch.float64).to(dtype))
                    getattr(getattr(made[1].tiles[pos], which), f).copy_(x)
            for tl in made:
                tl.exchange(which, groups)
            torch.cuda.synchronize()
            for pos, t in made[0].tiles.items():
                for f in FIELD_ORDER:
                    x = getattr(getattr(t, which), f)
                    if x is not None:
                        assert torch.equal(x, getattr(getattr(made[1].tiles[pos], which), f)), (which, groups, pos, f)


def test_tile_strip_height_changes_no_result(cuda):
    """The tiles' fused strip height (TiledExner.set_strip_rows, what the
    autotuner sets) moves work between wavefronts only: float32 tiled runs
    with 8-row and default strips are equal bit for bit."""
    mesh, arr, params = cases.radial_dambreak(nx=130, ny=140, dx=0.05, n_bumps=8, radius=2.0)
    sg, mask, solver, sc = gpu_case(mesh, arr, params, torch.float32)
    runs = []
    for rows in (None, 8):
        tl = TiledExner(sc.backend, sg, mask, sc.boundaries, (2, 2), 0.5)
        if rows:
            tl.set_strip_rows(rows)
        for _ in range(20):
            tl.ssprk2_step()
        torch.cuda.synchronize()
        runs.append((tl.dt.clone(), {f: tl.gather(f) for f in FIELDS_EXNER}))
    assert torch.equal(runs[0][0], runs[1][0])
    for f in FIELDS_EXNER:
        assert torch.equal(runs[0][1][f], runs[1][1][f]), f


@pytest.mark.parametrize("multi", [True, False])
def test_tile_static_flags_change_no_result(cuda, multi):
    """With no nodata cell and a uniform z the tiles' fused stage sweeps skip
    the mask and z streams (PXH_NO_NODATA | PXH_Z_UNIFORM in every tile's
    grid, through the multi-domain launch and through one launch per tile):
    float32 tiled runs with and without the promises are equal bit for bit."""
    mesh, arr, params = cases.radial_dambreak(nx=130, ny=140, dx=0.05, n_bumps=8, radius=2.0)
    sg, mask, solver, sc = gpu_case(mesh, arr, params, torch.float32)
    runs = []
    for static in (True, False):
        tl = TiledExner(sc.backend, sg, mask, sc.boundaries, (2, 2), 0.5, multi=multi, static=static)
        flags = {int(t.grid.flags) & (_hip.NO_NODATA | _hip.Z_UNIFORM) for t in tl.tiles.values()}
        assert flags == {_hip.NO_NODATA | _hip.Z_UNIFORM if static else 0}
        for _ in range(20):
            tl.ssprk2_step()
        torch.cuda.synchronize()
        runs.append((tl.dt.clone(), {f: tl.gather(f) for f in FIELDS_EXNER + ("z_b_lo",)}))
    assert torch.equal(runs[0][0], runs[1][0])
    for f in runs[0][1]:
        assert torch.equal(runs[0][1][f], runs[1][1][f]), f

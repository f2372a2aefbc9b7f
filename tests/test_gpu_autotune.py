"""The fused stage's launch shape (PXH_STRIP_ROWS, PXH_STAGE1_WAVES,
PXH_STAGE2_WAVES) and its autotuner.

Bar (written here, checked below): the strip height and the occupancy caps
change which wavefront computes a cell and when, not what it computes, so
every height and cap gives the default shape's state and dt bit for bit
(float64 and float32, single domain and ensemble); the tuner times every
candidate, applies the fastest, and answers later calls from its cache.
"""
import pytest
import torch

from pyexner_amd import autotune, cases
from pyexner_amd.ensemble import Ensemble, channel_members
from pyexner_amd.stepper import ExnerStepper
from tests.helpers import FIELDS_EXNER, gpu_case

pytestmark = pytest.mark.gpu


def _run(mesh, arr, params, dtype, rows, steps=8, waves=(0, 0)):
    sg, mask, solver, sc = gpu_case(mesh, arr, params, dtype)
    if rows is not None:
        sc.backend.set_strip_rows(rows)
    sc.backend.set_stage_waves(*waves)
    st = ExnerStepper(sc.backend, sg, mask, 0.5)
    for _ in range(steps):
        st.ssprk2_step()
    torch.cuda.synchronize()
    return st


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_strip_height_does_not_change_results(cuda, dtype):
    mesh, arr, params = cases.radial_dambreak(nx=190, ny=173, dx=0.05, n_bumps=8, radius=2.0)
    ref = _run(mesh, arr, params, dtype, None)
    for rows in (8, 12, 16, 24, 48, 64, 128):
        st = _run(mesh, arr, params, dtype, rows)
        assert torch.equal(st.dt, ref.dt), rows
        for f in FIELDS_EXNER:
            assert torch.equal(getattr(st.a, f), getattr(ref.a, f)), (rows, f)


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_occupancy_caps_do_not_change_results(cuda, dtype):
    mesh, arr, params = cases.radial_dambreak(nx=190, ny=173, dx=0.05, n_bumps=8, radius=2.0)
    ref = _run(mesh, arr, params, dtype, None)
    for waves in ((1, 1), (2, 3), (3, 2), (8, 8)):
        st = _run(mesh, arr, params, dtype, 16, waves=waves)
        assert st.be.stage_waves() == waves
        assert torch.equal(st.dt, ref.dt), waves
        for f in FIELDS_EXNER:
            assert torch.equal(getattr(st.a, f), getattr(ref.a, f)), (waves, f)


def test_tuner_times_candidates_and_caches(cuda):
    mesh, arr, params = cases.channel(nx=300, ny=200, n_bumps=6, seed=3)
    sg, mask, solver, sc = gpu_case(mesh, arr, params, torch.float32)
    be = sc.backend
    rows, times = autotune.tune_strip_rows(be, sg, mask, candidates=(16, 32, 48), steps=2, use_cache=False)
    print("strip rows", rows, {k: round(v, 4) for k, v in times.items()})
    heights = {k: v for k, v in times.items() if isinstance(k, int)}
    assert set(heights) == {16, 32, 48} and rows == min(heights, key=heights.get)
    caps = {k for k in times if not isinstance(k, int)}
    assert caps == {f"{rows}/w{a}{b}" for a, b in autotune.WAVE_CAPS}
    assert autotune.strip_rows_of(be.grid.flags) == rows
    kept = be.stage_waves()
    assert kept == (0, 0) or times[f"{rows}/w{kept[0]}{kept[1]}"] < times[rows]
    be.set_strip_rows(32)
    be.set_stage_waves(0, 0)
    again, t2 = autotune.tune_strip_rows(be, sg, mask, candidates=(16, 32, 48), steps=2)
    assert again == rows and t2 == {} and autotune.strip_rows_of(be.grid.flags) == rows
    assert be.stage_waves() == kept


def test_ensemble_strip_height_bitwise(cuda):
    members = channel_members(3, seed=9, nx_range=(61, 190), ny_range=(20, 75), bumps=6)
    a = Ensemble(members, dtype=torch.float32)
    b = Ensemble(members, dtype=torch.float32)
    b.set_strip_rows(16)
    for _ in range(6):
        a.ssprk2_step()
        b.ssprk2_step()
    torch.cuda.synchronize()
    assert torch.equal(a.dt, b.dt)
    for f in FIELDS_EXNER:
        assert torch.equal(getattr(a.a, f), getattr(b.a, f)), f

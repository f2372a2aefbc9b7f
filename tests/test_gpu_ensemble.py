"""Batched ensembles against the single-domain step, member by member.

Bar (written here, checked below): every member of a batched SSPRK2 run
equals the same member stepped alone with pxh_roe_exner_ssprk2 (fused stage),
state and dt bit for bit, in float64 and float32 (the batch runs the same
strip bodies on the same values); a batch given the static-field promises
(no nodata cell, uniform z) equals one streaming the mask and z, bit for bit."""
import numpy as np
import pytest
import torch

from pyexner_amd import _hip, cases
from pyexner_amd.ensemble import Ensemble, channel_members
from pyexner_amd.stepper import ExnerStepper
from tests.helpers import FIELDS_EXNER, gpu_case

pytestmark = pytest.mark.gpu


def _members():
    """Uneven members: seeded channels of several shapes, sediments and
    inflows, a strip-width edge case and one radial dam break."""
    ms = channel_members(4, seed=9, nx_range=(61, 190), ny_range=(20, 75), bumps=6)
    ms.append(cases.channel(nx=121, ny=33, dx=0.02, n_bumps=4, seed=5, h0=0.2, u0=0.8))
    return ms


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_ensemble_equals_members_stepped_alone(cuda, dtype):
    members = _members()
    ens = Ensemble(members, dtype=dtype, cfl=0.5)
    alone = []
    for mesh, arr, params in members:
        sg, mask, solver, sc = gpu_case(mesh, arr, params, dtype)
        alone.append(ExnerStepper(sc.backend, sg, mask, 0.5))
    names = FIELDS_EXNER + (("z_b_lo",) if ens.a.z_b_lo is not None else ())
    for k in range(40):
        ens.ssprk2_step()
        for st in alone:
            st.ssprk2_step()
        torch.cuda.synchronize()
        for m, st in enumerate(alone):
            assert torch.equal(ens.dt[m], st.dt), (k, m, float(ens.dt[m]), float(st.dt))
    for m, st in enumerate(alone):
        for f in names:
            got, want = ens.member(m, f), getattr(st.state, f)
            assert torch.equal(got, want), (m, f, float((got.double() - want.double()).abs().max()))
    assert len({tuple(s) for s in ens.shapes}) == len(members)


def test_ensemble_graph_replays_equal_eager_steps(cuda):
    members = _members()[:3]
    eager = Ensemble(members, dtype=torch.float32)
    graphed = Ensemble(members, dtype=torch.float32)
    graphed.capture()
    eager.ssprk2_step()
    for k in range(4):
        graphed.replay()
        eager.ssprk2_step()
        torch.cuda.synchronize()
        assert torch.equal(graphed.dt, eager.dt), k
        for f in FIELDS_EXNER:
            assert torch.equal(getattr(graphed.a, f), getattr(eager.a, f)), (k, f)


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
@pytest.mark.parametrize("uniform_z", [True, False])
def test_ensemble_static_flags_change_no_result(cuda, dtype, uniform_z):
    """The plan's static-field promises (PXH_NO_NODATA, and PXH_Z_UNIFORM when
    every member's z holds one value) let the stage sweeps skip the mask and z
    streams: 30 steps equal the run that streams them, dt and every field bit
    for bit.  uniform_z False gives one member a sloping fixed bed z, so only
    PXH_NO_NODATA applies."""
    members = _members()
    if not uniform_z:
        mesh, arr, params = members[1]
        arr = dict(arr)
        ny, nx = arr["h"].shape
        arr["z"] = np.broadcast_to(0.002 * np.arange(nx) / nx, (ny, nx)).copy()
        members[1] = (mesh, arr, params)
    runs = [Ensemble(members, dtype=dtype, cfl=0.5, static=s) for s in (True, False)]
    assert runs[0].no_nodata and runs[0].z_uniform == uniform_z
    assert runs[0]._flags & _hip.NO_NODATA and not runs[1]._flags & _hip.NO_NODATA
    assert bool(runs[0]._flags & _hip.Z_UNIFORM) == uniform_z
    names = FIELDS_EXNER + (("z_b_lo",) if runs[0].a.z_b_lo is not None else ())
    for k in range(30):
        for e in runs:
            e.ssprk2_step()
        torch.cuda.synchronize()
        assert torch.equal(runs[0].dt, runs[1].dt), k
    for f in names:
        assert torch.equal(getattr(runs[0].a, f), getattr(runs[1].a, f)), f


@pytest.mark.parametrize("static", [True, False])
def test_ensemble_packed_stage_changes_no_result(cuda, static):
    """PXH_PACKED_STAGE in the plan (float32 FAST) only changes how the first
    stage's output is laid out in b (four-float records over the arena): 30
    SSPRK2 steps with and without it are equal bit for bit, dt and every
    field, with and without the static-field promises."""
    members = _members()
    runs = [Ensemble(members, dtype=torch.float32, cfl=0.5, static=static, packed=p) for p in (True, False)]
    assert runs[0].packed and runs[0]._flags & _hip.PACKED_STAGE and not runs[1]._flags & _hip.PACKED_STAGE
    names = FIELDS_EXNER + (("z_b_lo",) if runs[0].a.z_b_lo is not None else ())
    for k in range(30):
        for e in runs:
            e.ssprk2_step()
        torch.cuda.synchronize()
        assert torch.equal(runs[0].dt, runs[1].dt), k
    for f in names:
        assert torch.equal(getattr(runs[0].a, f), getattr(runs[1].a, f)), f

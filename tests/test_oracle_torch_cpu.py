"""The torch CPU restatement (the multi-threaded CPU baseline of bench.py)
against the numpy restatement: float64 SSPRK2 steps of the channel and the
radial dam break agree to rounding, dt included."""
import numpy as np
import pytest
import torch

from oracle import steps as osteps
from oracle import torch_cpu as tc
from pyexner_amd import cases
from tests.helpers import oracle_exner


@pytest.mark.parametrize("name", ["channel", "radial", "periodic"])
def test_torch_cpu_matches_numpy_oracle_f64(name):
    if name == "channel":
        mesh, arr, params = cases.channel(nx=96, ny=40, n_bumps=4, seed=3)
    elif name == "radial":
        mesh, arr, params = cases.radial_dambreak(nx=64, ny=64, dx=0.05, n_bumps=3, radius=1.0)
    else:
        mesh, arr, params = cases.periodic_box(nx=40, ny=30)
    so, prob = oracle_exner(mesh, arr, params)
    st = tc.from_oracle_state(so, torch.float64)
    m0 = torch.as_tensor(prob.mask[0])
    fill = tc.Fill(prob.bcs)
    seds = np.asarray(so.seds, dtype=np.float64)
    for k in range(4):
        so, dto = osteps.ssprk2_step(prob, so)
        st, dtt = tc.ssprk2_step(st, m0, fill, seds, prob.constant, prob.dx, prob.cfl)
        assert abs(float(dtt) - float(dto)) <= 1e-13 * float(dto), k
    for f in ("h", "hu", "hv", "z_b", "G"):
        a, b = st[f].numpy(), getattr(so, f)
        scale = max(np.max(np.abs(b)), 1e-30)
        assert np.max(np.abs(a - b)) <= 1e-12 * scale, f
    assert np.max(np.abs(st["z_b"].numpy() - np.asarray(arr["z_b"]))) > 0  # the bed moved


def test_torch_cpu_f32_runs_close_to_f64():
    mesh, arr, params = cases.channel(nx=96, ny=40, n_bumps=4, seed=3)
    so, prob = oracle_exner(mesh, arr, params)
    m0 = torch.as_tensor(prob.mask[0])
    fill = tc.Fill(prob.bcs)
    seds = np.asarray(so.seds, dtype=np.float64)
    s64, s32 = tc.from_oracle_state(so, torch.float64), tc.from_oracle_state(so, torch.float32)
    for _ in range(10):
        s64, _ = tc.ssprk2_step(s64, m0, fill, seds, prob.constant, prob.dx, prob.cfl)
        s32, _ = tc.ssprk2_step(s32, m0, fill, seds, prob.constant, prob.dx, prob.cfl)
    for f in ("h", "hu", "hv"):
        a, b = s32[f].double().numpy(), s64[f].numpy()
        assert np.max(np.abs(a - b)) <= 1e-5 * np.max(np.abs(b)), f

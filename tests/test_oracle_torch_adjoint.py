"""The autograd twin of the oracle (oracle/torch_adjoint.py), pinned on the
CPU before it serves as the reference gradient of the GPU adjoint.

Bars (written here): its forward run equals the numpy oracle's init_fn and
SSPRK2 steps within 1e-12 relative per field (float64); its gradients with
respect to z_b0 (4 seeded directions), d50 and porosity equal central
differences of its own forward within 1e-6 relative (the differences'
truncation and round-off, not the gradient, set that bar).
"""
import numpy as np
import torch

from oracle import boundaries as ob
from oracle import steps as osteps
from oracle import torch_adjoint as ta
from oracle import torch_cpu as tc
from pyexner_amd import cases
from pyexner_amd.inputs import initial_G, seds_from_params
from tests.helpers import oracle_exner


def twin_setup(nx, ny):
    """The pre-init state of the calibration channel as float64 CPU tensors,
    its nodata mask, ring fill and sediment table."""
    mesh, arr, params = cases.calibration_channel(nx=nx, ny=ny)
    raw = {k: torch.as_tensor(np.asarray(arr[k], dtype=np.float64)) for k in ("h", "hu", "hv", "z", "z_b", "n")}
    raw["n_b"] = raw["n"].clone()
    raw["G"] = torch.as_tensor(np.asarray(initial_G(params, raw["h"].shape), dtype=np.float64))
    bcs = ob.Boundaries("Roe Exner", params["boundaries"], mesh.local_X, mesh.local_Y)
    mask = osteps.get_mask(np.asarray(arr["z"], dtype=np.float64), bcs.boundary_mask)
    return mesh, arr, params, raw, torch.as_tensor(mask[0]), tc.Fill(bcs), seds_from_params(params)


def test_twin_forward_is_the_oracle_run():
    mesh, arr, params, raw, m0, fill, seds = twin_setup(48, 24)
    so, prob = oracle_exner(mesh, arr, params)
    for _ in range(6):
        so, _ = osteps.ssprk2_step(prob, so)
    with torch.no_grad():
        st = ta.run(raw, m0, fill, seds, mesh.dh, params["cfl"], 6, raw["z_b"].clone(),
                    torch.tensor(seds[1, 0], dtype=torch.float64), torch.tensor(seds[6, 0], dtype=torch.float64))
    for f in ("h", "hu", "hv", "z_b", "G", "n_b"):
        a, b = st[f].numpy(), getattr(so, f)
        assert np.max(np.abs(a - b)) <= 1e-12 * np.max(np.abs(b)), f


def test_twin_gradient_matches_its_central_differences():
    mesh, arr, params, raw, m0, fill, seds = twin_setup(48, 24)
    n, cfl, dx = 8, params["cfl"], mesh.dh
    zb0 = raw["z_b"].clone()
    rng = np.random.default_rng(11)
    target = zb0 + torch.as_tensor(rng.uniform(-1e-3, 1e-3, zb0.shape))
    wh = torch.as_tensor(rng.uniform(-1e-2, 1e-2, zb0.shape))
    d50_0, por_0 = float(seds[1, 0]), float(seds[6, 0])

    def loss_fn(s):
        return 0.5 * ((s["z_b"] - target) ** 2).sum() + (wh * s["h"]).sum()

    _, gz, gd, gp = ta.gradients(raw, m0, fill, seds, dx, cfl, n, zb0, d50_0, por_0, loss_fn)
    assert torch.isfinite(gz).all()

    def fwd(zb, d, p):
        with torch.no_grad():
            return ta.run(raw, m0, fill, seds, dx, cfl, n, zb, torch.tensor(d, dtype=torch.float64),
                          torch.tensor(p, dtype=torch.float64))

    def diff(sp, sm, eps):
        # the difference taken cell by cell before the sums, so that the
        # loss's own round-off stays out of it
        zp, zm = sp["z_b"], sm["z_b"]
        return float((0.5 * (zp - zm) * (zp + zm - 2 * target)).sum() + (wh * (sp["h"] - sm["h"])).sum()) / (2 * eps)

    scale = float(gz.abs().sum()) * 1e-3
    for k in range(4):
        v = torch.as_tensor(rng.uniform(-1e-3, 1e-3, zb0.shape))
        ad = float((gz * v).sum())
        eps = 1e-6
        fd = diff(fwd(zb0 + eps * v, d50_0, por_0), fwd(zb0 - eps * v, d50_0, por_0), eps)
        assert abs(ad - fd) <= 1e-6 * max(abs(fd), scale), (k, ad, fd)
    eps = 1e-4 * d50_0
    fd = diff(fwd(zb0, d50_0 + eps, por_0), fwd(zb0, d50_0 - eps, por_0), eps)
    assert abs(gd - fd) <= 1e-6 * abs(fd), (gd, fd)
    eps = 1e-4
    fd = diff(fwd(zb0, d50_0, por_0 + eps), fwd(zb0, d50_0, por_0 - eps), eps)
    assert abs(gp - fd) <= 1e-6 * abs(fd), (gp, fd)

"""Adjoint kernels against central finite differences of the forward path.

Bar (written here, checked below): in float64, each adjoint directional
derivative agrees with the central difference of the forward entry point
along the same seeded direction to 1e-6 relative (the relative error is taken
against the larger of |FD| and 1e-3 of the gradient's scale along that field,
so a near-zero directional derivative is judged on an absolute scale).  The
differences use steps small enough that no branch switches, on the
calibration channel, where every cell is wet, carries sediment and differs in
bed from its neighbours by far more than SED_TOL (the forward step is smooth
there almost everywhere).

Nothing here has a reference vector: the reference has no gradient code, so
the derivatives are pinned by the forward path itself through central
differences (what jax.grad of the reference would have to agree with), and by
reverse-mode torch autograd of the float64 whole-array restatement
(oracle/torch_adjoint.py, pinned on the CPU by tests/test_oracle_torch_adjoint.py),
the stand-in for jax.grad: the run's gradients agree with it within 1e-12
(the z_b0 field relative to its largest entry; d50 and porosity relative).
"""
import numpy as np
import pytest
import torch

from pyexner_amd import _hip, cases
from pyexner_amd.adjoint import ExnerGradientRun, seeded_target
from pyexner_amd.state import fields_struct
from pyexner_amd.stepper import ExnerStepper
from tests.helpers import gpu_case

pytestmark = pytest.mark.gpu

TOL = 1e-6
HYDRO = ("h", "hu", "hv", "z_b", "G")


def _rel(ad, fd, scale):
    return abs(ad - fd) / max(abs(fd), 1e-3 * scale)


def _stream():
    return torch.cuda.current_stream().cuda_stream


@pytest.fixture(scope="module")
def setup(cuda):
    mesh, arr, params = cases.calibration_channel(nx=96, ny=40)
    sg, mask, solver, sc = gpu_case(mesh, arr, params, torch.float64)
    be = sc.backend
    stp = ExnerStepper(be, sg, mask, 0.5)
    for _ in range(5):
        stp.ssprk2_step()
    s = stp.a.clone()
    bc = sc.boundaries.adjoint_program(be.STEP_GROUPS)
    ws = torch.empty(int(_hip.load().pxh_adj_workspace_size(be.grid, bc)), dtype=torch.uint8, device="cuda")
    return mesh, arr, params, be, mask, s, ws, bc


def _gen(seed):
    return torch.Generator(device="cpu").manual_seed(seed)


def _rnd(shape, g):
    return torch.randn(shape, generator=g, dtype=torch.float64).cuda()


def test_stage_adjoint_matches_central_differences(setup):
    _, _, _, be, mask, s, ws, bc = setup
    m0 = mask[0].data_ptr()
    dt = torch.empty((), dtype=torch.float64, device="cuda")
    be.dt_into(fields_struct(s), m0, 0.5, dt.data_ptr(), _stream())
    g = _gen(11)
    W = {k: _rnd(s.h.shape, g) for k in HYDRO}

    def loss(sin, dtv):
        out = sin.replace(**{k: torch.empty_like(getattr(sin, k)) for k in HYDRO})
        be.step_into(fields_struct(sin), fields_struct(out), m0, dtv.data_ptr(), _stream())
        return sum(float((W[k] * getattr(out, k)).sum()) for k in W), out

    _, out = loss(s, dt)
    z = lambda: torch.zeros_like(s.h)  # noqa: E731
    lo = s.replace(**{k: W[k].clone() for k in W})
    li = s.replace(h=z(), hu=z(), hv=z(), z_b=z(), G=z(), n_b=z())
    lam_dt = torch.zeros(1, dtype=torch.float64, device="cuda")
    _hip.call("pxh_adj_roe_exner_stage", be.grid, fields_struct(s), fields_struct(out), m0, be.sed, bc,
              dt.data_ptr(), fields_struct(lo), fields_struct(li), lam_dt.data_ptr(), None, ws.data_ptr(),
              _stream())
    torch.cuda.synchronize()
    for k in ("h", "hu", "hv", "z_b", "n_b"):
        v = _rnd(s.h.shape, g) * 1e-3
        ad = float((getattr(li, k) * v).sum())
        # the friction term is smooth in n_b and small: a larger step keeps
        # round-off of the loss out of the difference
        eps = 1e-2 if k == "n_b" else 1e-4
        sp, sm = s.clone(), s.clone()
        setattr(sp, k, getattr(s, k) + eps * v)
        setattr(sm, k, getattr(s, k) - eps * v)
        fd = (loss(sp, dt)[0] - loss(sm, dt)[0]) / (2 * eps)
        scale = float(getattr(li, k).abs().sum() * 1e-3)
        print(f"stage d/d{k}: adjoint {ad:+.12e} FD {fd:+.12e}")
        assert _rel(ad, fd, scale) <= TOL, (k, ad, fd)
    eps = 1e-8
    fd = (loss(s, dt + eps)[0] - loss(s, dt - eps)[0]) / (2 * eps)
    print(f"stage d/ddt: adjoint {float(lam_dt[0]):+.12e} FD {fd:+.12e}")
    assert _rel(float(lam_dt[0]), fd, 0.0) <= TOL


def test_dt_adjoint_matches_central_differences(setup):
    """The CFL dt through the argmin edge, with the coupled cubic's roots
    differentiated implicitly (4 t^3 - 3 t = x  =>  dt/dx = 1 / (12 t^2 - 3))."""
    _, _, _, be, mask, s, ws, bc = setup
    m0 = mask[0].data_ptr()
    z = lambda: torch.zeros_like(s.h)  # noqa: E731
    ls = s.replace(h=z(), hu=z(), hv=z(), G=z())
    one = torch.ones(1, dtype=torch.float64, device="cuda")
    _hip.call("pxh_adj_roe_exner_dt", be.grid, fields_struct(s), m0, 0.5, one.data_ptr(), fields_struct(ls),
              ws.data_ptr(), _stream())
    torch.cuda.synchronize()
    nz = sum(int((getattr(ls, k) != 0).sum()) for k in ("h", "hu", "hv", "G"))
    assert 0 < nz <= 8  # the two cells of one edge

    def dtf(sx):
        d = torch.empty((), dtype=torch.float64, device="cuda")
        be.dt_into(fields_struct(sx), m0, 0.5, d.data_ptr(), _stream())
        return float(d)

    g = _gen(12)
    vs = {k: _rnd(s.h.shape, g) * 1e-2 for k in ("h", "hu", "hv", "G")}
    ad = sum(float((getattr(ls, k) * vs[k]).sum()) for k in vs)
    eps = 1e-4
    sp = s.replace(**{k: getattr(s, k) + eps * vs[k] for k in vs})
    sm = s.replace(**{k: getattr(s, k) - eps * vs[k] for k in vs})
    fd = (dtf(sp) - dtf(sm)) / (2 * eps)
    print(f"dt along all fields: adjoint {ad:+.12e} FD {fd:+.12e}")
    assert _rel(ad, fd, 0.0) <= TOL
    for k in vs:
        adk = float((getattr(ls, k) * vs[k]).sum())
        sp, sm = s.clone(), s.clone()
        setattr(sp, k, getattr(s, k) + eps * vs[k])
        setattr(sm, k, getattr(s, k) - eps * vs[k])
        fdk = (dtf(sp) - dtf(sm)) / (2 * eps)
        print(f"dt d/d{k}: adjoint {adk:+.12e} FD {fdk:+.12e}")
        assert _rel(adk, fdk, abs(fd)) <= TOL, k


def _run_case():
    mesh, arr, params = cases.calibration_channel(nx=96, ny=40)
    return mesh, arr, params


def test_run_gradient_matches_central_differences(cuda):
    """z_b0 along 16 seeded directions, d50 and porosity, through 24 SSPRK2
    steps with checkpoints every 7 steps (segments of 7, 7, 7, 3)."""
    mesh, arr, params = _run_case()
    run = ExnerGradientRun(mesh, arr, params, n_steps=24, checkpoint_every=7)
    zb0 = run.raw.z_b.clone()
    target = seeded_target(zb0, amp=1e-3)
    g = _gen(7)
    wh = _rnd(zb0.shape, g) * 1e-2
    d50_0, por_0 = float(run.seds[1, 0]), float(run.seds[6, 0])

    def loss_of(h, zb):
        return 0.5 * ((zb - target) ** 2).sum() + (wh * h).sum()

    zb = zb0.clone().requires_grad_(True)
    d50 = torch.tensor(d50_0, dtype=torch.float64, requires_grad=True)
    por = torch.tensor(por_0, dtype=torch.float64, requires_grad=True)
    h, hu, hv, zbT = run(zb, d50, por)
    loss_of(h, zbT).backward()
    grad = zb.grad.clone()
    gd, gp = float(d50.grad), float(por.grad)

    def outs(zbx, dd, pp):
        with torch.no_grad():
            a = run.forward(zbx, dd, pp)
            return a.h.clone(), a.z_b.clone()

    def diff(p, m, eps):
        # (L(p) - L(m)) / (2 eps), with the difference taken cell by cell
        # before the sums so that the loss's own round-off stays out of it
        (hp, zp), (hm, zm) = p, m
        return float((0.5 * (zp - zm) * (zp + zm - 2 * target)).sum() + (wh * (hp - hm)).sum()) / (2 * eps)

    scale = float(grad.abs().sum() * 1e-3)
    errs = []
    for k in range(16):
        v = _rnd(zb0.shape, g) * 1e-3
        ad = float((grad * v).sum())
        # the step is piecewise smooth (upwind selections, the dt argmin):
        # a perturbation of 1e-9 per cell crosses no switch
        eps = 1e-6
        fd = diff(outs(zb0 + eps * v, d50_0, por_0), outs(zb0 - eps * v, d50_0, por_0), eps)
        errs.append(_rel(ad, fd, scale))
        print(f"z_b0 direction {k}: adjoint {ad:+.12e} FD {fd:+.12e} rel {errs[-1]:.2e}")
    assert max(errs) <= TOL, errs
    # d50 and porosity move G in every cell: steps of 1e-4 (relative, and
    # absolute for porosity) keep the difference well above the outputs'
    # round-off while the Shields stress stays clear of critical
    eps = 1e-4 * d50_0
    fd = diff(outs(zb0, d50_0 + eps, por_0), outs(zb0, d50_0 - eps, por_0), eps)
    print(f"d50: adjoint {gd:+.12e} FD {fd:+.12e}")
    assert _rel(gd, fd, 0.0) <= TOL
    eps = 1e-4
    fd = diff(outs(zb0, d50_0, por_0 + eps), outs(zb0, d50_0, por_0 - eps), eps)
    print(f"porosity: adjoint {gp:+.12e} FD {fd:+.12e}")
    assert _rel(gp, fd, 0.0) <= TOL


def test_run_gradient_matches_torch_autograd(cuda):
    """24 SSPRK2 steps of the 96 x 40 calibration channel (checkpoints every
    7): the adjoint kernels' dL/dz_b0, dL/dd50, dL/dporosity against torch
    autograd through the CPU restatement of the same run."""
    from oracle import torch_adjoint as ta
    from tests.test_oracle_torch_adjoint import twin_setup
    mesh, arr, params, raw, m0, fill, seds = twin_setup(96, 40)
    run = ExnerGradientRun(mesh, arr, params, n_steps=24, checkpoint_every=7)
    zb0 = run.raw.z_b.clone()
    assert torch.equal(zb0.cpu(), raw["z_b"])
    target = seeded_target(zb0, amp=1e-3)
    wh = _rnd(zb0.shape, _gen(8)) * 1e-2
    d50_0, por_0 = float(run.seds[1, 0]), float(run.seds[6, 0])
    zb = zb0.clone().requires_grad_(True)
    d50 = torch.tensor(d50_0, dtype=torch.float64, requires_grad=True)
    por = torch.tensor(por_0, dtype=torch.float64, requires_grad=True)
    h, hu, hv, zbT = run(zb, d50, por)
    (0.5 * ((zbT - target) ** 2).sum() + (wh * h).sum()).backward()
    tgt, whc = target.cpu(), wh.cpu()

    def loss_cpu(s):
        return 0.5 * ((s["z_b"] - tgt) ** 2).sum() + (whc * s["h"]).sum()

    _, gz, gd, gp = ta.gradients(raw, m0, fill, seds, mesh.dh, params["cfl"], 24, raw["z_b"], d50_0, por_0,
                                 loss_cpu)
    ez = float((zb.grad.cpu() - gz).abs().max() / gz.abs().max())
    ed, ep = abs(float(d50.grad) - gd) / abs(gd), abs(float(por.grad) - gp) / abs(gp)
    print(f"z_b0 field {ez:.2e} (max |grad| {float(gz.abs().max()):.3e}), d50 {ed:.2e}, porosity {ep:.2e}")
    assert ez <= 1e-12 and ed <= 1e-12 and ep <= 1e-12, (ez, ed, ep)


def test_checkpoint_interval_does_not_change_the_gradient(cuda):
    """Replayed segments reproduce the forward exactly: the gradient is the
    same bit for bit whatever the checkpoint interval."""
    mesh, arr, params = _run_case()
    grads = []
    for every in (1, 5, 12):
        run = ExnerGradientRun(mesh, arr, params, n_steps=12, checkpoint_every=every)
        zb = run.raw.z_b.clone().requires_grad_(True)
        out = run(zb, float(run.seds[1, 0]), float(run.seds[6, 0]))
        (0.5 * (out[3] ** 2).sum()).backward()
        grads.append(zb.grad.clone())
    assert torch.equal(grads[0], grads[1]) and torch.equal(grads[0], grads[2])


def test_forward_of_gradient_run_is_the_solver_step(cuda):
    """The gradient run's forward is the ordinary fused SSPRK2 stepping, and
    its segment replay (stage by stage) reproduces it bit for bit."""
    mesh, arr, params = _run_case()
    run = ExnerGradientRun(mesh, arr, params, n_steps=9, checkpoint_every=9)
    a = run.forward(run.raw.z_b.clone(), float(run.seds[1, 0]), float(run.seds[6, 0]))
    stp = ExnerStepper(run.be, run.s0, run.mask, run.cfl)
    stp.compute_dt()
    for _ in range(9):
        stp.ssprk2_step()
    steps = run._replay(0)
    a9 = steps[-1][0]
    torch.cuda.synchronize()
    for k in HYDRO:
        assert torch.equal(getattr(a, k), getattr(stp.a, k)), k
    # the replayed state before the last step, stepped once more, is the final state
    last = ExnerStepper(run.be, a9, run.mask, run.cfl)
    last.compute_dt()
    assert torch.equal(last.dt, steps[-1][3])
    last.ssprk2_step()
    for k in HYDRO:
        assert torch.equal(getattr(last.a, k), getattr(a, k)), k


def test_erosion_deposition_coefficients_do_not_enter_the_step(cuda):
    """k_e and k_d (seds rows 3, 4) are read by neither compute_G nor the
    fluxes (roe_exner.py:85-130): their gradients are zero, which the forward
    shows by giving the same state bit for bit when they change."""
    mesh, arr, params = _run_case()
    outs = []
    for ke, kd in ((0.04, 0.01), (0.4, 0.2)):
        seds = {k: dict(v, erosion_flux=ke, deposition_flux=kd) for k, v in params["erosion"]["sediments"].items()}
        p = dict(params, erosion=dict(params["erosion"], sediments=seds))
        run = ExnerGradientRun(mesh, arr, p, n_steps=6, checkpoint_every=6)
        a = run.forward(run.raw.z_b.clone(), float(run.seds[1, 0]), float(run.seds[6, 0]))
        outs.append({k: getattr(a, k).clone() for k in HYDRO})
    for k in HYDRO:
        assert torch.equal(outs[0][k], outs[1][k]), k
    assert np.isfinite(float(outs[0]["z_b"].sum()))

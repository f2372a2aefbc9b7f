"""Gradients of a Roe-Exner SSPRK2 run by torch autograd through the
whole-array CPU restatement (oracle/torch_cpu.py), in float64.

Test infrastructure, like the rest of this package: the stand-in for jax.grad
of the reference, which is not runnable here (and which would give NaN
wherever an untaken branch divides 0 by 0, solvers/kernels/roe_exner.py:361,
370, 653; torch_cpu holds safe operands there, so such branches contribute
zero).  It differentiates the same map as pyexner_amd.adjoint.ExnerGradientRun:

* init_fn (solvers/roe_exner_solver.py:58-78): n_b = compute_n(n, z_b0)
  (roe_exner.py:64-73), G = compute_G of the initial hydro state
  (roe_exner.py:85-130), then the ring fill;
* n_steps SSPRK2 steps (integrators/SSPRK2.py:32-40), each with its CFL dt
  (roe_exner.py:211-310, roe_exner_solver.py:118-122);

as a function of the pre-init erodible bed z_b0 and the first sediment class's
d50 and porosity (seds rows 1 and 6, io/pnetcdf_reader.py:146-158).
"""
import numpy as np
import torch

from . import torch_cpu as tc
from .constants import SED_TOL


def sediment_rows(seds, d50, porosity):
    """The (7, nsed) table as rows, with class 0's d50 and every class's
    porosity replaced by the torch scalars d50 and porosity."""
    seds = np.asarray(seds, dtype=np.float64)
    rows = [torch.as_tensor(seds[i], dtype=torch.float64) for i in range(7)]
    rows[1] = torch.cat([d50.reshape(1), rows[1][1:]])
    rows[6] = porosity.reshape(1).expand(seds.shape[1])
    return rows


def init_state(raw, m0, fill, rows, z_b0):
    """init_fn of the raw state (dict of float64 tensors) with z_b = z_b0."""
    s = dict(raw)
    s["z_b"] = z_b0
    frac, d50 = rows[0], rows[1]
    nb = torch.mean(d50 * frac) ** (1 / 6) / 21.1
    s["n_b"] = torch.where(z_b0 > SED_TOL, nb, s["n"])
    s["G"] = tc.compute_g(s, m0, rows, False)
    return fill(s)


def run(raw, m0, fill, seds, dx, cfl, n_steps, z_b0, d50, porosity):
    """The final state (dict) of n_steps SSPRK2 steps from z_b0; differentiable
    in z_b0, d50 and porosity."""
    rows = sediment_rows(seds, d50, porosity)
    s = init_state(raw, m0, fill, rows, z_b0)
    for _ in range(n_steps):
        s, _ = tc.ssprk2_step(s, m0, fill, rows, False, dx, cfl)
    return s


def gradients(raw, m0, fill, seds, dx, cfl, n_steps, z_b0, d50, porosity, loss_fn):
    """(loss, dL/dz_b0, dL/dd50, dL/dporosity) of loss_fn(final state) by
    reverse-mode autograd."""
    zb = torch.as_tensor(z_b0, dtype=torch.float64).clone().requires_grad_(True)
    dd = torch.tensor(float(d50), dtype=torch.float64, requires_grad=True)
    pp = torch.tensor(float(porosity), dtype=torch.float64, requires_grad=True)
    loss = loss_fn(run(raw, m0, fill, seds, dx, cfl, n_steps, zb, dd, pp))
    loss.backward()
    return float(loss.detach()), zb.grad.detach(), float(dd.grad), float(pp.grad)

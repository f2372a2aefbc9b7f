"""Plain Roe kernels, restated from solvers/kernels/roe.py:13-315.

The as-coded quirks are kept: the Harten-Hyman lambda_E uses the already
modified lambda (roe.py:82-92), the source fix uses ``dx / 2*lambda`` ==
``(dx/2)*lambda`` (roe.py:180-181) and ``beta3min`` is built from
``hi - alpha_1`` (roe.py:173).  Fully dry edges produce NaN exactly as the
reference does (0/0 in the Roe average, roe.py:66).
"""
import numpy as np

from .constants import DRY_TOL, VEL_TOL, g
from .roe_exner import _sq, _edge_slices, _err


def compute_dt(state, mask, dx):
    """roe.py:13-30.  The mask is applied after the min, so it has no effect."""
    with np.errstate(**_err):
        zero = np.zeros((), dtype=state.h.dtype)
        wet = state.h > DRY_TOL
        h = np.where(wet, state.h, zero)
        hv = np.where(wet, state.hv, zero)
        hu = np.where(wet, state.hu, zero)
        u = hu / (h + DRY_TOL)
        v = hv / (h + DRY_TOL)
        c = np.sqrt(g * h)
        dt_x = np.min(dx / (np.abs(u) + c + VEL_TOL))
        dt_y = np.min(dx / (np.abs(v) + c + VEL_TOL))
        return np.minimum(dt_x, dt_y)


def compute_dt_roe(state, cfl, mask, dx):
    """compute_dt_roe (roe_solver.py:56-60): cfl * dt, no clamp."""
    return cfl * compute_dt(state, mask, dx)


def roe_edge(si, sj, nx, ny, dx):
    """roe.py:32-249."""
    hi, hui, hvi, zi, ni = si
    hj, huj, hvj, zj, nj = sj
    zero = np.zeros((), dtype=np.asarray(hi).dtype)
    with np.errstate(**_err):
        hi = np.where(hi > DRY_TOL, hi, zero)
        hj = np.where(hj > DRY_TOL, hj, zero)
        sqrt_i = np.sqrt(hi)
        sqrt_j = np.sqrt(hj)
        ui = np.where(hi > 0.0, hui / hi, zero)
        vi = np.where(hi > 0.0, hvi / hi, zero)
        uj = np.where(hj > 0.0, huj / hj, zero)
        vj = np.where(hj > 0.0, hvj / hj, zero)
        uhati = ui * nx + vi * ny
        vhati = -ui * ny + vi * nx
        uhatj = uj * nx + vj * ny
        vhatj = -uj * ny + vj * nx
        hav = 0.5 * (hi + hj)
        htilde = np.where(hav > DRY_TOL, 0.5 * (hi + hj), zero)
        ctilde = np.sqrt(g * htilde)
        utilde = (uhati * sqrt_i + uhatj * sqrt_j) / (sqrt_i + sqrt_j)
        vtilde = (vhati * sqrt_i + vhatj * sqrt_j) / (sqrt_i + sqrt_j)
        lambda_1 = utilde - ctilde
        lambda_2 = utilde
        lambda_3 = utilde + ctilde
        # Harten-Hyman: lambda replaced first, lambda_E from the replaced value
        ei = uhati - np.sqrt(g * hi)
        ej = uhatj - np.sqrt(g * hj)
        mask_i = (ei < 0.0) & (ej > 0.0)
        lambda_1 = np.where(mask_i, ei * (ej - lambda_1) / ((ej - ei) + VEL_TOL), lambda_1)
        lambda_E1 = np.where(mask_i, lambda_1 - ei * (ej - lambda_1) / ((ej - ei) + VEL_TOL), zero)
        ei = uhati + np.sqrt(g * hi)
        ej = uhatj + np.sqrt(g * hj)
        mask_j = (ei < 0.0) & (ej > 0.0)
        lambda_3 = np.where(mask_j, ej * (lambda_3 - ei) / ((ej - ei) + VEL_TOL), lambda_3)
        lambda_E3 = np.where(mask_j, lambda_3 - ej * (lambda_3 - ei) / ((ej - ei) + VEL_TOL), zero)
        lambda_E2 = np.zeros_like(lambda_E3)
        ones = np.ones_like(vtilde)
        zeros = np.zeros_like(vtilde)
        # P_inv rows (roe.py:107-113)
        denom = lambda_1 - lambda_3 + VEL_TOL
        Pinv = (
            (-lambda_3 / denom, ones / denom, zeros),
            (-vtilde / ctilde, zeros, ones / ctilde),
            (lambda_1 / denom, -ones / denom, zeros),
        )
        dh = hj - hi
        dhu = hj * uhatj - hi * uhati
        dhv = hj * vhatj - hi * vhati
        dU = (dh, dhu, dhv)
        # alphas = P_inv dU (einsum "...ji,...i->...j"), summed in index order
        alphas = [Pinv[r][0] * dU[0] + Pinv[r][1] * dU[1] + Pinv[r][2] * dU[2] for r in range(3)]
        # sources (roe.py:132-157)
        dz = zj - zi
        di = hi + zi
        dj = hj + zj
        dd = dj - di
        thrust_a = -g * htilde * dz
        mask1_dz = (dz >= 0) & (di < zj)
        mask2_dz = (dz < 0) & (dj < zi)
        dztilde = np.where(mask1_dz, hi, np.where(mask2_dz, hj, dz))
        hr = np.where(dz >= 0, hi, hj)
        thrust_b = -g * (hr - 0.5 * np.abs(dztilde)) * dztilde
        mask_thrust = (dz * dd >= 0.0) & (utilde * dz > 0.0)
        thrust = np.where(mask_thrust, np.maximum(thrust_a, thrust_b), thrust_b)
        ntilde = 0.5 * (ni + nj)
        sf = (_sq(ntilde) * np.sqrt(_sq(utilde) + _sq(vtilde)) * utilde) / (
            np.maximum(DRY_TOL, htilde ** (4 / 3)))
        tau = g * htilde * sf * dx
        Tn = (zeros, thrust - tau, zeros)
        betas = [Pinv[r][0] * Tn[0] + Pinv[r][1] * Tn[1] + Pinv[r][2] * Tn[2] for r in range(3)]
        lambdas = (lambda_1, lambda_2, lambda_3)
        # beta fix (roe.py:169-189)
        h_i1star = hi + alphas[0] - (betas[0] / lambdas[0])
        h_j3star = hj - alphas[2] + (betas[2] / lambdas[2])
        beta1min = -(hi + alphas[0]) * np.abs(lambdas[0])
        beta3min = -(hi - alphas[0]) * np.abs(lambdas[2])
        lam_abs_max = np.maximum(np.maximum(np.abs(lambda_1), np.abs(lambda_2)), np.abs(lambda_3))
        dt = dx / lam_abs_max
        dt1star = (dx / 2 * lambdas[0]) * (hi / (hi - h_i1star + VEL_TOL))
        dt3star = (dx / 2 * lambdas[2]) * (hj / (hj - h_j3star + VEL_TOL))
        m = (h_i1star < 0.0) & (h_j3star > 0.0) & (dt1star < dt)
        betas[0] = np.where(m, np.where(-beta1min >= beta3min, beta1min, betas[0]), betas[0])
        betas[2] = np.where(m, -betas[0], betas[2])
        m = (h_i1star > 0.0) & (h_j3star < 0.0) & (dt3star < dt)
        betas[2] = np.where(m, np.where(-beta3min >= beta3min, beta1min, betas[2]), betas[2])
        betas[0] = np.where(m, -betas[2], betas[0])
        mask_1 = (hi < DRY_TOL) & (h_i1star < 0.0)
        mask_2 = (hj < DRY_TOL) & (h_j3star < 0.0)
        P = [(ones, lambda_1, vtilde), (zeros, zeros, ctilde), (ones, lambda_3, vtilde)]
        lambdas_E = (lambda_E1, lambda_E2, lambda_E3)
        upwP = [np.zeros_like(vtilde) for _ in range(3)]
        upwM = [np.zeros_like(vtilde) for _ in range(3)]
        for k in range(3):
            lam = lambdas[k]
            for c in range(3):
                flux = (lam * alphas[k] - betas[k]) * P[k][c]
                fpos = np.where(mask_2, zero, np.where(mask_1, flux, np.where(lam > 0.0, flux, zero)))
                fneg = np.where(mask_2, flux, np.where(mask_1, zero, np.where(lam <= 0.0, flux, zero)))
                upwP[c] = upwP[c] + fpos
                upwM[c] = upwM[c] + fneg
        for k in range(3):
            lamE = lambdas_E[k]
            for c in range(3):
                flux = lamE * alphas[k] * P[k][c]
                upwP[c] = upwP[c] + np.where(lamE > 0.0, flux, zero)
                upwM[c] = upwM[c] + np.where(lamE <= 0.0, flux, zero)
        outP = (upwP[0], nx * upwP[1] + (-ny) * upwP[2], ny * upwP[1] + nx * upwP[2])
        outM = (upwM[0], nx * upwM[1] + (-ny) * upwM[2], ny * upwM[1] + nx * upwM[2])
    return outP, outM


def roe_solve_2D(state, dt, dx):
    """roe.py:251-315 (no active-edge masking)."""
    fields = (state.h, state.hu, state.hv, state.z, state.n)
    res = {}
    for axis, (nx, ny) in (("x", (1, 0)), ("y", (0, -1))):
        s1 = tuple(_edge_slices(f, axis)[0] for f in fields)
        s2 = tuple(_edge_slices(f, axis)[1] for f in fields)
        res[axis] = roe_edge(s1, s2, nx, ny, dx)
    Ny, Nx = state.h.shape
    out = []
    for c, q in enumerate((state.h, state.hu, state.hv)):
        F = np.zeros((Ny, Nx), dtype=state.h.dtype)
        F[:, :-1] += res["x"][1][c]
        F[:, 1:] += res["x"][0][c]
        F[:-1, :] += res["y"][1][c]
        F[1:, :] += res["y"][0][c]
        out.append(q - dt * F / dx)
    return state.replace(h=out[0], hu=out[1], hv=out[2])

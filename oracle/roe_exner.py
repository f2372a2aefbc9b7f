"""Roe-Exner kernels, restated from solvers/kernels/roe_exner.py.

Each function follows the cited reference lines operation by operation (same
operand order, same where/maximum structure), in the dtype of its inputs.
"""
import numpy as np

from .constants import DRY_TOL, SED_TOL, VEL_TOL, g

_err = dict(divide="ignore", invalid="ignore", over="ignore", under="ignore")


def _sq(x):
    # jax.lax.integer_pow(x, 2)
    return x * x


def _cube(x):
    # jax.lax.integer_pow(x, 3) == x * (x * x)
    return x * (x * x)


# ---------------------------------------------------------------------------
# momentum_corrections  (roe_exner.py:14-62)
# ---------------------------------------------------------------------------
def momentum_corrections(state, mask):
    h, z, z_b = state.h, state.z, state.z_b
    m = mask[0]
    C = np.s_[1:-1, 1:-1]
    W = np.s_[1:-1, 0:-2]
    E = np.s_[1:-1, 2:]
    N = np.s_[0:-2, 1:-1]
    S = np.s_[2:, 1:-1]
    hc = h[C]
    ztot = z + z_b
    eta = hc + ztot[C]
    dryW, dryE = h[W] < DRY_TOL, h[E] < DRY_TOL
    dryN, dryS = h[N] < DRY_TOL, h[S] < DRY_TOL
    stop_u = ((eta < ztot[E]) & dryE) | ((eta < ztot[W]) & dryW) | m[E] | m[W]
    stop_v = ((eta < ztot[N]) & dryN) | ((eta < ztot[S]) & dryS) | m[N] | m[S]
    active = (hc > DRY_TOL) & ~m[C]
    hu = state.hu.copy()
    hv = state.hv.copy()
    zero = np.zeros((), dtype=hu.dtype)
    hu[C] = np.where(active & ~stop_u, state.hu[C], zero)
    hv[C] = np.where(active & ~stop_v, state.hv[C], zero)
    return state.replace(hu=hu, hv=hv)


# ---------------------------------------------------------------------------
# compute_n  (roe_exner.py:64-73, factory 75-83)
# ---------------------------------------------------------------------------
def compute_n(n, z_b, seds):
    frac = seds[0][:, None, None]
    d50 = seds[1][:, None, None]
    d50 = np.mean(d50 * frac, axis=0)
    np_ = d50 ** (1 / 6) / 21.1
    return np.where(z_b > SED_TOL, np_, n).astype(n.dtype)


def compute_n_mode(n, z_b, seds, constant):
    return n if constant else compute_n(n, z_b, seds)


# ---------------------------------------------------------------------------
# compute_G  (roe_exner.py:85-130, constant-mode factory 132-144)
# ---------------------------------------------------------------------------
def compute_G(h, hu, hv, seds, mask):
    with np.errstate(**_err):
        h = np.maximum(h, DRY_TOL)
        u = hu / h
        v = hv / h
        vel2 = _sq(u) + _sq(v)
        frac, d50, density, k_d, k_e, theta_c, porosity = seds
        s = density / 1000.0
        np_ = d50 ** (1 / 6) / 21.1
        G = None
        cbrt_h = np.cbrt(h)
        sqrt_h = np.sqrt(h)
        sqrt_g = np.sqrt(np.asarray(g, dtype=h.dtype))
        for k in range(seds.shape[1]):
            theta_p = (_sq(np_[k]) * vel2) / ((s[k] - 1.0) * d50[k] * cbrt_h)
            theta_p = np.maximum(theta_p, 0.0)
            delta_theta = np.maximum(theta_p - theta_c[k], 0.0)
            gamma1 = _cube(np_[k]) * sqrt_g / ((s[k] - 1.0) * sqrt_h)
            gamma2 = 8.0 * np.sqrt(delta_theta) / (theta_p ** 1.5 + 1e-5)
            gamma3 = delta_theta
            contrib = frac[k] * gamma1 * gamma2 * gamma3
            G = contrib if G is None else G + contrib
        G = G / (1.0 - porosity.mean())
        zero = np.zeros((), dtype=h.dtype)
        G = np.where(mask[0], zero, G)
        G = np.where(h > DRY_TOL, G, zero)
    return G


def compute_G_constant(h, mask, state_G):
    zero = np.zeros((), dtype=state_G.dtype)
    G = np.where(mask[0], zero, state_G)
    return np.where(h > DRY_TOL, G, zero)


def compute_G_mode(h, hu, hv, seds, mask, state_G, constant):
    if constant:
        return compute_G_constant(h, mask, state_G)
    return compute_G(h, hu, hv, seds, mask)


# ---------------------------------------------------------------------------
# _get_lambda  (roe_exner.py:163-209): trigonometric roots of the coded cubic
# ---------------------------------------------------------------------------
def get_lambda(utilde, atilde, ctilde):
    with np.errstate(**_err):
        a = -1
        b = -2 * utilde
        c = (1 + atilde) * _sq(ctilde) - _sq(utilde)
        d = -_sq(ctilde) * atilde * utilde
        p = (3 * a * c - _sq(b)) / (3 * a ** 2)
        q = (2 * _cube(b) - 9 * a * b * c + 27 * a ** 2 * d) / (27 * a ** 3)
        cos_theta = np.clip((3 * q) / (2 * p) * np.sqrt(-3 / p), -1.0, 1.0)
        theta = np.arccos(cos_theta)
        sqrt_term = 2 * np.sqrt(-p / 3)
        shift = -b / (3 * a)
        lambda_max = sqrt_term * np.cos(theta / 3) + shift
        lambda_min = sqrt_term * np.cos((theta + 2 * np.pi) / 3) + shift
        small = atilde < 1e-5
        lambda_max = np.where(small, utilde + ctilde, lambda_max)
        lambda_min = np.where(small, utilde - ctilde, lambda_min)
    return lambda_min, lambda_max


# ---------------------------------------------------------------------------
# compute_dt (per edge) roe_exner.py:211-255 and compute_dt_2D 257-310
# ---------------------------------------------------------------------------
def compute_dt_edge(si, sj, nx, ny, dx, cubic_dtype=None):
    hi, hui, hvi, zi, gi = si
    hj, huj, hvj, zj, gj = sj
    with np.errstate(**_err):
        hi = np.maximum(hi, DRY_TOL)
        hj = np.maximum(hj, DRY_TOL)
        sqrt_i = np.sqrt(hi)
        sqrt_j = np.sqrt(hj)
        denom = sqrt_i + sqrt_j
        ui = hui / hi
        vi = hvi / hi
        uj = huj / hj
        vj = hvj / hj
        uhati = ui * nx + vi * ny
        vhati = -ui * ny + vi * nx
        uhatj = uj * nx + vj * ny
        vhatj = -uj * ny + vj * nx
        htilde = 0.5 * (hi + hj)
        utilde = (uhati * sqrt_i + uhatj * sqrt_j) / denom
        vtilde = (vhati * sqrt_i + vhatj * sqrt_j) / denom  # noqa: F841 (as coded)
        ctilde = np.sqrt(g * htilde)
        gtilde = 0.5 * (gi + gj)
        atilde = gtilde * (_sq(uhati) + uhati * uhatj + _sq(uhatj) + vhati * vhatj) / (sqrt_i * sqrt_j)
        lambda_1 = utilde - ctilde
        lambda_3 = utilde + ctilde
        if cubic_dtype is not None:
            lam_min, lam_max = get_lambda(utilde.astype(cubic_dtype), atilde.astype(cubic_dtype),
                                          ctilde.astype(cubic_dtype))
            lam_min = lam_min.astype(hi.dtype)
            lam_max = lam_max.astype(hi.dtype)
        else:
            lam_min, lam_max = get_lambda(utilde, atilde, ctilde)
        roe_speed = np.maximum(np.abs(lam_min), np.abs(lam_max))
        swe_speed = np.maximum(np.abs(lambda_1), np.abs(lambda_3))
        dt_exner = dx / roe_speed
        dt_roe = dx / swe_speed
        return np.where(roe_speed < VEL_TOL, dt_roe, dt_exner)


def _edge_slices(f, axis):
    if axis == "x":
        return f[:, :-1], f[:, 1:]
    return f[:-1, :], f[1:, :]


def active_edges(h, m0, axis):
    hi, hj = _edge_slices(h, axis)
    mi, mj = _edge_slices(m0, axis)
    return ((hi > DRY_TOL) | (hj > DRY_TOL)) & ~(mi | mj)


def compute_dt_2D(state, dx, mask, cubic_dtype=None):
    out = []
    for axis, (nx, ny) in (("x", (1, 0)), ("y", (0, -1))):
        s1 = tuple(_edge_slices(f, axis)[0] for f in (state.h, state.hu, state.hv, state.z, state.G))
        s2 = tuple(_edge_slices(f, axis)[1] for f in (state.h, state.hu, state.hv, state.z, state.G))
        act = active_edges(state.h, mask[0], axis)
        dte = compute_dt_edge(s1, s2, nx, ny, dx, cubic_dtype)
        inf = np.asarray(np.inf, dtype=state.h.dtype)
        out.append(np.where(act, dte, inf).min() if dte.size else inf)
    return np.minimum(out[0], out[1])


def compute_dt_roeexner(state, cfl, mask, dx, cubic_dtype=None):
    """compute_dt_roeexner (roe_exner_solver.py:118-122) on one rank."""
    local_dt = cfl * compute_dt_2D(state, dx, mask, cubic_dtype)
    return np.maximum(local_dt, 1e-6)


# ---------------------------------------------------------------------------
# roe_solver (per edge)  roe_exner.py:312-533
# ---------------------------------------------------------------------------
def roe_edge(si, sj, nx, ny, dx):
    hi, hui, hvi, z_bi, zi, ni = si
    hj, huj, hvj, z_bj, zj, nj = sj
    dt_ = np.asarray(hi).dtype
    zero = np.zeros((), dtype=dt_)
    with np.errstate(**_err):
        hi = np.maximum(hi, DRY_TOL)
        hj = np.maximum(hj, DRY_TOL)
        ui = np.where(hi > DRY_TOL, hui / hi, zero)
        vi = np.where(hi > DRY_TOL, hvi / hi, zero)
        uj = np.where(hj > DRY_TOL, huj / hj, zero)
        vj = np.where(hj > DRY_TOL, hvj / hj, zero)
        uhati = ui * nx + vi * ny
        vhati = -ui * ny + vi * nx
        uhatj = uj * nx + vj * ny
        vhatj = -uj * ny + vj * nx
        sqrt_i = np.sqrt(hi)
        sqrt_j = np.sqrt(hj)
        denom_sqrt = sqrt_i + sqrt_j
        htilde = 0.5 * (hi + hj)
        ctilde = np.sqrt(g * htilde)
        utilde = (uhati * sqrt_i + uhatj * sqrt_j) / denom_sqrt
        vtilde = (vhati * sqrt_i + vhatj * sqrt_j) / denom_sqrt
        lambda_1_roe = utilde - ctilde
        lambda_2 = utilde
        lambda_3_roe = utilde + ctilde
        # Harten-Hyman entropy fix (roe_exner.py:357-373)
        ei = uhati - np.sqrt(g * hi)
        ej = uhatj - np.sqrt(g * hj)
        mask_i = (ei < 0.0) & (ej > 0.0)
        lambda_E1 = np.where(mask_i, lambda_1_roe - ei * (ej - lambda_1_roe) / (ej - ei), zero)
        lambda_1 = np.where(mask_i, ei * (ej - lambda_1_roe) / (ej - ei), lambda_1_roe)
        ei = uhati + np.sqrt(g * hi)
        ej = uhatj + np.sqrt(g * hj)
        mask_j = (ei < 0.0) & (ej > 0.0)
        lambda_E3 = np.where(mask_j, lambda_3_roe - ej * (lambda_3_roe - ei) / (ej - ei), zero)
        lambda_3 = np.where(mask_j, ej * (lambda_3_roe - ei) / ((ej - ei) + VEL_TOL), lambda_3_roe)
        lambda_E2 = np.zeros_like(lambda_E3)
        # jumps and wave strengths (roe_exner.py:399-412)
        dh = hj - hi
        dhu = hj * uhatj - hi * uhati
        dhv = hj * vhatj - hi * vhati
        denom = -2.0 * ctilde
        alpha_1 = (-lambda_3_roe * dh + dhu) / denom
        alpha_2 = (-vtilde * dh + dhv) / ctilde
        alpha_3 = (lambda_1_roe * dh - dhu) / denom
        # bed source (roe_exner.py:421-432)
        z_bi = np.where(z_bi > SED_TOL, z_bi, SED_TOL * 0.01)
        z_bj = np.where(z_bj > SED_TOL, z_bj, SED_TOL * 0.01)
        dz = (zj + z_bj) - (zi + z_bi)
        thrust = -g * htilde * dz
        # friction (roe_exner.py:455-457)
        ntilde = 0.5 * (ni + nj)
        sf = (_sq(ntilde) * np.sqrt(_sq(utilde) + _sq(vtilde)) * utilde) / (
            np.maximum(DRY_TOL, htilde ** (4 / 3)))
        tau = g * htilde * sf * dx
        beta_1 = (thrust - tau) / denom
        beta_2 = np.zeros_like(beta_1)
        beta_3 = -beta_1
        # h_i1star / h_j3star dry masks (roe_exner.py:470-481) are dead: h >= DRY_TOL
        ones = np.ones_like(vtilde)
        zeros = np.zeros_like(vtilde)
        P = [(ones, lambda_1, vtilde), (zeros, zeros, ctilde), (ones, lambda_3, vtilde)]
        lambdas = (lambda_1, lambda_2, lambda_3)
        alphas = (alpha_1, alpha_2, alpha_3)
        betas = (beta_1, beta_2, beta_3)
        lambdas_E = (lambda_E1, lambda_E2, lambda_E3)
        upwP = [np.zeros_like(vtilde) for _ in range(3)]
        upwM = [np.zeros_like(vtilde) for _ in range(3)]
        for k in range(3):  # roe_body (roe_exner.py:484-507)
            lam = lambdas[k]
            pos = lam > 0.0
            neg = lam <= 0.0
            for c in range(3):
                flux = (lam * alphas[k] - betas[k]) * P[k][c]
                upwP[c] = upwP[c] + np.where(pos, flux, zero)
                upwM[c] = upwM[c] + np.where(neg, flux, zero)
        for k in range(3):  # entropy_body (roe_exner.py:510-524)
            lamE = lambdas_E[k]
            pos = lamE > 0.0
            neg = lamE <= 0.0
            for c in range(3):
                flux = lamE * alphas[k] * P[k][c]
                upwP[c] = upwP[c] + np.where(pos, flux, zero)
                upwM[c] = upwM[c] + np.where(neg, flux, zero)
        # rotate back with Tk_inv (roe_exner.py:526-531); zero terms dropped
        outP = (upwP[0], nx * upwP[1] + (-ny) * upwP[2], ny * upwP[1] + nx * upwP[2])
        outM = (upwM[0], nx * upwM[1] + (-ny) * upwM[2], ny * upwM[1] + nx * upwM[2])
    return outP, outM


def roe_solve_2D(state, dt, dx, mask):
    """roe_exner.py:535-612."""
    h = state.h
    fields = (state.h, state.hu, state.hv, state.z_b, state.z, state.n_b)
    zero = np.zeros((), dtype=h.dtype)
    res = {}
    for axis, (nx, ny) in (("x", (1, 0)), ("y", (0, -1))):
        s1 = tuple(_edge_slices(f, axis)[0] for f in fields)
        s2 = tuple(_edge_slices(f, axis)[1] for f in fields)
        P, M = roe_edge(s1, s2, nx, ny, dx)
        act = active_edges(h, mask[0], axis)
        res[axis] = ([np.where(act, p, zero) for p in P], [np.where(act, m, zero) for m in M])
    Ny, Nx = h.shape
    out = []
    for c, q in enumerate((state.h, state.hu, state.hv)):
        F = np.zeros((Ny, Nx), dtype=h.dtype)
        F[:, :-1] += res["x"][1][c]
        F[:, 1:] += res["x"][0][c]
        F[:-1, :] += res["y"][1][c]
        F[1:, :] += res["y"][0][c]
        out.append(q - dt * F / dx)
    return state.replace(h=out[0], hu=out[1], hv=out[2])


# ---------------------------------------------------------------------------
# exner_solver (per edge)  roe_exner.py:614-663 and exner_solve_2D 665-717
# ---------------------------------------------------------------------------
def exner_edge(si, sj, nx, ny, dx):
    hi, hui, hvi, z_bi, zi, gi = si
    hj, huj, hvj, z_bj, zj, gj = sj
    zero = np.zeros((), dtype=np.asarray(hi).dtype)
    with np.errstate(**_err):
        hi = np.where(hi > DRY_TOL, hi, DRY_TOL * 0.01)
        hj = np.where(hj > DRY_TOL, hj, DRY_TOL * 0.01)
        z_bi = np.where(z_bi > SED_TOL, z_bi, SED_TOL * 0.01)
        z_bj = np.where(z_bj > SED_TOL, z_bj, SED_TOL * 0.01)
        sqrt_i = np.sqrt(hi)
        sqrt_j = np.sqrt(hj)
        ui = np.where(hi > DRY_TOL, hui / hi, zero)
        vi = np.where(hi > DRY_TOL, hvi / hi, zero)
        uj = np.where(hj > DRY_TOL, huj / hj, zero)
        vj = np.where(hj > DRY_TOL, hvj / hj, zero)
        uhati = ui * nx + vi * ny
        uhatj = uj * nx + vj * ny
        utilde = (uhati * sqrt_i + uhatj * sqrt_j) / (sqrt_i + sqrt_j)
        umagi = _sq(ui) + _sq(vi)
        umagj = _sq(uj) + _sq(vj)
        qb_nhati = gi * umagi * uhati
        qb_nhatj = gj * umagj * uhatj
        gtilde = 0.5 * (gi + gj)
        dz = z_bj - z_bi
        dqbhat = gtilde * umagj * uhatj - gtilde * umagi * uhati
        lambda_4 = np.where(np.abs(dz) > SED_TOL, dqbhat / dz, utilde)
        corrector_i = (gtilde - gi) * umagi * uhati
        corrector_j = (gtilde - gj) * umagj * uhatj
        qbni = qb_nhati + corrector_i
        qbnj = qb_nhatj + corrector_j
        return np.where(lambda_4 >= 0, qbni, qbnj)


def exner_solve_2D(state, dt, dx, mask):
    """roe_exner.py:665-717."""
    h, z_b = state.h, state.z_b
    fields = (state.h, state.hu, state.hv, state.z_b, state.z, state.G)
    zero = np.zeros((), dtype=h.dtype)
    Fs = {}
    for axis, (nx, ny) in (("x", (1, 0)), ("y", (0, -1))):
        s1 = tuple(_edge_slices(f, axis)[0] for f in fields)
        s2 = tuple(_edge_slices(f, axis)[1] for f in fields)
        F = exner_edge(s1, s2, nx, ny, dx)
        mi, mj = _edge_slices(mask[0], axis)
        zbi, zbj = _edge_slices(z_b, axis)
        stop = (mi | mj) | ((zbi < SED_TOL) & (zbj < SED_TOL))
        hi, hj = _edge_slices(h, axis)
        depth = (hi > DRY_TOL) | (hj > DRY_TOL)
        Fs[axis] = np.where(depth & ~stop, F, zero)
    Ny, Nx = h.shape
    fl = np.zeros((Ny, Nx), dtype=h.dtype)
    fl[:, 1:] += -Fs["x"]
    fl[:, :-1] += Fs["x"]
    fl[1:, :] += -Fs["y"]
    fl[:-1, :] += Fs["y"]
    return state.replace(z_b=state.z_b - dt * fl / dx)

"""Whole-array torch restatement of the Roe-Exner SSPRK2 step on the CPU.

Test and benchmark infrastructure, like the rest of this package: it is the
multi-threaded CPU baseline (torch's intra-op thread pool over whole-array
slices, where and scatter-add, the shape of the reference's JAX code on a
CPU), not a product path.  It follows the numpy restatement in
oracle/roe_exner.py, which cites the reference line by line; the lines cited
here are the reference's:

* get_lambda            - solvers/kernels/roe_exner.py:163-209
* dt_2d                 - solvers/kernels/roe_exner.py:211-310, roe_exner_solver.py:118-122
* roe_solve_2d          - solvers/kernels/roe_exner.py:312-612
* momentum_corrections  - solvers/kernels/roe_exner.py:14-62
* compute_g             - solvers/kernels/roe_exner.py:85-144
* exner_solve_2d        - solvers/kernels/roe_exner.py:614-717
* fill                  - boundaries/roeexner_reflective.py:31-63, roeexner_transmissive.py:17-31,
                          roeexner_transmissivebed.py:17-27 and the periodic copy
* step / ssprk2_step    - solvers/roe_exner_solver.py:80-116, integrators/SSPRK2.py:32-40

Float64 runs agree with the numpy restatement to rounding (tests/
test_oracle_torch_cpu.py).  cbrt is h ** (1/3) here (torch has no cbrt).

The functions are differentiable by torch autograd (oracle/torch_adjoint.py):
a denominator or argument that only an untaken branch of a `where` uses is
replaced there by a safe value (1, or the branch's value held constant), so
that branch contributes a zero derivative instead of 0 * inf = NaN.  Forward
values are unchanged: the replaced operands only ever feed discarded
branches.  The sediment rows may hold torch scalars (d50 and porosity with
requires_grad).
"""
import numpy as np
import torch

from .constants import DRY_TOL, SED_TOL, VEL_TOL, g

FIELDS = ("h", "hu", "hv", "z", "z_b", "n", "n_b", "G")


def _rows(r):
    """A sediment row as a list of per-class scalars: Python floats from a
    numpy row, torch scalars (autograd) from a tensor."""
    if torch.is_tensor(r):
        return [r.reshape(-1)[k] for k in range(r.numel())]
    return [float(x) for x in np.asarray(r, dtype=np.float64).reshape(-1)]


def _mean(row):
    if any(torch.is_tensor(x) for x in row):
        return torch.stack([torch.as_tensor(x, dtype=torch.float64) for x in row]).mean()
    return float(np.mean(np.asarray(row, dtype=np.float64)))


def _xs(f, axis):
    return (f[:, :-1], f[:, 1:]) if axis == "x" else (f[:-1, :], f[1:, :])


def _to_lo(e, axis):
    """Edge values e onto the full grid at each edge's first cell (zeros
    elsewhere)."""
    return torch.nn.functional.pad(e, (0, 1) if axis == "x" else (0, 0, 0, 1))


def _to_hi(e, axis):
    """Edge values e onto the full grid at each edge's second cell."""
    return torch.nn.functional.pad(e, (1, 0) if axis == "x" else (0, 0, 1, 0))


def get_lambda(ut, at, ct):
    b = -2 * ut
    c = (1 + at) * (ct * ct) - ut * ut
    d = -(ct * ct) * at * ut
    # a = -1: p = (3ac - b^2) / 3, q = (2b^3 - 9abc + 27d) / -27
    p = (-3 * c - b * b) / 3
    q = (2 * (b * (b * b)) + 9 * b * c + 27 * d) / -27
    raw = (3 * q) / (2 * p) * torch.sqrt(-3 / p)
    inside = (raw > -1.0) & (raw < 1.0)
    # at the clip arccos(+-1) is held constant (the clipped argument does not
    # move), so its derivative is zero there, not arccos' infinite one
    th_clip = torch.arccos(torch.clamp(raw, -1.0, 1.0)).detach()
    th = torch.where(inside, torch.arccos(torch.where(inside, raw, torch.zeros_like(raw))), th_clip)
    st = 2 * torch.sqrt(-p / 3)
    shift = b / 3
    lmax = st * torch.cos(th / 3) + shift
    lmin = st * torch.cos((th + 2 * np.pi) / 3) + shift
    small = at < 1e-5
    return torch.where(small, ut - ct, lmin), torch.where(small, ut + ct, lmax)


def _dt_edge(si, sj, nx, ny, dx):
    hi, hui, hvi, gi = si
    hj, huj, hvj, gj = sj
    hi = torch.clamp(hi, min=DRY_TOL)
    hj = torch.clamp(hj, min=DRY_TOL)
    ri, rj = torch.sqrt(hi), torch.sqrt(hj)
    ui, vi, uj, vj = hui / hi, hvi / hi, huj / hj, hvj / hj
    uni, vni = ui * nx + vi * ny, -ui * ny + vi * nx
    unj, vnj = uj * nx + vj * ny, -uj * ny + vj * nx
    ut = (uni * ri + unj * rj) / (ri + rj)
    ct = torch.sqrt(g * (0.5 * (hi + hj)))
    at = 0.5 * (gi + gj) * (uni * uni + uni * unj + unj * unj + vni * vnj) / (ri * rj)
    lmin, lmax = get_lambda(ut, at, ct)
    roe = torch.maximum(lmin.abs(), lmax.abs())
    swe = torch.maximum((ut - ct).abs(), (ut + ct).abs())
    return dx / torch.where(roe < VEL_TOL, swe, roe)


def dt_2d(s, m0, dx, cfl):
    best = torch.tensor(float("inf"), dtype=s["h"].dtype)
    for axis, (nx, ny) in (("x", (1, 0)), ("y", (0, -1))):
        a = [_xs(s[k], axis) for k in ("h", "hu", "hv", "G")]
        dte = _dt_edge(tuple(x[0] for x in a), tuple(x[1] for x in a), nx, ny, dx)
        hi, hj = a[0]
        mi, mj = _xs(m0, axis)
        act = ((hi > DRY_TOL) | (hj > DRY_TOL)) & ~(mi | mj)
        if dte.numel():
            best = torch.minimum(best, torch.where(act, dte, torch.full_like(dte, float("inf"))).min())
    return torch.clamp(cfl * best, min=1e-6)


def _roe_edge(si, sj, nx, ny, dx):
    hi, hui, hvi, zbi, zi, ni = si
    hj, huj, hvj, zbj, zj, nj = sj
    zero = torch.zeros((), dtype=hi.dtype)
    hi = torch.clamp(hi, min=DRY_TOL)
    hj = torch.clamp(hj, min=DRY_TOL)
    ui, vi = torch.where(hi > DRY_TOL, hui / hi, zero), torch.where(hi > DRY_TOL, hvi / hi, zero)
    uj, vj = torch.where(hj > DRY_TOL, huj / hj, zero), torch.where(hj > DRY_TOL, hvj / hj, zero)
    uni, vni = ui * nx + vi * ny, -ui * ny + vi * nx
    unj, vnj = uj * nx + vj * ny, -uj * ny + vj * nx
    ri, rj = torch.sqrt(hi), torch.sqrt(hj)
    den = ri + rj
    ht = 0.5 * (hi + hj)
    ct = torch.sqrt(g * ht)
    ut = (uni * ri + unj * rj) / den
    vt = (vni * ri + vnj * rj) / den
    l1r, l3r = ut - ct, ut + ct
    ei, ej = uni - torch.sqrt(g * hi), unj - torch.sqrt(g * hj)
    one = torch.ones((), dtype=hi.dtype)
    f1 = (ei < 0.0) & (ej > 0.0)
    d1 = torch.where(f1, ej - ei, one)
    lE1 = torch.where(f1, l1r - ei * (ej - l1r) / d1, zero)
    l1 = torch.where(f1, ei * (ej - l1r) / d1, l1r)
    ei, ej = uni + torch.sqrt(g * hi), unj + torch.sqrt(g * hj)
    f3 = (ei < 0.0) & (ej > 0.0)
    d3 = torch.where(f3, ej - ei, one)
    lE3 = torch.where(f3, l3r - ej * (l3r - ei) / d3, zero)
    l3 = torch.where(f3, ej * (l3r - ei) / (d3 + VEL_TOL), l3r)
    dh = hj - hi
    dhu = hj * unj - hi * uni
    dhv = hj * vnj - hi * vni
    d2 = -2.0 * ct
    a1 = (-l3r * dh + dhu) / d2
    a2 = (-vt * dh + dhv) / ct
    a3 = (l1r * dh - dhu) / d2
    zbi = torch.where(zbi > SED_TOL, zbi, torch.full_like(zbi, SED_TOL * 0.01))
    zbj = torch.where(zbj > SED_TOL, zbj, torch.full_like(zbj, SED_TOL * 0.01))
    thrust = -g * ht * ((zj + zbj) - (zi + zbi))
    nt = 0.5 * (ni + nj)
    s2 = ut * ut + vt * vt
    speed = torch.where(s2 > 0.0, torch.sqrt(torch.where(s2 > 0.0, s2, one)), zero)
    sf = (nt * nt * speed * ut) / torch.clamp(ht ** (4 / 3), min=DRY_TOL)
    b1 = (thrust - g * ht * sf * dx) / d2
    ones = torch.ones_like(vt)
    zeros = torch.zeros_like(vt)
    E = ((ones, l1, vt), (zeros, zeros, ct), (ones, l3, vt))
    lam, lamE = (l1, ut, l3), (lE1, zeros, lE3)
    alpha, beta = (a1, a2, a3), (b1, zeros, -b1)
    P = [zeros, zeros, zeros]
    M = [zeros, zeros, zeros]
    for k in range(3):
        pos, neg = lam[k] > 0.0, lam[k] <= 0.0
        w = lam[k] * alpha[k] - beta[k]
        for c in range(3):
            f = w * E[k][c]
            P[c] = P[c] + torch.where(pos, f, zero)
            M[c] = M[c] + torch.where(neg, f, zero)
    for k in range(3):
        pos, neg = lamE[k] > 0.0, lamE[k] <= 0.0
        w = lamE[k] * alpha[k]
        for c in range(3):
            f = w * E[k][c]
            P[c] = P[c] + torch.where(pos, f, zero)
            M[c] = M[c] + torch.where(neg, f, zero)
    back = lambda u: (u[0], nx * u[1] + (-ny) * u[2], ny * u[1] + nx * u[2])  # noqa: E731
    return back(P), back(M)


def roe_solve_2d(s, m0, dt, dx):
    h = s["h"]
    flds = [s[k] for k in ("h", "hu", "hv", "z_b", "z", "n_b")]
    F = [torch.zeros_like(h) for _ in range(3)]
    zero = torch.zeros((), dtype=h.dtype)
    for axis, (nx, ny) in (("x", (1, 0)), ("y", (0, -1))):
        pairs = [_xs(f, axis) for f in flds]
        P, M = _roe_edge(tuple(p[0] for p in pairs), tuple(p[1] for p in pairs), nx, ny, dx)
        hi, hj = pairs[0]
        mi, mj = _xs(m0, axis)
        act = ((hi > DRY_TOL) | (hj > DRY_TOL)) & ~(mi | mj)
        for c in range(3):
            # F[i] += M on the edge's first cell, F[j] += P on its second
            # (written with padding, not in-place slices, for autograd)
            F[c] = F[c] + _to_lo(torch.where(act, M[c], zero), axis)
            F[c] = F[c] + _to_hi(torch.where(act, P[c], zero), axis)
    out = dict(s)
    for c, k in enumerate(("h", "hu", "hv")):
        out[k] = s[k] - dt * F[c] / dx
    return out


def momentum_corrections(s, m0):
    h, zt = s["h"], s["z"] + s["z_b"]
    C, W, E, N, S = ((slice(1, -1), slice(1, -1)), (slice(1, -1), slice(0, -2)), (slice(1, -1), slice(2, None)),
                     (slice(0, -2), slice(1, -1)), (slice(2, None), slice(1, -1)))
    eta = h[C] + zt[C]
    stop_u = ((eta < zt[E]) & (h[E] < DRY_TOL)) | ((eta < zt[W]) & (h[W] < DRY_TOL)) | m0[E] | m0[W]
    stop_v = ((eta < zt[N]) & (h[N] < DRY_TOL)) | ((eta < zt[S]) & (h[S] < DRY_TOL)) | m0[N] | m0[S]
    act = (h[C] > DRY_TOL) & ~m0[C]
    out = dict(s)
    zero = torch.zeros((), dtype=h.dtype)
    for k, stop in (("hu", stop_u), ("hv", stop_v)):
        q = s[k].clone()
        q[C] = torch.where(act & ~stop, s[k][C], zero)
        out[k] = q
    return out


def compute_g(s, m0, seds, constant):
    h = s["h"]
    zero = torch.zeros((), dtype=h.dtype)
    if constant:
        G = torch.where(m0, zero, s["G"])
        return torch.where(h > DRY_TOL, G, zero)
    hc = torch.clamp(h, min=DRY_TOL)
    u, v = s["hu"] / hc, s["hv"] / hc
    vel2 = u * u + v * v
    frac, d50, dens, theta_c, por = (_rows(seds[i]) for i in (0, 1, 2, 5, 6))
    cbrt_h = hc ** (1.0 / 3.0)
    sqrt_h = torch.sqrt(hc)
    G = torch.zeros_like(h)
    one = torch.ones((), dtype=h.dtype)
    for k in range(len(frac)):
        sk = dens[k] / 1000.0
        npk = d50[k] ** (1 / 6) / 21.1
        tp = torch.clamp((npk * npk * vel2) / ((sk - 1.0) * d50[k] * cbrt_h), min=0.0)
        dth = torch.clamp(tp - theta_c[k], min=0.0)
        g1 = npk * npk * npk * np.sqrt(g) / ((sk - 1.0) * sqrt_h)
        # sqrt(dth) at dth = 0 only feeds a term multiplied by dth = 0
        g2 = 8.0 * torch.sqrt(torch.where(dth > 0.0, dth, one)) / (tp ** 1.5 + 1e-5)
        G = G + frac[k] * g1 * g2 * dth
    G = G / (1.0 - _mean(por))
    G = torch.where(m0, zero, G)
    return torch.where(hc > DRY_TOL, G, zero)


def exner_solve_2d(s, m0, dt, dx):
    h, zb = s["h"], s["z_b"]
    fl = torch.zeros_like(h)
    zero = torch.zeros((), dtype=h.dtype)
    for axis, (nx, ny) in (("x", (1, 0)), ("y", (0, -1))):
        hi, hj = _xs(h, axis)
        hui, huj = _xs(s["hu"], axis)
        hvi, hvj = _xs(s["hv"], axis)
        zbi, zbj = _xs(zb, axis)
        gi, gj = _xs(s["G"], axis)
        mi, mj = _xs(m0, axis)
        hi2 = torch.where(hi > DRY_TOL, hi, torch.full_like(hi, DRY_TOL * 0.01))
        hj2 = torch.where(hj > DRY_TOL, hj, torch.full_like(hj, DRY_TOL * 0.01))
        zi2 = torch.where(zbi > SED_TOL, zbi, torch.full_like(zbi, SED_TOL * 0.01))
        zj2 = torch.where(zbj > SED_TOL, zbj, torch.full_like(zbj, SED_TOL * 0.01))
        ri, rj = torch.sqrt(hi2), torch.sqrt(hj2)
        ui, vi = torch.where(hi2 > DRY_TOL, hui / hi2, zero), torch.where(hi2 > DRY_TOL, hvi / hi2, zero)
        uj, vj = torch.where(hj2 > DRY_TOL, huj / hj2, zero), torch.where(hj2 > DRY_TOL, hvj / hj2, zero)
        uni, unj = ui * nx + vi * ny, uj * nx + vj * ny
        ut = (uni * ri + unj * rj) / (ri + rj)
        mi2, mj2 = ui * ui + vi * vi, uj * uj + vj * vj
        gt = 0.5 * (gi + gj)
        dz = zj2 - zi2
        dq = gt * mj2 * unj - gt * mi2 * uni
        big = dz.abs() > SED_TOL
        l4 = torch.where(big, dq / torch.where(big, dz, torch.ones_like(dz)), ut)
        qi = gi * mi2 * uni + (gt - gi) * mi2 * uni
        qj = gj * mj2 * unj + (gt - gj) * mj2 * unj
        F = torch.where(l4 >= 0, qi, qj)
        stop = (mi | mj) | ((zbi < SED_TOL) & (zbj < SED_TOL))
        F = torch.where(((hi > DRY_TOL) | (hj > DRY_TOL)) & ~stop, F, zero)
        fl = fl + _to_lo(F, axis)
        fl = fl - _to_hi(F, axis)
    out = dict(s)
    out["z_b"] = zb - dt * fl / dx
    return out


class Fill:
    """The ring fill of the oracle's Boundaries (index pairs computed there)
    for the Reflective, Transmissive, TransmissiveBed and Periodic kinds."""

    def __init__(self, oracle_bcs):
        self.ops = []
        for hd in oracle_bcs.handlers:
            b = tuple(torch.as_tensor(np.asarray(a), dtype=torch.long) for a in hd.b)
            i = tuple(torch.as_tensor(np.asarray(a), dtype=torch.long) for a in hd.i)
            if hd.kind not in ("Reflective", "Transmissive", "TransmissiveBed", "Periodic"):
                raise NotImplementedError(f"torch CPU fill of {hd.kind}")
            self.ops.append((hd.kind, b, i, (float(hd.normal[0]), float(hd.normal[1]))))

    def __call__(self, s):
        s = dict(s)
        for kind, b, i, (nx, ny) in self.ops:
            def copy(k):
                a = s[k].clone()
                a[b] = s[k][i]
                s[k] = a
            if kind == "Reflective":
                hu, hv = s["hu"][i], s["hv"][i]
                mn, mt = hu * nx + hv * ny, -hu * ny + hv * nx
                for k in ("h", "z_b", "n", "G"):
                    copy(k)
                a, c = s["hu"].clone(), s["hv"].clone()
                a[b] = -mn * nx - mt * ny
                c[b] = -mn * ny + mt * nx
                s["hu"], s["hv"] = a, c
            elif kind == "Transmissive":
                for k in ("h", "hu", "hv"):
                    copy(k)
            elif kind == "TransmissiveBed":
                copy("z_b")
            else:
                for k in FIELDS:
                    copy(k)
        return s


def step(s, m0, fill, seds, constant, dt, dx):
    s = roe_solve_2d(s, m0, dt, dx)
    s = momentum_corrections(s, m0)
    s = fill(s)
    s["G"] = compute_g(s, m0, seds, constant)
    s = exner_solve_2d(s, m0, dt, dx)
    return fill(s)


def ssprk2_step(s, m0, fill, seds, constant, dx, cfl):
    dt = dt_2d(s, m0, dx, cfl)
    s1 = step(s, m0, fill, seds, constant, dt, dx)
    s2 = step(s1, m0, fill, seds, constant, dt, dx)
    return {k: 0.5 * s[k] + 0.5 * s2[k] for k in FIELDS}, dt


def from_oracle_state(st, dtype):
    return {k: torch.as_tensor(np.ascontiguousarray(getattr(st, k)), dtype=dtype) for k in FIELDS}

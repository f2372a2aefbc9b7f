"""Gradients of a Roe-Exner SSPRK2 run by hand-written adjoint kernels.

`ExnerGradientRun` advances a case (mesh, arrays, YAML params) for a fixed
number of SSPRK2 steps in float64 and differentiates the final state with
respect to the initial erodible bed z_b0 and the sediment's d50 and porosity.
The forward pass keeps a checkpoint (h, hu, hv, z_b, G, dt) every
`checkpoint_every` steps; the backward pass replays one segment at a time from
its checkpoint, keeping that segment's stage states, and runs the adjoint of
each step in reverse (pxh_adj_roe_exner_ssprk2), then the adjoint of init_fn
(pxh_adj_roe_exner_init).  Memory: one checkpoint per segment plus twelve
fields per step of one segment.

The reference has no gradient code; these are the derivatives jax.grad would
give through its init_fn, compute_dt_fn and step_fn
(solvers/roe_exner_solver.py:58-122) chained by integrators/SSPRK2.py:32-40,
with untaken branches contributing zero.  The erosion and deposition
coefficients k_e, k_d (seds rows 3 and 4) never enter the path
(roe_exner.py:85-130), so their gradients are zero; d50 and porosity are the
sediment parameters the bed responds to.

    run = ExnerGradientRun(mesh, arrays, params, n_steps=200, checkpoint_every=20)
    h, hu, hv, z_b = run(z_b0, d50, porosity)      # torch.autograd.Function
    loss = 0.5 * ((z_b - target) ** 2).sum()
    loss.backward()                                 # z_b0.grad, d50.grad, porosity.grad
"""
import numpy as np
import torch

from . import _hip
from .constants import g
from .runtime.driver import setup_case
from .solvers.roe_exner_solver import sediment_table
from .state import fields_struct

EVOLVING = ("h", "hu", "hv", "z_b", "G")
OUTPUTS = ("h", "hu", "hv", "z_b")


def _stream():
    return torch.cuda.current_stream().cuda_stream


class ExnerGradientRun:
    """A Roe-Exner SSPRK2 run of n_steps (float64, MPM G) whose final state is
    differentiable with respect to z_b0 (pre-init bed), d50 and porosity of the
    (single) sediment class."""

    def __init__(self, mesh, arrays, params, n_steps, checkpoint_every=20, cfl=None, device="cuda"):
        params = dict(params, backend=dict(params.get("backend") or {}, fused_stage=True))
        state, bm, solver, sc, _, ic = setup_case(mesh, arrays, params, dtype=torch.float64, device=device)
        if sc.compute_G != "MPM":
            raise NotImplementedError("gradients need the MPM transport coefficient (grass_factor: MPM)")
        if bm.n_entries and not bm.idempotent:
            raise NotImplementedError("the adjoint ring fill needs a program whose targets are never sources")
        self.be = sc.backend
        self.solver, self.sc = solver, sc
        self.mask = solver.mask_fn(state, sc.mpi_handler, bm.boundary_mask)
        self.m0 = self.mask[0].data_ptr()
        # the forward sweeps' static-field promises (no nodata cell, a fixed
        # bed z of one value), checked once: z is static through init_fn
        self.no_nodata = not bool(self.mask[0].any())
        self.z_uniform = self.no_nodata and bool((state.z == state.z.reshape(-1)[0]).all())
        self.raw = state  # before init_fn
        self.seds = np.array(state.seds, dtype=np.float64)
        if self.seds.shape[1] != 1:
            raise NotImplementedError("gradients with respect to sediment parameters need one sediment class")
        self.cfl = float(params.get("cfl", 0.5) if cfl is None else cfl)
        self.n_steps = int(n_steps)
        self.every = max(1, int(checkpoint_every))
        self.device = device
        # the ring-fill programs of the adjoint: sorted by source, so that the
        # transposed fill sums in a fixed order and gradients are reproducible
        self.bc_all = bm.adjoint_program()
        self.bc_step = bm.adjoint_program(self.be.STEP_GROUPS)
        lib = _hip.load()
        ws = lib.pxh_adj_workspace_size(self.be.grid, self.bc_all)
        if ws < 0:
            raise ValueError("invalid grid for the adjoint")
        self.adj_ws = torch.empty(int(ws), dtype=torch.uint8, device=device)
        self.checkpoints = []
        self.s0 = None

    # -- forward ---------------------------------------------------------------
    def _set_sediment(self, d50, porosity):
        seds = self.seds.copy()
        seds[1, 0] = float(d50)
        seds[6, 0] = float(porosity)
        self.cur_seds = seds
        self.be.sed = sediment_table(seds, False, np.float64)

    def _init(self, z_b0):
        s = self.raw.clone().replace(seds=self.cur_seds)
        s.z_b.copy_(z_b0.reshape(s.z_b.shape))
        return self.solver.init_fn(s, self.mask, self.sc)

    def _buffers(self, like):
        return like.replace(**{n: torch.empty_like(getattr(like, n)) for n in EVOLVING})

    def _step(self, fa, fb, fc, dt):
        self.be.ssprk2_into(fa, fb, fc, self.m0, self.cfl, dt.data_ptr(), _stream(), dt_ready=True,
                            no_nodata=self.no_nodata, z_uniform=self.z_uniform)

    def forward(self, z_b0, d50, porosity):
        """Run n_steps from z_b0; returns the final state and keeps the
        checkpoints for backward()."""
        self._set_sediment(d50, porosity)
        self.s0 = self._init(z_b0)
        a = self.s0.clone()
        b, c = self._buffers(a), self._buffers(a)
        fa, fb, fc = fields_struct(a), fields_struct(b), fields_struct(c)
        dt = torch.empty((), dtype=torch.float64, device=self.device)
        self.be.dt_into(fa, self.m0, self.cfl, dt.data_ptr(), _stream())
        self.checkpoints = []
        # every step's dt, for the segment replays (a dt pass fewer per step)
        self.dts = torch.empty(self.n_steps, dtype=torch.float64, device=self.device)
        for n in range(self.n_steps):
            if n % self.every == 0:
                self.checkpoints.append({k: getattr(a, k).clone() for k in EVOLVING})
            self.dts[n].copy_(dt)
            self._step(fa, fb, fc, dt)
        self.final = a
        return a

    # -- backward --------------------------------------------------------------
    def _pool(self):
        """Buffers of one replayed segment: a for each of its steps and the
        state after the last, b and c for each step."""
        k = self.every
        return ([self._buffers(self.s0) for _ in range(k + 1)], [self._buffers(self.s0) for _ in range(k)],
                [self._buffers(self.s0) for _ in range(k)])

    def _replay(self, seg, pool=None):
        """States of the steps of segment seg: per step (a, b, c, dt) with b
        and c the full outputs of the two stages (the fused step keeps only
        the frame of c), recomputed stage by stage from the segment's
        checkpoint with the forward's dts: b = S(a), c = S(b),
        a' = 0.5 a + 0.5 c.  pool (from _pool): the buffers to write, else
        new ones.  Runs on the current stream."""
        n0 = seg * self.every
        n1 = min(self.n_steps, n0 + self.every)
        if pool is None:
            m = n1 - n0
            pool = ([self._buffers(self.s0) for _ in range(m + 1)], [self._buffers(self.s0) for _ in range(m)],
                    [self._buffers(self.s0) for _ in range(m)])
        aa, bb, cc = pool
        for k, v in self.checkpoints[seg].items():
            getattr(aa[0], k).copy_(v)
        out = []
        st = _stream()
        for j, n in enumerate(range(n0, n1)):
            dt = self.dts[n]
            a, b, c, nxt = aa[j], bb[j], cc[j], aa[j + 1]
            fa, fb, fc = fields_struct(a), fields_struct(b), fields_struct(c)
            self.be.step_into(fa, fb, self.m0, dt.data_ptr(), st, no_nodata=self.no_nodata, z_uniform=self.z_uniform)
            self.be.step_into(fb, fc, self.m0, dt.data_ptr(), st, no_nodata=self.no_nodata, z_uniform=self.z_uniform)
            self.be.combine_into(fa, fc, fields_struct(nxt), st)
            out.append((a, b, c, dt))
        return out

    def backward(self, lam):
        """lam: dict of adjoints of the final (h, hu, hv, z_b, G) (missing =
        zero).  Returns (dL/dz_b0, dL/dd50, dL/dporosity) with the adjoint of
        the initialised state in self.lam_init.

        The replay of segment s - 1 needs only its checkpoint, so it runs on
        a second stream while the adjoint steps of segment s run on the
        current one (two buffer pools, alternately; events order each replay
        after the adjoint that last read its pool, and each adjoint after its
        replay): on a grid too small to fill the GPU the two overlap."""
        shape = self.s0.h.shape
        z = lambda: torch.zeros(shape, dtype=torch.float64, device=self.device)  # noqa: E731
        L = self.s0.replace(**{k: (lam[k].detach().clone().to(torch.float64).contiguous() if lam.get(k) is not None
                                   else z()) for k in EVOLVING}, n_b=z())
        fl = fields_struct(L)
        lam_par = torch.zeros(8, dtype=torch.float64, device=self.device)
        ws = self.adj_ws.data_ptr()
        nseg = len(self.checkpoints)
        main = torch.cuda.current_stream()
        side = torch.cuda.Stream()
        pools = [self._pool() for _ in range(min(nseg, 2))]
        done = [None, None]  # per pool: the adjoint's last read of it
        pending = {}

        def replay_on_side(seg):
            k = seg % 2
            if done[k] is not None:
                side.wait_event(done[k])
            else:
                side.wait_stream(main)
            with torch.cuda.stream(side):
                steps = self._replay(seg, pools[k])
            ev = torch.cuda.Event()
            ev.record(side)
            pending[seg] = (steps, ev)

        replay_on_side(nseg - 1)
        for seg in reversed(range(nseg)):
            if seg > 0:
                replay_on_side(seg - 1)
            steps, ev = pending.pop(seg)
            main.wait_event(ev)
            for a_n, b_n, c_n, dt_n in reversed(steps):
                _hip.call("pxh_adj_roe_exner_ssprk2", self.be.grid, fields_struct(a_n), fields_struct(b_n),
                          fields_struct(c_n), self.m0, self.be.sed, self.bc_step, self.cfl, dt_n.data_ptr(), fl,
                          lam_par.data_ptr(), ws, _stream())
            done[seg % 2] = torch.cuda.Event()
            done[seg % 2].record(main)
        _hip.call("pxh_adj_roe_exner_init", self.be.grid, fields_struct(self.s0), self.m0, self.be.sed, self.bc_all,
                  fl, lam_par.data_ptr(), ws, _stream())
        self.lam_init = L
        self.lam_par = lam_par[:5].cpu().numpy()
        d_d50, d_por = chain_sediment_gradients(self.cur_seds, self.lam_par)
        return L.z_b.clone(), d_d50, d_por

    def __call__(self, z_b0, d50, porosity):
        d50 = torch.as_tensor(d50, dtype=torch.float64)
        porosity = torch.as_tensor(porosity, dtype=torch.float64)
        return _ExnerRunFn.apply(z_b0, d50, porosity, self)


def chain_sediment_gradients(seds, lam_par):
    """dL/dd50 and dL/dporosity of a one-class sediment table seds (7, 1) from
    lam_par, the derivatives with respect to the table entries A = np^2,
    B = (s-1) d50, C = np^3 sqrt(g), one_minus_por = 1 - porosity and
    nb_sed = (d50 frac)^(1/6) / 21.1, where np = d50^(1/6) / 21.1 and
    s = density / 1000 (roe_exner.py:64-116)."""
    frac, d50, density = float(seds[0, 0]), float(seds[1, 0]), float(seds[2, 0])
    lp = np.asarray(lam_par, dtype=np.float64)
    s = density / 1000.0
    np_ = d50 ** (1.0 / 6.0) / 21.1
    A, C = np_ * np_, np_ ** 3 * np.sqrt(g)
    d50m = d50 * frac
    nb = d50m ** (1.0 / 6.0) / 21.1
    d_d50 = (lp[0] * A / (3.0 * d50) + lp[1] * (s - 1.0) + lp[2] * C / (2.0 * d50)
             + lp[4] * nb * frac / (6.0 * d50m))
    return float(d_d50), float(-lp[3])


class _ExnerRunFn(torch.autograd.Function):
    """(z_b0, d50, porosity) -> final (h, hu, hv, z_b) of the run."""

    @staticmethod
    def forward(ctx, z_b0, d50, porosity, run):
        with torch.no_grad():
            a = run.forward(z_b0.detach().to(torch.float64), float(d50), float(porosity))
        ctx.run = run
        ctx.z_dtype, ctx.d_dtype, ctx.p_dtype = z_b0.dtype, d50.dtype, porosity.dtype
        ctx.p_dev = (d50.device, porosity.device)
        return tuple(getattr(a, k).clone() for k in OUTPUTS)

    @staticmethod
    def backward(ctx, gh, ghu, ghv, gzb):
        run = ctx.run
        lam = dict(zip(OUTPUTS, (gh, ghu, ghv, gzb)))
        with torch.no_grad():
            dz, dd, dp = run.backward(lam)
        return (dz.to(ctx.z_dtype), torch.tensor(dd, dtype=ctx.d_dtype, device=ctx.p_dev[0]),
                torch.tensor(dp, dtype=ctx.p_dtype, device=ctx.p_dev[1]), None)


def seeded_target(z_b0, seed=6, amp=1e-3):
    """The target bed of the calibration case: z_b0 plus a seeded smooth
    perturbation of amplitude amp."""
    rng = np.random.default_rng(seed)
    ny, nx = z_b0.shape
    y = np.linspace(0.0, 1.0, ny)[:, None]
    x = np.linspace(0.0, 1.0, nx)[None, :]
    k = rng.uniform(1.0, 4.0, size=4)
    ph = rng.uniform(0.0, 2 * np.pi, size=4)
    p = amp * (np.sin(2 * np.pi * k[0] * x + ph[0]) * np.cos(2 * np.pi * k[1] * y + ph[1])
               + 0.5 * np.sin(2 * np.pi * k[2] * (x + y) + ph[2]))
    return z_b0 + torch.as_tensor(p, dtype=z_b0.dtype, device=z_b0.device)


__all__ = ["ExnerGradientRun", "chain_sediment_gradients", "seeded_target"]

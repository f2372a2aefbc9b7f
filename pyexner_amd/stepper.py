"""Allocation-free time stepping on preallocated device buffers.

The functional bundle API (step_fn returns new tensors) matches the reference;
this module is the throughput path under it: buffers for u^n, u^1 and u^2 are
allocated once, the dt stays in a device scalar, and one SSPRK2 step
(dt -> stage 1 -> stage 2 -> combine, integrators/SSPRK2.py:32-40) or one
Forward-Euler step (integrators/forwardeuler.py:38-49) is a fixed sequence of
C-ABI calls on one stream, which can be captured into a hipGraph.
"""
import torch

from .state import fields_struct


def _like(state, names):
    kw = {n: torch.empty_like(getattr(state, n)) for n in names}
    return state.replace(**kw)


def _wide_h(state):
    """state with h moved to the front of a 4 ny nx buffer: the room the
    packed first-stage records (PXH_PACKED_STAGE) need when the buffer
    serves as b; Forward-Euler steps, which swap a and b, read the first
    ny nx elements as h."""
    h = state.h
    wide = torch.empty(4 * h.numel(), dtype=h.dtype, device=h.device)
    v = wide[:h.numel()].view(h.shape)
    v.copy_(h)
    return state.replace(h=v)


EVOLVING = ("h", "hu", "hv", "z_b", "G")


def record_fields(state):
    """pxh_fields of state with h, hu, hv, z_b as the view of the four-element
    records that PXH_PACKED_STAGE keeps at state.h (a _wide_h buffer): hu, hv,
    z_b one, two and three elements past h.  The separate stage / frame sweep
    entries and pxh_bc_apply take it where the fused step keeps its first
    stage's output."""
    f = fields_struct(state)
    es = state.h.element_size()
    f.hu, f.hv, f.z_b = f.h + es, f.h + 2 * es, f.h + 3 * es
    return f


class ExnerStepper:
    """Roe-Exner stepping on an initialised state (init_fn already applied)."""

    def __init__(self, backend, state, mask, cfl):
        self.be = backend
        self.cfl = float(cfl)
        self.mask = mask
        self.m0 = mask[0].data_ptr()
        # no nodata cell: the fused stage sweeps need not load the mask; a
        # fixed bed z of one value (a static field): they need not stream it
        self.no_nodata = not bool(mask[0].any())
        self.z_uniform = bool((state.z == state.z.reshape(-1)[0]).all())
        self.a = state.clone()
        names = EVOLVING + (("z_b_lo",) if state.z_b_lo is not None else ())
        self.b = _like(self.a, names)
        # float32: b's h holds the first stage's output as h, hu, hv, z_b
        # records (one stream into the second stage instead of four); a gets
        # the same room since Forward-Euler steps swap a and b
        self.packed = state.h.dtype == torch.float32
        if self.packed:
            self.a, self.b = _wide_h(self.a), _wide_h(self.b)
        self.c = _like(self.a, names)
        self.fa, self.fb, self.fc = (fields_struct(s) for s in (self.a, self.b, self.c))
        self.dt = torch.zeros((), dtype=state.h.dtype, device=state.h.device)
        self.dt_ready = False  # self.dt holds compute_dt_fn(self.a)
        self.time = 0.0
        self.graph = None

    @property
    def state(self):
        return self.a

    def _stream(self):
        return torch.cuda.current_stream().cuda_stream

    def compute_dt(self):
        self.be.dt_into(self.fa, self.m0, self.cfl, self.dt.data_ptr(), self._stream())
        self.dt_ready = True

    def ssprk2_step(self, time=None):
        """One SSPRK2 step; afterwards self.dt holds the dt of the new state.
        time: the step's start time for time-dependent value boundaries
        (default self.time; the stepper does not read dt back, so a caller
        with such boundaries advances self.time itself)."""
        self.be.ssprk2_into(self.fa, self.fb, self.fc, self.m0, self.cfl, self.dt.data_ptr(), self._stream(),
                            dt_ready=self.dt_ready, time=self.time if time is None else time,
                            states=(self.a, self.b, self.c), no_nodata=self.no_nodata,
                            z_uniform=self.z_uniform, packed=self.packed)
        self.dt_ready = True

    def fe_step(self):
        st = self._stream()
        dtp = self.dt.data_ptr()
        self.be.dt_into(self.fa, self.m0, self.cfl, dtp, st)
        self.be.step_into(self.fa, self.fb, self.m0, dtp, st, time=self.time, states=(self.a, self.b),
                          no_nodata=self.no_nodata, z_uniform=self.z_uniform)
        self.a, self.b = self.b, self.a
        self.fa, self.fb = self.fb, self.fa
        self.dt_ready = False

    def capture(self, nsteps=1, integrator="SSPRK2"):
        """Record nsteps steps into a hipGraph (FE needs an even count so the
        buffers return to their starting roles)."""
        if integrator != "SSPRK2" and nsteps % 2:
            raise ValueError("Forward-Euler capture needs an even number of steps")
        bm = self.be.boundaries
        if bm is not None and bm.time_dependent:
            # pxh_bc (and its time) is a kernel argument: every replay would
            # fill a moving bed at the time of capture
            raise NotImplementedError("hipGraph capture of a ring fill whose values move with time")
        fn = self.ssprk2_step if integrator == "SSPRK2" else self.fe_step
        s = torch.cuda.Stream()
        s.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(s):
            fn()  # warm-up outside capture; leaves dt_ready set for the captured SSPRK2 steps
        torch.cuda.current_stream().wait_stream(s)
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            for _ in range(nsteps):
                fn()
        self.graph = g
        return g

    def replay(self):
        self.graph.replay()

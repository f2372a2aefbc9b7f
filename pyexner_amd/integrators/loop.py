"""Helpers shared by the integrators' output-window loops.

The dt of every step is produced on the device; the host reads it through a
side stream while the step it feeds is already running, so the GPU is not
left idle between steps.
"""
import math

import numpy as np
import torch

from .. import _hip
from ..state import fields_struct


class DtReader:
    """Copies a device scalar to pinned host memory on a side stream."""

    def __init__(self, dtype, device):
        self.host = torch.empty((), dtype=dtype, pin_memory=True)
        self.side = torch.cuda.Stream(device=device)
        self.ready = torch.cuda.Event()
        self.done = torch.cuda.Event()

    def start(self, dt_dev):
        self.ready.record(torch.cuda.current_stream())
        with torch.cuda.stream(self.side):
            self.side.wait_event(self.ready)
            self.host.copy_(dt_dev, non_blocking=True)
            dt_dev.record_stream(self.side)
            self.done.record(self.side)

    def value(self):
        self.done.synchronize()
        return float(self.host.item())


class NonFiniteState(FloatingPointError):
    """A NaN or infinite dt, or state values, met by an output-window loop."""


class FiniteWatch:
    """Failure detection the reference lacks (its only guard is the dt clamp,
    solvers/roe_exner_solver.py:121, and a NaN dt simply ends its loop): each
    step's dt, already read on the host, must be finite, and at the end of
    every output window pxh_count_nonfinite scans the state (h, hu, hv, z_b,
    G) once, read back with the window.  Either failure raises
    NonFiniteState.  A backend with check_finite = False is not watched."""

    def __init__(self, backend, device):
        self.on = backend is not None and hasattr(backend, "grid") and getattr(backend, "check_finite", True)
        self.grid = backend.grid if self.on else None
        self.count = torch.zeros((), dtype=torch.int64, device=device) if self.on else None
        self.windows = 0

    def dt(self, d, step, time):
        if self.on and not math.isfinite(d):
            raise NonFiniteState(f"dt = {d} for step {step} at t = {time:.9g}")

    def window(self, state, time):
        if not self.on:
            return
        _hip.call("pxh_count_nonfinite", self.grid, fields_struct(state), self.count.data_ptr(),
                  torch.cuda.current_stream().cuda_stream)
        n = int(self.count.item())
        self.windows += 1
        if n:
            raise NonFiniteState(f"{n} cells hold NaN or infinite values at t = {time:.9g} "
                                 f"(end of output window {self.windows})")


def next_target(time, out_freq, end_time):
    """forwardeuler.py:33-34: min((int(time/out_freq) + 1) * out_freq, end_time)."""
    return min((int(np.int32(time / out_freq)) + 1) * out_freq, end_time)


class DtRing:
    """The dt each step of an ExnerStepper leaves on the device (the fused
    reduction's next dt), published to the host without stalling the step
    stream: a 4-byte copy of it straight into a pinned host slot on the step
    stream (part of the step's hipGraph), and an event after the step that
    the host waits on before reading the slot.  The host reads slot k before
    it queues the step that writes slot k + 1, so n slots are never
    overwritten unread."""

    def __init__(self, dt_dev, n=8):
        self.dt = dt_dev
        self.n = n
        self.host = torch.zeros(n, dtype=dt_dev.dtype, pin_memory=True)
        self.host_v = [self.host[k] for k in range(n)]
        self.host_np = self.host.numpy()
        self.done = [torch.cuda.Event() for _ in range(n)]

    def before(self, k):
        """Call on the step stream before the step whose dt goes to slot k
        (nothing to wait for: the slot was read before this step was queued)."""

    def stage(self, k):
        """The device part of publishing: copy dt into host slot k (captured
        into a step's graph, or issued eagerly after the step)."""
        self.host_v[k % self.n].copy_(self.dt, non_blocking=True)

    def publish(self, k):
        """The host part, after slot k was staged on the step stream."""
        self.done[k % self.n].record()

    def get(self, k):
        """The dt published from slot k (waits for the step that staged it)."""
        s = k % self.n
        self.done[s].synchronize()
        return float(self.host_np[s])


def run_exner_stepper(state, mask, backend, config, io, mesh, integrator, graphs=True, autotune=None):
    """The window loop of integrators/SSPRK2.py:53-97 (or forwardeuler.py:51-117)
    on an ExnerStepper: preallocated buffers, the fused SSPRK2 entry with its
    next dt, the autotuned strip height, and each regular step replayed from a
    hipGraph.

    The loop keeps the reference's decisions: a step is taken with the CFL dt
    while time + (the previous step's dt) < next output time - TIMESTEP_TOL,
    then one window-closing step whose dt is clipped to the output time
    (SSPRK2.py:82-87; Forward Euler writes a probe step of the clipped
    length instead and commits the CFL step, forwardeuler.py:91-117).  That
    decision for step j needs the dts of the steps before it only, and a
    regular step needs its own dt on the device only (the fused reduction of
    step j - 1 leaves it there), so the host queues step j, then waits for
    its dt (for step j - 1 to end) and decides step j + 1 while step j runs:
    the GPU always has the next step queued.  Only the window-closing step
    waits for its own input (its dt is the clipped CFL dt of the state before
    it).  Returns (state, iters, time)."""
    from ..autotune import tune_strip_rows
    from ..constants import TIMESTEP_TOL
    from ..stepper import ExnerStepper
    fe = integrator == "Forward Euler"
    if autotune is None:
        autotune = state.h.numel() >= (1 << 20)
    if autotune and not fe:
        tune_strip_rows(backend, state, mask, config.cfl)
    st = ExnerStepper(backend, state, mask, config.cfl)
    watch = FiniteWatch(backend, state.h.device)
    bm = backend.boundaries
    moving = bm is not None and bm.time_dependent
    ring = DtRing(st.dt)
    use_graphs = graphs and not moving and not fe
    g = None
    if use_graphs:
        st.compute_dt()
        keep = st.a.clone()
        keep_dt = st.dt.clone()
        # one eager step on a side stream before capture (library set-up),
        # then the state and dt restored
        s = torch.cuda.Stream()
        s.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(s):
            st.ssprk2_step()
            ring.stage(0)
        torch.cuda.current_stream().wait_stream(s)
        for name in st.a.field_names():
            getattr(st.a, name).copy_(getattr(keep, name))
        st.dt.copy_(keep_dt)
        # graph k: one SSPRK2 step and the copy of its next dt into slot k
        g = []
        for k in range(ring.n):
            gr = torch.cuda.CUDAGraph()
            with torch.cuda.graph(gr):
                st.ssprk2_step()
                ring.stage(k)
            g.append(gr)
        torch.cuda.synchronize()
        del keep, keep_dt

    def fe_dt(j):
        # Forward Euler: the step's dt is compute_dt_fn of its input
        st.compute_dt()
        ring.before(j)
        ring.stage(j)
        ring.publish(j)

    def step(j, t):
        """Launch step j (its dt is in st.dt); publish the dt it leaves."""
        if fe:
            st.be.step_into(st.fa, st.fb, st.m0, st.dt.data_ptr(), st._stream(), time=t, states=(st.a, st.b),
                            no_nodata=st.no_nodata, z_uniform=st.z_uniform)
            st.a, st.b = st.b, st.a
            st.fa, st.fb = st.fb, st.fa
            return
        ring.before(j)
        if g is not None:
            g[j % ring.n].replay()
        else:
            st.ssprk2_step(time=t)
            ring.stage(j)
        ring.publish(j)

    if fe:
        fe_dt(0)
    else:
        st.compute_dt()
        ring.before(0)
        ring.stage(0)
        ring.publish(0)
    # slot j holds the dt of step j + 1 (SSPRK2: left by step j; FE: computed
    # before step j + 1)
    time, d_prev, j = 0.0, 0.0, 0
    end, of = config.end_time, config.out_freq
    while time < end - TIMESTEP_TOL:
        while time + d_prev < next_target(time, of, end) - TIMESTEP_TOL:
            j += 1
            if fe and j > 1:
                fe_dt(j - 1)
            # a regular step needs its dt on the device only: it is queued
            # before the host waits for that dt (which waits for step j - 1),
            # so the GPU has step j in hand when step j - 1 ends
            step(j, time)
            d = ring.get(j - 1)  # the dt of step j
            watch.dt(d, j, time)
            time, d_prev = time + d, d
        # the window-closing step: the CFL dt of the current state, clipped
        target = next_target(time, of, end)
        if fe:
            fe_dt(j)
        d = ring.get(j)
        j += 1
        watch.dt(d, j, time)
        if time + d >= target - TIMESTEP_TOL:
            dc = target - time
            if fe:
                if io is not None:
                    probe = st.b
                    keep_dt = st.dt.clone()
                    st.dt.fill_(dc)
                    st.be.step_into(st.fa, st.fb, st.m0, st.dt.data_ptr(), st._stream(), time=time,
                                    states=(st.a, st.b), no_nodata=st.no_nodata, z_uniform=st.z_uniform)
                    io.write_state(probe, mesh, mask[0])
                    st.dt.copy_(keep_dt)
                step(j, time)
            else:
                st.dt.fill_(dc)
                step(j, time)
                d = dc
                if io is not None:
                    io.write_state(st.state, mesh, mask[0])
            time, d_prev = time + d, d
            watch.window(st.state, time)
            continue
        step(j, time)
        time, d_prev = time + d, d
    torch.cuda.synchronize()
    return st.state, j, time

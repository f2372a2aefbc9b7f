"""Step sequencing and time loops.

* get_mask           - solvers/roe_exner_solver.py:21-39 (one rank: every side physical)
* init_roeexner      - solvers/roe_exner_solver.py:58-78 (halo exchange is a no-op on one rank)
* step_roeexner      - solvers/roe_exner_solver.py:80-116
* dt_roeexner        - solvers/roe_exner_solver.py:118-122
* step_roe / dt_roe  - solvers/roe_solver.py:40-60 (as intended, see roe_solver.py:30,40-41)
* fe_step / ssprk2_step - integrators/forwardeuler.py:37-49, integrators/SSPRK2.py:32-40
* run_forward_euler  - integrators/forwardeuler.py:51-117 (output-window loop with probe steps)
* run_ssprk2         - integrators/SSPRK2.py:53-97 (output-window loop; the closing step is an
                       SSPRK2 step of the clipped dt, not the single step_fn call at :87)
"""
from dataclasses import fields

import numpy as np

from . import roe as roe_k
from . import roe_exner as rx
from .constants import TIMESTEP_TOL


def get_mask(z, b_mask):
    mask = np.isnan(z)
    mask[0, :] = True
    mask[-1, :] = True
    mask[:, 0] = True
    mask[:, -1] = True
    mask = np.where(b_mask, False, mask)
    return np.stack([mask, b_mask], axis=0)


def init_roeexner(state, mask, bcs, constant):
    n_b = rx.compute_n_mode(state.n, state.z_b, state.seds, constant)
    state = state.replace(n_b=n_b)
    G = rx.compute_G_mode(state.h, state.hu, state.hv, state.seds, mask, state.G, constant)
    state = state.replace(G=G)
    return bcs.apply(state)


def step_roeexner(state, dt, mask, bcs, dx, constant, time=0.0):
    state = rx.roe_solve_2D(state, dt, dx, mask)
    state = rx.momentum_corrections(state, mask)
    state = bcs.apply(state, time)
    G = rx.compute_G_mode(state.h, state.hu, state.hv, state.seds, mask, state.G, constant)
    state = state.replace(G=G)
    state = rx.exner_solve_2D(state, dt, dx, mask)
    state = bcs.apply(state, time)
    # n_b recomputed then discarded (roe_exner_solver.py:112-114)
    return state


def dt_roeexner(state, cfl, mask, dx, cubic_dtype=None):
    return rx.compute_dt_roeexner(state, cfl, mask, dx, cubic_dtype)


def step_roe(state, dt, bcs, dx):
    state = roe_k.roe_solve_2D(state, dt, dx)
    return bcs.apply(state)


def dt_roe(state, cfl, dx):
    return roe_k.compute_dt_roe(state, cfl, None, dx)


def combine_half(a, b):
    """tree_map(lambda a, b: 0.5 * a + 0.5 * b) over every field (SSPRK2.py:40)."""
    kw = {f.name: 0.5 * getattr(a, f.name) + 0.5 * getattr(b, f.name) for f in fields(a)}
    return type(a)(**kw)


class ExnerProblem:
    """Bundles mask, boundaries and parameters for the Roe-Exner step."""

    def __init__(self, mask, bcs, dx, cfl, constant, cubic_dtype=None):
        self.mask, self.bcs, self.dx, self.cfl = mask, bcs, dx, cfl
        self.constant, self.cubic_dtype = constant, cubic_dtype

    def dt(self, s):
        return dt_roeexner(s, self.cfl, self.mask, self.dx, self.cubic_dtype)

    def step(self, s, dt, time=0.0):
        return step_roeexner(s, dt, self.mask, self.bcs, self.dx, self.constant, time)


class RoeProblem:
    def __init__(self, bcs, dx, cfl):
        self.bcs, self.dx, self.cfl = bcs, dx, cfl

    def dt(self, s):
        return dt_roe(s, self.cfl, self.dx)

    def step(self, s, dt, time=0.0):
        return step_roe(s, dt, self.bcs, self.dx)


def fe_step(prob, s, time=0.0):
    dt = prob.dt(s)
    return prob.step(s, dt, time), dt


def ssprk2_step(prob, s, time=0.0):
    """Both stages see the step's start time (SSPRK2.py:34-35)."""
    dt = prob.dt(s)
    s1 = prob.step(s, dt, time)
    s2 = prob.step(s1, dt, time)
    return combine_half(s, s2), dt


def run_steps(prob, s, nsteps, integrator="Forward Euler"):
    """nsteps steps from time 0; the boundaries see each step's start time,
    accumulated as time + dt in the state's dtype (forwardeuler.py:43)."""
    f = fe_step if integrator == "Forward Euler" else ssprk2_step
    t = s.h.dtype.type(0.0)
    for _ in range(nsteps):
        s, dt = f(prob, s, t)
        t = s.h.dtype.type(t + dt)
    return s, float(t)


def run_forward_euler(prob, s, end_time, out_freq, on_output=None):
    """forwardeuler.py:51-117 with the device while-loop run on the host.

    The inner loop condition tests the previous step's dt (forwardeuler.py:32-35);
    at each output window a probe step of length next_target - time is taken
    and reported but not committed, then one committed step with the CFL dt.
    """
    dtype = s.h.dtype.type
    time = dtype(0.0)
    dt = dtype(0.0)
    iters = 0
    while time < end_time - TIMESTEP_TOL:
        while True:  # jax.lax.while_loop(cond_fn, body_fn, simstate)
            nxt = (int(np.int32(time / out_freq)) + 1) * out_freq
            target = min(nxt, end_time)
            if not (time + dt < target - TIMESTEP_TOL):
                break
            dt = prob.dt(s)
            s = prob.step(s, dt, time)
            time = dtype(time + dt)
            iters += 1
        dt = prob.dt(s)
        nxt = (int(time / out_freq) + 1) * out_freq
        target = min(nxt, end_time)
        if time + dt >= target - TIMESTEP_TOL:
            probe = prob.step(s, dtype(target - time), time)
            if on_output is not None:
                on_output(target, probe)
        s = prob.step(s, dt, time)
        time = dtype(time + dt)
        iters += 1
    return s, float(time), iters


def run_ssprk2(prob, s, end_time, out_freq, on_output=None):
    """SSPRK2.py:53-97 with the device while-loop (cond_fn :23-28, body_fn
    :32-49) run on the host.  The window-closing step (:79-87) takes the CFL
    dt clipped to the output time and is a whole SSPRK2 step; a clipped step
    is reported to on_output (the reference writes no output).  Time is
    accumulated in Python float."""
    time, dt, iters = 0.0, 0.0, 0
    while time < end_time - TIMESTEP_TOL:
        while True:
            target = min((int(np.int32(time / out_freq)) + 1) * out_freq, end_time)
            if not (time + dt < target - TIMESTEP_TOL):
                break
            dt = float(prob.dt(s))
            s1 = prob.step(s, dt, time)
            s = combine_half(s, prob.step(s1, dt, time))
            time = time + dt
            iters += 1
        target = min((int(time / out_freq) + 1) * out_freq, end_time)
        dt = float(prob.dt(s))
        clipped = time + dt >= target - TIMESTEP_TOL
        if clipped:
            dt = target - time
        s1 = prob.step(s, dt, time)
        s = combine_half(s, prob.step(s1, dt, time))
        if clipped and on_output is not None:
            on_output(target, s)
        time = time + dt
        iters += 1
    return s, time, iters

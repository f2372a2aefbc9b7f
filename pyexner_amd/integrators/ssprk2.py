"""SSPRK2 integrator (integrators/SSPRK2.py).

The reference module cannot be imported (it is not imported by
integrators/__init__.py:1, references undefined names at SSPRK2.py:42,53 and
calls the bundles with the wrong arity at :33,:36-37).  The step is its
documented intent (SSPRK2.py:32-40): one dt from u^n, u1 = S(u^n),
u2 = S(u1), u^{n+1} = 0.5 u^n + 0.5 u2 over every field.  The window loop
follows SSPRK2.py:53-97 with one change: the window-closing step, clipped to
the output time (SSPRK2.py:82-85), is an SSPRK2 step rather than the single
step_fn call at SSPRK2.py:87.  The state is written at t = 0 and after each
window-closing step that reaches its output time, as the Forward-Euler
run_fn writes it (forwardeuler.py:68-117; the reference's SSPRK2 writes none).
"""
import time as _time

import torch

from ..constants import TIMESTEP_TOL
from ..solvers.roe_exner_solver import ExnerBackend
from .loop import DtReader, FiniteWatch, next_target, run_exner_stepper
from .registry import IntegratorBundle, IntegratorConfig, register_integrator_bundle

last_run = {}


def config_fn_SSPRK2(cfl, end_time, out_freq, solver_bundle, solver_config):
    return IntegratorConfig(cfl=cfl, end_time=end_time, out_freq=out_freq, solver_config=solver_config,
                            solver_bundle=solver_bundle)


def combine(a, b):
    """tree_map(lambda a, b: 0.5*a + 0.5*b) (SSPRK2.py:40) on the device."""
    kw = {}
    for name in a.field_names():
        x, y = getattr(a, name), getattr(b, name)
        kw[name] = x if x is y else 0.5 * x + 0.5 * y
    return a.replace(**kw)


def ssprk2_step(sb, sc, state, time, dt, mask):
    s1 = sb.step_fn(state, time, dt, mask, sc)
    s2 = sb.step_fn(s1, time, dt, mask, sc)
    be = getattr(sc, "backend", None)
    if be is not None and hasattr(be, "combine_into"):
        from ..state import fields_struct
        from ..solvers.roe_exner_solver import _new_evolving
        out = _new_evolving(state)
        be.combine_into(fields_struct(state), fields_struct(s2), fields_struct(out),
                        torch.cuda.current_stream().cuda_stream)
        return out
    return combine(state, s2)


def run_fn_SSPRK2(state, config, io=None, mesh=None, b_mask=None):
    sb, sc = config.solver_bundle, config.solver_config
    mask = sb.mask_fn(state, sc.mpi_handler, b_mask)
    state = sb.init_fn(state, mask, sc)
    if io is not None:
        io.write_state(state, mesh, mask[0])
    be = getattr(sc, "backend", None)
    if isinstance(be, ExnerBackend) and getattr(be, "stepper_loop", True):
        a = _time.perf_counter()
        state, iters, time = run_exner_stepper(state, mask, be, config, io, mesh, "SSPRK2",
                                                 autotune=getattr(be, "autotune", None))
        last_run.update(time=time, iters=iters, wall=_time.perf_counter() - a, path="stepper")
        return state
    rd = DtReader(state.h.dtype, state.h.device)
    watch = FiniteWatch(be, state.h.device)
    time, dt, iters = 0.0, 0.0, 0
    a = _time.perf_counter()
    while time < config.end_time - TIMESTEP_TOL:
        while time + dt < next_target(time, config.out_freq, config.end_time) - TIMESTEP_TOL:
            dt_dev = sb.compute_dt_fn(state, config.cfl, mask, sc)
            rd.start(dt_dev)
            state = ssprk2_step(sb, sc, state, time, dt_dev, mask)
            dt = rd.value()
            watch.dt(dt, iters + 1, time)
            time = time + dt
            iters += 1
        target = next_target(time, config.out_freq, config.end_time)
        dt_dev = sb.compute_dt_fn(state, config.cfl, mask, sc)
        rd.start(dt_dev)
        dt = rd.value()
        watch.dt(dt, iters + 1, time)
        closing = time + dt >= target - TIMESTEP_TOL
        if closing:
            dt = target - time
            dt_dev = dt
        state = ssprk2_step(sb, sc, state, time, dt_dev, mask)
        if io is not None and closing:
            io.write_state(state, mesh, mask[0])
        time = time + dt
        iters += 1
        if closing:
            watch.window(state, time)
    torch.cuda.synchronize()
    last_run.update(time=time, iters=iters, dt=dt, wall=_time.perf_counter() - a, path="functional")
    return state


@register_integrator_bundle("SSPRK2")
def integrator_SSPRK2():
    return IntegratorBundle(name="SSPRK2", config=config_fn_SSPRK2, run_fn=run_fn_SSPRK2)

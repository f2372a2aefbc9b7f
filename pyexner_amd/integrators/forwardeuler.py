"""Forward Euler integrator (integrators/forwardeuler.py:14-140).

run_fn follows the reference loop: the inner loop advances while
``time + dt_prev < next_target - TIMESTEP_TOL`` (forwardeuler.py:32-35, 80);
at each output window a probe step of length ``next_target - time`` is taken
and handed to ``io`` but not committed (forwardeuler.py:91-109), then one
committed step with the CFL dt (forwardeuler.py:112-117).
"""
import time as _time

import torch

from ..constants import TIMESTEP_TOL
from ..solvers.roe_exner_solver import ExnerBackend
from .loop import DtReader, FiniteWatch, next_target, run_exner_stepper
from .registry import IntegratorBundle, IntegratorConfig, register_integrator_bundle

last_run = {}


def config_fn_forwardeuler(cfl, end_time, out_freq, solver_bundle, solver_config):
    return IntegratorConfig(cfl=cfl, end_time=end_time, out_freq=out_freq, solver_config=solver_config,
                            solver_bundle=solver_bundle)


def run_fn_forwardeuler(state, config, io=None, mesh=None, b_mask=None):
    sb, sc = config.solver_bundle, config.solver_config
    mask = sb.mask_fn(state, sc.mpi_handler, b_mask)
    state = sb.init_fn(state, mask, sc)
    if io is not None:
        io.write_state(state, mesh, mask[0])
    be = getattr(sc, "backend", None)
    if isinstance(be, ExnerBackend) and getattr(be, "stepper_loop", True):
        a = _time.perf_counter()
        state, iters, time = run_exner_stepper(state, mask, be, config, io, mesh, "Forward Euler",
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
            state = sb.step_fn(state, time, dt_dev, mask, sc)
            dt = rd.value()
            watch.dt(dt, iters + 1, time)
            time = time + dt
            iters += 1
        dt_dev = sb.compute_dt_fn(state, config.cfl, mask, sc)
        rd.start(dt_dev)
        dt = rd.value()
        watch.dt(dt, iters + 1, time)
        target = next_target(time, config.out_freq, config.end_time)
        closing = time + dt >= target - TIMESTEP_TOL
        if closing:
            probe = sb.step_fn(state, time, target - time, mask, sc)
            if io is not None:
                io.write_state(probe, mesh, mask[0])
        state = sb.step_fn(state, time, dt_dev, mask, sc)
        time = time + dt
        iters += 1
        if closing:
            watch.window(state, time)
    torch.cuda.synchronize()
    last_run.update(time=time, iters=iters, dt=dt, wall=_time.perf_counter() - a, path="functional")
    return state


@register_integrator_bundle("Forward Euler")
def integrator_forwardeuler():
    return IntegratorBundle(name="Forward Euler", config=config_fn_forwardeuler, run_fn=run_fn_forwardeuler)

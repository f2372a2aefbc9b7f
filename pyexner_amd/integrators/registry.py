"""Integrator bundle registry (integrators/registry.py:8-40)."""
from typing import Any, Callable, NamedTuple


class SimState(NamedTuple):
    time: float
    out_freq: float
    end_time: float
    dt: float
    state: Any
    cfl: float


class IntegratorConfig(NamedTuple):
    cfl: float
    end_time: float
    out_freq: float
    solver_config: Any
    solver_bundle: Any


class IntegratorBundle(NamedTuple):
    name: str
    config: Callable
    run_fn: Callable


INTEGRATOR_REGISTRY: dict = {}


def register_integrator_bundle(name: str):
    def decorator(bundle: Callable):
        INTEGRATOR_REGISTRY[name] = bundle()
        return bundle()
    return decorator


def create_integrator_bundle(name: str):
    if name not in INTEGRATOR_REGISTRY:
        raise KeyError(f"Unknown integration scheme '{name}'. Available options: {list(INTEGRATOR_REGISTRY.keys())}")
    return INTEGRATOR_REGISTRY[name]

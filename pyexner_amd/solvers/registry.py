"""Solver bundle registry (solvers/registry.py:10-38)."""
from typing import Any, Callable, NamedTuple


class SolverConfig(NamedTuple):
    mpi_handler: Any
    boundaries: Any
    dx: float
    halo_exchange: Callable
    compute_G: Any = None
    compute_n: Any = None
    backend: Any = None  # device-side program: grid, sediment table, ring-fill programs, dt workspace


class SolverBundle(NamedTuple):
    name: str
    config: Callable
    mask_fn: Callable
    init_fn: Callable
    step_fn: Callable
    compute_dt_fn: Callable


SOLVER_REGISTRY: dict = {}


def register_solver_bundle(name: str):
    def decorator(bundle: Callable):
        SOLVER_REGISTRY[name] = bundle()
        return bundle()
    return decorator


def create_solver_bundle(name: str):
    if name not in SOLVER_REGISTRY:
        raise KeyError(f"Unknown solver scheme '{name}'. Available options: {list(SOLVER_REGISTRY.keys())}")
    return SOLVER_REGISTRY[name]

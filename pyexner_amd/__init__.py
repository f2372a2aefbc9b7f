"""pyexner_amd - MI355X (gfx950) backend for PyExner's per-step finite-volume
hot path: Roe and Roe-Exner edge fluxes, Forward-Euler / SSPRK2 updates with
bed evolution, ghost-cell ring fill and the CFL dt reduction, behind the
reference's solver/integrator registry API (src/PyExner/__init__.py:1-39).
"""
from . import constants  # noqa: F401
from .boundaries import BOUNDARY_REGISTRY, BoundaryManager, register_boundary  # noqa: F401
from .integrators import INTEGRATOR_REGISTRY, create_integrator_bundle, register_integrator_bundle
from .mesh import Mesh2D
from .solvers import SOLVER_REGISTRY, create_solver_bundle, register_solver_bundle
from .state import STATE_REGISTRY, RoeExnerState, RoeState, create_state, register_state
from .runtime.driver import run_driver, setup_case, run_case

__all__ = [
    "Mesh2D", "run_driver", "setup_case", "run_case", "RoeState", "RoeExnerState", "STATE_REGISTRY",
    "create_state", "register_state", "SOLVER_REGISTRY", "create_solver_bundle", "register_solver_bundle",
    "INTEGRATOR_REGISTRY", "create_integrator_bundle", "register_integrator_bundle", "BoundaryManager",
    "BOUNDARY_REGISTRY", "register_boundary",
]

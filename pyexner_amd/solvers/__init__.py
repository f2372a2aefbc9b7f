from .roe_solver import solver_roe  # noqa: F401
from .roe_exner_solver import solver_roeexner  # noqa: F401
from .registry import SOLVER_REGISTRY, create_solver_bundle, register_solver_bundle

__all__ = ["SOLVER_REGISTRY", "create_solver_bundle", "register_solver_bundle"]

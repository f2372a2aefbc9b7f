from .forwardeuler import integrator_forwardeuler  # noqa: F401
from .ssprk2 import integrator_SSPRK2  # noqa: F401
from .registry import INTEGRATOR_REGISTRY, create_integrator_bundle, register_integrator_bundle

__all__ = ["INTEGRATOR_REGISTRY", "create_integrator_bundle", "register_integrator_bundle"]

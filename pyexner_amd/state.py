"""Structure-of-arrays states held in device tensors.

Mirrors state/base.py:8-47, state/roe_state.py:8-63, state/roe_exner_state.py:8-48
and the registry in state/registry.py:5-46.  Every field is a contiguous
(Ny, Nx) torch tensor on the GPU in the working dtype (float32 or float64);
``seds`` is the (7, nsed) sediment table (a host numpy array, read once when the
solver is configured).
"""
import traceback
from dataclasses import dataclass, fields, replace as dc_replace
from typing import Any, Optional

import numpy as np
import torch

from . import _hip

STATE_REGISTRY = {}


def register_state(name):
    def decorator(cls):
        STATE_REGISTRY[name] = cls
        return cls
    return decorator


def create_empty_state(name, mesh, rank=0, dtype=torch.float32, device="cuda"):
    try:
        state_cls = STATE_REGISTRY[name]
        return state_cls.empty(mesh, dtype=dtype, device=device)
    except KeyError:
        if rank == 0:
            print(f"[create_state] Unknown state '{name}'. Available options: {list(STATE_REGISTRY.keys())}")
            traceback.print_stack()
        raise


def create_state(name, params):
    return STATE_REGISTRY[name].from_params(params)


@dataclass
class BaseState:
    h: torch.Tensor
    hu: torch.Tensor
    hv: torch.Tensor

    def replace(self, **kwargs):
        return dc_replace(self, **kwargs)

    def to_host(self):
        kw = {}
        for f in fields(self):
            v = getattr(self, f.name)
            kw[f.name] = v.detach().cpu().numpy() if isinstance(v, torch.Tensor) else v
        return type(self)(**kw)

    def field_names(self):
        return [f.name for f in fields(self) if f.name != "seds" and getattr(self, f.name) is not None]

    @property
    def dtype(self):
        return self.h.dtype

    @property
    def shape(self):
        return tuple(self.h.shape)

    def clone(self):
        kw = {}
        for f in fields(self):
            v = getattr(self, f.name)
            kw[f.name] = v.clone() if isinstance(v, torch.Tensor) else v
        return type(self)(**kw)


def _dev_field(a, dtype, device):
    t = torch.as_tensor(np.asarray(a), device=device).to(dtype)
    return t.contiguous()


@register_state("Roe Exner")
@dataclass
class RoeExnerState(BaseState):
    z: torch.Tensor
    z_b: torch.Tensor
    n: torch.Tensor
    n_b: torch.Tensor
    G: torch.Tensor
    seds: Any
    # low part of a compensated float32 bed (z_b + z_b_lo); None when not used
    z_b_lo: Optional[torch.Tensor] = None

    @classmethod
    def empty(cls, mesh, dtype=torch.float32, device="cuda"):
        shape = mesh.local_shape
        z = lambda: torch.zeros(shape, dtype=dtype, device=device)  # noqa: E731
        return cls(h=z(), hu=z(), hv=z(), z=torch.ones(shape, dtype=dtype, device=device), z_b=z(), n=z(),
                   n_b=z(), G=z(), seds=np.zeros((1, 1)))

    @classmethod
    def from_arrays(cls, arrays, seds, dtype=torch.float32, device="cuda"):
        kw = {k: _dev_field(arrays[k], dtype, device) for k in ("h", "hu", "hv", "z", "z_b", "n")}
        kw["n_b"] = _dev_field(arrays.get("n_b", arrays["n"]), dtype, device)
        kw["G"] = _dev_field(arrays.get("G", np.zeros_like(np.asarray(arrays["h"]))), dtype, device)
        return cls(seds=np.asarray(seds, dtype=np.float64), **kw)

    def fields_struct(self):
        return fields_struct(self)


@register_state("Roe")
@dataclass
class RoeState(BaseState):
    z: torch.Tensor
    n: torch.Tensor

    @classmethod
    def empty(cls, mesh, dtype=torch.float32, device="cuda"):
        shape = mesh.local_shape
        z = lambda: torch.zeros(shape, dtype=dtype, device=device)  # noqa: E731
        return cls(h=z(), hu=z(), hv=z(), z=torch.ones(shape, dtype=dtype, device=device), n=z())

    @classmethod
    def from_params(cls, params, dtype=torch.float32, device="cuda"):
        h = np.asarray(params.get("h_init"))
        return cls(h=_dev_field(h, dtype, device), hu=_dev_field(h * params.get("u_init"), dtype, device),
                   hv=_dev_field(h * params.get("v_init"), dtype, device),
                   z=_dev_field(params.get("z_init"), dtype, device),
                   n=_dev_field(params.get("roughness"), dtype, device))

    @classmethod
    def from_arrays(cls, arrays, dtype=torch.float32, device="cuda"):
        return cls(**{k: _dev_field(arrays[k], dtype, device) for k in ("h", "hu", "hv", "z", "n")})


_FIELD_NAMES = _hip.FIELD_NAMES


def check_field(t, dtype, shape):
    if not isinstance(t, torch.Tensor) or t.device.type != "cuda":
        raise TypeError("state fields must be torch tensors on the GPU")
    if t.dtype != dtype or tuple(t.shape) != tuple(shape) or not t.is_contiguous():
        raise ValueError(f"field has dtype {t.dtype}, shape {tuple(t.shape)}; expected contiguous {dtype} {shape}")


def fields_struct(state):
    """pxh_fields of device pointers for a state (missing fields are NULL)."""
    f = _hip.Fields()
    dtype, shape = state.h.dtype, state.h.shape
    for name in _FIELD_NAMES:
        t = getattr(state, name, None)
        if t is None:
            continue
        check_field(t, dtype, shape)
        setattr(f, name, t.data_ptr())
    return f


def dtype_code(dtype):
    if dtype == torch.float32:
        return _hip.PXH_F32
    if dtype == torch.float64:
        return _hip.PXH_F64
    raise TypeError(f"unsupported dtype {dtype}; use torch.float32 or torch.float64")

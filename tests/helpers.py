"""Shared builders: the same synthetic input for the oracle (numpy) and the
GPU backend (torch tensors through the C-ABI)."""
import numpy as np

from oracle import boundaries as ob
from oracle import steps as osteps
from oracle.state import ExnerState, RoeState
from pyexner_amd.inputs import seds_from_params, initial_G


def oracle_exner(mesh, arrays, params, dtype=np.float64, cubic_dtype=None):
    seds = seds_from_params(params)
    a = {k: np.asarray(arrays[k], dtype=dtype) for k in ("h", "hu", "hv", "z", "z_b", "n")}
    G = np.asarray(arrays.get("G", initial_G(params, a["h"].shape)), dtype=dtype)
    s = ExnerState(n_b=a["n"].copy(), G=G, seds=seds.astype(dtype), **a)
    bcs = ob.Boundaries("Roe Exner", params["boundaries"], mesh.local_X, mesh.local_Y)
    mask = osteps.get_mask(s.z, bcs.boundary_mask)
    constant = not isinstance(params["erosion"]["grass_factor"], str)
    s = osteps.init_roeexner(s, mask, bcs, constant)
    prob = osteps.ExnerProblem(mask, bcs, mesh.dh, params.get("cfl", 0.5), constant, cubic_dtype)
    return s, prob


def oracle_roe(mesh, arrays, params, dtype=np.float64):
    s = RoeState(**{k: np.asarray(arrays[k], dtype=dtype) for k in ("h", "hu", "hv", "z", "n")})
    bcs = ob.Boundaries("Roe", params["boundaries"], mesh.local_X, mesh.local_Y)
    return s, osteps.RoeProblem(bcs, mesh.dh, params.get("cfl", 0.5))


def gpu_case(mesh, arrays, params, dtype):
    from pyexner_amd.runtime.driver import setup_case
    state, bm, solver, sc, integ, ic = setup_case(mesh, arrays, params, dtype=dtype)
    mask = solver.mask_fn(state, sc.mpi_handler, bm.boundary_mask)
    state = solver.init_fn(state, mask, sc)
    return state, mask, solver, sc


def gpu_fe(state, mask, solver, sc, cfl, time=0.0):
    dt = solver.compute_dt_fn(state, cfl, mask, sc)
    return solver.step_fn(state, time, dt, mask, sc), dt


def gpu_ssprk2(state, mask, solver, sc, cfl, time=0.0):
    from pyexner_amd.integrators.ssprk2 import ssprk2_step
    dt = solver.compute_dt_fn(state, cfl, mask, sc)
    return ssprk2_step(solver, sc, state, time, dt, mask), dt


FIELDS_EXNER = ("h", "hu", "hv", "z_b", "G")
FIELDS_ROE = ("h", "hu", "hv")


def rel_linf(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    scale = np.max(np.abs(b))
    err = np.max(np.abs(a - b))
    return err / scale if scale > 0 else err


def compare(gpu_state, ora_state, fields):
    host = gpu_state.to_host()
    return {f: rel_linf(getattr(host, f), getattr(ora_state, f)) for f in fields}

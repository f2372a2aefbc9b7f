#!/usr/bin/env python3
"""Throughput of the Roe-Exner SSPRK2 step on one MI355X.

Workload: Roe-Exner, SSPRK2, float32 (cubic wave speeds in float64), radial dam
break on one 8192 x 8192 domain, dx = 0.05 m, h = 2.0 m inside r < 50 m else
0.5 m, 256 seeded Gaussian bed bumps, MPM sediment (d = 1.61 mm, 2630 kg/m^3,
porosity 0.42), Reflective ring.  A step is one SSPRK2 step of the whole
domain (dt reduction, two Forward-Euler stages, combine) with the state
resident in HBM.  Prints one JSON line.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--size N] [--no-cpu]

The timed region is K SSPRK2 steps from the initial state (K = 500 by default,
the whole workload), after W warm-up steps on the same buffers.
"""
import argparse
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# Algorithmic HBM bytes per cell of each float32 kernel of the step (fields
# 4 B, nodata mask 1 B, compensated bed z_b + z_b_lo; neighbour re-reads and
# the ring fills not counted).
ALG_BYTES = {
    # roe_solve_2D + momentum_corrections: read h hu hv z_b z n_b + mask, write h hu hv
    "hydro_sweep": 6 * 4 + 1 + 3 * 4,
    # stage-1 compute_G + exner_solve_2D: read h hu hv z_b z_b_lo + mask, write z_b z_b_lo
    "bed_sweep": 5 * 4 + 1 + 2 * 4,
    # stage-2 bed sweep + SSPRK2 combine + next dt: read h hu hv z_b z_b_lo + mask and
    # h hu hv z_b z_b_lo G of u^n, write h hu hv z_b z_b_lo G
    "bed_sweep_combine_dt": 5 * 4 + 1 + 6 * 4 + 6 * 4,
    # standalone compute_dt (first step only): read h hu hv G + mask
    "compute_dt": 4 * 4 + 1,
}
PER_STEP = {"hydro_sweep": 2, "bed_sweep": 1, "bed_sweep_combine_dt": 1, "compute_dt": 0}
HBM_PEAK_GBS = 8000.0


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=500)
    p.add_argument("--warmup", type=int, default=5)
    p.add_argument("--size", type=int, default=8192)
    p.add_argument("--kernel-reps", type=int, default=10)
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--cpu-size", type=int, default=1024)
    p.add_argument("--cpu-steps", type=int, default=30)
    p.add_argument("--graph", type=int, default=1, help="capture each step into a hipGraph (1) or not (0)")
    return p.parse_args()


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def case(n):
    from pyexner_amd import cases
    scale = n / 8192.0
    return cases.radial_dambreak(nx=n, ny=n, dx=0.05, n_bumps=max(1, int(round(256 * scale * scale))),
                                 radius=50.0 * scale)


def cpu_baseline(n, steps):
    """The float32 oracle (numpy, one thread) on the same physics at n x n."""
    from tests.helpers import oracle_exner
    from oracle import steps as osteps
    mesh, arr, params = case(n)
    s, prob = oracle_exner(mesh, arr, params, dtype=np.float32, cubic_dtype=np.float64)
    s, _ = osteps.ssprk2_step(prob, s)  # warm caches
    t0 = time.perf_counter()
    for _ in range(steps):
        s, _ = osteps.ssprk2_step(prob, s)
    el = time.perf_counter() - t0
    return {"value": n * n * steps / el, "unit": "cell-steps/s", "cores": 1, "kind": "port",
            "sample": f"float32 numpy oracle, radial dam break {n}x{n} (same physics scaled), "
                      f"{steps} SSPRK2 steps, {el:.1f} s, cpu={platform.processor() or platform.machine()}, "
                      f"host cores={os.cpu_count()}"}


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {a.gpus} is not supported (every run here uses one "
              "MI355X: see the sharding design in DESIGN.md)", file=sys.stderr)
        sys.exit(2)
    import torch
    from pyexner_amd import _hip
    from pyexner_amd.runtime.driver import setup_case
    from pyexner_amd.state import fields_struct
    from pyexner_amd.stepper import ExnerStepper

    _hip.load()
    n = a.size
    log(f"building {n}x{n} radial dam break")
    t0 = time.perf_counter()
    mesh, arr, params = case(n)
    state, bm, solver, sc, integ, ic = setup_case(mesh, arr, params, dtype=torch.float32)
    del arr
    mask = solver.mask_fn(state, sc.mpi_handler, bm.boundary_mask)
    state = solver.init_fn(state, mask, sc)
    st = ExnerStepper(sc.backend, state, mask, params["cfl"])
    init = {f: getattr(state, f).clone() for f in ("h", "hu", "hv", "z_b", "z_b_lo", "G")}
    del state
    torch.cuda.synchronize()
    log(f"setup {time.perf_counter() - t0:.1f} s")

    # warm up (and capture one step into a hipGraph), then restart from the initial state
    for _ in range(a.warmup):
        st.ssprk2_step()
    use_graph = bool(a.graph)
    if use_graph:
        st.capture(1, "SSPRK2")
    torch.cuda.synchronize()
    for f, t in init.items():
        getattr(st.a, f).copy_(t)
    del init
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    st.compute_dt()  # the first step's dt; every step then leaves the next one
    for _ in range(a.steps):
        if use_graph:
            st.replay()
        else:
            st.ssprk2_step()
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1)
    ms_step = ms / a.steps
    value = n * n * a.steps / (ms / 1e3)
    log(f"{a.steps} steps in {ms:.2f} ms: {ms_step:.3f} ms/step, {value / 1e9:.3f} G cell-steps/s")
    h = st.state.h
    if not torch.isfinite(h).all():
        raise RuntimeError("non-finite depth after the timed steps")
    t_end = float(st.dt)  # next dt (sanity)
    log(f"next dt {t_end:.4e} s")

    # per-kernel durations: each kernel alone, HIP events on the stream it runs on
    be, m0 = sc.backend, st.m0
    stream = torch.cuda.current_stream().cuda_stream
    g = be.grid
    fa, fb, fc = st.fa, st.fb, st.fc
    dtp = st.dt.data_ptr()
    part = be.workspace.data_ptr()
    calls = {
        "hydro_sweep": lambda: _hip.call("pxh_roe_exner_hydro_sweep", g, fa, fb, m0, dtp, stream),
        "bed_sweep": lambda: _hip.call("pxh_roe_exner_bed_sweep", g, fb, fa, fb, None, m0, be.sed, dtp, None,
                                       stream),
        "bed_sweep_combine_dt": lambda: _hip.call("pxh_roe_exner_bed_sweep", g, fc, fb, fa, fa, m0, be.sed, dtp,
                                                  part, stream),
        "compute_dt": lambda: be.dt_into(fa, m0, st.cfl, dtp, stream),
    }
    kern = {}
    for name, fn in calls.items():
        fn()
        torch.cuda.synchronize()
        k0, k1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        k0.record()
        for _ in range(a.kernel_reps):
            fn()
        k1.record()
        torch.cuda.synchronize()
        us = k0.elapsed_time(k1) * 1e3 / a.kernel_reps
        gbs = ALG_BYTES[name] * n * n / (us * 1e-6) / 1e9
        kern[name] = {"us": round(us, 2), "alg_bytes_per_cell": ALG_BYTES[name], "alg_GBs": round(gbs, 1),
                      "per_step": PER_STEP[name]}
        log(f"  {name:22s} {us:9.1f} us  {gbs:8.1f} GB/s algorithmic")
    dom = max(kern, key=lambda k: kern[k]["us"] * PER_STEP[k])
    achieved = kern[dom]["alg_GBs"]
    roofline = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None, "kernel": dom,
                "alg_bytes_per_cell": ALG_BYTES[dom], "cells_per_launch": n * n}
    step_bytes = sum(ALG_BYTES[k] * c for k, c in PER_STEP.items())

    cpu = None
    if not a.no_cpu:
        log(f"cpu baseline: oracle float32 at {a.cpu_size}^2 x {a.cpu_steps} steps")
        cpu = cpu_baseline(a.cpu_size, a.cpu_steps)

    out = {
        "metric": "Roe-Exner SSPRK2 cell-steps/s at 8192^2 fp32",
        "value": value,
        "unit": "cell-steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "roe_exner_ssprk2_radial_dambreak", "nx": n, "ny": n, "dx": 0.05,
                   "integrator": "SSPRK2", "cfl": 0.5, "bumps": 256 if n == 8192 else None,
                   "steps_from_initial_state": a.steps, "hipgraph": use_graph,
                   "bed": "float32 z_b + z_b_lo (compensated)", "math": "float32 hardware rcp/rsq/log2/exp2",
                   "step_alg_bytes_per_cell": step_bytes, "step_alg_GBs": round(step_bytes * value / 1e9, 1)},
        "roofline": roofline,
        "kernels": kern,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

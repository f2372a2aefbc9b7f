#!/usr/bin/env python3
"""Throughput of the Roe-Exner SSPRK2 step on one MI355X.

Workload: Roe-Exner, SSPRK2, float32 (cubic wave speeds in float64), radial dam
break on one 8192 x 8192 domain, dx = 0.05 m, h = 2.0 m inside r < 50 m else
0.5 m, 256 seeded Gaussian bed bumps, MPM sediment (d = 1.61 mm, 2630 kg/m^3,
porosity 0.42), Reflective ring.  A step is one SSPRK2 step of the whole
domain (dt reduction, two Forward-Euler stages, combine) with the state
resident in HBM.  Prints one JSON line.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--size N] [--no-cpu]
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

# Algorithmic HBM bytes per cell for each float32 kernel of the step
# (state fields 4 B each, nodata mask 1 B; neighbour re-reads not counted).
ALG_BYTES = {
    "roe_solve_2D": 6 * 4 + 1 + 3 * 4,      # read h hu hv z_b z n_b + mask, write h hu hv
    "momentum_corrections": 3 * 4 + 1,      # read h, z, z_b + mask (writes only where zeroed)
    "compute_G": 3 * 4 + 1 + 4,             # read h hu hv + mask, write G
    "exner_solve_2D": 5 * 4 + 1 + 4,        # read h hu hv z_b G + mask, write z_b
    "compute_dt": 4 * 4 + 1,                # read h hu hv G + mask
    "ssprk2_combine": 10 * 4 + 5 * 4,       # read 2 x (h hu hv z_b G), write 5
}
HBM_PEAK_GBS = 8000.0


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=20)
    p.add_argument("--warmup", type=int, default=3)
    p.add_argument("--size", type=int, default=8192)
    p.add_argument("--kernel-reps", type=int, default=10)
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--cpu-size", type=int, default=512)
    p.add_argument("--cpu-steps", type=int, default=8)
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
    del state
    torch.cuda.synchronize()
    log(f"setup {time.perf_counter() - t0:.1f} s")

    for _ in range(a.warmup):
        st.ssprk2_step()
    torch.cuda.synchronize()
    use_graph = bool(a.graph)
    if use_graph:
        st.capture(1, "SSPRK2")
        st.replay()
        torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record()
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

    # per-kernel durations (HIP events on the stream the kernels run on)
    be, m0 = sc.backend, st.m0
    stream = torch.cuda.current_stream().cuda_stream
    g = be.grid
    fa, fb, fc = st.fa, st.fb, st.fc
    dtp = st.dt.data_ptr()
    calls = {
        "roe_solve_2D": lambda: _hip.call("pxh_roe_solve_2D", g, fa, fb, m0, dtp, stream),
        "momentum_corrections": lambda: _hip.call("pxh_momentum_corrections", g, fa, fb, m0, stream),
        "compute_G": lambda: _hip.call("pxh_compute_G", g, fa, fb, m0, be.sed, stream),
        "exner_solve_2D": lambda: _hip.call("pxh_exner_solve_2D", g, fa, fb, m0, dtp, stream),
        "compute_dt": lambda: be.dt_into(fa, m0, st.cfl, dtp, stream),
        "ssprk2_combine": lambda: be.combine_into(fa, fc, fb, stream),
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
        kern[name] = {"us": round(us, 2), "alg_GBs": round(gbs, 1)}
        log(f"  {name:22s} {us:9.1f} us  {gbs:8.1f} GB/s algorithmic")
    per_step_counts = {"roe_solve_2D": 2, "momentum_corrections": 2, "compute_G": 2, "exner_solve_2D": 2,
                       "compute_dt": 1, "ssprk2_combine": 1}
    dom = max(kern, key=lambda k: kern[k]["us"] * per_step_counts[k])
    achieved = kern[dom]["alg_GBs"]
    roofline = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None, "kernel": dom,
                "alg_bytes_per_cell": ALG_BYTES[dom], "cells_per_launch": n * n}
    step_bytes = sum(ALG_BYTES[k] * c for k, c in per_step_counts.items())

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
                   "hipgraph": use_graph, "step_alg_bytes_per_cell": step_bytes,
                   "step_alg_GBs": round(step_bytes * value / 1e9, 1)},
        "roofline": roofline,
        "kernels": kern,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

#!/usr/bin/env python3
"""Throughput of the Roe-Exner SSPRK2 step on one MI355X.

Workload: Roe-Exner, SSPRK2, float32 throughout (hardware-rate square roots,
reciprocals and transcendentals; the cubic wave speeds solved in float32),
radial dam break on one 8192 x 8192 domain, dx = 0.05 m, h = 2.0 m inside r < 50 m else
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
# 4 B, nodata mask 1 B, compensated bed z_b + z_b_lo; neighbour re-reads, the
# two-cell frame's small kernels and the ring fills not counted).
ALG_BYTES = {
    # stage 1, one launch: roe_solve_2D + momentum_corrections + compute_G +
    # exner_solve_2D.  read h hu hv z_b z_b_lo z n_b + mask, write h hu hv z_b
    # z_b_lo (MPM: G of stage 1 is recomputed by stage 2, so it is not stored)
    "stage_sweep": 7 * 4 + 1 + 5 * 4,
    # stage 2 + SSPRK2 combine + the next dt's minima, one launch: read the
    # stage-1 h hu hv z_b z_b_lo z n_b + mask and u^n's h hu hv z_b z_b_lo G,
    # write u^(n+1)'s h hu hv z_b z_b_lo G
    "stage_sweep_combine_dt": 7 * 4 + 1 + 6 * 4 + 6 * 4,
    # standalone compute_dt (first step only): read h hu hv G + mask
    "compute_dt": 4 * 4 + 1,
}
PER_STEP = {"stage_sweep": 1, "stage_sweep_combine_dt": 1, "compute_dt": 0}
HBM_PEAK_GBS = 8000.0
# The same kernels under the byte model of the hot-path survey (fields 4 B,
# mask 1 B; G recomputed, not stored; plain float32 bed): stage 1 reads h hu hv
# z_b z n_b + mask and writes h hu hv z_b (41 B); stage 2 + combine also reads
# u^n's h hu hv z_b (57 B); 98 B per SSPRK2 cell-step.  The roofline fraction
# is quoted on it; the bytes this build moves (G stored because the reference's
# SSPRK2 averages it and the next dt reads the average, compensated bed) are
# ALG_BYTES above.
SURVEY_BYTES = {"stage_sweep": 41, "stage_sweep_combine_dt": 57, "compute_dt": 13}


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "hbm_traffic_ssprk2_radial_dambreak_8192_fp32.txt")


def measured_traffic(n, kernel):
    """HBM bytes per launch of `kernel`, from the committed rocprofv3 counter
    summary (FETCH_SIZE and WRITE_SIZE per dispatch, each scaled by the factor
    that makes the counter read the known bytes of the SSPRK2 combine kernel),
    or None when the summary does not cover this grid."""
    if n != 8192 or not os.path.exists(TRAFFIC_FILE):
        return None
    with open(TRAFFIC_FILE) as f:
        for line in f:
            w = line.split()
            if len(w) >= 4 and w[0] == kernel and not line.startswith("#"):
                return int(float(w[3]))
    return None


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=500)
    p.add_argument("--warmup", type=int, default=5)
    p.add_argument("--size", type=int, default=8192)
    p.add_argument("--kernel-steps", type=int, default=10,
                   help="SSPRK2 steps issued sweep by sweep to time each kernel inside the step")
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--cpu-channel-steps", type=int, default=20,
                   help="timed steps of the 2048 x 512 channel in the CPU baseline")
    p.add_argument("--cpu-steps", type=int, default=3,
                   help="timed steps of the headline grid in the CPU baseline (0: skip it)")
    p.add_argument("--cpu-numpy-size", type=int, default=1024,
                   help="grid of the single-thread numpy oracle timing reported beside the baseline (0: skip)")
    p.add_argument("--graph", type=int, default=1, help="capture each step into a hipGraph (1) or not (0)")
    p.add_argument("--tile-steps", type=int, default=50,
                   help="steps of the same domain as a 4x4 decomposition (0: skip)")
    p.add_argument("--ens-members", type=int, default=256, help="members of the batched ensemble (0: skip)")
    p.add_argument("--ens-steps", type=int, default=20, help="timed steps of the ensemble and of its baseline")
    p.add_argument("--adj-steps", type=int, default=200,
                   help="steps of the gradient run (bed calibration, 1024x256 float64; 0: skip)")
    p.add_argument("--adj-every", type=int, default=20, help="checkpoint interval of the gradient run")
    p.add_argument("--adj-rows", type=int, default=0,
                   help="float64 strip height of the gradient run (0: tuned with --autotune 1, else 32)")
    p.add_argument("--accuracy", type=int, default=1,
                   help="also run the timed steps in float64 and report the float32 result's deviation (1) or not (0)")
    p.add_argument("--tile-rows", type=int, default=0,
                   help="strip height of the tiled run's fused stage (0: tuned with --autotune 1, else 32)")
    p.add_argument("--active-steps", type=int, default=20,
                   help="timed steps of an 8192^2 channel in which every cell moves and carries sediment (0: skip)")
    p.add_argument("--driver-steps", type=int, default=1000,
                   help="about this many steps of run_driver's SSPRK2 loop on the 2048x512 channel (0: skip)")
    p.add_argument("--autotune", type=int, default=1,
                   help="choose the fused stage's strip height by timing the candidates (1) or keep 32 (0)")
    return p.parse_args()


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def case(n):
    from pyexner_amd import cases
    scale = n / 8192.0
    return cases.radial_dambreak(nx=n, ny=n, dx=0.05, n_bumps=max(1, int(round(256 * scale * scale))),
                                 radius=50.0 * scale)


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or platform.machine()


def _torch_cpu_steps(mesh, arr, params, steps, torch, warm=True):
    """Median seconds per SSPRK2 step of the torch CPU restatement (float32),
    after one untimed step if warm."""
    from tests.helpers import oracle_exner
    from oracle import torch_cpu as tc
    so, prob = oracle_exner(mesh, arr, params, dtype=np.float32)
    st = tc.from_oracle_state(so, torch.float32)
    seds = np.asarray(so.seds, dtype=np.float64)
    del so
    m0 = torch.as_tensor(prob.mask[0])
    fill = tc.Fill(prob.bcs)
    if warm:
        st, _ = tc.ssprk2_step(st, m0, fill, seds, prob.constant, prob.dx, prob.cfl)
    ts = []
    for _ in range(steps):
        t0 = time.perf_counter()
        st, _ = tc.ssprk2_step(st, m0, fill, seds, prob.constant, prob.dx, prob.cfl)
        ts.append(time.perf_counter() - t0)
    return float(np.median(ts)), bool(torch.isfinite(st["h"]).all())


def cpu_baseline(a, n, torch):
    """Own CPU restatement, not the reference (its JAX CPU path is not
    runnable here): the whole-array torch restatement of oracle/torch_cpu.py
    in float32 on torch's intra-op pool, every thread this process may use
    (OMP_NUM_THREADS, else all cores).  Timed on the 2048 x 512 channel
    (median of --cpu-channel-steps steps) and on the headline grid (median of
    --cpu-steps steps); `value` is the headline grid's rate.  The numpy
    oracle on one thread is reported beside it for information."""
    from pyexner_amd import cases
    threads = min(int(os.environ.get("OMP_NUM_THREADS") or os.cpu_count()), os.cpu_count())
    prev = torch.get_num_threads()
    torch.set_num_threads(threads)
    try:
        mesh, arr, params = cases.channel()
        cells_c2 = mesh.local_shape[0] * mesh.local_shape[1]
        t_c2, ok2 = _torch_cpu_steps(mesh, arr, params, a.cpu_channel_steps, torch)
        del mesh, arr
        log(f"cpu baseline: channel 2048x512 {t_c2 * 1e3:.1f} ms/step on {threads} threads")
        out = {"unit": "cell-steps/s", "cores": threads, "kind": "port",
               "label": "own CPU restatement, not the reference",
               "cpu_model": _cpu_model(), "host_cores": os.cpu_count(),
               "channel_2048x512": {"steps": a.cpu_channel_steps, "s_per_step": t_c2,
                                    "value": cells_c2 / t_c2, "finite": ok2}}
        if a.cpu_steps > 0:
            mesh, arr, params = case(n)
            t_c3, ok3 = _torch_cpu_steps(mesh, arr, params, a.cpu_steps, torch, warm=False)
            del mesh, arr
            log(f"cpu baseline: radial {n}^2 {t_c3:.2f} s/step on {threads} threads")
            out.update(value=n * n / t_c3, sample=(
                f"torch CPU float32 restatement, {threads} threads, radial dam break {n}x{n}: median of "
                f"its first {a.cpu_steps} SSPRK2 steps ({t_c3:.2f} s/step); channel 2048x512: median of "
                f"{a.cpu_channel_steps} steps after one untimed"))
            out["radial"] = {"n": n, "steps": a.cpu_steps, "s_per_step": t_c3, "finite": ok3}
        else:
            out.update(value=cells_c2 / t_c2, sample=(
                f"torch CPU float32 restatement, {threads} threads, channel 2048x512, median of "
                f"{a.cpu_channel_steps} SSPRK2 steps"))
    finally:
        torch.set_num_threads(prev)
    if a.cpu_numpy_size > 0:
        out["numpy_oracle_info"] = numpy_oracle_rate(a.cpu_numpy_size, 2)
    return out


def numpy_oracle_rate(n, steps):
    """The float32 numpy oracle, one thread, on the same physics at n x n
    (information only)."""
    from tests.helpers import oracle_exner
    from oracle import steps as osteps
    mesh, arr, params = case(n)
    s, prob = oracle_exner(mesh, arr, params, dtype=np.float32)
    s, _ = osteps.ssprk2_step(prob, s)
    t0 = time.perf_counter()
    for _ in range(steps):
        s, _ = osteps.ssprk2_step(prob, s)
    el = time.perf_counter() - t0
    return {"value": n * n * steps / el, "unit": "cell-steps/s", "threads": 1,
            "sample": f"radial dam break {n}x{n}, {steps} SSPRK2 steps, {el:.1f} s"}


def tiled_run(a, n, params, torch):
    """The same domain as a 4 x 4 decomposition of (n/4)^2 owned tiles, all on
    this GPU: the fused stage per tile with three ghost rings per shared side,
    one exchange per stage (ghost strips copied straight from the neighbours'
    owned strips, one launch per phase), one hipGraph per step; with
    --autotune the tiles' strip height is timed first (on the same case,
    restarted from the initial state).  Returns its throughput (not the
    headline metric)."""
    from pyexner_amd.runtime.driver import setup_case
    from pyexner_amd.tiles import TiledExner
    log(f"tiled run: 4x4 tiles of {n // 4}^2")
    mesh, arr, _ = case(n)
    params = dict(params, backend={"fused_stage": False})
    state, bm, solver, sc, integ, ic = setup_case(mesh, arr, params, dtype=torch.float32)
    del arr
    mask = solver.mask_fn(state, sc.mpi_handler, bm.boundary_mask)
    state = solver.init_fn(state, mask, sc)
    rows, tune_ms = (a.tile_rows or None), {}
    if a.autotune and not rows:
        from pyexner_amd.autotune import tune_tile_strip_rows
        probe = TiledExner(sc.backend, state, mask, bm, (4, 4), params["cfl"])
        rows, tune_ms = tune_tile_strip_rows(probe)
        del probe
        torch.cuda.empty_cache()
        log("tile strip rows: " + ", ".join(f"{r}: {t:.3f} ms" for r, t in tune_ms.items()) + f" -> {rows}")
    tl = TiledExner(sc.backend, state, mask, bm, (4, 4), params["cfl"])
    if rows:
        tl.set_strip_rows(rows)
    del state
    tl.capture()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(a.tile_steps):
        tl.replay()
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / a.tile_steps
    if not torch.isfinite(tl.gather("h")).all():
        raise RuntimeError("non-finite depth in the tiled run")
    hb = tl.halo_bytes_per_step()
    log(f"tiled: {ms:.3f} ms/step, {n * n / (ms * 1e-3) / 1e9:.3f} G cell-steps/s, halo {hb / 1e6:.1f} MB/step")
    t = tl.tiles[(1, 1)]
    return {"ms_per_step": ms, "cell_steps_per_s": n * n / (ms * 1e-3), "steps": a.tile_steps,
            "tile_owned": [n // 4, n // 4], "tile_array": [t.grid.ny, t.grid.nx], "ghost_rings": tl.hw,
            "halo_bytes_per_step": hb, "strip_rows": rows or 32,
            "autotune_ms_per_step": {str(k): round(v, 4) for k, v in tune_ms.items()},
            "path": ("fused stage per tile, one exchange of every evolving field per stage, dt over tiles after it"
                     if tl.fused else "two-sweep stage, exchange after each sweep, dt reduced over tiles")}


def ensemble_run(a, torch):
    """The batched ensemble of uneven channel scenarios (nx ~ U{256..2048},
    ny ~ U{64..512}, seeded inflow, depth and sediment diameter): one hipGraph
    step for all members, against looping the single-domain step over them
    (eagerly, and with that whole loop captured in one hipGraph)."""
    from pyexner_amd.ensemble import Ensemble, channel_members
    from pyexner_amd.stepper import ExnerStepper
    t0 = time.perf_counter()
    members = channel_members(a.ens_members, seed=4)
    ens = Ensemble(members, dtype=torch.float32, cfl=0.5, keep_members=True)
    del members
    alone = [ExnerStepper(be, s, m, 0.5) for be, s, m in ens.alone]
    ens.alone = None
    torch.cuda.synchronize()
    log(f"ensemble: {ens.n} members, {ens.cells / 1e6:.1f} M cells, setup {time.perf_counter() - t0:.1f} s")
    k = a.ens_steps

    def timed(fn):
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(k):
            fn()
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / k

    rows, tune_ms = 32, {}
    if a.autotune:
        # the strip height of the members' fused stage sweeps, timed on the
        # ensemble itself and then restored to its initial state
        keep = {f: getattr(ens.a, f).clone() for f in ("h", "hu", "hv", "z_b", "z_b_lo", "G")
                if getattr(ens.a, f) is not None}
        for r in (16, 24, 32, 48):
            ens.set_strip_rows(r)
            ens.ssprk2_step()
            tune_ms[r] = timed(ens.ssprk2_step)
        rows = min(tune_ms, key=tune_ms.get)
        ens.set_strip_rows(rows)
        for f, t in keep.items():
            getattr(ens.a, f).copy_(t)
        ens.dt_ready = False
        del keep
        log("ensemble strip rows: " + ", ".join(f"{r}: {t:.3f} ms" for r, t in tune_ms.items()) + f" -> {rows}")
    ens.capture()
    ms_ens = timed(ens.replay)

    def loop():
        for s in alone:
            s.ssprk2_step()

    loop()
    ms_loop = timed(loop)
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        loop()
    ms_loop_graph = timed(g.replay)
    if not torch.isfinite(ens.a.h).all():
        raise RuntimeError("non-finite depth in the ensemble")
    cells = ens.cells
    log(f"ensemble: {ms_ens:.3f} ms/step ({cells / (ms_ens * 1e-3) / 1e9:.2f} G cell-steps/s); "
        f"looped single-domain steps {ms_loop:.3f} ms eager, {ms_loop_graph:.3f} ms in one hipGraph")
    return {"members": ens.n, "cells": cells, "steps": k, "ms_per_step": ms_ens,
            "cell_steps_per_s": cells / (ms_ens * 1e-3),
            "loop_ms_per_step": ms_loop, "loop_graph_ms_per_step": ms_loop_graph,
            "speedup_vs_loop": ms_loop / ms_ens, "speedup_vs_loop_graph": ms_loop_graph / ms_ens,
            "dtype": "f32", "scenario": "channel members, seed 4, dx 0.02, Transmissive in/out, Reflective walls",
            "strip_rows": rows, "autotune_ms_per_step": {str(k): round(v, 4) for k, v in tune_ms.items()}}


def active_run(a, n, torch):
    """The same step on an n x n channel (dx 0.02 m, h 0.1 m, u 0.5 m/s,
    256 seeded bumps, MPM sediment, Transmissive in/out, Reflective walls):
    every cell moves and carries sediment, so no still-water shortcut
    applies.  hipGraph replay, autotuned strip height."""
    from pyexner_amd import cases
    from pyexner_amd.autotune import tune_strip_rows
    from pyexner_amd.runtime.driver import setup_case
    from pyexner_amd.stepper import ExnerStepper
    mesh, arr, params = cases.channel(nx=n, ny=n, dx=0.02, n_bumps=256, seed=3)
    state, bm, solver, sc, integ, ic = setup_case(mesh, arr, params, dtype=torch.float32)
    del arr
    mask = solver.mask_fn(state, sc.mpi_handler, bm.boundary_mask)
    state = solver.init_fn(state, mask, sc)
    rows, _ = tune_strip_rows(sc.backend, state, mask, params["cfl"], steps=3, use_cache=False)
    st = ExnerStepper(sc.backend, state, mask, params["cfl"])
    del state
    moving = float((st.a.hu != 0).float().mean())
    st.capture(1, "SSPRK2")
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(a.active_steps):
        st.replay()
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / a.active_steps
    sediment = float((st.a.G > 0).float().mean())
    if not torch.isfinite(st.a.h).all():
        raise RuntimeError("non-finite depth in the active-flow run")
    log(f"active flow {n}^2: {ms:.3f} ms/step, {n * n / (ms * 1e-3) / 1e9:.2f} G cell-steps/s "
        f"(moving {moving:.3f}, G > 0 in {sediment:.3f} of the cells)")
    return {"workload": f"channel {n}x{n}, dx 0.02, h 0.1, u 0.5, 256 bumps", "ms_per_step": ms,
            "cell_steps_per_s": n * n / (ms * 1e-3), "steps": a.active_steps, "strip_rows": rows,
            "moving_fraction": moving, "sediment_fraction": sediment}


def driver_run(a, torch):
    """run_driver's SSPRK2 loop (integrators.ssprk2.run_fn_SSPRK2) on the
    2048 x 512 channel, float32, about a.driver_steps steps in 5 output
    windows without io: the ExnerStepper loop (hipGraph steps, dt read back
    one step behind), the functional loop (step_fn twice and the combine per
    step, dt read every step), and the stepper's bare graph replays."""
    from pyexner_amd import cases
    from pyexner_amd.integrators import ssprk2
    from pyexner_amd.runtime.driver import run_case, setup_case
    from pyexner_amd.stepper import ExnerStepper
    mesh, arr, params = cases.channel()
    ny, nx = mesh.local_X.shape
    state, bm, solver, sc, integ, ic = setup_case(mesh, arr, params, dtype=torch.float32)
    mask = solver.mask_fn(state, sc.mpi_handler, bm.boundary_mask)
    state = solver.init_fn(state, mask, sc)
    dt0 = float(solver.compute_dt_fn(state, params["cfl"], mask, sc))
    end = round(a.driver_steps * dt0, 3)
    out = {"workload": f"channel {nx}x{ny}, dx 0.02, SSPRK2, end_time {end} s, out_freq {end / 5:.4f} s"}
    for key, opts in (("stepper_loop", {}), ("functional_loop", {"stepper_loop": False})):
        p = dict(params, integrator="SSPRK2", end_time=end, out_freq=end / 5, backend=opts)
        run_case(mesh, arr, dict(p, end_time=end / 20, out_freq=end / 20), dtype=torch.float32)  # warm-up
        run_case(mesh, arr, p, dtype=torch.float32)
        r = dict(ssprk2.last_run)
        out[key] = {"iters": r["iters"], "wall_s": r["wall"], "cell_steps_per_s": nx * ny * r["iters"] / r["wall"],
                    "ms_per_step": r["wall"] * 1e3 / r["iters"]}
        log(f"run_driver {key}: {r['iters']} steps in {r['wall']:.3f} s, "
            f"{out[key]['ms_per_step']:.4f} ms/step")
    from pyexner_amd.autotune import tune_strip_rows
    rows, _ = tune_strip_rows(sc.backend, state, mask, params["cfl"])  # the height run_fn chose (cached)
    out["strip_rows"] = rows
    st = ExnerStepper(sc.backend, state, mask, params["cfl"])
    st.capture(1, "SSPRK2")
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(200):
        st.replay()
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / 200
    out["graph_replay_ms_per_step"] = ms
    log(f"stepper graph replay on the same grid: {ms:.4f} ms/step")
    return out


def accuracy_run(a, n, f32_end, torch):
    """The timed run's float32 state against the same a.steps SSPRK2 steps in
    float64 (which matches the float64 oracle to 1e-12): max-norm and L1
    deviations, relative to max|field| (hv: max|hu|) and, for the bed, to the
    bed change."""
    from pyexner_amd.runtime.driver import setup_case
    from pyexner_amd.stepper import ExnerStepper
    mesh, arr, params = case(n)
    state, bm, solver, sc, integ, ic = setup_case(mesh, arr, params, dtype=torch.float64)
    del arr
    mask = solver.mask_fn(state, sc.mpi_handler, bm.boundary_mask)
    state = solver.init_fn(state, mask, sc)
    zb0 = state.z_b.clone()
    st = ExnerStepper(sc.backend, state, mask, params["cfl"])
    del state
    t0 = time.perf_counter()
    st.compute_dt()
    for _ in range(a.steps):
        st.ssprk2_step()
    torch.cuda.synchronize()
    r = st.a
    out = {"steps": a.steps, "reference": "float64 run of the same steps (fused stage)"}
    hu_max = float(r.hu.abs().max())
    for f in ("h", "hu", "hv"):
        d = (f32_end[f].double() - getattr(r, f)).abs()
        scale = hu_max if f == "hv" else float(getattr(r, f).abs().max())
        out[f"{f}_max_rel"] = float(d.max()) / scale
        out[f"{f}_l1_rel"] = float(d.sum() / (r.hu.abs().sum() if f == "hv" else getattr(r, f).abs().sum()))
    bed = f32_end["z_b"].double() + (f32_end["z_b_lo"].double() if "z_b_lo" in f32_end else 0)
    d = (bed - r.z_b).abs()
    change = (r.z_b - zb0).abs()
    out["bed_max_over_max_dz"] = float(d.max() / change.max())
    out["bed_l1_over_l1_dz"] = float(d.sum() / change.sum())
    out["bed_cells_over_1e-4_max_dz"] = int((d > 1e-4 * change.max()).sum())
    log("float32 vs float64 after {} steps ({:.1f} s): ".format(a.steps, time.perf_counter() - t0)
        + ", ".join(f"{k} {v:.2e}" if isinstance(v, float) else f"{k} {v}" for k, v in out.items()
                    if k not in ("steps", "reference")))
    return out


def adjoint_run(a, torch):
    """Bed calibration: the gradient of 0.5 |z_b(T) - target|^2 with respect to
    z_b0, d50 and porosity over a.adj_steps SSPRK2 steps of the 1024x256
    calibration channel in float64, checkpoint every a.adj_every steps.  Times
    the forward pass (fused steps + checkpoints), the backward pass (segment
    replays + adjoint steps + init adjoint), and the same number of plain
    float64 forward steps; the float64 strip height (it sets the forward
    sweeps of both) is a.adj_rows, or tuned on this grid with --autotune."""
    from pyexner_amd import cases
    from pyexner_amd.adjoint import ExnerGradientRun, seeded_target
    from pyexner_amd.stepper import ExnerStepper
    mesh, arr, params = cases.calibration_channel(nx=1024, ny=256)
    run = ExnerGradientRun(mesh, arr, params, n_steps=a.adj_steps, checkpoint_every=a.adj_every)
    zb0 = run.raw.z_b.clone()
    target = seeded_target(zb0)
    d50_0, por_0 = float(run.seds[1, 0]), float(run.seds[6, 0])

    def once():
        zb = zb0.clone().requires_grad_(True)
        d50 = torch.tensor(d50_0, dtype=torch.float64, requires_grad=True)
        por = torch.tensor(por_0, dtype=torch.float64, requires_grad=True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        out = run(zb, d50, por)
        loss = 0.5 * ((out[3] - target) ** 2).sum()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        loss.backward()
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        return t1 - t0, t2 - t1, float(loss.detach()), zb.grad, float(d50.grad), float(por.grad)

    once()  # warm-up
    rows64 = a.adj_rows or 32
    if a.adj_rows:
        run.be.set_strip_rows(a.adj_rows)
    elif a.autotune:
        # 32-row strips give this 1024 x 256 grid 144 wavefronts for 1024 SIMDs
        from pyexner_amd.autotune import tune_strip_rows
        rows64, _ = tune_strip_rows(run.be, run.s0, run.mask, run.cfl, steps=5, use_cache=False)
    tf, tb, loss, g, gd, gp = once()
    # plain float64 forward steps of the same case (fused SSPRK2, no checkpoints)
    stp = ExnerStepper(run.be, run.s0, run.mask, run.cfl)
    stp.compute_dt()
    stp.ssprk2_step()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(a.adj_steps):
        stp.ssprk2_step()
    e1.record()
    torch.cuda.synchronize()
    ms_fwd = e0.elapsed_time(e1) / a.adj_steps
    ms_bwd = tb * 1e3 / a.adj_steps
    # the float32 forward step of the same case (the dtype the calibration
    # config names), strip height tuned, hipGraph replay
    from pyexner_amd.autotune import tune_strip_rows
    from pyexner_amd.runtime.driver import setup_case
    s32, bm32, sol32, sc32, _, _ = setup_case(mesh, arr, params, dtype=torch.float32)
    m32 = sol32.mask_fn(s32, sc32.mpi_handler, bm32.boundary_mask)
    s32 = sol32.init_fn(s32, m32, sc32)
    rows32, _ = tune_strip_rows(sc32.backend, s32, m32, run.cfl, steps=5, use_cache=False)
    st32 = ExnerStepper(sc32.backend, s32, m32, run.cfl)
    st32.capture(1, "SSPRK2")
    torch.cuda.synchronize()
    e0.record()
    for _ in range(a.adj_steps):
        st32.replay()
    e1.record()
    torch.cuda.synchronize()
    ms_fwd32 = e0.elapsed_time(e1) / a.adj_steps
    if not torch.isfinite(g).all():
        raise RuntimeError("non-finite bed gradient")
    log(f"adjoint: {a.adj_steps} steps 1024x256 f64: forward {tf * 1e3:.1f} ms, backward {tb * 1e3:.1f} ms; "
        f"plain forward step {ms_fwd:.3f} ms, backward per step {ms_bwd:.3f} ms ({ms_bwd / ms_fwd:.2f}x), "
        f"strip rows {rows64}; float32 forward step {ms_fwd32:.3f} ms ({ms_bwd / ms_fwd32:.2f}x)")
    return {"case": "calibration_channel 1024x256, dx 0.02, seed 5; target z_b0 + seeded perturbation (seed 6)",
            "dtype": "f64", "steps": a.adj_steps, "checkpoint_every": a.adj_every, "strip_rows": rows64,
            "forward_ms": tf * 1e3, "backward_ms": tb * 1e3, "plain_forward_ms_per_step": ms_fwd,
            "backward_ms_per_step": ms_bwd, "adjoint_step_vs_forward_step": ms_bwd / ms_fwd,
            "f32_forward_ms_per_step": ms_fwd32, "f32_forward_strip_rows": rows32,
            "adjoint_step_vs_f32_forward_step": ms_bwd / ms_fwd32,
            "loss": loss, "grad_zb0_norm": float(g.norm()), "grad_d50": gd, "grad_porosity": gp}


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {a.gpus} is not supported: the path runs on one "
              "MI355X (a decomposition across GPUs would give each GPU one tile of pyexner_amd.tiles)",
              file=sys.stderr)
        sys.exit(2)
    import torch
    from pyexner_amd import _hip
    from pyexner_amd.runtime.driver import setup_case
    from pyexner_amd.state import fields_struct
    from pyexner_amd.stepper import ExnerStepper, record_fields

    _hip.load()
    n = a.size
    log(f"building {n}x{n} radial dam break")
    t0 = time.perf_counter()
    mesh, arr, params = case(n)
    state, bm, solver, sc, integ, ic = setup_case(mesh, arr, params, dtype=torch.float32)
    del arr
    mask = solver.mask_fn(state, sc.mpi_handler, bm.boundary_mask)
    state = solver.init_fn(state, mask, sc)
    rows, tune_ms = 32, {}
    if a.autotune:
        from pyexner_amd.autotune import tune_strip_rows
        rows, tune_ms = tune_strip_rows(sc.backend, state, mask, params["cfl"], steps=3, use_cache=False)
        torch.cuda.empty_cache()
        log("strip rows: " + ", ".join(f"{r}: {t:.3f} ms" for r, t in tune_ms.items()) + f" -> {rows}, "
            f"wave caps {sc.backend.stage_waves()}")
    st = ExnerStepper(sc.backend, state, mask, params["cfl"])
    no_nodata, z_uniform = st.no_nodata, st.z_uniform
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
    f32_end = None
    if a.accuracy:
        f32_end = {f: getattr(st.a, f).clone() for f in ("h", "hu", "hv", "z_b", "z_b_lo")
                   if getattr(st.a, f) is not None}

    # Per-kernel durations inside the step: SSPRK2 steps issued launch group by
    # launch group through the C-ABI (the same launches pxh_roe_exner_ssprk2
    # makes for a frame-local ring, in the same order, on the same buffers),
    # with HIP events recorded on the stream the kernels run on around each
    # stage-sweep launch.
    be, m0 = sc.backend, st.m0
    stream = torch.cuda.current_stream().cuda_stream
    g = be.grid_for(st.no_nodata, st.z_uniform)  # the grid the stepper's steps run with
    fa, fb, fc = st.fa, st.fb, st.fc
    if st.packed:
        fb = record_fields(st.b)  # the first stage's output as the step keeps it (PXH_PACKED_STAGE)
    dtp = st.dt.data_ptr()
    part = be.workspace.data_ptr()
    G_H, G_HUV, G_ZB, G_G = _hip.G_H, _hip.G_HUV, _hip.G_ZB, _hip.G_G
    if not sc.boundaries.frame_local:
        raise RuntimeError("the bench ring is expected to be frame-local (fused stage)")
    bc_hy = sc.boundaries.program((G_H, G_HUV))
    bc_bed1 = sc.boundaries.program((G_ZB,) if be.sed.mpm else (G_ZB, G_G))
    bc_end = sc.boundaries.program((G_ZB, G_G))
    cfl = float(params["cfl"])
    sz = be.sed
    skip_g = _hip.SWEEP_SKIP_G

    def ev():
        e = torch.cuda.Event(enable_timing=True)
        e.record()
        return e

    # the bytes this build moves, less the streams the stage sweeps skip: the
    # mask on a grid without nodata cells (PXH_NO_NODATA), z when it holds one
    # value (PXH_Z_UNIFORM); the survey's byte model is kept as it stands
    alg_b, survey_b = dict(ALG_BYTES), dict(SURVEY_BYTES)
    for k in ("stage_sweep", "stage_sweep_combine_dt"):
        if st.no_nodata:
            alg_b[k] -= 1
        if st.no_nodata and st.z_uniform:
            alg_b[k] -= 4
    spans = {k: [] for k in ALG_BYTES}
    steps_span = []
    torch.cuda.synchronize()
    e0 = ev()
    be.dt_into(fa, m0, cfl, dtp, stream)
    spans["compute_dt"].append((e0, ev()))
    for _ in range(a.kernel_steps):
        es = ev()
        e0 = ev()
        _hip.call("pxh_roe_exner_stage_sweep", g, fa, fb, None, m0, sz, dtp, skip_g, None, stream)
        spans["stage_sweep"].append((e0, ev()))
        _hip.call("pxh_bc_apply", g, fb, bc_hy, stream)
        _hip.call("pxh_roe_exner_frame_sweep", g, fb, fa, fb, None, m0, sz, dtp, skip_g, 0.0, None, stream)
        _hip.call("pxh_bc_apply", g, fb, bc_bed1, stream)
        e0 = ev()
        _hip.call("pxh_roe_exner_stage_sweep", g, fb, fc, fa, m0, sz, dtp, 0, part, stream)
        spans["stage_sweep_combine_dt"].append((e0, ev()))
        _hip.call("pxh_bc_apply", g, fc, bc_hy, stream)
        _hip.call("pxh_roe_exner_frame_sweep", g, fc, fb, None, fa, m0, sz, dtp, 0, cfl, part, stream)
        _hip.call("pxh_bc_apply", g, fa, bc_end, stream)
        steps_span.append((es, ev()))
    # one dispatch of known size for the byte-counter calibration (profiling
    # runs only read it): the SSPRK2 combine over u^n and u^(1) into u^(2)
    be.combine_into(fa, st.fb, fc, stream)
    torch.cuda.synchronize()
    kern = {}
    for name, lst in spans.items():
        us = sum(x.elapsed_time(y) for x, y in lst) * 1e3 / len(lst)
        gbs = alg_b[name] * n * n / (us * 1e-6) / 1e9
        kern[name] = {"us": round(us, 2), "launches": len(lst), "alg_bytes_per_cell": alg_b[name],
                      "alg_GBs": round(gbs, 1), "per_step": PER_STEP[name]}
        log(f"  {name:22s} {us:9.1f} us  {gbs:8.1f} GB/s algorithmic  ({len(lst)} launches in the step sequence)")
    step_us = sum(x.elapsed_time(y) for x, y in steps_span) * 1e3 / len(steps_span)
    log(f"  instrumented step      {step_us:9.1f} us  (frame sweeps, ring fills and events included)")
    dom = max(kern, key=lambda k: kern[k]["us"] * PER_STEP[k])
    achieved = round(survey_b[dom] * n * n / (kern[dom]["us"] * 1e-6) / 1e9, 1)
    traffic = measured_traffic(n, dom)
    roofline = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic, "kernel": dom,
                "alg_bytes_per_cell": survey_b[dom], "cells_per_launch": n * n,
                "byte_model": "hot-path survey: G recomputed, plain float32 bed"}
    roofline_stored = {"achieved": kern[dom]["alg_GBs"], "frac": round(kern[dom]["alg_GBs"] / HBM_PEAK_GBS, 4),
                       "alg_bytes_per_cell": alg_b[dom],
                       "byte_model": "bytes this build moves: G stored, compensated bed z_b + z_b_lo"}
    step_bytes = sum(alg_b[k] * c for k, c in PER_STEP.items())
    step_survey = sum(survey_b[k] * c for k, c in PER_STEP.items())

    tile = None
    if a.tile_steps > 0 and n % 4 == 0:
        tile = tiled_run(a, n, params, torch)
    ens = None
    if a.ens_members > 0 and a.ens_steps > 0:
        torch.cuda.empty_cache()
        ens = ensemble_run(a, torch)

    active = None
    if a.active_steps > 0:
        torch.cuda.empty_cache()
        active = active_run(a, n, torch)
    drv = None
    if a.driver_steps > 0:
        torch.cuda.empty_cache()
        drv = driver_run(a, torch)

    acc = None
    if f32_end is not None:
        del st
        torch.cuda.empty_cache()
        acc = accuracy_run(a, n, f32_end, torch)
        del f32_end
    adj = None
    if a.adj_steps > 0:
        torch.cuda.empty_cache()
        adj = adjoint_run(a, torch)

    cpu = None
    if not a.no_cpu:
        log(f"cpu baseline: torch CPU restatement, channel x {a.cpu_channel_steps} and {n}^2 x {a.cpu_steps}")
        cpu = cpu_baseline(a, n, torch)

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
                   "stage": "fused (frame-local ring)", "strip_rows": rows,
                   "stage_waves_per_simd_cap": list(sc.backend.stage_waves()),
                   "autotune_ms_per_step": {str(k): round(v, 4) for k, v in tune_ms.items()}, "bed": "float32 z_b + z_b_lo (compensated)", "math": "float32 hardware rcp/rsq/log2/exp2",
                   "mask_loads": "skipped (no nodata cell)" if no_nodata else "per row",
                   "z_loads": "read once (z uniform)" if no_nodata and z_uniform else "per row",
                   "step_alg_bytes_per_cell": step_bytes, "step_alg_GBs": round(step_bytes * value / 1e9, 1),
                   "step_survey_bytes_per_cell": step_survey,
                   "step_survey_frac": round(step_survey * value / 1e9 / HBM_PEAK_GBS, 4)},
        "roofline": roofline,
        "roofline_stored_bytes": roofline_stored,
        "kernels": kern,
        "cpu_baseline": cpu,
        "active_flow": active,
        "run_driver_channel": drv,
        "tiles_4x4": tile,
        "ensemble": ens,
        "adjoint": adj,
        "accuracy_f32_vs_f64": acc,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

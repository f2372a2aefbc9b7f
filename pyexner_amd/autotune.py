"""Launch-shape autotuning for the fused Roe-Exner step.

The fused stage sweep gives each wavefront a strip of 60 owned columns and
`rows` owned rows (PXH_STRIP_ROWS of pxh_grid.flags).  Taller strips recompute
fewer halo rows; shorter strips give more wavefronts to the 256 CUs on small
grids and shorter register lifetimes.  The best height depends on the grid
and the GPU, so it is measured: `tune_strip_rows` times a few SSPRK2 steps of
the actual case for each candidate height on a copy of the state and keeps
the fastest.  Results are cached per (shape, dtype, math, sediment mode) for
the process, and in a JSON file when PYEXNER_TUNE_CACHE names one.

After the height, a few occupancy caps are timed (PXH_STAGE1_WAVES /
PXH_STAGE2_WAVES: the most wavefronts per SIMD of each stage's sweep, below
the occupancy the registers allow) and the fastest kept when it beats the
uncapped sweep by at least 1%.

The strip height and the caps change which wavefront computes a cell and
when, never what it computes: every candidate gives the same result bit for
bit.
"""
import json
import os

import torch

from . import _hip

# 8 and 12 rows give grids of about a million cells or fewer enough
# wavefronts to fill the 1024 SIMDs (2048 x 512 has 35 strip columns); taller
# strips recompute fewer halo rows on large grids
CANDIDATES = (8, 12, 16, 24, 32, 48, 64)
# (first, second) stage caps in wavefronts per SIMD tried after the height; the
# registers allow 4 and 3 (float32 sweeps), so each is one or two below that
WAVE_CAPS = ((3, 0), (2, 0), (0, 2))
_CACHE = {}


def _key(be):
    return f"{be.shape[0]}x{be.shape[1]}:{'f64' if be.dtype == torch.float64 else 'f32'}:" \
           f"{'exact' if be.exact_f32 else 'fast'}:{'mpm' if be.sed.mpm else 'const'}:" \
           f"{'lo' if be.compensated_bed else 'plain'}"


def _file():
    return os.environ.get("PYEXNER_TUNE_CACHE")


def _load_file():
    path = _file()
    if not path or not os.path.exists(path):
        return {}
    with open(path) as f:
        return json.load(f)


def _unpack(entry):
    """(rows, (w1, w2)) of a cache entry (an int: rows, no caps)."""
    if isinstance(entry, dict):
        w = entry.get("waves") or (0, 0)
        return int(entry["rows"]), (int(w[0]), int(w[1]))
    return int(entry), (0, 0)


def _store_file(key, rows):
    path = _file()
    if not path:
        return
    data = _load_file()
    data[key] = rows
    with open(path, "w") as f:
        json.dump(data, f, indent=1, sort_keys=True)


def time_steps(be, state, mask, cfl, steps):
    """Milliseconds per fused SSPRK2 step of `state` under the backend's
    current launch shape (one warm-up step; the state is not modified)."""
    from .stepper import ExnerStepper
    st = ExnerStepper(be, state, mask, cfl)
    st.compute_dt()
    st.ssprk2_step()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record()
    for _ in range(steps):
        st.ssprk2_step()
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / steps


def tune_strip_rows(be, state, mask, cfl=0.5, candidates=CANDIDATES, steps=4, use_cache=True, waves=True):
    """Choose the fused stage's strip height for backend `be` (an
    ExnerBackend) on `state` (initialised), then (waves=True) the occupancy
    caps of WAVE_CAPS, and apply them.  Returns (rows, {candidate: ms per
    step}): heights keyed by rows, caps by "rows/w<first><second>"; the
    timings are empty when the choice came from a cache.  be.stage_waves()
    gives the caps kept."""
    key = _key(be)
    if use_cache:
        hit = _CACHE.get(key) or _load_file().get(key)
        if hit:
            rows, w = _unpack(hit)
            be.set_strip_rows(rows)
            be.set_stage_waves(*w)
            _CACHE[key] = {"rows": rows, "waves": list(w)}
            return rows, {}
    times = {}
    be.set_stage_waves(0, 0)
    for rows in candidates:
        be.set_strip_rows(rows)
        times[rows] = time_steps(be, state, mask, cfl, steps)
    best = min(times, key=times.get)
    be.set_strip_rows(best)
    w_best, t_best = (0, 0), times[best]
    if waves and be.dtype == torch.float32:
        for w in WAVE_CAPS:
            be.set_stage_waves(*w)
            t = time_steps(be, state, mask, cfl, steps)
            times[f"{best}/w{w[0]}{w[1]}"] = t
            if t < 0.99 * t_best:
                w_best, t_best = w, t
    be.set_stage_waves(*w_best)
    entry = {"rows": best, "waves": list(w_best)}
    _CACHE[key] = entry
    _store_file(key, entry)
    return best, times


def strip_rows_of(flags):
    """The strip height a grid's flags select (PXH_STRIP_ROWS; 0 = 32)."""
    r = (int(flags) >> _hip.STRIP_ROWS_SHIFT) & 0xFF
    return r or 32


def tune_tile_strip_rows(tiled, candidates=(8, 12, 16, 24, 32), steps=3):
    """Choose the fused stage's strip height for the tiles of a TiledExner by
    timing `steps` SSPRK2 steps per candidate (after one warm-up step) and
    apply it.  The tiles advance while they are timed: tune before the run
    whose result counts, on a copy, or accept the extra steps.  Returns
    (rows, {rows: ms per step})."""
    times = {}
    for rows in candidates:
        tiled.set_strip_rows(rows)
        tiled.ssprk2_step()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        e0.record()
        for _ in range(steps):
            tiled.ssprk2_step()
        e1.record()
        torch.cuda.synchronize()
        times[rows] = e0.elapsed_time(e1) / steps
    best = min(times, key=times.get)
    tiled.set_strip_rows(best)
    return best, times

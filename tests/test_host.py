"""Host-side logic (no GPU): the C-ABI library's exports, the composed ring
fill against the oracle's sequential handlers, the sediment table, polygon
rasterisation and the input front end."""
import ctypes
import os
import re
import subprocess
import sys

import numpy as np
import pytest

from oracle import boundaries as ob
from oracle import roe_exner as rx
from oracle.state import ExnerState, RoeState
from pyexner_amd import _hip, cases
from pyexner_amd.boundaries import BoundaryManager, compose, compute_label_mask
from pyexner_amd.inputs import seds_from_params
from pyexner_amd.solvers.roe_exner_solver import sediment_table

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_declared_symbol():
    hdr = open(os.path.join(ROOT, "include", "pyexner_hip.h")).read()
    names = sorted(set(re.findall(r"\b(pxh_[A-Za-z0-9_]+)\s*\(", hdr)))
    assert len(names) >= 10
    lib = _hip.load()
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(_hip._SIGS), set(names) ^ set(_hip._SIGS)
    assert lib.pxh_version() >= 1
    assert lib.pxh_error_string(-1) == b"invalid argument"


def _emulate_program(entries, handlers, state, groups=None, values=(), time=0.0):
    """Apply the composed program the way k_bc does (every read before any write)."""
    out = state.copy()
    ny, nx = state.h.shape
    names = {_hip.G_H: "h", _hip.G_Z: "z", _hip.G_ZB: "z_b", _hip.G_N: "n", _hip.G_NB: "n_b", _hip.G_G: "G"}
    for e in entries:
        g_ = int(e["group"])
        if groups is not None and g_ not in groups:
            continue
        t, s = divmod(int(e["target"]), nx), divmod(int(e["src"]), nx)
        if int(e["nchain"]) & _hip.BC_VALUE:
            kind, a, b, c = values[int(e["chain"][0])]
            if kind == _hip.BV_CONST and g_ == _hip.G_HUV:
                if a == a:
                    out.hu[t] = a
                if b == b:
                    out.hv[t] = b
            elif kind == _hip.BV_CONST:
                getattr(out, names[g_])[t] = a
            elif kind == _hip.BV_AFFINE_T:
                getattr(out, names[g_])[t] = b * time + a
            elif kind == _hip.BV_OFFSET:
                getattr(out, names[g_])[t] = getattr(state, names[g_])[s] + a
            elif kind == _hip.BV_OUTFLUX:
                fl, gg = 1e-8, 9.81
                h_i = max(state.h[s], fl)
                un = state.hu[s] / h_i * b + state.hv[s] / h_i * c
                Rp = un + 2.0 * np.sqrt(gg * h_i)
                hb = h_i
                for _ in range(5):
                    hb = max(hb, fl)
                    cc = np.sqrt(gg * hb)
                    hb = max(hb - (a / hb + 2.0 * cc - Rp) / (-a / (hb * hb) + gg / max(cc, fl)), fl)
                out.h[t] = hb
            continue
        if g_ == _hip.G_HUV:
            hu, hv = state.hu[s], state.hv[s]
            for q in range(int(e["nchain"])):
                hd = handlers[int(e["chain"][q])]
                nxv, nyv = hd.normal
                if hd.kind == _hip.BC_REFLECT_RX:
                    mn = hu * nxv + hv * nyv
                    mt = -hu * nyv + hv * nxv
                    mr = -mn
                    hu, hv = mr * nxv - mt * nyv, mr * nyv + mt * nxv
                elif hd.kind == _hip.BC_REFLECT_ROE:
                    hu, hv = hu * (1 - 2 * nxv * nxv), hv * (1 - 2 * nyv * nyv)
            out.hu[t], out.hv[t] = hu, hv
        else:
            getattr(out, names[g_])[t] = getattr(state, names[g_])[s]
    return out


def _random_exner(shape, rng):
    return ExnerState(**{k: rng.uniform(-1, 1, shape) for k in ("h", "hu", "hv", "z", "z_b", "n", "n_b", "G")},
                      seds=np.zeros((7, 1)))


@pytest.mark.parametrize("builder,order", [
    ("channel", None), ("radial", None), ("periodic", None),
    ("radial", [3, 0, 2, 1]), ("channel", [5, 4, 0, 3, 1, 2]),
    ("berthon", None), ("berthon", [4, 2, 3, 1, 0]), ("outflux", None), ("outflux", [2, 0, 6, 4, 5, 1, 3]),
])
def test_composed_program_equals_sequential_handlers(builder, order):
    """The composed program (copies, reflections and value entries) against
    the handlers applied one after another, at t = 0.37 (time-dependent beds)."""
    rng = np.random.default_rng(7)
    if builder == "channel":
        mesh, arr, params = cases.channel(nx=14, ny=9, n_bumps=1)
    elif builder == "radial":
        mesh, arr, params = cases.radial_dambreak(nx=11, ny=13, n_bumps=1, radius=0.2)
    elif builder == "berthon":
        mesh, arr, params = cases.berthon_channel(nx=15, ny=8)
    elif builder == "outflux":
        mesh, arr, params = cases.outflux_channel(nx=13, ny=9, steep_fall=True)
    else:
        mesh, arr, params = cases.periodic_box(nx=12, ny=10)
    specs = params["boundaries"]
    if order is not None:
        keys = list(specs)
        specs = {keys[i]: specs[keys[i]] for i in order}
    bm = BoundaryManager({"flux_scheme": "Roe Exner", "boundaries": specs}, mesh.local_X, mesh.local_Y,
                         device="cpu")
    entries, idem, values, levels = compose(bm.boundary_handlers, mesh.local_shape)
    assert idem and len(levels) == 2 and (values == []) == (builder in ("channel", "radial", "periodic"))
    oc = ob.Boundaries("Roe Exner", specs, mesh.local_X, mesh.local_Y)
    np.testing.assert_array_equal(bm.boundary_mask, oc.boundary_mask)
    st = _random_exner(mesh.local_shape, rng)
    st = st.replace(h=np.abs(st.h) + 0.1)  # the outflux depth needs water
    got = _emulate_program(entries, bm.boundary_handlers, st, values=values, time=0.37)
    exp = oc.apply(st, 0.37)
    for f in ("h", "hu", "hv", "z", "z_b", "n", "n_b", "G"):
        np.testing.assert_array_equal(getattr(got, f), getattr(exp, f), err_msg=f)
        # -0.0 and 0.0 are told apart too
        np.testing.assert_array_equal(np.signbit(getattr(got, f)), np.signbit(getattr(exp, f)), err_msg=f)
    # idempotent: applying twice equals once
    again = _emulate_program(entries, bm.boundary_handlers, got, values=values, time=0.37)
    for f in ("h", "hu", "hv", "z_b", "G"):
        np.testing.assert_array_equal(getattr(again, f), getattr(got, f))


def test_composed_program_roe_reflective_and_oblique_normal():
    rng = np.random.default_rng(8)
    mesh, arr, params = cases.dambreak_1d(nx=10, ny=6)
    specs = dict(params["boundaries"])
    specs["north"] = dict(specs["north"], normal=[0.6, 0.8])  # oblique normal: rounding pinned too
    bm = BoundaryManager({"flux_scheme": "Roe", "boundaries": specs}, mesh.local_X, mesh.local_Y, device="cpu")
    entries, idem, values, levels = compose(bm.boundary_handlers, mesh.local_shape)
    oc = ob.Boundaries("Roe", specs, mesh.local_X, mesh.local_Y)
    st = RoeState(**{k: rng.uniform(-1, 1, mesh.local_shape) for k in ("h", "hu", "hv", "z", "n")})
    ex = ExnerState(h=st.h, hu=st.hu, hv=st.hv, z=st.z, z_b=st.z * 0, n=st.n, n_b=st.n * 0, G=st.n * 0,
                    seds=None)
    got = _emulate_program(entries, bm.boundary_handlers, ex)
    exp = oc.apply(st)
    for f in ("h", "hu", "hv", "z", "n"):
        np.testing.assert_array_equal(getattr(got, f), getattr(exp, f), err_msg=f)


def test_non_idempotent_layout_is_detected():
    # a two-cell-thick west polygon: the outer cell reads a cell the same handler writes
    mesh, arr, params = cases.channel(nx=12, ny=8, n_bumps=1)
    X, dh = mesh.local_X, mesh.dh
    poly = [[X[0, 0] - dh / 2, -10], [X[0, 1] + dh / 2, -10], [X[0, 1] + dh / 2, 10], [X[0, 0] - dh / 2, 10]]
    specs = {"w": {"type": "Transmissive", "polygon": poly, "values": [0], "normal": [-1.0, 0.0]}}
    bm = BoundaryManager({"flux_scheme": "Roe Exner", "boundaries": specs}, mesh.local_X, mesh.local_Y,
                         device="cpu")
    assert not bm.idempotent and bm.n_levels == 2
    # the outer column reads the inner one before the inner one is overwritten
    oc = ob.Boundaries("Roe Exner", specs, mesh.local_X, mesh.local_Y)
    st = _random_exner(mesh.local_shape, np.random.default_rng(5))
    got = _emulate_program(bm.entries, bm.boundary_handlers, st, values=bm.values)
    exp = oc.apply(st)
    for f in ("h", "hu", "hv", "z", "z_b", "n", "n_b", "G"):
        np.testing.assert_array_equal(getattr(got, f), getattr(exp, f), err_msg=f)


@pytest.mark.parametrize("dh", [0.1, 0.05])
def test_sequential_ring_fill_of_the_reference_berthon_inputs(dh):
    """The reference's Berthon convergence inputs (2D_const_flux/dh_*/input_*.yaml:
    bottom and top polygons two rows thick) compose into a program of several
    dependency levels that, applied level by level, equals the handlers
    applied one after another (boundary_registry.py:183-187), bit for bit,
    at a time where the inlet and outlet beds have moved."""
    mesh, arr, params = cases.berthon_convergence(dh)
    bm = BoundaryManager(params, mesh.local_X, mesh.local_Y, device="cpu")
    assert bm.n_levels > 1 and not bm.frame_local
    ls = bm.level_start
    assert ls[0] == 0 and ls[-1] == bm.n_entries and all(a <= b for a, b in zip(ls, ls[1:]))
    # a copy entry reads its source before the entry that writes that source
    # runs: the writer sits in a later level
    lev = np.repeat(np.arange(bm.n_levels), np.diff(ls))
    writer = {(int(e["target"]), int(e["group"])): int(lv) for e, lv in zip(bm.entries, lev)}
    for e, lv in zip(bm.entries, lev):
        if e["nchain"] & _hip.BC_VALUE:
            continue
        w = writer.get((int(e["src"]), int(e["group"])))
        assert w is None or w > lv
    oc = ob.Boundaries("Roe Exner", params["boundaries"], mesh.local_X, mesh.local_Y)
    st = _random_exner(mesh.local_shape, np.random.default_rng(11))
    got = _emulate_program(bm.entries, bm.boundary_handlers, st, values=bm.values, time=0.41)
    exp = oc.apply(st, 0.41)
    for f in ("h", "hu", "hv", "z", "z_b", "n", "n_b", "G"):
        np.testing.assert_array_equal(getattr(got, f), getattr(exp, f), err_msg=f)
        np.testing.assert_array_equal(np.signbit(getattr(got, f)), np.signbit(getattr(exp, f)), err_msg=f)


def test_multi_level_program_selects_the_sequential_stage():
    """A program of several levels turns the Roe-Exner backend to the
    single-kernel sequence (plain bed, no fused stage), and the plain-Roe
    backend, which fills once per step, refuses it."""
    import torch
    from pyexner_amd.solvers.roe_exner_solver import ExnerBackend
    from pyexner_amd.solvers.roe_solver import RoeBackend
    mesh, arr, params = cases.berthon_convergence(0.05)
    bm = BoundaryManager(params, mesh.local_X, mesh.local_Y, device="cpu")
    be = ExnerBackend(mesh.local_shape, torch.float32, mesh.dh, None, True, bm, device="cpu")
    assert be.sequential and not be.fused_stage and not be.compensated_bed
    assert be.bc_all.n_levels == bm.n_levels and tuple(be._bed.shape) == tuple(mesh.local_shape)
    with pytest.raises(NotImplementedError):
        RoeBackend(mesh.local_shape, torch.float64, mesh.dh, bm, device="cpu")


def test_sediment_table_matches_oracle_terms_float32():
    """The per-class constants precomputed on the host give the oracle's
    Shields number and gamma_1 bit for bit in float32."""
    params = cases.channel(nx=8, ny=8, n_bumps=1)[2]
    seds = seds_from_params(params).astype(np.float32)
    sd = sediment_table(seds, False, np.float32)
    rng = np.random.default_rng(3)
    h = rng.uniform(0.05, 1, 1000).astype(np.float32)
    vel2 = rng.uniform(0, 4, 1000).astype(np.float32)
    frac, d50, density, k_d, k_e, theta_c, porosity = seds
    s = density / np.float32(1000.0)
    np_ = d50 ** np.float32(1 / 6) / np.float32(21.1)
    th_or = (np_[0] * np_[0] * vel2) / ((s[0] - np.float32(1.0)) * d50[0] * np.cbrt(h))
    th_k = (np.float32(sd.A[0]) * vel2) / (np.float32(sd.B[0]) * np.cbrt(h))
    np.testing.assert_array_equal(th_or, th_k)
    g1_or = np_[0] * (np_[0] * np_[0]) * np.sqrt(np.float32(9.81)) / ((s[0] - np.float32(1.0)) * np.sqrt(h))
    g1_k = np.float32(sd.C[0]) / (np.float32(sd.D[0]) * np.sqrt(h))
    np.testing.assert_array_equal(g1_or, g1_k)
    nb = rx.compute_n(np.zeros((1, 1), np.float32), np.ones((1, 1), np.float32), seds)
    assert np.float32(sd.nb_sed) == nb[0, 0]


def test_polygon_rasterisation_against_matplotlib():
    mpl = pytest.importorskip("matplotlib.path")
    rng = np.random.default_rng(5)
    X, Y = np.meshgrid(np.linspace(0, 1, 101), np.linspace(1, 0, 77))
    poly = rng.uniform(0.1, 0.9, (7, 2))
    poly = poly[np.argsort(np.arctan2(poly[:, 1] - 0.5, poly[:, 0] - 0.5))]
    ours = compute_label_mask(X, Y, poly)
    ref = mpl.Path(poly).contains_points(np.stack([X.ravel(), Y.ravel()], -1)).reshape(X.shape)
    # compare away from the polygon outline
    from scipy.spatial.distance import cdist  # noqa: F401
    edge = np.zeros_like(ours)
    for a, b in zip(poly, np.roll(poly, 1, axis=0)):
        t = np.clip(((X - a[0]) * (b[0] - a[0]) + (Y - a[1]) * (b[1] - a[1])) / np.sum((b - a) ** 2), 0, 1)
        d = np.hypot(X - (a[0] + t * (b[0] - a[0])), Y - (a[1] + t * (b[1] - a[1])))
        edge |= d < 1e-6
    np.testing.assert_array_equal(ours[~edge], ref[~edge])


def test_netcdf_front_end(tmp_path):
    from scipy.io import netcdf_file
    from pyexner_amd.inputs import mesh_and_arrays_from_netcdf
    nx, ny, dh = 9, 5, 0.1
    x = np.arange(nx) * dh
    y = np.arange(ny - 1, -1, -1) * dh
    p = tmp_path / "in.nc"
    with netcdf_file(str(p), "w", version=2) as f:
        f.createDimension("x", nx)
        f.createDimension("y", ny)
        vx = f.createVariable("x", "f4", ("x",))
        vx[:] = x
        vy = f.createVariable("y", "f4", ("y",))
        vy[:] = y
        for k, val in (("h", 0.5), ("hu", 0.1), ("hv", 0.0), ("z", 0.0), ("z_b", 0.02), ("n", 0.0)):
            v = f.createVariable(k, "f8", ("y", "x"))
            v[:] = np.full((ny, nx), val)
    params = {"flux_scheme": "Roe Exner", "erosion": {"grass_factor": 0.005}}
    mesh, arrays = mesh_and_arrays_from_netcdf(str(p), params)
    assert mesh.local_shape == (ny, nx) and mesh.dh == 0.1
    np.testing.assert_allclose(mesh.local_X[0], x + 0.05, atol=1e-6)
    np.testing.assert_allclose(mesh.local_Y[:, 0], y - 0.05, atol=1e-6)
    assert np.all(arrays["G"] == 0.005) and np.all(arrays["h"] == 0.5)


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2 and "one GPU" in r.stderr


def test_frame_local_programs():
    """The fused stage needs every ring-fill target on the outer ring and every
    source within two cells of the edge: true for Reflective, Transmissive and
    Periodic rings, false when a boundary polygon reaches into the interior."""
    from pyexner_amd.boundaries import BoundaryManager
    for mk in (lambda: cases.channel(nx=40, ny=20, n_bumps=2), lambda: cases.radial_dambreak(nx=32, ny=30, n_bumps=2,
                                                                                            radius=0.4),
               lambda: cases.periodic_box(nx=24, ny=20)):
        mesh, arr, params = mk()
        bm = BoundaryManager(params, mesh.local_X, mesh.local_Y, device="cpu")
        assert bm.frame_local
    mesh, arr, params = cases.channel(nx=40, ny=20, n_bumps=2)
    X, Y, dh = mesh.local_X, mesh.local_Y, mesh.dh
    x0 = X[0, 5]
    # a three-cell-wide Transmissive strip: its targets leave the outer ring
    params["boundaries"]["inlet"]["polygon"] = [[X[0, 0] - dh / 2, Y[-1, 0] - dh / 2], [x0 - dh / 2, Y[-1, 0] - dh / 2],
                                                [x0 - dh / 2, Y[0, 0] + dh / 2], [X[0, 0] - dh / 2, Y[0, 0] + dh / 2]]
    bm = BoundaryManager(params, mesh.local_X, mesh.local_Y, device="cpu")
    assert not bm.frame_local


def test_ctypes_structs_match_the_c_header(tmp_path):
    """The ctypes mirrors in pyexner_amd._hip have the sizes and field offsets
    the C compiler gives the header's structs."""
    import ctypes
    mirrors = {"pxh_grid": _hip.Grid, "pxh_fields": _hip.Fields, "pxh_sediments": _hip.Sediments,
               "pxh_bc_entry": _hip.BCEntry, "pxh_bc_handler": _hip.BCHandler, "pxh_bc": _hip.BC}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "pyexner_hip.h"', "int main(void) {"]
    for cname, cls in mirrors.items():
        lines.append(f'    printf("{cname} sizeof %zu\\n", sizeof({cname}));')
        for fname, _ in cls._fields_:
            lines.append(f'    printf("{cname} {fname} %zu\\n", offsetof({cname}, {fname}));')
    lines += ["    return 0;", "}"]
    src = tmp_path / "abi.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "abi"
    subprocess.run(["gcc", "-std=c99", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True,
                   timeout=60)
    got = subprocess.run([str(exe)], capture_output=True, text=True, check=True, timeout=60).stdout.split("\n")
    seen = 0
    for ln in got:
        if not ln:
            continue
        cname, what, val = ln.split()
        cls = mirrors[cname]
        want = ctypes.sizeof(cls) if what == "sizeof" else getattr(cls, what).offset
        assert int(val) == want, (cname, what, int(val), want)
        seen += 1
    assert seen == sum(1 + len(c._fields_) for c in mirrors.values())


def test_tile_programs_partition_the_ring_fill():
    """Every fill of the single-domain program lands in exactly one tile, with
    target and source at the same global cells in the tile's local indices."""
    from pyexner_amd.tiles import HW, _tile_program, split
    mesh, arr, params = cases.channel(nx=47, ny=31, n_bumps=2)
    bm = BoundaryManager(params, mesh.local_X, mesh.local_Y, device="cpu")
    ny, nx = mesh.local_X.shape
    par = (2, 3)
    rows, cols = split(ny, par[0]), split(nx, par[1])
    assert rows[0][0] == 0 and rows[-1][1] == ny and cols[-1][1] == nx
    seen = []
    for ty in range(par[0]):
        for tx in range(par[1]):
            r0, r1 = rows[ty]
            c0, c1 = cols[tx]
            ext = (r0 - (HW if ty else 0), r1 + (HW if ty < par[0] - 1 else 0),
                   c0 - (HW if tx else 0), c1 + (HW if tx < par[1] - 1 else 0))
            sub, start = _tile_program(bm.entries, (ny, nx), (r0, r1, c0, c1), ext)
            tnx = ext[3] - ext[2]
            assert start[0] == 0 and start[-1] == len(sub)
            for e in sub:
                tr, tc = divmod(int(e["target"]), tnx)
                sr, sc = divmod(int(e["src"]), tnx)
                seen.append(((tr + ext[0]) * nx + tc + ext[2], (sr + ext[0]) * nx + sc + ext[2], int(e["group"])))
    glob = sorted((int(e["target"]), int(e["src"]), int(e["group"])) for e in bm.entries)
    assert sorted(seen) == glob


def test_halo_sizes_and_rejected_geometry():
    """pxh_halo_bytes (host-only: it reads no device memory): strip sizes, and
    the calls the strips cannot serve."""
    # stand-in addresses: the size query only checks which pointers are set
    fs = _hip.Fields(*[0x1000 * (k + 1) for k in range(8)], None)
    grid = _hip.Grid(8, 9, _hip.PXH_F32, 0, 0.1)
    lib = _hip.load()
    groups = (1 << _hip.G_H) | (1 << _hip.G_HUV) | (1 << _hip.G_ZB) | (1 << _hip.G_G)  # h hu hv z_b G
    assert lib.pxh_halo_bytes(grid, fs, 2, groups, _hip.HALO_N | _hip.HALO_S) == 2 * 2 * 9 * 5 * 4
    assert lib.pxh_halo_bytes(grid, fs, 1, groups, _hip.HALO_W) == 1 * 8 * 5 * 4
    assert lib.pxh_halo_bytes(grid, fs, 4, groups, _hip.HALO_W) < 0               # hw beyond 3
    assert lib.pxh_halo_bytes(grid, fs, 3, groups, _hip.HALO_W) == 3 * 8 * 5 * 4   # three rings (8 rows >= 7)
    assert lib.pxh_halo_bytes(grid, fs, 2, groups, 0) < 0                          # no side
    assert lib.pxh_halo_bytes(grid, fs, 2, groups, 15) < 0                         # both phases in one call
    assert lib.pxh_halo_bytes(grid, fs, 2, groups, _hip.HALO_W | _hip.HALO_N) < 0
    assert lib.pxh_halo_bytes(grid, fs, 2, 1 << 9, _hip.HALO_W) < 0                # no such group
    small = _hip.Grid(4, 9, _hip.PXH_F32, 0, 0.1)
    assert lib.pxh_halo_bytes(small, fs, 2, groups, _hip.HALO_N) < 0               # no owned row between rings
    fs.z_b_lo = 0x9000                                                              # compensated bed travels with z_b
    assert lib.pxh_halo_bytes(grid, fs, 1, groups, _hip.HALO_E) == 1 * 8 * 6 * 4


def test_ensemble_plan_sizes_and_rejections():
    """pxh_ensemble_plan (host-only): sizing call, filled plan and workspace,
    and the members it refuses."""
    import ctypes
    mem = (_hip.Member * 3)()
    shapes = [(40, 70), (33, 129), (65, 64)]
    off = 0
    for k, (ny, nx) in enumerate(shapes):
        mem[k].off, mem[k].ny, mem[k].nx, mem[k].dx = off, ny, nx, 0.02
        mem[k].sed.n_sed, mem[k].sed.mpm, mem[k].sed.one_minus_por = 1, 1, 0.58
        mem[k].sed.frac[0], mem[k].sed.AB[0], mem[k].sed.CD[0], mem[k].sed.theta_c[0] = 1.0, 0.1, 2.0, 0.047
        off += ny * nx
    lib = _hip.load()
    size = lib.pxh_ensemble_plan(_hip.PXH_F32, 0, mem, 3, None, 0)
    assert size > 0
    buf = (ctypes.c_uint8 * size)()
    assert lib.pxh_ensemble_plan(_hip.PXH_F32, 0, mem, 3, buf, size) == size
    ws = lib.pxh_ensemble_workspace_size(buf)
    assert ws > 0 and ws % 8 == 0
    assert lib.pxh_ensemble_workspace_size(None) < 0
    assert lib.pxh_ensemble_plan(_hip.PXH_F32, 0, mem, 0, None, 0) < 0           # no member
    mem[1].sed.mpm = 0                                                             # sediment modes differ
    assert lib.pxh_ensemble_plan(_hip.PXH_F32, 0, mem, 3, buf, size) < 0
    mem[1].sed.mpm = 1
    mem[2].ny = 1                                                                  # not a grid
    assert lib.pxh_ensemble_plan(_hip.PXH_F32, 0, mem, 3, buf, size) < 0


def test_adjoint_programs_and_workspace():
    """The adjoint's ring-fill program holds the forward entries sorted by
    (group, src), flagged PXH_BC_SRC_SORTED; pxh_adj_workspace_size (host-only)
    depends on the grid alone (the transposed fill moves each run of equal
    sources in one pass, with no scratch per entry) and refuses float32 grids
    (the adjoint is double only)."""
    for mk in (lambda: cases.calibration_channel(nx=40, ny=20), lambda: cases.periodic_box(nx=24, ny=20)):
        mesh, arr, params = mk()
        bm = BoundaryManager(params, mesh.local_X, mesh.local_Y, device="cpu")
        bc = bm.adjoint_program()
        ent = bm._adj_entries
        assert bc.flags == _hip.BC_GROUPED | _hip.BC_SRC_SORTED
        key = ent["group"].astype(np.int64) * (1 << 32) + ent["src"].astype(np.int64)
        assert np.all(np.diff(key) >= 0)
        as_rows = lambda e: sorted(map(bytes, e.view(np.uint8).reshape(len(e), -1)))  # noqa: E731
        assert as_rows(ent) == as_rows(bm.entries)
        assert list(bc.group_start) == bm.group_start
        lib = _hip.load()
        g64 = _hip.Grid(20, 40, _hip.PXH_F64, 0, 0.02)
        base = lib.pxh_adj_workspace_size(g64, None)
        assert base >= 11 * 8 * 20 * 40
        assert lib.pxh_adj_workspace_size(g64, bc) == base
        assert lib.pxh_adj_workspace_size(_hip.Grid(20, 40, _hip.PXH_F32, 0, 0.02), None) == -1


def test_sediment_gradient_chain_rule_against_table_differences():
    """chain_sediment_gradients maps derivatives with respect to the table
    entries (A, B, C, 1 - porosity, nb_sed) to d50 and porosity: checked
    against central differences of sediment_table itself."""
    from pyexner_amd.adjoint import chain_sediment_gradients
    seds = seds_from_params(cases.calibration_channel(nx=8, ny=8)[2]).astype(np.float64)
    rng = np.random.default_rng(3)
    lam = rng.normal(size=5)

    def functional(s):
        t = sediment_table(s, False, np.float64)
        return lam[0] * t.A[0] + lam[1] * t.B[0] + lam[2] * t.C[0] + lam[3] * t.one_minus_por + lam[4] * t.nb_sed

    d_d50, d_por = chain_sediment_gradients(seds, lam)
    for row, want in ((1, d_d50), (6, d_por)):
        eps = 1e-6 * abs(seds[row, 0])
        sp, sm = seds.copy(), seds.copy()
        sp[row, 0] += eps
        sm[row, 0] -= eps
        fd = (functional(sp) - functional(sm)) / (2 * eps)
        assert abs(fd - want) <= 1e-7 * abs(fd), (row, fd, want)


def test_netcdf_writer_round_trip(tmp_path):
    """NetCDFWriter (pnetcdf_reader.py:178-256 write side): (t, y, x) float32
    snapshots with NaN under the nodata mask, file coordinates recovered from
    the cell centres, read back with scipy."""
    from oracle.state import ExnerState
    from pyexner_amd.mesh import uniform_mesh
    from pyexner_amd.output import NetCDFWriter, read_snapshots
    ny, nx = 6, 9
    mesh = uniform_mesh(ny, nx, 0.1)
    rng = np.random.default_rng(5)
    f = lambda: rng.normal(size=(ny, nx))  # noqa: E731
    mask = np.zeros((ny, nx), dtype=bool)
    mask[0, :] = mask[:, -1] = True
    states = [ExnerState(h=f(), hu=f(), hv=f(), z=f(), z_b=f(), n=f(), n_b=f(), G=f(), seds=np.zeros((7, 1)))
              for _ in range(2)]
    path = str(tmp_path / "out.nc")
    with NetCDFWriter(path, mesh) as w:
        for s in states:
            w.write_state(s, mesh, mask)
    x, y, out = read_snapshots(path)
    np.testing.assert_allclose(x + 0.05, mesh.local_X[0], atol=1e-12)
    np.testing.assert_allclose(y - 0.05, mesh.local_Y[:, 0], atol=1e-12)
    assert set(out) == {"h", "hu", "hv", "z", "z_b", "n", "n_b", "G"}
    for name, arr in out.items():
        assert arr.shape == (2, ny, nx) and arr.dtype == np.float32
        for k, s in enumerate(states):
            want = np.where(mask, np.nan, getattr(s, name)).astype(np.float32)
            np.testing.assert_array_equal(arr[k], want)


def test_strip_height_flags_are_checked():
    """PXH_STRIP_ROWS (pxh_grid.flags bits 8-15): 0 or 8..128; the dt
    workspace grows with the strip count (host-only size queries)."""
    lib = _hip.load()
    sh = _hip.STRIP_ROWS_SHIFT
    sizes = {}
    for rows in (0, 8, 32, 128):
        sizes[rows] = lib.pxh_dt_workspace_size(_hip.Grid(4000, 3000, _hip.PXH_F32, rows << sh, 0.1))
        assert sizes[rows] > 0
    assert sizes[0] == sizes[32] and sizes[8] > sizes[32]
    for rows in (4, 129, 255):
        assert lib.pxh_dt_workspace_size(_hip.Grid(400, 300, _hip.PXH_F32, rows << sh, 0.1)) == -1


def test_grid_flag_fields_match_the_c_header(tmp_path):
    """The pxh_grid.flags fields pyexner_amd._hip packs (strip height,
    occupancy caps, NO_NODATA, EXACT_F32) are where the header's macros read
    them."""
    flags = (_hip.EXACT_F32 | _hip.NO_NODATA | (24 << _hip.STRIP_ROWS_SHIFT) | (3 << _hip.STAGE1_WAVES_SHIFT)
             | (2 << _hip.STAGE2_WAVES_SHIFT))
    src = tmp_path / "flags.c"
    src.write_text('#include <stdio.h>\n#include <stdint.h>\n#include "pyexner_hip.h"\nint main(void) {\n'
                   f'    int f = {flags};\n'
                   '    printf("%d %d %d %d %d\\n", PXH_STRIP_ROWS(f), PXH_STAGE1_WAVES(f), PXH_STAGE2_WAVES(f),\n'
                   '           (f & PXH_NO_NODATA) != 0, (f & PXH_EXACT_F32) != 0);\n    return 0;\n}\n')
    exe = tmp_path / "flags"
    subprocess.run(["gcc", "-std=c99", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True,
                   timeout=60)
    got = subprocess.run([str(exe)], capture_output=True, text=True, check=True, timeout=60).stdout.split()
    assert [int(x) for x in got] == [24, 3, 2, 1, 1]


def test_dt_workspace_holds_the_seam_columns():
    """The dt workspace ends with the combine sweeps' seam-column buffer: per
    60-column strip and row, a record of the combined h, hu, hv, G and the
    nodata flag of the strip's first and last owned column (10 values), in
    the grid's dtype."""
    lib = _hip.load()
    ny, nx = 4000, 3000
    f32 = lib.pxh_dt_workspace_size(_hip.Grid(ny, nx, _hip.PXH_F32, 0, 0.1))
    f64 = lib.pxh_dt_workspace_size(_hip.Grid(ny, nx, _hip.PXH_F64, 0, 0.1))
    cols = -(-nx // 60) * 2 * 5 * ny
    assert f64 - f32 == cols * 4
    assert f32 > cols * 4


def test_copy_batch_checks_its_arguments():
    """pxh_copy_batch (batched tile strips on one device) rejects a bad dtype,
    a negative count or total, a missing table and elements without jobs, and
    does nothing for an empty batch (host-only checks: no job runs)."""
    lib = _hip.load()
    assert lib.pxh_copy_batch(7, None, 0, 0, None) == _hip.PXH_E_ARG
    assert lib.pxh_copy_batch(_hip.PXH_F32, None, -1, 0, None) == _hip.PXH_E_ARG
    assert lib.pxh_copy_batch(_hip.PXH_F32, None, 1, 10, None) == _hip.PXH_E_ARG
    assert lib.pxh_copy_batch(_hip.PXH_F32, None, 0, 10, None) == _hip.PXH_E_ARG
    assert lib.pxh_copy_batch(_hip.PXH_F64, None, 0, -1, None) == _hip.PXH_E_ARG
    assert lib.pxh_copy_batch(_hip.PXH_F64, None, 0, 0, None) == _hip.PXH_OK
    assert ctypes.sizeof(_hip.CopyJob) == 48


def test_dt_multi_checks_its_arguments():
    """pxh_dt_multi (compute_dt over the tiles of one device) rejects a bad
    dtype, an empty domain table or block count and missing buffers; the block
    count of a domain is its dt sweep's (host-only checks: no sweep runs)."""
    lib = _hip.load()
    assert ctypes.sizeof(_hip.DtDomain) == 64
    assert lib.pxh_dt_multi_blocks(1, 10) == _hip.PXH_E_ARG
    nb = lib.pxh_dt_multi_blocks(2054, 2054)
    assert nb == -(-2054 // 62) * -(-(-(-2054 // 16)) // 2)
    p = ctypes.c_void_p(16)  # never dereferenced: every call below fails its checks first
    assert lib.pxh_dt_multi(5, 0, p, 1, nb, 0.5, p, p, None) == _hip.PXH_E_ARG
    assert lib.pxh_dt_multi(_hip.PXH_F32, 0, None, 1, nb, 0.5, p, p, None) == _hip.PXH_E_ARG
    assert lib.pxh_dt_multi(_hip.PXH_F32, 0, p, 0, nb, 0.5, p, p, None) == _hip.PXH_E_ARG
    assert lib.pxh_dt_multi(_hip.PXH_F32, 0, p, 1, 0, 0.5, p, p, None) == _hip.PXH_E_ARG
    assert lib.pxh_dt_multi(_hip.PXH_F64, 0, p, 1, nb, 0.5, None, p, None) == _hip.PXH_E_ARG
    assert lib.pxh_dt_multi(_hip.PXH_F64, 0, p, 1, nb, 0.5, p, None, None) == _hip.PXH_E_ARG


def test_multi_domain_sweeps_check_their_arguments():
    """The multi-domain fused sweeps (tiles on one device) size their launches
    with the single-domain geometry and reject bad scalars before any launch
    (host-only checks: no sweep runs)."""
    lib = _hip.load()
    assert ctypes.sizeof(_hip.SweepDomain) == 264
    sh = _hip.STRIP_ROWS_SHIFT
    assert lib.pxh_multi_stage_blocks(_hip.Grid(2054, 2054, _hip.PXH_F32, 12 << sh, 0.05)) == 35 * 172
    assert lib.pxh_multi_stage_blocks(_hip.Grid(2054, 2054, _hip.PXH_F32, 0, 0.05)) == 35 * 65
    assert lib.pxh_multi_stage_blocks(_hip.Grid(2054, 2054, _hip.PXH_F32, 4 << sh, 0.05)) == _hip.PXH_E_ARG
    assert lib.pxh_multi_frame_blocks(2054, 2054) == -(-(2 * 2 * 2054 + 2 * 2 * 2050) // 256)
    assert lib.pxh_multi_frame_blocks(1, 5) == _hip.PXH_E_ARG
    sed = _hip.Sediments()
    p = ctypes.c_void_p(16)  # never dereferenced: every call below fails its checks first
    assert lib.pxh_roe_exner_stage_sweep_multi(9, 0, p, 1, 8, 0, 0, sed, p, 0, None, None, None) == _hip.PXH_E_ARG
    assert lib.pxh_roe_exner_stage_sweep_multi(_hip.PXH_F32, 0, None, 1, 8, 0, 0, sed, p, 0, None, None, None) == \
        _hip.PXH_E_ARG
    assert lib.pxh_roe_exner_stage_sweep_multi(_hip.PXH_F32, 0, p, 0, 8, 0, 0, sed, p, 0, None, None, None) == \
        _hip.PXH_E_ARG
    assert lib.pxh_roe_exner_stage_sweep_multi(_hip.PXH_F32, 200 << sh, p, 1, 8, 0, 0, sed, p, 0, None, None, None) == \
        _hip.PXH_E_ARG
    assert lib.pxh_roe_exner_stage_sweep_multi(_hip.PXH_F32, 0, p, 1, 8, 0, 0, sed, None, 0, None, None, None) == \
        _hip.PXH_E_ARG
    # the dt buffers: both or neither, and only for the second stage
    assert lib.pxh_roe_exner_stage_sweep_multi(_hip.PXH_F32, 0, p, 1, 8, 1, 0, sed, p, 0, p, None, None) == \
        _hip.PXH_E_ARG
    assert lib.pxh_roe_exner_stage_sweep_multi(_hip.PXH_F32, 0, p, 1, 8, 0, 0, sed, p, 0, p, p, None) == \
        _hip.PXH_E_ARG
    # the tiles' frame dt: the frame's cells and the strip seams' edges (more
    # seams with shorter strips); seam records of 10 values per strip column
    # and row
    g24 = _hip.Grid(2054, 2054, _hip.PXH_F32, 24 << sh, 0.05)
    g12 = _hip.Grid(2054, 2054, _hip.PXH_F32, 12 << sh, 0.05)
    assert lib.pxh_multi_dt_frame_blocks(g12) > lib.pxh_multi_dt_frame_blocks(g24) > \
        lib.pxh_multi_frame_blocks(2054, 2054)
    assert lib.pxh_multi_seam_elems(2054, 2054) == -(-2054 // 60) * 2054 * 10
    assert lib.pxh_multi_seam_elems(1, 9) == _hip.PXH_E_ARG
    assert lib.pxh_dt_frame_multi(_hip.PXH_F32, 0, p, 1, 8, 4, 0.5, None, p, p, None) == _hip.PXH_E_ARG
    assert lib.pxh_dt_frame_multi(_hip.PXH_F32, 0, p, 1, 8, 4, 0.5, p, None, p, None) == _hip.PXH_E_ARG
    assert lib.pxh_dt_frame_multi(_hip.PXH_F32, 0, p, 1, 0, 4, 0.5, p, p, p, None) == _hip.PXH_E_ARG
    assert lib.pxh_dt_frame_multi(_hip.PXH_F32, 0, None, 1, 8, 4, 0.5, p, p, p, None) == _hip.PXH_E_ARG
    assert lib.pxh_roe_exner_frame_sweep_multi(_hip.PXH_F64, 0, p, 1, 0, 0, sed, p, 0, None) == _hip.PXH_E_ARG
    assert lib.pxh_roe_exner_frame_sweep_multi(_hip.PXH_F64, 0, None, 1, 8, 0, sed, p, 0, None) == _hip.PXH_E_ARG
    sed.mpm, sed.n_sed = 1, 0  # MPM without a sediment class
    assert lib.pxh_roe_exner_frame_sweep_multi(_hip.PXH_F64, 0, p, 1, 8, 0, sed, p, 0, None) == _hip.PXH_E_RANGE


def _record_fields(base, es, packed):
    """pxh_fields of made-up addresses (never dereferenced): h, hu, hv, z_b as
    one record view (one, two, three elements past h) or as separate arrays."""
    f = _hip.Fields()
    step = es if packed else 1 << 20
    f.h, f.hu, f.hv, f.z_b = base, base + step, base + 2 * step, base + 3 * step
    f.z, f.n, f.n_b, f.G = (base + (k << 24) for k in range(4, 8))
    return f


def test_record_views_only_where_the_fused_step_keeps_them():
    """A pxh_fields whose hu, hv, z_b sit one, two, three elements past h (the
    first stage's records of PXH_PACKED_STAGE) is accepted by the stage and
    frame sweeps only as the first stage's output or the second stage's input
    on the float32 hardware-math path, and refused by the other entries
    (host-only checks: every call below fails before any launch)."""
    lib = _hip.load()
    sed = _hip.Sediments()
    p = ctypes.c_void_p(16)
    g32 = _hip.Grid(64, 64, _hip.PXH_F32, 0, 0.05)
    gx = _hip.Grid(64, 64, _hip.PXH_F32, _hip.EXACT_F32, 0.05)
    g64 = _hip.Grid(64, 64, _hip.PXH_F64, 0, 0.05)
    rec, a, c = _record_fields(1 << 40, 4, True), _record_fields(2 << 40, 4, False), _record_fields(3 << 40, 4, False)
    rec64 = _record_fields(1 << 40, 8, True)
    E = _hip.PXH_E_ARG
    stage = lib.pxh_roe_exner_stage_sweep
    frame = lib.pxh_roe_exner_frame_sweep
    # stage sweep: records in without old (first stage), records out with old
    # (second stage), records in and out, old as records
    assert stage(g32, rec, a, None, p, sed, p, 0, None, None) == E
    assert stage(g32, a, rec, c, p, sed, p, 0, None, None) == E
    assert stage(g32, rec, rec, c, p, sed, p, 0, None, None) == E
    assert stage(g32, a, c, rec, p, sed, p, 0, None, None) == E
    # float64 and exact float32 keep no records
    assert stage(g64, a, rec64, None, p, sed, p, 0, None, None) == E
    assert stage(gx, a, rec, None, p, sed, p, 0, None, None) == E
    # frame sweep: first stage bed as records, hyd records with a plain out;
    # second stage hyd as records, old as records
    assert frame(g32, rec, rec, rec, None, p, sed, p, 0, 0.5, None, None) == E
    assert frame(g32, rec, a, c, None, p, sed, p, 0, 0.5, None, None) == E
    assert frame(g32, rec, a, None, c, p, sed, p, 0, 0.5, None, None) == E
    assert frame(g32, c, a, None, rec, p, sed, p, 0, 0.5, None, None) == E
    # entries that never read records
    assert lib.pxh_roe_exner_hydro_sweep(g32, rec, a, p, p, None) == E
    assert lib.pxh_roe_exner_hydro_sweep(g32, a, rec, p, p, None) == E
    assert lib.pxh_roe_exner_bed_sweep(g32, rec, a, c, None, p, sed, p, None, None) == E

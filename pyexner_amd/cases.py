"""Synthetic workloads (initial states, boundaries, parameters).

All draws come from numpy.random.default_rng(seed) in the order listed in each
builder.  Bed bumps are z_b = sum_k A_k exp(-((x-xc_k)^2 + (y-yc_k)^2) / (2 sigma_k^2)),
each bump evaluated on the cells within 8 sigma_k of its centre (exactly zero
beyond; exp(-32) ~ 1e-14).

* dambreak_1d  - plain Roe dam break on a wet bed, 512 x 32, dx = 0.1 m,
  h = 1.0 for x < 25.6 m else 0.5, Reflective ring (the Stoker case).
* channel      - Roe-Exner channel, 2048 x 512, dx = 0.02 m, 32 bumps, h = 0.1,
  u = 0.5, MPM sediment, Transmissive in/outflow, Reflective side walls.
* radial_dambreak - Roe-Exner radial dam break, 8192 x 8192, dx = 0.05 m,
  h = 2.0 inside r < 50 m else 0.5, 256 bumps, MPM sediment, Reflective ring.
Every builder takes (nx, ny) so tests can shrink the same physics.
"""
import numpy as np

from .mesh import uniform_mesh

MPM_SEDIMENT = {"sed1": {"fraction": 1.0, "diameter": 1.61e-3, "density": 2630.0, "erosion_flux": 0.04,
                         "deposition_flux": 0.01}}


def ring_boundaries(mesh, kinds):
    """Axis-aligned boundary polygons around the outer ring of cells.

    kinds: list of (name, type, side) with side in W, E, S, N; polygons span
    half a cell either side of the ring's cell centres and the full other
    extent (corners included), as tests/runtime/2D_dambreak/build_scaling_cases.py:7-55 lays them out."""
    X, Y, dh = mesh.local_X, mesh.local_Y, mesh.dh
    x0, x1 = X[0, 0], X[0, -1]
    yb, yt = Y[-1, 0], Y[0, 0]
    hh = dh / 2
    full_x = (x0 - hh, x1 + hh)
    full_y = (yb - hh, yt + hh)
    rect = lambda xa, xb, ya, yb_: [[xa, ya], [xb, ya], [xb, yb_], [xa, yb_]]  # noqa: E731
    side = {
        "W": (rect(x0 - hh, x0 + hh, *full_y), [-1.0, 0.0]),
        "E": (rect(x1 - hh, x1 + hh, *full_y), [1.0, 0.0]),
        "S": (rect(*full_x, yb - hh, yb + hh), [0.0, -1.0]),
        "N": (rect(*full_x, yt - hh, yt + hh), [0.0, 1.0]),
    }
    out = {}
    for name, typ, sd in kinds:
        poly, normal = side[sd]
        out[name] = {"type": typ, "polygon": [[float(a), float(b)] for a, b in poly],
                     "values": [0.0, 0.0, 0.0, 0.0], "normal": normal}
    return out


def gaussian_bumps(X, Y, rng, nb, amp=(0.0, 0.02), sigma=(0.05, 0.3)):
    dh = abs(X[0, 1] - X[0, 0])
    x0, x1 = X[0, 0], X[0, -1]
    ymin, ymax = Y[-1, 0], Y[0, 0]
    xc = rng.uniform(x0, x1, nb)
    yc = rng.uniform(ymin, ymax, nb)
    A = rng.uniform(amp[0], amp[1], nb)
    sg = rng.uniform(sigma[0], sigma[1], nb)
    zb = np.zeros(X.shape)
    ny, nx = X.shape
    for k in range(nb):
        r = 8.0 * sg[k]
        c0 = max(int(np.floor((xc[k] - r - x0) / dh)), 0)
        c1 = min(int(np.ceil((xc[k] + r - x0) / dh)) + 1, nx)
        r0 = max(int(np.floor((ymax - (yc[k] + r)) / dh)), 0)
        r1 = min(int(np.ceil((ymax - (yc[k] - r)) / dh)) + 1, ny)
        if c0 >= c1 or r0 >= r1:
            continue
        xs, ys = X[r0:r1, c0:c1], Y[r0:r1, c0:c1]
        d2 = (xs - xc[k]) ** 2 + (ys - yc[k]) ** 2
        zb[r0:r1, c0:c1] += np.where(d2 <= r * r, A[k] * np.exp(-d2 / (2 * sg[k] ** 2)), 0.0)
    return zb


def dambreak_1d(nx=512, ny=32, dx=0.1, x_split=25.6, hl=1.0, hr=0.5):
    mesh = uniform_mesh(ny, nx, dx)
    X = mesh.local_X
    h = np.where(X < x_split, hl, hr)
    zeros = np.zeros_like(h)
    arrays = {"h": h, "hu": zeros.copy(), "hv": zeros.copy(), "z": zeros.copy(), "n": zeros.copy()}
    params = {"flux_scheme": "Roe", "integrator": "Forward Euler", "cfl": 0.5, "end_time": 2.0,
              "out_freq": 2.0, "parNx": 1, "parNy": 1,
              "boundaries": ring_boundaries(mesh, [("west", "Reflective", "W"), ("east", "Reflective", "E"),
                                                   ("south", "Reflective", "S"), ("north", "Reflective", "N")])}
    return mesh, arrays, params


def roe_stepped_bed(nx=64, ny=40, dx=0.05, seed=21, level=0.5, end_time=0.05, film_range=(0.05, 0.12)):
    """Plain Roe over a bed of flat blocks (steps at block edges), a still
    level with thin films on the blocks that stand above it and a gentle
    random current: the edges take the step-face and wall thrusts and the
    source fix of roe.py:132-189.  Every cell stays wet (two dry neighbours
    give 0/0 in the Roe average, roe.py:66)."""
    mesh = uniform_mesh(ny, nx, dx)
    rng = np.random.default_rng(seed)
    by, bx = 5, 8
    blocks = rng.uniform(0.0, 0.8, (ny // by + 1, nx // bx + 1))
    z = np.kron(blocks, np.ones((by, bx)))[:ny, :nx]
    film = rng.uniform(film_range[0], film_range[1], (ny, nx))
    h = np.maximum(level - z, film)
    still = (level - z) < film  # films stand still
    u = np.where(still, 0.0, rng.uniform(-0.3, 0.3, (ny, nx)) * np.sqrt(9.81 * h))
    v = np.where(still, 0.0, rng.uniform(-0.3, 0.3, (ny, nx)) * np.sqrt(9.81 * h))
    arrays = {"h": h, "hu": h * u, "hv": h * v, "z": z, "n": np.full_like(h, 0.02)}
    params = {"flux_scheme": "Roe", "integrator": "Forward Euler", "cfl": 0.5, "end_time": end_time,
              "out_freq": end_time, "parNx": 1, "parNy": 1,
              "boundaries": ring_boundaries(mesh, [("west", "Reflective", "W"), ("east", "Reflective", "E"),
                                                   ("south", "Reflective", "S"), ("north", "Reflective", "N")])}
    return mesh, arrays, params


def _exner_params(mesh, bounds, integrator="SSPRK2"):
    return {"flux_scheme": "Roe Exner", "integrator": integrator, "cfl": 0.5, "end_time": 1.0, "out_freq": 1.0,
            "parNx": 1, "parNy": 1,
            "erosion": {"grass_factor": "MPM", "bulk_porosity": 0.42, "sediments": MPM_SEDIMENT},
            "boundaries": bounds}


def channel(nx=2048, ny=512, dx=0.02, n_bumps=32, seed=2, h0=0.1, u0=0.5):
    mesh = uniform_mesh(ny, nx, dx)
    rng = np.random.default_rng(seed)
    zb = gaussian_bumps(mesh.local_X, mesh.local_Y, rng, n_bumps)
    h = np.full(zb.shape, h0)
    arrays = {"h": h, "hu": h * u0, "hv": np.zeros_like(h), "z": np.zeros_like(h), "z_b": zb,
              "n": np.zeros_like(h), "G": np.zeros_like(h)}
    bounds = ring_boundaries(mesh, [("inlet", "Transmissive", "W"), ("outlet", "Transmissive", "E"),
                                    ("inlet_bed", "TransmissiveBed", "W"), ("outlet_bed", "TransmissiveBed", "E"),
                                    ("south", "Reflective", "S"), ("north", "Reflective", "N")])
    return mesh, arrays, _exner_params(mesh, bounds)


def radial_dambreak(nx=8192, ny=8192, dx=0.05, n_bumps=256, seed=3, radius=50.0, h_in=2.0, h_out=0.5):
    mesh = uniform_mesh(ny, nx, dx)
    rng = np.random.default_rng(seed)
    X, Y = mesh.local_X, mesh.local_Y
    zb = gaussian_bumps(X, Y, rng, n_bumps)
    xm = 0.5 * (X[0, 0] + X[0, -1])
    ym = 0.5 * (Y[0, 0] + Y[-1, 0])
    h = np.where((X - xm) ** 2 + (Y - ym) ** 2 < radius * radius, h_in, h_out)
    zeros = np.zeros_like(h)
    arrays = {"h": h, "hu": zeros.copy(), "hv": zeros.copy(), "z": zeros.copy(), "z_b": zb, "n": zeros.copy(),
              "G": zeros.copy()}
    bounds = ring_boundaries(mesh, [("west", "Reflective", "W"), ("east", "Reflective", "E"),
                                    ("south", "Reflective", "S"), ("north", "Reflective", "N")])
    return mesh, arrays, _exner_params(mesh, bounds)


def wetdry_obstacle(nx=150, ny=70, dx=0.05, seed=9):
    """Dam break onto a dry bed with an emergent block and a sloping erodible
    layer: wet/dry fronts, cells whose water surface is below a neighbour's
    bed, and a width/height that are not multiples of the sweep strips."""
    mesh = uniform_mesh(ny, nx, dx)
    rng = np.random.default_rng(seed)
    X, Y = mesh.local_X, mesh.local_Y
    Lx = X[0, -1] - X[0, 0]
    h = np.where(X < X[0, 0] + Lx / 3, 0.4, 0.0)
    z = np.zeros_like(h)
    blk = (np.abs(X - (X[0, 0] + 0.55 * Lx)) < 0.3) & (np.abs(Y - Y[:, 0].mean()) < 0.5)
    z[blk] = 0.8
    zb = gaussian_bumps(X, Y, rng, 6, amp=(0.0, 0.05), sigma=(0.1, 0.4)) + 0.02 * (X - X[0, 0]) / Lx
    h = np.where(blk, 0.0, h)
    arrays = {"h": h, "hu": np.zeros_like(h), "hv": np.zeros_like(h), "z": z, "z_b": zb,
              "n": np.full_like(h, 0.02), "G": np.zeros_like(h)}
    bounds = ring_boundaries(mesh, [("west", "Reflective", "W"), ("east", "Transmissive", "E"),
                                    ("east_bed", "TransmissiveBed", "E"), ("south", "Reflective", "S"),
                                    ("north", "Reflective", "N")])
    return mesh, arrays, _exner_params(mesh, bounds)


def periodic_box(nx=64, ny=48, dx=0.05, seed=11, exner=True):
    """Smooth periodic-compatible state for translation/conservation checks."""
    mesh = uniform_mesh(ny, nx, dx)
    rng = np.random.default_rng(seed)
    X, Y = mesh.local_X, mesh.local_Y
    Lx, Ly = (nx - 2) * dx, (ny - 2) * dx
    kx, ky = 2 * np.pi / Lx, 2 * np.pi / Ly
    ph = rng.uniform(0, 2 * np.pi, 4)
    h = 0.3 + 0.05 * np.sin(kx * X + ph[0]) * np.cos(ky * Y + ph[1])
    u = 0.2 + 0.1 * np.cos(kx * X + ph[2])
    v = 0.1 * np.sin(ky * Y + ph[3])
    zb = 0.01 + 0.005 * np.cos(kx * X) * np.sin(ky * Y)
    zeros = np.zeros_like(h)
    arrays = {"h": h, "hu": h * u, "hv": h * v, "z": zeros.copy(), "z_b": zb, "n": zeros.copy(),
              "G": zeros.copy()}
    kinds = [("west", "Periodic", "W"), ("east", "Periodic", "E"), ("south", "Periodic", "S"),
             ("north", "Periodic", "N")]
    bounds = ring_boundaries(mesh, kinds)
    if exner:
        return mesh, arrays, _exner_params(mesh, bounds)
    params = {"flux_scheme": "Roe", "integrator": "Forward Euler", "cfl": 0.5, "end_time": 1.0,
              "out_freq": 1.0, "parNx": 1, "parNy": 1, "boundaries": bounds}
    return mesh, {k: arrays[k] for k in ("h", "hu", "hv", "z", "n")}, params


def calibration_channel(nx=1024, ny=256, dx=0.02, n_bumps=8, seed=5):
    """The channel of the bed-calibration case: channel boundaries and MPM
    sediment, with depth >= 0.05 everywhere (0.2 +- 0.03 m), a smoothly
    varying subcritical flow (0.8 +- 0.1 m/s, Froude <= 0.7) whose Shields
    stress stays well above critical, Manning n = 0.02, and an erodible layer
    on a tilted base (0.05 m plus slopes in x and y plus low bumps), so that
    every cell carries sediment and neighbouring beds differ by more than
    SED_TOL: the regime where the step is differentiable almost everywhere."""
    mesh = uniform_mesh(ny, nx, dx)
    rng = np.random.default_rng(seed)
    X, Y = mesh.local_X, mesh.local_Y
    Lx = X[0, -1] - X[0, 0]
    Ly = Y[0, 0] - Y[-1, 0]
    xs, ys = (X - X[0, 0]) / Lx, (Y - Y[-1, 0]) / Ly
    ph = rng.uniform(0, 2 * np.pi, 4)
    h = 0.2 + 0.03 * np.sin(2 * np.pi * xs + ph[0]) * np.cos(np.pi * ys + ph[1])
    u = 0.8 + 0.1 * np.cos(2 * np.pi * xs + ph[2])
    v = 0.05 * np.sin(np.pi * ys + ph[3])
    # base slopes 1% (x, falling downstream) and 0.8% (y): 2e-4 and 1.6e-4 m
    # per cell at dx = 0.02, above any bump's slope, so |dz_b| >> SED_TOL
    zb = (0.05 + 0.01 * (X[0, -1] - X) + 0.008 * (Y - Y[-1, 0])
          + gaussian_bumps(X, Y, rng, n_bumps, amp=(0.0, 0.001), sigma=(0.3, 0.6)))
    arrays = {"h": h, "hu": h * u, "hv": h * v, "z": np.zeros_like(h), "z_b": zb, "n": np.full_like(h, 0.02),
              "G": np.zeros_like(h)}
    bounds = ring_boundaries(mesh, [("inlet", "Transmissive", "W"), ("outlet", "Transmissive", "E"),
                                    ("inlet_bed", "TransmissiveBed", "W"), ("outlet_bed", "TransmissiveBed", "E"),
                                    ("south", "Reflective", "S"), ("north", "Reflective", "N")])
    return mesh, arrays, _exner_params(mesh, bounds)


def _with_values(bounds, name, values):
    bounds[name] = dict(bounds[name], values=[float(v) for v in values])
    return bounds


def berthon_channel(nx=140, ny=22, dx=0.05, gamma=1.2, grass=0.005, integrator="Forward Euler"):
    """The Berthon et al. steady state of tests/runtime/2D_const_flux
    (2D_build_num_sol.py:8-56: q = 1, A = alpha = beta = 0.005, constant G)
    on one-cell ring polygons: Berthon bed and ConstantInflux at the inlet,
    NormalFlowDepth at the outlet, Reflective walls.  gamma lifts the bed so
    that it carries sediment (the reference's gamma = 1 leaves it below
    SED_TOL, where the Exner update is off)."""
    mesh = uniform_mesh(ny, nx, dx)
    X = mesh.local_X
    A = alpha = beta = 0.005
    q, g = 1.0, 9.81
    xf = X - dx / 2  # file x, where the reference evaluates the profile
    u = ((alpha * xf + beta) / A) ** (1.0 / 3.0)
    h = q / u
    zb = -(u ** 3 + 2 * g * q) / (2 * g * u) + gamma
    arrays = {"h": h, "hu": h * u, "hv": np.zeros_like(h), "z": np.zeros_like(h), "z_b": zb,
              "n": np.zeros_like(h)}
    bounds = ring_boundaries(mesh, [("bottom", "Reflective", "S"), ("inlet_bed", "Berthon", "W"),
                                    ("inlet_flux", "ConstantInflux", "W"), ("outlet_bed", "NormalFlowDepth", "E"),
                                    ("top", "Reflective", "N")])
    prof = [q, A, alpha, beta, gamma]
    _with_values(bounds, "inlet_bed", prof + [float(xf[0, 0])])
    _with_values(bounds, "inlet_flux", [q, 1.0, 0.0])
    _with_values(bounds, "outlet_bed", prof + [float(xf[0, -1])])
    params = {"flux_scheme": "Roe Exner", "integrator": integrator, "cfl": 0.5, "end_time": 1.0, "out_freq": 1.0,
              "parNx": 1, "parNy": 1, "erosion": {"grass_factor": grass}, "boundaries": bounds}
    return mesh, arrays, params


def outflux_channel(nx=120, ny=30, dx=0.02, seed=12, q=0.05, steep_fall=False):
    """An MPM channel with a prescribed discharge: ConstantInflux and
    TransmissiveBed at the inlet, ConstantOutflux and TransmissiveBed at the
    outlet (with steep_fall, SteepFall instead: the bed one metre below the
    interior's, which drains this shallow subcritical channel within a few
    hundred steps), a Reflective north wall, and a south wall that reflects and
    then zeroes the momentum (ZeroMomentum)."""
    mesh = uniform_mesh(ny, nx, dx)
    rng = np.random.default_rng(seed)
    zb = 0.02 + gaussian_bumps(mesh.local_X, mesh.local_Y, rng, 6, amp=(0.0, 0.01))
    h = np.full(zb.shape, 0.1)
    arrays = {"h": h, "hu": h * 0.5, "hv": np.zeros_like(h), "z": np.zeros_like(h), "z_b": zb,
              "n": np.full_like(h, 0.02), "G": np.zeros_like(h)}
    bounds = ring_boundaries(mesh, [("inlet", "ConstantInflux", "W"), ("inlet_bed", "TransmissiveBed", "W"),
                                    ("outlet", "ConstantOutflux", "E"),
                                    ("outlet_bed", "SteepFall" if steep_fall else "TransmissiveBed", "E"),
                                    ("north", "Reflective", "N"), ("south", "Reflective", "S"),
                                    ("south_wall", "ZeroMomentum", "S")])
    _with_values(bounds, "inlet", [q, 1.0, 0.0])
    _with_values(bounds, "outlet", [q, 1.0, 0.0])
    return mesh, arrays, _exner_params(mesh, bounds)


def nodata_island(nx=96, ny=40, dx=0.02, seed=13):
    """The calibration channel's flow around a nodata island: z = NaN on an
    interior block (isnan(z) is nodata, roe_exner_solver.py:21-39), as the
    reference's erodible channel marks the cells outside its flume
    (erodible_channel.py:124-129)."""
    mesh, arrays, params = calibration_channel(nx=nx, ny=ny, dx=dx, seed=seed)
    z = np.array(arrays["z"], dtype=np.float64)
    r0, c0 = ny // 3, nx // 3
    z[r0:r0 + ny // 4, c0:c0 + nx // 6] = np.nan
    arrays = dict(arrays, z=z)
    return mesh, arrays, params


def berthon_convergence(dh=0.05, end_time=10.0, out_freq=1.0, integrator="Forward Euler"):
    """The reference's Berthon convergence case as its own inputs give it
    (tests/runtime/2D_const_flux/2D_build_num_sol.py:8-56 and
    dh_*/input_*.yaml): x = 0..7 and y = 1..0 in steps of dh, stored as
    float32 file coordinates; q = 1, A = alpha = beta = 0.005, gamma = 1,
    constant G = 0.005; Reflective bottom and top polygons two rows of cell
    centres thick (y within dh of 0 and of 1), Berthon bed and ConstantInflux
    at the inlet column, NormalFlowDepth at the outlet.  The thick polygons
    make a ring fill that reads cells it writes: the handlers apply one
    after another (boundary_registry.py:183-187), which the composed program
    keeps with two dependency levels."""
    from .mesh import mesh_from_coords
    x = np.arange(0.0, 7.0 + dh, dh)
    y = np.arange(1.0, 0.0 - dh, -dh)
    X, _ = np.meshgrid(x, y, indexing="xy")
    A = alpha = beta = 0.005
    gamma, q0, g = 1.0, 1.0, 9.81
    u = ((alpha * X + beta) / A) ** (1.0 / 3.0)
    q = q0 * np.ones_like(X)
    h = q / u
    zb = -(u ** 3 + 2 * g * q) / (2 * g * u) + gamma
    arrays = {"h": h, "hu": h * u, "hv": np.zeros_like(h), "z": np.zeros_like(h), "z_b": zb,
              "n": np.zeros_like(h)}
    mesh = mesh_from_coords(x.astype(np.float32), y.astype(np.float32))
    d = float(dh)
    top, right = 1.0, 7.0 + d
    prof = [1.0, 0.005, 0.005, 0.005, 1]
    bounds = {
        "bottom": {"type": "Reflective", "polygon": [[0.0, -d], [right, -d], [right, d], [0.0, d]],
                   "values": [0, 0, 0, 0], "normal": [0.0, -1.0]},
        "inlet_bed": {"type": "Berthon", "polygon": [[0.0, 0.0], [d, 0.0], [d, top], [0.0, top]],
                      "values": prof + [d / 2], "normal": [-1.0, 0.0]},
        "inlet_flux": {"type": "ConstantInflux", "polygon": [[0.0, 0.0], [d, 0.0], [d, top], [0.0, top]],
                       "values": [1.0, 1.0, 0.0], "normal": [-1.0, 0.0]},
        "outlet_bed": {"type": "NormalFlowDepth",
                       "polygon": [[7.0 - d, 0.0], [right, 0.0], [right, top + d], [7.0 - d, top + d]],
                       "values": prof + [7.0 - d / 2], "normal": [1.0, 0.0]},
        "top": {"type": "Reflective", "polygon": [[0.0, top + d], [right, top + d], [right, top - d], [0.0, top - d]],
                "values": [0, 0, 0, 0], "normal": [0.0, 1.0]},
    }
    params = {"flux_scheme": "Roe Exner", "integrator": integrator, "cfl": 0.5, "end_time": end_time,
              "out_freq": out_freq, "parNx": 1, "parNy": 1, "erosion": {"grass_factor": 0.005},
              "boundaries": bounds}
    return mesh, arrays, params

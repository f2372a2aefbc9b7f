"""ctypes binding of libpyexner_hip.so (include/pyexner_hip.h).

The library is built in-tree by ``pyexner_amd.build.build()`` into
``pyexner_amd/lib/``.  There is no fallback: if the library is missing or a
call fails, an exception is raised.
"""
import ctypes
import os

# PYEXNER_HIP_LIB names another build of the same sources (the host-sanitized
# one of pyexner_amd.build.build_host_sanitized, for the tests run under it)
_LIB_PATH = os.environ.get("PYEXNER_HIP_LIB") or os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib",
                                                              "libpyexner_hip.so")

PXH_F32 = 0
PXH_F64 = 1
MAX_SED = 8
MAX_HANDLERS = 16
MAX_CHAIN = 6

PXH_OK, PXH_E_ARG, PXH_E_RANGE, PXH_E_UNSUPPORTED = 0, -1, -2, -3
STRIP_ROWS_SHIFT = 8  # pxh_grid.flags bits 8-15: owned rows per fused strip (0: 32)
# pxh_grid.flags bits 16-19 / 20-23: wavefronts per SIMD of the first / second
# stage's fused sweep (0: no cap)
STAGE1_WAVES_SHIFT, STAGE2_WAVES_SHIFT = 16, 20
BC_COPY, BC_REFLECT_RX, BC_REFLECT_ROE = 0, 1, 2
BC_FRAME = 1  # pxh_bc.flags: every target on the outer ring, every source within two cells of the edge
BC_GROUPED = 2  # pxh_bc.flags: entries ordered by group, group_start[] valid
BC_SRC_SORTED = 4  # pxh_bc.flags: within a group, entries ordered by src (adjoint programs)
BC_VALUE = 0x80  # pxh_bc_entry.nchain flag: a value entry, chain[0] indexes pxh_bc.values
BV_CONST, BV_AFFINE_T, BV_OFFSET, BV_OUTFLUX = 1, 2, 3, 4
MAX_VALUES = 32
MAX_LEVELS = 16  # PXH_MAX_LEVELS: dependency levels of a ring-fill program
N_GROUPS = 7
NOALLOC_SCRATCH_BYTES = 64 << 20  # PXH_NOALLOC_SCRATCH_BYTES
HALO_W, HALO_N, HALO_E, HALO_S = 1, 2, 4, 8  # pxh_halo_* sides
SWEEP_SKIP_G = 1  # flags of the stage/frame sweeps: leave out.G unwritten in MPM mode
G_H, G_HUV, G_Z, G_ZB, G_N, G_NB, G_G = range(7)
GROUP_FIELDS = {G_H: ("h",), G_HUV: ("hu", "hv"), G_Z: ("z",), G_ZB: ("z_b",), G_N: ("n",),
                G_NB: ("n_b",), G_G: ("G",)}


EXACT_F32 = 1
NO_NODATA = 2  # pxh_grid.flags: mask0 is zero in every cell
Z_UNIFORM = 4  # pxh_grid.flags: z holds one value (z[0]) in every cell
PACKED_STAGE = 8  # pxh_grid.flags: b.h holds 4 ny nx elements, the first stage's output as h, hu, hv, z_b records


class Grid(ctypes.Structure):
    _fields_ = [("ny", ctypes.c_int32), ("nx", ctypes.c_int32), ("dtype", ctypes.c_int32),
                ("flags", ctypes.c_int32), ("dx", ctypes.c_double)]


FIELD_NAMES = ("h", "hu", "hv", "z", "z_b", "n", "n_b", "G", "z_b_lo")


class Fields(ctypes.Structure):
    _fields_ = [(n, ctypes.c_void_p) for n in FIELD_NAMES]


class Sediments(ctypes.Structure):
    _fields_ = [("n_sed", ctypes.c_int32), ("mpm", ctypes.c_int32), ("one_minus_por", ctypes.c_double),
                ("nb_sed", ctypes.c_double)] + [
        (n, ctypes.c_double * MAX_SED) for n in ("frac", "A", "B", "C", "D", "theta_c", "AB", "CD")]


class Member(ctypes.Structure):      # pxh_member
    _fields_ = [("off", ctypes.c_int64), ("ny", ctypes.c_int32), ("nx", ctypes.c_int32), ("dx", ctypes.c_double),
                ("sed", Sediments)]


class BCEntry(ctypes.Structure):
    _fields_ = [("target", ctypes.c_int32), ("src", ctypes.c_int32), ("group", ctypes.c_uint8),
                ("nchain", ctypes.c_uint8), ("chain", ctypes.c_uint8 * MAX_CHAIN)]


class BCHandler(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("reserved", ctypes.c_int32), ("nx", ctypes.c_double),
                ("ny", ctypes.c_double)]


class SweepDomain(ctypes.Structure):  # pxh_sweep_domain ("in" is a Python keyword)
    _fields_ = [("in_", Fields), ("out", Fields), ("old", Fields), ("mask0", ctypes.c_void_p),
                ("ny", ctypes.c_int32), ("nx", ctypes.c_int32), ("sblk0", ctypes.c_int32), ("fblk0", ctypes.c_int32),
                ("dx", ctypes.c_double), ("xs0", ctypes.c_int64), ("tblk0", ctypes.c_int32),
                ("shared", ctypes.c_uint32)]


class DtDomain(ctypes.Structure):    # pxh_dt_domain
    _fields_ = [("h", ctypes.c_void_p), ("hu", ctypes.c_void_p), ("hv", ctypes.c_void_p), ("G", ctypes.c_void_p),
                ("mask0", ctypes.c_void_p), ("ny", ctypes.c_int32), ("nx", ctypes.c_int32), ("blk0", ctypes.c_int32),
                ("reserved", ctypes.c_int32), ("dx", ctypes.c_double)]


class CopyJob(ctypes.Structure):     # pxh_copy_job
    _fields_ = [("src", ctypes.c_void_p), ("dst", ctypes.c_void_p), ("src_ld", ctypes.c_int64),
                ("dst_ld", ctypes.c_int64), ("start", ctypes.c_int64), ("rows", ctypes.c_int32),
                ("cols", ctypes.c_int32)]


class BCValue(ctypes.Structure):     # pxh_bc_value
    _fields_ = [("kind", ctypes.c_int32), ("reserved", ctypes.c_int32), ("a", ctypes.c_double),
                ("b", ctypes.c_double), ("c", ctypes.c_double)]


class BC(ctypes.Structure):
    _fields_ = [("entries", ctypes.c_void_p), ("n_entries", ctypes.c_int32), ("n_handlers", ctypes.c_int32),
                ("group_mask", ctypes.c_uint32), ("flags", ctypes.c_int32),
                ("group_start", ctypes.c_int32 * (N_GROUPS + 1)),
                ("handlers", BCHandler * MAX_HANDLERS),
                ("time", ctypes.c_double), ("n_values", ctypes.c_int32), ("reserved", ctypes.c_int32),
                ("values", BCValue * MAX_VALUES), ("n_levels", ctypes.c_int32),
                ("level_start", ctypes.c_int32 * (MAX_LEVELS + 1))]


assert ctypes.sizeof(BCEntry) == 16

_P = ctypes.POINTER
_SIGS = {
    "pxh_version": (ctypes.c_int, []),
    "pxh_error_string": (ctypes.c_char_p, [ctypes.c_int]),
    "pxh_roe_exner_init": (ctypes.c_int, [_P(Grid), _P(Fields), ctypes.c_void_p, _P(Sediments), ctypes.c_void_p]),
    "pxh_roe_exner_step": (ctypes.c_int, [_P(Grid), _P(Fields), _P(Fields), ctypes.c_void_p, _P(Sediments), _P(BC),
                                          ctypes.c_void_p, ctypes.c_void_p]),
    "pxh_roe_exner_dt": (ctypes.c_int, [_P(Grid), _P(Fields), ctypes.c_void_p, ctypes.c_double, ctypes.c_void_p,
                                        ctypes.c_void_p, ctypes.c_void_p]),
    "pxh_dt_workspace_size": (ctypes.c_int64, [_P(Grid)]),
    "pxh_ssprk2_combine": (ctypes.c_int, [_P(Grid), _P(Fields), _P(Fields), _P(Fields), ctypes.c_void_p]),
    "pxh_roe_step": (ctypes.c_int, [_P(Grid), _P(Fields), _P(Fields), _P(BC), ctypes.c_void_p, ctypes.c_void_p]),
    "pxh_roe_dt": (ctypes.c_int, [_P(Grid), _P(Fields), ctypes.c_double, ctypes.c_void_p, ctypes.c_void_p,
                                  ctypes.c_void_p]),
    "pxh_bc_apply": (ctypes.c_int, [_P(Grid), _P(Fields), _P(BC), ctypes.c_void_p]),
    "pxh_roe_exner_ssprk2": (ctypes.c_int, [_P(Grid), _P(Fields), _P(Fields), _P(Fields), ctypes.c_void_p,
                                            _P(Sediments), _P(BC), ctypes.c_double, ctypes.c_void_p,
                                            ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]),
    "pxh_count_nonfinite": (ctypes.c_int, [_P(Grid), _P(Fields), ctypes.c_void_p, ctypes.c_void_p]),
    "pxh_roe_exner_ssprk2_noalloc": (ctypes.c_int, [_P(Grid), _P(Fields), _P(Fields), _P(Fields), ctypes.c_void_p,
                                                    _P(Sediments), _P(BC), ctypes.c_double, ctypes.c_void_p,
                                                    ctypes.c_int, ctypes.c_void_p]),
    "pxh_roe_exner_hydro_sweep": (ctypes.c_int, [_P(Grid), _P(Fields), _P(Fields), ctypes.c_void_p,
                                                 ctypes.c_void_p, ctypes.c_void_p]),
    "pxh_roe_exner_bed_sweep": (ctypes.c_int, [_P(Grid), _P(Fields), _P(Fields), _P(Fields), _P(Fields),
                                               ctypes.c_void_p, _P(Sediments), ctypes.c_void_p, ctypes.c_void_p,
                                               ctypes.c_void_p]),
    "pxh_dt_finalize": (ctypes.c_int, [_P(Grid), ctypes.c_void_p, ctypes.c_double, ctypes.c_void_p,
                                       ctypes.c_void_p]),
    "pxh_roe_exner_stage_sweep": (ctypes.c_int, [_P(Grid), _P(Fields), _P(Fields), _P(Fields), ctypes.c_void_p,
                                                 _P(Sediments), ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p,
                                                 ctypes.c_void_p]),
    "pxh_roe_exner_frame_sweep": (ctypes.c_int, [_P(Grid), _P(Fields), _P(Fields), _P(Fields), _P(Fields),
                                                 ctypes.c_void_p, _P(Sediments), ctypes.c_void_p, ctypes.c_int,
                                                 ctypes.c_double, ctypes.c_void_p, ctypes.c_void_p]),
    "pxh_halo_bytes": (ctypes.c_int64, [_P(Grid), _P(Fields), ctypes.c_int, ctypes.c_uint32, ctypes.c_int]),
    "pxh_halo_pack": (ctypes.c_int, [_P(Grid), _P(Fields), ctypes.c_int, ctypes.c_uint32, ctypes.c_int,
                                     ctypes.c_void_p, ctypes.c_void_p]),
    "pxh_halo_unpack": (ctypes.c_int, [_P(Grid), _P(Fields), ctypes.c_int, ctypes.c_uint32, ctypes.c_int,
                                       ctypes.c_void_p, ctypes.c_void_p]),
    "pxh_multi_stage_blocks": (ctypes.c_int64, [_P(Grid)]),
    "pxh_multi_frame_blocks": (ctypes.c_int64, [ctypes.c_int32, ctypes.c_int32]),
    "pxh_multi_seam_elems": (ctypes.c_int64, [ctypes.c_int32, ctypes.c_int32]),
    "pxh_multi_dt_frame_blocks": (ctypes.c_int64, [_P(Grid)]),
    "pxh_roe_exner_stage_sweep_multi": (ctypes.c_int, [ctypes.c_int32, ctypes.c_int32, ctypes.c_void_p, ctypes.c_int32,
                                                       ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, _P(Sediments),
                                                       ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p,
                                                       ctypes.c_void_p, ctypes.c_void_p]),
    "pxh_dt_frame_multi": (ctypes.c_int, [ctypes.c_int32, ctypes.c_int32, ctypes.c_void_p, ctypes.c_int32,
                                          ctypes.c_int32, ctypes.c_int32, ctypes.c_double, ctypes.c_void_p,
                                          ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    "pxh_roe_exner_frame_sweep_multi": (ctypes.c_int, [ctypes.c_int32, ctypes.c_int32, ctypes.c_void_p, ctypes.c_int32,
                                                       ctypes.c_int32, ctypes.c_int32, _P(Sediments), ctypes.c_void_p,
                                                       ctypes.c_int, ctypes.c_void_p]),
    "pxh_dt_multi_blocks": (ctypes.c_int64, [ctypes.c_int32, ctypes.c_int32]),
    "pxh_dt_multi": (ctypes.c_int, [ctypes.c_int32, ctypes.c_int32, ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32,
                                    ctypes.c_double, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    "pxh_copy_batch": (ctypes.c_int, [ctypes.c_int32, ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64,
                                      ctypes.c_void_p]),
    "pxh_ensemble_plan": (ctypes.c_int64, [ctypes.c_int32, ctypes.c_int32, _P(Member), ctypes.c_int32,
                                           ctypes.c_void_p, ctypes.c_int64]),
    "pxh_ensemble_workspace_size": (ctypes.c_int64, [ctypes.c_void_p]),
    "pxh_ensemble_dt": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, _P(Fields), ctypes.c_void_p, ctypes.c_double,
                                       ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    "pxh_ensemble_ssprk2": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, _P(Fields), _P(Fields), _P(Fields),
                                           ctypes.c_void_p, _P(BC), ctypes.c_double, ctypes.c_void_p,
                                           ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]),
    "pxh_roe_solve_2D": (ctypes.c_int, [_P(Grid), _P(Fields), _P(Fields), ctypes.c_void_p, ctypes.c_void_p,
                                        ctypes.c_void_p]),
    "pxh_momentum_corrections": (ctypes.c_int, [_P(Grid), _P(Fields), _P(Fields), ctypes.c_void_p, ctypes.c_void_p]),
    "pxh_compute_G": (ctypes.c_int, [_P(Grid), _P(Fields), _P(Fields), ctypes.c_void_p, _P(Sediments),
                                     ctypes.c_void_p]),
    "pxh_exner_solve_2D": (ctypes.c_int, [_P(Grid), _P(Fields), _P(Fields), ctypes.c_void_p, ctypes.c_void_p,
                                          ctypes.c_void_p]),
    "pxh_adj_workspace_size": (ctypes.c_int64, [_P(Grid), _P(BC)]),
    "pxh_adj_roe_exner_stage": (ctypes.c_int, [_P(Grid), _P(Fields), _P(Fields), ctypes.c_void_p, _P(Sediments),
                                               _P(BC), ctypes.c_void_p, _P(Fields), _P(Fields), ctypes.c_void_p,
                                               ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    "pxh_adj_roe_exner_dt": (ctypes.c_int, [_P(Grid), _P(Fields), ctypes.c_void_p, ctypes.c_double, ctypes.c_void_p,
                                            _P(Fields), ctypes.c_void_p, ctypes.c_void_p]),
    "pxh_adj_roe_exner_ssprk2": (ctypes.c_int, [_P(Grid), _P(Fields), _P(Fields), _P(Fields), ctypes.c_void_p,
                                                _P(Sediments), _P(BC), ctypes.c_double, ctypes.c_void_p, _P(Fields),
                                                ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    "pxh_adj_roe_exner_init": (ctypes.c_int, [_P(Grid), _P(Fields), ctypes.c_void_p, _P(Sediments), _P(BC),
                                              _P(Fields), ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
}

_lib = None


def lib_path():
    return _LIB_PATH


def load():
    """Load the HIP library (once).  Raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB_PATH):
        raise ImportError(f"{_LIB_PATH} is missing: build it with pyexner_amd.build.build() "
                          "(hipcc --offload-arch=gfx950); there is no CPU fallback")
    lib = ctypes.CDLL(_LIB_PATH)
    for name, (res, args) in _SIGS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


class HipError(RuntimeError):
    pass


def check(status, what):
    if status != 0:
        msg = load().pxh_error_string(status).decode()
        raise HipError(f"{what} failed with status {status}: {msg}")


def call(name, *args):
    check(getattr(load(), name)(*args), name)

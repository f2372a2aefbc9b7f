/*
 * pyexner_hip.h - C-ABI of the MI355X (gfx950) backend for PyExner's per-step
 * finite-volume hot path.
 *
 * Every entry point is extern "C", takes plain pointers and sizes, never
 * allocates device memory, never synchronises the host, and returns an int
 * status (0 = success; >0 = hipError_t of the failing launch; <0 = PXH_E_*).
 * All work is ordered on the stream passed in (a hipStream_t passed as void*;
 * NULL = the default stream), so a sequence of calls can be captured into a
 * hipGraph.  Arrays are (ny, nx) row-major and contiguous, x along the
 * fastest axis, +y towards decreasing row index.  Field data is float
 * (PXH_F32) or double (PXH_F64) as grid->dtype says; masks are uint8.
 *
 * Reference interfaces replaced (paths relative to src/PyExner/):
 *   step_fn   "Roe Exner"  solvers/roe_exner_solver.py:80-116  -> pxh_roe_exner_step
 *   compute_dt_fn "Roe Exner" solvers/roe_exner_solver.py:118-122 -> pxh_roe_exner_dt
 *   init_fn   "Roe Exner"  solvers/roe_exner_solver.py:58-78   -> pxh_roe_exner_init (+ pxh_bc_apply)
 *   step_fn   "Roe"        solvers/roe_solver.py:40-54         -> pxh_roe_step
 *   compute_dt_fn "Roe"    solvers/roe_solver.py:56-60         -> pxh_roe_dt
 *   BoundaryManager.apply  domain/boundary_registry.py:183-187 -> pxh_bc_apply
 *   SSPRK2 combine          integrators/SSPRK2.py:40            -> pxh_ssprk2_combine
 *   make_halo_exchange      solvers/kernels/roe_exner.py:719-803 -> pxh_halo_pack / pxh_halo_unpack
 *                                                                   (pxh_copy_batch: tiles on one device)
 *   (no counterpart: reverse mode of step_fn / compute_dt_fn / init_fn) -> pxh_adj_roe_exner_*
 */
#ifndef PYEXNER_HIP_H
#define PYEXNER_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PXH_F32 0
#define PXH_F64 1

#define PXH_OK 0
#define PXH_E_ARG (-1)      /* invalid argument (null pointer, bad size, bad dtype) */
#define PXH_E_RANGE (-2)    /* size beyond a compiled-in limit */
#define PXH_E_UNSUPPORTED (-3)

#define PXH_MAX_SED 8
#define PXH_MAX_HANDLERS 16
#define PXH_MAX_CHAIN 6

/* Grid geometry. */
#define PXH_EXACT_F32 1  /* flags bit: float32 with IEEE division/sqrt/libm (default: hardware rcp/rsq/log2/exp2) */
/* flags bit: the caller promises that mask0 is zero in every cell (no nodata
 * cell; the ring covered by boundary polygons), so the fused stage sweeps do
 * not load it.  Results are those of the same call with the mask loaded.
 * In the flags of pxh_ensemble_plan and pxh_roe_exner_stage_sweep_multi the
 * promise holds for every member / domain. */
#define PXH_NO_NODATA 2
/* flags bit: the caller promises that the fixed bed z holds one value (z[0])
 * in every cell, so the fused stage sweeps read it once instead of streaming
 * z (honoured together with PXH_NO_NODATA, on the float64 and the float32
 * hardware-math paths).  Results are those of the same call without it.
 * For an ensemble or a multi-domain sweep: each member's / domain's z holds
 * one value, its own first cell's. */
#define PXH_Z_UNIFORM 4
/* flags bit, honoured by pxh_roe_exner_ssprk2 (and _noalloc) on the float32
 * hardware-math path with a frame-local ring fill: b->h holds 4 * ny * nx
 * elements, and the first stage's output keeps h, hu, hv, z_b there as one
 * record of four per cell (b->hu, hv, z_b are not read or written: b is
 * scratch), which the second stage reads as one stream instead of four.
 * Results are those of the same call without it; other entries and paths
 * ignore it (b->h's first ny * nx elements are then b's h).  In an
 * ensemble plan's flags it asks the same of pxh_ensemble_ssprk2: b->h holds
 * 4 elements per arena cell, the records of every member.  The separate
 * stage / frame sweep entries take the same records as a view
 * (pxh_roe_exner_stage_sweep). */
#define PXH_PACKED_STAGE 8
/* flags bits 8-15: owned rows per strip of the fused stage sweep (0: 32;
 * otherwise 8..128), the launch shape pyexner_amd.autotune chooses per grid. */
#define PXH_STRIP_ROWS_SHIFT 8
#define PXH_STRIP_ROWS(flags) (((flags) >> PXH_STRIP_ROWS_SHIFT) & 0xff)
/* flags bits 16-19 / 20-23: the most wavefronts per SIMD of the first /
 * second SSPRK2 stage's fused sweep (the Forward-Euler stage counts as a
 * first stage), 0: no cap (registers set the occupancy), else 1..8.  Another
 * launch shape pyexner_amd.autotune chooses per grid; results do not depend
 * on it. */
#define PXH_STAGE1_WAVES_SHIFT 16
#define PXH_STAGE2_WAVES_SHIFT 20
#define PXH_STAGE1_WAVES(flags) (((flags) >> PXH_STAGE1_WAVES_SHIFT) & 0xf)
#define PXH_STAGE2_WAVES(flags) (((flags) >> PXH_STAGE2_WAVES_SHIFT) & 0xf)

typedef struct {
    int32_t ny, nx;   /* array shape, ghost ring included */
    int32_t dtype;    /* PXH_F32 or PXH_F64 */
    int32_t flags;    /* PXH_EXACT_F32, PXH_NO_NODATA, PXH_Z_UNIFORM, PXH_PACKED_STAGE, PXH_STRIP_ROWS,
                         PXH_STAGE*_WAVES */
    double dx;        /* cell size (dx == dy) */
} pxh_grid;

/* Device pointers to the structure-of-arrays state (state/roe_exner_state.py:8-48;
 * the plain Roe state uses h, hu, hv, z, n only).  Each is (ny, nx).
 * z_b_lo (float32 only, may be NULL): the low part of a compensated bed,
 * z_b + z_b_lo, |z_b_lo| <= ulp(z_b)/2; the Exner update and the SSPRK2
 * combine then accumulate the bed without float32 round-off drift while every
 * other kernel reads z_b. */
typedef struct {
    void *h, *hu, *hv, *z, *z_b, *n, *n_b, *G;
    void *z_b_lo;
} pxh_fields;

/* Sediment table, precomputed on the host in the working dtype from
 * seds = [frac, d50, density, k_d, k_e, theta_c, porosity] (io/pnetcdf_reader.py:146-158):
 *   A = np^2, B = (s-1)*d50, C = np^3*sqrt(g), D = s-1, with s = density/1000,
 *   np = d50^(1/6)/21.1 (solvers/kernels/roe_exner.py:95-116);
 *   one_minus_por = 1 - mean(porosity); nb_sed = mean(d50*frac)^(1/6)/21.1 (roe_exner.py:64-73). */
typedef struct {
    int32_t n_sed;
    int32_t mpm;              /* 1: Meyer-Peter-Mueller G; 0: constant G (grass_factor number) */
    double one_minus_por;
    double nb_sed;
    double frac[PXH_MAX_SED], A[PXH_MAX_SED], B[PXH_MAX_SED], C[PXH_MAX_SED],
           D[PXH_MAX_SED], theta_c[PXH_MAX_SED];
    double AB[PXH_MAX_SED], CD[PXH_MAX_SED];  /* A/B and C/D for the hardware-math float32 path */
} pxh_sediments;

/* Composed ghost-cell program (domain/boundary_registry.py:183-187 applied in
 * YAML order, resolved on the host into one gather/scatter per target cell). */
#define PXH_BC_COPY 0          /* Transmissive, TransmissiveBed, Periodic */
#define PXH_BC_REFLECT_RX 1    /* domain/boundaries/roeexner_reflective.py:31-59 */
#define PXH_BC_REFLECT_ROE 2   /* domain/boundaries/roe_reflective.py:34-39 */

#define PXH_G_H 0
#define PXH_G_HUV 1
#define PXH_G_Z 2
#define PXH_G_ZB 3
#define PXH_G_N 4
#define PXH_G_NB 5
#define PXH_G_G 6

typedef struct {
    int32_t target, src;            /* linear cell indices (row*nx + col) */
    uint8_t group;                  /* PXH_G_* */
    uint8_t nchain;                 /* handlers applied to src, in order (momentum only); with
                                       PXH_BC_VALUE set, a value entry: chain[0] indexes pxh_bc.values
                                       (an index >= n_values makes the entry a no-op) */
    uint8_t chain[PXH_MAX_CHAIN];
} pxh_bc_entry;                     /* 16 bytes */

/* Value entries: the target of a value boundary (domain/boundaries/roeexner_*.py
 * beyond the copy handlers) takes a value computed from the spec rather than
 * a copy of its source.  Kinds (a, b, c as in pxh_bc_value; t = pxh_bc.time):
 *   PXH_BV_CONST    field = a; for PXH_G_HUV hu = a, hv = b (a NaN component
 *                   keeps the cell's value: ZeroMomentum with a NaN value)
 *   PXH_BV_AFFINE_T field = b * t + a (Berthon / NormalFlowDepth bed, zb0 - alpha t)
 *   PXH_BV_OFFSET   field = field[src] + a (SteepFall: z_b of the interior - 1)
 *   PXH_G_H only, PXH_BV_OUTFLUX: h from the outgoing Riemann invariant of the
 *                   source cell at the prescribed discharge q = a through the
 *                   normal (b, c), five Newton steps (roeexner_constantoutflux.py:24-56) */
#define PXH_BC_VALUE 0x80
#define PXH_BV_CONST 1
#define PXH_BV_AFFINE_T 2
#define PXH_BV_OFFSET 3
#define PXH_BV_OUTFLUX 4
#define PXH_MAX_VALUES 32
#define PXH_MAX_LEVELS 16
typedef struct {
    int32_t kind;                   /* PXH_BV_* */
    int32_t reserved;
    double a, b, c;
} pxh_bc_value;

typedef struct {
    int32_t kind;                   /* PXH_BC_* */
    int32_t reserved;
    double nx, ny;                  /* outward normal as given in the YAML */
} pxh_bc_handler;

/* pxh_bc.flags: PXH_BC_FRAME promises that every entry's target lies on the
 * outer ring of cells (row 0 or ny-1, column 0 or nx-1) and its source within
 * two cells of the grid edge.  The step entries then fuse the hydro and bed
 * updates of each stage and update the two-cell frame after the fill; without
 * the flag they run the two sweeps with the fill between them. */
#define PXH_BC_FRAME 1
/* pxh_bc.flags: PXH_BC_GROUPED promises that the entries are ordered by group
 * and that group g occupies entries [group_start[g], group_start[g+1]); a call
 * then visits only the entries of the groups in group_mask.  Without it every
 * call visits every entry. */
#define PXH_BC_GROUPED 2
/* pxh_bc.flags: PXH_BC_SRC_SORTED promises that within each group the entries
 * are ordered by src, so that the entries filled from one source are adjacent.
 * The adjoint fill (pxh_adj_*) then sums their adjoints in a fixed order and
 * its results do not depend on scheduling; without the flag it accumulates
 * with atomics (the same sums up to round-off). */
#define PXH_BC_SRC_SORTED 4
#define PXH_N_GROUPS 7

typedef struct {
    const pxh_bc_entry *entries;    /* device array, n_entries long */
    int32_t n_entries;
    int32_t n_handlers;
    uint32_t group_mask;            /* which PXH_G_* groups this call writes (bit i = group i) */
    int32_t flags;                  /* PXH_BC_* flags */
    int32_t group_start[PXH_N_GROUPS + 1];  /* with PXH_BC_GROUPED; group_start[7] = n_entries */
    pxh_bc_handler handlers[PXH_MAX_HANDLERS];
    double time;                    /* simulation time of the fill (PXH_BV_AFFINE_T values) */
    int32_t n_values;
    int32_t reserved;
    pxh_bc_value values[PXH_MAX_VALUES];
    /* Dependency levels: with n_levels > 1 the entries are ordered by level and
     * level L is entries [level_start[L], level_start[L+1]), applied one level
     * after another (level_start[n_levels] = n_entries; group_start and
     * PXH_BC_GROUPED are then not used).  Every entry reads its source as it
     * was before the whole program, as when the handlers of
     * domain/boundary_registry.py:183-187 run one after another and each reads
     * its sources before writing: an entry that reads a cell which another
     * entry writes sits in an earlier level than that writer.  n_levels 0 or
     * 1: one level (no entry reads a cell the program writes). */
    int32_t n_levels;
    int32_t level_start[PXH_MAX_LEVELS + 1];
} pxh_bc;

/* ---- library ------------------------------------------------------------ */
int pxh_version(void);
const char *pxh_error_string(int status);

/* ---- Roe-Exner ----------------------------------------------------------- */
/* init_fn minus the boundary fill: n_b = compute_n (roe_exner.py:64-73) unless
 * constant mode, then G = compute_G (roe_exner.py:85-144). In place on s. */
int pxh_roe_exner_init(const pxh_grid *grid, const pxh_fields *s, const uint8_t *mask0,
                       const pxh_sediments *sed, void *stream);

/* One Forward-Euler step_fn: out = S(in, dt).  out->h, hu, hv, z_b, G must not
 * alias in's; z, n, n_b are static and may be shared.  dt is read from the
 * device scalar dt_dev (dtype of the grid).  bc must hold a program of one
 * dependency level (sources never targets) covering groups h, hu/hv, z_b and
 * G: the step fills (h, hu, hv) after the hydro update and (z_b, G) after the
 * bed update, which equals the reference's two fills of every group only for
 * such programs.  A program of several levels returns PXH_E_UNSUPPORTED
 * (likewise pxh_roe_exner_ssprk2, pxh_roe_step and the ensemble); its step is
 * the single-kernel entries below sequenced as roe_exner_solver.py:80-116
 * does, with pxh_bc_apply of every group at both fills. */
int pxh_roe_exner_step(const pxh_grid *grid, const pxh_fields *in, const pxh_fields *out,
                       const uint8_t *mask0, const pxh_sediments *sed, const pxh_bc *bc,
                       const void *dt_dev, void *stream);

/* The four kernels step_fn sequences, one call each, for callers that
 * sequence them themselves (e.g. with halo exchanges in between):
 *   roe_solve_2D        solvers/kernels/roe_exner.py:535-612  out.{h,hu,hv} from in
 *   momentum_corrections roe_exner.py:14-62                   in place on out.{hu,hv}; bed from in
 *   compute_G           roe_exner.py:85-144                   out.G from out.{h,hu,hv} (constant mode: in.G)
 *   exner_solve_2D      roe_exner.py:665-717                  out.z_b from in.z_b and out.{h,hu,hv,G} */
int pxh_roe_solve_2D(const pxh_grid *grid, const pxh_fields *in, const pxh_fields *out, const uint8_t *mask0,
                     const void *dt_dev, void *stream);
int pxh_momentum_corrections(const pxh_grid *grid, const pxh_fields *in, const pxh_fields *out,
                             const uint8_t *mask0, void *stream);
int pxh_compute_G(const pxh_grid *grid, const pxh_fields *in, const pxh_fields *out, const uint8_t *mask0,
                  const pxh_sediments *sed, void *stream);
int pxh_exner_solve_2D(const pxh_grid *grid, const pxh_fields *in, const pxh_fields *out, const uint8_t *mask0,
                       const void *dt_dev, void *stream);

/* One whole SSPRK2 step on three caller-owned buffer sets (integrators/SSPRK2.py:32-40):
 * dt = compute_dt_fn(a) into dt_dev (skipped when dt_ready != 0: dt_dev already
 * holds it); b = step_fn(a, dt); c = step_fn(b, dt); a = 0.5 a + 0.5 c.  On
 * return dt_dev holds compute_dt_fn(a) of the new a (reduced inside the last
 * sweep), so consecutive calls pass dt_ready = 1.  a is updated in place; b
 * and c are scratch whose z, n, n_b may share a's (static fields); the evolving
 * fields of a, b, c must not alias.  workspace as for pxh_roe_exner_dt.
 * The ring is averaged like every other cell (SSPRK2.py:40): with bed values
 * that are not fixed points of the average (PXH_BV_AFFINE_T, PXH_BV_OFFSET)
 * the call runs two whole steps with their full fills, the combine and a dt
 * pass instead of the fused sweeps (slower, same result as the reference). */
int pxh_roe_exner_ssprk2(const pxh_grid *grid, const pxh_fields *a, const pxh_fields *b, const pxh_fields *c,
                         const uint8_t *mask0, const pxh_sediments *sed, const pxh_bc *bc, double cfl,
                         void *dt_dev, void *workspace, int dt_ready, void *stream);

/* pxh_roe_exner_ssprk2 without the workspace argument, for callers that cannot
 * allocate device memory (code under hipGraph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  No counterpart in the reference.  The dt
 * workspace lives in a static device array of PXH_NOALLOC_SCRATCH_BYTES per
 * GPU, held until the process ends; grids whose pxh_dt_workspace_size exceeds
 * it are refused with PXH_E_RANGE (float32 up to about 8192 x 8192 at any
 * strip height).  The call allocates nothing; the caller still passes a, b, c
 * and dt_dev as for pxh_roe_exner_ssprk2, whose results it gives bit for bit.
 * Every call on one GPU shares that GPU's array: calls that overlap on one GPU
 * (concurrent streams, threads) race and must be serialised by the caller. */
#define PXH_NOALLOC_SCRATCH_BYTES (64LL << 20)
int pxh_roe_exner_ssprk2_noalloc(const pxh_grid *grid, const pxh_fields *a, const pxh_fields *b,
                                 const pxh_fields *c, const uint8_t *mask0, const pxh_sediments *sed,
                                 const pxh_bc *bc, double cfl, void *dt_dev, int dt_ready, void *stream);

/* Failure detection, no counterpart in the reference (its only guard is the dt
 * clamp, solvers/roe_exner_solver.py:121): the number of cells of the grid
 * whose h, hu, hv, z_b, z_b_lo or G (null pointers skipped) is NaN or
 * infinite, written to *count_dev (device memory, one uint64_t) in stream
 * order.  Two launches (a zeroing kernel, then the scan); the integrators'
 * output-window loops call it once per window and read the count back with
 * the window's output. */
int pxh_count_nonfinite(const pxh_grid *grid, const pxh_fields *s, void *count_dev, void *stream);

/* The two fused sweeps a step is built from, one launch each, for callers that
 * sequence the step themselves (ring fills or halo exchanges in between):
 *   hydro sweep: roe_solve_2D + momentum_corrections (roe_exner.py:535-612, 14-62):
 *                out.{h,hu,hv} from in.
 *   bed sweep:   compute_G + exner_solve_2D (roe_exner.py:85-144, 665-717) on the
 *                filled hyd.{h,hu,hv}, bed from bed.{z_b,z_b_lo,G}:
 *                old == NULL: out.{z_b,z_b_lo,G};
 *                old != NULL: out = 0.5 old + 0.5 result over h,hu,hv,z_b,G
 *                (integrators/SSPRK2.py:40), and when dt_partial != NULL the
 *                per-block CFL minima of the combined state, which
 *                pxh_dt_finalize turns into compute_dt_fn's value. */
int pxh_roe_exner_hydro_sweep(const pxh_grid *grid, const pxh_fields *in, const pxh_fields *out,
                              const uint8_t *mask0, const void *dt_dev, void *stream);
int pxh_roe_exner_bed_sweep(const pxh_grid *grid, const pxh_fields *hyd, const pxh_fields *bed,
                            const pxh_fields *out, const pxh_fields *old, const uint8_t *mask0,
                            const pxh_sediments *sed, const void *dt_dev, void *dt_partial, void *stream);
int pxh_dt_finalize(const pxh_grid *grid, const void *dt_partial, double cfl, void *dt_out, void *stream);

/* The fused stage that pxh_roe_exner_step and pxh_roe_exner_ssprk2 run when
 * bc->flags has PXH_BC_FRAME, for callers that sequence the step themselves.
 * Valid only when the caller's ring fill is frame-local (PXH_BC_FRAME): a
 * stage (roe_exner_solver.py:80-116) is then
 *   stage sweep (one launch) -> fill (h, hu, hv) of the stage result -> frame sweep
 * and the caller fills (z_b, G) after the frame sweep.  The "frame" is the
 * cells within two cells of the grid edge.
 *   stage sweep, old == NULL: roe_solve_2D + momentum_corrections of in into
 *     out.{h,hu,hv} (every cell), compute_G + exner_solve_2D into
 *     out.{z_b,z_b_lo,G} of the cells off the frame.
 *   stage sweep, old != NULL (second SSPRK2 stage, integrators/SSPRK2.py:38-40):
 *     out.{h,hu,hv} of the cells within three cells of the edge (what the fill
 *     and the frame sweep read), old = 0.5 old + 0.5 S(in) over h, hu, hv,
 *     z_b, z_b_lo, G off the frame; with dt_partial, the per-block CFL minima of
 *     that combined state over the edges inside each strip.
 *   frame sweep: compute_G + exner_solve_2D of the frame's cells from the
 *     filled hyd.{h,hu,hv} and bed.{z_b,z_b_lo,G} (the stage input):
 *     old == NULL: into out.{z_b,z_b_lo,G}; old != NULL: combined into old as
 *     above, and with dt_partial the minima of the frame's edges and of the
 *     edges across strip seams, and the final reduction: dt_dev =
 *     compute_dt_fn(old), so it holds the stage's dt on entry and the next
 *     step's on return.
 *   Pairing: the final reduction folds in the strip minima and seam records
 *     that the stage sweep left in dt_partial.  A frame sweep given dt_partial
 *     is valid only after a second-stage sweep (old != NULL) that was given
 *     the same dt_partial, the same grid (strip height included) and the same
 *     old, with nothing else writing dt_partial in between; otherwise it
 *     reduces stale minima.  Pass dt_partial to both calls or to neither.
 * flags: PXH_SWEEP_SKIP_G leaves out.G unwritten in MPM mode (the next stage
 * recomputes G; the first SSPRK2 stage uses it).  dt_partial holds
 * pxh_dt_workspace_size(grid) bytes.
 * Records (float32 without PXH_EXACT_F32): a pxh_fields whose hu, hv, z_b
 * point one, two and three elements past h (h's buffer holding 4 ny nx
 * elements) views the four-float records of PXH_PACKED_STAGE, and these
 * entries accept one where pxh_roe_exner_ssprk2 keeps them: the first
 * stage's out (stage sweep, then frame sweep with hyd = out) and the second
 * stage's in (stage sweep) and bed (frame sweep); pxh_bc_apply fills such a
 * view too.  Every other entry refuses one (PXH_E_ARG). */
#define PXH_SWEEP_SKIP_G 1
int pxh_roe_exner_stage_sweep(const pxh_grid *grid, const pxh_fields *in, const pxh_fields *out,
                              const pxh_fields *old, const uint8_t *mask0, const pxh_sediments *sed,
                              const void *dt_dev, int flags, void *dt_partial, void *stream);
int pxh_roe_exner_frame_sweep(const pxh_grid *grid, const pxh_fields *hyd, const pxh_fields *bed,
                              const pxh_fields *out, const pxh_fields *old, const uint8_t *mask0,
                              const pxh_sediments *sed, void *dt_dev, int flags, double cfl, void *dt_partial,
                              void *stream);

/* compute_dt_fn: *dt_out = max(cfl * min over active edges, 1e-6) written to
 * device memory.  workspace must hold pxh_dt_workspace_size(grid) bytes. */
int pxh_roe_exner_dt(const pxh_grid *grid, const pxh_fields *s, const uint8_t *mask0,
                     double cfl, void *dt_out, void *workspace, void *stream);
int64_t pxh_dt_workspace_size(const pxh_grid *grid);

/* SSPRK2 combination out = 0.5*a + 0.5*b over h, hu, hv, z_b, G (the static
 * fields are equal in a and b).  out may alias a or b. */
int pxh_ssprk2_combine(const pxh_grid *grid, const pxh_fields *a, const pxh_fields *b,
                       const pxh_fields *out, void *stream);

/* ---- ensembles ------------------------------------------------------------ */
/* Many independent Roe-Exner domains (members of uneven shape, each with its
 * own sediment table and its own CFL dt) advanced together: each SSPRK2 step is
 * one launch per sweep for all members, not one step per member.  No reference
 * counterpart (the reference runs one domain per process); it is the batched
 * form of pxh_roe_exner_ssprk2 on frame-local rings.
 * Every field of every member lives in one arena per field (pxh_fields of
 * arena pointers), member m's (ny, nx) arrays at element offset off; the mask
 * arena likewise.  The ring-fill program addresses arena elements (targets and
 * sources offset by their member's off) and must be frame-local for every
 * member (PXH_BC_FRAME).  dt_dev holds one dt per member.
 * pxh_ensemble_plan fills, on the host, the plan the step reads (member
 * table and float32 sediment folds); called with plan == NULL or too small a
 * plan_bytes it returns the size.  The caller copies the plan to device memory
 * and passes both copies.  workspace: pxh_ensemble_workspace_size bytes. */
typedef struct {
    int64_t off;              /* first element of the member's arrays in the arenas */
    int32_t ny, nx;
    double dx;
    pxh_sediments sed;
} pxh_member;
int64_t pxh_ensemble_plan(int32_t dtype, int32_t flags, const pxh_member *members, int32_t n, void *plan,
                          int64_t plan_bytes);
int64_t pxh_ensemble_workspace_size(const void *plan);
/* compute_dt_fn of every member into dt_dev[m]. */
int pxh_ensemble_dt(const void *plan, const void *plan_dev, const pxh_fields *a, const uint8_t *mask0, double cfl,
                    void *dt_dev, void *workspace, void *stream);
/* One SSPRK2 step of every member on three arena sets, as pxh_roe_exner_ssprk2:
 * dt from a unless dt_ready, b = S(a), c = S(b), a = 0.5 a + 0.5 c, and dt_dev
 * left holding each member's dt of the new a. */
int pxh_ensemble_ssprk2(const void *plan, const void *plan_dev, const pxh_fields *a, const pxh_fields *b,
                        const pxh_fields *c, const uint8_t *mask0, const pxh_bc *bc, double cfl, void *dt_dev,
                        void *workspace, int dt_ready, void *stream);

/* ---- adjoint (double only) ------------------------------------------------ */
/* Reverse-mode derivatives of the Roe-Exner step: each entry is the transpose
 * of the forward entry it names, evaluated at the forward's inputs, for
 * gradients of a scalar loss of the final state.  The reference has no
 * gradient code; these are what jax.grad through its step_fn / compute_dt_fn /
 * init_fn (solvers/roe_exner_solver.py:58-122) would give, with the
 * derivative of every untaken branch taken as zero.  Adjoint fields ("lam") are (ny, nx) double arrays in
 * pxh_fields slots of the same names; MPM G and ring fills without
 * PXH_BV_OUTFLUX values only (PXH_E_UNSUPPORTED otherwise).
 * lam_dt: one device double; lam_par: five device doubles, the derivatives
 * with respect to the first sediment class's A, B, C (pxh_sediments),
 * one_minus_por and nb_sed, accumulated (may be NULL).  The ring-fill program
 * must not read a cell it writes.  workspace: pxh_adj_workspace_size bytes,
 * which depend on the grid only (largest_bc, may be NULL, is validated). */
int64_t pxh_adj_workspace_size(const pxh_grid *grid, const pxh_bc *largest_bc);
/* One stage (pxh_roe_exner_step in -> out at dt): lam_out (h, hu, hv, z_b, G)
 * is the adjoint of out and is consumed; the adjoint of in accumulates into
 * lam_in (h, hu, hv, z_b, and n_b when set); d/d dt into *lam_dt. */
int pxh_adj_roe_exner_stage(const pxh_grid *grid, const pxh_fields *in, const pxh_fields *out,
                            const uint8_t *mask0, const pxh_sediments *sed, const pxh_bc *bc, const void *dt_dev,
                            const pxh_fields *lam_out, const pxh_fields *lam_in, void *lam_dt, void *lam_par,
                            void *workspace, void *stream);
/* compute_dt_fn (pxh_roe_exner_dt): lam_s (h, hu, hv, G) += lam_dt * d dt / d s. */
int pxh_adj_roe_exner_dt(const pxh_grid *grid, const pxh_fields *s, const uint8_t *mask0, double cfl,
                         const void *lam_dt, const pxh_fields *lam_s, void *workspace, void *stream);
/* One SSPRK2 step (pxh_roe_exner_ssprk2 from a, with dt = compute_dt_fn(a)):
 * a the step's input state, b the stage-1 output (h, hu, hv, z_b), c the
 * stage-2 output's hydro (h, hu, hv).  lam holds the adjoint of the step's
 * output (h, hu, hv, z_b, G) on entry and of a on return; lam->n_b (may be
 * NULL) accumulates the adjoint of the bed Manning coefficient n_b. */
int pxh_adj_roe_exner_ssprk2(const pxh_grid *grid, const pxh_fields *a, const pxh_fields *b, const pxh_fields *c,
                             const uint8_t *mask0, const pxh_sediments *sed, const pxh_bc *bc, double cfl,
                             const void *dt_dev, const pxh_fields *lam, void *lam_par, void *workspace,
                             void *stream);
/* init_fn (pxh_roe_exner_init + pxh_bc_apply of the full program bc): s the
 * initialised state; lam the adjoint of s on entry (h, hu, hv, z_b, G, n_b)
 * and of the state before initialisation on return (h, hu, hv, z_b). */
int pxh_adj_roe_exner_init(const pxh_grid *grid, const pxh_fields *s, const uint8_t *mask0,
                           const pxh_sediments *sed, const pxh_bc *bc, const pxh_fields *lam, void *lam_par,
                           void *workspace, void *stream);

/* ---- plain Roe ----------------------------------------------------------- */
int pxh_roe_step(const pxh_grid *grid, const pxh_fields *in, const pxh_fields *out,
                 const pxh_bc *bc, const void *dt_dev, void *stream);
int pxh_roe_dt(const pxh_grid *grid, const pxh_fields *s, double cfl, void *dt_out,
               void *workspace, void *stream);

/* ---- boundaries / halos -------------------------------------------------- */
/* Apply the composed boundary program in place (groups bc->group_mask); s
 * may be a view of four-element records (pxh_roe_exner_stage_sweep). */
int pxh_bc_apply(const pxh_grid *grid, const pxh_fields *s, const pxh_bc *bc, void *stream);

/* Halo strips of a subdomain tile: make_halo_exchange (solvers/kernels/roe_exner.py:719-803),
 * with the MPI sendrecv left to the caller.  The tile is the grid: its owned
 * block plus hw ghost rings (hw = 1, 2 or 3).  pack copies the send strips (the hw
 * owned rows/columns next to each selected side) of the fields of `groups`
 * (bits 1 << PXH_G_*; PXH_G_ZB carries z_b_lo too when present) into buf;
 * unpack copies buf into the receive strips (the hw ghost rows/columns).  Depth
 * 0 is the row/column nearest the shared boundary on both tiles, so the buffer
 * a neighbour packed for its east side unpacks as it is into this tile's west
 * side, and so on.  West/east strips span every row, north/south strips every
 * column: exchange west/east first, unpack, then pack north/south, and the
 * corner blocks travel with the second phase.  One call handles one phase:
 * `sides` is a subset of west|east or of north|south.  Layout of buf: per field (PXH_G_*
 * order, z_b_lo after z_b), per selected side (west, north, east, south), hw
 * strips depth-major; pxh_halo_bytes gives its size. */
#define PXH_HALO_W 1
#define PXH_HALO_N 2
#define PXH_HALO_E 4
#define PXH_HALO_S 8
int64_t pxh_halo_bytes(const pxh_grid *grid, const pxh_fields *s, int hw, uint32_t groups, int sides);
int pxh_halo_pack(const pxh_grid *grid, const pxh_fields *s, int hw, uint32_t groups, int sides, void *buf,
                  void *stream);
int pxh_halo_unpack(const pxh_grid *grid, const pxh_fields *s, int hw, uint32_t groups, int sides,
                    const void *buf, void *stream);

/* Batched strip copies on one device: the exchange of the tiles of a
 * decomposition that share one GPU (make_halo_exchange, roe_exner.py:719-803,
 * has no such form: its ranks send packed strips).  Each ghost strip is copied
 * straight from the neighbour's owned strip, every strip of an exchange phase
 * in one launch, in place of a pack and an unpack per side of every tile.
 * jobs: device memory, n entries; job k copies rows x cols elements of dtype
 * from src (row stride src_ld elements) to dst (row stride dst_ld); start is
 * the sum of rows * cols over the jobs before it (job 0: 0) and total the sum
 * over all n.  No destination may overlap a source of the same call.  The
 * table lives on the device, so only n, total and dtype are checked here. */
/* The fused stage's two sweeps over several domains, one launch each: the
 * tiles of a decomposition that share one device (the reference runs one
 * tile per rank and has no such form).  Domain k is what
 * pxh_roe_exner_stage_sweep and pxh_roe_exner_frame_sweep take for one tile:
 * the stage input `in`, `out`, and for the second SSPRK2 stage (second = 1)
 * `old`; the frame sweep reads hyd = out and bed = in and writes out (first
 * stage) or the combine into old (second).  sblk0 / fblk0: the domain's first
 * stage / frame block, the running sums of pxh_multi_stage_blocks /
 * pxh_multi_frame_blocks over the domains before it; nblocks: the sum over
 * all n of the sweep launched.  lo = 1 when the fields carry z_b_lo (float32
 * compensated bed).  Every domain must be valid for the single-domain entry
 * with grid (ny, nx, dtype, flags, dx): the table lives on the device, so
 * only the scalars are checked here (flags, sed, dt_dev and
 * PXH_SWEEP_SKIP_G in sweep_flags as in the single-domain entries).
 * The next dt, reduced inside the second stage as the single-domain fused
 * step does: a second-stage sweep given dt_partial and xs writes one CFL
 * minimum per block into dt_partial[0, nblocks) (over the edges between the
 * cells off each domain's two-cell frame) and each domain's seam records at
 * xs + xs0 (elements of the dtype; xs0: the running sum of
 * pxh_multi_seam_elems over the domains before it); then, after the frame
 * sweep, pxh_dt_frame_multi adds each domain's frame edges and strip-seam
 * edges and reduces everything into *dt_out = max(cfl * min, 1e-6).
 * `shared` (bit PXH_HALO_N/S/W/E): the domain's sides that are ghost rings
 * of a neighbouring tile; the frame cells there hold no valid combined
 * state, and edges touching them are left to the neighbour, whose owned
 * cells they are.  tblk0: the domain's first block of pxh_dt_frame_multi,
 * the running sum of pxh_multi_dt_frame_blocks.  With three ghost rings per
 * shared side (the fused stage's halo) every edge of the decomposed grid is
 * an edge between valid cells of some domain, so the result is the dt of
 * the whole grid. */
typedef struct {
    pxh_fields in, out, old;
    const uint8_t *mask0;
    int32_t ny, nx;
    int32_t sblk0, fblk0;
    double dx;
    int64_t xs0;
    int32_t tblk0;
    uint32_t shared;
} pxh_sweep_domain;
int64_t pxh_multi_stage_blocks(const pxh_grid *grid);
int64_t pxh_multi_frame_blocks(int32_t ny, int32_t nx);
int64_t pxh_multi_seam_elems(int32_t ny, int32_t nx);
int64_t pxh_multi_dt_frame_blocks(const pxh_grid *grid);
int pxh_roe_exner_stage_sweep_multi(int32_t dtype, int32_t flags, const pxh_sweep_domain *doms, int32_t n,
                                    int32_t nblocks, int32_t second, int32_t lo, const pxh_sediments *sed,
                                    const void *dt_dev, int sweep_flags, void *dt_partial, void *xs,
                                    void *stream);
/* nblocks: the sum of pxh_multi_dt_frame_blocks over the n domains; nstage:
 * the second-stage sweep's block count, whose minima lie in
 * dt_partial[0, nstage); this call writes its own partials after them
 * (dt_partial holds nstage + nblocks elements of the dtype) and the dt. */
int pxh_dt_frame_multi(int32_t dtype, int32_t flags, const pxh_sweep_domain *doms, int32_t n, int32_t nblocks,
                       int32_t nstage, double cfl, void *dt_partial, const void *xs, void *dt_out, void *stream);
int pxh_roe_exner_frame_sweep_multi(int32_t dtype, int32_t flags, const pxh_sweep_domain *doms, int32_t n,
                                    int32_t nblocks, int32_t second, const pxh_sediments *sed, const void *dt_dev,
                                    int sweep_flags, void *stream);

/* compute_dt_fn over several domains in one sweep launch and one reduction:
 * the tiles of a decomposition on one device (compute_dt_roeexner,
 * roe_exner_solver.py:118-122, whose MIN allreduce over ranks becomes the
 * minimum over the domains).  *dt_out = max(cfl * min over the active edges of
 * every domain, 1e-6).  doms: device memory, n entries; domain k holds the
 * fields h, hu, hv, G and the nodata mask of a ny x nx grid with spacing dx,
 * and its sweep blocks start at blk0 = the sum of pxh_dt_multi_blocks over the
 * domains before it; nblocks is the sum over all n, and partial holds nblocks
 * * 8 bytes.  flags: PXH_EXACT_F32 as in pxh_grid.flags. */
typedef struct {
    const void *h, *hu, *hv, *G;
    const uint8_t *mask0;
    int32_t ny, nx;
    int32_t blk0;
    int32_t reserved;
    double dx;
} pxh_dt_domain;
int64_t pxh_dt_multi_blocks(int32_t ny, int32_t nx);
int pxh_dt_multi(int32_t dtype, int32_t flags, const pxh_dt_domain *doms, int32_t n, int32_t nblocks, double cfl,
                 void *dt_out, void *partial, void *stream);

typedef struct {
    const void *src;
    void *dst;
    int64_t src_ld, dst_ld;
    int64_t start;
    int32_t rows, cols;
} pxh_copy_job;
int pxh_copy_batch(int32_t dtype, const pxh_copy_job *jobs, int32_t n, int64_t total, void *stream);

#ifdef __cplusplus
}
#endif
#endif

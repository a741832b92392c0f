/*
 * fluid_projection.h -- C ABI of the MI355X MAC-grid pressure projection.
 *
 * This is the drop-in boundary a host binds (N-API addon, ctypes, C++): plain
 * pointers and sizes, C linkage, no exceptions, no torch types.  It replaces,
 * for one GPU, the projection stage that the reference encodes in WebGPU:
 *
 *   fp_grid_create   <- new WebGPUUniformPressureMultigrid(device, dims, spacing)
 *                       lib/methods/uniform/webgpu-uniform-pressure-multigrid.ts:228-309
 *                       (throws RangeError on an unplannable lattice, :251-256)
 *                       and the restricted tall-cell layout of lib/core/tall-cell-grid.ts:309-436
 *   fp_project       <- WebGPUUniformPressureMultigrid.encode (…multigrid.ts:327-364)
 *                       followed by the `project` kernel
 *                       (lib/methods/uniform/webgpu-uniform-reference.ts:1280-1282,
 *                        webgpu-uniform-reference.wgsl.ts:1052-1099)
 *   fp_report        <- the 60-byte multigrid diagnostics buffer decoded by readStats
 *                       (webgpu-uniform-reference.ts:1358-1421) and the MGPCG
 *                       error bits of webgpu-octree-pipelined-mgpcg.ts:66-73
 *   fp_grid_destroy  <- WebGPUUniformPressureMultigrid.destroy (…multigrid.ts:595-606)
 *   fp_submit/fp_poll   no reference counterpart (the reference solves one grid per
 *                       frame); they serve many independent projections.
 *
 * Precision fallback: an FP16 multigrid whose values overflow (a non-finite
 * preconditioned residual with a finite right-hand side) is solved again with
 * the BF16 multigrid and reported with FP_FLAG_PRECISION_FALLBACK
 * (fp_project, fp_project_batch, fp_submit).
 *
 * Conventions
 *   Arrays are C order with z fastest.  Uniform grid (nx,ny,nz):
 *     cells  phi, pressure, target_divergence : (nx, ny, nz)
 *     u (nx+1, ny, nz), v (nx, ny+1, nz), w (nx, ny, nz+1); face u(i,j,k) lies
 *     between cells i-1 and i.  Domain-boundary faces are closed walls: their
 *     input values are ignored and they are written as 0.
 *   Tall-cell grid (nx, fine_ny, nz, band B): packed samples (nx, B+2, nz);
 *     packed row 0 is the tall-cell bottom sample, 1 its top sample, 2+k the
 *     cubic band cell base+k (lib/core/tall-cell-grid.ts:171-190 indexing,
 *     transposed to z-fastest).  Velocities live in face slots: v (nx, B+2, nz)
 *     with slot 0 the velocity of the base-1 faces inside the tall cell, 1 the
 *     face tall top -> band row 0, 1+k the face band k-1 -> k, B+1 the face
 *     band top -> air; u (nx+1, S, nz) and w (nx, S, nz+1), S = B+2+D (D the
 *     max neighbour base delta), between columns A (-) and N (+): slots 0 and 1
 *     the mean and per-row slope (about row (m-1)/2, m = min base) of an affine
 *     velocity over fine rows 0..m-1, slot 2+j the single fine row m+j up to
 *     the higher band top; other slots are unused and written as 0.  Pressure
 *     is linear between the two samples of a tall cell.
 *   Liquid iff phi < 0.  Pressure is in Pa, velocity in m/s, h in m, dt in s,
 *   rho in kg/m^3.
 *
 * Errors: every entry point returns FP_OK (0) or a negative FP_ERR_* code and
 * then fp_last_error() describes it (thread-local).  Numerical outcomes are
 * never errors: they are FP_FLAG_* bits in fp_report.flags.
 * Threading: a context is used by one host thread at a time.  A grid has at
 * most one projection in flight: fp_project, fp_project_batch, fp_submit and
 * the probes return FP_ERR_STATE for a grid whose submitted projection has
 * not been polled to completion.
 */
#ifndef FLUID_PROJECTION_H
#define FLUID_PROJECTION_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FP_ABI_VERSION 3

/* return codes */
#define FP_OK 0
#define FP_ERR_INVALID_ARGUMENT (-1)
#define FP_ERR_DEVICE (-2)
#define FP_ERR_OUT_OF_MEMORY (-3)
#define FP_ERR_LAYOUT (-4)
#define FP_ERR_BUDGET (-5)
#define FP_ERR_STATE (-6)
#define FP_ERR_UNSUPPORTED (-7)

/* fp_report.flags (after webgpu-octree-pipelined-mgpcg.ts:66-73) */
#define FP_FLAG_NONFINITE 0x1
#define FP_FLAG_NONPOSITIVE_CURVATURE 0x2
#define FP_FLAG_NONPOSITIVE_PRECONDITIONER 0x4
#define FP_FLAG_NONCONVERGENCE 0x8
#define FP_FLAG_UNANCHORED_COMPONENT 0x10
#define FP_FLAG_PRECISION_FALLBACK 0x20
#define FP_FLAG_COARSE_SINGULAR 0x40

#define FP_LAYOUT_UNIFORM 0
#define FP_LAYOUT_TALL_CELL 1

#define FP_MEMORY_HOST 0
#define FP_MEMORY_DEVICE 1

#define FP_PRECISION_FP32 0        /* FP32 multigrid and CG, FP64 refinement */
#define FP_PRECISION_MIXED_FP16 1  /* FP16 multigrid (per-level scaling), FP32 CG, FP64 refinement */
#define FP_PRECISION_MIXED_BF16 2  /* BF16 multigrid, FP32 CG, FP64 refinement */

/* fp_options.solver */
#define FP_SOLVER_MGPCG 0  /* FP64-refined MGPCG of the Neumann-wall system (default) */
#define FP_SOLVER_CM11A 1  /* the reference's CM11a separating-wall LCP multigrid
                              (lib/methods/uniform/webgpu-uniform-pressure-multigrid.ts:488-576):
                              box walls hold p >= 0, so liquid may pull away from them; FP32
                              with a fixed cycle schedule; uniform grids, fp_project only */

typedef struct fp_ctx fp_ctx;
typedef struct fp_grid fp_grid;

typedef struct {
  int32_t layout;   /* FP_LAYOUT_* */
  int32_t nx;       /* cells along x */
  int32_t ny;       /* uniform: cells along y; tall-cell: fine rows (fineNy) */
  int32_t nz;       /* cells along z */
  int32_t band;     /* tall-cell: cubic band rows B (regularLayers); else 0 */
  int32_t max_delta; /* tall-cell: max neighbour base delta D (0: 4, maximumNeighborDelta,
                        lib/core/tall-cell-grid.ts:55); else 0 */
  double h;         /* cubic cell size (m) */
  const int32_t* column_bases; /* tall-cell: (nx, nz) host array, tall height per column in [2, ny],
                                   neighbours within D (limitNeighboringTallCellBases) */
} fp_grid_desc;

typedef struct {
  int32_t memory;   /* FP_MEMORY_HOST or FP_MEMORY_DEVICE for every array below */
  int32_t stream_order; /* 1: order the projection after `stream` below; 0: no ordering */
  const float* phi; /* level set at samples, metres, negative in liquid */
  const float* u;
  const float* v;
  const float* w;
  const float* target_divergence; /* optional desired divergence (1/s) per sample, or NULL */
  double dt;
  double rho;
  void* stream;     /* hipStream_t of the caller, used when stream_order is 1 (NULL is the
                       null stream).  The projection's device work then starts after the
                       work queued on it before the call (fp_project, fp_project_batch,
                       fp_submit), so device arrays produced or reused there are read and
                       written in order. */
  const float* density; /* optional (ABI 3), uniform grids: the liquid volume fraction per cell, the
                           reference's surface density.  When given, phi is not read and the level
                           set is pressurePhi's phi = -(density - 1/2) h (webgpu-uniform-reference.
                           wgsl.ts:215-242 with open fraction 1: box containers, no embedded solids) */
} fp_inputs;

typedef struct {
  float* pressure;  /* optional (NULL): pressure per sample (Pa) */
  float* u;         /* projected face velocities; may alias fp_inputs u/v/w */
  float* v;
  float* w;
} fp_outputs;

typedef struct {
  double relative_tolerance; /* FP64 ||b - A p||_2 / ||b||_2 target (default 1e-6) */
  int32_t max_iterations;    /* inner MGPCG iterations per refinement pass (default 100) */
  int32_t max_outer;         /* FP64 refinement passes (default 4) */
  int32_t precision;         /* FP_PRECISION_* (default MIXED_FP16) */
  int32_t pre_sweeps;        /* red-black GS sweeps before restriction (default 2) */
  int32_t post_sweeps;       /* ... after prolongation (default 2) */
  int32_t warm_start;        /* 1: start from the grid's resident pressure of the last solve */
  int32_t phi_preserve_levels; /* CM11a C (default 2) */
  int32_t check_every;       /* host convergence poll period in iterations (default 1) */
  int32_t timed_kernel;      /* FP_KERNEL_* whose launches are timed with HIP events (0: none) */
  int32_t graphs;            /* 1: replay each pair of CG + V-cycle iterations as one hipGraph (default) */
  int32_t precision_test_shift; /* test hook (0 in production): scales the multigrid input normalisation
                                   by 2^k, k in [0, 64], to force an FP16 overflow from the host */
  int32_t persistent;        /* 1: the column-tile sweeps run as persistent launches (a fixed grid of
                                workgroups walking each stage's table) instead of one workgroup per tile */
  int32_t solver;            /* FP_SOLVER_* (ABI 3) */
  int32_t open_top;          /* FP_SOLVER_CM11A: 1 opens the +y side of the box to air (p = 0 beyond it,
                                params.boundary.w of the reference) */
  int32_t full_cycles;       /* FP_SOLVER_CM11A: full multigrid cycles (0: 3, UNIFORM_CM11A_FULL_CYCLES) */
  int32_t v_cycles;          /* FP_SOLVER_CM11A: V-cycles after them (0: 4) */
  int32_t lcp_sweeps;        /* FP_SOLVER_CM11A: projected red-black sweeps before and after each
                                coarse correction (0: 6) */
} fp_options;

/* kernel classes for fp_options.timed_kernel */
#define FP_KERNEL_NONE 0
#define FP_KERNEL_SMOOTH_FINE 1    /* red-black colour pass on the finest level */
#define FP_KERNEL_RESIDUAL_FINE 2  /* r - A e on the finest level */
#define FP_KERNEL_CG_DIRECTION 3   /* d = z + beta d, q = A d, d.q */
#define FP_KERNEL_CG_UPDATE 4      /* x += alpha d, r -= alpha q, r.r */
#define FP_KERNEL_TRUE_RESIDUAL 5  /* FP64 b - A x */
#define FP_KERNEL_PROLONG_FINE 6   /* trilinear prolongation into the finest level */
#define FP_KERNEL_RESTRICT_FINE 7  /* restriction out of the finest level */

typedef struct {
  int32_t flags;             /* FP_FLAG_* */
  int32_t iterations;        /* inner MGPCG iterations, all passes */
  int32_t outer_iterations;  /* FP64 refinement passes */
  int32_t levels;            /* multigrid levels */
  double relative_residual;  /* final FP64 ||b - A p|| / ||b|| */
  double rhs_norm;           /* ||b||_2 (h^2-scaled units) */
  double max_div_before;     /* max |div u| over liquid samples before projection (1/s) */
  double max_div_after;      /* ... after, from the FP32 output velocities */
  double device_ms;          /* device time of the whole projection */
  int64_t liquid_samples;
  int64_t samples;
  double timed_kernel_ms;        /* summed device time of the timed kernel's launches */
  int32_t timed_kernel_launches; /* how many launches were timed */
  int32_t coarse_sweeps;         /* FP_SOLVER_CM11A: most sweeps of one coarsest LCP solve */
  double lcp_residual;           /* FP_SOLVER_CM11A: max projected residual in 1/s (mgMeasureFineResidual,
                                    …multigrid.wgsl.ts:473-481); relative_residual is then the
                                    2-norm of the projected residual over ||b|| */
  int32_t coarse_failures;       /* FP_SOLVER_CM11A: coarsest solves that reached the 4096-sweep cap
                                    (FP_FLAG_NONCONVERGENCE) */
  int32_t reserved;
} fp_report;

/* ---- context ------------------------------------------------------------ */
const char* fp_last_error(void);
int fp_abi_version(void);
int fp_ctx_create(int32_t device, uint64_t hbm_budget_bytes, fp_ctx** out);
int fp_ctx_destroy(fp_ctx* ctx);
int fp_ctx_bytes_in_use(const fp_ctx* ctx, uint64_t* bytes);

/* ---- grids --------------------------------------------------------------- */
int fp_default_options(fp_options* options);
/* Bytes a grid of this description would hold on the device. */
int fp_grid_bytes(const fp_grid_desc* desc, int32_t precision, uint64_t* bytes);
int fp_grid_create(fp_ctx* ctx, const fp_grid_desc* desc, fp_grid** out);
int fp_grid_destroy(fp_grid* grid);
/* Multigrid plan: writes levels*3 dims (and strides) into the arrays of
 * capacity 3*max_levels; returns the level count through *levels. */
int fp_grid_levels(const fp_grid* grid, int32_t max_levels, int32_t* dims, int32_t* strides, int32_t* levels);
/* Face counts of the grid's three velocity arrays. */
int fp_grid_face_counts(const fp_grid* grid, int64_t* nu, int64_t* nv, int64_t* nw);

/* Launch shapes of a uniform grid's column-tile sweeps, picked per grid size
 * by the autotuner (fluid_amd/autotune.py).  No reference counterpart (the
 * reference's dispatch shapes are fixed 4x4x4 workgroups). */
typedef struct {
  int32_t sweep_tile_z; /* 0: automatic (32 when nz >= 32, else 16); 16 or 32 (32 needs nz >= 32) */
  int32_t sweep_groups; /* 0: automatic (8192); workgroups a level's sweep is split into, in [256, 65536],
                           which sets the x chunk each workgroup marches */
  int32_t sweep_pairs;  /* 0 (default): one launch per finest-level sweep; 1: pairs of finest-level sweeps
                           in one launch that reads each plane once (FP16/BF16 multigrid; the same bits,
                           slower on MI355X at 384^3: 0.358 ms per pair against 2 x 0.146 ms) */
  int32_t reserved;
} fp_tuning;
int fp_grid_set_tuning(fp_grid* grid, const fp_tuning* tuning);

/* ---- projection ------------------------------------------------------------ */
/* One complete projection: labels + theta assembly -> divergence/rhs ->
 * MGPCG to the FP64 relative tolerance -> gradient subtract.  Synchronous. */
int fp_project(fp_grid* grid, const fp_inputs* in, fp_outputs* out, const fp_options* options, fp_report* report);

/* Independent projections of several grids of one context in one batched
 * solve: every solver stage is one launch over all the grids, each grid keeps
 * its own scalars and convergence test, and its result is the same as when
 * projected alone.  in/out/reports are arrays of `count`; the options are
 * shared.  Synchronous.  (No reference counterpart: the reference projects one
 * grid per frame.) */
int fp_project_batch(fp_grid* const* grids, int32_t count, const fp_inputs* in, fp_outputs* out,
                     const fp_options* options, fp_report* reports);

/* fp_project for one uniform grid, allocating nothing: no hipMalloc, hipFree,
 * pinned-host allocation, event or graph is created by the call.  No
 * counterpart in the reference; it is for callers that cannot allocate
 * (code under HIP graph capture setup, tight serving loops, harnesses that
 * forbid allocation).  The launch tables, the solver state it gathers and the
 * poll counters live in a static device array of 4096 bytes per GPU (looked
 * up at fp_ctx_create); the grid's own buffers were allocated by
 * fp_grid_create.  Its kernels and results are those of fp_project without
 * graph replay (bitwise equal).  Every call on a GPU shares that GPU's array:
 * calls that overlap on one GPU (from several contexts or threads) race.
 * A launch table that would not fit the array is rejected
 * (FP_ERR_INVALID_ARGUMENT; a uniform grid has at most 120 parts, 2320 bytes
 * with the state and counters, so none is).  Tall-cell grids:
 * FP_ERR_INVALID_ARGUMENT.  device_ms reports the call's host wall time. */
int fp_project_noalloc(fp_grid* grid, const fp_inputs* in, fp_outputs* out, const fp_options* options,
                       fp_report* report);

/* Serving (continuous batching).  fp_submit queues one projection and returns
 * at once; the context's worker thread advances every queued and running
 * projection by one stage per round (set-up, one CG + V-cycle iteration,
 * refinement, gradient subtract), each stage of all of them in one batched
 * launch, so requests arriving while others run join them on the device.
 * The input and output arrays must stay valid until the ticket completes.
 * fp_poll waits up to timeout_ms (-1: forever, 0: no wait) and returns FP_OK
 * with the report (and the submit-to-completion latency) once the ticket is
 * done, FP_PENDING while it is not.  A grid with running tickets cannot be
 * destroyed (FP_ERR_STATE).  Synchronous calls on the context wait for the
 * worker's current round.  (No reference counterpart: the reference runs one
 * projection per frame on its render thread.) */
#define FP_PENDING 1
int fp_submit(fp_grid* grid, const fp_inputs* in, fp_outputs* out, const fp_options* options, uint64_t* ticket);
int fp_poll(fp_ctx* ctx, uint64_t ticket, int32_t timeout_ms, fp_report* report, double* latency_ms);

/* ---- tall-cell remeshing ---- */
/* Re-lays a tall-cell grid out around the surface of phi and moves phi and the
 * velocity face slots to the new layout, on the device.  Per column: the
 * lowest and highest fine rows where liquid meets air vertically (phi read
 * linearly inside the tall cell, from the band samples above it, 5 h beyond
 * the band; isInitialSurfaceCell, lib/core/tall-cell-grid.ts:263-279), the base
 * chooseTallCellBase picks for them with halos min(halo, band)
 * (tall-cell-grid.ts:61-82,327; a dry column takes the highest base,
 * :378-386), then base <= min(4-neighbour base) + max_delta to convergence
 * (limitNeighboringTallCellBases, :84-106, run for nx + nz passes at :395).
 * Quantities move as docs/papers/tallCells.txt:240-252 states: a band sample
 * copies the old value at its fine row (linear inside the old tall cell), a
 * tall cell's bottom and top samples are the least-squares line through the
 * old values of its rows; v slot 0 is the mean of the old faces inside the new
 * tall cell; u/w slots 0 and 1 are the least-squares affine fit over the rows
 * both columns' new tall cells share, slot 2+j the old value at row m+j up to
 * the higher band top (0 above).  Columns (faces) whose base(s) did not change
 * are copied unchanged.  FP64 arithmetic without contraction.
 * The grid keeps the new bases (fp_project then uses them); its resident
 * warm-start pressure is cleared.  Outputs may alias the inputs; a NULL
 * output is not written.  Device arrays must be complete when the call is
 * made (it orders itself after no other stream); it returns when the outputs
 * are written.  The first call allocates staging for one copy of
 * the inputs and one of the outputs (FP_ERR_BUDGET if the context's budget
 * cannot hold it).  Grids with fine_ny - band < 2 have no restricted layout
 * (the reference then keeps an ordinary grid): FP_ERR_LAYOUT.
 * No reference entry point: the reference builds a layout once per scene
 * (createTallCellLayout, tall-cell-grid.ts:309-436). */
typedef struct {
  int32_t memory;       /* FP_MEMORY_HOST or FP_MEMORY_DEVICE for the arrays below */
  int32_t liquid_halo;  /* G_L, cells of liquid kept below the surface (0: 16, tallCellSettings.high) */
  int32_t air_halo;     /* G_A, cells of air kept above it (0: 8) */
  int32_t reserved;
  const float* phi;     /* (nx, B+2, nz) in the grid's current layout */
  const float* u;       /* (nx+1, S, nz) */
  const float* v;       /* (nx, B+2, nz) */
  const float* w;       /* (nx, S, nz+1) */
  float* phi_out;       /* the same shapes, in the new layout */
  float* u_out;
  float* v_out;
  float* w_out;
  int32_t* column_bases_out; /* (nx, nz) host array of the new bases, or NULL */
} fp_tall_remesh_args;
int fp_tall_remesh(fp_grid* grid, const fp_tall_remesh_args* args, int64_t* changed_columns);

/* ---- velocity extrapolation (uniform grids) ---- */
/* Extends every velocity component from the faces next to liquid into the
 * air, the stage after the projection in the reference's uniform solver
 * (WebGPUUniformVelocityExtrapolator.encode,
 * lib/methods/uniform/webgpu-uniform-velocity-extrapolation.ts:312-392, kernels
 * in …extrapolation.wgsl.ts).  Each component lives on its lattice of
 * positive faces (face u(i+1,j,k) of cell (i,j,k)); a face is open when it is
 * interior or, with open_top, the top face (faceOpenFraction,
 * webgpu-uniform-reference.wgsl.ts:511-519); a source face is open with liquid
 * on either side (liquid iff phi < 0, or density > 1/2) and keeps its input
 * value.  Jeong-Ross-Whitaker passes of the Godunov distance clipped to two
 * cells, each updated face taking the first-order upwind extension of CM11b
 * Eq. 7 (at most min(max(nx,ny,nz),16) passes, or max_passes), then the
 * hierarchy fill of CM11b Sec. 3.3.1 (levels ceil-halved until the smallest
 * extent is 1; renormalised trilinear restriction of known values, the
 * vertical component's corresponding-cell fallback, prolongation of the
 * unknown).  Outputs: known values on open faces, 0 on walls and on open
 * faces the hierarchy leaves unknown (packTransportShell, :491-512).  FP32.
 * Outputs may alias the inputs.  The first call on a grid allocates the
 * extrapolation state (about 57 bytes per cell; FP_ERR_BUDGET beyond the
 * context's budget).  Device arrays must be complete when the call is made;
 * it returns when the outputs are written. */
typedef struct {
  int32_t memory;       /* FP_MEMORY_HOST or FP_MEMORY_DEVICE for the arrays below */
  int32_t open_top;     /* 1: the top (+y) faces are open to the atmosphere */
  int32_t max_passes;   /* band passes (0: min(max(nx,ny,nz), 16), the reference's ceiling) */
  int32_t reserved;
  const float* phi;     /* (nx,ny,nz) level set, or NULL when density is given */
  const float* density; /* (nx,ny,nz) surface density (liquid iff > 1/2), used when phi is NULL */
  const float* u;       /* (nx+1,ny,nz) */
  const float* v;       /* (nx,ny+1,nz) */
  const float* w;       /* (nx,ny,nz+1) */
  float* u_out;
  float* v_out;
  float* w_out;
} fp_extrapolate_args;
typedef struct {
  int32_t passes;        /* band passes that ran */
  int32_t active_left;   /* band faces still active after the last pass (0: the front converged) */
  int32_t levels;        /* hierarchy levels */
  int32_t reserved;
  int64_t band_faces;    /* faces known after the band (sources included) */
  int64_t unknown_faces; /* open faces left unknown (written 0) */
  double device_ms;      /* host wall time of the call */
} fp_extrapolate_report;
int fp_extrapolate_velocity(fp_grid* grid, const fp_extrapolate_args* args, fp_extrapolate_report* report);

/* ---- velocity advection and body forces (uniform grids) ---- */
/* The step before the projection in the reference's uniform solver
 * (Algorithm 1 step 3; lib/methods/uniform/webgpu-uniform-reference.wgsl.ts):
 * every positive face is traced back along an RK2 characteristic through the
 * transported field (the extrapolated velocity, a sample outside the domain
 * reading 0; sampleVelocityComponent :261-270), in pieces of at most 1.5
 * cells, at most 32, each stage clamped into the box (departurePoint :324-334),
 * then:
 *   FP_ADVECT_SEMI_LAGRANGIAN  the sample there (semiLagrangianAdvection
 *                              :922-936; a positive wall face keeps
 *                              min(sample, its current velocity));
 *   FP_ADVECT_MACCORMACK       bounded MacCormack (advect :938-952, the
 *                              prediction extrapolated, reverseAdvection
 *                              :954-961, correctAdvection :993-1004 with
 *                              boundedMacCormack :963-991);
 * and body forces (applyVelocityForces :872-915): dt * viscosity * Laplacian
 * of the current velocity where the cell holds liquid (the wall reflecting the
 * velocity with no_slip), gravity * dt on the +y face where either cell holds
 * liquid above 1e-5 occupancy (clamp(1/2 - phi/4h), or clamp(density)).
 * Scope: box containers without embedded solids, inflow or surface tension.
 * The negative domain faces are copied.  transport: the extrapolation of
 * (u, v, w) (fp_extrapolate_velocity), or NULL for the library to compute it
 * from phi / density, which also extrapolate the MacCormack prediction.
 * FP32, in the reference's operation order.  Outputs may alias the inputs.
 * The first call on a grid allocates its state (FP_ERR_BUDGET beyond the
 * context's budget); device arrays must be complete when the call is made.
 * Grids of 2^30 or more faces per component, or of (nx+1)(ny+1) >= 2^24, are
 * refused with FP_ERR_UNSUPPORTED (32-bit offsets, 24-bit multiplies). */
#define FP_ADVECT_SEMI_LAGRANGIAN 0
#define FP_ADVECT_MACCORMACK 1
typedef struct {
  int32_t memory;       /* FP_MEMORY_HOST or FP_MEMORY_DEVICE for the arrays below */
  int32_t scheme;       /* FP_ADVECT_* */
  int32_t open_top;     /* 1: the top (+y) side is open (traces may leave it, no wall clamp) */
  int32_t no_slip;      /* 1: viscosity sees walls reflecting the velocity (fluidWallMode no-slip) */
  double dt;            /* s */
  double gravity;       /* m/s^2 along +y (e.g. -9.80665) */
  double viscosity;     /* kinematic, m^2/s (dynamicViscosity / density); 0: none */
  const float* phi;     /* (nx,ny,nz) level set of the advected surface, or NULL when density is given */
  const float* density; /* (nx,ny,nz) surface density, used when phi is NULL */
  const float* u;       /* current (projected) velocity: (nx+1,ny,nz) */
  const float* v;       /* (nx,ny+1,nz) */
  const float* w;       /* (nx,ny,nz+1) */
  const float* tu;      /* its extrapolation, the same shapes, or all three NULL */
  const float* tv;
  const float* tw;
  float* u_out;
  float* v_out;
  float* w_out;
} fp_advect_args;
typedef struct {
  double device_ms;     /* host wall time of the call */
  int32_t reserved[4];
} fp_advect_report;
int fp_advect_velocity(fp_grid* grid, const fp_advect_args* args, fp_advect_report* report);

/* ---- parity probes (device work on host arrays; used by the test suite) ---- */
/* out = A p (FP64) on the finest level, h^2-scaled. */
int fp_probe_apply(fp_grid* grid, const float* phi, const double* p, double* out);
/* b (FP64, h^2-scaled) from host inputs. */
int fp_probe_rhs(fp_grid* grid, const fp_inputs* in, double* b);
/* z = M^-1 r, one V-cycle of the preconditioner in the given precision. */
int fp_probe_vcycle(fp_grid* grid, const float* phi, const fp_options* options, const float* r, float* z);
/* The level-l level set of the multigrid hierarchy (FP32). */
int fp_probe_level_phi(fp_grid* grid, const float* phi, int32_t level, int32_t phi_preserve_levels, float* out);
/* Transfer taps between level-1 (fine) and level (coarse) along one axis, as
 * the kernels evaluate them: per fine index f its two prolongation taps
 * (prolong_coarse[2f+q], prolong_weight[2f+q]; coarse -1 where unused), per
 * coarse index c its four restriction taps (restrict_fine[4c+q],
 * restrict_weight[4c+q]; fine -1 where unused).  The reference's map is
 * mgFineChild / mgTrilinearPressure (…multigrid.wgsl.ts:71-76,217-241). */
int fp_probe_taps(fp_grid* grid, int32_t level, int32_t axis, int32_t* prolong_coarse, float* prolong_weight,
                  int32_t* restrict_fine, float* restrict_weight);
/* Tall-cell layout maps: the fine row of every packed sample and whether it
 * is an unknown (liquid and inside the domain).  No reference counterpart
 * (the reference computes these in buildInitialPhi, tall-cell-grid.ts:171-190). */
int fp_probe_tall_layout(fp_grid* grid, const float* phi, int32_t* rows, uint8_t* unknown);

#ifdef __cplusplus
}
#endif
#endif

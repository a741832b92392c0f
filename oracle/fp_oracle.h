/*
 * fp_oracle.h -- FP64 CPU oracle for the MAC-grid pressure projection.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the shipped library links, loads or
 * calls this code: only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg use it, and only as the checker.
 *
 * The oracle is a plain-C restatement of the reference's algebra:
 *   - ghost-fluid theta ........ lib/core/cm12-numerics.ts:37-43
 *   - A = G^T W G / theta ....... lib/methods/adaptive-mass/sparse-atlas-composite-projection.ts:1078-1082,1171-1192
 *   - rhs = G^T W u ............. sparse-atlas-composite-projection.ts:1238-1261
 *   - Jacobi-PCG + nullspace .... sparse-atlas-composite-projection.ts:1194-1211,1346-1457
 *   - Neumaier/Kahan dots ........ sparse-atlas-composite-projection.ts:1213-1225,
 *                                  lib/methods/octree-shared/octree-pipelined-pcg.ts:146-157
 *   - generic direct-curvature PCG lib/methods/octree-shared/octree-pipelined-pcg.ts:305-594
 *   - level planning ............ lib/methods/uniform/webgpu-uniform-pressure-multigrid.ts:99-169
 *   - coarse phi (CM11a Eq.15-16) lib/methods/uniform/webgpu-uniform-pressure-multigrid.wgsl.ts:169-195
 *   - gradient subtract ......... lib/methods/uniform/webgpu-uniform-reference.wgsl.ts:1052-1099
 *   - tall-cell layout .......... lib/core/tall-cell-grid.ts:61-106,309-436
 *
 * Storage order everywhere is C order with z fastest:
 *   cell (i,j,k)            -> (i*ny + j)*nz + k
 *   x-face u(i,j,k), i<=nx  -> (i*ny + j)*nz + k          (between cells i-1 and i)
 *   y-face v(i,j,k), j<=ny  -> (i*(ny+1) + j)*nz + k
 *   z-face w(i,j,k), k<=nz  -> (i*ny + j)*(nz+1) + k
 */
#ifndef FP_ORACLE_H
#define FP_ORACLE_H
#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_THETA_MIN 0.05
#define ORC_THETA_EPS 1e-9
#define ORC_MAX_LEVELS 16

/* ---- shared numerics ---------------------------------------------------- */
/* Sets the OpenMP thread count (n > 0) and returns the count in effect. */
int orc_set_threads(int n);
double orc_ghost_theta(double liquid_phi, double air_phi, double eps);
double orc_dot_kahan(const double* a, const double* b, size_t n);     /* sparse-atlas dot() */
double orc_dot_neumaier(const double* a, const double* b, size_t n);  /* octree-pipelined-pcg dot() */

/* ---- multigrid level planning ------------------------------------------- */
/* Returns the number of levels (>=1) or -1.  dims[l*3+a], strides[l*3+a]
 * (stride of level l relative to l-1; level 0 strides are 1).  `extended`
 * (out, may be NULL) is 1 when the reference planner rejects the lattice and
 * the ceil-halving extension produced the plan. */
int orc_plan_levels(int nx, int ny, int nz, int* dims, int* strides, int* extended);
/* The reference planner alone: returns levels or -1 on rejection. */
int orc_plan_levels_reference(int nx, int ny, int nz, int* dims);

/* ---- uniform MAC grid --------------------------------------------------- */
/* Face weights (h^2-scaled coefficients): 0 on walls and air-air faces,
 * 1 between two liquid cells, 1/theta across a liquid-air face. */
void orc_face_weights(int nx, int ny, int nz, const float* phi,
                      double* wx, double* wy, double* wz);
/* Level-l variant used by the multigrid: phi in double, per-axis geometric
 * factor g[a] multiplies every weight of axis a. */
void orc_face_weights_d(int nx, int ny, int nz, const double* phi, const double g[3],
                        double* wx, double* wy, double* wz);
/* b = -(rho h/dt) * sum(u_out - u_in) + (rho h^2/dt) q  on liquid cells. */
void orc_rhs(int nx, int ny, int nz, double h, double dt, double rho,
             const float* phi, const float* u, const float* v, const float* w,
             const float* q, double* b);
/* out = A p on liquid cells (0 elsewhere). */
void orc_apply(int nx, int ny, int nz, const float* phi,
               const double* wx, const double* wy, const double* wz,
               const double* p, double* out);
void orc_apply_d(int nx, int ny, int nz, const double* phi,
                 const double* wx, const double* wy, const double* wz,
                 const double* p, double* out);
void orc_diagonal(int nx, int ny, int nz, const float* phi,
                  const double* wx, const double* wy, const double* wz, double* diag);
/* Liquid components through liquid-liquid faces.  comp[c] = component id or
 * -1; anchored[id] = 1 when the component has a liquid-air face.  Returns the
 * component count. */
int orc_components(int nx, int ny, int nz, const float* phi,
                   const double* wx, const double* wy, const double* wz,
                   int32_t* comp, uint8_t* anchored);
/* Velocity update u -= dt/(rho h) * w * (p+ - p-), air p = 0; faces with no
 * liquid side and wall faces become 0.  Outputs may alias the inputs. */
void orc_project(int nx, int ny, int nz, double h, double dt, double rho,
                 const float* phi, const double* wx, const double* wy, const double* wz,
                 const double* p, const float* u, const float* v, const float* w,
                 double* uo, double* vo, double* wo);
/* max over liquid cells of |(1/h) sum(u_out - u_in) - q|. */
double orc_max_divergence(int nx, int ny, int nz, double h, const float* phi,
                          const double* u, const double* v, const double* w, const float* q);

/* ---- solvers ------------------------------------------------------------ */
typedef struct {
  int32_t iterations;
  int32_t converged;
  int32_t components;
  int32_t unanchored;
  double  rhs_norm;
  double  relative_residual;   /* ||b - A p|| / ||b||, recomputed, after nullspace removal */
  double  max_residual;
} orc_solve_report;

/* Sparse-atlas Jacobi-PCG (sparse-atlas-composite-projection.ts:1346-1457). */
int orc_jacobi_pcg(int nx, int ny, int nz, const float* phi, const double* b,
                   double rel_tol, double abs_tol, int max_iter,
                   double* p, orc_solve_report* rep);

typedef struct {
  int32_t pre_sweeps;
  int32_t post_sweeps;
  int32_t phi_preserve_levels;  /* C of CM11a (2) */
} orc_mg_options;

/* FP64 MGPCG with the multigrid V-cycle the GPU path uses. */
int orc_mgpcg(int nx, int ny, int nz, const float* phi, const double* b,
              const orc_mg_options* opt, double rel_tol, int max_iter,
              double* p, orc_solve_report* rep);
/* One application of the V-cycle preconditioner z = M^-1 r (for parity). */
int orc_vcycle_apply(int nx, int ny, int nz, const float* phi, const orc_mg_options* opt,
                     const double* r, double* z);
/* Coarse-level phi and weights of the hierarchy, for parity of the
 * coarsening maps: writes the level-l phi into out (size of level l). */
int orc_level_phi(int nx, int ny, int nz, const float* phi, int phi_preserve_levels,
                  int level, double* out);
/* 1-D cell-centred prolongation taps of fine index f from a coarse axis of
 * nc cells and stride 1 or 2 (mgTrilinearPressure, multigrid.wgsl.ts:217-241:
 * weights 3/4 and 1/4, an out-of-grid tap dropped and the rest renormalised);
 * returns the tap count (1 or 2).  Restriction is its transpose. */
int orc_taps1d(int f, int stride, int nc, int* idx, double* wt);

/* Generic direct-curvature PCG on a dense row-major SPD matrix with an
 * optional diagonal (Jacobi) preconditioner -- octree-pipelined-pcg.ts:305-594.
 * Returns 0 on acceptance, otherwise a failure code:
 *   11/12/13 operator not linear/symmetric/positive (audit),
 *   21/22/23 preconditioner not linear/symmetric/positive,
 *   3 non-positive preconditioned residual, 4 non-positive curvature,
 *   5 maximum iterations, 6 true-residual validation, 7 non-finite. */
int orc_dense_pcg(int n, const double* A, const double* b, const double* x0,
                  int use_jacobi, int max_iter, double rel_tol, double abs_tol,
                  int validate, double* x, int* iterations, int* reductions,
                  double* residual_norm);

/* ---- restricted tall-cell layout ---------------------------------------- */
/* chooseTallCellBase (tall-cell-grid.ts:61-82). */
int orc_tall_choose_base(int lowest_surface, int highest_surface, int fine_ny,
                         int regular_layers, int liquid_halo, int air_halo);
/* limitNeighboringTallCellBases (tall-cell-grid.ts:84-106), in place. */
void orc_tall_limit_bases(int32_t* bases, int nx, int nz, int max_delta, int passes);

/* Tall-cell projection (variational reduction of the uniform operator; the
 * layout and slot conventions are described in fp_oracle.c).  Samples are
 * packed [nx][B+2][nz]; u faces [nx+1][S][nz], w faces [nx][S][nz+1], v faces
 * [nx][B+2][nz], S = B + 2 + max_delta. */
typedef struct {
  int nx, ny, nz, band, max_delta;
  const int32_t* bases; /* [nx][nz], 2 <= base <= ny */
  const float* phi;     /* packed samples */
  double h;
} orc_tall;
#define ORC_TALL_BAD_BASE 31
int orc_tall_check(const orc_tall* t);
void orc_tall_unknowns(const orc_tall* t, uint8_t* mask);
int orc_tall_apply(const orc_tall* t, const double* p, double* out);
int orc_tall_rhs(const orc_tall* t, const double* u, const double* v, const double* w, double dt, double rho,
                 double* b);
int orc_tall_subtract(const orc_tall* t, const double* p, const double* u, const double* v, const double* w,
                      double dt, double rho, double* uo, double* vo, double* wo);
int orc_tall_diagonal(const orc_tall* t, double* d);
int orc_tall_solve(const orc_tall* t, const double* b, double rel_tol, int max_iter, double* p, int* iterations,
                   double* relative_residual);
double orc_tall_max_divergence(const orc_tall* t, const double* u, const double* v, const double* w);
int64_t orc_tall_coo(const orc_tall* t, int64_t* rows, int64_t* cols, double* vals, int64_t cap);

/* ---- CM11a separating-wall LCP solve (fp_oracle_cm11a.c) ---- */
typedef struct {
  int full_cycles, v_cycles, pre, post, constraint_levels, preserve, coarse_cap;
  float coarse_tol; /* coarsest projected residual, 1/s */
  int open_top;
} orc_cm11a_options;
typedef struct {
  int levels;
  int coarse_sweeps;   /* most sweeps of one coarsest solve */
  int coarse_failed;   /* coarsest solves that reached the cap */
  int coarse_solves;
  float lcp_residual;  /* finest projected residual max, 1/s */
  float coarse_residual;
} orc_cm11a_report;
void orc_cm11a_default_options(orc_cm11a_options* o);
int orc_cm11a_project(int nx, int ny, int nz, double h, const float* phi, const float* u, const float* v,
                      const float* w, const float* q, double dt, double rho, const orc_cm11a_options* opt,
                      float* p_padded, float* u_out, float* v_out, float* w_out, orc_cm11a_report* rep);
int orc_cm11a_system(int nx, int ny, int nz, double h, const float* phi, const float* u, const float* v,
                     const float* w, double dt, double rho, int open_top, float* coef, float* b, float* pmin);

/* ---- uniform velocity extrapolation (fp_oracle_extrap.c) ---- */
typedef struct {
  int passes;          /* FIM passes executed */
  int active_left;     /* active faces after the last pass (0: the front converged) */
  int levels;          /* hierarchy levels */
  int reserved;
  int64_t band_known;  /* faces known after the narrow band, sources included */
  int64_t unknown;     /* open faces left unknown */
} orc_extrap_report;
/* phi (or density when phi is NULL) per cell; u, v, w face arrays; outputs of
 * the same shapes.  band_dist / band_val (3, nx, ny, nz) and band_known (nx, ny,
 * nz; bit c) receive the narrow-band state when not NULL.  max_passes 0: the
 * reference's min(max(dims), 16). */
int orc_extrapolate(int nx, int ny, int nz, double h, int open_top, int max_passes, const float* phi,
                    const float* density, const float* u, const float* v, const float* w, float* uo, float* vo,
                    float* wo, float* band_dist, float* band_val, uint8_t* band_known, orc_extrap_report* rep);

/* ---- uniform velocity advection and body forces (fp_oracle_advect.c) ---- */
/* scheme 0: semi-Lagrangian, 1: bounded MacCormack (the predicted field is
 * extrapolated with orc_extrapolate on the same phi / density).  u, v, w: the
 * current (projected) face velocities; tu, tv, tw: their extrapolation (the
 * transported field); outputs of the face shapes. */
int orc_advect_velocity(int nx, int ny, int nz, double h, double dt, double gravity, double nu, int scheme,
                        int open_top, int no_slip, const float* phi, const float* density, const float* u,
                        const float* v, const float* w, const float* tu, const float* tv, const float* tw,
                        float* uo, float* vo, float* wo);

#ifdef __cplusplus
}
#endif
#endif

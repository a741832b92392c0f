/*
 * fp_oracle.c -- FP64 CPU oracle for the MAC-grid pressure projection.
 *
 * TEST INFRASTRUCTURE ONLY: used by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py as the checker.  The product path never links
 * or loads it.  Every function names the reference file:line it restates.
 *
 * Parity status of each function (how it is pinned) is listed in the tests
 * that exercise it (tests/test_oracle_*.py); nothing here is "parity
 * unpinned" except where a test says so.
 */
#include "fp_oracle.h"

#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

#define CIDX(i, j, k) (((size_t)(i) * ny + (j)) * nz + (k))
#define UIDX(i, j, k) (((size_t)(i) * ny + (j)) * nz + (k))
#define VIDX(i, j, k) (((size_t)(i) * (ny + 1) + (j)) * nz + (k))
#define WIDX(i, j, k) (((size_t)(i) * ny + (j)) * (nz + 1) + (k))

/* OpenMP threads of the stencil loops (the CPU baseline times 1 and all). */
int orc_set_threads(int n) {
  if (n > 0) omp_set_num_threads(n);
  return omp_get_max_threads();
}

/* cm12GhostFluidTheta, lib/core/cm12-numerics.ts:37-43. */
double orc_ghost_theta(double liquid_phi, double air_phi, double eps) {
  double den = fabs(liquid_phi) + fabs(air_phi);
  if (den < eps) den = eps;
  double t = fabs(liquid_phi) / den;
  if (t < ORC_THETA_MIN) t = ORC_THETA_MIN;
  if (t > 1.0) t = 1.0;
  return t;
}

/* Kahan dot of sparse-atlas-composite-projection.ts:1213-1225. */
double orc_dot_kahan(const double* a, const double* b, size_t n) {
  double sum = 0.0, correction = 0.0;
  for (size_t i = 0; i < n; ++i) {
    double product = a[i] * b[i];
    double adjusted = product - correction;
    double next = sum + adjusted;
    correction = (next - sum) - adjusted;
    sum = next;
  }
  return sum;
}

/* Neumaier dot of octree-pipelined-pcg.ts:146-157. */
double orc_dot_neumaier(const double* a, const double* b, size_t n) {
  double sum = 0.0, comp = 0.0;
  for (size_t i = 0; i < n; ++i) {
    double product = a[i] * b[i];
    double next = sum + product;
    comp += fabs(sum) >= fabs(product) ? (sum - next) + product : (product - next) + sum;
    sum = next;
  }
  return sum + comp;
}

/* ------------------------------------------------------------------------ */
/* Level planning: planUniformCM11aHierarchy and its two rules,             */
/* webgpu-uniform-pressure-multigrid.ts:99-169.                             */
/* ------------------------------------------------------------------------ */
static long haloed(const int* d) { return (long)(d[0] + 2) * (d[1] + 2) * (d[2] + 2); }

static int lockstep_levels(const int* d0, int* dims) {
  int minimum = d0[0];
  if (d0[1] < minimum) minimum = d0[1];
  if (d0[2] < minimum) minimum = d0[2];
  if (minimum <= 0 || (minimum & (minimum - 1)) != 0) return -1;
  int count = 0;
  while ((1 << (count + 1)) <= minimum) ++count; /* floor(log2(minimum)) */
  if (count < 1) return -1;
  int coarsening = 1 << (count - 1);
  for (int a = 0; a < 3; ++a)
    if (d0[a] % coarsening != 0) return -1;
  if (count > ORC_MAX_LEVELS) return -1;
  for (int l = 0; l < count; ++l)
    for (int a = 0; a < 3; ++a) dims[l * 3 + a] = d0[a] >> l;
  if (haloed(&dims[(count - 1) * 3]) > 256) return -1;
  return count;
}

static int semi_levels(const int* d0, int* dims) {
  int count = 1;
  for (int a = 0; a < 3; ++a) dims[a] = d0[a];
  for (;;) {
    const int* prev = &dims[(count - 1) * 3];
    int next[3], same = 1;
    for (int a = 0; a < 3; ++a) {
      next[a] = (prev[a] % 2 == 0 && prev[a] > 2) ? prev[a] / 2 : prev[a];
      if (next[a] != prev[a]) same = 0;
    }
    if (same) return count;
    if (count >= ORC_MAX_LEVELS) return count;
    for (int a = 0; a < 3; ++a) dims[count * 3 + a] = next[a];
    ++count;
  }
}

int orc_plan_levels_reference(int nx, int ny, int nz, int* dims) {
  int d0[3] = {nx, ny, nz};
  for (int a = 0; a < 3; ++a)
    if (d0[a] < 2) return -1;
  int count = lockstep_levels(d0, dims);
  if (count < 0) count = semi_levels(d0, dims);
  if (haloed(&dims[(count - 1) * 3]) > 256) return -1;
  return count;
}

/* Extension for lattices the reference rejects (odd edges such as 37^3):
 * ceil-halve every axis longer than two cells until the haloed coarsest
 * lattice fits 256 cells. */
static int ceil_levels(const int* d0, int* dims) {
  int count = 1;
  for (int a = 0; a < 3; ++a) dims[a] = d0[a];
  while (haloed(&dims[(count - 1) * 3]) > 256 && count < ORC_MAX_LEVELS) {
    const int* prev = &dims[(count - 1) * 3];
    int same = 1;
    for (int a = 0; a < 3; ++a) {
      int v = prev[a] > 2 ? (prev[a] + 1) / 2 : prev[a];
      if (v != prev[a]) same = 0;
      dims[count * 3 + a] = v;
    }
    if (same) break;
    ++count;
  }
  return count;
}

int orc_plan_levels(int nx, int ny, int nz, int* dims, int* strides, int* extended) {
  int d0[3] = {nx, ny, nz};
  for (int a = 0; a < 3; ++a)
    if (d0[a] < 2) return -1;
  int ext = 0;
  int count = orc_plan_levels_reference(nx, ny, nz, dims);
  if (count < 0) {
    ext = 1;
    count = ceil_levels(d0, dims);
    if (haloed(&dims[(count - 1) * 3]) > 256) return -1;
  }
  if (extended) *extended = ext;
  if (strides) {
    for (int l = 0; l < count; ++l)
      for (int a = 0; a < 3; ++a)
        strides[l * 3 + a] = (l == 0 || dims[l * 3 + a] == dims[(l - 1) * 3 + a]) ? 1 : 2;
  }
  return count;
}

/* ------------------------------------------------------------------------ */
/* Uniform MAC operator                                                     */
/* ------------------------------------------------------------------------ */
static double face_weight(double pa, double pb, double g) {
  int la = pa < 0.0, lb = pb < 0.0; /* mgLiquid: phi < 0, multigrid.wgsl.ts:80 */
  if (!la && !lb) return 0.0;
  if (la && lb) return g;
  /* mgCoefficientRaw, multigrid.wgsl.ts:93-115: theta from the liquid side. */
  double t = la ? orc_ghost_theta(pa, pb, ORC_THETA_EPS) : orc_ghost_theta(pb, pa, ORC_THETA_EPS);
  return g / t;
}

void orc_face_weights_d(int nx, int ny, int nz, const double* phi, const double g[3],
                        double* wx, double* wy, double* wz) {
  for (int i = 0; i <= nx; ++i)
    for (int j = 0; j < ny; ++j)
      for (int k = 0; k < nz; ++k)
        wx[UIDX(i, j, k)] = (i == 0 || i == nx) ? 0.0 : face_weight(phi[CIDX(i - 1, j, k)], phi[CIDX(i, j, k)], g[0]);
  for (int i = 0; i < nx; ++i)
    for (int j = 0; j <= ny; ++j)
      for (int k = 0; k < nz; ++k)
        wy[VIDX(i, j, k)] = (j == 0 || j == ny) ? 0.0 : face_weight(phi[CIDX(i, j - 1, k)], phi[CIDX(i, j, k)], g[1]);
  for (int i = 0; i < nx; ++i)
    for (int j = 0; j < ny; ++j)
      for (int k = 0; k <= nz; ++k)
        wz[WIDX(i, j, k)] = (k == 0 || k == nz) ? 0.0 : face_weight(phi[CIDX(i, j, k - 1)], phi[CIDX(i, j, k)], g[2]);
}

void orc_face_weights(int nx, int ny, int nz, const float* phi, double* wx, double* wy, double* wz) {
  size_t n = (size_t)nx * ny * nz;
  double* pd = (double*)malloc(n * sizeof(double));
  for (size_t c = 0; c < n; ++c) pd[c] = phi[c];
  const double g[3] = {1.0, 1.0, 1.0};
  orc_face_weights_d(nx, ny, nz, pd, g, wx, wy, wz);
  free(pd);
}

/* divergenceAt (webgpu-uniform-reference.wgsl.ts:1010-1027) with no solids:
 * V = 1 on every interior face, closed walls carry no flux; the sign and
 * scale follow mgBuildFinestRhs (multigrid.wgsl.ts:152-167), multiplied by
 * h^2 so that every coefficient of A is dimensionless. */
void orc_rhs(int nx, int ny, int nz, double h, double dt, double rho,
             const float* phi, const float* u, const float* v, const float* w,
             const float* q, double* b) {
  const double s = rho * h / dt, sq = rho * h * h / dt;
  for (int i = 0; i < nx; ++i)
    for (int j = 0; j < ny; ++j)
      for (int k = 0; k < nz; ++k) {
        size_t c = CIDX(i, j, k);
        if (!(phi[c] < 0.0f)) { b[c] = 0.0; continue; }
        double up = (i + 1 < nx) ? u[UIDX(i + 1, j, k)] : 0.0;
        double um = (i > 0) ? u[UIDX(i, j, k)] : 0.0;
        double vp = (j + 1 < ny) ? v[VIDX(i, j + 1, k)] : 0.0;
        double vm = (j > 0) ? v[VIDX(i, j, k)] : 0.0;
        double wp = (k + 1 < nz) ? w[WIDX(i, j, k + 1)] : 0.0;
        double wm = (k > 0) ? w[WIDX(i, j, k)] : 0.0;
        double flux = (up - um) + (vp - vm) + (wp - wm);
        b[c] = -s * flux + (q ? sq * (double)q[c] : 0.0);
      }
}

static void apply_generic(int nx, int ny, int nz, const double* phid, const float* phif,
                          const double* wx, const double* wy, const double* wz,
                          const double* p, double* out) {
#define LIQ(c) (phid ? (phid[c] < 0.0) : (phif[c] < 0.0f))
#pragma omp parallel for collapse(2) schedule(static)
  for (int i = 0; i < nx; ++i)
    for (int j = 0; j < ny; ++j)
      for (int k = 0; k < nz; ++k) {
        size_t c = CIDX(i, j, k);
        if (!LIQ(c)) { out[c] = 0.0; continue; }
        double pc = p[c], acc = 0.0, wf;
        /* sparse-atlas applyLiquidOperator: sum_rows c (W/theta) (sum c p). */
        wf = wx[UIDX(i, j, k)];     if (wf != 0.0) { size_t n = CIDX(i - 1, j, k); acc += wf * (pc - (LIQ(n) ? p[n] : 0.0)); }
        wf = wx[UIDX(i + 1, j, k)]; if (wf != 0.0) { size_t n = CIDX(i + 1, j, k); acc += wf * (pc - (LIQ(n) ? p[n] : 0.0)); }
        wf = wy[VIDX(i, j, k)];     if (wf != 0.0) { size_t n = CIDX(i, j - 1, k); acc += wf * (pc - (LIQ(n) ? p[n] : 0.0)); }
        wf = wy[VIDX(i, j + 1, k)]; if (wf != 0.0) { size_t n = CIDX(i, j + 1, k); acc += wf * (pc - (LIQ(n) ? p[n] : 0.0)); }
        wf = wz[WIDX(i, j, k)];     if (wf != 0.0) { size_t n = CIDX(i, j, k - 1); acc += wf * (pc - (LIQ(n) ? p[n] : 0.0)); }
        wf = wz[WIDX(i, j, k + 1)]; if (wf != 0.0) { size_t n = CIDX(i, j, k + 1); acc += wf * (pc - (LIQ(n) ? p[n] : 0.0)); }
        out[c] = acc;
      }
#undef LIQ
}

void orc_apply(int nx, int ny, int nz, const float* phi, const double* wx, const double* wy,
               const double* wz, const double* p, double* out) {
  apply_generic(nx, ny, nz, NULL, phi, wx, wy, wz, p, out);
}

void orc_apply_d(int nx, int ny, int nz, const double* phi, const double* wx, const double* wy,
                 const double* wz, const double* p, double* out) {
  apply_generic(nx, ny, nz, phi, NULL, wx, wy, wz, p, out);
}

static void diagonal_generic(int nx, int ny, int nz, const double* phid, const float* phif,
                             const double* wx, const double* wy, const double* wz, double* diag) {
  for (int i = 0; i < nx; ++i)
    for (int j = 0; j < ny; ++j)
      for (int k = 0; k < nz; ++k) {
        size_t c = CIDX(i, j, k);
        int liq = phid ? (phid[c] < 0.0) : (phif[c] < 0.0f);
        diag[c] = liq ? (wx[UIDX(i, j, k)] + wx[UIDX(i + 1, j, k)]) + (wy[VIDX(i, j, k)] + wy[VIDX(i, j + 1, k)]) +
                            (wz[WIDX(i, j, k)] + wz[WIDX(i, j, k + 1)])
                      : 0.0;
      }
}

void orc_diagonal(int nx, int ny, int nz, const float* phi, const double* wx, const double* wy,
                  const double* wz, double* diag) {
  diagonal_generic(nx, ny, nz, NULL, phi, wx, wy, wz, diag);
}

/* Connected liquid components and their anchoring, after
 * sparse-atlas-composite-projection.ts:1087-1146 (a component is anchored when
 * one of its rows is a cut liquid-air row). */
int orc_components(int nx, int ny, int nz, const float* phi, const double* wx, const double* wy,
                   const double* wz, int32_t* comp, uint8_t* anchored) {
  size_t n = (size_t)nx * ny * nz;
  for (size_t c = 0; c < n; ++c) comp[c] = -1;
  size_t* stack = (size_t*)malloc(n * sizeof(size_t));
  int count = 0;
  for (size_t start = 0; start < n; ++start) {
    if (!(phi[start] < 0.0f) || comp[start] >= 0) continue;
    int id = count++;
    int anchor = 0;
    size_t top = 0;
    stack[top++] = start;
    comp[start] = id;
    while (top > 0) {
      size_t c = stack[--top];
      int k = (int)(c % nz), j = (int)((c / nz) % ny), i = (int)(c / ((size_t)nz * ny));
      struct { int di, dj, dk; double w; } nb[6] = {
          {-1, 0, 0, wx[UIDX(i, j, k)]}, {1, 0, 0, wx[UIDX(i + 1, j, k)]},
          {0, -1, 0, wy[VIDX(i, j, k)]}, {0, 1, 0, wy[VIDX(i, j + 1, k)]},
          {0, 0, -1, wz[WIDX(i, j, k)]}, {0, 0, 1, wz[WIDX(i, j, k + 1)]}};
      for (int f = 0; f < 6; ++f) {
        if (nb[f].w == 0.0) continue; /* wall */
        size_t m = CIDX(i + nb[f].di, j + nb[f].dj, k + nb[f].dk);
        if (!(phi[m] < 0.0f)) { anchor = 1; continue; } /* liquid-air row */
        if (comp[m] < 0) { comp[m] = id; stack[top++] = m; }
      }
    }
    anchored[id] = (uint8_t)anchor;
  }
  free(stack);
  return count;
}

/* Gradient subtract, webgpu-uniform-reference.wgsl.ts:1052-1099 (interior
 * faces) with the no-flow walls of sparse-atlas-composite-projection.ts:12-13;
 * inactive rows are zeroed as in sparse-atlas-composite-projection.ts:1468-1494. */
void orc_project(int nx, int ny, int nz, double h, double dt, double rho, const float* phi,
                 const double* wx, const double* wy, const double* wz, const double* p,
                 const float* u, const float* v, const float* w, double* uo, double* vo, double* wo) {
  const double s = dt / (rho * h);
#define PL(c) ((phi[c] < 0.0f) ? p[c] : 0.0)
  for (int i = 0; i <= nx; ++i)
    for (int j = 0; j < ny; ++j)
      for (int k = 0; k < nz; ++k) {
        size_t f = UIDX(i, j, k);
        double wf = wx[f];
        uo[f] = (wf == 0.0) ? 0.0 : (double)u[f] - s * wf * (PL(CIDX(i, j, k)) - PL(CIDX(i - 1, j, k)));
      }
  for (int i = 0; i < nx; ++i)
    for (int j = 0; j <= ny; ++j)
      for (int k = 0; k < nz; ++k) {
        size_t f = VIDX(i, j, k);
        double wf = wy[f];
        vo[f] = (wf == 0.0) ? 0.0 : (double)v[f] - s * wf * (PL(CIDX(i, j, k)) - PL(CIDX(i, j - 1, k)));
      }
  for (int i = 0; i < nx; ++i)
    for (int j = 0; j < ny; ++j)
      for (int k = 0; k <= nz; ++k) {
        size_t f = WIDX(i, j, k);
        double wf = wz[f];
        wo[f] = (wf == 0.0) ? 0.0 : (double)w[f] - s * wf * (PL(CIDX(i, j, k)) - PL(CIDX(i, j, k - 1)));
      }
#undef PL
}

double orc_max_divergence(int nx, int ny, int nz, double h, const float* phi, const double* u,
                          const double* v, const double* w, const float* q) {
  double m = 0.0;
  for (int i = 0; i < nx; ++i)
    for (int j = 0; j < ny; ++j)
      for (int k = 0; k < nz; ++k) {
        size_t c = CIDX(i, j, k);
        if (!(phi[c] < 0.0f)) continue;
        double flux = (u[UIDX(i + 1, j, k)] - u[UIDX(i, j, k)]) + (v[VIDX(i, j + 1, k)] - v[VIDX(i, j, k)]) +
                      (w[WIDX(i, j, k + 1)] - w[WIDX(i, j, k)]);
        double d = fabs(flux / h - (q ? (double)q[c] : 0.0));
        if (d > m) m = d;
      }
  return m;
}

/* ------------------------------------------------------------------------ */
/* Nullspace handling (sparse-atlas-composite-projection.ts:1194-1211)       */
/* ------------------------------------------------------------------------ */
typedef struct {
  int count;
  int32_t* comp;
  uint8_t* anchored;
  double* sums;
  long* sizes;
} nullspace_t;

static void nullspace_project(const nullspace_t* ns, double* x, size_t n) {
  if (ns->count == 0) return;
  int any = 0;
  for (int c = 0; c < ns->count; ++c)
    if (!ns->anchored[c]) { any = 1; break; }
  if (!any) return;
  for (int c = 0; c < ns->count; ++c) ns->sums[c] = 0.0;
  for (size_t i = 0; i < n; ++i)
    if (ns->comp[i] >= 0 && !ns->anchored[ns->comp[i]]) ns->sums[ns->comp[i]] += x[i];
  for (size_t i = 0; i < n; ++i) {
    int c = ns->comp[i];
    if (c >= 0 && !ns->anchored[c]) x[i] -= ns->sums[c] / (double)ns->sizes[c];
  }
}

static void nullspace_init(nullspace_t* ns, int nx, int ny, int nz, const float* phi, const double* wx,
                           const double* wy, const double* wz) {
  size_t n = (size_t)nx * ny * nz;
  ns->comp = (int32_t*)malloc(n * sizeof(int32_t));
  ns->anchored = (uint8_t*)malloc(n > 0 ? n : 1);
  ns->count = orc_components(nx, ny, nz, phi, wx, wy, wz, ns->comp, ns->anchored);
  ns->sums = (double*)calloc((size_t)ns->count + 1, sizeof(double));
  ns->sizes = (long*)calloc((size_t)ns->count + 1, sizeof(long));
  for (size_t i = 0; i < n; ++i)
    if (ns->comp[i] >= 0) ns->sizes[ns->comp[i]]++;
}

static void nullspace_free(nullspace_t* ns) {
  free(ns->comp); free(ns->anchored); free(ns->sums); free(ns->sizes);
}

static int nullspace_unanchored(const nullspace_t* ns) {
  int u = 0;
  for (int c = 0; c < ns->count; ++c) u += !ns->anchored[c];
  return u;
}

static double norm2(const double* x, size_t n) {
  double d = orc_dot_kahan(x, x, n);
  return sqrt(d > 0 ? d : 0);
}

/* ------------------------------------------------------------------------ */
/* Jacobi-PCG exactly as sparse-atlas-composite-projection.ts:1371-1457.    */
/* ------------------------------------------------------------------------ */
int orc_jacobi_pcg(int nx, int ny, int nz, const float* phi, const double* b_in, double rel_tol,
                   double abs_tol, int max_iter, double* p, orc_solve_report* rep) {
  size_t n = (size_t)nx * ny * nz;
  double* wx = (double*)malloc((size_t)(nx + 1) * ny * nz * sizeof(double));
  double* wy = (double*)malloc((size_t)nx * (ny + 1) * nz * sizeof(double));
  double* wz = (double*)malloc((size_t)nx * ny * (nz + 1) * sizeof(double));
  orc_face_weights(nx, ny, nz, phi, wx, wy, wz);
  double* diag = (double*)malloc(n * sizeof(double));
  orc_diagonal(nx, ny, nz, phi, wx, wy, wz, diag);
  nullspace_t ns;
  nullspace_init(&ns, nx, ny, nz, phi, wx, wy, wz);
  double* b = (double*)malloc(n * sizeof(double));
  double* r = (double*)malloc(n * sizeof(double));
  double* z = (double*)malloc(n * sizeof(double));
  double* d = (double*)malloc(n * sizeof(double));
  double* ad = (double*)malloc(n * sizeof(double));
  memcpy(b, b_in, n * sizeof(double));
  nullspace_project(&ns, b, n);
  double bnorm = norm2(b, n);
  /* p holds the initial guess (zero for a cold start). */
  nullspace_project(&ns, p, n);
  orc_apply(nx, ny, nz, phi, wx, wy, wz, p, ad);
  for (size_t i = 0; i < n; ++i) r[i] = b[i] - ad[i];
  nullspace_project(&ns, r, n);
  for (size_t i = 0; i < n; ++i) z[i] = diag[i] > 0 ? r[i] / diag[i] : 0.0;
  nullspace_project(&ns, z, n);
  memcpy(d, z, n * sizeof(double));
  double rz = orc_dot_kahan(r, z, n);
  double rnorm = norm2(r, n);
  double target = abs_tol > rel_tol * bnorm ? abs_tol : rel_tol * bnorm;
  int it = 0, status = 0;
  while (it < max_iter && rnorm > target) {
    orc_apply(nx, ny, nz, phi, wx, wy, wz, d, ad);
    nullspace_project(&ns, ad, n);
    double curv = orc_dot_kahan(d, ad, n);
    if (!(curv > 0) || !isfinite(curv)) { status = 4; break; }
    double alpha = rz / curv;
    for (size_t i = 0; i < n; ++i) { p[i] += alpha * d[i]; r[i] -= alpha * ad[i]; }
    nullspace_project(&ns, p, n);
    nullspace_project(&ns, r, n);
    ++it;
    rnorm = norm2(r, n);
    if (rnorm <= target) break;
    for (size_t i = 0; i < n; ++i) z[i] = diag[i] > 0 ? r[i] / diag[i] : 0.0;
    nullspace_project(&ns, z, n);
    double rz_next = orc_dot_kahan(r, z, n);
    double beta = rz_next / rz;
    for (size_t i = 0; i < n; ++i) d[i] = z[i] + beta * d[i];
    nullspace_project(&ns, d, n);
    rz = rz_next;
  }
  /* true residual, sparse-atlas-composite-projection.ts:1459-1466 */
  orc_apply(nx, ny, nz, phi, wx, wy, wz, p, ad);
  for (size_t i = 0; i < n; ++i) r[i] = b[i] - ad[i];
  nullspace_project(&ns, r, n);
  double tn = norm2(r, n), mx = 0.0;
  for (size_t i = 0; i < n; ++i) if (fabs(r[i]) > mx) mx = fabs(r[i]);
  if (rep) {
    rep->iterations = it;
    rep->converged = tn <= target;
    rep->components = ns.count;
    rep->unanchored = nullspace_unanchored(&ns);
    rep->rhs_norm = bnorm;
    rep->relative_residual = bnorm > 0 ? tn / bnorm : 0.0;
    rep->max_residual = mx;
  }
  if (status == 0 && !(tn <= target)) status = 5;
  free(wx); free(wy); free(wz); free(diag); free(b); free(r); free(z); free(d); free(ad);
  nullspace_free(&ns);
  return status;
}

/* ------------------------------------------------------------------------ */
/* Multigrid hierarchy (FP64) used by the MGPCG oracle.                     */
/* ------------------------------------------------------------------------ */
typedef struct {
  int d[3];
  int s[3];       /* stride relative to the finer level */
  double g[3];    /* (h0 / h_l,a)^2 */
  size_t n;
  double* phi;
  double *wx, *wy, *wz;
  double* diag;
  double *r, *e, *t;
} orc_level;

typedef struct {
  int L;
  orc_level lv[ORC_MAX_LEVELS];
  /* coarsest dense Cholesky */
  int nc;
  int* cmap;      /* coarse liquid index -> cell */
  double* chol;   /* nc x nc lower factor */
  double* cdiag_ok;
  orc_mg_options opt;
} orc_mg;

/* FP32 sum of eight corner terms in the pairing of mgD4Sum8
 * (webgpu-uniform-pressure-multigrid.wgsl.ts:50-54). */
static float d4_sum8f(const float* t) {
  float y0 = (t[0] + t[5]) + (t[1] + t[4]);
  float y1 = (t[2] + t[7]) + (t[3] + t[6]);
  return y0 + y1;
}

/* mgDownsampleTopology phi rule (multigrid.wgsl.ts:169-195), in FP32 as the
 * reference's r32float level sets: the D4-ordered sum of the eight corner
 * children (mgFineChild, :71-76: an axis of stride 1 repeats its child) over
 * 8, or the sum of the non-negative children over their count when their
 * signs are mixed and the destination is one of the C finest coarse levels.
 * The last coarse cell of an odd axis (lattices the reference rejects, which
 * this build ceil-halves) repeats its one child the same way. */
static void downsample_phi(const orc_level* f, orc_level* c, int level, int preserve) {
  for (int I = 0; I < c->d[0]; ++I)
    for (int J = 0; J < c->d[1]; ++J)
      for (int K = 0; K < c->d[2]; ++K) {
        float v[8], ps[8], pn[8], ng[8];
        for (int o = 0; o < 8; ++o) {
          int i = I * c->s[0] + (o & 1) * (c->s[0] - 1);
          int j = J * c->s[1] + ((o >> 1) & 1) * (c->s[1] - 1);
          int k = K * c->s[2] + ((o >> 2) & 1) * (c->s[2] - 1);
          if (i > f->d[0] - 1) i = f->d[0] - 1;
          if (j > f->d[1] - 1) j = f->d[1] - 1;
          if (k > f->d[2] - 1) k = f->d[2] - 1;
          v[o] = (float)f->phi[((size_t)i * f->d[1] + j) * f->d[2] + k];
          ps[o] = v[o] >= 0.0f ? v[o] : 0.0f;
          pn[o] = v[o] >= 0.0f ? 1.0f : 0.0f;
          ng[o] = v[o] >= 0.0f ? 0.0f : 1.0f;
        }
        float sum = d4_sum8f(v), pos = d4_sum8f(ps), npos = d4_sum8f(pn), nneg = d4_sum8f(ng);
        int mixed = npos > 0.0f && nneg > 0.0f;
        float out = (mixed && level <= preserve) ? pos / (npos > 1.0f ? npos : 1.0f) : sum / 8.0f;
        c->phi[((size_t)I * c->d[1] + J) * c->d[2] + K] = (double)out;
      }
}

static void level_alloc(orc_level* l) {
  int nx = l->d[0], ny = l->d[1], nz = l->d[2];
  l->n = (size_t)nx * ny * nz;
  l->phi = (double*)calloc(l->n, sizeof(double));
  l->wx = (double*)calloc((size_t)(nx + 1) * ny * nz, sizeof(double));
  l->wy = (double*)calloc((size_t)nx * (ny + 1) * nz, sizeof(double));
  l->wz = (double*)calloc((size_t)nx * ny * (nz + 1), sizeof(double));
  l->diag = (double*)calloc(l->n, sizeof(double));
  l->r = (double*)calloc(l->n, sizeof(double));
  l->e = (double*)calloc(l->n, sizeof(double));
  l->t = (double*)calloc(l->n, sizeof(double));
}

static void level_free(orc_level* l) {
  free(l->phi); free(l->wx); free(l->wy); free(l->wz); free(l->diag);
  free(l->r); free(l->e); free(l->t);
}

/* A singular component of the coarsest block (no air face: an all-liquid
 * level) leaves a pivot that vanishes to rounding; it gets its row's
 * original diagonal added (the factor of A + a_jj e_j e_j^T, SPD). */
static int cholesky(double* a, int n) {
  for (int j = 0; j < n; ++j) {
    const double d0 = a[j * n + j];
    double s = d0;
    for (int k = 0; k < j; ++k) s -= a[j * n + k] * a[j * n + k];
    if (s <= 1e-9 * d0 && s > -1e-9 * d0) s += d0;
    if (!(s > 0.0)) return -1;
    double l = sqrt(s);
    a[j * n + j] = l;
    for (int i = j + 1; i < n; ++i) {
      double t = a[i * n + j];
      for (int k = 0; k < j; ++k) t -= a[i * n + k] * a[j * n + k];
      a[i * n + j] = t / l;
    }
    for (int i = 0; i < j; ++i) a[i * n + j] = 0.0;
  }
  return 0;
}

static int mg_build(orc_mg* mg, int nx, int ny, int nz, const float* phi, const orc_mg_options* opt) {
  memset(mg, 0, sizeof(*mg));
  mg->opt = *opt;
  int dims[ORC_MAX_LEVELS * 3], strides[ORC_MAX_LEVELS * 3];
  int L = orc_plan_levels(nx, ny, nz, dims, strides, NULL);
  if (L < 1) return -1;
  mg->L = L;
  for (int l = 0; l < L; ++l) {
    orc_level* lv = &mg->lv[l];
    for (int a = 0; a < 3; ++a) {
      lv->d[a] = dims[l * 3 + a];
      lv->s[a] = strides[l * 3 + a];
      lv->g[a] = l == 0 ? 1.0 : mg->lv[l - 1].g[a] / (double)(lv->s[a] * lv->s[a]);
    }
    level_alloc(lv);
    if (l == 0) {
      for (size_t c = 0; c < lv->n; ++c) lv->phi[c] = phi[c];
    } else {
      downsample_phi(&mg->lv[l - 1], lv, l, opt->phi_preserve_levels);
    }
    orc_face_weights_d(lv->d[0], lv->d[1], lv->d[2], lv->phi, lv->g, lv->wx, lv->wy, lv->wz);
    diagonal_generic(lv->d[0], lv->d[1], lv->d[2], lv->phi, NULL, lv->wx, lv->wy, lv->wz, lv->diag);
  }
  /* coarsest: dense SPD matrix over its liquid cells */
  orc_level* c = &mg->lv[L - 1];
  mg->cmap = (int*)malloc(c->n * sizeof(int));
  int* rev = (int*)malloc(c->n * sizeof(int));
  int nc = 0;
  for (size_t i = 0; i < c->n; ++i) {
    rev[i] = -1;
    if (c->phi[i] < 0.0 && c->diag[i] > 0.0) { rev[i] = nc; mg->cmap[nc++] = (int)i; }
  }
  mg->nc = nc;
  mg->chol = (double*)calloc((size_t)nc * nc + 1, sizeof(double));
  double* unit = (double*)calloc(c->n, sizeof(double));
  double* col = (double*)calloc(c->n, sizeof(double));
  for (int q = 0; q < nc; ++q) {
    memset(unit, 0, c->n * sizeof(double));
    unit[mg->cmap[q]] = 1.0;
    orc_apply_d(c->d[0], c->d[1], c->d[2], c->phi, c->wx, c->wy, c->wz, unit, col);
    for (int r = 0; r < nc; ++r) mg->chol[r * nc + q] = col[mg->cmap[r]];
  }
  free(unit); free(col); free(rev);
  if (nc > 0 && cholesky(mg->chol, nc) != 0) {
    /* A singular coarsest component: fall back to a diagonal scaling so the
     * preconditioner stays a fixed SPD map. */
    for (int r = 0; r < nc; ++r)
      for (int q = 0; q < nc; ++q) mg->chol[r * nc + q] = (r == q) ? sqrt(c->diag[mg->cmap[r]]) : 0.0;
  }
  return 0;
}

static void mg_free(orc_mg* mg) {
  for (int l = 0; l < mg->L; ++l) level_free(&mg->lv[l]);
  free(mg->cmap); free(mg->chol);
}

static void gs_colour(orc_level* l, int colour) {
  int nx = l->d[0], ny = l->d[1], nz = l->d[2];
  const double *wx = l->wx, *wy = l->wy, *wz = l->wz, *phi = l->phi;
  double* e = l->e;
#pragma omp parallel for collapse(2) schedule(static)
  for (int i = 0; i < nx; ++i)
    for (int j = 0; j < ny; ++j)
      for (int k = 0; k < nz; ++k) {
        if (((i + j + k) & 1) != colour) continue;
        size_t c = CIDX(i, j, k);
        if (!(phi[c] < 0.0) || !(l->diag[c] > 0.0)) { e[c] = 0.0; continue; }
        double acc = l->r[c], wf;
#define NB(W, ni, nj, nk) wf = (W); if (wf != 0.0) { size_t m = CIDX(ni, nj, nk); if (phi[m] < 0.0) acc += wf * e[m]; }
        NB(wx[UIDX(i, j, k)], i - 1, j, k)
        NB(wx[UIDX(i + 1, j, k)], i + 1, j, k)
        NB(wy[VIDX(i, j, k)], i, j - 1, k)
        NB(wy[VIDX(i, j + 1, k)], i, j + 1, k)
        NB(wz[WIDX(i, j, k)], i, j, k - 1)
        NB(wz[WIDX(i, j, k + 1)], i, j, k + 1)
#undef NB
        e[c] = acc / l->diag[c];
      }
}

/* 1-D prolongation weights of fine index f from coarse cells: returns the
 * number of taps (1 or 2). */
static int taps1d(int f, int stride, int nc, int* idx, double* wt) {
  if (stride == 1) { idx[0] = f; wt[0] = 1.0; return 1; }
  int I = f >> 1;
  int J = (f & 1) ? I + 1 : I - 1;
  if (J < 0 || J >= nc) { idx[0] = I; wt[0] = 1.0; return 1; }
  idx[0] = I; wt[0] = 0.75; idx[1] = J; wt[1] = 0.25;
  return 2;
}

int orc_taps1d(int f, int stride, int nc, int* idx, double* wt) { return taps1d(f, stride, nc, idx, wt); }

/* e_f += P e_c on liquid fine cells. */
static void prolong_add(orc_level* f, const orc_level* c) {
  for (int i = 0; i < f->d[0]; ++i) {
    int xi[2]; double xw[2]; int nxt = taps1d(i, c->s[0], c->d[0], xi, xw);
    for (int j = 0; j < f->d[1]; ++j) {
      int yi[2]; double yw[2]; int nyt = taps1d(j, c->s[1], c->d[1], yi, yw);
      for (int k = 0; k < f->d[2]; ++k) {
        size_t fc = ((size_t)i * f->d[1] + j) * f->d[2] + k;
        if (!(f->phi[fc] < 0.0)) continue;
        int zi[2]; double zw[2]; int nzt = taps1d(k, c->s[2], c->d[2], zi, zw);
        double acc = 0.0;
        for (int a = 0; a < nxt; ++a)
          for (int b = 0; b < nyt; ++b)
            for (int q = 0; q < nzt; ++q) {
              size_t cc = ((size_t)xi[a] * c->d[1] + yi[b]) * c->d[2] + zi[q];
              if (c->phi[cc] < 0.0) acc += xw[a] * yw[b] * zw[q] * c->e[cc];
            }
        f->e[fc] += acc;
      }
    }
  }
}

/* r_c = (P^T t_f) / 2^(coarsened axes), masked to liquid coarse cells. */
static void restrict_residual(const orc_level* f, orc_level* c) {
  memset(c->r, 0, c->n * sizeof(double));
  int k2 = (c->s[0] == 2) + (c->s[1] == 2) + (c->s[2] == 2);
  double scale = 1.0 / (double)(1 << k2);
  for (int i = 0; i < f->d[0]; ++i) {
    int xi[2]; double xw[2]; int nxt = taps1d(i, c->s[0], c->d[0], xi, xw);
    for (int j = 0; j < f->d[1]; ++j) {
      int yi[2]; double yw[2]; int nyt = taps1d(j, c->s[1], c->d[1], yi, yw);
      for (int k = 0; k < f->d[2]; ++k) {
        size_t fc = ((size_t)i * f->d[1] + j) * f->d[2] + k;
        if (!(f->phi[fc] < 0.0)) continue;
        double rv = f->t[fc];
        int zi[2]; double zw[2]; int nzt = taps1d(k, c->s[2], c->d[2], zi, zw);
        for (int a = 0; a < nxt; ++a)
          for (int b = 0; b < nyt; ++b)
            for (int q = 0; q < nzt; ++q) {
              size_t cc = ((size_t)xi[a] * c->d[1] + yi[b]) * c->d[2] + zi[q];
              c->r[cc] += scale * xw[a] * yw[b] * zw[q] * rv;
            }
      }
    }
  }
  for (size_t cc = 0; cc < c->n; ++cc)
    if (!(c->phi[cc] < 0.0)) c->r[cc] = 0.0;
}

static void coarse_solve(orc_mg* mg) {
  orc_level* c = &mg->lv[mg->L - 1];
  int nc = mg->nc;
  memset(c->e, 0, c->n * sizeof(double));
  if (nc == 0) return;
  double* y = (double*)malloc((size_t)nc * sizeof(double));
  for (int r = 0; r < nc; ++r) {
    double s = c->r[mg->cmap[r]];
    for (int k = 0; k < r; ++k) s -= mg->chol[r * nc + k] * y[k];
    y[r] = s / mg->chol[r * nc + r];
  }
  for (int r = nc - 1; r >= 0; --r) {
    double s = y[r];
    for (int k = r + 1; k < nc; ++k) s -= mg->chol[k * nc + r] * y[k];
    y[r] = s / mg->chol[r * nc + r];
  }
  for (int r = 0; r < nc; ++r) c->e[mg->cmap[r]] = y[r];
  free(y);
}

static void vcycle(orc_mg* mg, int l) {
  if (l == mg->L - 1) { coarse_solve(mg); return; }
  orc_level* f = &mg->lv[l];
  orc_level* c = &mg->lv[l + 1];
  memset(f->e, 0, f->n * sizeof(double));
  for (int s = 0; s < mg->opt.pre_sweeps; ++s) { gs_colour(f, 0); gs_colour(f, 1); }
  orc_apply_d(f->d[0], f->d[1], f->d[2], f->phi, f->wx, f->wy, f->wz, f->e, f->t);
  for (size_t i = 0; i < f->n; ++i) f->t[i] = (f->phi[i] < 0.0) ? f->r[i] - f->t[i] : 0.0;
  restrict_residual(f, c);
  vcycle(mg, l + 1);
  prolong_add(f, c);
  for (int s = 0; s < mg->opt.post_sweeps; ++s) { gs_colour(f, 1); gs_colour(f, 0); }
}

static void mg_apply(orc_mg* mg, const double* r, double* z) {
  orc_level* f = &mg->lv[0];
  memcpy(f->r, r, f->n * sizeof(double));
  for (size_t i = 0; i < f->n; ++i) if (!(f->phi[i] < 0.0)) f->r[i] = 0.0;
  vcycle(mg, 0);
  memcpy(z, f->e, f->n * sizeof(double));
}

int orc_vcycle_apply(int nx, int ny, int nz, const float* phi, const orc_mg_options* opt, const double* r,
                     double* z) {
  orc_mg mg;
  if (mg_build(&mg, nx, ny, nz, phi, opt) != 0) return -1;
  mg_apply(&mg, r, z);
  mg_free(&mg);
  return 0;
}

int orc_level_phi(int nx, int ny, int nz, const float* phi, int preserve, int level, double* out) {
  orc_mg mg;
  orc_mg_options opt = {1, 1, preserve};
  if (mg_build(&mg, nx, ny, nz, phi, &opt) != 0) return -1;
  if (level < 0 || level >= mg.L) { mg_free(&mg); return -1; }
  memcpy(out, mg.lv[level].phi, mg.lv[level].n * sizeof(double));
  mg_free(&mg);
  return 0;
}

/* MGPCG: PCG with the V-cycle as M^-1 (CG control flow of
 * webgpu-octree-pipelined-mgpcg.ts:973-1166; FP64 throughout). */
int orc_mgpcg(int nx, int ny, int nz, const float* phi, const double* b_in, const orc_mg_options* opt,
              double rel_tol, int max_iter, double* p, orc_solve_report* rep) {
  size_t n = (size_t)nx * ny * nz;
  orc_mg mg;
  if (mg_build(&mg, nx, ny, nz, phi, opt) != 0) return -1;
  orc_level* f = &mg.lv[0];
  nullspace_t ns;
  float* phif = (float*)phi;
  nullspace_init(&ns, nx, ny, nz, phif, f->wx, f->wy, f->wz);
  double* b = (double*)malloc(n * sizeof(double));
  double* r = (double*)malloc(n * sizeof(double));
  double* z = (double*)malloc(n * sizeof(double));
  double* d = (double*)malloc(n * sizeof(double));
  double* ad = (double*)malloc(n * sizeof(double));
  memcpy(b, b_in, n * sizeof(double));
  nullspace_project(&ns, b, n);
  double bnorm = norm2(b, n);
  nullspace_project(&ns, p, n);
  orc_apply_d(nx, ny, nz, f->phi, f->wx, f->wy, f->wz, p, ad);
  for (size_t i = 0; i < n; ++i) r[i] = b[i] - ad[i];
  nullspace_project(&ns, r, n);
  double target = rel_tol * bnorm;
  double rnorm = norm2(r, n);
  int it = 0, status = 0;
  double rz = 0.0;
  while (it < max_iter && rnorm > target) {
    mg_apply(&mg, r, z);
    nullspace_project(&ns, z, n);
    double rz_next = orc_dot_neumaier(r, z, n);
    if (!(rz_next > 0) || !isfinite(rz_next)) { status = 3; break; }
    if (it == 0) memcpy(d, z, n * sizeof(double));
    else {
      double beta = rz_next / rz;
      for (size_t i = 0; i < n; ++i) d[i] = z[i] + beta * d[i];
    }
    rz = rz_next;
    orc_apply_d(nx, ny, nz, f->phi, f->wx, f->wy, f->wz, d, ad);
    nullspace_project(&ns, ad, n);
    double curv = orc_dot_neumaier(d, ad, n);
    if (!(curv > 0) || !isfinite(curv)) { status = 4; break; }
    double alpha = rz / curv;
    for (size_t i = 0; i < n; ++i) { p[i] += alpha * d[i]; r[i] -= alpha * ad[i]; }
    ++it;
    rnorm = norm2(r, n);
  }
  orc_apply_d(nx, ny, nz, f->phi, f->wx, f->wy, f->wz, p, ad);
  for (size_t i = 0; i < n; ++i) r[i] = b[i] - ad[i];
  nullspace_project(&ns, r, n);
  double tn = norm2(r, n), mx = 0.0;
  for (size_t i = 0; i < n; ++i) if (fabs(r[i]) > mx) mx = fabs(r[i]);
  if (rep) {
    rep->iterations = it;
    rep->converged = tn <= target;
    rep->components = ns.count;
    rep->unanchored = nullspace_unanchored(&ns);
    rep->rhs_norm = bnorm;
    rep->relative_residual = bnorm > 0 ? tn / bnorm : 0.0;
    rep->max_residual = mx;
  }
  if (status == 0 && !(tn <= target)) status = 5;
  free(b); free(r); free(z); free(d); free(ad);
  nullspace_free(&ns);
  mg_free(&mg);
  return status;
}

/* ------------------------------------------------------------------------ */
/* Generic dense PCG: solveOctreePipelinedPCG, octree-pipelined-pcg.ts:305-594 */
/* ------------------------------------------------------------------------ */
static void dense_mv(int n, const double* A, const double* x, double* y) {
  for (int i = 0; i < n; ++i) {
    double s = 0.0;
    for (int j = 0; j < n; ++j) s += A[i * n + j] * x[j];
    y[i] = s;
  }
}

static int close_rel(double a, double b, double tol) {
  double m = 1.0;
  if (fabs(a) > m) m = fabs(a);
  if (fabs(b) > m) m = fabs(b);
  return fabs(a - b) <= tol * m;
}

/* auditLinearSPDMap (octree-pipelined-pcg.ts:184-230) for a matrix map. */
static int audit_map(int n, const double* M, double tol) {
  double *p1 = malloc(n * sizeof(double)), *p2 = malloc(n * sizeof(double));
  double *i1 = malloc(n * sizeof(double)), *i2 = malloc(n * sizeof(double));
  double *cmb = malloc(n * sizeof(double)), *ic = malloc(n * sizeof(double));
  for (int i = 0; i < n; ++i) {
    p1[i] = (double)((i * 17 + 5) % 23) - 11.0;
    p2[i] = (double)((i * 29 + 3) % 19) - 9.0;
    cmb[i] = p1[i] - 0.375 * p2[i];
  }
  dense_mv(n, M, p1, i1); dense_mv(n, M, p2, i2); dense_mv(n, M, cmb, ic);
  int fail = 0;
  for (int i = 0; i < n && !fail; ++i)
    if (!isfinite(i1[i]) || !isfinite(i2[i])) fail = 7;
  for (int i = 0; i < n && !fail; ++i)
    if (!close_rel(ic[i], i1[i] - 0.375 * i2[i], tol)) fail = 1;   /* not linear */
  if (!fail && !close_rel(orc_dot_neumaier(p1, i2, n), orc_dot_neumaier(p2, i1, n), tol)) fail = 2;  /* not symmetric */
  if (!fail && (!(orc_dot_neumaier(p1, i1, n) > 0) || !(orc_dot_neumaier(p2, i2, n) > 0))) fail = 3;  /* not positive */
  free(p1); free(p2); free(i1); free(i2); free(cmb); free(ic);
  return fail;
}

int orc_dense_pcg(int n, const double* A, const double* b, const double* x0, int use_jacobi, int max_iter,
                  double rel_tol, double abs_tol, int validate, double* x, int* iterations, int* reductions,
                  double* residual_norm) {
  double* Minv = (double*)calloc((size_t)n * n, sizeof(double));
  for (int i = 0; i < n; ++i) Minv[i * n + i] = use_jacobi ? 1.0 / A[i * n + i] : 1.0;
  *iterations = 0;
  *reductions = 0;
  if (validate) {
    int a = audit_map(n, A, 2e-12);
    if (a) { free(Minv); for (int i = 0; i < n; ++i) x[i] = x0 ? x0[i] : 0.0; return a == 7 ? 7 : 10 + a; }
    a = audit_map(n, Minv, 2e-12);
    if (a) { free(Minv); for (int i = 0; i < n; ++i) x[i] = x0 ? x0[i] : 0.0; return a == 7 ? 7 : 20 + a; }
  }
  double *r = malloc(n * sizeof(double)), *z = malloc(n * sizeof(double)), *az = malloc(n * sizeof(double));
  double *d = malloc(n * sizeof(double)), *ad = malloc(n * sizeof(double)), *tmp = malloc(n * sizeof(double));
  for (int i = 0; i < n; ++i) x[i] = x0 ? x0[i] : 0.0;
  dense_mv(n, A, x, tmp);
  for (int i = 0; i < n; ++i) r[i] = b[i] - tmp[i];
  double bn = sqrt(fmax(0, orc_dot_neumaier(b, b, n)));
  double stop = fmax(abs_tol, rel_tol * fmax(bn, 4.9406564584124654e-324));
  dense_mv(n, Minv, r, z);
  dense_mv(n, A, z, az);
  /* mergeReduction: gamma = r.z, delta = z.Az, rr = r.r */
  double gamma = orc_dot_neumaier(r, z, n), delta = orc_dot_neumaier(z, az, n);
  double rr = orc_dot_neumaier(r, r, n);
  int red = 1, it = 0, status = 0;
  double rec = sqrt(fmax(0, rr));
  if (rec <= stop) { *reductions = red; *residual_norm = rec; goto done; }
  if (!(gamma > 0) || !isfinite(gamma)) { status = 3; goto done; }
  if (!(delta > 0) || !isfinite(delta)) { status = 4; goto done; }
  double alpha = gamma / delta;
  memcpy(d, z, n * sizeof(double));
  memcpy(ad, az, n * sizeof(double));
  while (it < max_iter) {
    for (int i = 0; i < n; ++i) { x[i] += alpha * d[i]; r[i] -= alpha * ad[i]; }
    ++it;
    dense_mv(n, Minv, r, z);
    double g2 = orc_dot_neumaier(r, z, n);
    double rr2 = orc_dot_neumaier(r, r, n);
    ++red;
    rec = sqrt(fmax(0, rr2));
    if (rec <= stop) {
      dense_mv(n, A, x, tmp);
      for (int i = 0; i < n; ++i) tmp[i] = b[i] - tmp[i];
      double vn = sqrt(fmax(0, orc_dot_neumaier(tmp, tmp, n)));
      *residual_norm = vn;
      if (vn > stop) status = 6;
      goto done;
    }
    if (!(g2 > 0) || !isfinite(g2)) { status = 3; goto done; }
    if (it >= max_iter) { status = 5; break; }
    double beta = g2 / gamma;
    for (int i = 0; i < n; ++i) d[i] = z[i] + beta * d[i];
    dense_mv(n, A, d, ad);
    double den = orc_dot_neumaier(d, ad, n);
    ++red;
    if (!(den > 0) || !isfinite(den)) { status = 4; goto done; }
    alpha = g2 / den;
    gamma = g2;
  }
  if (status == 0) status = 5;
done:
  if (status != 0 && status != 6) {
    dense_mv(n, A, x, tmp);
    for (int i = 0; i < n; ++i) tmp[i] = b[i] - tmp[i];
    *residual_norm = sqrt(fmax(0, orc_dot_neumaier(tmp, tmp, n)));
  }
  *iterations = it;
  *reductions = red;
  free(r); free(z); free(az); free(d); free(ad); free(tmp); free(Minv);
  return status;
}

/* ------------------------------------------------------------------------ */
/* Restricted tall-cell layout                                              */
/* ------------------------------------------------------------------------ */
/* chooseTallCellBase, lib/core/tall-cell-grid.ts:61-82 (JS Math.round is
 * round-half-up, reproduced with floor(x + 0.5)). */
int orc_tall_choose_base(int lowest, int highest, int fine_ny, int regular_layers, int liquid_halo,
                         int air_halo) {
  int maximum_base = fine_ny - regular_layers;
  if (maximum_base < 0) maximum_base = 0;
  int lower = highest + 1 + air_halo - regular_layers;
  int upper = lowest + 1 - liquid_halo;
  int desired = lower <= upper ? (int)floor((lower + upper) / 2.0 + 0.5) : lower;
  int clamped = desired;
  if (clamped > maximum_base) clamped = maximum_base;
  if (clamped < 0) clamped = 0;
  return maximum_base >= 2 ? (clamped < 2 ? 2 : clamped) : 0;
}

/* limitNeighboringTallCellBases, lib/core/tall-cell-grid.ts:84-106 (Jacobi
 * passes of base <= min(neighbour) + delta). */
void orc_tall_limit_bases(int32_t* bases, int nx, int nz, int max_delta, int passes) {
  size_t n = (size_t)nx * nz;
  int32_t* next = (int32_t*)malloc(n * sizeof(int32_t));
  for (int pass = 0; pass < passes; ++pass) {
    for (int z = 0; z < nz; ++z)
      for (int x = 0; x < nx; ++x) {
        size_t idx = (size_t)x * nz + z; /* columns stored [x][z], z fastest */
        int32_t upper = bases[idx];
        if (x > 0 && bases[idx - nz] + max_delta < upper) upper = bases[idx - nz] + max_delta;
        if (x + 1 < nx && bases[idx + nz] + max_delta < upper) upper = bases[idx + nz] + max_delta;
        if (z > 0 && bases[idx - 1] + max_delta < upper) upper = bases[idx - 1] + max_delta;
        if (z + 1 < nz && bases[idx + 1] + max_delta < upper) upper = bases[idx + 1] + max_delta;
        next[idx] = upper;
      }
    memcpy(bases, next, n * sizeof(int32_t));
  }
  free(next);
}

/* ------------------------------------------------------------------------ */
/* Restricted tall-cell projection                                          */
/* ------------------------------------------------------------------------ */
/* The tall-cell pressure space is p_fine = E p.  Along a column (x, z) of
 * base b the fine rows 0..b-1 interpolate linearly between the bottom sample
 * (row 0) and the top sample (row b-1) -- the represented-field rule of
 * representedInitialPhiVolumeCells, tall-cell-grid.ts:206-249, applied to phi
 * and p alike -- rows b..b+B-1 are the cubic band samples, and rows above the
 * band are air (phi = +5h, the narrow-band clamp of buildInitialPhi,
 * tall-cell-grid.ts:171-190; p = 0).
 *
 * Face velocities live in slots.  v: slot 0 is the one velocity of the b-1
 * faces inside a tall cell, slot 1 the face tall top -> band row 0, slot 1+k
 * the face band k-1 -> k, slot B+1 the face band top -> air (the lid when the
 * band reaches it).  u and w: slots 0 and 1 are the mean and the slope (per
 * fine row, about row (m-1)/2) of an affine velocity over rows 0..m-1, m the
 * smaller base of the two columns; slot 2+j is the single fine row m+j.
 *
 * The operator, right-hand side and gradient subtract are the variational
 * reduction of the uniform ghost-fluid projection on the expanded fine grid
 * (docs/papers/sg2025narita.md:120-172; the restricted form of
 * tallCells.txt Eq. 13-18): A = E^T A_u E, b = E^T b_u(expanded u), and the
 * subtracted fine velocities are projected back onto the slots by least
 * squares (identity for single-row slots, mean for v slot 0, mean and slope
 * for u/w slots 0-1).  Because each pressure coefficient c(y) is affine in y
 * over a slot's rows, E^T D of the projected field equals E^T D of the fine
 * field, so the reduced divergence vanishes exactly where the fine one does.
 * This restatement walks every fine face; the device kernels evaluate the
 * same sums in closed form from per-face weight moments. */
#define TALL_P(t) ((t)->band + 2)
#define TALL_S(t) ((t)->band + 2 + (t)->max_delta)

enum { TALL_ROLE_ONE = 0, TALL_ROLE_MEAN = 1, TALL_ROLE_AFFINE = 2 };

typedef struct {
  int n;           /* g = sum_q c[q] p[idx[q]] (p of air samples = 0) */
  size_t idx[4];
  double c[4];
  double w;        /* ghost-fluid weight of the fine face */
  int array;       /* 0 u, 1 v, 2 w */
  int role;
  size_t slot, slot1; /* AFFINE: mean slot, slope slot */
  double t;        /* AFFINE: row - (m-1)/2 */
  double count;    /* MEAN: rows sharing the slot; AFFINE: m */
  double q;        /* AFFINE: sum of t^2 over the m rows */
} tall_face;

static size_t tall_sample(const orc_tall* t, int x, int py, int z) {
  return ((size_t)x * TALL_P(t) + py) * t->nz + z;
}

static int tall_base(const orc_tall* t, int x, int z) { return t->bases[(size_t)x * t->nz + z]; }

/* phi of fine cell (x, y, z) and p there as sum c[q] p[idx[q]]. */
static double tall_cell(const orc_tall* t, int x, int z, int y, int* nd, size_t* idx, double* c) {
  const int b = tall_base(t, x, z);
  *nd = 0;
  if (y < b) {
    const double k = y / (double)(b - 1);
    const double pb = t->phi[tall_sample(t, x, 0, z)], pt = t->phi[tall_sample(t, x, 1, z)];
    if (y == 0 || y == b - 1) {
      idx[0] = tall_sample(t, x, y == 0 ? 0 : 1, z);
      c[0] = 1.0;
      *nd = 1;
      return y == 0 ? pb : pt;
    }
    idx[0] = tall_sample(t, x, 0, z);
    c[0] = 1.0 - k;
    idx[1] = tall_sample(t, x, 1, z);
    c[1] = k;
    *nd = 2;
    return pb + (pt - pb) * k;
  }
  if (y - b < t->band && y < t->ny) {
    idx[0] = tall_sample(t, x, 2 + y - b, z);
    c[0] = 1.0;
    *nd = 1;
    return t->phi[idx[0]];
  }
  return 5.0 * t->h;
}

typedef void (*tall_face_cb)(void* ctx, const tall_face* f);

/* The fine face between cells a (-) and b (+); walls (n = 0) are skipped. */
static void tall_emit(const orc_tall* t, tall_face* f, int xa, int za, int xb, int zb, int ya, int yb,
                      tall_face_cb cb, void* ctx) {
  int na, nb;
  size_t ia[2], ib[2];
  double ca[2], cbv[2];
  const double phia = tall_cell(t, xa, za, ya, &na, ia, ca);
  const double phib = tall_cell(t, xb, zb, yb, &nb, ib, cbv);
  f->n = 0;
  for (int q = 0; q < nb; ++q) { f->idx[f->n] = ib[q]; f->c[f->n] = cbv[q]; ++f->n; }
  for (int q = 0; q < na; ++q) {
    int merged = 0; /* a face inside one tall cell names its samples twice */
    for (int r = 0; r < f->n; ++r)
      if (f->idx[r] == ia[q]) { f->c[r] -= ca[q]; merged = 1; }
    if (!merged) { f->idx[f->n] = ia[q]; f->c[f->n] = -ca[q]; ++f->n; }
  }
  f->w = face_weight(phia, phib, 1.0);
  cb(ctx, f);
}

/* Every fine face of the domain interior (walls carry no flux) once. */
static int tall_faces(const orc_tall* t, tall_face_cb cb, void* ctx) {
  const int nx = t->nx, ny = t->ny, nz = t->nz, B = t->band, P = TALL_P(t), S = TALL_S(t);
  for (size_t i = 0; i < (size_t)nx * nz; ++i)
    if (t->bases[i] < 2 || t->bases[i] > ny) return ORC_TALL_BAD_BASE;
  for (int x = 0; x < nx; ++x)
    for (int z = 0; z < nz; ++z) {
      const int b = tall_base(t, x, z);
      tall_face f;
      memset(&f, 0, sizeof f);
      /* vertical faces y -> y+1 below the lid and below the band top + 1 */
      f.array = 1;
      for (int y = 0; y + 1 < ny && y < b + B; ++y) {
        if (y < b - 1) {
          f.role = TALL_ROLE_MEAN;
          f.slot = ((size_t)x * P + 0) * nz + z;
          f.count = b - 1;
        } else {
          f.role = TALL_ROLE_ONE;
          f.slot = ((size_t)x * P + (y - b + 2)) * nz + z;
        }
        tall_emit(t, &f, x, z, x, z, y, y + 1, cb, ctx);
      }
      /* horizontal faces to the +x and +z neighbours */
      for (int dir = 0; dir < 2; ++dir) {
        const int xn = x + (dir == 0), zn = z + (dir == 1);
        if (xn >= nx || zn >= nz) continue;
        const int bn = tall_base(t, xn, zn);
        if (abs(bn - b) > t->max_delta) return ORC_TALL_BAD_BASE;
        const int m = b < bn ? b : bn, M = b > bn ? b : bn;
        const size_t slot0 = dir == 0 ? ((size_t)(x + 1) * S) * nz + z : ((size_t)x * S) * (nz + 1) + (z + 1);
        const size_t sstride = dir == 0 ? (size_t)nz : (size_t)(nz + 1);
        int ytop = M + B;
        if (ytop > ny) ytop = ny;
        f.array = dir == 0 ? 0 : 2;
        for (int y = 0; y < ytop; ++y) {
          if (y < m) {
            f.role = TALL_ROLE_AFFINE;
            f.slot = slot0;
            f.slot1 = slot0 + sstride;
            f.t = y - 0.5 * (m - 1);
            f.count = m;
            f.q = m * ((double)m * m - 1.0) / 12.0;
          } else {
            f.role = TALL_ROLE_ONE;
            f.slot = slot0 + (size_t)(2 + y - m) * sstride;
          }
          tall_emit(t, &f, x, z, xn, zn, y, y, cb, ctx);
        }
      }
    }
  return 0;
}

static double tall_face_velocity(const tall_face* f, const double* U) {
  if (f->role == TALL_ROLE_AFFINE) return U[f->slot] + f->t * U[f->slot1];
  return U[f->slot];
}

int orc_tall_check(const orc_tall* t) {
  for (int x = 0; x < t->nx; ++x)
    for (int z = 0; z < t->nz; ++z) {
      const int b = tall_base(t, x, z);
      if (b < 2 || b > t->ny) return ORC_TALL_BAD_BASE;
      if (x + 1 < t->nx && abs(tall_base(t, x + 1, z) - b) > t->max_delta) return ORC_TALL_BAD_BASE;
      if (z + 1 < t->nz && abs(tall_base(t, x, z + 1) - b) > t->max_delta) return ORC_TALL_BAD_BASE;
    }
  return 0;
}

/* Unknown samples: liquid, inside the domain. */
void orc_tall_unknowns(const orc_tall* t, uint8_t* mask) {
  const int P = TALL_P(t);
  for (int x = 0; x < t->nx; ++x)
    for (int z = 0; z < t->nz; ++z) {
      const int b = tall_base(t, x, z);
      for (int py = 0; py < P; ++py) {
        const size_t i = tall_sample(t, x, py, z);
        const int active = py < 2 ? 1 : (b + py - 2 < t->ny);
        mask[i] = active && t->phi[i] < 0.0f;
      }
    }
}

static size_t tall_samples(const orc_tall* t) { return (size_t)t->nx * TALL_P(t) * t->nz; }

typedef struct {
  const double* p;
  double* out;
  const uint8_t* mask;
} tall_apply_ctx;

static double tall_grad(const tall_face* f, const double* p, const uint8_t* mask) {
  double g = 0.0;
  for (int q = 0; q < f->n; ++q)
    if (mask[f->idx[q]]) g += f->c[q] * p[f->idx[q]];
  return g;
}

static void tall_apply_cb(void* vctx, const tall_face* f) {
  tall_apply_ctx* c = (tall_apply_ctx*)vctx;
  if (f->w == 0.0) return;
  const double s = f->w * tall_grad(f, c->p, c->mask);
  for (int q = 0; q < f->n; ++q)
    if (c->mask[f->idx[q]]) c->out[f->idx[q]] += s * f->c[q];
}

int orc_tall_apply(const orc_tall* t, const double* p, double* out) {
  const size_t n = tall_samples(t);
  uint8_t* mask = (uint8_t*)malloc(n);
  orc_tall_unknowns(t, mask);
  memset(out, 0, n * sizeof(double));
  tall_apply_ctx c = {p, out, mask};
  const int rc = tall_faces(t, tall_apply_cb, &c);
  free(mask);
  return rc;
}

typedef struct {
  const double* U[3];
  double* b;
  const uint8_t* mask;
  double s;
} tall_rhs_ctx;

static void tall_rhs_cb(void* vctx, const tall_face* f) {
  tall_rhs_ctx* c = (tall_rhs_ctx*)vctx;
  const double u = tall_face_velocity(f, c->U[f->array]);
  for (int q = 0; q < f->n; ++q)
    if (c->mask[f->idx[q]]) c->b[f->idx[q]] += c->s * f->c[q] * u;
}

/* b = (rho h / dt) E^T (-div u_fine) on unknown samples; the uniform rhs
 * (orc_rhs) of the expanded velocity, gathered through E^T. */
int orc_tall_rhs(const orc_tall* t, const double* u, const double* v, const double* w, double dt, double rho,
                 double* b) {
  const size_t n = tall_samples(t);
  uint8_t* mask = (uint8_t*)malloc(n);
  orc_tall_unknowns(t, mask);
  memset(b, 0, n * sizeof(double));
  tall_rhs_ctx c = {{u, v, w}, b, mask, rho * t->h / dt};
  const int rc = tall_faces(t, tall_rhs_cb, &c);
  free(mask);
  return rc;
}

typedef struct {
  const double* p;
  const double* U[3];
  double* O[3];
  uint8_t* live[3];
  const uint8_t* mask;
  double s;
} tall_sub_ctx;

static void tall_sub_cb(void* vctx, const tall_face* f) {
  tall_sub_ctx* c = (tall_sub_ctx*)vctx;
  double u = tall_face_velocity(f, c->U[f->array]);
  int live = f->w != 0.0;
  if (live) u -= c->s * f->w * tall_grad(f, c->p, c->mask);
  for (int q = 0; q < f->n; ++q) live |= c->mask[f->idx[q]] && f->c[q] != 0.0;
  double* o = c->O[f->array];
  if (live) c->live[f->array][f->slot] = 1;
  switch (f->role) {
    case TALL_ROLE_ONE: o[f->slot] = u; break;
    case TALL_ROLE_MEAN: o[f->slot] += u / f->count; break;
    default:
      o[f->slot] += u / f->count;
      o[f->slot1] += u * f->t / f->q;
      if (live) c->live[f->array][f->slot1] = 1;
  }
}

/* Fine velocities u - dt/(rho h) w grad p, projected back onto the slots.  A
 * slot none of whose faces has a liquid side or an unknown sample in its
 * expansion is set to 0 (the faces orc_project zeroes); unused slots are 0.
 * Faces without a liquid side keep their velocity inside a live slot, so the
 * reduced divergence of the result is exactly b - A p. */
int orc_tall_subtract(const orc_tall* t, const double* p, const double* u, const double* v, const double* w,
                      double dt, double rho, double* uo, double* vo, double* wo) {
  const size_t n = tall_samples(t);
  const size_t nu = (size_t)(t->nx + 1) * TALL_S(t) * t->nz, nw = (size_t)t->nx * TALL_S(t) * (t->nz + 1);
  memset(uo, 0, nu * sizeof(double));
  memset(vo, 0, n * sizeof(double));
  memset(wo, 0, nw * sizeof(double));
  uint8_t* mask = (uint8_t*)malloc(n);
  uint8_t* lu = (uint8_t*)calloc(nu, 1);
  uint8_t* lv = (uint8_t*)calloc(n, 1);
  uint8_t* lw = (uint8_t*)calloc(nw, 1);
  orc_tall_unknowns(t, mask);
  tall_sub_ctx c = {p, {u, v, w}, {uo, vo, wo}, {lu, lv, lw}, mask, dt / (rho * t->h)};
  const int rc = tall_faces(t, tall_sub_cb, &c);
  for (size_t i = 0; i < nu; ++i) if (!lu[i]) uo[i] = 0.0;
  for (size_t i = 0; i < n; ++i) if (!lv[i]) vo[i] = 0.0;
  for (size_t i = 0; i < nw; ++i) if (!lw[i]) wo[i] = 0.0;
  free(mask); free(lu); free(lv); free(lw);
  return rc;
}

typedef struct {
  double* d;
  const uint8_t* mask;
} tall_diag_ctx;

static void tall_diag_cb(void* vctx, const tall_face* f) {
  tall_diag_ctx* c = (tall_diag_ctx*)vctx;
  for (int q = 0; q < f->n; ++q)
    if (c->mask[f->idx[q]]) c->d[f->idx[q]] += f->w * f->c[q] * f->c[q];
}

int orc_tall_diagonal(const orc_tall* t, double* d) {
  const size_t n = tall_samples(t);
  uint8_t* mask = (uint8_t*)malloc(n);
  orc_tall_unknowns(t, mask);
  memset(d, 0, n * sizeof(double));
  tall_diag_ctx c = {d, mask};
  const int rc = tall_faces(t, tall_diag_cb, &c);
  free(mask);
  return rc;
}

/* FP64 Jacobi-PCG on the unknown samples, from p = 0. */
int orc_tall_solve(const orc_tall* t, const double* b, double rel_tol, int max_iter, double* p, int* iterations,
                   double* relative_residual) {
  const size_t n = tall_samples(t);
  uint8_t* mask = (uint8_t*)malloc(n);
  orc_tall_unknowns(t, mask);
  double* d = (double*)calloc(n, sizeof(double));
  double *r = (double*)calloc(n, sizeof(double)), *z = (double*)calloc(n, sizeof(double));
  double *s = (double*)calloc(n, sizeof(double)), *q = (double*)calloc(n, sizeof(double));
  int rc = orc_tall_diagonal(t, d);
  memset(p, 0, n * sizeof(double));
  for (size_t i = 0; i < n; ++i) r[i] = mask[i] ? b[i] : 0.0;
  const double bb = orc_dot_neumaier(r, r, n);
  int it = 0;
  if (rc == 0 && bb > 0.0) {
    for (size_t i = 0; i < n; ++i) {
      z[i] = (mask[i] && d[i] > 0.0) ? r[i] / d[i] : 0.0;
      s[i] = z[i];
    }
    double rz = orc_dot_neumaier(r, z, n);
    while (it < max_iter) {
      orc_tall_apply(t, s, q);
      const double sq = orc_dot_neumaier(s, q, n);
      if (!(sq > 0.0)) { rc = 4; break; }
      const double a = rz / sq;
      for (size_t i = 0; i < n; ++i) { p[i] += a * s[i]; r[i] -= a * q[i]; }
      ++it;
      if (orc_dot_neumaier(r, r, n) <= rel_tol * rel_tol * bb) break;
      for (size_t i = 0; i < n; ++i) z[i] = (mask[i] && d[i] > 0.0) ? r[i] / d[i] : 0.0;
      const double rz2 = orc_dot_neumaier(r, z, n);
      const double beta = rz2 / rz;
      rz = rz2;
      for (size_t i = 0; i < n; ++i) s[i] = z[i] + beta * s[i];
    }
  }
  double rel = 0.0;
  if (rc == 0 && bb > 0.0) {
    orc_tall_apply(t, p, q);
    double rr = 0.0;
    for (size_t i = 0; i < n; ++i)
      if (mask[i]) rr += (b[i] - q[i]) * (b[i] - q[i]);
    rel = sqrt(rr / bb);
  }
  if (iterations) *iterations = it;
  if (relative_residual) *relative_residual = rel;
  free(mask); free(d); free(r); free(z); free(s); free(q);
  return rc;
}

/* max over unknown samples of |E^T div u_fine| / (h * support), support the
 * fine rows of the sample's hat function (b/2 for a tall end, 1 in the band):
 * a per-volume divergence in 1/s comparable with the uniform diagnostic. */
double orc_tall_max_divergence(const orc_tall* t, const double* u, const double* v, const double* w) {
  const size_t n = tall_samples(t);
  double* dv = (double*)calloc(n, sizeof(double));
  if (orc_tall_rhs(t, u, v, w, 1.0, 1.0, dv) != 0) { free(dv); return -1.0; }
  uint8_t* mask = (uint8_t*)malloc(n);
  orc_tall_unknowns(t, mask);
  double m = 0.0;
  for (int x = 0; x < t->nx; ++x)
    for (int z = 0; z < t->nz; ++z) {
      const int b = tall_base(t, x, z);
      for (int py = 0; py < TALL_P(t); ++py) {
        const size_t i = tall_sample(t, x, py, z);
        if (!mask[i]) continue;
        const double support = py < 2 ? 0.5 * b : 1.0;
        const double d = fabs(dv[i]) / (t->h * t->h) / support;
        if (d > m) m = d;
      }
    }
  free(dv);
  free(mask);
  return m;
}

typedef struct {
  int64_t* ri;
  int64_t* ci;
  double* v;
  size_t n, cap;
  const uint8_t* mask;
} tall_coo_ctx;

static void tall_coo_cb(void* vctx, const tall_face* f) {
  tall_coo_ctx* c = (tall_coo_ctx*)vctx;
  if (f->w == 0.0) return;
  for (int a = 0; a < f->n; ++a) {
    if (!c->mask[f->idx[a]]) continue;
    for (int b = 0; b < f->n; ++b) {
      if (!c->mask[f->idx[b]]) continue;
      if (c->n < c->cap) {
        c->ri[c->n] = (int64_t)f->idx[a];
        c->ci[c->n] = (int64_t)f->idx[b];
        c->v[c->n] = f->w * f->c[a] * f->c[b];
      }
      ++c->n;
    }
  }
}

/* The entries w c_a c_b of every fine face (duplicates to be summed): the
 * operator orc_tall_apply applies, as coordinates.  Returns the entry count;
 * only the first `cap` are written. */
int64_t orc_tall_coo(const orc_tall* t, int64_t* rows, int64_t* cols, double* vals, int64_t cap) {
  const size_t n = tall_samples(t);
  uint8_t* mask = (uint8_t*)malloc(n);
  orc_tall_unknowns(t, mask);
  tall_coo_ctx c = {rows, cols, vals, 0, (size_t)cap, mask};
  const int rc = tall_faces(t, tall_coo_cb, &c);
  free(mask);
  return rc ? -(int64_t)rc : (int64_t)c.n;
}

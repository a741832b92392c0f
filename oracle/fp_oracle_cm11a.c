/*
 * fp_oracle_cm11a.c -- CPU restatement of the reference's CM11a separating-wall
 * pressure solve (test infrastructure; never linked into the product).
 *
 * The reference solves the uniform grid's pressure as a linear complementarity
 * problem on a hierarchy of halo-padded levels (Chentanez & Mueller 2011,
 * "A multigrid fluid pressure solver handling separating solid boundary
 * conditions"):
 *   lib/methods/uniform/webgpu-uniform-pressure-multigrid.wgsl.ts
 *     :49-59   D4-ordered sums (mgD4Sum6 / mgD4Sum8)
 *     :71-76   coarse -> fine child map (mgFineChild)
 *     :93-115  face coefficients V/(h^2 theta) with the open-top exterior (mgCoefficientRaw)
 *     :116-133 baked coefficient lookup and A p (mgCoefficient, mgApply)
 *     :135-150 finest topology: cell/face dual volumes, halo phi = h/2 (mgBuildFinestTopology)
 *     :152-167 finest rhs b = -rho (div - corr) / dt, p = 0, p_min (mgBuildFinestRhs)
 *     :169-195 coarse topology and level set, wall dual volumes 1/2 (mgDownsampleTopology)
 *     :197-205 residual on pressure unknowns (mgResidual)
 *     :210-215 restriction (mgRestrictResidual)
 *     :217-253 trilinear prolongation without halo taps (mgTrilinearPressure, Add, Assign)
 *     :255-286 pressure copies and p_min shifts
 *     :288-300 one-cell phi continuation into solid cells (mgExtrapolatePhiOneCell)
 *     :311-323 coarse p_min (mgDownsampleSubtract, mgDownsampleMinimum)
 *     :325-344 projected red-black Gauss-Seidel (mgSmoothColour)
 *     :346-471 double-single coarsest LCP solve (mgTwoSum .. mgSolveCoarsest)
 *     :473-481 projected residual (mgMeasureFineResidual)
 *   lib/methods/uniform/webgpu-uniform-pressure-multigrid.ts
 *     :5-19    schedule constants (3 full cycles, 4 V-cycles, 6 + 6 sweeps,
 *              3 constraint levels, C = 2, coarse tolerance 1e-4 1/s, cap 4096)
 *     :488-576 the schedule (vCycle, fullCycle)
 *   lib/methods/uniform/webgpu-uniform-reference.wgsl.ts
 *     :536-561 dual volume 1/2 at a grid-aligned box wall, 1 at the open top
 *     :1010-1027 CM11a divergence with dual volumes (divergenceAt)
 *     :1052-1099 velocity update with the halo pressure (project)
 *   lib/core/cm12-numerics.ts:37-43 ghost-fluid theta
 *
 * Scope: a box container (six grid-aligned walls, the +y side optionally
 * open to air), the level set given directly (liquid iff phi < 0), no
 * embedded solids.  Arithmetic is FP32 in the reference's operation order
 * (compiled with -ffp-contract=off), so the GPU build of the same algorithm
 * can be compared bit for bit.  Levels are stored z fastest with a one-cell
 * halo: index (x * Dy + y) * Dz + z over the padded dims D = n + 2.
 *
 * The restatement is pinned by what the mathematics fixes (tests/
 * test_oracle_cm11a.py): with no active constraint (pressure above p_min
 * everywhere) the halo rows copy their interior neighbour and the interior
 * equations are the Neumann-wall system of orc_mgpcg, so enough cycles reach
 * that solution; with separating walls the result satisfies the LCP's
 * complementarity conditions and matches a brute-force projected
 * Gauss-Seidel solve of the same assembled system on small grids.
 */
#include "fp_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define CM_NEG_BIG (-3.402823e38f) /* the reference's "no constraint" p_min */
#define CM_THETA_MIN 0.05f

typedef struct {
  int D[3]; /* padded dims */
  int n;
  float h[3];
  float *phi_raw, *phi, *topo, *coef; /* topo, coef: 4 per cell */
  float *p, *rhs0, *rhs1, *res, *pmin;
} cml;

static inline int cidx(const cml* L, int x, int y, int z) { return (x * L->D[1] + y) * L->D[2] + z; }
static inline int cvalid(const cml* L, int x, int y, int z) {
  return x >= 0 && y >= 0 && z >= 0 && x < L->D[0] && y < L->D[1] && z < L->D[2];
}
static inline int cinterior(const cml* L, int x, int y, int z) {
  return x >= 1 && y >= 1 && z >= 1 && x < L->D[0] - 1 && y < L->D[1] - 1 && z < L->D[2] - 1;
}
static const int NB[6][3] = {{-1, 0, 0}, {1, 0, 0}, {0, -1, 0}, {0, 1, 0}, {0, 0, -1}, {0, 0, 1}};

static inline float d4s6(const float* t) { return ((t[0] + t[1]) + (t[4] + t[5])) + (t[2] + t[3]); }
static inline float d4s8(const float* t) {
  const float y0 = (t[0] + t[5]) + (t[1] + t[4]);
  const float y1 = (t[2] + t[7]) + (t[3] + t[6]);
  return y0 + y1;
}
static inline float cm_theta(float l, float a, float eps) {
  const float r = fabsf(l) / fmaxf(fabsf(l) + fabsf(a), eps);
  return fminf(1.0f, fmaxf(CM_THETA_MIN, r));
}

/* ---- double-single arithmetic (…multigrid.wgsl.ts:364-381) ---- */
typedef struct { float x, y; } ds;
static inline ds two_sum(float a, float b) {
  const float s = a + b;
  const float bb = s - a;
  ds r = {s, (a - (s - bb)) + (b - bb)};
  return r;
}
static inline ds ds_add(ds a, ds b) {
  const ds s = two_sum(a.x, b.x), t = two_sum(a.y, b.y), u = two_sum(s.y, t.x), v = two_sum(s.x, u.x);
  ds r = {v.x, (v.y + u.y) + t.y};
  return r;
}
static inline ds ds_neg(ds a) { ds r = {-a.x, -a.y}; return r; }
static inline ds ds_scale(ds a, float b) {
  const float product = a.x * b;
  const float error = fmaf(a.x, b, -product) + a.y * b;
  return two_sum(product, error);
}
static inline ds ds_divide(ds a, float b) {
  const float q = a.x / b;
  const ds qq = {q, 0.0f};
  const ds remainder = ds_add(a, ds_neg(ds_scale(qq, b)));
  const float correction = (remainder.x + remainder.y) / b;
  return two_sum(q, correction);
}
static inline ds d4s6_ds(const ds* v) {
  return ds_add(ds_add(ds_add(v[0], v[1]), ds_add(v[4], v[5])), ds_add(v[2], v[3]));
}

/* ---- levels ---- */
static int cml_alloc(cml* L, const int* phys, const float* h) {
  for (int a = 0; a < 3; ++a) {
    L->D[a] = phys[a] + 2;
    L->h[a] = h[a];
  }
  L->n = L->D[0] * L->D[1] * L->D[2];
  const size_t n = (size_t)L->n;
  L->phi_raw = calloc(n, 4);
  L->phi = calloc(n, 4);
  L->topo = calloc(4 * n, 4);
  L->coef = calloc(4 * n, 4);
  L->p = calloc(n, 4);
  L->rhs0 = calloc(n, 4);
  L->rhs1 = calloc(n, 4);
  L->res = calloc(n, 4);
  L->pmin = calloc(n, 4);
  return (L->phi_raw && L->phi && L->topo && L->coef && L->p && L->rhs0 && L->rhs1 && L->res && L->pmin) ? 0 : -1;
}
static void cml_free(cml* L) {
  free(L->phi_raw);
  free(L->phi);
  free(L->topo);
  free(L->coef);
  free(L->p);
  free(L->rhs0);
  free(L->rhs1);
  free(L->res);
  free(L->pmin);
}

/* positive-face dual volume of an interior cell of the finest level
 * (pressureFaceData, webgpu-uniform-reference.wgsl.ts:536-561): 1 inside, 1/2 at a
 * box wall, 1 at the open top. */
static float face_volume(const int* n, int x, int y, int z, int axis, int open_top) {
  const int c[3] = {x, y, z};
  if (c[axis] + 1 < n[axis]) return 1.0f;
  if (axis == 1 && open_top) return 1.0f;
  return 0.5f;
}

/* mgBuildFinestTopology + mgBuildFinestRhs (…multigrid.wgsl.ts:135-167) */
static void finest_setup(cml* L, const int* n, const float* phi, const float* u, const float* v, const float* w,
                         const float* q, float dt, float rho, int open_top, float* b, float* pmin) {
  const float* h = L->h;
  const float hmin = fminf(h[0], fminf(h[1], h[2]));
  for (int X = 0; X < L->D[0]; ++X)
    for (int Y = 0; Y < L->D[1]; ++Y)
      for (int Z = 0; Z < L->D[2]; ++Z) {
        const int i = cidx(L, X, Y, Z);
        float* t = L->topo + 4 * (size_t)i;
        L->p[i] = 0.0f;
        if (cinterior(L, X, Y, Z)) {
          const int x = X - 1, y = Y - 1, z = Z - 1;
          const size_t s = ((size_t)x * n[1] + y) * n[2] + z;
          t[0] = 1.0f;
          for (int a = 0; a < 3; ++a) t[1 + a] = face_volume(n, x, y, z, a, open_top);
          L->phi_raw[i] = phi[s];
          pmin[i] = CM_NEG_BIG;
          float rhs = 0.0f;
          if (phi[s] < 0.0f) {
            /* divergenceAt (webgpu-uniform-reference.wgsl.ts:1010-1027), cell open fraction 1 */
            const float vi = 1.0f;
            float terms[6];
            for (int a = 0; a < 3; ++a) {
              const int c[3] = {x, y, z};
              const float vp = face_volume(n, x, y, z, a, open_top);
              const float vm = c[a] == 0 ? 0.5f : 1.0f;
              float up, um;
              if (a == 0) {
                up = u[((size_t)(x + 1) * n[1] + y) * n[2] + z];
                um = u[((size_t)x * n[1] + y) * n[2] + z];
              } else if (a == 1) {
                up = v[((size_t)x * (n[1] + 1) + y + 1) * n[2] + z];
                um = v[((size_t)x * (n[1] + 1) + y) * n[2] + z];
              } else {
                up = w[((size_t)x * n[1] + y) * (n[2] + 1) + z + 1];
                um = w[((size_t)x * n[1] + y) * (n[2] + 1) + z];
              }
              /* the solid-velocity terms (V - Vi) u_s with u_s = 0 on static box walls */
              terms[2 * a] = (vp * up) / h[a] + (vp - vi) * 0.0f;
              terms[2 * a + 1] = -(vm * um) / h[a] - (vm - vi) * 0.0f;
            }
            const float div = d4s6(terms);
            const float corr = q ? q[s] : 0.0f;
            rhs = rho * (div - corr) / dt;
          }
          b[i] = -rhs;
        } else {
          const int open = open_top && Y == L->D[1] - 1 && X > 0 && X < L->D[0] - 1 && Z > 0 && Z < L->D[2] - 1;
          const float o = open ? 1.0f : 0.0f;
          t[0] = t[1] = t[2] = t[3] = o;
          const int yi = Y > 0 && Y < L->D[1] - 1, zi = Z > 0 && Z < L->D[2] - 1, xi = X > 0 && X < L->D[0] - 1;
          if (X == 0 && yi && zi) t[1] = 0.5f;
          if (Y == 0 && xi && zi) t[2] = 0.5f;
          if (Z == 0 && xi && yi) t[3] = 0.5f;
          L->phi_raw[i] = 0.5f * hmin;
          b[i] = 0.0f;
          pmin[i] = open ? CM_NEG_BIG : 0.0f;
        }
      }
}

/* mgFineChild (…multigrid.wgsl.ts:71-76): fine padded index of corner o of coarse cell c */
static void fine_child(const cml* F, const cml* Cc, int cx, int cy, int cz, int o, int* out) {
  const int c[3] = {cx, cy, cz};
  for (int a = 0; a < 3; ++a) {
    const int fp = F->D[a] - 2;
    int cp = Cc->D[a] - 2;
    if (cp < 1) cp = 1;
    int stride = fp / cp;
    if (stride < 1) stride = 1;
    const int oa = (o >> a) & 1;
    int v = stride * (c[a] - 1) + oa * (stride - 1);
    if (v < -1) v = -1;
    if (v > fp) v = fp;
    out[a] = v + 1;
  }
}

/* mgDownsampleTopology (…multigrid.wgsl.ts:169-195); dest: destination level index */
static void down_topology(const cml* F, cml* Cc, int dest, int levels, int preserve, int open_top) {
  for (int X = 0; X < Cc->D[0]; ++X)
    for (int Y = 0; Y < Cc->D[1]; ++Y)
      for (int Z = 0; Z < Cc->D[2]; ++Z) {
        float tt[4][8], ph[8], pos[8], pf[8], nf[8];
        for (int o = 0; o < 8; ++o) {
          int q[3];
          fine_child(F, Cc, X, Y, Z, o, q);
          const int j = cidx(F, q[0], q[1], q[2]);
          for (int k = 0; k < 4; ++k) tt[k][o] = F->topo[4 * (size_t)j + k];
          const float phi = F->phi_raw[j];
          ph[o] = phi;
          pos[o] = phi >= 0.0f ? phi : 0.0f;
          pf[o] = phi >= 0.0f ? 1.0f : 0.0f;
          nf[o] = phi >= 0.0f ? 0.0f : 1.0f;
        }
        float vsum[4];
        for (int k = 0; k < 4; ++k) vsum[k] = d4s8(tt[k]);
        const float phi_sum = d4s8(ph), pos_sum = d4s8(pos), pc = d4s8(pf), nc = d4s8(nf);
        const int mixed = pc > 0.0f && nc > 0.0f;
        const int use_positive = mixed && (levels - dest) >= (levels - preserve);
        const float cphi = use_positive ? pos_sum / fmaxf(pc, 1.0f) : phi_sum / 8.0f;
        const int i = cidx(Cc, X, Y, Z);
        float* t = Cc->topo + 4 * (size_t)i;
        for (int k = 0; k < 4; ++k) t[k] = vsum[k] / 8.0f;
        if (X == Cc->D[0] - 2) t[1] = 0.5f;
        if (Y == Cc->D[1] - 2) t[2] = open_top ? 1.0f : 0.5f;
        if (Z == Cc->D[2] - 2) t[3] = 0.5f;
        Cc->phi_raw[i] = cphi;
      }
}

/* mgExtrapolatePhiOneCell (…multigrid.wgsl.ts:288-300) */
static void extrapolate_phi(cml* L) {
  for (int X = 0; X < L->D[0]; ++X)
    for (int Y = 0; Y < L->D[1]; ++Y)
      for (int Z = 0; Z < L->D[2]; ++Z) {
        const int i = cidx(L, X, Y, Z);
        if (L->topo[4 * (size_t)i] > 1e-5f) {
          L->phi[i] = L->phi_raw[i];
          continue;
        }
        float terms[6], wts[6];
        for (int nb = 0; nb < 6; ++nb) {
          const int x = X + NB[nb][0], y = Y + NB[nb][1], z = Z + NB[nb][2];
          terms[nb] = 0.0f;
          wts[nb] = 0.0f;
          if (!cvalid(L, x, y, z)) continue;
          const int j = cidx(L, x, y, z);
          const float v = L->topo[4 * (size_t)j];
          if (v > 1e-5f && L->phi_raw[j] < 0.0f) {
            terms[nb] = v * L->phi_raw[j];
            wts[nb] = v;
          }
        }
        const float sum = d4s6(terms), weight = d4s6(wts);
        L->phi[i] = weight > 0.0f ? sum / fmaxf(weight, 1e-9f) : L->phi_raw[i];
      }
}

/* mgCoefficientRaw (…multigrid.wgsl.ts:93-115) for the +axis face of cell (X,Y,Z) */
static float coefficient_raw(const cml* L, int X, int Y, int Z, int axis, int open_top) {
  int q[3] = {X, Y, Z};
  q[axis] += 1;
  const int i = cidx(L, X, Y, Z);
  const float h = L->h[axis];
  if (!cvalid(L, q[0], q[1], q[2])) {
    if (axis == 1 && Y == L->D[1] - 1 && open_top) {
      const float theta = cm_theta(L->phi[i], 0.5f * h, 1e-9f);
      return L->topo[4 * (size_t)i + 2] / (h * h * theta);
    }
    return 0.0f;
  }
  const float vf = L->topo[4 * (size_t)i + 1 + axis];
  if (vf <= 1e-6f) return 0.0f;
  const int j = cidx(L, q[0], q[1], q[2]);
  const int il = L->phi[i] < 0.0f, ql = L->phi[j] < 0.0f;
  float theta = 1.0f;
  if (il && !ql) theta = cm_theta(L->phi[i], L->phi[j], 1e-9f);
  if (!il && ql) theta = cm_theta(L->phi[j], L->phi[i], 1e-9f);
  return vf / (h * h * theta);
}

/* mgBakeCoefficients (…multigrid.wgsl.ts:302-309) */
static void bake(cml* L, int open_top) {
  for (int X = 0; X < L->D[0]; ++X)
    for (int Y = 0; Y < L->D[1]; ++Y)
      for (int Z = 0; Z < L->D[2]; ++Z) {
        const int i = cidx(L, X, Y, Z);
        float* c = L->coef + 4 * (size_t)i;
        for (int a = 0; a < 3; ++a) c[a] = coefficient_raw(L, X, Y, Z, a, open_top);
        c[3] = L->phi[i] < 0.0f ? 1.0f : 0.0f;
      }
}

static inline int baked(const cml* L, int x, int y, int z) {
  return cvalid(L, x, y, z) && L->coef[4 * (size_t)cidx(L, x, y, z) + 3] > 0.5f;
}
/* mgCoefficient (…multigrid.wgsl.ts:116-122): the coefficient of neighbour nb */
static inline float coefficient(const cml* L, int X, int Y, int Z, int nb) {
  const int axis = nb / 2, x = X + NB[nb][0], y = Y + NB[nb][1], z = Z + NB[nb][2];
  if (nb & 1) return L->coef[4 * (size_t)cidx(L, X, Y, Z) + axis];
  if (!cvalid(L, x, y, z)) return 0.0f;
  return L->coef[4 * (size_t)cidx(L, x, y, z) + axis];
}
static inline float pval(const cml* L, const float* p, int x, int y, int z) {
  if (x < 0) x = 0;
  if (y < 0) y = 0;
  if (z < 0) z = 0;
  if (x >= L->D[0]) x = L->D[0] - 1;
  if (y >= L->D[1]) y = L->D[1] - 1;
  if (z >= L->D[2]) z = L->D[2] - 1;
  return p[cidx(L, x, y, z)];
}
static float apply_at(const cml* L, const float* p, int X, int Y, int Z) {
  if (!baked(L, X, Y, Z)) return 0.0f;
  const float centre = p[cidx(L, X, Y, Z)];
  float terms[6];
  for (int nb = 0; nb < 6; ++nb) {
    const int x = X + NB[nb][0], y = Y + NB[nb][1], z = Z + NB[nb][2];
    const float a = coefficient(L, X, Y, Z, nb);
    const float nv = baked(L, x, y, z) ? pval(L, p, x, y, z) : 0.0f;
    terms[nb] = a * (centre - nv);
  }
  return d4s6(terms);
}

/* mgSmoothColour (…multigrid.wgsl.ts:325-344): the cells of one colour; in
 * place, because a pass reads only the other colour and writes only its own
 * (the pass-through projection of the other colour happens in its own pass). */
static void smooth_colour(cml* L, const float* rhs, const float* pmin, int colour) {
  for (int X = 0; X < L->D[0]; ++X)
    for (int Y = 0; Y < L->D[1]; ++Y)
      for (int Z = 0; Z < L->D[2]; ++Z) {
        if (((X + Y + Z) & 1) != colour) continue;
        const int i = cidx(L, X, Y, Z);
        if (!baked(L, X, Y, Z)) {
          L->p[i] = fmaxf(L->p[i], pmin[i]);
          continue;
        }
        float dt_[6], st[6];
        for (int nb = 0; nb < 6; ++nb) {
          const int x = X + NB[nb][0], y = Y + NB[nb][1], z = Z + NB[nb][2];
          const float a = coefficient(L, X, Y, Z, nb);
          dt_[nb] = a;
          st[nb] = baked(L, x, y, z) ? a * pval(L, L->p, x, y, z) : 0.0f;
        }
        const float diag = d4s6(dt_), sum = d4s6(st);
        const float pn = diag > 0.0f ? (sum + rhs[i]) / diag : 0.0f;
        L->p[i] = fmaxf(pn, pmin[i]);
      }
}
static void sweep(cml* L, const float* rhs, const float* pmin) {
  smooth_colour(L, rhs, pmin, 0);
  smooth_colour(L, rhs, pmin, 1);
}

/* mgResidual (…multigrid.wgsl.ts:197-205) */
static void residual(const cml* L, const float* rhs, float* out) {
  for (int X = 0; X < L->D[0]; ++X)
    for (int Y = 0; Y < L->D[1]; ++Y)
      for (int Z = 0; Z < L->D[2]; ++Z) {
        const int i = cidx(L, X, Y, Z);
        out[i] = baked(L, X, Y, Z) ? rhs[i] - apply_at(L, L->p, X, Y, Z) : 0.0f;
      }
}

/* mgRestrictResidual (…multigrid.wgsl.ts:210-215) */
static void restrict_to(const cml* F, const float* in, cml* Cc, float* out) {
  for (int X = 0; X < Cc->D[0]; ++X)
    for (int Y = 0; Y < Cc->D[1]; ++Y)
      for (int Z = 0; Z < Cc->D[2]; ++Z) {
        float t[8];
        for (int o = 0; o < 8; ++o) {
          int q[3];
          fine_child(F, Cc, X, Y, Z, o, q);
          t[o] = in[cidx(F, q[0], q[1], q[2])];
        }
        out[cidx(Cc, X, Y, Z)] = d4s8(t) / 8.0f;
      }
}

/* mgTrilinearPressure (…multigrid.wgsl.ts:217-241) at fine padded cell */
static float trilinear(const cml* Cc, const cml* F, int X, int Y, int Z) {
  const int id[3] = {X, Y, Z};
  float f[3];
  int base[3];
  for (int a = 0; a < 3; ++a) {
    const float fine_phys = (float)(F->D[a] - 2), coarse_phys = (float)(Cc->D[a] - 2);
    const float scale = coarse_phys / fmaxf(fine_phys, 1.0f);
    const float q = ((float)id[a] - 0.5f) * scale + 0.5f;
    const float fl = floorf(q);
    base[a] = (int)fl;
    f[a] = q - fl;
  }
  float values[8], weights[8];
  for (int o = 0; o < 8; ++o) {
    const int ox = o & 1, oy = (o >> 1) & 1, oz = (o >> 2) & 1;
    const int px = base[0] + ox, py = base[1] + oy, pz = base[2] + oz;
    values[o] = 0.0f;
    weights[o] = 0.0f;
    if (!cinterior(Cc, px, py, pz)) continue;
    const float w = ((ox ? f[0] : 1.0f - f[0]) * (oy ? f[1] : 1.0f - f[1])) * (oz ? f[2] : 1.0f - f[2]);
    values[o] = w * pval(Cc, Cc->p, px, py, pz);
    weights[o] = w;
  }
  const float value = d4s8(values), total = d4s8(weights);
  return total > 0.0f ? value / total : 0.0f;
}
static void prolong(const cml* Cc, cml* F, int add) {
  for (int X = 0; X < F->D[0]; ++X)
    for (int Y = 0; Y < F->D[1]; ++Y)
      for (int Z = 0; Z < F->D[2]; ++Z) {
        const int i = cidx(F, X, Y, Z);
        const float t = trilinear(Cc, F, X, Y, Z);
        F->p[i] = add ? F->p[i] + t : t;
      }
}

/* mgDownsampleSubtract / mgDownsampleMinimum (…multigrid.wgsl.ts:311-323) */
static void down_minimum(const cml* F, const float* fmin, int subtract, cml* Cc) {
  for (int X = 0; X < Cc->D[0]; ++X)
    for (int Y = 0; Y < Cc->D[1]; ++Y)
      for (int Z = 0; Z < Cc->D[2]; ++Z) {
        float lower = CM_NEG_BIG;
        for (int o = 0; o < 8; ++o) {
          int q[3];
          fine_child(F, Cc, X, Y, Z, o, q);
          const int j = cidx(F, q[0], q[1], q[2]);
          lower = fmaxf(lower, subtract ? fmin[j] - F->p[j] : fmin[j]);
        }
        Cc->pmin[cidx(Cc, X, Y, Z)] = lower;
      }
}
static void fill(float* a, int n, float v) {
  for (int i = 0; i < n; ++i) a[i] = v;
}

/* mgSolveCoarsest (…multigrid.wgsl.ts:404-471): coefficients from the coarse
 * level set and dual volumes, double-single projected red-black sweeps until
 * the projected residual (in 1/s) is at most tol or cap sweeps. */
static float coarse_coefficient(const cml* L, int X, int Y, int Z, int nb, int open_top) {
  const int axis = nb / 2, x = X + NB[nb][0], y = Y + NB[nb][1], z = Z + NB[nb][2];
  const int i = cidx(L, X, Y, Z);
  const float h = L->h[axis];
  if (!cvalid(L, x, y, z)) {
    if (axis == 1 && Y == L->D[1] - 1 && y == L->D[1] && open_top) {
      const float theta = cm_theta(L->phi[i], 0.5f * h, 1e-9f);
      return L->topo[4 * (size_t)i + 2] / (h * h * theta);
    }
    return 0.0f;
  }
  const int j = cidx(L, x, y, z);
  const float vf = (nb & 1) ? L->topo[4 * (size_t)i + 1 + axis] : L->topo[4 * (size_t)j + 1 + axis];
  if (vf <= 1e-6f) return 0.0f;
  float theta = 1.0f;
  if (L->phi[j] >= 0.0f) theta = cm_theta(L->phi[i], L->phi[j], 1e-9f);
  return vf / (h * h * theta);
}

static void coarse_solve(cml* L, const float* rhs, const float* pmin, float dt, float rho, float tol, int cap,
                         int open_top, orc_cm11a_report* rep) {
  const int n = L->n;
  ds* P = malloc(sizeof(ds) * (size_t)n);
  for (int i = 0; i < n; ++i) {
    P[i].x = L->p[i];
    P[i].y = 0.0f;
  }
  int converged = 0, iterations = 0;
  float worst = 0.0f;
  for (int it = 0; it < cap; ++it) {
    iterations = it + 1;
    for (int colour = 0; colour < 2; ++colour)
      for (int X = 0; X < L->D[0]; ++X)
        for (int Y = 0; Y < L->D[1]; ++Y)
          for (int Z = 0; Z < L->D[2]; ++Z) {
            const int i = cidx(L, X, Y, Z);
            if (!(L->phi[i] < 0.0f) || ((X + Y + Z) & 1) != colour) continue;
            float dg[6];
            ds st[6];
            for (int nb = 0; nb < 6; ++nb) {
              const int x = X + NB[nb][0], y = Y + NB[nb][1], z = Z + NB[nb][2];
              const float a = coarse_coefficient(L, X, Y, Z, nb, open_top);
              dg[nb] = a;
              st[nb].x = st[nb].y = 0.0f;
              if (cvalid(L, x, y, z)) {
                const int j = cidx(L, x, y, z);
                if (L->phi[j] < 0.0f) st[nb] = ds_scale(P[j], a);
              }
            }
            const float diag = d4s6(dg);
            const ds sum = d4s6_ds(st);
            if (diag > 0.0f) {
              const ds r0 = {rhs[i], 0.0f};
              const ds next = ds_divide(ds_add(sum, r0), diag);
              if (next.x + next.y < pmin[i]) {
                P[i].x = pmin[i];
                P[i].y = 0.0f;
              } else {
                P[i] = next;
              }
            }
          }
    for (int i = 0; i < n; ++i)
      if (P[i].x + P[i].y < pmin[i]) {
        P[i].x = pmin[i];
        P[i].y = 0.0f;
      }
    float mx = 0.0f;
    for (int X = 0; X < L->D[0]; ++X)
      for (int Y = 0; Y < L->D[1]; ++Y)
        for (int Z = 0; Z < L->D[2]; ++Z) {
          const int i = cidx(L, X, Y, Z);
          if (!(L->phi[i] < 0.0f)) continue;
          ds ap[6];
          float dg[6];
          for (int nb = 0; nb < 6; ++nb) {
            const int x = X + NB[nb][0], y = Y + NB[nb][1], z = Z + NB[nb][2];
            const float a = coarse_coefficient(L, X, Y, Z, nb, open_top);
            dg[nb] = a;
            ds nv = {0.0f, 0.0f};
            if (cvalid(L, x, y, z)) {
              const int j = cidx(L, x, y, z);
              if (L->phi[j] < 0.0f) nv = P[j];
            }
            ap[nb] = ds_scale(ds_add(P[i], ds_neg(nv)), a);
          }
          const ds applied = d4s6_ds(ap);
          const float diag = d4s6(dg);
          const ds r0 = {rhs[i], 0.0f};
          const ds lrp = ds_add(r0, ds_neg(applied));
          const float lr = lrp.x + lrp.y;
          if (diag > 0.0f) {
            const float pressure = P[i].x + P[i].y;
            const float gap = fmaxf(0.0f, pressure - pmin[i]);
            const int projects = lr < 0.0f && -lr >= gap * diag;
            const float lcp = projects ? gap * diag : fabsf(lr);
            const float divres = lcp * dt / rho;
            if (divres > mx) mx = divres;
          }
        }
    worst = mx;
    if (mx <= tol) {
      converged = 1;
      break;
    }
  }
  for (int i = 0; i < n; ++i) L->p[i] = P[i].x + P[i].y;
  free(P);
  if (iterations > rep->coarse_sweeps) rep->coarse_sweeps = iterations;
  if (worst > rep->coarse_residual) rep->coarse_residual = worst;
  if (!converged) rep->coarse_failed += 1;
  rep->coarse_solves += 1;
}

typedef struct {
  cml* lv;
  int L;
  float dt, rho, tol;
  int pre, post, cap, constraint_levels, open_top;
  orc_cm11a_report* rep;
} cm_ctx;

/* vCycle (…multigrid.ts:518-533); pmin0: the finest level's current p_min */
static void vcycle(cm_ctx* c, int l, const float* rhs, const float* pmin0) {
  cml* lv = c->lv;
  const float* pmin = l == 0 ? pmin0 : lv[l].pmin;
  if (l == c->L - 1) {
    coarse_solve(&lv[l], rhs, pmin, c->dt, c->rho, c->tol, c->cap, c->open_top, c->rep);
    return;
  }
  for (int i = 0; i < c->pre; ++i) sweep(&lv[l], rhs, pmin);
  residual(&lv[l], rhs, lv[l].res);
  restrict_to(&lv[l], lv[l].res, &lv[l + 1], lv[l + 1].rhs0);
  fill(lv[l + 1].p, lv[l + 1].n, 0.0f);
  if (l < c->constraint_levels)
    down_minimum(&lv[l], pmin, 1, &lv[l + 1]);
  else
    fill(lv[l + 1].pmin, lv[l + 1].n, CM_NEG_BIG);
  vcycle(c, l + 1, lv[l + 1].rhs0, pmin0);
  prolong(&lv[l + 1], &lv[l], 1);
  for (int i = 0; i < c->post; ++i) sweep(&lv[l], rhs, pmin);
}

/* fullCycle (…multigrid.ts:534-557) */
static void full_cycle(cm_ctx* c, const float* b, const float* pmin_orig, float* pmin_shift, float* backup) {
  cml* lv = c->lv;
  const int L = c->L;
  memcpy(backup, lv[0].p, sizeof(float) * (size_t)lv[0].n);
  for (int i = 0; i < lv[0].n; ++i) pmin_shift[i] = pmin_orig[i] - lv[0].p[i];
  residual(&lv[0], b, lv[0].rhs1);
  for (int l = 0; l + 1 < L; ++l) {
    restrict_to(&lv[l], lv[l].rhs1, &lv[l + 1], lv[l + 1].rhs1);
    if (l < c->constraint_levels)
      down_minimum(&lv[l], l == 0 ? pmin_shift : lv[l].pmin, 0, &lv[l + 1]);
    else
      fill(lv[l + 1].pmin, lv[l + 1].n, CM_NEG_BIG);
  }
  fill(lv[L - 1].p, lv[L - 1].n, 0.0f);
  coarse_solve(&lv[L - 1], lv[L - 1].rhs1, L - 1 == 0 ? pmin_shift : lv[L - 1].pmin, c->dt, c->rho, c->tol, c->cap,
               c->open_top, c->rep);
  for (int l = L - 2; l >= 0; --l) {
    prolong(&lv[l + 1], &lv[l], 0);
    vcycle(c, l, lv[l].rhs1, pmin_shift);
  }
  for (int i = 0; i < lv[0].n; ++i) lv[0].p[i] = lv[0].p[i] + backup[i];
}

/* mgMeasureFineResidual (…multigrid.wgsl.ts:473-481): max projected residual (1/s) */
static float measure(const cml* L, const float* b, const float* pmin, float dt, float rho) {
  float mx = 0.0f;
  for (int X = 0; X < L->D[0]; ++X)
    for (int Y = 0; Y < L->D[1]; ++Y)
      for (int Z = 0; Z < L->D[2]; ++Z) {
        if (!baked(L, X, Y, Z)) continue;
        float dg[6];
        for (int nb = 0; nb < 6; ++nb) dg[nb] = coefficient(L, X, Y, Z, nb);
        const float diag = d4s6(dg);
        if (diag <= 0.0f) continue;
        const int i = cidx(L, X, Y, Z);
        const float r = b[i] - apply_at(L, L->p, X, Y, Z);
        const float gap = fmaxf(0.0f, L->p[i] - pmin[i]);
        const int projects = r < 0.0f && -r >= gap * diag;
        const float pr = projects ? gap * diag : fabsf(r);
        const float v = pr * dt / rho;
        if (v > mx) mx = v;
      }
  return mx;
}

/* project (webgpu-uniform-reference.wgsl.ts:1052-1099) on a box container:
 * interior faces with a liquid side, the walls with the halo pressure of a
 * liquid cell (0 beside a cell without liquid), the open top against p = 0. */
static void project_velocities(const cml* L, const int* n, const float* phi, float dt, float rho, int open_top,
                               float* u, float* v, float* w) {
  const float scale = dt / rho;
  const float* h = L->h;
  const float hmin = fminf(h[0], fminf(h[1], h[2]));
  for (int x = 0; x < n[0]; ++x)
    for (int y = 0; y < n[1]; ++y)
      for (int z = 0; z < n[2]; ++z) {
        const size_t s = ((size_t)x * n[1] + y) * n[2] + z;
        const int c[3] = {x, y, z};
        const int liquid = phi[s] < 0.0f;
        const float p0 = liquid ? L->p[cidx(L, x + 1, y + 1, z + 1)] : 0.0f;
        for (int a = 0; a < 3; ++a) {
          float* face;
          size_t lo, hi;
          if (a == 0) {
            face = u;
            lo = ((size_t)x * n[1] + y) * n[2] + z;
            hi = ((size_t)(x + 1) * n[1] + y) * n[2] + z;
          } else if (a == 1) {
            face = v;
            lo = ((size_t)x * (n[1] + 1) + y) * n[2] + z;
            hi = ((size_t)x * (n[1] + 1) + y + 1) * n[2] + z;
          } else {
            face = w;
            lo = ((size_t)x * n[1] + y) * (n[2] + 1) + z;
            hi = ((size_t)x * n[1] + y) * (n[2] + 1) + z + 1;
          }
          int hx = x + 1, hy = y + 1, hz = z + 1;
          if (c[a] == 0) {
            /* the low wall: dual volume 1/2 > 0 */
            if (a == 0) hx = 0;
            else if (a == 1) hy = 0;
            else hz = 0;
            if (liquid) face[lo] -= scale * (p0 - L->p[cidx(L, hx, hy, hz)]) / h[a];
            else face[lo] = 0.0f;
            hx = x + 1, hy = y + 1, hz = z + 1;
          }
          if (c[a] == n[a] - 1) {
            if (a == 1 && open_top) {
              if (liquid) {
                const float theta = cm_theta(phi[s], 0.5f * hmin, 1e-6f);
                face[hi] -= scale * (0.0f - p0) / (h[1] * theta);
              } else {
                face[hi] = 0.0f;
              }
            } else if (liquid) {
              if (a == 0) hx = n[0] + 1;
              else if (a == 1) hy = n[1] + 1;
              else hz = n[2] + 1;
              face[hi] -= scale * (L->p[cidx(L, hx, hy, hz)] - p0) / h[a];
            } else {
              /* a cell without liquid has no pressure row: the wall face keeps u_s = 0
                 (the reference's thin-density case v <= u_s needs the density field) */
              face[hi] = 0.0f;
            }
            continue;
          }
          int d[3] = {x, y, z};
          d[a] += 1;
          const size_t sn = ((size_t)d[0] * n[1] + d[1]) * n[2] + d[2];
          const int nl = phi[sn] < 0.0f;
          if (liquid || nl) {
            const float p1 = nl ? L->p[cidx(L, d[0] + 1, d[1] + 1, d[2] + 1)] : 0.0f;
            float theta = 1.0f;
            if (liquid && !nl) theta = cm_theta(phi[s], phi[sn], 1e-6f);
            if (!liquid && nl) theta = cm_theta(phi[sn], phi[s], 1e-6f);
            face[hi] -= scale * (p1 - p0) / (h[a] * theta);
          } else {
            face[hi] = 0.0f;
          }
        }
      }
}

void orc_cm11a_default_options(orc_cm11a_options* o) {
  o->full_cycles = 3;
  o->v_cycles = 4;
  o->pre = 6;
  o->post = 6;
  o->constraint_levels = 3;
  o->preserve = 2;
  o->coarse_cap = 4096;
  o->coarse_tol = 1e-4f;
  o->open_top = 0;
}

int orc_cm11a_project(int nx, int ny, int nz, double h, const float* phi, const float* u, const float* v,
                      const float* w, const float* q, double dt, double rho, const orc_cm11a_options* opt,
                      float* p_out, float* u_out, float* v_out, float* w_out, orc_cm11a_report* rep) {
  int dims[ORC_MAX_LEVELS * 3], strides[ORC_MAX_LEVELS * 3], ext = 0;
  const int L = orc_plan_levels(nx, ny, nz, dims, strides, &ext);
  if (L <= 0) return -1;
  /* the reference's hierarchy halves axes exactly (mgFineChild's stride);
     lattices it rejects are not solved here either */
  for (int l = 1; l < L; ++l)
    for (int a = 0; a < 3; ++a)
      if (dims[3 * l + a] != dims[3 * (l - 1) + a] && 2 * dims[3 * l + a] != dims[3 * (l - 1) + a]) return -3;
  memset(rep, 0, sizeof(*rep));
  rep->levels = L;
  cml lv[ORC_MAX_LEVELS];
  memset(lv, 0, sizeof(lv));
  const int n[3] = {nx, ny, nz};
  int ok = 1;
  for (int l = 0; l < L; ++l) {
    float hl[3];
    for (int a = 0; a < 3; ++a) hl[a] = (float)(h * ((double)n[a] / (double)dims[3 * l + a]));
    ok = ok && cml_alloc(&lv[l], dims + 3 * l, hl) == 0;
  }
  if (!ok || lv[L - 1].n > 1024) {
    for (int l = 0; l < L; ++l) cml_free(&lv[l]);
    return -2;
  }
  const size_t n0 = (size_t)lv[0].n;
  float* b = calloc(n0, 4);
  float* pmin0 = calloc(n0, 4);
  float* pmin_shift = calloc(n0, 4);
  float* backup = calloc(n0, 4);
  const float fdt = (float)dt, frho = (float)rho;
  finest_setup(&lv[0], n, phi, u, v, w, q, fdt, frho, opt->open_top, b, pmin0);
  for (int l = 0; l + 1 < L; ++l) down_topology(&lv[l], &lv[l + 1], l + 1, L, opt->preserve, opt->open_top);
  for (int l = 0; l < L; ++l) {
    extrapolate_phi(&lv[l]);
    bake(&lv[l], opt->open_top);
  }
  cm_ctx c = {lv, L, fdt, frho, opt->coarse_tol, opt->pre, opt->post, opt->coarse_cap, opt->constraint_levels,
              opt->open_top, rep};
  for (int k = 0; k < opt->full_cycles; ++k) full_cycle(&c, b, pmin0, pmin_shift, backup);
  for (int k = 0; k < opt->v_cycles; ++k) vcycle(&c, 0, b, pmin0);
  rep->lcp_residual = measure(&lv[0], b, pmin0, fdt, frho);
  if (p_out) memcpy(p_out, lv[0].p, sizeof(float) * n0);
  const size_t nu = (size_t)(nx + 1) * ny * nz, nv = (size_t)nx * (ny + 1) * nz, nw = (size_t)nx * ny * (nz + 1);
  if (u_out && v_out && w_out) {
    memcpy(u_out, u, 4 * nu);
    memcpy(v_out, v, 4 * nv);
    memcpy(w_out, w, 4 * nw);
    project_velocities(&lv[0], n, phi, fdt, frho, opt->open_top, u_out, v_out, w_out);
  }
  free(b);
  free(pmin0);
  free(pmin_shift);
  free(backup);
  for (int l = 0; l < L; ++l) cml_free(&lv[l]);
  return L;
}

/* The finest level's assembled LCP (coefficients, b, p_min) for the
 * brute-force checks: coef (4 per padded cell: +x, +y, +z, liquid flag), b,
 * pmin; returns the padded cell count. */
int orc_cm11a_system(int nx, int ny, int nz, double h, const float* phi, const float* u, const float* v,
                     const float* w, double dt, double rho, int open_top, float* coef, float* b, float* pmin) {
  const int n[3] = {nx, ny, nz};
  const int phys[3] = {nx, ny, nz};
  const float hl[3] = {(float)h, (float)h, (float)h};
  cml L0;
  memset(&L0, 0, sizeof(L0));
  if (cml_alloc(&L0, phys, hl)) {
    cml_free(&L0);
    return -1;
  }
  finest_setup(&L0, n, phi, u, v, w, NULL, (float)dt, (float)rho, open_top, b, pmin);
  extrapolate_phi(&L0);
  bake(&L0, open_top);
  memcpy(coef, L0.coef, sizeof(float) * 4 * (size_t)L0.n);
  const int cells = L0.n;
  cml_free(&L0);
  return cells;
}


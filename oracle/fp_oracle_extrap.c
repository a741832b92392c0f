/* fp_oracle_extrap.c -- CPU restatement of the uniform velocity extrapolation
 * (TEST INFRASTRUCTURE ONLY: the checker of fp_extrapolate_velocity, never the
 * measured or shipped path).
 *
 * Follows lib/methods/uniform/webgpu-uniform-velocity-extrapolation.wgsl.ts
 * and its schedule (webgpu-uniform-velocity-extrapolation.ts:312-392):
 *   - each velocity component lives on its own lattice of positive faces (the
 *     face between cell p and p + e_c); a face is open when it is interior, or
 *     the top face of an open-top box (faceOpenFraction,
 *     webgpu-uniform-reference.wgsl.ts:511-519, no embedded solids);
 *   - source faces: open with liquid on either side (sourceFace, :100-104);
 *     liquid iff phi < 0, or density > 1/2 (CM12_LIQUID_ISOVALUE);
 *   - seed (seedActiveFront, :148-170): sources known at distance 0, open faces
 *     next to a source active;
 *   - synchronous Jeong-Ross-Whitaker passes (updateActiveFront, :296-331) of
 *     the Godunov distance (godunovDistance, :188-215) clipped to a two-cell band
 *     (:66-71), each updated face taking the first-order upwind extension of
 *     CM11b Eq. 7 (upwindExtensionValue, :221-248), convergence within one
 *     local ulp-scaled epsilon (fimConvergenceEpsilon, :114-119), activation of
 *     faces next to a converged neighbour (activatedByConvergedUpwindNeighbor,
 *     :266-283); at most min(max(dims), 16) passes, stopping at an empty
 *     active set (the indirect chain, :335-357 and …extrapolation.ts:146);
 *   - the hierarchy fill (CM11b Sec. 3.3.1): levels ceil-halved until the
 *     smallest extent is 1 (…extrapolation.ts:231-247); restriction by
 *     renormalised trilinear interpolation of known values with the
 *     component axis face-centred (hierarchyComponentSample, :385-413) and, for
 *     the vertical component only, the 2x2x2 corresponding-cell fallback
 *     (:430-447); prolongation of the unknown components (:473-488);
 *   - the result: known values on open faces, 0 elsewhere (packTransportShell,
 *     :491-512).
 * FP32 in the kernels' operation order (D4-ordered sums, :83-89), compiled
 * with -ffp-contract=off, so the device kernels give the same bits.
 */
#include "fp_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define EX_INF 65504.0f

typedef struct {
  int d[3];
  float h;
  int open_top;
  const float* phi;
  const float* density;
  uint8_t* cls;  /* bits 0-2 open(c), 3-5 source(c) */
} ex_grid;

static inline int64_t ex_idx(const int* d, int x, int y, int z) { return ((int64_t)x * d[1] + y) * d[2] + z; }

static inline int ex_in(const int* d, const int* p) {
  return p[0] >= 0 && p[1] >= 0 && p[2] >= 0 && p[0] < d[0] && p[1] < d[1] && p[2] < d[2];
}

static int ex_liquid(const ex_grid* g, const int* p) {
  if (!ex_in(g->d, p)) return 0;
  const int64_t i = ex_idx(g->d, p[0], p[1], p[2]);
  return g->phi ? g->phi[i] < 0.0f : g->density[i] > 0.5f;
}

static int ex_open(const ex_grid* g, const int* p, int c) {
  if (!ex_in(g->d, p)) return 0;
  if (p[c] + 1 < g->d[c]) return 1;
  return c == 1 && g->open_top;
}

static inline int cls_open(const ex_grid* g, const int* p, int c) {
  return ex_in(g->d, p) && (g->cls[ex_idx(g->d, p[0], p[1], p[2])] >> c & 1);
}
static inline int cls_source(const ex_grid* g, const int* p, int c) {
  return ex_in(g->d, p) && (g->cls[ex_idx(g->d, p[0], p[1], p[2])] >> (3 + c) & 1);
}

/* One ping-pong side of the band state. */
typedef struct {
  float* val[3];
  float* dist[3];
  uint8_t* known;   /* bit c */
  uint8_t* active;  /* bit c */
} ex_state;

static float ex_eps(const ex_grid* g, float reference) {
  const float minimum = 2.0f * g->h;
  const float dx = (float)g->d[0] * g->h, dy = (float)g->d[1] * g->h, dz = (float)g->d[2] * g->h;
  const float diagonal = sqrtf((dx * dx + dy * dy) + dz * dz);
  float a = fabsf(reference);
  a = a < minimum ? minimum : (a > diagonal ? diagonal : a);
  return a * 1.1920929e-7f;
}

static float nb_dist(const ex_grid* g, const ex_state* s, const int* q, int c) {
  if (!cls_open(g, q, c)) return EX_INF;
  return s->dist[c][ex_idx(g->d, q[0], q[1], q[2])];
}

static float nb_val(const ex_grid* g, const ex_state* s, const int* q, int c) {
  if (!cls_open(g, q, c)) return 0.0f;
  const int64_t i = ex_idx(g->d, q[0], q[1], q[2]);
  return (s->known[i] >> c & 1) ? s->val[c][i] : 0.0f;
}

static float godunov(const ex_grid* g, const ex_state* s, const int* p, int c) {
  float m[3], sp[3];
  for (int a = 0; a < 3; ++a) {
    int lo[3] = {p[0], p[1], p[2]}, hi[3] = {p[0], p[1], p[2]};
    lo[a] -= 1;
    hi[a] += 1;
    const float dl = nb_dist(g, s, lo, c), dh = nb_dist(g, s, hi, c);
    m[a] = dl < dh ? dl : dh;
    sp[a] = g->h;
  }
  for (int i = 0; i < 2; ++i)
    for (int j = i + 1; j < 3; ++j)
      if (m[j] < m[i]) {
        float t = m[i]; m[i] = m[j]; m[j] = t;
        t = sp[i]; sp[i] = sp[j]; sp[j] = t;
      }
  if (m[0] >= 0.5f * EX_INF) return EX_INF;
  float root = m[0] + sp[0];
  for (int count = 2; count <= 3; ++count) {
    if (m[count - 1] >= 0.5f * EX_INF || root <= m[count - 1]) break;
    float A = 0.0f, B = 0.0f, Cc = -1.0f;
    for (int k = 0; k < count; ++k) {
      const float inv = 1.0f / (sp[k] * sp[k]);
      A += inv;
      B += m[k] * inv;
      Cc += m[k] * m[k] * inv;
    }
    const float disc = B * B - A * Cc;
    root = (B + sqrtf(disc > 0.0f ? disc : 0.0f)) / A;
  }
  return root;
}

static float upwind(const ex_grid* g, const ex_state* s, const int* p, int c, float solved) {
  float wt[3], ws[3];
  const float eps = ex_eps(g, solved);
  for (int a = 0; a < 3; ++a) {
    wt[a] = 0.0f;
    ws[a] = 0.0f;
    int lo[3] = {p[0], p[1], p[2]}, hi[3] = {p[0], p[1], p[2]};
    lo[a] -= 1;
    hi[a] += 1;
    const float dl = nb_dist(g, s, lo, c), dh = nb_dist(g, s, hi, c);
    const float mn = dl < dh ? dl : dh;
    if (mn >= solved - eps) continue;
    float mv = 0.0f, mc = 0.0f;
    if (fabsf(dl - mn) <= eps) { mv += nb_val(g, s, lo, c); mc += 1.0f; }
    if (fabsf(dh - mn) <= eps) { mv += nb_val(g, s, hi, c); mc += 1.0f; }
    if (mc <= 0.0f) continue;
    const float weight = (solved - mn) / (g->h * g->h);
    wt[a] = weight * mv / mc;
    ws[a] = weight;
  }
  const float value = (wt[0] + wt[2]) + wt[1], sum = (ws[0] + ws[2]) + ws[1];
  return sum > 0.0f ? value / sum : 0.0f;
}

static int converges(const ex_grid* g, const ex_state* s, const int* q, int c) {
  if (!cls_open(g, q, c) || cls_source(g, q, c)) return 0;
  const int64_t i = ex_idx(g->d, q[0], q[1], q[2]);
  if (!(s->active[i] >> c & 1)) return 0;
  const float old = s->dist[c][i];
  if (old >= 0.5f * EX_INF) return 0;
  const float gd = godunov(g, s, q, c);
  const float upd = old < gd ? old : gd;
  if (upd > 2.0f * g->h) return 0;
  return fabsf(upd - old) <= ex_eps(g, old);
}

static int activated(const ex_grid* g, const ex_state* s, const int* p, int c) {
  const float cand = godunov(g, s, p, c);
  if (!cls_open(g, p, c) || cand > 2.0f * g->h) return 0;
  const float own = s->dist[c][ex_idx(g->d, p[0], p[1], p[2])];
  const float eps = ex_eps(g, own);
  for (int a = 0; a < 3; ++a)
    for (int side = -1; side <= 1; side += 2) {
      int q[3] = {p[0], p[1], p[2]};
      q[a] += side;
      if (converges(g, s, q, c)) {
        const float nd = nb_dist(g, s, q, c), gd = godunov(g, s, q, c);
        const float un = nd < gd ? nd : gd;
        if (own > un + eps) return 1;
      }
    }
  return 0;
}

/* One synchronous pass src -> dst; returns the active faces of dst. */
static int64_t fim_pass(const ex_grid* g, const ex_state* src, ex_state* dst) {
  const int* d = g->d;
  int64_t active = 0;
  for (int x = 0; x < d[0]; ++x)
    for (int y = 0; y < d[1]; ++y)
      for (int z = 0; z < d[2]; ++z) {
        const int p[3] = {x, y, z};
        const int64_t i = ex_idx(d, x, y, z);
        uint8_t known = src->known[i], act = src->active[i];
        for (int c = 0; c < 3; ++c) {
          dst->val[c][i] = src->val[c][i];
          dst->dist[c][i] = src->dist[c][i];
          if (!cls_open(g, p, c) || cls_source(g, p, c)) continue;
          if (src->active[i] >> c & 1) {
            const float old = src->dist[c][i];
            const float eps = ex_eps(g, old);
            const float cand = godunov(g, src, p, c);
            const float upd = old < cand ? old : cand;
            if (upd <= 2.0f * g->h) {
              dst->dist[c][i] = upd;
              dst->val[c][i] = upwind(g, src, p, c, upd);
              known |= (uint8_t)(1u << c);
            }
            const int conv = old < 0.5f * EX_INF && fabsf(upd - old) <= eps;
            if (!conv && upd <= 2.0f * g->h) act |= (uint8_t)(1u << c);
            else act &= (uint8_t)~(1u << c);
          } else if (activated(g, src, p, c)) {
            act |= (uint8_t)(1u << c);
          }
        }
        dst->known[i] = known;
        dst->active[i] = act;
        active += __builtin_popcount(act & 7u);
      }
  return active;
}

/* ---- hierarchy ---- */
typedef struct {
  int d[3];
  float* val[3];
  uint8_t* known;
} ex_level;

static int ex_level_alloc(ex_level* L, const int* d) {
  memcpy(L->d, d, sizeof(L->d));
  const size_t n = (size_t)d[0] * d[1] * d[2];
  for (int c = 0; c < 3; ++c) {
    L->val[c] = (float*)calloc(n, sizeof(float));
    if (!L->val[c]) return -1;
  }
  L->known = (uint8_t*)calloc(n, 1);
  return L->known ? 0 : -1;
}

static void ex_level_free(ex_level* L) {
  for (int c = 0; c < 3; ++c) free(L->val[c]);
  free(L->known);
}

/* renormalised trilinear sample of known values at target position p; returns the weight sum */
static float hier_sample(const ex_level* S, const int* T, const int* p, int c, float* value) {
  float pos[3];
  for (int a = 0; a < 3; ++a) pos[a] = ((float)p[a] + 0.5f) * (float)S->d[a] / (float)T[a] - 0.5f;
  pos[c] = (float)(p[c] + 1) * (float)S->d[c] / (float)T[c] - 1.0f;
  int lower[3];
  float fr[3];
  for (int a = 0; a < 3; ++a) {
    const float f = floorf(pos[a]);
    lower[a] = (int)f;
    fr[a] = pos[a] - f;
  }
  float cx[8], cw[8];
  for (int oz = 0; oz <= 1; ++oz)
    for (int oy = 0; oy <= 1; ++oy)
      for (int ox = 0; ox <= 1; ++ox) {
        const int off[3] = {ox, oy, oz};
        int q[3];
        float wgt[3];
        for (int a = 0; a < 3; ++a) {
          q[a] = lower[a] + off[a];
          q[a] = q[a] < 0 ? 0 : (q[a] > S->d[a] - 1 ? S->d[a] - 1 : q[a]);
          wgt[a] = off[a] ? fr[a] : 1.0f - fr[a];
        }
        const float weight = wgt[0] * wgt[1] * wgt[2];
        const int corner = ox + 2 * oy + 4 * oz;
        const int64_t i = ex_idx(S->d, q[0], q[1], q[2]);
        cx[corner] = 0.0f;
        cw[corner] = 0.0f;
        if (weight > 0.0f && (S->known[i] >> c & 1)) {
          cx[corner] = weight * S->val[c][i];
          cw[corner] = weight;
        }
      }
  const float x = ((cx[0] + cx[5]) + (cx[1] + cx[4])) + ((cx[2] + cx[7]) + (cx[3] + cx[6]));
  const float w = ((cw[0] + cw[5]) + (cw[1] + cw[4])) + ((cw[2] + cw[7]) + (cw[3] + cw[6]));
  *value = w > 0.0f ? x / w : 0.0f;
  return w;
}

static void known_at(const ex_level* S, int x, int y, int z, int c, float* v, float* w) {
  const int q[3] = {x, y, z};
  *v = 0.0f;
  *w = 0.0f;
  if (!ex_in(S->d, q)) return;
  const int64_t i = ex_idx(S->d, x, y, z);
  if (S->known[i] >> c & 1) {
    *v = S->val[c][i];
    *w = 1.0f;
  }
}

static float corresponding_sample(const ex_level* S, const int* p, int c, float* value) {
  const int lx = 2 * p[0], ly = 2 * p[1], lz = 2 * p[2];
  float sx = 0.0f, sw = 0.0f;
  for (int oy = 0; oy <= 1; ++oy) {
    float v00, w00, v11, w11, v10, w10, v01, w01;
    known_at(S, lx, ly + oy, lz, c, &v00, &w00);
    known_at(S, lx + 1, ly + oy, lz + 1, c, &v11, &w11);
    known_at(S, lx + 1, ly + oy, lz, c, &v10, &w10);
    known_at(S, lx, ly + oy, lz + 1, c, &v01, &w01);
    sx += (v00 + v11) + (v10 + v01);
    sw += (w00 + w11) + (w10 + w01);
  }
  *value = sw > 0.0f ? sx / sw : 0.0f;
  return sw;
}

static void restrict_level(const ex_level* S, ex_level* T) {
  for (int x = 0; x < T->d[0]; ++x)
    for (int y = 0; y < T->d[1]; ++y)
      for (int z = 0; z < T->d[2]; ++z) {
        const int p[3] = {x, y, z};
        const int64_t i = ex_idx(T->d, x, y, z);
        uint8_t known = 0;
        for (int c = 0; c < 3; ++c) {
          float v;
          float w = hier_sample(S, T->d, p, c, &v);
          if (w <= 0.0f && c == 1) w = corresponding_sample(S, p, c, &v);
          T->val[c][i] = w > 0.0f ? v : 0.0f;
          if (w > 0.0f) known |= (uint8_t)(1u << c);
        }
        T->known[i] = known;
      }
}

/* out = existing with its unknown components sampled from the coarser level S */
static void prolong_level(const ex_level* S, const ex_level* existing, ex_level* out) {
  for (int x = 0; x < out->d[0]; ++x)
    for (int y = 0; y < out->d[1]; ++y)
      for (int z = 0; z < out->d[2]; ++z) {
        const int p[3] = {x, y, z};
        const int64_t i = ex_idx(out->d, x, y, z);
        uint8_t known = existing->known[i];
        for (int c = 0; c < 3; ++c) {
          out->val[c][i] = existing->val[c][i];
          if (known >> c & 1) continue;
          float v;
          const float w = hier_sample(S, out->d, p, c, &v);
          if (w > 0.0f) {
            out->val[c][i] = v;
            known |= (uint8_t)(1u << c);
          }
        }
        out->known[i] = known;
      }
}

int orc_extrapolate(int nx, int ny, int nz, double h, int open_top, int max_passes, const float* phi,
                    const float* density, const float* u, const float* v, const float* w, float* uo, float* vo,
                    float* wo, float* band_dist, float* band_val, uint8_t* band_known, orc_extrap_report* rep) {
  if (nx < 1 || ny < 1 || nz < 1 || (!phi && !density)) return -1;
  ex_grid g;
  g.d[0] = nx; g.d[1] = ny; g.d[2] = nz;
  g.h = (float)h;
  g.open_top = open_top;
  g.phi = phi;
  g.density = phi ? NULL : density;
  const int64_t n = (int64_t)nx * ny * nz;
  g.cls = (uint8_t*)calloc((size_t)n, 1);
  ex_state s[2];
  int bad = !g.cls;
  for (int k = 0; k < 2; ++k) {
    for (int c = 0; c < 3; ++c) {
      s[k].val[c] = (float*)calloc((size_t)n, sizeof(float));
      s[k].dist[c] = (float*)calloc((size_t)n, sizeof(float));
      bad |= !s[k].val[c] || !s[k].dist[c];
    }
    s[k].known = (uint8_t*)calloc((size_t)n, 1);
    s[k].active = (uint8_t*)calloc((size_t)n, 1);
    bad |= !s[k].known || !s[k].active;
  }
  int rc = bad ? -2 : 0;
  const float* in[3] = {u, v, w};
  /* classes */
  for (int x = 0; x < nx && !rc; ++x)
    for (int y = 0; y < ny; ++y)
      for (int z = 0; z < nz; ++z) {
        const int p[3] = {x, y, z};
        uint8_t b = 0;
        for (int c = 0; c < 3; ++c) {
          if (!ex_open(&g, p, c)) continue;
          b |= (uint8_t)(1u << c);
          int q[3] = {x, y, z};
          q[c] += 1;
          if (ex_liquid(&g, p) || ex_liquid(&g, q)) b |= (uint8_t)(8u << c);
        }
        g.cls[ex_idx(g.d, x, y, z)] = b;
      }
  /* seed */
  int64_t active = 0;
  for (int x = 0; x < nx && !rc; ++x)
    for (int y = 0; y < ny; ++y)
      for (int z = 0; z < nz; ++z) {
        const int p[3] = {x, y, z};
        const int64_t i = ex_idx(g.d, x, y, z);
        uint8_t known = 0, act = 0;
        for (int c = 0; c < 3; ++c) {
          s[0].val[c][i] = 0.0f;
          s[0].dist[c][i] = EX_INF;
          if (!cls_open(&g, p, c)) continue;
          if (cls_source(&g, p, c)) {
            /* positive face of cell p: index p + e_c of the (n_c + 1)-long face array */
            const int64_t fi = c == 0 ? ((int64_t)(x + 1) * ny + y) * nz + z
                               : c == 1 ? ((int64_t)x * (ny + 1) + y + 1) * nz + z
                                        : ((int64_t)x * ny + y) * (nz + 1) + z + 1;
            s[0].val[c][i] = in[c][fi];
            s[0].dist[c][i] = 0.0f;
            known |= (uint8_t)(1u << c);
          } else {
            int touches = 0;
            for (int a = 0; a < 3 && !touches; ++a)
              for (int side = -1; side <= 1; side += 2) {
                int q[3] = {x, y, z};
                q[a] += side;
                if (cls_source(&g, q, c)) touches = 1;
              }
            if (touches) act |= (uint8_t)(1u << c);
          }
        }
        s[0].known[i] = known;
        s[0].active[i] = act;
        active += __builtin_popcount(act);
      }
  int ceiling = nx > ny ? nx : ny;
  ceiling = nz > ceiling ? nz : ceiling;
  ceiling = ceiling < 16 ? ceiling : 16;
  if (max_passes > 0) ceiling = max_passes;
  int latest = 0, passes = 0;
  for (int pass = 0; pass < ceiling && active > 0 && !rc; ++pass) {
    active = fim_pass(&g, &s[latest], &s[latest ^ 1]);
    latest ^= 1;
    passes += 1;
  }
  const ex_state* res = &s[latest];
  int64_t band_known_count = 0;
  if (!rc) {
    for (int64_t i = 0; i < n; ++i) band_known_count += __builtin_popcount(res->known[i] & 7u);
    for (int c = 0; c < 3; ++c) {
      if (band_dist) memcpy(band_dist + c * n, res->dist[c], (size_t)n * sizeof(float));
      if (band_val) memcpy(band_val + c * n, res->val[c], (size_t)n * sizeof(float));
    }
    if (band_known) memcpy(band_known, res->known, (size_t)n);
  }
  /* hierarchy */
  ex_level base;
  memcpy(base.d, g.d, sizeof(base.d));
  for (int c = 0; c < 3; ++c) base.val[c] = res->val[c];
  base.known = res->known;
  ex_level down[32], up[32];
  int L = 0;
  int dims[3] = {nx, ny, nz};
  while (!rc && (dims[0] < dims[1] ? (dims[0] < dims[2] ? dims[0] : dims[2]) : (dims[1] < dims[2] ? dims[1] : dims[2])) > 1) {
    for (int a = 0; a < 3; ++a) dims[a] = (dims[a] + 1) / 2;
    if (ex_level_alloc(&down[L], dims) || ex_level_alloc(&up[L], dims)) rc = -2;
    L += 1;
  }
  ex_level fin;
  int have_fin = 0;
  if (!rc && L > 0) {
    for (int l = 0; l < L; ++l) restrict_level(l == 0 ? &base : &down[l - 1], &down[l]);
    const ex_level* coarser = &down[L - 1];
    for (int l = L - 2; l >= 0; --l) {
      prolong_level(coarser, &down[l], &up[l]);
      coarser = &up[l];
    }
    if (ex_level_alloc(&fin, g.d)) rc = -2;
    else {
      have_fin = 1;
      prolong_level(coarser, &base, &fin);
    }
  }
  const ex_level* out = have_fin ? &fin : &base;
  int64_t unknown = 0;
  float* o[3] = {uo, vo, wo};
  if (!rc) {
    memset(uo, 0, sizeof(float) * (size_t)(nx + 1) * ny * nz);
    memset(vo, 0, sizeof(float) * (size_t)nx * (ny + 1) * nz);
    memset(wo, 0, sizeof(float) * (size_t)nx * ny * (nz + 1));
    for (int x = 0; x < nx; ++x)
      for (int y = 0; y < ny; ++y)
        for (int z = 0; z < nz; ++z) {
          const int p[3] = {x, y, z};
          const int64_t i = ex_idx(g.d, x, y, z);
          for (int c = 0; c < 3; ++c) {
            if (!cls_open(&g, p, c)) continue;
            const int64_t fi = c == 0 ? ((int64_t)(x + 1) * ny + y) * nz + z
                               : c == 1 ? ((int64_t)x * (ny + 1) + y + 1) * nz + z
                                        : ((int64_t)x * ny + y) * (nz + 1) + z + 1;
            if (out->known[i] >> c & 1) o[c][fi] = out->val[c][i];
            else unknown += 1;
          }
        }
  }
  if (rep) {
    rep->passes = passes;
    rep->active_left = (int)active;
    rep->levels = L;
    rep->band_known = band_known_count;
    rep->unknown = unknown;
  }
  for (int l = 0; l < L; ++l) {
    ex_level_free(&down[l]);
    ex_level_free(&up[l]);
  }
  if (have_fin) ex_level_free(&fin);
  for (int k = 0; k < 2; ++k) {
    for (int c = 0; c < 3; ++c) {
      free(s[k].val[c]);
      free(s[k].dist[c]);
    }
    free(s[k].known);
    free(s[k].active);
  }
  free(g.cls);
  return rc;
}

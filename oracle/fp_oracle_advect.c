/* fp_oracle_advect.c -- CPU restatement of the uniform velocity advection and
 * body forces (TEST INFRASTRUCTURE ONLY: the checker of fp_advect_velocity,
 * never the measured or shipped path).
 *
 * Follows lib/methods/uniform/webgpu-uniform-reference.wgsl.ts for box
 * containers without embedded solids, inflow or surface tension:
 *   - the transported field is the extrapolated velocity (the transport shell
 *     of packTransportShell): a face sample outside the domain reads 0
 *     (sampleVelocityComponent, :261-270, trilinear with D4-ordered corners,
 *     d4Sum8 :103-106);
 *   - departure points by RK2 in pieces of at most 1.5 cells, up to 32, each
 *     stage clamped into the domain (the top stays open for an open-top box)
 *     (departurePoint :324-334, clampVelocityTraceToDomain :317-322);
 *   - semi-Lagrangian advection (semiLagrangianAdvection :922-936): the
 *     positive wall faces keep min(advected, current face velocity), then
 *     body forces; bounded MacCormack (advect :938-952 prediction with wall
 *     faces restored, the predicted field extrapolated, reverseAdvection
 *     :954-961 with -dt, correctAdvection :993-1004 with boundedMacCormack
 *     :963-991: the correction reverts to the prediction outside the range
 *     of the departure cell's donors, read with clampCell, sampledFaceVelocity
 *     :247-250), then body forces;
 *   - body forces (applyVelocityForces :872-915): dt nu laplace(v) where the
 *     cell holds liquid (velocityLaplacian :866-869, the wall reflecting the
 *     velocity for no-slip walls, diffusionVelocity :860), gravity on the +y
 *     face where either adjacent cell holds liquid above 1e-5 occupancy
 *     (surfaceOccupancy :169-173: clamp(1/2 - phi/(4h)) for a level set,
 *     clamp(density) otherwise).
 * FP32 in the kernels' operation order, compiled with -ffp-contract=off, so
 * the device kernels give the same bits.
 */
#include "fp_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
  int d[3];
  float h;
  int open_top;
  const float* t[3];  /* transported (extrapolated) face arrays u, v, w */
} ad_field;

static inline int64_t fidx(const int* d, int c, int x, int y, int z) {
  /* face index p (the face between cells p-1 and p along c) in the (n_c + 1)-long array */
  return c == 0 ? ((int64_t)x * d[1] + y) * d[2] + z
       : c == 1 ? ((int64_t)x * (d[1] + 1) + y) * d[2] + z
                : ((int64_t)x * d[1] + y) * (d[2] + 1) + z;
}

/* the transport shell at padded index T (cell T - 1's positive face); 0 outside */
static float shell(const ad_field* F, int c, int tx, int ty, int tz) {
  const int T[3] = {tx, ty, tz};
  for (int a = 0; a < 3; ++a) {
    const int lo = a == c ? 0 : 1;
    if (T[a] < lo || T[a] > F->d[a]) return 0.0f;
  }
  /* padded T along c is face T of the array; along the others, cell T - 1 */
  const int x = c == 0 ? tx : tx - 1, y = c == 1 ? ty : ty - 1, z = c == 2 ? tz : tz - 1;
  return F->t[c][fidx(F->d, c, x, y, z)];
}

static float d4_8(const float* v) {
  const float y0 = (v[0] + v[5]) + (v[1] + v[4]);
  const float y1 = (v[2] + v[7]) + (v[3] + v[6]);
  return y0 + y1;
}

static float sample_comp(const ad_field* F, const float* p, int c) {
  float q[3];
  for (int a = 0; a < 3; ++a) {
    const float off = a == c ? 1.0f : 0.5f, lo = a == c ? -1.0f : 0.0f, hi = (float)(F->d[a] - 1);
    float v = p[a] - off;
    v = v < lo ? lo : (v > hi ? hi : v);
    q[a] = v;
  }
  int b[3];
  float fr[3];
  for (int a = 0; a < 3; ++a) {
    const float f = floorf(q[a]);
    b[a] = (int)f;
    fr[a] = q[a] - f;
  }
  float terms[8];
  for (int corner = 0; corner < 8; ++corner) {
    const int o[3] = {corner & 1, corner >> 1 & 1, corner >> 2 & 1};
    float w[3];
    for (int a = 0; a < 3; ++a) w[a] = o[a] ? fr[a] : 1.0f - fr[a];
    terms[corner] = w[0] * w[1] * w[2] * shell(F, c, b[0] + o[0] + 1, b[1] + o[1] + 1, b[2] + o[2] + 1);
  }
  return d4_8(terms);
}

static void clamp_trace(const ad_field* F, float* q) {
  q[0] = q[0] < 0.0f ? 0.0f : (q[0] > (float)F->d[0] ? (float)F->d[0] : q[0]);
  q[2] = q[2] < 0.0f ? 0.0f : (q[2] > (float)F->d[2] ? (float)F->d[2] : q[2]);
  q[1] = q[1] < 0.0f ? 0.0f : q[1];
  if (!F->open_top) q[1] = q[1] > (float)F->d[1] ? (float)F->d[1] : q[1];
}

static void departure(const ad_field* F, const float* pos, float dt, float* out) {
  float point[3] = {pos[0], pos[1], pos[2]};
  float remaining = fabsf(dt);
  const float direction = dt >= 0.0f ? 1.0f : -1.0f;
  const float h = F->h;
  for (int step = 0; step < 32; ++step) {
    if (remaining <= 1e-7f) break;
    float first[3];
    for (int c = 0; c < 3; ++c) first[c] = sample_comp(F, point, c);
    const float ax = fabsf(first[0]) / h, ay = fabsf(first[1]) / h, az = fabsf(first[2]) / h;
    const float rate = ax > (ay > az ? ay : az) ? ax : (ay > az ? ay : az);
    const float lim = 1.5f / (rate > 1e-6f ? rate : 1e-6f);
    const float step_s = remaining < lim ? remaining : lim;
    const float signed_step = direction * step_s;
    float mid[3];
    for (int a = 0; a < 3; ++a) mid[a] = point[a] - 0.5f * first[a] * signed_step / h;
    clamp_trace(F, mid);
    float second[3];
    for (int c = 0; c < 3; ++c) second[c] = sample_comp(F, mid, c);
    for (int a = 0; a < 3; ++a) point[a] = point[a] - second[a] * signed_step / h;
    clamp_trace(F, point);
    remaining -= step_s;
  }
  out[0] = point[0];
  out[1] = point[1];
  out[2] = point[2];
}

static float advect_comp(const ad_field* F, int x, int y, int z, int c, float dt) {
  float pos[3] = {(float)x + 0.5f, (float)y + 0.5f, (float)z + 0.5f};
  pos[c] = (float)(c == 0 ? x : c == 1 ? y : z) + 1.0f;
  float dep[3];
  departure(F, pos, dt, dep);
  return sample_comp(F, dep, c);
}

/* donor value (sampledFaceVelocity): 0 past the component's extent, the
 * clamped cell otherwise */
static float donor(const ad_field* F, const int* p, int c) {
  if (p[c] < 0 || p[c] >= F->d[c]) return 0.0f;
  int q[3];
  for (int a = 0; a < 3; ++a) q[a] = p[a] < 0 ? 0 : (p[a] > F->d[a] - 1 ? F->d[a] - 1 : p[a]);
  return shell(F, c, q[0] + 1, q[1] + 1, q[2] + 1);
}

static float bounded(const ad_field* F, int x, int y, int z, int c, float dt, float predicted, float original,
                     float reversed) {
  float pos[3] = {(float)x + 0.5f, (float)y + 0.5f, (float)z + 0.5f};
  pos[c] = (float)(c == 0 ? x : c == 1 ? y : z) + 1.0f;
  float dep[3];
  departure(F, pos, dt, dep);
  float q[3];
  for (int a = 0; a < 3; ++a) {
    const float off = a == c ? 1.0f : 0.5f, lo = a == c ? -1.0f : 0.0f, hi = (float)(F->d[a] - 1);
    float v = dep[a] - off;
    q[a] = v < lo ? lo : (v > hi ? hi : v);
  }
  int b[3];
  float fr[3];
  for (int a = 0; a < 3; ++a) {
    const float f = floorf(q[a]);
    b[a] = (int)f;
    fr[a] = q[a] - f;
  }
  float lower = 1e30f, upper = -1e30f;
  for (int corner = 0; corner < 8; ++corner) {
    const int o[3] = {corner & 1, corner >> 1 & 1, corner >> 2 & 1};
    const int p[3] = {b[0] + o[0], b[1] + o[1], b[2] + o[2]};
    float w[3];
    for (int a = 0; a < 3; ++a) w[a] = o[a] ? fr[a] : 1.0f - fr[a];
    const float weight = w[0] * w[1] * w[2];
    const float value = donor(F, p, c);
    if (weight > 0.0f) {
      lower = value < lower ? value : lower;
      upper = value > upper ? value : upper;
    }
  }
  const float corrected = predicted + 0.5f * (original - reversed);
  const int revert = corrected < lower || corrected > upper;
  return revert ? predicted : corrected;
}

/* per-cell positive-face velocity of an authoritative (projected) field */
static float own_face(const int* d, const float* const* f, int c, int x, int y, int z) {
  return f[c][fidx(d, c, x + (c == 0), y + (c == 1), z + (c == 2))];
}

static float occupancy(const int* d, const float* phi, const float* dens, float h, int x, int y, int z) {
  if (x < 0 || y < 0 || z < 0 || x >= d[0] || y >= d[1] || z >= d[2]) return 0.0f;
  const int64_t i = ((int64_t)x * d[1] + y) * d[2] + z;
  float v = phi ? 0.5f - phi[i] / (4.0f * h) : dens[i];
  return v < 0.0f ? 0.0f : (v > 1.0f ? 1.0f : v);
}

/* diffusionVelocity: the cell's positive faces, clamped into the domain, the
 * velocity reflected outside the domain for no-slip walls */
static void diff_vel(const int* d, const float* const* f, int no_slip, int x, int y, int z, float* out) {
  const int inside = x >= 0 && y >= 0 && z >= 0 && x < d[0] && y < d[1] && z < d[2];
  const int cx = x < 0 ? 0 : (x >= d[0] ? d[0] - 1 : x);
  const int cy = y < 0 ? 0 : (y >= d[1] ? d[1] - 1 : y);
  const int cz = z < 0 ? 0 : (z >= d[2] ? d[2] - 1 : z);
  for (int c = 0; c < 3; ++c) {
    const float v = own_face(d, f, c, cx, cy, cz);
    out[c] = no_slip && !inside ? -v : v;
  }
}

static void forces(const int* d, const float* const* cur, const float* phi, const float* dens, float h, float dt,
                   float gravity, float nu, int no_slip, int x, int y, int z, float* v) {
  const float occ = occupancy(d, phi, dens, h, x, y, z);
  if (occ > 0.0f && nu > 0.0f) {
    float ce[3], xp[3], xm[3], yp[3], ym[3], zp[3], zm[3];
    diff_vel(d, cur, no_slip, x, y, z, ce);
    diff_vel(d, cur, no_slip, x + 1, y, z, xp);
    diff_vel(d, cur, no_slip, x - 1, y, z, xm);
    diff_vel(d, cur, no_slip, x, y + 1, z, yp);
    diff_vel(d, cur, no_slip, x, y - 1, z, ym);
    diff_vel(d, cur, no_slip, x, y, z + 1, zp);
    diff_vel(d, cur, no_slip, x, y, z - 1, zm);
    const float h2 = h * h;
    for (int c = 0; c < 3; ++c) {
      const float lap = ((xp[c] - 2.0f * ce[c] + xm[c]) / h2 + (yp[c] - 2.0f * ce[c] + ym[c]) / h2) +
                        (zp[c] - 2.0f * ce[c] + zm[c]) / h2;
      v[c] += dt * nu * lap;
    }
  }
  const float occ_y = occupancy(d, phi, dens, h, x, y + 1, z);
  if (occ > 1e-5f || occ_y > 1e-5f) v[1] += gravity * dt;
}

int orc_advect_velocity(int nx, int ny, int nz, double h, double dt, double gravity, double nu, int scheme,
                        int open_top, int no_slip, const float* phi, const float* density, const float* u,
                        const float* v, const float* w, const float* tu, const float* tv, const float* tw,
                        float* uo, float* vo, float* wo) {
  if (nx < 1 || ny < 1 || nz < 1 || (!phi && !density)) return -1;
  const int d[3] = {nx, ny, nz};
  const float hf = (float)h, dtf = (float)dt, gf = (float)gravity, nuf = (float)nu;
  const float* cur[3] = {u, v, w};
  float* out[3] = {uo, vo, wo};
  const size_t nf[3] = {(size_t)(nx + 1) * ny * nz, (size_t)nx * (ny + 1) * nz, (size_t)nx * ny * (nz + 1)};
  ad_field F = {{nx, ny, nz}, hf, open_top, {tu, tv, tw}};
  /* the negative domain faces carry over unchanged (carryBoundaryVelocity) */
  for (int c = 0; c < 3; ++c) memcpy(out[c], cur[c], nf[c] * sizeof(float));
  float* pred[3] = {NULL, NULL, NULL};
  float* rev[3] = {NULL, NULL, NULL};
  if (scheme == 1) {
    for (int c = 0; c < 3; ++c) {
      pred[c] = (float*)malloc(nf[c] * sizeof(float));
      rev[c] = (float*)malloc(nf[c] * sizeof(float));
      if (!pred[c] || !rev[c]) return -2;
      memcpy(pred[c], cur[c], nf[c] * sizeof(float));
      memcpy(rev[c], cur[c], nf[c] * sizeof(float));
    }
    /* prediction: walls restored (advect) */
    for (int x = 0; x < nx; ++x)
      for (int y = 0; y < ny; ++y)
        for (int z = 0; z < nz; ++z)
          for (int c = 0; c < 3; ++c) {
            const int p[3] = {x, y, z};
            const int64_t fi = fidx(d, c, x + (c == 0), y + (c == 1), z + (c == 2));
            float val = advect_comp(&F, x, y, z, c, dtf);
            if (p[c] == d[c] - 1 && !(c == 1 && open_top)) val = cur[c][fi];
            pred[c][fi] = val;
          }
    /* the predicted field's transport shell */
    float* te[3];
    for (int c = 0; c < 3; ++c) {
      te[c] = (float*)malloc(nf[c] * sizeof(float));
      if (!te[c]) return -2;
    }
    orc_extrap_report er;
    int rc = orc_extrapolate(nx, ny, nz, h, open_top, 0, phi, density, pred[0], pred[1], pred[2], te[0], te[1],
                             te[2], NULL, NULL, NULL, &er);
    if (rc) return rc;
    ad_field P = {{nx, ny, nz}, hf, open_top, {te[0], te[1], te[2]}};
    /* reverse advection of the prediction with -dt, walls from the prediction */
    for (int x = 0; x < nx; ++x)
      for (int y = 0; y < ny; ++y)
        for (int z = 0; z < nz; ++z)
          for (int c = 0; c < 3; ++c) {
            const int p[3] = {x, y, z};
            const int64_t fi = fidx(d, c, x + (c == 0), y + (c == 1), z + (c == 2));
            float val = advect_comp(&P, x, y, z, c, -dtf);
            if (p[c] == d[c] - 1 && !(c == 1 && open_top)) val = pred[c][fi];
            rev[c][fi] = val;
          }
    for (int c = 0; c < 3; ++c) free(te[c]);
  }
  for (int x = 0; x < nx; ++x)
    for (int y = 0; y < ny; ++y)
      for (int z = 0; z < nz; ++z) {
        const int p[3] = {x, y, z};
        float val[3];
        for (int c = 0; c < 3; ++c) {
          const int64_t fi = fidx(d, c, x + (c == 0), y + (c == 1), z + (c == 2));
          if (scheme == 1) {
            val[c] = bounded(&F, x, y, z, c, dtf, pred[c][fi], cur[c][fi], rev[c][fi]);
          } else {
            val[c] = advect_comp(&F, x, y, z, c, dtf);
            if (p[c] == d[c] - 1 && !(c == 1 && open_top)) val[c] = val[c] < cur[c][fi] ? val[c] : cur[c][fi];
          }
        }
        forces(d, cur, phi, density, hf, dtf, gf, nuf, no_slip, x, y, z, val);
        for (int c = 0; c < 3; ++c) out[c][fidx(d, c, x + (c == 0), y + (c == 1), z + (c == 2))] = val[c];
      }
  for (int c = 0; c < 3; ++c) {
    free(pred[c]);
    free(rev[c]);
  }
  return 0;
}

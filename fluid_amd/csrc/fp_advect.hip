// fp_advect.hip -- velocity advection and body forces of a uniform grid on MI355X.
//
// Kernels restate lib/methods/uniform/webgpu-uniform-reference.wgsl.ts (cited
// per function) for box containers without solids, inflow or surface tension.
// One thread per cell advances the cell's three positive faces; each face's
// characteristic is an RK2 trace through the transported (extrapolated)
// field with trilinear samples in D4 corner order.  The face arrays keep this
// library's layout (z fastest, u (nx+1,ny,nz) ...): the reference's padded
// transport shell is the same data with the outside read as 0.  No
// contraction: results equal oracle/fp_oracle_advect.c bit for bit.
#pragma clang fp contract(off)

#include "fp_advect.hpp"
#include "fp_extrap.hpp"
#include "fp_uniform.hpp"

#include <algorithm>
#include <chrono>
#include <cstring>

namespace fp {

namespace {

constexpr int kThreads = 256;

struct AdGeo {
  int d[3];
  float h;
  int open_top;
  int no_slip;
  float dt, gravity, nu;
};

struct Faces {
  const float* f[3];
};

// v clamped into [lo, hi] (lo < hi) in one instruction.  For finite v this
// equals v < lo ? lo : (v > hi ? hi : v) up to the sign of a zero result;
// positions are only ever used as p - off with off >= 1/2, or compared, so
// that sign never reaches a result.
__device__ __forceinline__ float med3(float v, float lo, float hi) { return __builtin_amdgcn_fmed3f(v, lo, hi); }

// a * b for a, b < 2^24 and a product below 2^32: one full-rate 24-bit
// multiply (written out: the compiler keeps a quarter-rate 32-bit multiply
// when the operands' ranges are not visible inside the trace loop)
__device__ __forceinline__ unsigned mul24(unsigned a, unsigned b) {
  unsigned r;
  asm("v_mul_u32_u24 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
  return r;
}

// face (x, y, z) of component c (fewer than 2^24 rows and 2^30 faces per
// component, checked at set-up)
__device__ __forceinline__ unsigned fidx(const AdGeo& g, int c, int x, int y, int z) {
  const unsigned ey = g.d[1] + (c == 1), ez = g.d[2] + (c == 2);
  return mul24(mul24((unsigned)x, ey) + (unsigned)y, ez) + (unsigned)z;
}

// sampleVelocityComponent (…wgsl.ts:261-270) with d4Sum8 (:103-106)
__device__ float sample_comp(const AdGeo& g, const Faces& t, const float* p, int c) {
  float q[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const float off = a == c ? 1.0f : 0.5f, lo = a == c ? -1.0f : 0.0f, hi = (float)(g.d[a] - 1);
    const float v = p[a] - off;
    q[a] = med3(v, lo, hi);
  }
  int b[3];
  float fr[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const float f = floorf(q[a]);
    // q lies in [lo, hi], so this clamp changes nothing; it keeps the
    // indices below inside the arrays whatever the input
    b[a] = min(max((int)f, a == c ? -1 : 0), g.d[a] - 1);
    fr[a] = q[a] - f;
  }
  // All eight corners' values first, then the weights.  Corner o is face
  // array element e = b + o (+1 along c): e >= 0 always, and only an o = 1
  // corner at the top of an axis lies past the array (it reads 0, as the
  // reference's padded shell does; its weight is 0).  Corners k and k + 4
  // differ only in z and come from one 8-byte load at a row index clamped so
  // that both words lie in the row (every row holds at least 2 faces: grids
  // are at least 2 cells on each axis).  Byte offsets fit 32 bits (fewer than
  // 2^30 faces per component), so the loads take a scalar base.
  const int ex = g.d[0] + (c == 0), ey = g.d[1] + (c == 1), ez = g.d[2] + (c == 2);
  const int ix0 = b[0] + (c == 0), iy0 = b[1] + (c == 1), iz0 = b[2] + (c == 2);
  const bool inx1 = ix0 + 1 < ex, iny1 = iy0 + 1 < ey, topz = iz0 > ez - 2;
  const int ix1 = min(ix0 + 1, ex - 1), iy1 = min(iy0 + 1, ey - 1), zb = min(iz0, ez - 2);
  // (24-bit multiplies: fewer than 2^24 rows per component)
  const unsigned r0 = mul24((unsigned)ix0, (unsigned)ey), r1 = mul24((unsigned)ix1, (unsigned)ey);
  const unsigned rows[4] = {r0 + iy0, r1 + iy0, r0 + iy1, r1 + iy1};
  const char* base = (const char*)t.f[c];
  float vals[8];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    float2 pair;
    __builtin_memcpy(&pair, base + ((mul24(rows[k], (unsigned)ez) + (unsigned)zb) << 2), sizeof(pair));
    const bool in = ((k & 1) == 0 || inx1) && ((k & 2) == 0 || iny1);
    vals[k] = in ? (topz ? pair.y : pair.x) : 0.0f;
    vals[k + 4] = in && !topz ? pair.y : 0.0f;
  }
  float terms[8];
#pragma unroll
  for (int corner = 0; corner < 8; ++corner) {
    const int o[3] = {corner & 1, corner >> 1 & 1, corner >> 2 & 1};
    float w[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) w[a] = o[a] ? fr[a] : 1.0f - fr[a];
    terms[corner] = w[0] * w[1] * w[2] * vals[corner];
  }
  const float y0 = (terms[0] + terms[5]) + (terms[1] + terms[4]);
  const float y1 = (terms[2] + terms[7]) + (terms[3] + terms[6]);
  return y0 + y1;
}

// clampVelocityTraceToDomain (…wgsl.ts:317-322)
__device__ __forceinline__ void clamp_trace(const AdGeo& g, float* q) {
  q[0] = med3(q[0], 0.0f, (float)g.d[0]);
  q[2] = med3(q[2], 0.0f, (float)g.d[2]);
  q[1] = g.open_top ? fmaxf(q[1], 0.0f) : med3(q[1], 0.0f, (float)g.d[1]);
}

// departurePoint (…wgsl.ts:324-334), RK2 in pieces of at most 1.5 cells, for
// the cell's three faces at once: the same operations per face, the three
// traces advanced together (a finished trace keeps its point), so each round
// of samples serves three faces
__device__ void departure3(const AdGeo& g, const Faces& t, const float (*pos)[3], float dt, float (*out)[3]) {
  float point[3][3], remaining[3];
#pragma unroll
  for (int f = 0; f < 3; ++f) {
    point[f][0] = pos[f][0];
    point[f][1] = pos[f][1];
    point[f][2] = pos[f][2];
    remaining[f] = fabsf(dt);
  }
  const float direction = dt >= 0.0f ? 1.0f : -1.0f;
  const float h = g.h;
  for (int step = 0; step < 32; ++step) {
    bool live[3];
#pragma unroll
    for (int f = 0; f < 3; ++f) live[f] = !(remaining[f] <= 1e-7f);
    if (!live[0] && !live[1] && !live[2]) break;
    float first[3][3];
#pragma unroll
    for (int f = 0; f < 3; ++f)
#pragma unroll
      for (int c = 0; c < 3; ++c) first[f][c] = sample_comp(g, t, point[f], c);
    float mid[3][3], step_s[3], signed_step[3];
#pragma unroll
    for (int f = 0; f < 3; ++f) {
      // max(|u|/h, |v|/h, |w|/h) as max(|u|, |v|, |w|)/h: a correctly rounded
      // division by h > 0 is monotonic, so both pick the same value
      const float ax = fabsf(first[f][0]), ay = fabsf(first[f][1]), az = fabsf(first[f][2]);
      const float myz = ay > az ? ay : az;
      const float rate = (ax > myz ? ax : myz) / h;
      const float lim = 1.5f / (rate > 1e-6f ? rate : 1e-6f);
      step_s[f] = remaining[f] < lim ? remaining[f] : lim;
      signed_step[f] = direction * step_s[f];
#pragma unroll
      for (int a = 0; a < 3; ++a) mid[f][a] = point[f][a] - 0.5f * first[f][a] * signed_step[f] / h;
      clamp_trace(g, mid[f]);
    }
    float second[3][3];
#pragma unroll
    for (int f = 0; f < 3; ++f)
#pragma unroll
      for (int c = 0; c < 3; ++c) second[f][c] = sample_comp(g, t, mid[f], c);
#pragma unroll
    for (int f = 0; f < 3; ++f) {
      float next[3];
#pragma unroll
      for (int a = 0; a < 3; ++a) next[a] = point[f][a] - second[f][a] * signed_step[f] / h;
      clamp_trace(g, next);
#pragma unroll
      for (int a = 0; a < 3; ++a) point[f][a] = live[f] ? next[a] : point[f][a];
      remaining[f] = live[f] ? remaining[f] - step_s[f] : remaining[f];
    }
  }
#pragma unroll
  for (int f = 0; f < 3; ++f) {
    out[f][0] = point[f][0];
    out[f][1] = point[f][1];
    out[f][2] = point[f][2];
  }
}

__device__ __forceinline__ void face_position(int x, int y, int z, int c, float* pos) {
  pos[0] = (float)x + 0.5f;
  pos[1] = (float)y + 0.5f;
  pos[2] = (float)z + 0.5f;
  pos[c] = (float)(c == 0 ? x : c == 1 ? y : z) + 1.0f;
}

// advectVelocityComponent (…wgsl.ts:415-423) without solids or the
// liquid-only toggle, for the three faces of cell (x,y,z) together
__device__ void advect3(const AdGeo& g, const Faces& t, int x, int y, int z, float dt, float* val) {
  float pos[3][3], dep[3][3];
#pragma unroll
  for (int c = 0; c < 3; ++c) face_position(x, y, z, c, pos[c]);
  departure3(g, t, pos, dt, dep);
#pragma unroll
  for (int c = 0; c < 3; ++c) val[c] = sample_comp(g, t, dep[c], c);
}


// boundedMacCormack (…wgsl.ts:963-991)
// (dep: the face's departure point, the forward trace of the prediction)
__device__ float bounded(const AdGeo& g, const Faces& t, const float* dep, int c, float predicted, float original,
                         float reversed) {
  float q[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const float off = a == c ? 1.0f : 0.5f, lo = a == c ? -1.0f : 0.0f, hi = (float)(g.d[a] - 1);
    const float v = dep[a] - off;
    q[a] = med3(v, lo, hi);
  }
  int b[3];
  float fr[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const float f = floorf(q[a]);
    b[a] = min(max((int)f, a == c ? -1 : 0), g.d[a] - 1);  // changes nothing; keeps indices in the arrays
    fr[a] = q[a] - f;
  }
  // The donors (sampledFaceVelocity, …wgsl.ts:247-250): corner o reads face
  // array element clamp(b + o, 0, d - 1) (+1 along c), or 0 when b + o
  // leaves [0, d - 1] along c.  The two z corners of each (x, y) come from
  // one 8-byte load, as in sample_comp.
  const int ey = g.d[1] + (c == 1), ez = g.d[2] + (c == 2);
  int idx[3][2];
  bool valid[2] = {true, true};
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int o = 0; o < 2; ++o) {
      const int p = b[a] + o;
      idx[a][o] = min(max(p, 0), g.d[a] - 1) + (a == c);
      if (a == c) valid[o] = p >= 0 && p < g.d[a];
    }
  const int zb = min(idx[2][0], ez - 2);
  const bool z0hi = idx[2][0] > zb, z1hi = idx[2][1] > zb;
  const char* base = (const char*)t.f[c];
  float values[8];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const unsigned row = mul24((unsigned)idx[0][k & 1], (unsigned)ey) + (unsigned)idx[1][k >> 1];
    float2 pair;
    __builtin_memcpy(&pair, base + ((mul24(row, (unsigned)ez) + (unsigned)zb) << 2), sizeof(pair));
    const int o[3] = {k & 1, k >> 1, 0};
    values[k] = valid[o[c]] ? (z0hi ? pair.y : pair.x) : 0.0f;
    const int o4[3] = {k & 1, k >> 1, 1};
    values[k + 4] = valid[o4[c]] ? (z1hi ? pair.y : pair.x) : 0.0f;
  }
  float lower = 1e30f, upper = -1e30f;
#pragma unroll
  for (int corner = 0; corner < 8; ++corner) {
    const int o[3] = {corner & 1, corner >> 1 & 1, corner >> 2 & 1};
    float w[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) w[a] = o[a] ? fr[a] : 1.0f - fr[a];
    const float weight = w[0] * w[1] * w[2];
    const float value = values[corner];
    if (weight > 0.0f) {
      lower = value < lower ? value : lower;
      upper = value > upper ? value : upper;
    }
  }
  const float corrected = predicted + 0.5f * (original - reversed);
  const bool revert = corrected < lower || corrected > upper;
  return revert ? predicted : corrected;
}

__device__ __forceinline__ float occupancy(const AdGeo& g, const float* phi, const float* dens, int x, int y, int z) {
  if (x < 0 || y < 0 || z < 0 || x >= g.d[0] || y >= g.d[1] || z >= g.d[2]) return 0.0f;
  const unsigned i = mul24(mul24((unsigned)x, (unsigned)g.d[1]) + (unsigned)y, (unsigned)g.d[2]) + (unsigned)z;
  const float v = phi ? 0.5f - phi[i] / (4.0f * g.h) : dens[i];
  return v < 0.0f ? 0.0f : (v > 1.0f ? 1.0f : v);
}

// diffusionVelocity (…wgsl.ts:860): the cell's positive faces, clamped into
// the domain, reflected outside it for no-slip walls
__device__ __forceinline__ void diff_vel(const AdGeo& g, const Faces& cur, int x, int y, int z, float* out) {
  const bool inside = x >= 0 && y >= 0 && z >= 0 && x < g.d[0] && y < g.d[1] && z < g.d[2];
  const int cx = min(max(x, 0), g.d[0] - 1), cy = min(max(y, 0), g.d[1] - 1), cz = min(max(z, 0), g.d[2] - 1);
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const float v = cur.f[c][fidx(g, c, cx + (c == 0), cy + (c == 1), cz + (c == 2))];
    out[c] = g.no_slip && !inside ? -v : v;
  }
}

// applyVelocityForces (…wgsl.ts:872-915) without surface tension or inflow
__device__ void forces(const AdGeo& g, const Faces& cur, const float* phi, const float* dens, int x, int y, int z,
                       float* v) {
  const float occ = occupancy(g, phi, dens, x, y, z);
  if (occ > 0.0f && g.nu > 0.0f) {
    float ce[3], xp[3], xm[3], yp[3], ym[3], zp[3], zm[3];
    diff_vel(g, cur, x, y, z, ce);
    diff_vel(g, cur, x + 1, y, z, xp);
    diff_vel(g, cur, x - 1, y, z, xm);
    diff_vel(g, cur, x, y + 1, z, yp);
    diff_vel(g, cur, x, y - 1, z, ym);
    diff_vel(g, cur, x, y, z + 1, zp);
    diff_vel(g, cur, x, y, z - 1, zm);
    const float h2 = g.h * g.h;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const float lap = ((xp[c] - 2.0f * ce[c] + xm[c]) / h2 + (yp[c] - 2.0f * ce[c] + ym[c]) / h2) +
                        (zp[c] - 2.0f * ce[c] + zm[c]) / h2;
      v[c] += g.dt * g.nu * lap;
    }
  }
  const float occ_y = occupancy(g, phi, dens, x, y + 1, z);
  if (occ > 1e-5f || occ_y > 1e-5f) v[1] += g.gravity * g.dt;
}

__device__ __forceinline__ bool cell_of(const AdGeo& g, int& x, int& y, int& z) {
  const unsigned i = blockIdx.x * kThreads + threadIdx.x;
  const unsigned n = (unsigned)g.d[0] * g.d[1] * g.d[2];
  if (i >= n) return false;
  z = (int)(i % (unsigned)g.d[2]);
  const unsigned t = i / (unsigned)g.d[2];
  y = (int)(t % (unsigned)g.d[1]);
  x = (int)(t / (unsigned)g.d[1]);
  return true;
}

// the negative domain faces carry over (carryBoundaryVelocity, …wgsl.ts:196)
__device__ __forceinline__ void carry_walls(const AdGeo& g, const Faces& cur, float* const* out, int x, int y, int z) {
  if (x == 0) out[0][fidx(g, 0, 0, y, z)] = cur.f[0][fidx(g, 0, 0, y, z)];
  if (y == 0) out[1][fidx(g, 1, x, 0, z)] = cur.f[1][fidx(g, 1, x, 0, z)];
  if (z == 0) out[2][fidx(g, 2, x, y, 0)] = cur.f[2][fidx(g, 2, x, y, 0)];
}

// semiLagrangianAdvection (…wgsl.ts:922-936)
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(1, 4))) void k_adv_sl(AdGeo g, Faces t, Faces cur, const float* __restrict__ phi,
                                                    const float* __restrict__ dens, float* uo, float* vo, float* wo) {
  int x, y, z;
  if (!cell_of(g, x, y, z)) return;
  float* out[3] = {uo, vo, wo};
  const int p[3] = {x, y, z};
  float val[3];
  advect3(g, t, x, y, z, g.dt, val);
  for (int c = 0; c < 3; ++c) {
    if (p[c] == g.d[c] - 1 && !(c == 1 && g.open_top)) {
      const float old = cur.f[c][fidx(g, c, x + (c == 0), y + (c == 1), z + (c == 2))];
      val[c] = val[c] < old ? val[c] : old;
    }
  }
  forces(g, cur, phi, dens, x, y, z, val);
  for (int c = 0; c < 3; ++c) out[c][fidx(g, c, x + (c == 0), y + (c == 1), z + (c == 2))] = val[c];
  carry_walls(g, cur, out, x, y, z);
}

// advect (prediction, …wgsl.ts:938-952) and reverseAdvection (:954-961):
// the wall faces keep the field the trace starts from
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(1, 4))) void k_adv_trace(AdGeo g, Faces t, Faces walls, float dt, float* uo, float* vo,
                                                       float* wo) {
  int x, y, z;
  if (!cell_of(g, x, y, z)) return;
  float* out[3] = {uo, vo, wo};
  const int p[3] = {x, y, z};
  float adv[3];
  advect3(g, t, x, y, z, dt, adv);
  for (int c = 0; c < 3; ++c) {
    const unsigned fi = fidx(g, c, x + (c == 0), y + (c == 1), z + (c == 2));
    float val = adv[c];
    if (p[c] == g.d[c] - 1 && !(c == 1 && g.open_top)) val = walls.f[c][fi];
    out[c][fi] = val;
  }
  carry_walls(g, walls, out, x, y, z);
}

// correctAdvection (…wgsl.ts:993-1004)
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(1, 4))) void k_adv_correct(AdGeo g, Faces t, Faces cur, Faces pred, Faces rev,
                                                         const float* __restrict__ phi, const float* __restrict__ dens,
                                                         float* uo, float* vo, float* wo) {
  int x, y, z;
  if (!cell_of(g, x, y, z)) return;
  float* out[3] = {uo, vo, wo};
  float val[3], pos[3][3], dep[3][3];
  for (int c = 0; c < 3; ++c) face_position(x, y, z, c, pos[c]);
  departure3(g, t, pos, g.dt, dep);
  for (int c = 0; c < 3; ++c) {
    const unsigned fi = fidx(g, c, x + (c == 0), y + (c == 1), z + (c == 2));
    val[c] = bounded(g, t, dep[c], c, pred.f[c][fi], cur.f[c][fi], rev.f[c][fi]);
  }
  forces(g, cur, phi, dens, x, y, z, val);
  for (int c = 0; c < 3; ++c) out[c][fidx(g, c, x + (c == 0), y + (c == 1), z + (c == 2))] = val[c];
  carry_walls(g, cur, out, x, y, z);
}

}  // namespace

Advector::~Advector() {
  for (void* p : allocations_) (void)hipFree(p);
}

int Advector::alloc(void** p, size_t bytes) {
  bytes = (bytes + 255) & ~(size_t)255;
  if (hipMalloc(p, bytes) != hipSuccess) {
    set_error("velocity advection: hipMalloc(%zu) failed", bytes);
    return FP_ERR_OUT_OF_MEMORY;
  }
  allocations_.push_back(*p);
  bytes_ += bytes;
  return FP_OK;
}

size_t Advector::estimate_bytes(int nx, int ny, int nz) {
  const size_t faces = (size_t)(nx + 1) * ny * nz + (size_t)nx * (ny + 1) * nz + (size_t)nx * ny * (nz + 1);
  return 5 * ((faces * 4 + 255) & ~(size_t)255);
}

int Advector::create(int nx, int ny, int nz, double h) {
  // the paired z loads need 2 faces per row; byte offsets of 32 bits
  if (nx < 2 || ny < 2 || nz < 2) return FP_ERR_UNSUPPORTED;
  const size_t most = (size_t)(std::max(nx, std::max(ny, nz)) + 1) * nx * ny * nz / std::min(nx, std::min(ny, nz));
  if (most >= ((size_t)1 << 30) || (size_t)(nx + 1) * (ny + 1) >= ((size_t)1 << 24)) {
    set_error("velocity advection: %d x %d x %d has 2^30 or more faces or 2^24 or more rows per component", nx, ny,
              nz);
    return FP_ERR_UNSUPPORTED;
  }
  d_[0] = nx;
  d_[1] = ny;
  d_[2] = nz;
  h_ = (float)h;
  nf_[0] = (size_t)(nx + 1) * ny * nz;
  nf_[1] = (size_t)nx * (ny + 1) * nz;
  nf_[2] = (size_t)nx * ny * (nz + 1);
  const size_t faces = nf_[0] + nf_[1] + nf_[2];
  int rc;
  if ((rc = alloc((void**)&transport_, faces * 4))) return rc;
  if ((rc = alloc((void**)&pred_, faces * 4))) return rc;
  if ((rc = alloc((void**)&rev_, faces * 4))) return rc;
  if ((rc = alloc((void**)&pred_ext_, faces * 4))) return rc;
  if ((rc = alloc((void**)&out_, faces * 4))) return rc;
  return FP_OK;
}

int Advector::stage_host(size_t floats) {
  if (host_floats_ >= floats) return FP_OK;
  int rc = alloc((void**)&host_, floats * 4);
  if (!rc) host_floats_ = floats;
  return rc;
}

int Advector::run(hipStream_t s, Extrapolator* ex, const fp_advect_args* a, fp_advect_report* rep) {
  const auto t0 = std::chrono::steady_clock::now();
  const int nx = d_[0], ny = d_[1], nz = d_[2];
  const size_t n = (size_t)nx * ny * nz;
  const bool host = a->memory == FP_MEMORY_HOST;
  const float* field = a->phi ? a->phi : a->density;
  const float* cur_in[3] = {a->u, a->v, a->w};
  const float* tr_in[3] = {a->tu, a->tv, a->tw};
  const bool have_transport = a->tu && a->tv && a->tw;
  // inputs on the device
  const float* f = field;
  const float* cur[3] = {cur_in[0], cur_in[1], cur_in[2]};
  const float* tr[3] = {tr_in[0], tr_in[1], tr_in[2]};
  if (host) {
    const size_t need = n + 2 * (nf_[0] + nf_[1] + nf_[2]);
    int rc = stage_host(need);
    if (rc) return rc;
    float* p = host_;
    HIPCHK(hipMemcpyAsync(p, field, n * 4, hipMemcpyHostToDevice, s));
    f = p;
    p += n;
    for (int c = 0; c < 3; ++c) {
      HIPCHK(hipMemcpyAsync(p, cur_in[c], nf_[c] * 4, hipMemcpyHostToDevice, s));
      cur[c] = p;
      p += nf_[c];
    }
    for (int c = 0; c < 3; ++c) {
      if (have_transport) HIPCHK(hipMemcpyAsync(p, tr_in[c], nf_[c] * 4, hipMemcpyHostToDevice, s));
      tr[c] = p;
      p += nf_[c];
    }
  }
  float* split[4][3];
  float* bases[4] = {transport_, pred_, rev_, pred_ext_};
  for (int k = 0; k < 4; ++k) {
    split[k][0] = bases[k];
    split[k][1] = bases[k] + nf_[0];
    split[k][2] = bases[k] + nf_[0] + nf_[1];
  }
  const int open_top = a->open_top ? 1 : 0;
  if (!have_transport) {
    fp_extrapolate_args e;
    std::memset(&e, 0, sizeof(e));
    e.memory = FP_MEMORY_DEVICE;
    e.open_top = open_top;
    e.phi = a->phi ? f : nullptr;
    e.density = a->phi ? nullptr : f;
    e.u = cur[0];
    e.v = cur[1];
    e.w = cur[2];
    e.u_out = split[0][0];
    e.v_out = split[0][1];
    e.w_out = split[0][2];
    int rc = ex->run(s, &e, nullptr);
    if (rc) return rc;
    for (int c = 0; c < 3; ++c) tr[c] = split[0][c];
  }
  // outputs: written directly unless host memory or an overlap with an input
  float* out[3] = {a->u_out, a->v_out, a->w_out};
  float* dst[3];
  float* stage[3] = {out_, out_ + nf_[0], out_ + nf_[0] + nf_[1]};
  for (int c = 0; c < 3; ++c) {
    bool overlaps = host;
    const size_t ob = nf_[c] * 4;
    const void* ins[7] = {field, cur_in[0], cur_in[1], cur_in[2], tr_in[0], tr_in[1], tr_in[2]};
    const size_t ib[7] = {n * 4, nf_[0] * 4, nf_[1] * 4, nf_[2] * 4, nf_[0] * 4, nf_[1] * 4, nf_[2] * 4};
    for (int k = 0; k < 7 && !overlaps; ++k) {
      if (!ins[k]) continue;
      const char *o0 = (const char*)out[c], *i0 = (const char*)ins[k];
      overlaps = o0 < i0 + ib[k] && i0 < o0 + ob;
    }
    dst[c] = overlaps ? stage[c] : out[c];
  }
  AdGeo g;
  g.d[0] = nx;
  g.d[1] = ny;
  g.d[2] = nz;
  g.h = h_;
  g.open_top = open_top;
  g.no_slip = a->no_slip ? 1 : 0;
  g.dt = (float)a->dt;
  g.gravity = (float)a->gravity;
  g.nu = (float)a->viscosity;
  const Faces T{{tr[0], tr[1], tr[2]}}, C{{cur[0], cur[1], cur[2]}};
  const float* phi = a->phi ? f : nullptr;
  const float* dens = a->phi ? nullptr : f;
  const dim3 grid((unsigned)((n + kThreads - 1) / kThreads)), block(kThreads);
  if (a->scheme == FP_ADVECT_SEMI_LAGRANGIAN) {
    hipLaunchKernelGGL(k_adv_sl, grid, block, 0, s, g, T, C, phi, dens, dst[0], dst[1], dst[2]);
  } else {
    hipLaunchKernelGGL(k_adv_trace, grid, block, 0, s, g, T, C, g.dt, split[1][0], split[1][1], split[1][2]);
    fp_extrapolate_args e;
    std::memset(&e, 0, sizeof(e));
    e.memory = FP_MEMORY_DEVICE;
    e.open_top = open_top;
    e.phi = phi;
    e.density = dens;
    e.u = split[1][0];
    e.v = split[1][1];
    e.w = split[1][2];
    e.u_out = split[3][0];
    e.v_out = split[3][1];
    e.w_out = split[3][2];
    // the same surface as the transport's extrapolation when this call made it
    int rc = ex->run(s, &e, nullptr, !have_transport);
    if (rc) return rc;
    const Faces P{{split[1][0], split[1][1], split[1][2]}}, PE{{split[3][0], split[3][1], split[3][2]}};
    const Faces R{{split[2][0], split[2][1], split[2][2]}};
    hipLaunchKernelGGL(k_adv_trace, grid, block, 0, s, g, PE, P, -g.dt, split[2][0], split[2][1], split[2][2]);
    hipLaunchKernelGGL(k_adv_correct, grid, block, 0, s, g, T, C, P, R, phi, dens, dst[0], dst[1], dst[2]);
  }
  for (int c = 0; c < 3; ++c)
    if (dst[c] != out[c])
      HIPCHK(hipMemcpyAsync(out[c], dst[c], nf_[c] * 4, host ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice, s));
  HIPCHK(hipStreamSynchronize(s));
  if (rep) {
    std::memset(rep, 0, sizeof(*rep));
    rep->device_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  }
  return FP_OK;
}

}  // namespace fp

// fp_cm11a.hip -- the CM11a separating-wall LCP pressure solve on MI355X.
//
// Kernels restate lib/methods/uniform/webgpu-uniform-pressure-multigrid.wgsl.ts
// (cited per kernel) and the schedule of …multigrid.ts:488-576.  Every level
// is a halo-padded FP32 lattice in HBM, z fastest, with packed float4
// dual volumes and coefficients (one 16-byte load per cell).  Red-black
// passes run in place: a pass reads only the other colour and writes only its
// own, which is what the reference's ping-pong textures compute.  The whole
// file is compiled without floating-point contraction, and sums follow the
// reference's D4 order, so results equal oracle/fp_oracle_cm11a.c bit for bit.
#pragma clang fp contract(off)

#include "fp_cm11a.hpp"
#include "fp_common.hpp"
#include "fp_uniform.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>

namespace fp {

namespace {

constexpr float kNegBig = -3.402823e38f;  // the reference's "no constraint" p_min
constexpr float kThetaMin = 0.05f;
constexpr int kCoarseLanes = 1024;        // most padded cells of the coarsest level

__device__ __forceinline__ float d4s6(const float* t) { return ((t[0] + t[1]) + (t[4] + t[5])) + (t[2] + t[3]); }
__device__ __forceinline__ float d4s8(const float* t) {
  const float y0 = (t[0] + t[5]) + (t[1] + t[4]);
  const float y1 = (t[2] + t[7]) + (t[3] + t[6]);
  return y0 + y1;
}
__device__ __forceinline__ float cm_theta(float l, float a, float eps) {
  const float r = fabsf(l) / fmaxf(fabsf(l) + fabsf(a), eps);
  return fminf(1.0f, fmaxf(kThetaMin, r));
}
__constant__ int kNb[6][3] = {{-1, 0, 0}, {1, 0, 0}, {0, -1, 0}, {0, 1, 0}, {0, 0, -1}, {0, 0, 1}};

__device__ __forceinline__ int cidx(const CmLevel& L, int x, int y, int z) { return (x * L.D1 + y) * L.D2 + z; }
__device__ __forceinline__ bool cvalid(const CmLevel& L, int x, int y, int z) {
  return x >= 0 && y >= 0 && z >= 0 && x < L.D0 && y < L.D1 && z < L.D2;
}
__device__ __forceinline__ bool cinterior(const CmLevel& L, int x, int y, int z) {
  return x >= 1 && y >= 1 && z >= 1 && x < L.D0 - 1 && y < L.D1 - 1 && z < L.D2 - 1;
}
__device__ __forceinline__ void cdecode(const CmLevel& L, int i, int& x, int& y, int& z) {
  z = i % L.D2;
  const int t = i / L.D2;
  y = t % L.D1;
  x = t / L.D1;
}
__device__ __forceinline__ float hof(const CmLevel& L, int a) { return a == 0 ? L.h0 : (a == 1 ? L.h1 : L.h2); }
__device__ __forceinline__ float comp(const float4& v, int k) {
  return k == 0 ? v.x : (k == 1 ? v.y : (k == 2 ? v.z : v.w));
}

__device__ __forceinline__ int wave_sum_int(int v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  return v;
}

#define CM_FOR(i, n) for (int i = (int)(blockIdx.x * blockDim.x + threadIdx.x); i < (n); i += (int)(gridDim.x * blockDim.x))

// ---- double-single arithmetic (…multigrid.wgsl.ts:364-381) ----
struct DS {
  float x, y;
};
__device__ __forceinline__ DS two_sum(float a, float b) {
  const float s = a + b;
  const float bb = s - a;
  return DS{s, (a - (s - bb)) + (b - bb)};
}
__device__ __forceinline__ DS ds_add(DS a, DS b) {
  const DS s = two_sum(a.x, b.x), t = two_sum(a.y, b.y), u = two_sum(s.y, t.x), v = two_sum(s.x, u.x);
  return DS{v.x, (v.y + u.y) + t.y};
}
__device__ __forceinline__ DS ds_neg(DS a) { return DS{-a.x, -a.y}; }
__device__ __forceinline__ DS ds_scale(DS a, float b) {
  const float product = a.x * b;
  const float error = __fmaf_rn(a.x, b, -product) + a.y * b;
  return two_sum(product, error);
}
__device__ __forceinline__ DS ds_divide(DS a, float b) {
  const float q = a.x / b;
  const DS remainder = ds_add(a, ds_neg(ds_scale(DS{q, 0.0f}, b)));
  const float correction = (remainder.x + remainder.y) / b;
  return two_sum(q, correction);
}
__device__ __forceinline__ DS d4s6_ds(const DS* v) {
  return ds_add(ds_add(ds_add(v[0], v[1]), ds_add(v[4], v[5])), ds_add(v[2], v[3]));
}

// dual volume of the positive face of an interior finest cell
// (webgpu-uniform-reference.wgsl.ts:536-561): 1 inside, 1/2 at a box wall, 1 at the open top
__device__ __forceinline__ float face_volume(int n, int c, int axis, int open_top) {
  if (c + 1 < n) return 1.0f;
  if (axis == 1 && open_top) return 1.0f;
  return 0.5f;
}

// mgBuildFinestTopology + mgBuildFinestRhs (…multigrid.wgsl.ts:135-167), the
// divergence of divergenceAt (webgpu-uniform-reference.wgsl.ts:1010-1027)
__global__ void k_cm_finest(CmLevel L, int nx, int ny, int nz, const float* __restrict__ phi,
                            const float* __restrict__ u, const float* __restrict__ v, const float* __restrict__ w,
                            const float* __restrict__ q, float dt, float rho, int open_top, float* __restrict__ b,
                            float* __restrict__ pmin, CmStats* stats) {
  const float hmin = fminf(L.h0, fminf(L.h1, L.h2));
  int liquid = 0;
  CM_FOR(i, L.n) {
    int X, Y, Z;
    cdecode(L, i, X, Y, Z);
    L.p[i] = 0.0f;
    if (cinterior(L, X, Y, Z)) {
      const int x = X - 1, y = Y - 1, z = Z - 1;
      const size_t s = ((size_t)x * ny + y) * nz + z;
      const float ph = phi[s];
      L.topo[i] = make_float4(1.0f, face_volume(nx, x, 0, open_top), face_volume(ny, y, 1, open_top),
                              face_volume(nz, z, 2, open_top));
      L.phi_raw[i] = ph;
      pmin[i] = kNegBig;
      float rhs = 0.0f;
      if (ph < 0.0f) {
        ++liquid;
        const float vi = 1.0f;
        const int c[3] = {x, y, z}, n[3] = {nx, ny, nz};
        float terms[6];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
          const float vp = face_volume(n[a], c[a], a, open_top);
          const float vm = c[a] == 0 ? 0.5f : 1.0f;
          float up, um;
          if (a == 0) {
            up = u[((size_t)(x + 1) * ny + y) * nz + z];
            um = u[((size_t)x * ny + y) * nz + z];
          } else if (a == 1) {
            up = v[((size_t)x * (ny + 1) + y + 1) * nz + z];
            um = v[((size_t)x * (ny + 1) + y) * nz + z];
          } else {
            up = w[((size_t)x * ny + y) * (nz + 1) + z + 1];
            um = w[((size_t)x * ny + y) * (nz + 1) + z];
          }
          const float h = hof(L, a);
          terms[2 * a] = (vp * up) / h + (vp - vi) * 0.0f;
          terms[2 * a + 1] = -(vm * um) / h - (vm - vi) * 0.0f;
        }
        const float div = d4s6(terms);
        const float corr = q ? q[s] : 0.0f;
        rhs = rho * (div - corr) / dt;
      }
      b[i] = -rhs;
    } else {
      const bool xi = X > 0 && X < L.D0 - 1, yi = Y > 0 && Y < L.D1 - 1, zi = Z > 0 && Z < L.D2 - 1;
      const bool open = open_top && Y == L.D1 - 1 && xi && zi;
      const float o = open ? 1.0f : 0.0f;
      float4 t = make_float4(o, o, o, o);
      if (X == 0 && yi && zi) t.y = 0.5f;
      if (Y == 0 && xi && zi) t.z = 0.5f;
      if (Z == 0 && xi && yi) t.w = 0.5f;
      L.topo[i] = t;
      L.phi_raw[i] = 0.5f * hmin;
      b[i] = 0.0f;
      pmin[i] = open ? kNegBig : 0.0f;
    }
  }
  liquid = wave_sum_int(liquid);
  if ((threadIdx.x & 63) == 0 && liquid) atomicAdd(&stats->liquid, liquid);
}

// mgFineChild (…multigrid.wgsl.ts:71-76)
__device__ __forceinline__ int fine_child(const CmLevel& F, const CmLevel& C, int X, int Y, int Z, int o) {
  const int c[3] = {X, Y, Z}, fd[3] = {F.D0, F.D1, F.D2}, cd[3] = {C.D0, C.D1, C.D2};
  int out[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const int fp = fd[a] - 2;
    const int cp = max(cd[a] - 2, 1);
    const int stride = max(fp / cp, 1);
    const int oa = (o >> a) & 1;
    out[a] = min(max(stride * (c[a] - 1) + oa * (stride - 1), -1), fp) + 1;
  }
  return cidx(F, out[0], out[1], out[2]);
}

// mgDownsampleTopology (…multigrid.wgsl.ts:169-195)
__global__ void k_cm_down_topology(CmLevel F, CmLevel C, int dest, int levels, int preserve, int open_top) {
  CM_FOR(i, C.n) {
    int X, Y, Z;
    cdecode(C, i, X, Y, Z);
    float t0[8], t1[8], t2[8], t3[8], ph[8], pos[8], pf[8], nf[8];
#pragma unroll
    for (int o = 0; o < 8; ++o) {
      const int j = fine_child(F, C, X, Y, Z, o);
      const float4 t = F.topo[j];
      t0[o] = t.x;
      t1[o] = t.y;
      t2[o] = t.z;
      t3[o] = t.w;
      const float phi = F.phi_raw[j];
      ph[o] = phi;
      pos[o] = phi >= 0.0f ? phi : 0.0f;
      pf[o] = phi >= 0.0f ? 1.0f : 0.0f;
      nf[o] = phi >= 0.0f ? 0.0f : 1.0f;
    }
    const float phi_sum = d4s8(ph), pos_sum = d4s8(pos), pc = d4s8(pf), nc = d4s8(nf);
    const bool mixed = pc > 0.0f && nc > 0.0f;
    const bool use_positive = mixed && (levels - dest) >= (levels - preserve);
    float4 t = make_float4(d4s8(t0) / 8.0f, d4s8(t1) / 8.0f, d4s8(t2) / 8.0f, d4s8(t3) / 8.0f);
    if (X == C.D0 - 2) t.y = 0.5f;
    if (Y == C.D1 - 2) t.z = open_top ? 1.0f : 0.5f;
    if (Z == C.D2 - 2) t.w = 0.5f;
    C.topo[i] = t;
    C.phi_raw[i] = use_positive ? pos_sum / fmaxf(pc, 1.0f) : phi_sum / 8.0f;
  }
}

// mgExtrapolatePhiOneCell (…multigrid.wgsl.ts:288-300)
__global__ void k_cm_extrapolate(CmLevel L) {
  CM_FOR(i, L.n) {
    if (L.topo[i].x > 1e-5f) {
      L.phi[i] = L.phi_raw[i];
      continue;
    }
    int X, Y, Z;
    cdecode(L, i, X, Y, Z);
    float terms[6], wts[6];
#pragma unroll
    for (int nb = 0; nb < 6; ++nb) {
      const int x = X + kNb[nb][0], y = Y + kNb[nb][1], z = Z + kNb[nb][2];
      terms[nb] = 0.0f;
      wts[nb] = 0.0f;
      if (!cvalid(L, x, y, z)) continue;
      const int j = cidx(L, x, y, z);
      const float vq = L.topo[j].x, pq = L.phi_raw[j];
      if (vq > 1e-5f && pq < 0.0f) {
        terms[nb] = vq * pq;
        wts[nb] = vq;
      }
    }
    const float sum = d4s6(terms), weight = d4s6(wts);
    L.phi[i] = weight > 0.0f ? sum / fmaxf(weight, 1e-9f) : L.phi_raw[i];
  }
}

// mgCoefficientRaw (…multigrid.wgsl.ts:93-115), +axis face
__device__ __forceinline__ float coefficient_raw(const CmLevel& L, int i, int X, int Y, int Z, int axis,
                                                 int open_top) {
  int q[3] = {X, Y, Z};
  q[axis] += 1;
  const float h = hof(L, axis);
  if (!cvalid(L, q[0], q[1], q[2])) {
    if (axis == 1 && Y == L.D1 - 1 && open_top) {
      const float theta = cm_theta(L.phi[i], 0.5f * h, 1e-9f);
      return L.topo[i].z / (h * h * theta);
    }
    return 0.0f;
  }
  const float vf = comp(L.topo[i], 1 + axis);
  if (vf <= 1e-6f) return 0.0f;
  const int j = cidx(L, q[0], q[1], q[2]);
  const float pi = L.phi[i], pj = L.phi[j];
  const bool il = pi < 0.0f, ql = pj < 0.0f;
  float theta = 1.0f;
  if (il && !ql) theta = cm_theta(pi, pj, 1e-9f);
  if (!il && ql) theta = cm_theta(pj, pi, 1e-9f);
  return vf / (h * h * theta);
}

// mgBakeCoefficients (…multigrid.wgsl.ts:302-309)
__global__ void k_cm_bake(CmLevel L, int open_top) {
  CM_FOR(i, L.n) {
    int X, Y, Z;
    cdecode(L, i, X, Y, Z);
    L.coef[i] = make_float4(coefficient_raw(L, i, X, Y, Z, 0, open_top), coefficient_raw(L, i, X, Y, Z, 1, open_top),
                            coefficient_raw(L, i, X, Y, Z, 2, open_top), L.phi[i] < 0.0f ? 1.0f : 0.0f);
  }
}

// The cell's six coefficients (mgCoefficient, …multigrid.wgsl.ts:116-122),
// neighbour flags (mgBakedLiquid) and indices
struct Stencil {
  float a[6];
  int j[6];
  bool liq[6];
};
__device__ __forceinline__ void stencil(const CmLevel& L, int i, int X, int Y, int Z, const float4& own,
                                        Stencil& s) {
#pragma unroll
  for (int nb = 0; nb < 6; ++nb) {
    const int axis = nb / 2, x = X + kNb[nb][0], y = Y + kNb[nb][1], z = Z + kNb[nb][2];
    const bool in = cvalid(L, x, y, z);
    const int j = in ? cidx(L, x, y, z) : i;
    s.j[nb] = j;
    if (nb & 1) {
      s.a[nb] = comp(own, axis);
      s.liq[nb] = in && L.coef[j].w > 0.5f;
    } else {
      const float4 cq = in ? L.coef[j] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
      s.a[nb] = comp(cq, axis);
      s.liq[nb] = in && cq.w > 0.5f;
    }
  }
}
__device__ __forceinline__ float apply_at(const float* __restrict__ p, int i, const Stencil& s) {
  const float centre = p[i];
  float terms[6];
#pragma unroll
  for (int nb = 0; nb < 6; ++nb) terms[nb] = s.a[nb] * (centre - (s.liq[nb] ? p[s.j[nb]] : 0.0f));
  return d4s6(terms);
}

// mgSmoothColour (…multigrid.wgsl.ts:325-344): the cells of one colour, in place
__global__ void k_cm_smooth(CmLevel L, const float* __restrict__ rhs, const float* __restrict__ pmin, int colour) {
  CM_FOR(i, L.n) {
    int X, Y, Z;
    cdecode(L, i, X, Y, Z);
    if (((X + Y + Z) & 1) != colour) continue;
    const float4 own = L.coef[i];
    if (!(own.w > 0.5f)) {
      L.p[i] = fmaxf(L.p[i], pmin[i]);
      continue;
    }
    Stencil s;
    stencil(L, i, X, Y, Z, own, s);
    float st[6];
#pragma unroll
    for (int nb = 0; nb < 6; ++nb) st[nb] = s.liq[nb] ? s.a[nb] * L.p[s.j[nb]] : 0.0f;
    const float diag = d4s6(s.a), sum = d4s6(st);
    const float pn = diag > 0.0f ? (sum + rhs[i]) / diag : 0.0f;
    L.p[i] = fmaxf(pn, pmin[i]);
  }
}

// mgResidual (…multigrid.wgsl.ts:197-205)
__global__ void k_cm_residual(CmLevel L, const float* __restrict__ rhs, float* __restrict__ out) {
  CM_FOR(i, L.n) {
    const float4 own = L.coef[i];
    if (!(own.w > 0.5f)) {
      out[i] = 0.0f;
      continue;
    }
    int X, Y, Z;
    cdecode(L, i, X, Y, Z);
    Stencil s;
    stencil(L, i, X, Y, Z, own, s);
    out[i] = rhs[i] - apply_at(L.p, i, s);
  }
}

// mgRestrictResidual (…multigrid.wgsl.ts:210-215)
__global__ void k_cm_restrict(CmLevel F, const float* __restrict__ in, CmLevel C, float* __restrict__ out) {
  CM_FOR(i, C.n) {
    int X, Y, Z;
    cdecode(C, i, X, Y, Z);
    float t[8];
#pragma unroll
    for (int o = 0; o < 8; ++o) t[o] = in[fine_child(F, C, X, Y, Z, o)];
    out[i] = d4s8(t) / 8.0f;
  }
}

// mgTrilinearPressure + mgProlongateAdd / Assign (…multigrid.wgsl.ts:217-253)
__global__ void k_cm_prolong(CmLevel C, CmLevel F, int add) {
  CM_FOR(i, F.n) {
    int id[3];
    cdecode(F, i, id[0], id[1], id[2]);
    const int fd[3] = {F.D0, F.D1, F.D2}, cd[3] = {C.D0, C.D1, C.D2};
    float f[3];
    int base[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const float fine_phys = (float)(fd[a] - 2), coarse_phys = (float)(cd[a] - 2);
      const float scale = coarse_phys / fmaxf(fine_phys, 1.0f);
      const float q = ((float)id[a] - 0.5f) * scale + 0.5f;
      const float fl = floorf(q);
      base[a] = (int)fl;
      f[a] = q - fl;
    }
    float values[8], weights[8];
#pragma unroll
    for (int o = 0; o < 8; ++o) {
      const int ox = o & 1, oy = (o >> 1) & 1, oz = (o >> 2) & 1;
      const int px = base[0] + ox, py = base[1] + oy, pz = base[2] + oz;
      values[o] = 0.0f;
      weights[o] = 0.0f;
      if (!cinterior(C, px, py, pz)) continue;
      const float w = ((ox ? f[0] : 1.0f - f[0]) * (oy ? f[1] : 1.0f - f[1])) * (oz ? f[2] : 1.0f - f[2]);
      values[o] = w * C.p[cidx(C, px, py, pz)];
      weights[o] = w;
    }
    const float value = d4s8(values), total = d4s8(weights);
    const float t = total > 0.0f ? value / total : 0.0f;
    F.p[i] = add ? F.p[i] + t : t;
  }
}

// mgDownsampleSubtract / mgDownsampleMinimum (…multigrid.wgsl.ts:311-323)
__global__ void k_cm_down_min(CmLevel F, const float* __restrict__ fmin, int subtract, CmLevel C) {
  CM_FOR(i, C.n) {
    int X, Y, Z;
    cdecode(C, i, X, Y, Z);
    float lower = kNegBig;
#pragma unroll
    for (int o = 0; o < 8; ++o) {
      const int j = fine_child(F, C, X, Y, Z, o);
      lower = fmaxf(lower, subtract ? fmin[j] - F.p[j] : fmin[j]);
    }
    C.pmin[i] = lower;
  }
}

__global__ void k_cm_fill(float* __restrict__ a, int n, float v) {
  CM_FOR(i, n) a[i] = v;
}
// mgCopyPressure + mgShiftMinimum at the start of a full cycle
__global__ void k_cm_backup_shift(const float* __restrict__ p, const float* __restrict__ pmin0,
                                  float* __restrict__ backup, float* __restrict__ shift, int n) {
  CM_FOR(i, n) {
    const float v = p[i];
    backup[i] = v;
    shift[i] = pmin0[i] - v;
  }
}
// mgAddPressure
__global__ void k_cm_add(float* __restrict__ p, const float* __restrict__ backup, int n) {
  CM_FOR(i, n) p[i] = p[i] + backup[i];
}

// mgCoarseCoefficient (…multigrid.wgsl.ts:384-397): from the coarse level set and dual volumes
__device__ __forceinline__ float coarse_coefficient(const CmLevel& L, int i, int X, int Y, int Z, int nb,
                                                    int open_top) {
  const int axis = nb / 2, x = X + kNb[nb][0], y = Y + kNb[nb][1], z = Z + kNb[nb][2];
  const float h = hof(L, axis);
  if (!cvalid(L, x, y, z)) {
    if (axis == 1 && Y == L.D1 - 1 && y == L.D1 && open_top) {
      const float theta = cm_theta(L.phi[i], 0.5f * h, 1e-9f);
      return L.topo[i].z / (h * h * theta);
    }
    return 0.0f;
  }
  const int j = cidx(L, x, y, z);
  const float vf = (nb & 1) ? comp(L.topo[i], 1 + axis) : comp(L.topo[j], 1 + axis);
  if (vf <= 1e-6f) return 0.0f;
  float theta = 1.0f;
  if (L.phi[j] >= 0.0f) theta = cm_theta(L.phi[i], L.phi[j], 1e-9f);
  return vf / (h * h * theta);
}

// mgSolveCoarsest (…multigrid.wgsl.ts:404-471): one workgroup, a lane per
// padded cell, double-single projected red-black sweeps in LDS until the
// projected residual (1/s) is at most tol, or cap sweeps.
__global__ __launch_bounds__(kCoarseLanes) void k_cm_coarse(CmLevel L, const float* __restrict__ rhs,
                                                            const float* __restrict__ pmin, float dt, float rho,
                                                            float tol, int cap, int open_top, CmStats* stats) {
  __shared__ float sx[kCoarseLanes], sy[kCoarseLanes];
  __shared__ unsigned s_max;
  const int lane = threadIdx.x;
  const bool live = lane < L.n;
  int X = 0, Y = 0, Z = 0;
  float a[6] = {0, 0, 0, 0, 0, 0};
  int j[6] = {0, 0, 0, 0, 0, 0};
  bool nl[6] = {false, false, false, false, false, false};
  bool liquid = false;
  float r = 0.0f, pm = 0.0f;
  int colour = 0;
  if (live) {
    cdecode(L, lane, X, Y, Z);
    colour = (X + Y + Z) & 1;
    liquid = L.phi[lane] < 0.0f;
    r = rhs[lane];
    pm = pmin[lane];
    sx[lane] = L.p[lane];
    sy[lane] = 0.0f;
    if (liquid)
#pragma unroll
      for (int nb = 0; nb < 6; ++nb) {
        a[nb] = coarse_coefficient(L, lane, X, Y, Z, nb, open_top);
        const int x = X + kNb[nb][0], y = Y + kNb[nb][1], z = Z + kNb[nb][2];
        if (cvalid(L, x, y, z)) {
          j[nb] = cidx(L, x, y, z);
          nl[nb] = L.phi[j[nb]] < 0.0f;
        }
      }
  }
  const float diag = d4s6(a);
  __syncthreads();
  int iterations = 0;
  bool converged = false;
  float worst = 0.0f;
  for (int it = 0; it < cap; ++it) {
    iterations = it + 1;
#pragma unroll
    for (int c = 0; c < 2; ++c) {
      if (live && liquid && colour == c) {
        DS st[6];
#pragma unroll
        for (int nb = 0; nb < 6; ++nb) st[nb] = nl[nb] ? ds_scale(DS{sx[j[nb]], sy[j[nb]]}, a[nb]) : DS{0.0f, 0.0f};
        const DS sum = d4s6_ds(st);
        if (diag > 0.0f) {
          const DS next = ds_divide(ds_add(sum, DS{r, 0.0f}), diag);
          if (next.x + next.y < pm) {
            sx[lane] = pm;
            sy[lane] = 0.0f;
          } else {
            sx[lane] = next.x;
            sy[lane] = next.y;
          }
        }
      }
      __syncthreads();
    }
    if (live && sx[lane] + sy[lane] < pm) {
      sx[lane] = pm;
      sy[lane] = 0.0f;
    }
    if (lane == 0) s_max = 0u;
    __syncthreads();
    if (live && liquid && diag > 0.0f) {
      const DS P{sx[lane], sy[lane]};
      DS ap[6];
#pragma unroll
      for (int nb = 0; nb < 6; ++nb) {
        const DS nv = nl[nb] ? DS{sx[j[nb]], sy[j[nb]]} : DS{0.0f, 0.0f};
        ap[nb] = ds_scale(ds_add(P, ds_neg(nv)), a[nb]);
      }
      const DS applied = d4s6_ds(ap);
      const DS lrp = ds_add(DS{r, 0.0f}, ds_neg(applied));
      const float lr = lrp.x + lrp.y;
      const float pressure = P.x + P.y;
      const float gap = fmaxf(0.0f, pressure - pm);
      const bool projects = lr < 0.0f && -lr >= gap * diag;
      const float lcp = projects ? gap * diag : fabsf(lr);
      const float divres = lcp * dt / rho;
      atomicMax(&s_max, __float_as_uint(divres));
    }
    __syncthreads();
    worst = __uint_as_float(s_max);
    if (worst <= tol) {
      converged = true;
      break;
    }
    __syncthreads();  // s_max is reset by the next iteration
  }
  if (live) L.p[lane] = sx[lane] + sy[lane];
  if (lane == 0) {
    stats->coarse_sweeps = max(stats->coarse_sweeps, iterations);
    stats->coarse_residual_bits = max(stats->coarse_residual_bits, __float_as_uint(worst));
    stats->coarse_failed += converged ? 0 : 1;
    stats->coarse_solves += 1;
  }
}

// mgMeasureFineResidual (…multigrid.wgsl.ts:473-481): max projected residual
// (1/s) and, per workgroup, the sum of squares of the projected residual and
// of b (the relative LCP residual of the report)
__global__ void k_cm_measure(CmLevel L, const float* __restrict__ b, const float* __restrict__ pmin, float dt,
                             float rho, CmStats* stats, double* partials) {
  double acc[2] = {0.0, 0.0};
  float mx = 0.0f;
  CM_FOR(i, L.n) {
    const float4 own = L.coef[i];
    if (!(own.w > 0.5f)) continue;
    int X, Y, Z;
    cdecode(L, i, X, Y, Z);
    Stencil s;
    stencil(L, i, X, Y, Z, own, s);
    const float diag = d4s6(s.a);
    if (diag <= 0.0f) continue;
    const float r = b[i] - apply_at(L.p, i, s);
    const float gap = fmaxf(0.0f, L.p[i] - pmin[i]);
    const bool projects = r < 0.0f && -r >= gap * diag;
    const float pr = projects ? gap * diag : fabsf(r);
    mx = fmaxf(mx, pr * dt / rho);
    acc[0] += (double)pr * (double)pr;
    acc[1] += (double)b[i] * (double)b[i];
  }
  mx = wave_max(mx);
  if ((threadIdx.x & 63) == 0) atomicMax(&stats->lcp_residual_bits, __float_as_uint(mx));
  block_sum<2>(acc);
  if (threadIdx.x == 0) {
    partials[2 * blockIdx.x] = acc[0];
    partials[2 * blockIdx.x + 1] = acc[1];
  }
}

// max |div u| over liquid cells with the dual-volume divergence (1/s)
__global__ void k_cm_max_div(int nx, int ny, int nz, float h, const float* __restrict__ phi,
                             const float* __restrict__ u, const float* __restrict__ v, const float* __restrict__ w,
                             int open_top, unsigned* out_bits) {
  const int n = nx * ny * nz;
  float mx = 0.0f;
  CM_FOR(s, n) {
    if (!(phi[s] < 0.0f)) continue;
    const int z = s % nz, t = s / nz, y = t % ny, x = t / ny;
    const int c[3] = {x, y, z}, nn[3] = {nx, ny, nz};
    float terms[6];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const float vp = face_volume(nn[a], c[a], a, open_top);
      const float vm = c[a] == 0 ? 0.5f : 1.0f;
      float up, um;
      if (a == 0) {
        up = u[((size_t)(x + 1) * ny + y) * nz + z];
        um = u[((size_t)x * ny + y) * nz + z];
      } else if (a == 1) {
        up = v[((size_t)x * (ny + 1) + y + 1) * nz + z];
        um = v[((size_t)x * (ny + 1) + y) * nz + z];
      } else {
        up = w[((size_t)x * ny + y) * (nz + 1) + z + 1];
        um = w[((size_t)x * ny + y) * (nz + 1) + z];
      }
      terms[2 * a] = (vp * up) / h;
      terms[2 * a + 1] = -(vm * um) / h;
    }
    mx = fmaxf(mx, fabsf(d4s6(terms)));
  }
  mx = wave_max(mx);
  if ((threadIdx.x & 63) == 0) atomicMax(out_bits, __float_as_uint(mx));
}

// project (webgpu-uniform-reference.wgsl.ts:1052-1099) on a box container:
// each face is written by the thread of the cell below it (and the low wall
// faces by the cells at index 0), so outputs may alias the inputs
__global__ void k_cm_project(CmLevel L, int nx, int ny, int nz, const float* __restrict__ phi, const float* uin,
                             const float* vin, const float* win, float* uo, float* vo, float* wo, float dt, float rho,
                             int open_top) {
  const float scale = dt / rho;
  const float hmin = fminf(L.h0, fminf(L.h1, L.h2));
  const int n = nx * ny * nz;
  CM_FOR(s, n) {
    const int z = s % nz, t = s / nz, y = t % ny, x = t / ny;
    const int c[3] = {x, y, z}, nn[3] = {nx, ny, nz};
    const float ph = phi[s];
    const bool liquid = ph < 0.0f;
    const float p0 = liquid ? L.p[cidx(L, x + 1, y + 1, z + 1)] : 0.0f;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const float* fin = a == 0 ? uin : (a == 1 ? vin : win);
      float* fo = a == 0 ? uo : (a == 1 ? vo : wo);
      if (!fo) continue;
      size_t lo, hi;
      if (a == 0) {
        lo = ((size_t)x * ny + y) * nz + z;
        hi = ((size_t)(x + 1) * ny + y) * nz + z;
      } else if (a == 1) {
        lo = ((size_t)x * (ny + 1) + y) * nz + z;
        hi = ((size_t)x * (ny + 1) + y + 1) * nz + z;
      } else {
        lo = ((size_t)x * ny + y) * (nz + 1) + z;
        hi = ((size_t)x * ny + y) * (nz + 1) + z + 1;
      }
      const float h = hof(L, a);
      if (c[a] == 0) {
        int hc[3] = {x + 1, y + 1, z + 1};
        hc[a] = 0;
        fo[lo] = liquid ? fin[lo] - scale * (p0 - L.p[cidx(L, hc[0], hc[1], hc[2])]) / h : 0.0f;
      }
      if (c[a] == nn[a] - 1) {
        if (a == 1 && open_top) {
          if (liquid) {
            const float theta = cm_theta(ph, 0.5f * hmin, 1e-6f);
            fo[hi] = fin[hi] - scale * (0.0f - p0) / (h * theta);
          } else {
            fo[hi] = 0.0f;
          }
        } else if (liquid) {
          int hc[3] = {x + 1, y + 1, z + 1};
          hc[a] = nn[a] + 1;
          fo[hi] = fin[hi] - scale * (L.p[cidx(L, hc[0], hc[1], hc[2])] - p0) / h;
        } else {
          fo[hi] = 0.0f;
        }
        continue;
      }
      int d[3] = {x, y, z};
      d[a] += 1;
      const size_t sn = ((size_t)d[0] * ny + d[1]) * nz + d[2];
      const float pn = phi[sn];
      const bool nlq = pn < 0.0f;
      if (liquid || nlq) {
        const float p1 = nlq ? L.p[cidx(L, d[0] + 1, d[1] + 1, d[2] + 1)] : 0.0f;
        float theta = 1.0f;
        if (liquid && !nlq) theta = cm_theta(ph, pn, 1e-6f);
        if (!liquid && nlq) theta = cm_theta(pn, ph, 1e-6f);
        fo[hi] = fin[hi] - scale * (p1 - p0) / (h * theta);
      } else {
        fo[hi] = 0.0f;
      }
    }
  }
}

// interior pressure of the liquid cells (0 elsewhere) into the caller's array
__global__ void k_cm_pressure_out(CmLevel L, int nx, int ny, int nz, const float* __restrict__ phi,
                                  float* __restrict__ out) {
  const int n = nx * ny * nz;
  CM_FOR(s, n) {
    const int z = s % nz, t = s / nz, y = t % ny, x = t / ny;
    out[s] = phi[s] < 0.0f ? L.p[cidx(L, x + 1, y + 1, z + 1)] : 0.0f;
  }
}

int grid_for(long long n) { return (int)std::min<long long>(std::max<long long>((n + 255) / 256, 1), 65536); }
constexpr int kMeasureBlocks = 1024;

}  // namespace

#define CM_LAUNCH(kern, n, s, ...) hipLaunchKernelGGL(kern, dim3(grid_for(n)), dim3(256), 0, s, ##__VA_ARGS__)

Cm11a::~Cm11a() {
  for (void* p : allocations_) (void)hipFree(p);
}

int Cm11a::alloc(void** p, size_t bytes) {
  bytes = (bytes + 255) & ~(size_t)255;
  if (hipMalloc(p, bytes) != hipSuccess) {
    set_error("CM11a: hipMalloc(%zu) failed", bytes);
    return FP_ERR_OUT_OF_MEMORY;
  }
  allocations_.push_back(*p);
  bytes_ += bytes;
  return FP_OK;
}

size_t Cm11a::estimate_bytes(int nx, int ny, int nz) {
  int dims[kMaxLevels][3], strides[kMaxLevels][3];
  const int L = plan_uniform_levels(nx, ny, nz, dims, strides);
  size_t total = 0;
  auto up = [](size_t b) { return (b + 255) & ~(size_t)255; };
  for (int l = 0; l < L; ++l) {
    const size_t n = (size_t)(dims[l][0] + 2) * (dims[l][1] + 2) * (dims[l][2] + 2);
    total += 7 * up(n * 4) + 2 * up(n * 16);
    if (l == 0) total += 4 * up(n * 4);
  }
  return total + up((size_t)kMeasureBlocks * 16) + up(sizeof(CmStats));
}

int Cm11a::create(int nx, int ny, int nz, double h) {
  int dims[kMaxLevels][3], strides[kMaxLevels][3];
  const int L = plan_uniform_levels(nx, ny, nz, dims, strides);
  if (L <= 0) {
    set_error("CM11a: no level plan for %dx%dx%d", nx, ny, nz);
    return FP_ERR_LAYOUT;
  }
  // the reference's hierarchy halves axes exactly (mgFineChild's stride): a
  // lattice its planner rejects (the library ceil-halves those) has no CM11a form
  for (int l = 1; l < L; ++l)
    for (int a = 0; a < 3; ++a)
      if (dims[l][a] != dims[l - 1][a] && 2 * dims[l][a] != dims[l - 1][a]) {
        set_error("CM11a: %dx%dx%d has no hierarchy of exact halvings (planUniformCM11aHierarchy rejects it)", nx,
                  ny, nz);
        return FP_ERR_LAYOUT;
      }
  const long long coarsest = (long long)(dims[L - 1][0] + 2) * (dims[L - 1][1] + 2) * (dims[L - 1][2] + 2);
  if (coarsest > kCoarseLanes) {
    set_error("CM11a: the coarsest level of %dx%dx%d has %lld cells with its halo (at most %d)", nx, ny, nz,
              coarsest, kCoarseLanes);
    return FP_ERR_UNSUPPORTED;
  }
  nx_ = nx, ny_ = ny, nz_ = nz, h_ = h;
  const int n[3] = {nx, ny, nz};
  lv_.resize(L);
  for (int l = 0; l < L; ++l) {
    CmLevel& c = lv_[l];
    c.D0 = dims[l][0] + 2;
    c.D1 = dims[l][1] + 2;
    c.D2 = dims[l][2] + 2;
    c.n = c.D0 * c.D1 * c.D2;
    // spacing of the level: h * (finest cells / level cells) per axis (…multigrid.ts:583-588)
    c.h0 = (float)(h * ((double)n[0] / (double)dims[l][0]));
    c.h1 = (float)(h * ((double)n[1] / (double)dims[l][1]));
    c.h2 = (float)(h * ((double)n[2] / (double)dims[l][2]));
    const size_t b4 = (size_t)c.n * 4, b16 = (size_t)c.n * 16;
    int rc = alloc((void**)&c.phi_raw, b4);
    if (!rc) rc = alloc((void**)&c.phi, b4);
    if (!rc) rc = alloc((void**)&c.topo, b16);
    if (!rc) rc = alloc((void**)&c.coef, b16);
    if (!rc) rc = alloc((void**)&c.p, b4);
    if (!rc) rc = alloc((void**)&c.rhs0, b4);
    if (!rc) rc = alloc((void**)&c.rhs1, b4);
    if (!rc) rc = alloc((void**)&c.res, b4);
    if (!rc) rc = alloc((void**)&c.pmin, b4);
    if (rc) return rc;
  }
  const size_t b0 = (size_t)lv_[0].n * 4;
  int rc = alloc((void**)&b_, b0);
  if (!rc) rc = alloc((void**)&pmin0_, b0);
  if (!rc) rc = alloc((void**)&pshift_, b0);
  if (!rc) rc = alloc((void**)&backup_, b0);
  if (!rc) rc = alloc((void**)&partials_, (size_t)kMeasureBlocks * 16);
  if (!rc) rc = alloc((void**)&stats_, sizeof(CmStats));
  return rc;
}

void Cm11a::coarse_solve(hipStream_t s, const float* rhs, const float* pmin) {
  const CmLevel& c = lv_.back();
  const int threads = std::max(64, ((c.n + 63) / 64) * 64);
  hipLaunchKernelGGL(k_cm_coarse, dim3(1), dim3(threads), 0, s, c, rhs, pmin, dt_, rho_, tol_, cap_, open_top_,
                     stats_);
}

// vCycle (…multigrid.ts:518-533); pmin0: the finest level's current p_min
void Cm11a::vcycle(hipStream_t s, int l, const float* rhs, const float* pmin0) {
  const int L = (int)lv_.size();
  CmLevel& c = lv_[l];
  const float* pmin = l == 0 ? pmin0 : c.pmin;
  if (l == L - 1) {
    coarse_solve(s, rhs, pmin);
    return;
  }
  CmLevel& d = lv_[l + 1];
  for (int i = 0; i < pre_; ++i)
    for (int colour = 0; colour < 2; ++colour) CM_LAUNCH(k_cm_smooth, c.n, s, c, rhs, pmin, colour);
  CM_LAUNCH(k_cm_residual, c.n, s, c, rhs, c.res);
  CM_LAUNCH(k_cm_restrict, d.n, s, c, (const float*)c.res, d, d.rhs0);
  CM_LAUNCH(k_cm_fill, d.n, s, d.p, d.n, 0.0f);
  if (l < constraint_levels_)
    CM_LAUNCH(k_cm_down_min, d.n, s, c, pmin, 1, d);
  else
    CM_LAUNCH(k_cm_fill, d.n, s, d.pmin, d.n, kNegBig);
  vcycle(s, l + 1, d.rhs0, pmin0);
  CM_LAUNCH(k_cm_prolong, c.n, s, d, c, 1);
  for (int i = 0; i < post_; ++i)
    for (int colour = 0; colour < 2; ++colour) CM_LAUNCH(k_cm_smooth, c.n, s, c, rhs, pmin, colour);
}

// fullCycle (…multigrid.ts:534-557)
void Cm11a::full_cycle(hipStream_t s, const float* pmin0) {
  const int L = (int)lv_.size();
  CmLevel& f = lv_[0];
  CM_LAUNCH(k_cm_backup_shift, f.n, s, (const float*)f.p, pmin0, backup_, pshift_, f.n);
  CM_LAUNCH(k_cm_residual, f.n, s, f, (const float*)b_, f.rhs1);
  for (int l = 0; l + 1 < L; ++l) {
    CmLevel& c = lv_[l];
    CmLevel& d = lv_[l + 1];
    CM_LAUNCH(k_cm_restrict, d.n, s, c, (const float*)c.rhs1, d, d.rhs1);
    if (l < constraint_levels_)
      CM_LAUNCH(k_cm_down_min, d.n, s, c, (const float*)(l == 0 ? pshift_ : c.pmin), 0, d);
    else
      CM_LAUNCH(k_cm_fill, d.n, s, d.pmin, d.n, kNegBig);
  }
  CmLevel& cc = lv_[L - 1];
  CM_LAUNCH(k_cm_fill, cc.n, s, cc.p, cc.n, 0.0f);
  coarse_solve(s, cc.rhs1, L == 1 ? pshift_ : cc.pmin);
  for (int l = L - 2; l >= 0; --l) {
    CM_LAUNCH(k_cm_prolong, lv_[l].n, s, lv_[l + 1], lv_[l], 0);
    vcycle(s, l, lv_[l].rhs1, pshift_);
  }
  CM_LAUNCH(k_cm_add, f.n, s, f.p, (const float*)backup_, f.n);
}

int Cm11a::run(hipStream_t s, UniformGrid* g, const fp_options* opt, double dt, double rho, fp_report* rep) {
  const auto t0 = std::chrono::steady_clock::now();
  GridRec& R = *g->host_rec();
  const int L = (int)lv_.size();
  pre_ = opt->lcp_sweeps > 0 ? opt->lcp_sweeps : 6;
  post_ = pre_;
  const int full = opt->full_cycles > 0 ? opt->full_cycles : 3;
  const int vcyc = opt->v_cycles > 0 ? opt->v_cycles : 4;
  open_top_ = opt->open_top ? 1 : 0;
  dt_ = (float)dt;
  rho_ = (float)rho;
  const float* phi = R.lv[0].phi;  // the staged level set
  HIPCHK(hipMemsetAsync(stats_, 0, sizeof(CmStats), s));
  CmLevel& f = lv_[0];
  CM_LAUNCH(k_cm_max_div, (long long)nx_ * ny_ * nz_, s, nx_, ny_, nz_, (float)h_, phi, R.in_u, R.in_v, R.in_w,
            open_top_, &stats_->div_before_bits);
  CM_LAUNCH(k_cm_finest, f.n, s, f, nx_, ny_, nz_, phi, R.in_u, R.in_v, R.in_w, R.in_q, dt_, rho_, open_top_, b_,
            pmin0_, stats_);
  for (int l = 0; l + 1 < L; ++l)
    CM_LAUNCH(k_cm_down_topology, lv_[l + 1].n, s, lv_[l], lv_[l + 1], l + 1, L, 2, open_top_);
  for (int l = 0; l < L; ++l) {
    CM_LAUNCH(k_cm_extrapolate, lv_[l].n, s, lv_[l]);
    CM_LAUNCH(k_cm_bake, lv_[l].n, s, lv_[l], open_top_);
  }
  for (int k = 0; k < full; ++k) full_cycle(s, pmin0_);
  for (int k = 0; k < vcyc; ++k) vcycle(s, 0, b_, pmin0_);
  hipLaunchKernelGGL(k_cm_measure, dim3(kMeasureBlocks), dim3(256), 0, s, f, (const float*)b_, (const float*)pmin0_,
                     dt_, rho_, stats_, partials_);
  // outputs: the projected velocities (in place when they alias the inputs) and the pressure
  hipLaunchKernelGGL(k_cm_project, dim3(grid_for((long long)nx_ * ny_ * nz_)), dim3(256), 0, s, f, nx_, ny_, nz_,
                     phi, R.in_u, R.in_v, R.in_w, R.out_u, R.out_v, R.out_w, dt_, rho_, open_top_);
  if (R.out_p)
    CM_LAUNCH(k_cm_pressure_out, (long long)nx_ * ny_ * nz_, s, f, nx_, ny_, nz_, phi, R.out_p);
  if (R.out_u && R.out_v && R.out_w)
    CM_LAUNCH(k_cm_max_div, (long long)nx_ * ny_ * nz_, s, nx_, ny_, nz_, (float)h_, phi, (const float*)R.out_u,
              (const float*)R.out_v, (const float*)R.out_w, open_top_, &stats_->div_after_bits);
  CmStats st;
  std::vector<double> part(2 * kMeasureBlocks);
  HIPCHK(hipMemcpyAsync(&st, stats_, sizeof(CmStats), hipMemcpyDeviceToHost, s));
  HIPCHK(hipMemcpyAsync(part.data(), partials_, part.size() * 8, hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  HIPCHK(hipGetLastError());
  if (rep) {
    std::memset(rep, 0, sizeof(*rep));
    double rr = 0.0, bb = 0.0;
    for (int k = 0; k < kMeasureBlocks; ++k) {
      rr += part[2 * k];
      bb += part[2 * k + 1];
    }
    float lcp, cres, db, da;
    std::memcpy(&lcp, &st.lcp_residual_bits, 4);
    std::memcpy(&cres, &st.coarse_residual_bits, 4);
    std::memcpy(&db, &st.div_before_bits, 4);
    std::memcpy(&da, &st.div_after_bits, 4);
    rep->iterations = full + vcyc;
    rep->outer_iterations = 0;
    rep->levels = L;
    rep->relative_residual = bb > 0.0 ? std::sqrt(rr / bb) : 0.0;
    rep->rhs_norm = std::sqrt(bb);
    rep->max_div_before = db;
    rep->max_div_after = (R.out_u && R.out_v && R.out_w) ? (double)da : std::nan("");
    rep->lcp_residual = lcp;
    rep->coarse_sweeps = st.coarse_sweeps;
    rep->coarse_failures = st.coarse_failed;
    if (st.coarse_failed > 0) rep->flags |= FP_FLAG_NONCONVERGENCE;
    if (!std::isfinite(lcp) || !std::isfinite(rep->relative_residual)) rep->flags |= FP_FLAG_NONFINITE;
    rep->samples = (int64_t)nx_ * ny_ * nz_;
    rep->liquid_samples = st.liquid;
    (void)cres;
  }
  const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  if (rep) rep->device_ms = ms;
  return FP_OK;
}

}  // namespace fp

// fp_march.hpp -- 7-point stencil passes as column-tile x-marches (gfx950).
//
// A 256-lane workgroup owns a (TY x TZ) tile of (y, z) columns (z fastest: one
// 64-lane wave per row when TZ = 64) and walks a chunk of x planes.  Each plane
// is loaded once into LDS with a one-cell in-plane ring, three planes rotate,
// and every lane evaluates the stencil of its own column at the middle plane.
// No integer division, coalesced row loads, and the +-x neighbours come from
// LDS instead of a second trip to L2.
//
// An operation `Op` supplies
//   using Cell = ...;                                   LDS record per cell
//   void prepare();                                     reads device scalars once per workgroup
//   Cell load(int x, int y, int z, bool inside) const;  staging (inside: in the domain)
//   void apply(int x, int y, int z, const Cell& c, const Cell* nb, const bool* has,
//              double* acc, float* mx) const;           nb/has: -x,+x,-y,+y,-z,+z
//   static constexpr int NV;                            sum-reduction width
//   static constexpr bool WITH_MAX;                     also reduce a max (stored after the sums)
//   double* partials;                                   per-workgroup partials (a finisher
//                                                       kernel completes them)
#pragma once

#include "fp_common.hpp"
#include "fp_sweep.hpp"

namespace fp {

template <int TZ, class Op>
__global__ __launch_bounds__(256) void k_march(int nx, int ny, int nz, Op op_in, int chunk, int tiles_z,
                                               int tiles_y, int ngroups, const State* st, int honour_done) {
  if (honour_done && st->done) return;
  constexpr int TY = 256 / TZ;
  constexpr int RZ = TZ + 2, RY = TY + 2;
  using Cell = typename Op::Cell;
  __shared__ Cell s[3][RY][RZ];
  Op op = op_in;
  op.prepare();
  float mx = 0.0f;
  int tzi, tyi, cxi;
  tile_of_group(blockIdx.x, tiles_z, tiles_y, ngroups, tzi, tyi, cxi);
  const int z0 = tzi * TZ, y0 = tyi * TY;
  const int x0 = cxi * chunk;
  const int x1 = min(x0 + chunk, nx);
  const int tid = threadIdx.x;
  const int ly = tid / TZ, lz = tid - ly * TZ;
  const int gy = y0 + ly, gz = z0 + lz;
  double acc[Op::NV > 0 ? Op::NV : 1];
  for (int k = 0; k < (Op::NV > 0 ? Op::NV : 1); ++k) acc[k] = 0.0;
  constexpr int NC = (RY * RZ + 255) / 256;
  Cell pre[NC];
  auto fetch = [&](int L) {
    const bool pin = (L >= 0 && L < nx);
#pragma unroll
    for (int k = 0; k < NC; ++k) {
      const int q = tid + 256 * k;
      if (q < RY * RZ) {
        const int ry = q / RZ, rz = q - ry * RZ;
        const int yy = y0 + ry - 1, zz = z0 + rz - 1;
        const bool inside = pin && yy >= 0 && yy < ny && zz >= 0 && zz < nz;
        pre[k] = op.load(L, yy, zz, inside);
      }
    }
  };
  fetch(x0 - 1);
  for (int L = x0 - 1; L <= x1; ++L) {
    const int slot = (L + 3) % 3;
#pragma unroll
    for (int k = 0; k < NC; ++k) {
      const int q = tid + 256 * k;
      if (q < RY * RZ) {
        const int ry = q / RZ, rz = q - ry * RZ;
        s[slot][ry][rz] = pre[k];
      }
    }
    lds_barrier();
    if (L + 1 <= x1) fetch(L + 1);
    const int B = L - 1;
    if (B >= x0 && B < x1 && gy < ny && gz < nz) {
      const int sB = (B + 3) % 3, sBm = (B + 2) % 3, sBp = (B + 4) % 3;
      const int cy = ly + 1, cz = lz + 1;
      Cell nb[6] = {s[sBm][cy][cz], s[sBp][cy][cz], s[sB][cy - 1][cz], s[sB][cy + 1][cz], s[sB][cy][cz - 1],
                    s[sB][cy][cz + 1]};
      const bool has[6] = {B > 0, B + 1 < nx, gy > 0, gy + 1 < ny, gz > 0, gz + 1 < nz};
      op.apply(B, gy, gz, s[sB][cy][cz], nb, has, acc, &mx);
    }
    lds_barrier();
  }
  if constexpr (Op::NV > 0) {
    if constexpr (Op::WITH_MAX) block_partials_max<Op::NV>(acc, mx, op.partials);
    else block_partials<Op::NV>(acc, op.partials);
  }
}

struct MarchGrid {
  int ngroups;
  int chunk;
  int tz, ty;     // tiles along z and y
  int TZ;         // tile width along z (64 or 32)
};

#ifndef FP_MARCH_TZ
#define FP_MARCH_TZ 32
#endif

// Tiles of TZ x 256/TZ columns, x chunks of at least 8 planes, about
// target_blocks workgroups (at most the 16384 partials of the reductions).
inline MarchGrid march_grid(int nx, int ny, int nz, long target_blocks = 4096) {
  const int TZ = nz >= FP_MARCH_TZ ? FP_MARCH_TZ : 32;
  const int TY = 256 / TZ;
  const int tz = (nz + TZ - 1) / TZ, ty = (ny + TY - 1) / TY;
  const long tiles = (long)tz * ty;
  long c = ((long)nx * tiles + target_blocks - 1) / target_blocks;
  if (c < 8) c = 8;
  if (c > nx) c = nx;
  while ((long)tiles * ((nx + c - 1) / c) > 16384) ++c;
  MarchGrid m;
  m.chunk = (int)c;
  m.tz = tz;
  m.ty = ty;
  m.TZ = TZ;
  m.ngroups = (int)(tiles * ((nx + c - 1) / c));
  return m;
}

// ---------------------------------------------------------------------------
// CG direction: d = z + beta d_old ; q = A d (FP32 operator) ; delta = d . q.
// ---------------------------------------------------------------------------
template <typename T>
struct OpDirection {
  struct Cell { float d; float p; };
  static constexpr int NV = 1;
  static constexpr bool WITH_MAX = false;
  int ny, nz;
  const float* phi;
  const T* z;
  const float* dold;
  float* dnew;
  float* q;
  float zmul;  // static part of the z scale; divided by the device rscale
  float zm, beta;
  float gx, gy, gz;
  State* st;
  double* partials;
  __device__ __forceinline__ void prepare() {
    zm = zmul / st->rscale;
    beta = (float)st->beta;
  }
  __device__ __forceinline__ Cell load(int x, int y, int zz, bool inside) const {
    Cell c{0.0f, 1.0f};
    if (inside) {
      const size_t i = ((size_t)x * ny + y) * nz + zz;
      c.p = phi[i];
      if (c.p < 0.0f) c.d = Store<T>::ld(z, i) * zm + beta * dold[i];
    }
    return c;
  }
  __device__ __forceinline__ void apply(int x, int y, int zz, const Cell& c, const Cell* nb, const bool* has,
                                        double* acc, float*) const {
    const size_t i = ((size_t)x * ny + y) * nz + zz;
    float qc = 0.0f;
    if (c.p < 0.0f) {
      const float g[6] = {gx, gx, gy, gy, gz, gz};
#pragma unroll
      for (int f = 0; f < 6; ++f)
        if (has[f]) qc += face_weight_fast(c.p, nb[f].p, g[f]) * (c.d - nb[f].d);
      acc[0] += (double)c.d * (double)qc;
    }
    dnew[i] = c.d;
    q[i] = qc;
  }
};

// ---------------------------------------------------------------------------
// V-cycle residual: t = (r*rm - emul * A e) * tinv on liquid cells.
// ---------------------------------------------------------------------------
template <typename TR, typename T>
struct OpResidual {
  struct Cell { float e; float p; };
  static constexpr int NV = 0;
  static constexpr bool WITH_MAX = false;
  int ny, nz;
  const float* phi;
  const TR* r;
  const T* e;
  T* t;
  float rmul, emul, tinv;
  int dyn;  // right-hand side scaled by the device rscale (finest level)
  float rm;
  float gx, gy, gz;
  const State* st;
  double* partials;
  __device__ __forceinline__ void prepare() { rm = dyn ? rmul * st->rscale : rmul; }
  __device__ __forceinline__ Cell load(int x, int y, int zz, bool inside) const {
    Cell c{0.0f, 1.0f};
    if (inside) {
      const size_t i = ((size_t)x * ny + y) * nz + zz;
      c.p = phi[i];
      if (c.p < 0.0f) c.e = Store<T>::ld(e, i);
    }
    return c;
  }
  __device__ __forceinline__ void apply(int x, int y, int zz, const Cell& c, const Cell* nb, const bool* has,
                                        double*, float*) const {
    const size_t i = ((size_t)x * ny + y) * nz + zz;
    float res = 0.0f;
    if (c.p < 0.0f) {
      const float g[6] = {gx, gx, gy, gy, gz, gz};
      float ap = 0.0f;
#pragma unroll
      for (int f = 0; f < 6; ++f)
        if (has[f]) ap += face_weight_fast(c.p, nb[f].p, g[f]) * (c.e - nb[f].e);
      res = (Store<TR>::ld(r, i) * rm - emul * ap) * tinv;
    }
    Store<T>::st(t, i, res);
  }
};

// ---------------------------------------------------------------------------
// V-cycle residual on the packed coefficients of a level (the operator its
// fused sweeps relax): t = (r*rm - emul * A e) * tinv.
// ---------------------------------------------------------------------------
template <typename TR, typename T>
struct OpResidualW {
  struct Cell { W4<T> w; float e; };
  static constexpr int NV = 0;
  static constexpr bool WITH_MAX = false;
  int ny, nz;
  const W4<T>* wts;
  const TR* r;
  const T* e;
  T* t;
  float rmul, emul, tinv;
  int dyn;
  float rm;
  const State* st;
  double* partials;
  __device__ __forceinline__ void prepare() { rm = dyn ? rmul * st->rscale : rmul; }
  __device__ __forceinline__ Cell load(int x, int y, int zz, bool inside) const {
    Cell c;
    c.e = 0.0f;
    Store<T>::st(&c.w.wx, 0, 0.0f);
    Store<T>::st(&c.w.wy, 0, 0.0f);
    Store<T>::st(&c.w.wz, 0, 0.0f);
    Store<T>::st(&c.w.id, 0, 0.0f);
    if (inside) {
      const size_t i = ((size_t)x * ny + y) * nz + zz;
      c.w = wts[i];
      if (wf(c.w.id) > 0.0f) c.e = Store<T>::ld(e, i);
    }
    return c;
  }
  __device__ __forceinline__ void apply(int x, int y, int zz, const Cell& c, const Cell* nb, const bool*,
                                        double*, float*) const {
    const size_t i = ((size_t)x * ny + y) * nz + zz;
    float res = 0.0f;
    if (wf(c.w.id) > 0.0f) {
      const float ap = wf(c.w.wx) * (c.e - nb[1].e) + wf(nb[0].w.wx) * (c.e - nb[0].e) +
                       wf(c.w.wy) * (c.e - nb[3].e) + wf(nb[2].w.wy) * (c.e - nb[2].e) +
                       wf(c.w.wz) * (c.e - nb[5].e) + wf(nb[4].w.wz) * (c.e - nb[4].e);
      res = (Store<TR>::ld(r, i) * rm - emul * ap) * tinv;
    }
    Store<T>::st(t, i, res);
  }
};

// ---------------------------------------------------------------------------
// FP64 true residual r = b - A x, ||r||^2, max |r| (refinement control).
// ---------------------------------------------------------------------------
struct OpTrueResidual {
  struct Cell { double x; float p; };
  static constexpr int NV = 1;
  static constexpr bool WITH_MAX = true;
  int ny, nz;
  const float* phi;
  const double* b;
  const double* xv;
  float* r;
  State* st;
  double* partials;
  __device__ __forceinline__ void prepare() {}
  __device__ __forceinline__ Cell load(int x, int y, int zz, bool inside) const {
    Cell c{0.0, 1.0f};
    if (inside) {
      const size_t i = ((size_t)x * ny + y) * nz + zz;
      c.p = phi[i];
      if (c.p < 0.0f) c.x = xv[i];
    }
    return c;
  }
  __device__ __forceinline__ void apply(int x, int y, int zz, const Cell& c, const Cell* nb, const bool* has,
                                        double* acc, float* mx) const {
    const size_t i = ((size_t)x * ny + y) * nz + zz;
    double res = 0.0;
    if (c.p < 0.0f) {
      double ap = 0.0;
#pragma unroll
      for (int f = 0; f < 6; ++f)
        if (has[f]) ap += face_weight_d(c.p, nb[f].p) * (c.x - nb[f].x);
      res = b[i] - ap;
      acc[0] += res * res;
      *mx = fmaxf(*mx, (float)fabs(res));
    }
    r[i] = (float)res;
  }
};

// Launches the march and returns its workgroup count (the partials to finish).
template <class Op>
inline unsigned launch_march(hipStream_t s, int nx, int ny, int nz, const Op& op, const State* st, int honour_done) {
  const MarchGrid m = march_grid(nx, ny, nz);
  if (m.TZ == 64)
    hipLaunchKernelGGL((k_march<64, Op>), dim3(m.ngroups), dim3(256), 0, s, nx, ny, nz, op, m.chunk, m.tz, m.ty,
                       m.ngroups, st, honour_done);
  else
    hipLaunchKernelGGL((k_march<32, Op>), dim3(m.ngroups), dim3(256), 0, s, nx, ny, nz, op, m.chunk, m.tz, m.ty,
                       m.ngroups, st, honour_done);
  return (unsigned)m.ngroups;
}

}  // namespace fp

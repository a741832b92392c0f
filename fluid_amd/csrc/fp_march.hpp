// fp_march.hpp -- 7-point stencil passes as column-tile x-marches (gfx950).
//
// A 256-lane workgroup owns a (TY x TZ) = 8 x 32 tile of (y, z) columns (z
// fastest) and walks a chunk of x planes.  Each plane
// is loaded once into LDS with a one-cell in-plane ring, three planes rotate,
// and every lane evaluates the stencil of its own column at the middle plane.
// No integer division, coalesced row loads, and the +-x neighbours come from
// LDS instead of a second trip to L2.
//
// An operation `Op` supplies
//   using Cell = ...;                                   LDS record per cell
//   using Raw = ...;                                    a cell's loads as fetched
//   static Op make(const GridRec&, const MarchLaunch&); binds the grid's arrays
//   void prepare();                                     reads device scalars once per workgroup
//   Raw fetch(int x, int y, int z, bool inside) const;  every load of a cell (inside: in the
//                                                       domain), none depending on another
//   Cell finish(const Raw&) const;                      masks and scaling, at the LDS commit
//   static constexpr int DEPTH;                         planes of loads in flight (1 or 2)
// A plane's loads are issued one plane step ahead, all at once, and masked
// only when they are committed, so no load waits for another.
//   using Own = ...;                                    the lane's own-cell loads (not staged)
//   Own own(int x, int y, int z, bool inside) const;    fetched one plane step ahead
//   void apply(int x, int y, int z, const Cell& c, const Cell* nb, const bool* has,
//              const Own& o, double* acc, float* mx) const;  nb/has: -x,+x,-y,+y,-z,+z
//   void zero(int x, int y, int z) const;               the outputs of an air cell
//   static constexpr bool ZERO_OUTSIDE;                 store zeros outside the liquid hulls
//                                                       (false: every reader of the outputs
//                                                       stays inside the hulls or on liquid)
//   static constexpr int NV;                            sum-reduction width
//   static constexpr bool WITH_MAX;                     also reduce a max (stored after the sums)
// Reductions go to the grid's partials (a finisher kernel completes them).
#pragma once

#include "fp_common.hpp"
#include "fp_sweep.hpp"
#include "fp_uniform.hpp"

namespace fp {

// Launch-uniform parameters of a march (the same for every grid of a batch).
struct MarchLaunch {
  int level;    // multigrid level (V-cycle residual)
  int which_e;  // correction buffer (0: e, 1: e2) the residual reads; FP64
                // true residual: 1 on the first refinement pass (x = 0 on
                // grids that are not warm-started)
};

constexpr int kMarchTZ = kHullTZ;

// honour: 0 always runs, 1 skips grids whose inner solve is done, 2 skips
// grids whose refinement has converged.
template <class Op>
__global__ __launch_bounds__(256) void k_march(const Part* parts, int np, int honour, MarchLaunch la) {
  FP_PART(parts, np);
  const State* st_ = cload(&Gp->st);
  if (honour == 1 && cload(&st_->done)) return;
  if (honour == 2 && cload(&st_->outer_done)) return;
  constexpr int TZ = kMarchTZ;
  constexpr int TY = 256 / TZ;
  constexpr int RZ = TZ + 2, RY = TY + 2;
  using Cell = typename Op::Cell;
  __shared__ Cell s[3][RY][RZ];
  const LevelRec lv = cload(&Gp->lv[la.level]);
  const int nx = lv.d[0], ny = lv.d[1], nz = lv.d[2];
  const int chunk = lv.mg.chunk;
  Op op = Op::make(Gp, la);
  op.prepare();
  float mx = 0.0f;
  int tzi, tyi, cxi;
  tile_of_group(blk, lv.mg.tz, lv.mg.ty, nblk, tzi, tyi, cxi);
  const int z0 = tzi * TZ, y0 = tyi * TY;
  int x0 = cxi * chunk;
  int x1 = min(x0 + chunk, nx);
  const int tid = threadIdx.x;
  const int ly = tid / TZ, lz = tid - ly * TZ;
  const int gy = y0 + ly, gz = z0 + lz;
  if (lv.hull) {
    // planes of the chunk outside the tile's liquid hull: zeros
    const int* hp = lv.hull + 2 * (tyi * lv.mg.tz + tzi);
    const int xs = max(x0, cload(hp)), xe = min(x1, cload(hp + 1) + 1);
    if (Op::ZERO_OUTSIDE && gy < ny && gz < nz)
      for (int x = x0; x < x1; ++x)
        if (x < xs || x >= xe) op.zero(x, gy, gz);
    x0 = xs;
    x1 = xe;
  }
  double acc[Op::NV > 0 ? Op::NV : 1];
  for (int k = 0; k < (Op::NV > 0 ? Op::NV : 1); ++k) acc[k] = 0.0;
  // own-cell loads: plane L's are issued at step L and used at step L + 1
  using Own = typename Op::Own;
  const bool mine = gy < ny && gz < nz;
  Own own_cur = Own(), own_next = Own();
  constexpr int NC = (RY * RZ + 255) / 256;
  using Raw = typename Op::Raw;
  // Op::DEPTH planes of loads in flight: plane L's loads are issued at step
  // L - DEPTH into register set (L - x0 + 1) % DEPTH (a compile-time index:
  // the plane loop is unrolled by DEPTH)
  constexpr int DP = Op::DEPTH;
  static_assert(DP == 1 || DP == 2, "march prefetch depth");
  Raw pre[DP][NC];
  auto where = [&](int L, int k, int& yy, int& zz) {
    const int q = tid + 256 * k;
    const int ry = q / RZ, rz = q - ry * RZ;
    yy = y0 + ry - 1;
    zz = z0 + rz - 1;
    return q < RY * RZ && L >= 0 && L < nx && yy >= 0 && yy < ny && zz >= 0 && zz < nz;
  };
  auto fetch = [&](int L, int b) {
#pragma unroll
    for (int k = 0; k < NC; ++k) {
      int yy, zz;
      const bool inside = where(L, k, yy, zz);
      pre[b][k] = op.fetch(L, yy, zz, inside);
    }
  };
  if (x0 < x1) {
#pragma unroll
    for (int d = 0; d < DP; ++d) fetch(x0 - 1 + d, d);
  }
  auto step = [&](int L, int b) {
    const int slot = (L + 3) % 3;
#pragma unroll
    for (int k = 0; k < NC; ++k) {
      const int q = tid + 256 * k;
      if (q < RY * RZ) {
        const int ry = q / RZ, rz = q - ry * RZ;
        s[slot][ry][rz] = op.finish(pre[b][k]);
      }
    }
    lds_barrier();
    if (L + DP <= x1) fetch(L + DP, b);
    if (L >= x0 && L < x1) own_next = op.own(L, gy, gz, mine);
    const int B = L - 1;
    if (B >= x0 && B < x1 && mine) {
      const int sB = (B + 3) % 3, sBm = (B + 2) % 3, sBp = (B + 4) % 3;
      const int cy = ly + 1, cz = lz + 1;
      Cell nb[6] = {s[sBm][cy][cz], s[sBp][cy][cz], s[sB][cy - 1][cz], s[sB][cy + 1][cz], s[sB][cy][cz - 1],
                    s[sB][cy][cz + 1]};
      const bool has[6] = {B > 0, B + 1 < nx, gy > 0, gy + 1 < ny, gz > 0, gz + 1 < nz};
      op.apply(B, gy, gz, s[sB][cy][cz], nb, has, own_cur, acc, &mx);
    }
    own_cur = own_next;
    lds_barrier();
  };
  if (x0 < x1)
    for (int L = x0 - 1; L <= x1; L += DP) {
#pragma unroll
      for (int d = 0; d < DP; ++d)
        if (L + d <= x1) step(L + d, d);
    }
  if constexpr (Op::NV > 0) {
    double* partials = cload(&Gp->partials);
    if constexpr (Op::WITH_MAX) block_partials_max<Op::NV>(acc, mx, partials, blk, nblk);
    else block_partials<Op::NV>(acc, partials, blk, nblk);
  }
}

// Coarse levels (level >= 2) of at most this many cells stay in the caches:
// their marches are bound by the latency of the plane chain, not by HBM, so
// they take short x chunks (more, shorter chains).  The other levels take
// chunks of at least 8 planes (a march carries up to 4 planes of start-up).
// A batched launch covers the level of every grid of the batch; from level 2
// on that is at most 1/64 of the finest levels' cells, so the launch stays
// cache-sized too (level 1 of a large batch is not).
constexpr long long kSmallLevelCells = 1ll << 21;
inline long min_march_chunk(long long cells, int level) {
  return (level >= 2 && cells <= kSmallLevelCells) ? 2 : 8;
}

// Tiles of 32 x 8 columns, about target_blocks workgroups (at most the 16384
// partials of the reductions).
#ifndef FP_MARCH_GROUPS
#define FP_MARCH_GROUPS 8192
#endif
inline TileGeo march_geometry(int nx, int ny, int nz, int level, long target_blocks = FP_MARCH_GROUPS) {
  const int TZ = kMarchTZ, TY = 256 / TZ;
  const int tz = (nz + TZ - 1) / TZ, ty = (ny + TY - 1) / TY;
  const long tiles = (long)tz * ty;
  long c = ((long)nx * tiles + target_blocks - 1) / target_blocks;
  const long cmin = min_march_chunk((long long)nx * ny * nz, level);
  if (c < cmin) c = cmin;
  if (c > nx) c = nx;
  while ((long)tiles * ((nx + c - 1) / c) > (long)kMaxBlocks) ++c;
  TileGeo m;
  m.chunk = (int)c;
  m.tz = tz;
  m.ty = ty;
  m.ngroups = (int)(tiles * ((nx + c - 1) / c));
  return m;
}

// ---------------------------------------------------------------------------
// Liquid hull of the march tiles along x.  A tile's planes outside
// [hull[0], hull[1]] ([nx, -1] when it has none) hold no liquid in its 32 x 8 columns, so every stencil
// pass on the level (sweeps, residual, CG direction, true residual) yields
// zeros there: those kernels store zeros for such planes and march only over
// the hull, which gives the same bits as marching over the whole chunk (their
// results do not depend on where a chunk starts or ends).
// ---------------------------------------------------------------------------
__global__ void k_hull_reset(const Part* parts, int np, int level) {
  FP_PART(parts, np);
  int* hull = G.lv[level].hull;
  if (!hull) return;
  const size_t tiles = (size_t)G.lv[level].mg.tz * G.lv[level].mg.ty;
  const int nx = G.lv[level].d[0];
  FP_STRIDE(t, tiles) {  // empty: [nx, -1]
    hull[2 * t] = nx;
    hull[2 * t + 1] = -1;
  }
}

// One workgroup per (tile, x chunk) of the march geometry: the chunk's first
// and last liquid planes, merged into the tile's hull with one atomic min and
// max.
__global__ __launch_bounds__(256) void k_tile_hull(const Part* parts, int np, int level) {
  FP_PART(parts, np);
  const LevelRec lv = cload(&Gp->lv[level]);
  if (!lv.hull) return;
  constexpr int TZ = kMarchTZ, TY = 256 / TZ;
  const int nx = lv.d[0], ny = lv.d[1], nz = lv.d[2];
  int tzi, tyi, cxi;
  tile_of_group(blk, lv.mg.tz, lv.mg.ty, nblk, tzi, tyi, cxi);
  const int x0 = cxi * lv.mg.chunk, x1 = min(x0 + lv.mg.chunk, nx);
  const int tid = threadIdx.x;
  const int gy = tyi * TY + tid / TZ, gz = tzi * TZ + tid % TZ;
  const bool inside = gy < ny && gz < nz;
  const size_t col = (size_t)gy * nz + gz;
  const long long plane = (long long)ny * nz;
  int lo = nx, hi = -1;
  for (int xb = x0; xb < x1; xb += 8) {
    float v[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) v[k] = (inside && xb + k < x1) ? ldg(lv.phi, (size_t)((xb + k) * plane) + col) : 1.0f;
#pragma unroll
    for (int k = 0; k < 8; ++k)
      if (__any(v[k] < 0.0f)) {
        lo = min(lo, xb + k);
        hi = max(hi, xb + k);
      }
  }
  __shared__ int s_lo[4], s_hi[4];
  if ((tid & 63) == 0) {
    s_lo[tid >> 6] = lo;
    s_hi[tid >> 6] = hi;
  }
  __syncthreads();
  if (tid == 0) {
    for (int w = 1; w < 4; ++w) {
      lo = min(lo, s_lo[w]);
      hi = max(hi, s_hi[w]);
    }
    if (lo <= hi) {
      int* h = lv.hull + 2 * (tyi * lv.mg.tz + tzi);
      atomicMin(h, lo);
      atomicMax(h + 1, hi);
    }
  }
}

// Planes of a march-geometry workgroup's chunk that can hold liquid in its
// tile (the chunk clamped to the tile's hull; the whole chunk on levels
// without hulls) and the lane's cell column.
struct TileSpan {
  int x0, x1;
  size_t col;    // y * nz + z of the lane's column
  bool inside;   // the column lies in the grid
};
__device__ __forceinline__ TileSpan tile_span(const LevelRec& lv, int blk, int nblk) {
  constexpr int TZ = kMarchTZ, TY = 256 / TZ;
  int tzi, tyi, cxi;
  tile_of_group(blk, lv.mg.tz, lv.mg.ty, nblk, tzi, tyi, cxi);
  TileSpan t;
  t.x0 = cxi * lv.mg.chunk;
  t.x1 = min(t.x0 + lv.mg.chunk, lv.d[0]);
  if (lv.hull) {
    const int* hp = lv.hull + 2 * (tyi * lv.mg.tz + tzi);
    t.x0 = max(t.x0, cload(hp));
    t.x1 = min(t.x1, cload(hp + 1) + 1);
    if (t.x0 > t.x1) t.x0 = t.x1;  // no liquid: an empty span
  }
  const int gy = tyi * TY + (int)threadIdx.x / TZ, gz = tzi * TZ + (int)threadIdx.x % TZ;
  t.inside = gy < lv.d[1] && gz < lv.d[2];
  t.col = (size_t)gy * lv.d[2] + gz;
  return t;
}

// ---------------------------------------------------------------------------
// CG direction: d = z + beta d_old ; q = A d (FP32 operator) ; delta = d . q.
// ---------------------------------------------------------------------------
template <typename T>
struct OpDirection {
  struct Cell { float d; float p; };
  static constexpr int DEPTH = 2;  // 145 us against 165 at depth 1 (384^3 hulls)
  static constexpr int NV = 1;
  static constexpr bool WITH_MAX = false;
  // d and q are read on liquid cells (the next direction) and over the hulls
  // (the CG update), never on the air planes outside them
  static constexpr bool ZERO_OUTSIDE = false;
  int ny, nz;
  const float* phi;
  const T* z;
  const float* dold;
  float* dnew;
  float* q;
  float zmul;  // static part of the z scale; divided by the device rscale
  float zm, beta;
  float gx, gy, gz;
  const State* st;
  static __device__ __forceinline__ OpDirection make(const GridRec* Gp, const MarchLaunch&) {
    const LevelRec l0 = cload(&Gp->lv[0]);
    OpDirection o;
    o.ny = l0.d[1];
    o.nz = l0.d[2];
    o.phi = l0.phi;
    o.z = (const T*)(l0.efinal ? l0.e2 : l0.e);
    o.st = cload(&Gp->st);
    // CG direction buffers alternate with the grid's own iteration count
    const int parity = cload(&o.st->iter) & 1;
    o.dold = cload(&Gp->d[parity]);
    o.dnew = cload(&Gp->d[parity ^ 1]);
    o.q = cload(&Gp->q);
    o.zmul = l0.Se;
    o.gx = l0.g[0];
    o.gy = l0.g[1];
    o.gz = l0.g[2];
    return o;
  }
  __device__ __forceinline__ void prepare() {
    zm = zmul / cload(&st->rscale);
    beta = (float)cload(&st->beta);
  }
  struct Raw { float p; T z; float d; };  // phi (air outside the domain), z, d_old
  __device__ __forceinline__ Raw fetch(int x, int y, int zz, bool inside) const {
    Raw w{1.0f, to_store<T>(0.0f), 0.0f};
    if (inside) {
      const size_t i = ((size_t)x * ny + y) * nz + zz;
      w.p = ldg(phi, i);
      w.z = ldg(z, i);
      w.d = ldg(dold, i);
    }
    return w;
  }
  __device__ __forceinline__ Cell finish(const Raw& w) const {
    return Cell{w.p < 0.0f ? Store<T>::ld(&w.z, 0) * zm + beta * w.d : 0.0f, w.p};
  }
  struct Own {};
  __device__ __forceinline__ Own own(int, int, int, bool) const { return Own(); }
  __device__ __forceinline__ void apply(int x, int y, int zz, const Cell& c, const Cell* nb, const bool* has,
                                        const Own&, double* acc, float*) const {
    const size_t i = ((size_t)x * ny + y) * nz + zz;
    float qc = 0.0f;
    if (c.p < 0.0f) {
      const float g[6] = {gx, gx, gy, gy, gz, gz};
#pragma unroll
      for (int f = 0; f < 6; ++f)
        if (has[f]) qc += face_weight_fast(c.p, nb[f].p, g[f]) * (c.d - nb[f].d);
      acc[0] += (double)c.d * (double)qc;
    }
    stg(dnew, i, c.d);
    stg(q, i, qc);
  }
  __device__ __forceinline__ void zero(int x, int y, int zz) const {
    const size_t i = ((size_t)x * ny + y) * nz + zz;
    stg(dnew, i, 0.0f);
    stg(q, i, 0.0f);
  }
};

// ---------------------------------------------------------------------------
// V-cycle residual on the packed coefficients of a level (the operator its
// fused sweeps relax): t = (r*rm - emul * A e) * tinv.
// ---------------------------------------------------------------------------
template <typename TR, typename T>
struct OpResidualW {
  struct Cell { W4<T> w; float e; };
  static constexpr int DEPTH = 1;  // 143 us against 165 at depth 2
  static constexpr int NV = 0;
  static constexpr bool WITH_MAX = false;
  // the restriction reads t only inside the fine hulls (zero elsewhere)
  static constexpr bool ZERO_OUTSIDE = false;
  int ny, nz;
  const W4<T>* wts;
  const TR* r;
  const T* e;
  T* t;
  float rmul, emul, tinv;
  int dyn;  // right-hand side scaled by the device rscale (finest level)
  float rm;
  const State* st;
  static __device__ __forceinline__ OpResidualW make(const GridRec* Gp, const MarchLaunch& la) {
    const LevelRec l = cload(&Gp->lv[la.level]);
    const bool top = la.level == 0;
    OpResidualW o;
    o.ny = l.d[1];
    o.nz = l.d[2];
    o.wts = (const W4<T>*)l.w;
    o.r = (const TR*)(top ? (const void*)cload(&Gp->r) : l.r);
    o.e = (const T*)(la.which_e ? l.e2 : l.e);
    o.t = (T*)l.t;
    o.rmul = (top ? 1.0f : l.Sr) / l.St;
    o.emul = l.Se / l.St;
    o.tinv = 1.0f;
    o.dyn = top;
    o.st = cload(&Gp->st);
    return o;
  }
  __device__ __forceinline__ void prepare() { rm = dyn ? rmul * cload(&st->rscale) : rmul; }
  struct Raw { W4<T> w; T e; };  // coefficients (zero outside the domain), correction
  __device__ __forceinline__ Raw fetch(int x, int y, int zz, bool inside) const {
    Raw r;
    Store<T>::st(&r.w.wx, 0, 0.0f);
    Store<T>::st(&r.w.wy, 0, 0.0f);
    Store<T>::st(&r.w.wz, 0, 0.0f);
    Store<T>::st(&r.w.id, 0, 0.0f);
    r.e = to_store<T>(0.0f);
    if (inside) {
      const size_t i = ((size_t)x * ny + y) * nz + zz;
      r.w = ldg(wts, i);
      r.e = ldg(e, i);
    }
    return r;
  }
  __device__ __forceinline__ Cell finish(const Raw& r) const {
    Cell c;
    c.w = r.w;
    c.e = wf(r.w.id) > 0.0f ? Store<T>::ld(&r.e, 0) : 0.0f;
    return c;
  }
  using Own = TR;  // right-hand side of the lane's cell
  __device__ __forceinline__ Own own(int x, int y, int zz, bool inside) const {
    return inside ? ldg(r, ((size_t)x * ny + y) * nz + zz) : to_store<TR>(0.0f);
  }
  __device__ __forceinline__ void apply(int x, int y, int zz, const Cell& c, const Cell* nb, const bool*,
                                        const Own& rv, double*, float*) const {
    const size_t i = ((size_t)x * ny + y) * nz + zz;
    float res = 0.0f;
    if (wf(c.w.id) > 0.0f) {
      const float ap = wf(c.w.wx) * (c.e - nb[1].e) + wf(nb[0].w.wx) * (c.e - nb[0].e) +
                       wf(c.w.wy) * (c.e - nb[3].e) + wf(nb[2].w.wy) * (c.e - nb[2].e) +
                       wf(c.w.wz) * (c.e - nb[5].e) + wf(nb[4].w.wz) * (c.e - nb[4].e);
      res = (Store<TR>::ld(&rv, 0) * rm - emul * ap) * tinv;
    }
    Store<T>::gst(t, i, res);
  }
  __device__ __forceinline__ void zero(int x, int y, int zz) const {
    Store<T>::gst(t, ((size_t)x * ny + y) * nz + zz, 0.0f);
  }
};

// ---------------------------------------------------------------------------
// FP64 true residual r = b - A x, ||r||^2, max |r| (refinement control).
// ---------------------------------------------------------------------------
struct OpTrueResidual {
  struct Cell { double x; float p; };
  static constexpr int DEPTH = 1;  // depth 2 is no faster
  static constexpr int NV = 2;     // ||b - A x||^2 and the sum of x (null-space mean)
  static constexpr bool WITH_MAX = true;
  // r is read on liquid cells (the sweeps' own right-hand side) and over the
  // hulls (CG update, r.z)
  static constexpr bool ZERO_OUTSIDE = false;
  int ny, nz;
  const float* phi;
  const double* b;
  const double* xv;
  float* r;
  double bmean;  // mean of b removed on an all-liquid grid (else 0)
  bool x_zero;  // the iterate is still zero: A x is not loaded
  static __device__ __forceinline__ OpTrueResidual make(const GridRec* Gp, const MarchLaunch& la) {
    const LevelRec l0 = cload(&Gp->lv[0]);
    OpTrueResidual o;
    o.ny = l0.d[1];
    o.nz = l0.d[2];
    o.phi = l0.phi;
    o.b = cload(&Gp->b);
    o.xv = cload(&Gp->x);
    o.r = cload(&Gp->r);
    o.bmean = cload(&cload(&Gp->st)->bmean);
    o.x_zero = la.which_e == 1 && !cload(&Gp->warm);
    return o;
  }
  __device__ __forceinline__ void prepare() {}
  using Raw = Cell;  // phi (air outside the domain) and x as loaded
  __device__ __forceinline__ Raw fetch(int x, int y, int zz, bool inside) const {
    Raw c{0.0, 1.0f};
    if (inside) {
      const size_t i = ((size_t)x * ny + y) * nz + zz;
      c.p = ldg(phi, i);
      if (!x_zero) c.x = ldg(xv, i);
    }
    return c;
  }
  // x of an air cell (stale after the liquid moved) is the zero ghost value
  __device__ __forceinline__ Cell finish(const Raw& c) const { return Cell{c.p < 0.0f ? c.x : 0.0, c.p}; }
  using Own = double;  // b of the lane's cell
  __device__ __forceinline__ Own own(int x, int y, int zz, bool inside) const {
    return inside ? ldg(b, ((size_t)x * ny + y) * nz + zz) : 0.0;
  }
  __device__ __forceinline__ void apply(int x, int y, int zz, const Cell& c, const Cell* nb, const bool* has,
                                        const Own& bv, double* acc, float* mx) const {
    const size_t i = ((size_t)x * ny + y) * nz + zz;
    double res = 0.0;
    if (c.p < 0.0f) {
      double ap = 0.0;
#pragma unroll
      for (int f = 0; f < 6; ++f)
        if (has[f]) ap += face_weight_d(c.p, nb[f].p) * (c.x - nb[f].x);
      res = (bv - bmean) - ap;
      acc[0] += res * res;
      acc[1] += c.x;
      *mx = fmaxf(*mx, (float)fabs(res));
    }
    stg(r, i, (float)res);
  }
  __device__ __forceinline__ void zero(int x, int y, int zz) const {
    stg(r, ((size_t)x * ny + y) * nz + zz, 0.0f);
  }
};

}  // namespace fp

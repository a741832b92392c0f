// fp_tall.hip -- restricted tall-cell pressure projection on gfx950, batched.
//
// Pressure space: the packed samples of lib/core/tall-cell-grid.ts (one tall
// cell + a cubic band per column, indexing of tall-cell-grid.ts:171-190,
// transposed to z-fastest); inside a tall cell the fine-row pressure and
// level set are linear between the bottom and top samples (the represented
// field of representedInitialPhiVolumeCells, tall-cell-grid.ts:228-238).  The
// projection is the variational reduction of the uniform ghost-fluid
// projection on the expanded fine grid (docs/papers/sg2025narita.md:120-172,
// the SPD form of docs/papers/tallCells.txt Eq. 13-18):
//   A = E^T D^T W D E,  b = (rho h/dt) E^T D^T u,  u <- P_slots(u - dt/(rho h) W D E p)
// evaluated per sample in closed form: the rows two neighbouring tall cells
// share enter through the weight moments sum w, sum w t, sum w t^2 of their
// face (t = row - (m-1)/2), every other fine face is one explicit row.
//
// Solver: FP64 iterative refinement around FP32 CG preconditioned by a
// multigrid V-cycle.  Sample level: damped (bottom, top)-block Jacobi on the
// true operator (each face touches two blocks, so A <= 2 blockdiag(A) and any
// damping below 1 keeps the cycle SPD).  Coarser levels: a 7-point lattice
// surrogate (tall rows kept as two layers, band rows aggregated in pairs,
// columns in 2 x 2), whose couplings are the aggregation (Galerkin) sums
// halved along coarsened axes, smoothed by red-black Gauss-Seidel, with
// cell-centred linear transfers across columns; dense FP64 inverse on the
// coarsest lattice.
#include "fp_tall.hpp"
#include "fp_state.hpp"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

namespace fp {

// --------------------------------------------------------------------------
// geometry
// --------------------------------------------------------------------------
struct TallGeo {
  int nx, ny, nz, B, P, S;
  double h;
  __device__ __forceinline__ size_t sidx(int x, int py, int z) const { return ((size_t)x * P + py) * nz + z; }
  __device__ __forceinline__ size_t col(int x, int z) const { return (size_t)x * nz + z; }
  __device__ __forceinline__ size_t uidx(int x, int s, int z) const { return ((size_t)x * S + s) * nz + z; }
  __device__ __forceinline__ size_t widx(int x, int s, int z) const { return ((size_t)x * S + s) * (nz + 1) + z; }
};

__device__ __forceinline__ TallGeo geo_of(const TallRec& R) {
  return TallGeo{R.nx, R.ny, R.nz, R.band, R.P, R.S, R.h};
}

template <typename Real> __device__ __forceinline__ Real fw(float a, float b);
template <> __device__ __forceinline__ float fw<float>(float a, float b) { return face_weight(a, b, 1.0f); }
template <> __device__ __forceinline__ double fw<double>(float a, float b) { return face_weight_d(a, b); }
// face weight with interpolated (non-sample) level sets
template <typename Real> __device__ __forceinline__ Real fwr(Real a, Real b);
template <> __device__ __forceinline__ float fwr<float>(float a, float b) { return face_weight(a, b, 1.0f); }
template <> __device__ __forceinline__ double fwr<double>(double a, double b) {
  const bool la = a < 0.0, lb = b < 0.0;
  if (la && lb) return 1.0;
  if (!(la || lb)) return 0.0;
  return la ? 1.0 / ghost_theta_d(a, b) : 1.0 / ghost_theta_d(b, a);
}

// Fine row y of column (x, z) with base b: level set and p = c0 v[i0] + c1 v[i1].
template <typename Real>
struct RowRep {
  int n;
  size_t i0, i1;
  Real c0, c1;
  Real phi;
};

template <typename Real>
__device__ __forceinline__ RowRep<Real> row_rep(const TallGeo& g, const float* __restrict__ phi, int b, int x, int z,
                                                int y) {
  RowRep<Real> r;
  r.i0 = r.i1 = 0;
  r.c0 = r.c1 = (Real)0;
  if (y < b) {
    const size_t ib = g.sidx(x, 0, z), it = g.sidx(x, 1, z);
    const Real pb = (Real)phi[ib], pt = (Real)phi[it];
    if (y == 0) {
      r.n = 1; r.i0 = ib; r.c0 = (Real)1; r.phi = pb;
    } else if (y == b - 1) {
      r.n = 1; r.i0 = it; r.c0 = (Real)1; r.phi = pt;
    } else {
      const Real k = (Real)y / (Real)(b - 1);
      r.n = 2; r.i0 = ib; r.i1 = it; r.c0 = (Real)1 - k; r.c1 = k;
      r.phi = pb + (pt - pb) * k;
    }
  } else if (y - b < g.B && y < g.ny) {
    r.n = 1;
    r.i0 = g.sidx(x, y - b + 2, z);
    r.c0 = (Real)1;
    r.phi = (Real)phi[r.i0];
  } else {
    r.n = 0;
    r.phi = (Real)(5.0 * g.h);
  }
  return r;
}

template <typename Real, class V>
__device__ __forceinline__ Real row_value(const RowRep<Real>& r, const V& vec) {
  if (r.n == 0) return (Real)0;
  Real v = r.c0 * vec(r.i0);
  if (r.n == 2) v += r.c1 * vec(r.i1);
  return v;
}

// neighbour columns: -x, +x, -z, +z
__device__ __forceinline__ bool neighbour(const TallGeo& g, int x, int z, int dir, int& xn, int& zn) {
  xn = x + (dir == 1) - (dir == 0);
  zn = z + (dir == 3) - (dir == 2);
  return xn >= 0 && xn < g.nx && zn >= 0 && zn < g.nz;
}

// Moments of the face between (x, z) and its neighbour in direction dir.
__device__ __forceinline__ const double* face_moments(const TallGeo& g, const TallCol* __restrict__ cols, int x, int z,
                                                      int dir) {
  switch (dir) {
    case 0: return cols[g.col(x - 1, z)].mx;
    case 1: return cols[g.col(x, z)].mx;
    case 2: return cols[g.col(x, z - 1)].mz;
    default: return cols[g.col(x, z)].mz;
  }
}

// --------------------------------------------------------------------------
// operator rows.  Vector accessors return the value at a sample (vectors are
// zero on samples that are not unknowns, so air samples read as p = 0).
// --------------------------------------------------------------------------
// Band unknown (x, py >= 2, z) at fine row y = b + py - 2: (A p)_i and A_ii.
template <typename Real, class V>
__device__ __forceinline__ void band_row(const TallGeo& g, const int* __restrict__ bases,
                                         const float* __restrict__ phi, int x, int py, int z, int b, const V& vec,
                                         Real& ap, Real& dg) {
  const int y = b + py - 2;
  const size_t i = g.sidx(x, py, z);
  const float phic = phi[i];
  Real acc = (Real)0, d = (Real)0;
  {
    const size_t j = g.sidx(x, py - 1, z);
    const Real w = fw<Real>(phic, phi[j]);
    d += w;
    acc += w * vec(j);
  }
  if (y + 1 < g.ny) {
    if (py + 1 < g.P) {
      const size_t j = g.sidx(x, py + 1, z);
      const Real w = fw<Real>(phic, phi[j]);
      d += w;
      acc += w * vec(j);
    } else {
      d += fwr<Real>((Real)phic, (Real)(5.0 * g.h));
    }
  }
#pragma unroll
  for (int dir = 0; dir < 4; ++dir) {
    int xn, zn;
    if (!neighbour(g, x, z, dir, xn, zn)) continue;
    const int bn = bases[g.col(xn, zn)];
    const RowRep<Real> r = row_rep<Real>(g, phi, bn, xn, zn, y);
    const Real w = fwr<Real>((Real)phic, r.phi);
    d += w;
    acc += w * row_value(r, vec);
  }
  ap = d * vec(i) - acc;
  dg = d;
}

// The (bottom, top) pair of column (x, z): (A p) on both and the 2 x 2
// diagonal block.
template <typename Real>
struct PairOut {
  Real ab, at;        // (A p) bottom, top
  Real bb, bt, tt;    // block entries
};

template <typename Real, class V>
__device__ __forceinline__ PairOut<Real> tall_pair(const TallGeo& g, const int* __restrict__ bases,
                                                   const float* __restrict__ phi, const TallCol* __restrict__ cols,
                                                   int x, int z, int b, const V& vec) {
  PairOut<Real> o;
  const Real kap = (Real)1 / (Real)(b - 1);
  const size_t ib = g.sidx(x, 0, z), it = g.sidx(x, 1, z);
  const Real pb = vec(ib), pt = vec(it);
  const float phib = phi[ib], phit = phi[it];
  {
    const Real c = (Real)cols[g.col(x, z)].wint * kap * kap;
    o.ab = c * (pb - pt);
    o.at = c * (pt - pb);
    o.bb = c;
    o.tt = c;
    o.bt = -c;
  }
  if (b < g.ny) {
    const size_t j = g.sidx(x, 2, z);
    const Real w = fw<Real>(phit, phi[j]);
    o.at += w * (pt - vec(j));
    o.tt += w;
  }
#pragma unroll
  for (int dir = 0; dir < 4; ++dir) {
    int xn, zn;
    if (!neighbour(g, x, z, dir, xn, zn)) continue;
    const int bn = bases[g.col(xn, zn)];
    const double* M = face_moments(g, cols, x, z, dir);
    const Real M0 = (Real)M[0], M1 = (Real)M[1], M2 = (Real)M[2];
    const int m = b < bn ? b : bn;
    const Real ym = (Real)0.5 * (Real)(m - 1);
    const Real kn = (Real)1 / (Real)(bn - 1);
    const Real pbn = vec(g.sidx(xn, 0, zn)), ptn = vec(g.sidx(xn, 1, zn));
    const Real G0 = (pbn + (ptn - pbn) * ym * kn) - (pb + (pt - pb) * ym * kap);
    const Real G1 = (ptn - pbn) * kn - (pt - pb) * kap;
    const Real F0 = G0 * M0 + G1 * M1, F1 = G0 * M1 + G1 * M2;
    const Real c0 = (Real)1 - ym * kap, c1 = -kap, d0 = ym * kap, d1 = kap;
    o.ab -= c0 * F0 + c1 * F1;
    o.at -= d0 * F0 + d1 * F1;
    o.bb += c0 * c0 * M0 + (Real)2 * c0 * c1 * M1 + c1 * c1 * M2;
    o.tt += d0 * d0 * M0 + (Real)2 * d0 * d1 * M1 + d1 * d1 * M2;
    o.bt += c0 * d0 * M0 + (c0 * d1 + c1 * d0) * M1 + c1 * d1 * M2;
    // rows of this tall cell above the neighbour's base (at most max_delta)
    for (int y = m; y < b; ++y) {
      const Real k = (Real)y * kap;
      const Real phia = (Real)phib + ((Real)phit - (Real)phib) * k;
      const Real pa = pb + (pt - pb) * k;
      const RowRep<Real> r = row_rep<Real>(g, phi, bn, xn, zn, y);
      const Real w = fwr<Real>(phia, r.phi);
      const Real gg = row_value(r, vec) - pa;
      o.ab -= w * gg * ((Real)1 - k);
      o.at -= w * gg * k;
      o.bb += w * ((Real)1 - k) * ((Real)1 - k);
      o.tt += w * k * k;
      o.bt += w * k * ((Real)1 - k);
    }
  }
  return o;
}

__device__ __forceinline__ bool is_unknown(const TallGeo& g, const float* __restrict__ phi, int b, int x, int py,
                                           int z) {
  if (py >= 2 && b + py - 2 >= g.ny) return false;
  return phi[g.sidx(x, py, z)] < 0.0f;
}

// Vector accessors
struct VecF {
  const float* p;
  __device__ __forceinline__ float operator()(size_t i) const { return p[i]; }
};
struct VecD {
  const double* p;
  __device__ __forceinline__ double operator()(size_t i) const { return p[i]; }
};
// z + beta d (the CG direction, formed on the fly for every neighbour read)
struct VecDir {
  const float* z;
  const float* d;
  float beta;
  __device__ __forceinline__ float operator()(size_t i) const { return z[i] + beta * d[i]; }
};

// Item it of a pass over nx columns x ni rows (or segments) x zch z chunks:
// (x, k, zc).  With xcd, workgroup b (XCD b % 8) takes the items of XCD
// b % 8's block of nx / 8 columns, x fastest within each (k, zc); otherwise
// zc fastest, then k, then x.
__device__ __forceinline__ void tall_item_of(int it, int nx, int ni, int zch, bool xcd, int& x, int& k, int& zc) {
  if (xcd) {
    const int x8 = nx >> 3, l = it >> 3;
    const int sl = l / x8;
    x = (it & 7) * x8 + (l - sl * x8);
    k = sl / zch;
    zc = sl - k * zch;
  } else {
    const int xk = it / zch;
    zc = it - xk * zch;
    k = xk % ni;
    x = xk / ni;
  }
}

// Sample loops run over lines of FP_BLOCK samples along z at one (x, packed
// row): the line index is uniform across the workgroup (scalar division),
// each lane takes one z, and a wave shares its packed row (no divergence
// between tall and band rows).  As the segment passes below, the lines are
// dealt to the XCDs in blocks of nx / 8 columns walked x fastest (when nx and
// the stride divide by 8), so that neighbour columns share an L2.
#define TALL_LINES(g) for (int line_ = blk, zch_ = ((g).nz + FP_BLOCK - 1) / FP_BLOCK, \
                           xcd_ = ((g).nx & 7) == 0 && (nblk & 7) == 0;                 \
                           line_ < (g).nx * (g).P * zch_; line_ += nblk)
#define TALL_LINE(g, idx, x, py, z)                       \
  int x, py, zc_;                                         \
  tall_item_of(line_, (g).nx, (g).P, zch_, xcd_, x, py, zc_); \
  const int z = zc_ * FP_BLOCK + (int)threadIdx.x;        \
  if (z >= (g).nz) continue;                              \
  const size_t idx = ((size_t)x * (g).P + py) * (g).nz + z;

// FP32 band segments.  The FP32 passes of the solve (block Jacobi, V-cycle
// residual, CG direction) take a column's band rows kSegRows at a time: a
// lane loads its column's base, the four neighbour bases and the
// neighbours' tall bottom / top samples once for the segment, and fetches
// every neighbour's value at the segment's fine rows without branches (the
// band candidate at a clamped packed row, the tall interpolation and air are
// evaluated and selected), so all loads of a segment are in flight together
// instead of one dependent chain per sample.  The rows' (A p) and diagonal
// follow band_row<float> term by term.
constexpr int kSegRows = 4;

template <class V>
__device__ __forceinline__ void band_segment(const TallGeo& g, const int* __restrict__ bases,
                                             const float* __restrict__ phi, int x, int py0, int z, int b,
                                             const V& vec, float (&ap)[kSegRows], float (&dg)[kSegRows]) {
  const float phi_air = (float)(5.0 * g.h);
  float pc[kSegRows + 2], vc[kSegRows + 2];  // rows py0-1 .. py0+kSegRows
#pragma unroll
  for (int j = 0; j < kSegRows + 2; ++j) {
    const int py = min(py0 - 1 + j, g.P - 1);
    const size_t i = g.sidx(x, py, z);
    pc[j] = phi[i];
    vc[j] = vec(i);
  }
  float d[kSegRows], acc[kSegRows];
#pragma unroll
  for (int j = 0; j < kSegRows; ++j) {
    const int py = py0 + j, y = b + py - 2;
    const float p0 = pc[j + 1];
    // below: the previous band row or the tall top (always a sample)
    const float wd = fw<float>(p0, pc[j]);
    d[j] = wd;
    acc[j] = wd * vc[j];
    // above: the next band row, air past the band, nothing past the grid
    if (y + 1 < g.ny) {
      if (py + 1 < g.P) {
        const float wu = fw<float>(p0, pc[j + 2]);
        d[j] += wu;
        acc[j] += wu * vc[j + 2];
      } else {
        d[j] += fwr<float>(p0, phi_air);
      }
    }
  }
#pragma unroll
  for (int dir = 0; dir < 4; ++dir) {
    int xn, zn;
    if (!neighbour(g, x, z, dir, xn, zn)) continue;
    const int bn = bases[g.col(xn, zn)];
    const size_t ib = g.sidx(xn, 0, zn), it = g.sidx(xn, 1, zn);
    const float pb = phi[ib], pt = phi[it];
    const float vb = vec(ib), vt = vec(it);
    float pband[kSegRows], vband[kSegRows];
#pragma unroll
    for (int j = 0; j < kSegRows; ++j) {
      const int y = b + py0 + j - 2;
      const int pr = min(max(y - bn + 2, 2), g.P - 1);  // in bounds whichever case applies
      const size_t i = g.sidx(xn, pr, zn);
      pband[j] = phi[i];
      vband[j] = vec(i);
    }
#pragma unroll
    for (int j = 0; j < kSegRows; ++j) {
      const int y = b + py0 + j - 2;
      float pn, vn;
      if (y < bn) {  // inside the neighbour's tall cell (row_rep: rows 0, bn-1 exact)
        const float k = (float)y / (float)(bn - 1);
        pn = y == 0 ? pb : (y == bn - 1 ? pt : pb + (pt - pb) * k);
        vn = y == 0 ? vb : (y == bn - 1 ? vt : (1.0f - k) * vb + k * vt);
      } else if (y - bn < g.B && y < g.ny) {
        pn = pband[j];
        vn = vband[j];
      } else {
        pn = phi_air;
        vn = 0.0f;
      }
      const float w = fwr<float>(pc[j + 1], pn);
      d[j] += w;
      acc[j] += w * vn;
    }
  }
#pragma unroll
  for (int j = 0; j < kSegRows; ++j) {
    ap[j] = d[j] * vc[j + 1] - acc[j];
    dg[j] = d[j];
  }
}

// Work items of the segment passes: per (x, z chunk of FP_BLOCK), item 0 is
// the column's (bottom, top) pair and items 1.. its band segments (py0 =
// 2 + kSegRows (k-1)); the item index is uniform across the workgroup and
// each lane takes one z.  Workgroup b runs on XCD b % 8 (a part starts at a
// multiple of 8 and the stride is a multiple of 8), so when nx and the stride
// divide by 8 the items are dealt as eight contiguous blocks of nx / 8
// columns, one per XCD, walked x fastest within each (segment, z chunk): the
// +-x neighbour columns an item reads were just read by the workgroups before
// it on the same XCD, whose L2 still holds them.
#define TALL_ITEMS(g) for (int item_ = blk, zch_ = ((g).nz + FP_BLOCK - 1) / FP_BLOCK, \
                           ni_ = 1 + ((g).B + kSegRows - 1) / kSegRows,                \
                           xcd_ = ((g).nx & 7) == 0 && (nblk & 7) == 0;                \
                           item_ < (g).nx * ni_ * zch_; item_ += nblk)
#define TALL_ITEM(g, x, k, z)                                   \
  int x, k, zc_;                                                \
  tall_item_of(item_, (g).nx, ni_, zch_, xcd_, x, k, zc_);      \
  const int z = zc_ * FP_BLOCK + (int)threadIdx.x;              \
  if (z >= (g).nz) continue;

// --------------------------------------------------------------------------
// set-up
// --------------------------------------------------------------------------
// Face moments of the +x and +z faces of each column and the internal weight.
__device__ void moments_of(const TallGeo& g, const float* __restrict__ phi, int x, int z, int b, int xn, int zn, int bn,
                           double* M) {
  const int m = b < bn ? b : bn;
  const float pa0 = phi[g.sidx(x, 0, z)], pa1 = phi[g.sidx(x, 1, z)];
  const float pn0 = phi[g.sidx(xn, 0, zn)], pn1 = phi[g.sidx(xn, 1, zn)];
  const bool all_liquid = pa0 < 0.0f && pa1 < 0.0f && pn0 < 0.0f && pn1 < 0.0f;
  const bool all_air = !(pa0 < 0.0f) && !(pa1 < 0.0f) && !(pn0 < 0.0f) && !(pn1 < 0.0f);
  if (all_liquid) {
    M[0] = (double)m;
    M[1] = 0.0;
    M[2] = (double)m * ((double)m * m - 1.0) / 12.0;
    return;
  }
  M[0] = M[1] = M[2] = 0.0;
  if (all_air) return;
  const double ym = 0.5 * (m - 1);
  for (int y = 0; y < m; ++y) {
    const RowRep<double> ra = row_rep<double>(g, phi, b, x, z, y);
    const RowRep<double> rn = row_rep<double>(g, phi, bn, xn, zn, y);
    const double w = fwr<double>(ra.phi, rn.phi);
    const double t = y - ym;
    M[0] += w;
    M[1] += w * t;
    M[2] += w * t * t;
  }
}

__global__ void k_tall_columns(const TallPart* parts, int np) {
  FP_PART(parts, np);
  const TallGeo g = geo_of(G);
  const int* __restrict__ bases = G.bases;
  const float* __restrict__ phi = G.phi;
  TallCol* __restrict__ cols = G.cols;
  FP_STRIDE(c, (size_t)G.ncol) {
    const int x = (int)(c / (size_t)g.nz), z = (int)(c % (size_t)g.nz);
    const int b = bases[c];
    TallCol out;
    out.pad = 0.0;
    for (int a = 0; a < 3; ++a) out.mx[a] = out.mz[a] = 0.0;
    if (x + 1 < g.nx) moments_of(g, phi, x, z, b, x + 1, z, bases[g.col(x + 1, z)], out.mx);
    if (z + 1 < g.nz) moments_of(g, phi, x, z, b, x, z + 1, bases[g.col(x, z + 1)], out.mz);
    const float pb = phi[g.sidx(x, 0, z)], pt = phi[g.sidx(x, 1, z)];
    if (pb < 0.0f && pt < 0.0f) {
      out.wint = (double)(b - 1);
    } else if (!(pb < 0.0f) && !(pt < 0.0f)) {
      out.wint = 0.0;
    } else {
      double s = 0.0;
      for (int y = 0; y + 1 < b; ++y) {
        const RowRep<double> r0 = row_rep<double>(g, phi, b, x, z, y);
        const RowRep<double> r1 = row_rep<double>(g, phi, b, x, z, y + 1);
        s += fwr<double>(r0.phi, r1.phi);
      }
      out.wint = s;
    }
    cols[c] = out;
    int tp = 1;
    for (int py = g.P - 1; py >= 2; --py)
      if (b + py - 2 < g.ny && phi[g.sidx(x, py, z)] < 0.0f) {
        tp = py;
        break;
      }
    G.top[c] = tp;
  }
}

// Block-Jacobi inverses (band diagonal, tall 2 x 2 blocks), the lattice
// surrogate of the sample level, and the warm-start mask of x.
template <typename Real>
struct ZeroVec {
  __device__ __forceinline__ Real operator()(size_t) const { return (Real)0; }
};

__device__ __forceinline__ bool sample_unknown(const TallGeo& g, const int* __restrict__ bases,
                                               const float* __restrict__ phi, int x, int py, int z) {
  return is_unknown(g, phi, bases[g.col(x, z)], x, py, z);
}

__global__ void k_tall_diag(const TallPart* parts, int np) {
  FP_PART(parts, np);
  const TallGeo g = geo_of(G);
  const int* __restrict__ bases = G.bases;
  const float* __restrict__ phi = G.phi;
  const TallCol* __restrict__ cols = G.cols;
  float* __restrict__ dinv = G.dinv;
  fp_f32x4* __restrict__ binv = G.blk;
  double* __restrict__ xv = G.x;
  TALL_LINES(g) {
    TALL_LINE(g, idx, x, py, z);
    const int b = bases[g.col(x, z)];
    const bool unk = is_unknown(g, phi, b, x, py, z);
    if (!unk) xv[idx] = 0.0;
    float di = 0.0f;
    if (unk && py >= 2) {
      float ap, dg;
      band_row<float>(g, bases, phi, x, py, z, b, ZeroVec<float>{}, ap, dg);
      di = dg > 0.0f ? 1.0f / dg : 0.0f;
    }
    dinv[idx] = di;
    if (py == 0) {
      const PairOut<double> o = tall_pair<double>(g, bases, phi, cols, x, z, b, ZeroVec<double>{});
      const bool ub = unk, ut = is_unknown(g, phi, b, x, 1, z);
      fp_f32x4 inv = {0.0f, 0.0f, 0.0f, 0.0f};
      const double det = o.bb * o.tt - o.bt * o.bt;
      if (ub && ut && det > 0.0) {
        inv.x = (float)(o.tt / det);
        inv.y = (float)(-o.bt / det);
        inv.z = (float)(o.bb / det);
      } else {
        if (ub && o.bb > 0.0) inv.x = (float)(1.0 / o.bb);
        if (ut && o.tt > 0.0) inv.z = (float)(1.0 / o.tt);
      }
      binv[g.col(x, z)] = inv;
    }
  }
}

}  // namespace fp

#include "fp_tall_galerkin.hpp"

namespace fp {

// --------------------------------------------------------------------------
// right-hand side and divergence: b_i = (rho h/dt) sum_f c_f,i u_f (E^T of
// the fine divergence).  mode 0: b, ||b||^2, liquid count and max |div| of
// the inputs; mode 1: max |div| of the outputs only.
// --------------------------------------------------------------------------
struct Slots {
  const float* u;
  const float* v;
  const float* w;
};

// Flux sums (sum_f c_f u_f) of a band sample and of the tall pair.
__device__ double band_flux(const TallGeo& g, const int* __restrict__ bases, const Slots& s, int x, int py, int z,
                            int b) {
  const int y = b + py - 2;
  double f = (double)s.v[g.sidx(x, py - 1, z)];
  if (y + 1 < g.ny) f -= (double)s.v[g.sidx(x, py, z)];
#pragma unroll
  for (int dir = 0; dir < 4; ++dir) {
    int xn, zn;
    if (!neighbour(g, x, z, dir, xn, zn)) continue;
    const int m = min(b, bases[g.col(xn, zn)]);
    const int slot = 2 + y - m;
    switch (dir) {
      case 0: f += (double)s.u[g.uidx(x, slot, z)]; break;
      case 1: f -= (double)s.u[g.uidx(x + 1, slot, z)]; break;
      case 2: f += (double)s.w[g.widx(x, slot, z)]; break;
      default: f -= (double)s.w[g.widx(x, slot, z + 1)]; break;
    }
  }
  return f;
}

__device__ void pair_flux(const TallGeo& g, const int* __restrict__ bases, const Slots& s, int x, int z, int b,
                          double& fb, double& ft) {
  const double kap = 1.0 / (double)(b - 1);
  const double v0 = (double)s.v[g.sidx(x, 0, z)];
  fb = -v0;
  ft = v0;
  if (b < g.ny) ft -= (double)s.v[g.sidx(x, 1, z)];
#pragma unroll
  for (int dir = 0; dir < 4; ++dir) {
    int xn, zn;
    if (!neighbour(g, x, z, dir, xn, zn)) continue;
    const int bn = bases[g.col(xn, zn)];
    const int m = min(b, bn);
    const double ym = 0.5 * (m - 1), Q = (double)m * ((double)m * m - 1.0) / 12.0;
    const double sign = (dir == 0 || dir == 2) ? 1.0 : -1.0;  // this column on the + side of -x/-z faces
    size_t s0, sstride;
    const float* arr;
    if (dir < 2) {
      const int fx = dir == 0 ? x : x + 1;
      s0 = g.uidx(fx, 0, z);
      sstride = (size_t)g.nz;
      arr = s.u;
    } else {
      const int fz = dir == 2 ? z : z + 1;
      s0 = g.widx(x, 0, fz);
      sstride = (size_t)g.nz + 1;
      arr = s.w;
    }
    const double U0 = (double)arr[s0], U1 = (double)arr[s0 + sstride];
    const double c0 = 1.0 - ym * kap, c1 = -kap, d0 = ym * kap, d1 = kap;
    fb += sign * (c0 * m * U0 + c1 * Q * U1);
    ft += sign * (d0 * m * U0 + d1 * Q * U1);
    for (int y = m; y < b; ++y) {
      const double k = (double)y * kap;
      const double U = (double)arr[s0 + (size_t)(2 + y - m) * sstride];
      fb += sign * (1.0 - k) * U;
      ft += sign * k * U;
    }
  }
}

__global__ void k_tall_rhs(const TallPart* parts, int np, int mode) {
  FP_PART(parts, np);
  const TallGeo g = geo_of(G);
  const int* __restrict__ bases = G.bases;
  const float* __restrict__ phi = G.phi;
  const Slots s = mode == 0 ? Slots{G.in_u, G.in_v, G.in_w} : Slots{G.out_u, G.out_v, G.out_w};
  if (mode == 1 && (!s.u || !s.v || !s.w)) return;
  double* __restrict__ bv = G.b;
  const double sf = G.s_flux, inv_h = G.inv_h;
  double acc[3] = {0.0, 0.0, 0.0};
  float dmax = 0.0f;
  TALL_LINES(g) {
    TALL_LINE(g, idx, x, py, z);
    const int b = bases[g.col(x, z)];
    if (py == 1) continue;  // written with the bottom sample
    if (py >= 2) {
      double val = 0.0;
      if (is_unknown(g, phi, b, x, py, z)) {
        const double f = band_flux(g, bases, s, x, py, z, b);
        val = sf * f;
        dmax = fmaxf(dmax, (float)(fabs(f) * inv_h));
        acc[0] += val * val;
        acc[1] += 1.0;
        acc[2] += val;
      }
      if (mode == 0) bv[idx] = val;
    } else {
      double fb, ft;
      pair_flux(g, bases, s, x, z, b, fb, ft);
      const bool ub = is_unknown(g, phi, b, x, 0, z), ut = is_unknown(g, phi, b, x, 1, z);
      const double vb = ub ? sf * fb : 0.0, vt = ut ? sf * ft : 0.0;
      const double support = 0.5 * (double)b;
      if (ub) dmax = fmaxf(dmax, (float)(fabs(fb) * inv_h / support));
      if (ut) dmax = fmaxf(dmax, (float)(fabs(ft) * inv_h / support));
      acc[0] += vb * vb + vt * vt;
      acc[1] += (ub ? 1.0 : 0.0) + (ut ? 1.0 : 0.0);
      acc[2] += vb + vt;
      if (mode == 0) {
        bv[idx] = vb;
        bv[g.sidx(x, 1, z)] = vt;
      }
    }
  }
  block_partials_max<3>(acc, dmax, G.partials, blk, nblk);
}

// max |div| of the projected field into State::divmax_after.
template <class P>
__global__ void k_tall_fin_divafter(const P* parts) {
  const P p = parts[blockIdx.x];
  if (!p.g->out_u || !p.g->out_v || !p.g->out_w) return;
  double out[3];
  float mx;
  sum_partials_max<3>(p.g->partials, p.nblocks, out, mx);
  if (threadIdx.x == 0) p.g->st->divmax_after = __float_as_uint(mx);
}

// --------------------------------------------------------------------------
// operator applications at the sample level
// --------------------------------------------------------------------------
// FP64 true residual r = b - A x (FP32 copy for the inner solve), ||r||^2 and max|r|.
__global__ void k_tall_true(const TallPart* parts, int np) {
  FP_PART(parts, np);
  if (G.st->outer_done) return;
  const TallGeo g = geo_of(G);
  const int* __restrict__ bases = G.bases;
  const float* __restrict__ phi = G.phi;
  const double* __restrict__ bv = G.b;
  float* __restrict__ r = G.r;
  const VecD xv{G.x};
  const double bmean = G.st->bmean;  // removed on an unanchored grid (else 0)
  double acc[2] = {0.0, 0.0};
  float rmax = 0.0f;
  TALL_LINES(g) {
    TALL_LINE(g, idx, x, py, z);
    if (py == 1) continue;
    const int b = bases[g.col(x, z)];
    if (py >= 2) {
      double res = 0.0;
      if (is_unknown(g, phi, b, x, py, z)) {
        double ap, dg;
        band_row<double>(g, bases, phi, x, py, z, b, xv, ap, dg);
        res = (bv[idx] - bmean) - ap;
        acc[1] += xv(idx);
      }
      r[idx] = (float)res;
      acc[0] += res * res;
      rmax = fmaxf(rmax, fabsf((float)res));
    } else {
      const PairOut<double> o = tall_pair<double>(g, bases, phi, G.cols, x, z, b, xv);
      const size_t it = g.sidx(x, 1, z);
      const bool ub = is_unknown(g, phi, b, x, 0, z), ut = is_unknown(g, phi, b, x, 1, z);
      const double rb = ub ? (bv[idx] - bmean) - o.ab : 0.0;
      const double rt = ut ? (bv[it] - bmean) - o.at : 0.0;
      r[idx] = (float)rb;
      r[it] = (float)rt;
      acc[0] += rb * rb + rt * rt;
      acc[1] += (ub ? xv(idx) : 0.0) + (ut ? xv(it) : 0.0);
      rmax = fmaxf(rmax, fmaxf(fabsf((float)rb), fabsf((float)rt)));
    }
  }
  block_partials_max<2>(acc, rmax, G.partials, blk, nblk);
}

// Block-Jacobi sweep: out = in + omega B^-1 (r - A in); first: in = 0.
// sel: 0 reads z writes z2, 1 reads z2 writes z.
__global__ void k_tall_jacobi(const TallPart* parts, int np, int sel, int first) {
  FP_PART(parts, np);
  if (G.st->done) return;
  const TallGeo g = geo_of(G);
  const int* __restrict__ bases = G.bases;
  const float* __restrict__ phi = G.phi;
  const float* __restrict__ r = G.r;
  const float* __restrict__ zin = sel ? G.z2 : G.z;
  float* __restrict__ zout = first ? G.z : (sel ? G.z : G.z2);
  const float om = G.omega;
  const VecF zv{zin};
  TALL_ITEMS(g) {
    TALL_ITEM(g, x, k, z);
    const int b = bases[g.col(x, z)];
    if (k > 0) {
      const int py0 = 2 + kSegRows * (k - 1);
      if (py0 > G.top[g.col(x, z)]) {  // all air: the zero correction
#pragma unroll
        for (int j = 0; j < kSegRows; ++j)
          if (py0 + j < g.P) zout[g.sidx(x, py0 + j, z)] = 0.0f;
        continue;
      }
      float ap[kSegRows], dg[kSegRows];
      if (!first) band_segment(g, bases, phi, x, py0, z, b, zv, ap, dg);
#pragma unroll
      for (int j = 0; j < kSegRows; ++j) {
        if (py0 + j >= g.P) break;
        const size_t idx = g.sidx(x, py0 + j, z);
        const float di = G.dinv[idx];
        float out = 0.0f;
        if (di != 0.0f) out = first ? om * di * r[idx] : zin[idx] + om * di * (r[idx] - ap[j]);
        zout[idx] = out;
      }
    } else {
      const size_t idx = g.sidx(x, 0, z);
      const fp_f32x4 inv = G.blk[g.col(x, z)];
      const size_t it = g.sidx(x, 1, z);
      float rb = r[idx], rt = r[it];
      float ob = 0.0f, ot = 0.0f;
      if (!first) {
        const PairOut<float> o = tall_pair<float>(g, bases, phi, G.cols, x, z, b, zv);
        rb -= o.ab;
        rt -= o.at;
        ob = zin[idx];
        ot = zin[it];
      }
      zout[idx] = inv.x != 0.0f || inv.y != 0.0f ? ob + om * (inv.x * rb + inv.y * rt) : 0.0f;
      zout[it] = inv.z != 0.0f || inv.y != 0.0f ? ot + om * (inv.y * rb + inv.z * rt) : 0.0f;
    }
  }
}

// t = r - A z (sel: which buffer holds z).
__global__ void k_tall_residual(const TallPart* parts, int np, int sel) {
  FP_PART(parts, np);
  if (G.st->done) return;
  const TallGeo g = geo_of(G);
  const int* __restrict__ bases = G.bases;
  const float* __restrict__ phi = G.phi;
  const float* __restrict__ r = G.r;
  const VecF zv{sel ? G.z2 : G.z};
  float* __restrict__ t = G.t;
  TALL_ITEMS(g) {
    TALL_ITEM(g, x, k, z);
    const int b = bases[g.col(x, z)];
    if (k > 0) {
      const int py0 = 2 + kSegRows * (k - 1);
      if (py0 > G.top[g.col(x, z)]) {  // all air: no residual
#pragma unroll
        for (int j = 0; j < kSegRows; ++j)
          if (py0 + j < g.P) t[g.sidx(x, py0 + j, z)] = 0.0f;
        continue;
      }
      float ap[kSegRows], dg[kSegRows];
      band_segment(g, bases, phi, x, py0, z, b, zv, ap, dg);
#pragma unroll
      for (int j = 0; j < kSegRows; ++j) {
        if (py0 + j >= g.P) break;
        const size_t idx = g.sidx(x, py0 + j, z);
        t[idx] = is_unknown(g, phi, b, x, py0 + j, z) ? r[idx] - ap[j] : 0.0f;
      }
    } else {
      const size_t idx = g.sidx(x, 0, z);
      const PairOut<float> o = tall_pair<float>(g, bases, phi, G.cols, x, z, b, zv);
      const size_t it = g.sidx(x, 1, z);
      t[idx] = is_unknown(g, phi, b, x, 0, z) ? r[idx] - o.ab : 0.0f;
      t[it] = is_unknown(g, phi, b, x, 1, z) ? r[it] - o.at : 0.0f;
    }
  }
}

// CG direction: d = z + beta d_old (formed on the fly), q = A d, d.q.
__global__ void k_tall_direction(const TallPart* parts, int np) {
  FP_PART(parts, np);
  const State* st = G.st;
  if (st->done) return;
  const TallGeo g = geo_of(G);
  const int* __restrict__ bases = G.bases;
  const float* __restrict__ phi = G.phi;
  const int par = st->iter & 1;
  const float* __restrict__ zz = G.zfinal ? G.z2 : G.z;
  const float* __restrict__ dold = G.d[par];
  float* __restrict__ dnew = G.d[par ^ 1];
  float* __restrict__ q = G.q;
  const VecDir dv{zz, dold, (float)st->beta};
  double acc[1] = {0.0};
  TALL_ITEMS(g) {
    TALL_ITEM(g, x, k, z);
    const int b = bases[g.col(x, z)];
    if (k > 0) {
      const int py0 = 2 + kSegRows * (k - 1);
      if (py0 > G.top[g.col(x, z)]) {  // all air: d and q vanish
#pragma unroll
        for (int j = 0; j < kSegRows; ++j)
          if (py0 + j < g.P) {
            dnew[g.sidx(x, py0 + j, z)] = 0.0f;
            q[g.sidx(x, py0 + j, z)] = 0.0f;
          }
        continue;
      }
      float ap[kSegRows], dg[kSegRows];
      band_segment(g, bases, phi, x, py0, z, b, dv, ap, dg);
#pragma unroll
      for (int j = 0; j < kSegRows; ++j) {
        if (py0 + j >= g.P) break;
        const size_t idx = g.sidx(x, py0 + j, z);
        const bool u = is_unknown(g, phi, b, x, py0 + j, z);
        const float dd = u ? dv(idx) : 0.0f, qa = u ? ap[j] : 0.0f;
        dnew[idx] = dd;
        q[idx] = qa;
        acc[0] += (double)dd * (double)qa;
      }
    } else {
      const size_t idx = g.sidx(x, 0, z);
      const PairOut<float> o = tall_pair<float>(g, bases, phi, G.cols, x, z, b, dv);
      const size_t it = g.sidx(x, 1, z);
      const bool ub = is_unknown(g, phi, b, x, 0, z), ut = is_unknown(g, phi, b, x, 1, z);
      const float db = ub ? dv(idx) : 0.0f, dt = ut ? dv(it) : 0.0f;
      const float qb = ub ? o.ab : 0.0f, qt = ut ? o.at : 0.0f;
      dnew[idx] = db;
      dnew[it] = dt;
      q[idx] = qb;
      q[it] = qt;
      acc[0] += (double)db * qb + (double)dt * qt;
    }
  }
  block_partials<1>(acc, G.partials, blk, nblk);
}

// ---- plain vector kernels ---------------------------------------------------
__global__ void k_tall_rz(const TallPart* parts, int np) {
  FP_PART(parts, np);
  if (G.st->done) return;
  const float* __restrict__ r = G.r;
  const float* __restrict__ zz = G.zfinal ? G.z2 : G.z;
  double acc[1] = {0.0};
  FP_STRIDE(i, (size_t)G.n) acc[0] += (double)r[i] * (double)zz[i];
  block_partials<1>(acc, G.partials, blk, nblk);
}

__global__ void k_tall_update(const TallPart* parts, int np) {
  FP_PART(parts, np);
  const State* st = G.st;
  if (st->done) return;
  const float alpha = (float)st->alpha;
  const float* __restrict__ d = G.d[(st->iter & 1) ^ 1];
  const float* __restrict__ q = G.q;
  float* __restrict__ e = G.e;
  float* __restrict__ r = G.r;
  double acc[1] = {0.0};
  float rmax = 0.0f;
  FP_STRIDE(i, (size_t)G.n) {
    e[i] += alpha * d[i];
    const float rn = r[i] - alpha * q[i];
    r[i] = rn;
    acc[0] += (double)rn * (double)rn;
    rmax = fmaxf(rmax, fabsf(rn));
  }
  block_partials_max<1>(acc, rmax, G.partials, blk, nblk);
}

// which: 0 x (unless warm), 1 e, 2 d[0]
__global__ void k_tall_zero(const TallPart* parts, int np, int which, int skip_converged) {
  FP_PART(parts, np);
  if (skip_converged && G.st->outer_done) return;
  if (which == 0) {
    if (G.warm) return;
    FP_STRIDE(i, (size_t)G.n) G.x[i] = 0.0;
    return;
  }
  float* p = which == 1 ? G.e : G.d[0];
  FP_STRIDE(i, (size_t)G.n) p[i] = 0.0f;
}

__global__ void k_tall_accumulate(const TallPart* parts, int np) {
  FP_PART(parts, np);
  const State* st = G.st;
  if (st->outer_done) return;
  const bool keep = !(st->flags & 0x7);
  double* __restrict__ x = G.x;
  float* __restrict__ e = G.e;
  FP_STRIDE(i, (size_t)G.n) {
    if (keep) x[i] += (double)e[i];
    e[i] = 0.0f;
  }
}

// z += P e1 on unknown samples (piecewise constant: the level-1 node holding
// the sample; sel: z buffer).
__global__ void k_tall_prolong0(const TallPart* parts, int np, int sel) {
  FP_PART(parts, np);
  if (G.st->done) return;
  const TallGeo g = geo_of(G);
  const TallLevel& C = G.lv[1];
  float* __restrict__ zz = sel ? G.z2 : G.z;
  const float* __restrict__ ec = final_e(G, 1);
  const float* __restrict__ dinv = G.dinv;
  const fp_f32x4* __restrict__ binv = G.blk;
  TALL_LINES(g) {
    TALL_LINE(g, idx, x, py, z);
    bool unk;
    if (py >= 2) {
      unk = dinv[idx] != 0.0f;
    } else {
      const fp_f32x4 iv = binv[g.col(x, z)];
      unk = py == 0 ? (iv.x != 0.0f || iv.y != 0.0f) : (iv.z != 0.0f || iv.y != 0.0f);
    }
    if (!unk) continue;
    const int Y = py < 2 ? py : 2 + (py - 2) / C.cy;
    zz[idx] += ec[((size_t)(x >> 1) * C.ny + Y) * C.nz + (z >> 1)];
  }
}

// --------------------------------------------------------------------------
// outputs: pressure and the slot subtract (FP64)
// --------------------------------------------------------------------------
__global__ void k_tall_pressure_out(const TallPart* parts, int np) {
  FP_PART(parts, np);
  float* __restrict__ p = G.out_p;
  if (!p) return;
  const TallGeo g = geo_of(G);
  // an unanchored grid's pressure is defined up to a constant: its mean is 0
  const State* st = G.st;
  const double xmean = st->nullspace ? st->xsum / (double)st->liquid : 0.0;
  TALL_LINES(g) {
    TALL_LINE(g, idx, x, py, z);
    p[idx] = is_unknown(g, G.phi, G.bases[g.col(x, z)], x, py, z) ? (float)(G.x[idx] - xmean) : 0.0f;
  }
}

// Whether a fine row carries an unknown sample with a nonzero coefficient.
__device__ __forceinline__ bool row_live(const TallGeo& g, const float* __restrict__ phi, const int* __restrict__ bases,
                                         const RowRep<double>& r, int x, int z) {
  if (r.n == 0) return false;
  const int b = bases[g.col(x, z)];
  auto unk = [&](size_t i) {
    const int py = (int)((i / (size_t)g.nz) % (size_t)g.P);
    return is_unknown(g, phi, b, x, py, z);
  };
  bool live = r.c0 != 0.0 && unk(r.i0);
  if (r.n == 2) live = live || (r.c1 != 0.0 && unk(r.i1));
  return live;
}

// axis 0: u (nx+1, S, nz); 1: v (nx, P, nz); 2: w (nx, S, nz+1)
__global__ void k_tall_subtract(const TallPart* parts, int np, int axis) {
  FP_PART(parts, np);
  const float* __restrict__ vin = axis == 0 ? G.in_u : (axis == 1 ? G.in_v : G.in_w);
  float* __restrict__ vout = axis == 0 ? G.out_u : (axis == 1 ? G.out_v : G.out_w);
  if (!vout) return;
  const TallGeo g = geo_of(G);
  const int* __restrict__ bases = G.bases;
  const float* __restrict__ phi = G.phi;
  const double* __restrict__ xv = G.x;
  const double sc = G.sub_scale;
  const int fx = g.nx + (axis == 0), fs = axis == 1 ? g.P : g.S, fz = g.nz + (axis == 2);
  const size_t nf = (size_t)fx * fs * fz;
  FP_STRIDE(f, nf) {
    const int k = (int)(f % (size_t)fz);
    const size_t tt = f / (size_t)fz;
    const int s = (int)(tt % (size_t)fs);
    const int i = (int)(tt / (size_t)fs);
    double out = 0.0;
    const double U = (double)vin[f];
    if (axis == 1) {
      const int x = i, z = k;
      const int b = bases[g.col(x, z)];
      const size_t ib = g.sidx(x, 0, z), it = g.sidx(x, 1, z);
      if (s == 0) {
        const double kap = 1.0 / (double)(b - 1);
        const double wint = G.cols[g.col(x, z)].wint;
        const bool live = wint > 0.0 || is_unknown(g, phi, b, x, 0, z) || is_unknown(g, phi, b, x, 1, z);
        if (live) out = U - sc * kap * (xv[it] - xv[ib]) * wint / (double)(b - 1);
      } else {
        const int ylo = b + s - 2;
        if (ylo + 1 < g.ny) {
          const RowRep<double> lo = row_rep<double>(g, phi, b, x, z, ylo);
          const RowRep<double> hi = row_rep<double>(g, phi, b, x, z, ylo + 1);
          const double w = fwr<double>(lo.phi, hi.phi);
          const bool live = w != 0.0 || row_live(g, phi, bases, lo, x, z) || row_live(g, phi, bases, hi, x, z);
          if (live) out = w != 0.0 ? U - sc * w * (row_value(hi, VecD{xv}) - row_value(lo, VecD{xv})) : U;
        }
      }
    } else {
      // face between A (-) and N (+)
      int xa, za, xn, zn;
      if (axis == 0) {
        xa = i - 1; xn = i; za = zn = k;
        if (i == 0 || i == g.nx) { vout[f] = 0.0f; continue; }
      } else {
        za = k - 1; zn = k; xa = xn = i;
        if (k == 0 || k == g.nz) { vout[f] = 0.0f; continue; }
      }
      const int ba = bases[g.col(xa, za)], bn = bases[g.col(xn, zn)];
      const int m = min(ba, bn), M = max(ba, bn);
      if (s < 2) {
        const double* Mo = axis == 0 ? G.cols[g.col(xa, za)].mx : G.cols[g.col(xa, za)].mz;
        const double ka = 1.0 / (double)(ba - 1), kn = 1.0 / (double)(bn - 1);
        const double ym = 0.5 * (m - 1);
        const double pba = xv[g.sidx(xa, 0, za)], pta = xv[g.sidx(xa, 1, za)];
        const double pbn = xv[g.sidx(xn, 0, zn)], ptn = xv[g.sidx(xn, 1, zn)];
        const double G0 = (pbn + (ptn - pbn) * ym * kn) - (pba + (pta - pba) * ym * ka);
        const double G1 = (ptn - pbn) * kn - (pta - pba) * ka;
        const bool live = Mo[0] > 0.0 || is_unknown(g, phi, ba, xa, 0, za) || is_unknown(g, phi, ba, xa, 1, za) ||
                          is_unknown(g, phi, bn, xn, 0, zn) || is_unknown(g, phi, bn, xn, 1, zn);
        if (live) {
          if (s == 0) {
            out = U - sc * (G0 * Mo[0] + G1 * Mo[1]) / (double)m;
          } else {
            const double Q = (double)m * ((double)m * m - 1.0) / 12.0;
            out = U - sc * (G0 * Mo[1] + G1 * Mo[2]) / Q;
          }
        }
      } else {
        const int y = m + s - 2;
        if (y < min(g.ny, M + g.B)) {
          const RowRep<double> ra = row_rep<double>(g, phi, ba, xa, za, y);
          const RowRep<double> rn = row_rep<double>(g, phi, bn, xn, zn, y);
          const double w = fwr<double>(ra.phi, rn.phi);
          const bool live = w != 0.0 || row_live(g, phi, bases, ra, xa, za) || row_live(g, phi, bases, rn, xn, zn);
          if (live) out = w != 0.0 ? U - sc * w * (row_value(rn, VecD{xv}) - row_value(ra, VecD{xv})) : U;
        }
      }
    }
    vout[f] = (float)out;
  }
}

// Layout probe: fine row of each sample and its unknown flag.
__global__ void k_tall_layout(const TallPart* parts, int np, int* rows, unsigned char* unk) {
  FP_PART(parts, np);
  const TallGeo g = geo_of(G);
  TALL_LINES(g) {
    TALL_LINE(g, idx, x, py, z);
    const int b = G.bases[g.col(x, z)];
    rows[idx] = py == 0 ? 0 : (py == 1 ? b - 1 : b + py - 2);
    unk[idx] = is_unknown(g, G.phi, b, x, py, z) ? 1 : 0;
  }
}

// FP64 operator probe: out = A p.
__global__ void k_tall_apply_d(const TallPart* parts, int np, const double* p, double* out) {
  FP_PART(parts, np);
  const TallGeo g = geo_of(G);
  const VecD pv{p};
  TALL_LINES(g) {
    TALL_LINE(g, idx, x, py, z);
    if (py == 1) continue;
    const int b = G.bases[g.col(x, z)];
    if (py >= 2) {
      double ap = 0.0, dg;
      if (is_unknown(g, G.phi, b, x, py, z)) band_row<double>(g, G.bases, G.phi, x, py, z, b, pv, ap, dg);
      out[idx] = ap;
    } else {
      const PairOut<double> o = tall_pair<double>(g, G.bases, G.phi, G.cols, x, z, b, pv);
      out[idx] = is_unknown(g, G.phi, b, x, 0, z) ? o.ab : 0.0;
      out[g.sidx(x, 1, z)] = is_unknown(g, G.phi, b, x, 1, z) ? o.at : 0.0;
    }
  }
}

// ==========================================================================
// host side
// ==========================================================================
int plan_tall_levels(int nx, int band, int nz, int dims[][3], int cy[]) {
  int L = 0;
  int cx = nx, cb = band, cz = nz;
  dims[0][0] = cx;
  dims[0][1] = cb + 2;
  dims[0][2] = cz;
  cy[0] = 1;
  L = 1;
  // always at least one lattice level; coarsen until the lattice is small
  while (L < kTallMaxLevels) {
    const long long n = (long long)dims[L - 1][0] * dims[L - 1][1] * dims[L - 1][2];
    if (L >= 2 && n <= kTallMaxCoarse) break;
    const int nxc = (cx + 1) / 2, nzc = (cz + 1) / 2, nbc = cb > 1 ? (cb + 1) / 2 : cb;
    if (L >= 2 && nxc == cx && nzc == cz && nbc == cb) break;
    cy[L] = cb > 1 ? 2 : 1;
    cx = nxc;
    cz = nzc;
    cb = nbc;
    dims[L][0] = cx;
    dims[L][1] = cb + 2;
    dims[L][2] = cz;
    ++L;
  }
  const long long nc = (long long)dims[L - 1][0] * dims[L - 1][1] * dims[L - 1][2];
  return nc <= kTallMaxCoarse ? L : -1;
}

static std::atomic<int> g_next_tall_id{1 << 20};

TallGrid::~TallGrid() { release(); }

void TallGrid::release() {
  for (void* p : allocations_) (void)hipFree(p);
  allocations_.clear();
  if (rec_) (void)hipHostFree(rec_);
  rec_ = nullptr;
}

int TallGrid::alloc(void** p, size_t bytes) {
  if (bytes == 0) bytes = 16;
  bytes = (bytes + 255) & ~(size_t)255;
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) {
    set_error("hipMalloc(%zu): %s", bytes, hipGetErrorString(e));
    return FP_ERR_OUT_OF_MEMORY;
  }
  allocations_.push_back(*p);
  bytes_ += bytes;
  return FP_OK;
}

size_t TallGrid::estimate_bytes(int nx, int ny, int nz, int band, int max_delta) {
  int dims[kTallMaxLevels][3], cy[kTallMaxLevels];
  const int L = plan_tall_levels(nx, band, nz, dims, cy);
  if (L < 2) return 0;
  (void)ny;
  const int P = band + 2, S = band + 2 + max_delta;
  const size_t n = (size_t)nx * P * nz, ncol = (size_t)nx * nz;
  size_t total = 0;
  auto add = [&](size_t b) { total += (b + 255) & ~(size_t)255; };
  add(ncol * 4);                 // bases
  add(n * 4);                    // phi
  add(ncol * sizeof(TallCol));
  add(ncol * sizeof(int));        // top
  add(ncol * 16);                // blk
  add(n * 4);                    // dinv
  add(n * 8);                    // b
  add(n * 8);                    // x
  for (int i = 0; i < 8; ++i) add(n * 4);  // r e q d0 d1 z z2 t
  add((size_t)(nx + 1) * S * nz * 4);
  add(n * 4);
  add((size_t)nx * S * (nz + 1) * 4);
  const int K = (max_delta + 5) / 2;
  for (int l = 1; l < L; ++l) {
    const size_t nl = (size_t)dims[l][0] * dims[l][1] * dims[l][2];
    add(nl * 4 * (5 * K + 3));     // banded couplings
    add(nl * 4 * 5);               // dinv r e e2 t
  }
  add(2 * kTallMaxLevels * sizeof(int));  // reach
  add(kTallMaxCoarse * 4 + 16);
  add(16);
  add((size_t)kTallMaxCoarse * kTallMaxCoarse * 8);
  add(sizeof(State));
  add(4 * (size_t)kMaxBlocks * 8);
  add(sizeof(TallRec));
  return total;
}

int TallGrid::create(int nx, int ny, int nz, int band, int max_delta, double h, const int* bases_host) {
  int dims[kTallMaxLevels][3], cy[kTallMaxLevels];
  const int L = plan_tall_levels(nx, band, nz, dims, cy);
  if (L < 2) {
    set_error("no tall-cell multigrid plan for %d x %d columns, band %d", nx, nz, band);
    return FP_ERR_LAYOUT;
  }
  for (int i = 0; i < nx * nz; ++i) {
    const int b = bases_host[i];
    if (b < 2 || b > ny) {
      set_error("column %d: base %d outside [2, %d]", i, b, ny);
      return FP_ERR_LAYOUT;
    }
  }
  for (int x = 0; x < nx; ++x)
    for (int z = 0; z < nz; ++z) {
      const int b = bases_host[x * nz + z];
      if ((x + 1 < nx && std::abs(bases_host[(x + 1) * nz + z] - b) > max_delta) ||
          (z + 1 < nz && std::abs(bases_host[x * nz + z + 1] - b) > max_delta)) {
        set_error("columns (%d,%d): neighbouring bases differ by more than %d", x, z, max_delta);
        return FP_ERR_LAYOUT;
      }
    }
  HIPCHK(hipHostMalloc((void**)&rec_, sizeof(TallRec), hipHostMallocDefault));
  std::memset(rec_, 0, sizeof(TallRec));
  TallRec& R = *rec_;
  id_ = g_next_tall_id.fetch_add(1);
  R.nx = nx; R.ny = ny; R.nz = nz; R.band = band; R.P = band + 2; R.S = band + 2 + max_delta; R.D = max_delta;
  R.L = L;
  R.h = h;
  R.n = (long long)nx * R.P * nz;
  R.ncol = (long long)nx * nz;
  // no air row above any band: all in-domain samples liquid leave the
  // pressure unanchored (rows past the lid are outside the domain)
  R.ns_liquid = 0;
  for (int i = 0; i < nx * nz && R.ns_liquid >= 0; ++i)
    R.ns_liquid = bases_host[i] + band >= ny ? R.ns_liquid + 2 + (ny - bases_host[i]) : -1;
  R.nu = (long long)(nx + 1) * R.S * nz;
  R.nv = R.n;
  R.nw = (long long)nx * R.S * (nz + 1);
  R.omega = 0.8f;
  R.lb = L - 1;
  for (int l = 1; l < L; ++l)
    if ((long long)dims[l][0] * dims[l][1] * dims[l][2] <= kTallBottomNodes) {
      R.lb = l;
      break;
    }
  R.K = (max_delta + 5) / 2;
  if (R.K > kTallMaxK) {
    set_error("tall-cell max neighbour delta %d exceeds %d", max_delta, 2 * kTallMaxK - 5);
    return FP_ERR_LAYOUT;
  }
  const size_t n = (size_t)R.n, ncol = (size_t)R.ncol;
  int rc;
  int* bases = nullptr;
  if ((rc = alloc((void**)&bases, ncol * 4))) return rc;
  R.bases = bases;
  HIPCHK(hipMemcpy(bases, bases_host, ncol * 4, hipMemcpyHostToDevice));
  if ((rc = alloc((void**)&R.phi, n * 4))) return rc;
  if ((rc = alloc((void**)&R.cols, ncol * sizeof(TallCol)))) return rc;
  if ((rc = alloc((void**)&R.top, (size_t)R.ncol * sizeof(int)))) return rc;
  if ((rc = alloc((void**)&R.blk, ncol * 16))) return rc;
  if ((rc = alloc((void**)&R.dinv, n * 4))) return rc;
  if ((rc = alloc((void**)&R.b, n * 8))) return rc;
  if ((rc = alloc((void**)&R.x, n * 8))) return rc;
  float** vecs[] = {&R.r, &R.e, &R.q, &R.d[0], &R.d[1], &R.z, &R.z2, &R.t};
  for (float** v : vecs)
    if ((rc = alloc((void**)v, n * 4))) return rc;
  if ((rc = alloc((void**)&in_u_, (size_t)R.nu * 4))) return rc;
  if ((rc = alloc((void**)&in_v_, (size_t)R.nv * 4))) return rc;
  if ((rc = alloc((void**)&in_w_, (size_t)R.nw * 4))) return rc;
  for (int l = 1; l < L; ++l) {
    TallLevel& lv = R.lv[l];
    lv.nx = dims[l][0];
    lv.ny = dims[l][1];
    lv.nz = dims[l][2];
    lv.cy = cy[l];
    lv.n = (long long)lv.nx * lv.ny * lv.nz;
    const size_t nl = (size_t)lv.n;
    if ((rc = alloc((void**)&lv.w, nl * 4 * (size_t)(5 * R.K + 3)))) return rc;
    float** lvecs[] = {&lv.dinv, &lv.r, &lv.e, &lv.e2, &lv.t};
    for (float** v : lvecs)
      if ((rc = alloc((void**)v, nl * 4))) return rc;
  }
  R.lv[0].nx = nx;
  R.lv[0].ny = R.P;
  R.lv[0].nz = nz;
  R.lv[0].n = R.n;
  R.lv[0].cy = 1;
  R.max_coarse = kTallMaxCoarse;
  if ((rc = alloc((void**)&R.reach, 2 * kTallMaxLevels * sizeof(int)))) return rc;
  if ((rc = alloc((void**)&R.cmap, (size_t)kTallMaxCoarse * 4 + 16))) return rc;
  if ((rc = alloc((void**)&R.ncoarse, 16))) return rc;
  if ((rc = alloc((void**)&R.minv, (size_t)kTallMaxCoarse * kTallMaxCoarse * 8))) return rc;
  if ((rc = alloc((void**)&R.st, sizeof(State)))) return rc;
  if ((rc = alloc((void**)&R.partials, 4 * (size_t)kMaxBlocks * 8))) return rc;
  if ((rc = alloc((void**)&dev_rec_, sizeof(TallRec)))) return rc;
  HIPCHK(hipMemset(R.st, 0, sizeof(State)));
  HIPCHK(hipMemset(R.x, 0, n * 8));
  // vectors start at zero: unknown-free samples must read as 0 everywhere
  for (float** v : vecs) HIPCHK(hipMemset(*v, 0, n * 4));
  R.inv_h = 1.0 / h;
  R.max_iter = 1;
  HIPCHK(hipMemcpy(dev_rec_, rec_, sizeof(TallRec), hipMemcpyHostToDevice));
  return FP_OK;
}

int TallGrid::bind(hipStream_t s, const fp_inputs* in, fp_outputs* out, const fp_options* opt) {
  TallRec& R = *rec_;
  const size_t n = (size_t)R.n;
  host_io_ = in->memory == FP_MEMORY_HOST;
  if (host_io_) {
    HIPCHK(hipMemcpyAsync(R.phi, in->phi, n * 4, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(in_u_, in->u, (size_t)R.nu * 4, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(in_v_, in->v, (size_t)R.nv * 4, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(in_w_, in->w, (size_t)R.nw * 4, hipMemcpyHostToDevice, s));
    R.in_u = in_u_;
    R.in_v = in_v_;
    R.in_w = in_w_;
    R.out_u = out && out->u ? in_u_ : nullptr;
    R.out_v = out && out->v ? in_v_ : nullptr;
    R.out_w = out && out->w ? in_w_ : nullptr;
    R.out_p = out && out->pressure ? R.t : nullptr;
  } else {
    HIPCHK(hipMemcpyAsync(R.phi, in->phi, n * 4, hipMemcpyDeviceToDevice, s));
    R.in_u = in->u;
    R.in_v = in->v;
    R.in_w = in->w;
    R.out_u = out ? out->u : nullptr;
    R.out_v = out ? out->v : nullptr;
    R.out_w = out ? out->w : nullptr;
    R.out_p = out ? out->pressure : nullptr;
  }
  const double rho = in->rho, dt = in->dt, h = R.h;
  R.s_flux = rho * h / dt;
  R.inv_h = 1.0 / h;
  R.sub_scale = dt / (rho * h);
  R.tol2 = opt ? opt->relative_tolerance * opt->relative_tolerance : 0.0;
  R.max_iter = opt ? std::max(1, opt->max_iterations) : 1;
  R.warm = opt ? (opt->warm_start ? 1 : 0) : 0;
  R.rscale_bias = 1.0f;
  const int pre = opt ? opt->pre_sweeps : 2, post = opt ? opt->post_sweeps : 2;
  // z / z2 ping-pong: sweep 1 writes z, every later sweep flips
  const int sweeps = std::max(pre, 1) + post;
  R.zfinal = (sweeps - 1) & 1;
  return FP_OK;
}

int TallGrid::upload(hipStream_t s) {
  HIPCHK(hipMemcpyAsync(dev_rec_, rec_, sizeof(TallRec), hipMemcpyHostToDevice, s));
  return FP_OK;
}

int TallGrid::finish_outputs(hipStream_t s, const fp_inputs* in, fp_outputs* out) {
  (void)in;
  if (!host_io_ || !out) return FP_OK;
  const TallRec& R = *rec_;
  if (out->u) HIPCHK(hipMemcpyAsync(out->u, in_u_, (size_t)R.nu * 4, hipMemcpyDeviceToHost, s));
  if (out->v) HIPCHK(hipMemcpyAsync(out->v, in_v_, (size_t)R.nv * 4, hipMemcpyDeviceToHost, s));
  if (out->w) HIPCHK(hipMemcpyAsync(out->w, in_w_, (size_t)R.nw * 4, hipMemcpyDeviceToHost, s));
  if (out->pressure) HIPCHK(hipMemcpyAsync(out->pressure, R.t, (size_t)R.n * 4, hipMemcpyDeviceToHost, s));
  return FP_OK;
}

// ---- projector ---------------------------------------------------------------
struct TallProjector::Plan {
  std::vector<TallGrid*> grids;
  int pre = 2, post = 2, maxL = 2;
  std::vector<TallPart> host;
  TallPart* arena = nullptr;
  struct Pending {
    TallTable* t;
    size_t off;
  };
  std::vector<Pending> pending;
  TallTable one, samples, columns, faces[3];
  // grids for which lattice level l is not the coarsest: level-l nodes
  // (smoothing, residual, prolongation into l) and level-(l+1) nodes
  // (restriction and coarsening into l+1)
  // cnodes / ccoarser: the same for the levels the V-cycle smooths with grid
  // kernels (below the grid's bottom level lb)
  struct Level {
    TallTable nodes;
    TallTable coarser;
    TallTable cnodes;
    TallTable ccoarser;
  } lv[kTallMaxLevels];
  TallTable coarse;      // one workgroup per grid (coarsest lattice)
  hipGraphExec_t graph = nullptr;
  State* gathered = nullptr;
  State* gathered_host = nullptr;

  static int blocks_for(size_t items) {
    const size_t b = (items + FP_BLOCK - 1) / FP_BLOCK;
    return (int)std::max<size_t>(1, std::min<size_t>(b, 4096));
  }
  void add(TallTable& t, const std::vector<std::pair<TallGrid*, int>>& items) {
    t.n = (int)items.size();
    pending.push_back({&t, host.size()});
    int b0 = 0;
    for (const auto& it : items) {
      host.push_back(TallPart{b0, it.second, it.first->dev_rec()});
      b0 += (it.second + 7) & ~7;
    }
    t.blocks = b0;
  }
};

TallProjector::~TallProjector() {
  if (round_arena_) (void)hipFree(round_arena_);
  if (round_staging_) (void)hipHostFree(round_staging_);
  if (round_states_) (void)hipFree(round_states_);
  if (round_states_host_) (void)hipHostFree(round_states_host_);
  for (auto& kv : plans_) {
    Plan* p = kv.second;
    if (p->graph) (void)hipGraphExecDestroy(p->graph);
    if (p->arena) (void)hipFree(p->arena);
    if (p->gathered) (void)hipFree(p->gathered);
    if (p->gathered_host) (void)hipHostFree(p->gathered_host);
    delete p;
  }
  for (hipEvent_t e : timer_events_) (void)hipEventDestroy(e);
  if (dev_counts_) (void)hipFree(dev_counts_);
  if (host_counts_) (void)hipHostFree(host_counts_);
}

void TallProjector::forget(int grid_id) {
  for (auto it = plans_.begin(); it != plans_.end();) {
    bool uses = false;
    for (TallGrid* g : it->second->grids) uses = uses || g->id() == grid_id;
    if (uses) {
      Plan* p = it->second;
      if (p->graph) (void)hipGraphExecDestroy(p->graph);
      if (p->arena) (void)hipFree(p->arena);
      if (p->gathered) (void)hipFree(p->gathered);
      if (p->gathered_host) (void)hipHostFree(p->gathered_host);
      delete p;
      it = plans_.erase(it);
    } else {
      ++it;
    }
  }
}

void TallProjector::fill_tables(Plan* p, const std::vector<TallGrid*>& grids, int pre, int post) {
  p->grids = grids;
  p->pre = pre;
  p->post = post;
  p->maxL = 0;
  p->host.clear();
  p->pending.clear();
  std::vector<std::pair<TallGrid*, int>> one, samples, columns, faces[3];
  for (TallGrid* g : grids) {
    const TallRec& R = *g->host_rec();
    p->maxL = std::max(p->maxL, R.L);
    one.push_back({g, 1});
    samples.push_back({g, Plan::blocks_for((size_t)R.n)});
    columns.push_back({g, Plan::blocks_for((size_t)R.ncol)});
    faces[0].push_back({g, Plan::blocks_for((size_t)R.nu)});
    faces[1].push_back({g, Plan::blocks_for((size_t)R.nv)});
    faces[2].push_back({g, Plan::blocks_for((size_t)R.nw)});
  }
  p->add(p->one, one);
  p->add(p->samples, samples);
  p->add(p->columns, columns);
  for (int a = 0; a < 3; ++a) p->add(p->faces[a], faces[a]);
  p->add(p->coarse, one);
  for (int l = 1; l < p->maxL; ++l) {
    std::vector<std::pair<TallGrid*, int>> nodes, coarser, cnodes, ccoarser;
    for (TallGrid* g : grids) {
      const TallRec& R = *g->host_rec();
      if (l + 1 < R.L) {
        nodes.push_back({g, Plan::blocks_for((size_t)R.lv[l].n)});
        coarser.push_back({g, Plan::blocks_for((size_t)R.lv[l + 1].n)});
        if (l < R.lb) {
          cnodes.push_back(nodes.back());
          ccoarser.push_back(coarser.back());
        }
      }
    }
    p->add(p->lv[l].nodes, nodes);
    p->add(p->lv[l].coarser, coarser);
    p->add(p->lv[l].cnodes, cnodes);
    p->add(p->lv[l].ccoarser, ccoarser);
  }
  std::vector<std::pair<TallGrid*, int>> coarser;
  for (TallGrid* g : grids) coarser.push_back({g, Plan::blocks_for((size_t)g->host_rec()->lv[1].n)});
  p->add(p->lv[0].coarser, coarser);
}

TallProjector::Plan* TallProjector::plan_for(const std::vector<TallGrid*>& grids, const fp_options* opt) {
  std::vector<long long> key;
  for (TallGrid* g : grids) key.push_back(g->id());
  key.push_back(opt->pre_sweeps);
  key.push_back(opt->post_sweeps);
  auto it = plans_.find(key);
  if (it != plans_.end()) return it->second;
  Plan* p = new Plan();
  fill_tables(p, grids, opt->pre_sweeps, opt->post_sweeps);
  bool ok = hipMalloc((void**)&p->arena, std::max<size_t>(1, p->host.size()) * sizeof(TallPart)) == hipSuccess;
  ok = ok && hipMemcpy(p->arena, p->host.data(), p->host.size() * sizeof(TallPart), hipMemcpyHostToDevice) == hipSuccess;
  ok = ok && hipMalloc((void**)&p->gathered, grids.size() * sizeof(State)) == hipSuccess;
  ok = ok && hipHostMalloc((void**)&p->gathered_host, grids.size() * sizeof(State), hipHostMallocDefault) == hipSuccess;
  if (!ok) {
    set_error("tall-cell plan: device allocation failed");
    if (p->arena) (void)hipFree(p->arena);
    if (p->gathered) (void)hipFree(p->gathered);
    if (p->gathered_host) (void)hipHostFree(p->gathered_host);
    delete p;
    return nullptr;
  }
  for (const auto& pd : p->pending) pd.t->dev = p->arena + pd.off;
  plans_[key] = p;
  return p;
}

void TallProjector::timer_begin(hipStream_t s) {
  if ((size_t)(2 * timer_used_ + 2) > timer_events_.size()) {
    hipEvent_t a, b;
    if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) return;
    timer_events_.push_back(a);
    timer_events_.push_back(b);
  }
  (void)hipEventRecord(timer_events_[2 * timer_used_], s);
}

void TallProjector::timer_end(hipStream_t s) {
  if ((size_t)(2 * timer_used_ + 2) > timer_events_.size()) return;
  (void)hipEventRecord(timer_events_[2 * timer_used_ + 1], s);
  ++timer_used_;
}

#define TTIMED(cls, stmt)                     \
  do {                                        \
    const bool _t = (cls) == timed_kernel_;   \
    if (_t) timer_begin(s);                   \
    stmt;                                     \
    if (_t) timer_end(s);                     \
  } while (0)

int TallProjector::setup(hipStream_t s, Plan& p) {
  PLAUNCH(k_tall_columns, p.columns, s);
  PLAUNCH(k_tall_diag, p.samples, s);
  FLAUNCH(k_tall_reach_reset, p.one, s);
  PLAUNCH(k_tall_galerkin1, p.lv[0].coarser, s);
  PLAUNCH(k_tall_l1diag, p.lv[0].coarser, s, 1);
  for (int l = 1; l + 1 < p.maxL; ++l) {
    PLAUNCH(k_tall_galerkin, p.lv[l].coarser, s, l);
    PLAUNCH(k_tall_l1diag, p.lv[l].coarser, s, l + 1);
  }
  FLAUNCH(k_tall_coarse_setup, p.one, s);
  return FP_OK;
}

int TallProjector::enqueue_vcycle(hipStream_t s, Plan& p) {
  const int pre = std::max(p.pre, 1), post = p.post;
  // sample level: pre-smoothing (sweep 1 from zero writes z)
  int cur = 0;  // 0: z, 1: z2
  for (int sw = 0; sw < pre; ++sw) {
    TTIMED(FP_KERNEL_SMOOTH_FINE, PLAUNCH(k_tall_jacobi, p.samples, s, cur, sw == 0 ? 1 : 0));
    if (sw > 0) cur ^= 1;
  }
  TTIMED(FP_KERNEL_RESIDUAL_FINE, PLAUNCH(k_tall_residual, p.samples, s, cur));
  TTIMED(FP_KERNEL_RESTRICT_FINE, PLAUNCH(k_lat_restrict, p.lv[0].coarser, s, 0));
  // lattice levels: kTallCoarseSweeps sweeps before (the first from zero
  // writes e, then e and e2 alternate) and after the coarse correction; the
  // post-sweeps start where the pre-sweeps ended, so with equal counts the
  // correction of a smoothed level always ends in e2 (final_e); the coarsest
  // solve writes e
  const int cpre = (kTallCoarseSweeps - 1) & 1;  // buffer of e after the pre-sweeps
  // (levels from each grid's lb on run inside k_lat_bottom, one workgroup per
  // grid, with the coarsest solve)
  for (int l = 1; l + 1 < p.maxL; ++l) {
    PLAUNCH(k_lat_jacobi, p.lv[l].cnodes, s, l, 0, 1);
    for (int sw = 1, c = 0; sw < kTallCoarseSweeps; ++sw, c ^= 1) PLAUNCH(k_lat_jacobi, p.lv[l].cnodes, s, l, c, 0);
    PLAUNCH(k_lat_residual, p.lv[l].cnodes, s, l, cpre);
    PLAUNCH(k_lat_restrict, p.lv[l].ccoarser, s, l);
  }
  if (p.coarse.n > 0)
    hipLaunchKernelGGL(k_lat_bottom, dim3(p.coarse.n), dim3(kTallBottomThreads), 0, s, p.coarse.dev);
  for (int l = p.maxL - 2; l >= 1; --l) {
    PLAUNCH(k_lat_prolong, p.lv[l].cnodes, s, l, cpre);
    for (int sw = 0, c = cpre; sw < kTallCoarseSweeps; ++sw, c ^= 1) PLAUNCH(k_lat_jacobi, p.lv[l].cnodes, s, l, c, 0);
  }
  TTIMED(FP_KERNEL_PROLONG_FINE, PLAUNCH(k_tall_prolong0, p.samples, s, cur));
  for (int sw = 0; sw < post; ++sw) {
    TTIMED(FP_KERNEL_SMOOTH_FINE, PLAUNCH(k_tall_jacobi, p.samples, s, cur, 0));
    cur ^= 1;
  }
  return FP_OK;
}

int TallProjector::enqueue_iteration(hipStream_t s, Plan& p) {
  enqueue_vcycle(s, p);
  PLAUNCH(k_tall_rz, p.samples, s);
  FLAUNCH(k_fin_rz<TallPart>, p.samples, s);
  TTIMED(FP_KERNEL_CG_DIRECTION, PLAUNCH(k_tall_direction, p.samples, s));
  FLAUNCH(k_fin_dq<TallPart>, p.samples, s);
  TTIMED(FP_KERNEL_CG_UPDATE, PLAUNCH(k_tall_update, p.samples, s));
  FLAUNCH(k_fin_update<TallPart>, p.samples, s);
  return FP_OK;
}

int TallProjector::ensure_graph(hipStream_t s, Plan& p) {
  if (p.graph) return FP_OK;
  hipGraph_t g = nullptr;
  const int saved = timed_kernel_;
  timed_kernel_ = 0;
  HIPCHK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
  enqueue_iteration(s, p);
  enqueue_iteration(s, p);
  hipError_t e = hipStreamEndCapture(s, &g);
  timed_kernel_ = saved;
  if (e != hipSuccess) {
    set_error("hipStreamEndCapture: %s", hipGetErrorString(e));
    return FP_ERR_DEVICE;
  }
  e = hipGraphInstantiate(&p.graph, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (e != hipSuccess) {
    p.graph = nullptr;
    set_error("hipGraphInstantiate: %s", hipGetErrorString(e));
    return FP_ERR_DEVICE;
  }
  return FP_OK;
}

int TallProjector::poll(hipStream_t s, Plan& p, int* counts) {
  if (!dev_counts_) {
    HIPCHK(hipMalloc((void**)&dev_counts_, 16));
    HIPCHK(hipHostMalloc((void**)&host_counts_, 16, hipHostMallocDefault));
  }
  hipLaunchKernelGGL(k_status<TallPart>, dim3(1), dim3(FP_BLOCK), 0, s, p.one.dev, p.one.n, dev_counts_);
  HIPCHK(hipMemcpyAsync(host_counts_, dev_counts_, 8, hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  counts[0] = host_counts_[0];
  counts[1] = host_counts_[1];
  return FP_OK;
}

int TallProjector::run(hipStream_t s, const std::vector<TallGrid*>& grids, const fp_options* opt, fp_report* reps) {
  if (grids.empty()) return FP_OK;
  for (TallGrid* g : grids) {
    int rc = g->upload(s);
    if (rc) return rc;
  }
  Plan* pp = plan_for(grids, opt);
  if (!pp) return FP_ERR_OUT_OF_MEMORY;
  Plan& p = *pp;
  timed_kernel_ = opt->timed_kernel;
  timer_used_ = 0;
  hipEvent_t ev0, ev1;
  HIPCHK(hipEventCreate(&ev0));
  HIPCHK(hipEventCreate(&ev1));
  HIPCHK(hipEventRecord(ev0, s));
  FLAUNCH(k_reset_state<TallPart>, p.one, s);
  PLAUNCH(k_tall_zero, p.samples, s, 0, 0);
  setup(s, p);
  PLAUNCH(k_tall_rhs, p.samples, s, 0);
  FLAUNCH(k_fin_rhs<TallPart>, p.samples, s);
  PLAUNCH(k_tall_zero, p.samples, s, 1, 0);
  PLAUNCH(k_tall_zero, p.samples, s, 2, 0);
  const int max_outer = std::max(1, opt->max_outer);
  const int max_iter = std::max(1, opt->max_iterations);
  const int check = std::max(1, opt->check_every);
  const bool graphs = opt->graphs && timed_kernel_ == 0;
  int counts[2] = {0, 0};
  for (int pass = 0; pass <= max_outer; ++pass) {
    const int new_pass = pass < max_outer;
    TTIMED(FP_KERNEL_TRUE_RESIDUAL, PLAUNCH(k_tall_true, p.samples, s));
    FLAUNCH(k_fin_true<TallPart>, p.samples, s, new_pass);
    int rc = poll(s, p, counts);
    if (rc) return rc;
    if (counts[0] == 0 || !new_pass) break;
    if (graphs) {
      rc = ensure_graph(s, p);
      if (rc) return rc;
      const int period = std::max(2, check + (check & 1));
      for (int it = 0; it < max_iter; it += 2) {
        HIPCHK(hipGraphLaunch(p.graph, s));
        if ((it + 2) % period == 0 || it + 2 >= max_iter) {
          if ((rc = poll(s, p, counts))) return rc;
          if (counts[1] == 0) break;
        }
      }
    } else {
      for (int it = 0; it < max_iter; ++it) {
        enqueue_iteration(s, p);
        if ((it + 1) % check == 0 || it + 1 == max_iter) {
          if ((rc = poll(s, p, counts))) return rc;
          if (counts[1] == 0) break;
        }
      }
    }
    PLAUNCH(k_tall_accumulate, p.samples, s);
    PLAUNCH(k_tall_zero, p.samples, s, 2, 1);
  }
  for (int a = 0; a < 3; ++a) PLAUNCH(k_tall_subtract, p.faces[a], s, a);
  PLAUNCH(k_tall_rhs, p.samples, s, 1);
  FLAUNCH(k_tall_fin_divafter<TallPart>, p.samples, s);
  PLAUNCH(k_tall_pressure_out, p.samples, s);
  HIPCHK(hipEventRecord(ev1, s));
  hipLaunchKernelGGL(k_gather_states<TallPart>, dim3((unsigned)((grids.size() + 255) / 256)), dim3(256), 0, s,
                     p.one.dev, p.one.n, p.gathered);
  HIPCHK(hipMemcpyAsync(p.gathered_host, p.gathered, grids.size() * sizeof(State), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  float ms = 0.0f;
  HIPCHK(hipEventElapsedTime(&ms, ev0, ev1));
  (void)hipEventDestroy(ev0);
  (void)hipEventDestroy(ev1);
  double timed_ms = 0.0;
  for (int i = 0; i < timer_used_; ++i) {
    float t = 0.0f;
    HIPCHK(hipEventElapsedTime(&t, timer_events_[2 * i], timer_events_[2 * i + 1]));
    timed_ms += t;
  }
  const int timed_launches = timer_used_;
  timed_kernel_ = 0;
  timer_used_ = 0;
  if (reps) {
    for (size_t i = 0; i < grids.size(); ++i) {
      const State& S = p.gathered_host[i];
      const TallRec& R = *grids[i]->host_rec();
      fp_report* rep = &reps[i];
      std::memset(rep, 0, sizeof(*rep));
      int flags = S.flags;
      if (!S.outer_done || (S.flags & 0x1)) flags |= 0x8;
      rep->flags = flags;
      rep->iterations = S.total_iter;
      rep->outer_iterations = S.outer;
      rep->levels = R.L;
      rep->rhs_norm = std::sqrt(std::max(0.0, S.bb));
      rep->relative_residual = S.bb > 0 ? std::sqrt(std::max(0.0, S.rr64) / S.bb) : 0.0;
      float fa, fb;
      std::memcpy(&fa, &S.divmax_before, 4);
      std::memcpy(&fb, &S.divmax_after, 4);
      rep->max_div_before = fa;
      rep->max_div_after = fb;
      rep->device_ms = ms;
      rep->samples = (int64_t)R.n;
      rep->liquid_samples = (int64_t)S.liquid;
      rep->timed_kernel_ms = timed_ms;
      rep->timed_kernel_launches = timed_launches;
    }
  }
  return FP_OK;
}

// ---- continuous batching ---------------------------------------------------
static double tall_now_seconds() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

int TallProjector::serve_round(hipStream_t s, std::vector<Job*>& jobs) {
  struct Group {
    fp_options opt;
    std::vector<Job*> jobs;
  };
  auto key_of = [](const fp_options& o) { return ((long long)o.pre_sweeps << 16) | (long long)o.post_sweeps; };
  std::map<long long, Group> setups, accum, iter, resid[2];
  std::vector<Job*> finish, live;
  for (Job* j : jobs) {
    if (j->stage == Job::DONE) continue;
    live.push_back(j);
    const long long k = key_of(j->opt);
    Group* g = nullptr;
    if (j->stage == Job::SETUP) g = &setups[k];
    else if (j->stage == Job::ACCUMULATE) g = &accum[k];
    else if (j->stage == Job::ITERATE) g = &iter[k];
    else finish.push_back(j);
    if (g) {
      g->opt = j->opt;
      g->jobs.push_back(j);
    }
    if (j->stage == Job::SETUP || j->stage == Job::ACCUMULATE) {
      Group& r = resid[j->passes < std::max(1, j->opt.max_outer) ? 1 : 0][k];
      r.opt = j->opt;
      r.jobs.push_back(j);
    }
  }
  if (live.empty()) return FP_OK;
  std::vector<std::unique_ptr<Plan>> plans;
  auto make = [&](const std::vector<Job*>& js, const fp_options& o) {
    std::vector<TallGrid*> gs;
    for (Job* j : js) gs.push_back(j->tall);
    plans.emplace_back(new Plan());
    fill_tables(plans.back().get(), gs, o.pre_sweeps, o.post_sweeps);
    return plans.back().get();
  };
  std::vector<std::pair<Plan*, const Group*>> setup_p, accum_p, iter_p, resid_p[2];
  for (auto& kv : setups) setup_p.push_back({make(kv.second.jobs, kv.second.opt), &kv.second});
  for (auto& kv : accum) accum_p.push_back({make(kv.second.jobs, kv.second.opt), &kv.second});
  for (int np = 0; np < 2; ++np)
    for (auto& kv : resid[np]) resid_p[np].push_back({make(kv.second.jobs, kv.second.opt), &kv.second});
  for (auto& kv : iter) iter_p.push_back({make(kv.second.jobs, kv.second.opt), &kv.second});
  fp_options plain;
  std::memset(&plain, 0, sizeof(plain));
  Plan* fin_p = finish.empty() ? nullptr : make(finish, plain);
  Plan* all_p = make(live, plain);
  size_t total = 0;
  for (auto& p : plans) total += p->host.size();
  if (total > round_arena_cap_) {
    if (round_arena_) (void)hipFree(round_arena_);
    if (round_staging_) (void)hipHostFree(round_staging_);
    round_arena_ = nullptr;
    round_staging_ = nullptr;
    const size_t cap = std::max<size_t>(total * 2, 1024);
    HIPCHK(hipMalloc((void**)&round_arena_, cap * sizeof(TallPart)));
    HIPCHK(hipHostMalloc((void**)&round_staging_, cap * sizeof(TallPart), hipHostMallocDefault));
    round_arena_cap_ = cap;
  }
  if (live.size() > round_states_cap_) {
    if (round_states_) (void)hipFree(round_states_);
    if (round_states_host_) (void)hipHostFree(round_states_host_);
    round_states_ = nullptr;
    round_states_host_ = nullptr;
    const size_t cap = std::max<size_t>(live.size() * 2, 64);
    HIPCHK(hipMalloc((void**)&round_states_, cap * sizeof(State)));
    HIPCHK(hipHostMalloc((void**)&round_states_host_, cap * sizeof(State), hipHostMallocDefault));
    round_states_cap_ = cap;
  }
  size_t off = 0;
  for (auto& p : plans) {
    std::memcpy(round_staging_ + off, p->host.data(), p->host.size() * sizeof(TallPart));
    for (const auto& pd : p->pending) pd.t->dev = round_arena_ + off + pd.off;
    off += p->host.size();
  }
  HIPCHK(hipMemcpyAsync(round_arena_, round_staging_, std::max<size_t>(total, 1) * sizeof(TallPart),
                        hipMemcpyHostToDevice, s));
  // ---- set-up: records, face moments, block inverses, lattice, rhs ----
  for (auto& sp : setup_p)
    for (Job* j : sp.second->jobs) {
      j->t_start = tall_now_seconds();
      int rc = j->tall->bind(s, &j->in, &j->out, &j->opt);
      if (rc) return rc;
      if ((rc = j->tall->upload(s))) return rc;
    }
  for (auto& sp : setup_p) {
    Plan& p = *sp.first;
    FLAUNCH(k_reset_state<TallPart>, p.one, s);
    PLAUNCH(k_tall_zero, p.samples, s, 0, 0);
    setup(s, p);
    PLAUNCH(k_tall_rhs, p.samples, s, 0);
    FLAUNCH(k_fin_rhs<TallPart>, p.samples, s);
    PLAUNCH(k_tall_zero, p.samples, s, 1, 0);
    PLAUNCH(k_tall_zero, p.samples, s, 2, 0);
  }
  for (auto& ap : accum_p) {
    PLAUNCH(k_tall_accumulate, ap.first->samples, s);
    PLAUNCH(k_tall_zero, ap.first->samples, s, 2, 1);
  }
  for (int np = 0; np < 2; ++np)
    for (auto& rp : resid_p[np]) {
      PLAUNCH(k_tall_true, rp.first->samples, s);
      FLAUNCH(k_fin_true<TallPart>, rp.first->samples, s, np);
    }
  for (auto& ip : iter_p) enqueue_iteration(s, *ip.first);
  if (fin_p) {
    for (int a = 0; a < 3; ++a) PLAUNCH(k_tall_subtract, fin_p->faces[a], s, a);
    PLAUNCH(k_tall_rhs, fin_p->samples, s, 1);
    FLAUNCH(k_tall_fin_divafter<TallPart>, fin_p->samples, s);
    PLAUNCH(k_tall_pressure_out, fin_p->samples, s);
    for (Job* j : finish) {
      int rc = j->tall->finish_outputs(s, &j->in, &j->out);
      if (rc) return rc;
    }
  }
  hipLaunchKernelGGL(k_gather_states<TallPart>, dim3((unsigned)((live.size() + 255) / 256)), dim3(256), 0, s,
                     all_p->one.dev, all_p->one.n, round_states_);
  HIPCHK(hipMemcpyAsync(round_states_host_, round_states_, live.size() * sizeof(State), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  const double t = tall_now_seconds();
  for (size_t i = 0; i < live.size(); ++i) {
    Job* j = live[i];
    const State& S = round_states_host_[i];
    switch (j->stage) {
      case Job::SETUP:
      case Job::ACCUMULATE: {
        const bool more = j->passes < std::max(1, j->opt.max_outer);
        j->passes += 1;
        j->stage = (S.outer_done || !more) ? Job::FINISH : Job::ITERATE;
        break;
      }
      case Job::ITERATE:
        if (S.done) j->stage = Job::ACCUMULATE;
        break;
      case Job::FINISH: {
        const TallRec& R = *j->tall->host_rec();
        fp_report* rep = &j->report;
        std::memset(rep, 0, sizeof(*rep));
        int flags = S.flags;
        if (!S.outer_done || (S.flags & 0x1)) flags |= 0x8;
        rep->flags = flags;
        rep->iterations = S.total_iter;
        rep->outer_iterations = S.outer;
        rep->levels = R.L;
        rep->rhs_norm = std::sqrt(std::max(0.0, S.bb));
        rep->relative_residual = S.bb > 0 ? std::sqrt(std::max(0.0, S.rr64) / S.bb) : 0.0;
        float fa, fb;
        std::memcpy(&fa, &S.divmax_before, 4);
        std::memcpy(&fb, &S.divmax_after, 4);
        rep->max_div_before = fa;
        rep->max_div_after = fb;
        rep->device_ms = (t - j->t_start) * 1e3;
        rep->samples = (int64_t)R.n;
        rep->liquid_samples = (int64_t)S.liquid;
        j->t_done = t;
        j->stage = Job::DONE;
        break;
      }
      default:
        break;
    }
  }
  return FP_OK;
}

// ---- probes ----------------------------------------------------------------
int TallProjector::probe_apply(hipStream_t s, TallGrid* g, const float* phi, const double* p, double* out) {
  TallRec& R = *g->host_rec();
  const size_t n = (size_t)R.n;
  double *dp = nullptr, *dout = nullptr;
  HIPCHK(hipMemcpy(R.phi, phi, n * 4, hipMemcpyHostToDevice));
  HIPCHK(hipMalloc((void**)&dp, n * 8));
  HIPCHK(hipMalloc((void**)&dout, n * 8));
  hipError_t e = hipMemcpy(dp, p, n * 8, hipMemcpyHostToDevice);
  fp_options o;
  std::memset(&o, 0, sizeof(o));
  o.pre_sweeps = o.post_sweeps = 2;
  Plan* pp = e == hipSuccess ? plan_for({g}, &o) : nullptr;
  if (pp) {
    g->upload(s);
    PLAUNCH(k_tall_columns, pp->columns, s);
    PLAUNCH(k_tall_apply_d, pp->samples, s, dp, dout);
    e = hipStreamSynchronize(s);
    if (e == hipSuccess) e = hipMemcpy(out, dout, n * 8, hipMemcpyDeviceToHost);
  }
  (void)hipFree(dp);
  (void)hipFree(dout);
  if (!pp) return FP_ERR_OUT_OF_MEMORY;
  if (e != hipSuccess) {
    set_error("tall apply probe: %s", hipGetErrorString(e));
    return FP_ERR_DEVICE;
  }
  return FP_OK;
}

int TallProjector::probe_rhs(hipStream_t s, TallGrid* g, const fp_inputs* in, double* b) {
  int rc = g->bind(s, in, nullptr, nullptr);
  if (rc) return rc;
  fp_options o;
  std::memset(&o, 0, sizeof(o));
  o.pre_sweeps = o.post_sweeps = 2;
  Plan* pp = plan_for({g}, &o);
  if (!pp) return FP_ERR_OUT_OF_MEMORY;
  g->upload(s);
  PLAUNCH(k_tall_rhs, pp->samples, s, 0);
  HIPCHK(hipStreamSynchronize(s));
  HIPCHK(hipMemcpy(b, g->host_rec()->b, (size_t)g->host_rec()->n * 8, hipMemcpyDeviceToHost));
  return FP_OK;
}

int TallProjector::probe_layout(hipStream_t s, TallGrid* g, const float* phi, int* rows, unsigned char* unknown) {
  TallRec& R = *g->host_rec();
  const size_t n = (size_t)R.n;
  HIPCHK(hipMemcpy(R.phi, phi, n * 4, hipMemcpyHostToDevice));
  int* drows = nullptr;
  unsigned char* dunk = nullptr;
  HIPCHK(hipMalloc((void**)&drows, n * 4));
  hipError_t e = hipMalloc((void**)&dunk, n);
  fp_options o;
  std::memset(&o, 0, sizeof(o));
  o.pre_sweeps = o.post_sweeps = 2;
  Plan* pp = e == hipSuccess ? plan_for({g}, &o) : nullptr;
  if (pp) {
    g->upload(s);
    PLAUNCH(k_tall_layout, pp->samples, s, drows, dunk);
    e = hipStreamSynchronize(s);
    if (e == hipSuccess) e = hipMemcpy(rows, drows, n * 4, hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(unknown, dunk, n, hipMemcpyDeviceToHost);
  }
  (void)hipFree(drows);
  if (dunk) (void)hipFree(dunk);
  if (!pp) return FP_ERR_OUT_OF_MEMORY;
  if (e != hipSuccess) {
    set_error("tall layout probe: %s", hipGetErrorString(e));
    return FP_ERR_DEVICE;
  }
  return FP_OK;
}

}  // namespace fp

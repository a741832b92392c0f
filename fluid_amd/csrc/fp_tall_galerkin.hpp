// fp_tall_galerkin.hpp -- coarse levels of the tall-cell multigrid.
//
// Lattice level l >= 1 has nodes (X, Y, Z): X, Z aggregate 2 columns of level
// l-1, Y < 2 keeps the tall bottom / top rows, Y >= 2 aggregates `cy` band
// rows (piecewise-constant prolongation P, restriction P^T).  Its operator is
// the Galerkin product s * P^T A_{l-1} P (s = 0.6: the halved-per-axis
// scaling that turns the aggregation sums into a rediscretisation-like
// operator, tuned on rough and smooth layouts), stored per node in banded
// form:
//   +x couplings to (X+1, Y+dy, Z), dy in [-K, K]   slots [0, 2K]
//   +z couplings to (X, Y+dy, Z+1), dy in [-K, K]   slots [2K+1, 4K+1]
//   same-column couplings to (X, Y+dy, Z), dy in [1, K]  slots [4K+2, 5K+1]
//   diagonal                                          slot 5K+2
// (slot-major: w[slot * n + node]).  The -x / -z / below couplings of a node
// are the transposed entries its neighbours store.  K bounds the packed-row
// offset of any coupling: neighbouring columns differ by at most D rows, so a
// level-1 coupling spans at most ceil((D+1)/2) + 1 aggregated rows; K = (D+5)/2
// covers it, and coarser levels only shrink offsets.  A coupling beyond K
// (none by this bound) would be folded into both diagonals as |a|, which keeps
// the operator SPD.
//
// Smoother: kTallCoarseSweeps scaled l1-Jacobi sweeps before and after the
// coarse correction, e += theta (r - A e) / sum_j |a_ij| with theta = 1.8
// (on the 512^2 wave grid: 29 PCG iterations at 1.4, 28 at 1.6, 27 at 1.7,
// 26 at 1.8, 27 at 1.9).
// The coarse operators are not diagonally dominant (the affine profile inside
// tall cells gives positive couplings; lambda_max(D^-1 A) reaches 3), so plain
// damped Jacobi has no safe fixed weight.  With the l1 row sums,
// lambda_max(D_l1^-1 A) <= 1 by Gershgorin, so any theta < 2 keeps
// 2 D_l1 / theta - A positive definite: each sweep converges and the V-cycle
// stays an SPD map for any layout.
#pragma once

#include "fp_tall.hpp"

namespace fp {

constexpr float kTallGalerkinScale = 0.6f;
constexpr float kTallL1Theta = 1.8f;
constexpr int kTallCoarseSweeps = 3;
constexpr int kTallBottomThreads = 1024;

__device__ __forceinline__ int slot_x(int K, int dy) { return dy + K; }
__device__ __forceinline__ int slot_z(int K, int dy) { return 2 * K + 1 + dy + K; }
__device__ __forceinline__ int slot_y(int K, int dy) { return 4 * K + 2 + dy - 1; }
__device__ __forceinline__ int slot_d(int K) { return 5 * K + 2; }

// Coarse row of a finer row.
__device__ __forceinline__ int coarse_row(int y, int cy) { return y < 2 ? y : 2 + (y - 2) / cy; }

// Accumulates entries a_ij of the finer operator, j in coarse node C(j), into
// the couplings of coarse node C = (X, Y, Z).  The slots live in LDS (column
// threadIdx.x of a slot-major block array): dynamically indexed private
// arrays would go to scratch memory.
struct CoarseAcc {
  float* a;  // slot s at a[s * FP_BLOCK]
  int K, X, Y, Z;
  __device__ float& at(int s) { return a[s * FP_BLOCK]; }
  __device__ void init(float* col, int K_, int X_, int Y_, int Z_) {
    a = col;
    K = K_; X = X_; Y = Y_; Z = Z_;
    for (int s = 0; s < 5 * K + 3; ++s) at(s) = 0.0f;
  }
  __device__ void add(int Xj, int Yj, int Zj, float v) {
    const int dX = Xj - X, dY = Yj - Y, dZ = Zj - Z;
    if (dX == 0 && dY == 0 && dZ == 0) {
      at(slot_d(K)) += v;
      return;
    }
    if (dY > K || dY < -K) {
      at(slot_d(K)) += fabsf(v);
      return;
    }
    if (dX == 1 && dZ == 0) at(slot_x(K, dY)) += v;
    else if (dZ == 1 && dX == 0) at(slot_z(K, dY)) += v;
    else if (dX == 0 && dZ == 0 && dY > 0) at(slot_y(K, dY)) += v;
    // the -x, -z and below couplings are stored by the neighbour
  }
};

// ---- level 1 from the sample operator -------------------------------------
// Row i of A on a band unknown: emits (j, a_ij) for unknown samples j (incl. i).
template <class Emit>
__device__ void band_row_entries(const TallGeo& g, const int* __restrict__ bases, const float* __restrict__ phi,
                                 int x, int py, int z, int b, Emit& emit) {
  const int y = b + py - 2;
  const size_t i = g.sidx(x, py, z);
  const float phic = phi[i];
  float dg = 0.0f;
  {
    const size_t j = g.sidx(x, py - 1, z);
    const float w = fw<float>(phic, phi[j]);
    dg += w;
    if (phi[j] < 0.0f) emit(x, py - 1, z, -w);
  }
  if (y + 1 < g.ny) {
    if (py + 1 < g.P) {
      const size_t j = g.sidx(x, py + 1, z);
      const float w = fw<float>(phic, phi[j]);
      dg += w;
      if (phi[j] < 0.0f) emit(x, py + 1, z, -w);
    } else {
      dg += fwr<float>(phic, (float)(5.0 * g.h));
    }
  }
  for (int dir = 0; dir < 4; ++dir) {
    int xn, zn;
    if (!neighbour(g, x, z, dir, xn, zn)) continue;
    const int bn = bases[g.col(xn, zn)];
    const RowRep<float> r = row_rep<float>(g, phi, bn, xn, zn, y);
    const float w = fwr<float>(phic, r.phi);
    dg += w;
    if (r.n >= 1 && phi[r.i0] < 0.0f) emit(xn, (int)((r.i0 / (size_t)g.nz) % (size_t)g.P), zn, -w * r.c0);
    if (r.n == 2 && phi[r.i1] < 0.0f) emit(xn, 1, zn, -w * r.c1);
  }
  emit(x, py, z, dg);
}

// Row of the tall bottom (which = 0) or top (which = 1) unknown of column (x, z).
template <class Emit>
__device__ void tall_row_entries(const TallGeo& g, const int* __restrict__ bases, const float* __restrict__ phi,
                                 const TallCol* __restrict__ cols, int x, int z, int b, int which, Emit& emit) {
  const float kap = 1.0f / (float)(b - 1);
  const bool ub = phi[g.sidx(x, 0, z)] < 0.0f, ut = phi[g.sidx(x, 1, z)] < 0.0f;
  const float phib = phi[g.sidx(x, 0, z)], phit = phi[g.sidx(x, 1, z)];
  float self = 0.0f, other = 0.0f;  // a_ii and a_i,(other tall sample)
  {
    const float c = (float)cols[g.col(x, z)].wint * kap * kap;
    self += c;
    other -= c;
  }
  if (which == 1 && b < g.ny) {
    const size_t j = g.sidx(x, 2, z);
    const float w = fw<float>(phit, phi[j]);
    self += w;
    if (phi[j] < 0.0f) emit(x, 2, z, -w);
  }
  for (int dir = 0; dir < 4; ++dir) {
    int xn, zn;
    if (!neighbour(g, x, z, dir, xn, zn)) continue;
    const int bn = bases[g.col(xn, zn)];
    const double* M = face_moments(g, cols, x, z, dir);
    const float M0 = (float)M[0], M1 = (float)M[1], M2 = (float)M[2];
    const int m = b < bn ? b : bn;
    const float ym = 0.5f * (float)(m - 1);
    const float kn = 1.0f / (float)(bn - 1);
    // this side's coefficient function of the chosen sample: a0 + a1 t
    const float c0 = 1.0f - ym * kap, c1 = -kap, d0 = ym * kap, d1 = kap;
    const float a0 = which == 0 ? c0 : d0, a1 = which == 0 ? c1 : d1;
    const float o0 = which == 0 ? d0 : c0, o1 = which == 0 ? d1 : c1;
    self += a0 * a0 * M0 + 2.0f * a0 * a1 * M1 + a1 * a1 * M2;
    other += a0 * o0 * M0 + (a0 * o1 + a1 * o0) * M1 + a1 * o1 * M2;
    const float e0 = 1.0f - ym * kn, e1 = -kn, f0 = ym * kn, f1 = kn;
    if (phi[g.sidx(xn, 0, zn)] < 0.0f) emit(xn, 0, zn, -(a0 * e0 * M0 + (a0 * e1 + a1 * e0) * M1 + a1 * e1 * M2));
    if (phi[g.sidx(xn, 1, zn)] < 0.0f) emit(xn, 1, zn, -(a0 * f0 * M0 + (a0 * f1 + a1 * f0) * M1 + a1 * f1 * M2));
    for (int y = m; y < b; ++y) {
      const float k = (float)y * kap;
      const float ca = which == 0 ? 1.0f - k : k, co = which == 0 ? k : 1.0f - k;
      const float phia = phib + (phit - phib) * k;
      const RowRep<float> r = row_rep<float>(g, phi, bn, xn, zn, y);
      const float w = fwr<float>(phia, r.phi);
      self += w * ca * ca;
      other += w * ca * co;
      if (r.n >= 1 && phi[r.i0] < 0.0f) emit(xn, (int)((r.i0 / (size_t)g.nz) % (size_t)g.P), zn, -w * ca * r.c0);
      if (r.n == 2 && phi[r.i1] < 0.0f) emit(xn, 1, zn, -w * ca * r.c1);
    }
  }
  emit(x, which, z, self);
  if (which == 0 ? ut : ub) emit(x, 1 - which, z, other);
}

__global__ __launch_bounds__(FP_BLOCK) void k_tall_galerkin1(const TallPart* parts, int np) {
  __shared__ float s_acc[kTallSlotsMax * FP_BLOCK];
  FP_PART(parts, np);
  const TallGeo g = geo_of(G);
  const int* __restrict__ bases = G.bases;
  const float* __restrict__ phi = G.phi;
  const TallLevel& C = G.lv[1];
  const int K = G.K, cy = C.cy;
  const size_t n = (size_t)C.n;
  FP_STRIDE(ci, n) {
    const int Z = (int)(ci % (size_t)C.nz);
    const size_t tt = ci / (size_t)C.nz;
    const int Y = (int)(tt % (size_t)C.ny);
    const int X = (int)(tt / (size_t)C.ny);
    CoarseAcc acc;
    acc.init(s_acc + threadIdx.x, K, X, Y, Z);
    auto emit = [&](int xj, int pyj, int zj, float v) { acc.add(xj >> 1, coarse_row(pyj, cy), zj >> 1, v); };
    int y0, y1;
    if (Y < 2) {
      y0 = Y;
      y1 = Y + 1;
    } else {
      y0 = 2 + cy * (Y - 2);
      y1 = min(y0 + cy, g.P);
    }
    for (int x = 2 * X; x < min(2 * X + 2, g.nx); ++x)
      for (int z = 2 * Z; z < min(2 * Z + 2, g.nz); ++z) {
        const int b = bases[g.col(x, z)];
        for (int py = y0; py < y1; ++py) {
          if (!is_unknown(g, phi, b, x, py, z)) continue;
          if (py < 2) tall_row_entries(g, bases, phi, G.cols, x, z, b, py, emit);
          else band_row_entries(g, bases, phi, x, py, z, b, emit);
        }
      }
    const int NS = 5 * K + 3;
    for (int s = 0; s < NS; ++s) C.w[(size_t)s * n + ci] = kTallGalerkinScale * acc.at(s);
  }
}

// ---- level l+1 from level l (both banded) -----------------------------------
// Every coupling of a level-l node i, emitted as (neighbour X, Y, Z, a).
// kh / kv bound the row offsets visited for the x/z and the same-column
// couplings (K, K visit the whole band; a level's measured reach skips slots
// that are zero on every node).
template <class Emit>
__device__ void lattice_row_entries(const TallLevel& L, int K, int kh, int kv, int x, int y, int z, Emit& emit) {
  const size_t n = (size_t)L.n;
  const float* __restrict__ w = L.w;
  auto at = [&](int xx, int yy, int zz) { return ((size_t)xx * L.ny + yy) * L.nz + zz; };
  const size_t i = at(x, y, z);
  emit(x, y, z, w[(size_t)slot_d(K) * n + i]);
  for (int dy = -kh; dy <= kh; ++dy) {
    const int yy = y + dy;
    if (yy < 0 || yy >= L.ny) continue;
    if (x + 1 < L.nx) {
      const float v = w[(size_t)slot_x(K, dy) * n + i];
      if (v != 0.0f) emit(x + 1, yy, z, v);
    }
    if (x > 0) {  // stored by (x-1, yy) towards (x, y): offset -dy
      const float v = w[(size_t)slot_x(K, -dy) * n + at(x - 1, yy, z)];
      if (v != 0.0f) emit(x - 1, yy, z, v);
    }
    if (z + 1 < L.nz) {
      const float v = w[(size_t)slot_z(K, dy) * n + i];
      if (v != 0.0f) emit(x, yy, z + 1, v);
    }
    if (z > 0) {
      const float v = w[(size_t)slot_z(K, -dy) * n + at(x, yy, z - 1)];
      if (v != 0.0f) emit(x, yy, z - 1, v);
    }
  }
  for (int dy = 1; dy <= kv; ++dy) {
    if (y + dy < L.ny) {
      const float v = w[(size_t)slot_y(K, dy) * n + i];
      if (v != 0.0f) emit(x, y + dy, z, v);
    }
    if (y - dy >= 0) {
      const float v = w[(size_t)slot_y(K, dy) * n + at(x, y - dy, z)];
      if (v != 0.0f) emit(x, y - dy, z, v);
    }
  }
}

__global__ __launch_bounds__(FP_BLOCK) void k_tall_galerkin(const TallPart* parts, int np, int level) {
  __shared__ float s_acc[kTallSlotsMax * FP_BLOCK];
  FP_PART(parts, np);
  const TallLevel& F = G.lv[level];
  const TallLevel& C = G.lv[level + 1];
  const int K = G.K, cy = C.cy;
  const size_t n = (size_t)C.n;
  const size_t nf = (size_t)F.n;
  FP_STRIDE(ci, n) {
    const int Z = (int)(ci % (size_t)C.nz);
    const size_t tt = ci / (size_t)C.nz;
    const int Y = (int)(tt % (size_t)C.ny);
    const int X = (int)(tt / (size_t)C.ny);
    CoarseAcc acc;
    acc.init(s_acc + threadIdx.x, K, X, Y, Z);
    auto emit = [&](int xj, int yj, int zj, float v) { acc.add(xj >> 1, coarse_row(yj, cy), zj >> 1, v); };
    int y0, y1;
    if (Y < 2) {
      y0 = Y;
      y1 = Y + 1;
    } else {
      y0 = 2 + cy * (Y - 2);
      y1 = min(y0 + cy, F.ny);
    }
    for (int x = 2 * X; x < min(2 * X + 2, F.nx); ++x)
      for (int y = y0; y < y1; ++y)
        for (int z = 2 * Z; z < min(2 * Z + 2, F.nz); ++z) {
          const size_t fi = ((size_t)x * F.ny + y) * F.nz + z;
          if (!(F.w[(size_t)slot_d(K) * nf + fi] > 0.0f)) continue;
          lattice_row_entries(F, K, K, K, x, y, z, emit);
        }
    const int NS = 5 * K + 3;
    for (int s = 0; s < NS; ++s) C.w[(size_t)s * n + ci] = kTallGalerkinScale * acc.at(s);
  }
}

// theta / sum_j |a_ij| (0 on inactive nodes): the scaled l1-Jacobi weight.
__global__ void k_tall_l1diag(const TallPart* parts, int np, int level) {
  FP_PART(parts, np);
  const TallLevel& L = G.lv[level];
  const int K = G.K;
  const size_t n = (size_t)L.n;
  int kh = 0, kv = 0;
  FP_STRIDE(i, n) {
    const int z = (int)(i % (size_t)L.nz);
    const size_t tt = i / (size_t)L.nz;
    const int y = (int)(tt % (size_t)L.ny);
    const int x = (int)(tt / (size_t)L.ny);
    float s = 0.0f, d = 0.0f;
    auto emit = [&](int xj, int yj, int zj, float v) {
      if (xj == x && yj == y && zj == z) {
        d = v;
        return;
      }
      s += fabsf(v);
      const int dy = yj > y ? yj - y : y - yj;
      if (xj == x && zj == z) kv = max(kv, dy);
      else kh = max(kh, dy);
    };
    lattice_row_entries(L, K, K, K, x, y, z, emit);
    L.dinv[i] = d > 0.0f ? kTallL1Theta / (d + s) : 0.0f;
  }
  // the level's reach: the largest row offset of a nonzero coupling
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    kh = max(kh, __shfl_xor(kh, o, 64));
    kv = max(kv, __shfl_xor(kv, o, 64));
  }
  if ((threadIdx.x & 63) == 0) {
    atomicMax(G.reach + 2 * level, kh);
    atomicMax(G.reach + 2 * level + 1, kv);
  }
}

// Zero reach before the levels' l1 diagonals measure it (one workgroup per grid).
__global__ void k_tall_reach_reset(const TallPart* parts) {
  const TallRec& G = *parts[blockIdx.x].g;
  for (int i = threadIdx.x; i < 2 * kTallMaxLevels; i += blockDim.x) G.reach[i] = 0;
}

// (A e)_i on a lattice level within the level's reach (kh, kv).  Every weight
// and value load is independent of the others: a zero coupling contributes
// zero, and rows past the level's edge are clamped and masked, so the loads of
// a node are all in flight together.
__device__ __forceinline__ float lattice_apply(const TallLevel& L, int K, int kh, int kv, const float* __restrict__ e,
                                               int x, int y, int z) {
  const size_t n = (size_t)L.n;
  const float* __restrict__ w = L.w;
  const size_t sy = (size_t)L.nz, sx = (size_t)L.ny * L.nz;
  const size_t i = (size_t)x * sx + (size_t)y * sy + z;
  float acc = w[(size_t)slot_d(K) * n + i] * e[i];
  const bool xp = x + 1 < L.nx, xm = x > 0, zp = z + 1 < L.nz, zm = z > 0;
  for (int dy = -kh; dy <= kh; ++dy) {
    const int yy = y + dy;
    const bool row = yy >= 0 && yy < L.ny;
    const size_t j = i + (size_t)((long long)(row ? dy : 0) * (long long)sy);  // (x, yy, z), clamped
    const float a0 = xp ? w[(size_t)slot_x(K, dy) * n + i] : 0.0f;
    const float e0 = xp ? e[j + sx] : 0.0f;
    const float a1 = xm ? w[(size_t)slot_x(K, -dy) * n + j - sx] : 0.0f;
    const float e1 = xm ? e[j - sx] : 0.0f;
    const float a2 = zp ? w[(size_t)slot_z(K, dy) * n + i] : 0.0f;
    const float e2 = zp ? e[j + 1] : 0.0f;
    const float a3 = zm ? w[(size_t)slot_z(K, -dy) * n + j - 1] : 0.0f;
    const float e3 = zm ? e[j - 1] : 0.0f;
    const float s = (a0 * e0 + a1 * e1) + (a2 * e2 + a3 * e3);
    acc += row ? s : 0.0f;
  }
  for (int dy = 1; dy <= kv; ++dy) {
    const bool up = y + dy < L.ny, dn = y - dy >= 0;
    const float a0 = up ? w[(size_t)slot_y(K, dy) * n + i] : 0.0f;
    const float e0 = up ? e[i + (size_t)dy * sy] : 0.0f;
    const float a1 = dn ? w[(size_t)slot_y(K, dy) * n + i - (size_t)dy * sy] : 0.0f;
    const float e1 = dn ? e[i - (size_t)dy * sy] : 0.0f;
    acc += a0 * e0 + a1 * e1;
  }
  return acc;
}

// ---- per-node steps of the lattice V-cycle (grid kernels below, and the
// one-workgroup bottom of the cycle, k_lat_bottom) ----
__device__ __forceinline__ void node_xyz(const TallLevel& L, size_t i, int& x, int& y, int& z) {
  z = (int)(i % (size_t)L.nz);
  const size_t tt = i / (size_t)L.nz;
  y = (int)(tt % (size_t)L.ny);
  x = (int)(tt / (size_t)L.ny);
}

// Scaled l1-Jacobi: out = in + dinv (r - A in) (first: in = 0).  sel 0 reads e
// writes e2; 1 reads e2 writes e.
__device__ __forceinline__ void lat_jacobi_node(const TallRec& G, int level, int sel, int first, int kh, int kv,
                                                size_t i) {
  const TallLevel& L = G.lv[level];
  const float* __restrict__ ein = sel ? L.e2 : L.e;
  float* __restrict__ eout = first ? L.e : (sel ? L.e : L.e2);
  const float di = L.dinv[i];
  float out = 0.0f;
  if (di != 0.0f) {
    if (first) {
      out = L.r[i] * di;
    } else {
      int x, y, z;
      node_xyz(L, i, x, y, z);
      out = ein[i] + (L.r[i] - lattice_apply(L, G.K, kh, kv, ein, x, y, z)) * di;
    }
  }
  eout[i] = out;
}

// t = r - A e (sel: which buffer holds e).
__device__ __forceinline__ void lat_residual_node(const TallRec& G, int level, int sel, int kh, int kv, size_t i) {
  const TallLevel& L = G.lv[level];
  const float* __restrict__ e = sel ? L.e2 : L.e;
  float res = 0.0f;
  if (L.dinv[i] != 0.0f) {
    int x, y, z;
    node_xyz(L, i, x, y, z);
    res = L.r[i] - lattice_apply(L, G.K, kh, kv, e, x, y, z);
  }
  L.t[i] = res;
}

// r_{l+1} = P^T t_l at coarse node ci (level 0: the sample residual).
__device__ __forceinline__ void lat_restrict_node(const TallRec& G, int level, size_t ci) {
  const TallLevel& C = G.lv[level + 1];
  const int fnx = level == 0 ? G.nx : G.lv[level].nx;
  const int fny = level == 0 ? G.P : G.lv[level].ny;
  const int fnz = level == 0 ? G.nz : G.lv[level].nz;
  const float* __restrict__ t = level == 0 ? G.t : G.lv[level].t;
  const int cy = C.cy;
  int X, Y, Z;
  node_xyz(C, ci, X, Y, Z);
  float s = 0.0f;
  if (C.dinv[ci] != 0.0f) {
    int y0, y1;
    if (Y < 2) {
      y0 = Y;
      y1 = Y + 1;
    } else {
      y0 = 2 + cy * (Y - 2);
      y1 = min(y0 + cy, fny);
    }
    for (int x = 2 * X; x < min(2 * X + 2, fnx); ++x)
      for (int y = y0; y < y1; ++y)
        for (int z = 2 * Z; z < min(2 * Z + 2, fnz); ++z) s += t[((size_t)x * fny + y) * fnz + z];
  }
  C.r[ci] = s;
}

// Buffer holding the final correction of lattice level l: the coarsest
// level's solve writes e; the smoothed levels end their post-sweeps in e2.
__device__ __forceinline__ const float* final_e(const TallRec& G, int l) {
  return l == G.L - 1 ? G.lv[l].e : G.lv[l].e2;
}

// e_l += P e_{l+1} on active level-l node i (l >= 1; sel: e_l buffer).
__device__ __forceinline__ void lat_prolong_node(const TallRec& G, int level, int sel, size_t i) {
  const TallLevel& F = G.lv[level];
  const TallLevel& C = G.lv[level + 1];
  if (F.dinv[i] == 0.0f) return;
  float* __restrict__ ef = sel ? F.e2 : F.e;
  const float* __restrict__ ec = final_e(G, level + 1);
  int x, y, z;
  node_xyz(F, i, x, y, z);
  ef[i] += ec[((size_t)(x >> 1) * C.ny + coarse_row(y, C.cy)) * C.nz + (z >> 1)];
}

__global__ void k_lat_jacobi(const TallPart* parts, int np, int level, int sel, int first) {
  FP_PART(parts, np);
  if (G.st->done) return;
  const int kh = cload(G.reach + 2 * level), kv = cload(G.reach + 2 * level + 1);
  FP_STRIDE(i, (size_t)G.lv[level].n) lat_jacobi_node(G, level, sel, first, kh, kv, i);
}

__global__ void k_lat_residual(const TallPart* parts, int np, int level, int sel) {
  FP_PART(parts, np);
  if (G.st->done) return;
  const int kh = cload(G.reach + 2 * level), kv = cload(G.reach + 2 * level + 1);
  FP_STRIDE(i, (size_t)G.lv[level].n) lat_residual_node(G, level, sel, kh, kv, i);
}

__global__ void k_lat_restrict(const TallPart* parts, int np, int level) {
  FP_PART(parts, np);
  if (G.st->done) return;
  FP_STRIDE(ci, (size_t)G.lv[level + 1].n) lat_restrict_node(G, level, ci);
}

__global__ void k_lat_prolong(const TallPart* parts, int np, int level, int sel) {
  FP_PART(parts, np);
  if (G.st->done) return;
  FP_STRIDE(i, (size_t)G.lv[level].n) lat_prolong_node(G, level, sel, i);
}

// Dense FP64 inverse of the coarsest lattice (one workgroup per grid).  As on
// uniform grids, the vanishing last pivot of a component without an air face
// (a grid that is liquid everywhere, its band up to the lid) gets its row's
// original diagonal added: the inverse of A + a_pp e_p e_p^T, SPD.
__global__ void k_tall_coarse_setup(const TallPart* parts) {
  const TallRec& G = *parts[blockIdx.x].g;
  const TallLevel& L = G.lv[G.L - 1];
  const int K = G.K;
  int* __restrict__ cmap = G.cmap;
  double* __restrict__ a = G.minv;
  __shared__ int s_n, s_fail;
  __shared__ double s_diag[kTallMaxCoarse];
  const size_t n = (size_t)L.n;
  if (threadIdx.x == 0) {
    int c = 0;
    for (size_t i = 0; i < n && c < G.max_coarse; ++i)
      if (L.w[(size_t)slot_d(K) * n + i] > 0.0f) cmap[c++] = (int)i;
    s_n = c;
    s_fail = 0;
    *G.ncoarse = c;
  }
  __syncthreads();
  const int m = s_n;
  for (int e = threadIdx.x; e < m * m; e += blockDim.x) a[e] = 0.0;
  __syncthreads();
  // scatter each active row's couplings (the row sees all its entries)
  for (int r = threadIdx.x; r < m; r += blockDim.x) {
    const size_t ir = (size_t)cmap[r];
    const int z = (int)(ir % (size_t)L.nz);
    const size_t tt = ir / (size_t)L.nz;
    const int y = (int)(tt % (size_t)L.ny);
    const int x = (int)(tt / (size_t)L.ny);
    auto emit = [&](int xj, int yj, int zj, float v) {
      const size_t j = ((size_t)xj * L.ny + yj) * L.nz + zj;
      for (int c = 0; c < m; ++c)
        if ((size_t)cmap[c] == j) {
          a[(size_t)r * m + c] += (double)v;
          break;
        }
    };
    lattice_row_entries(L, K, K, K, x, y, z, emit);
  }
  __syncthreads();
  for (int r = threadIdx.x; r < m; r += blockDim.x) s_diag[r] = a[(size_t)r * m + r];
  __syncthreads();
  for (int p = 0; p < m; ++p) {
    double piv = a[p * m + p];
    if (piv <= 1e-9 * s_diag[p] && piv > -1e-9 * s_diag[p]) piv += s_diag[p];  // singular component
    if (!(piv > 0.0) && threadIdx.x == 0) s_fail = 1;
    __syncthreads();
    if (s_fail) break;
    const double inv = 1.0 / piv;
    for (int e = threadIdx.x; e < m * m; e += blockDim.x) {
      const int r = e / m, c = e % m;
      if (r == p || c == p) continue;
      a[e] -= a[r * m + p] * a[p * m + c] * inv;
    }
    __syncthreads();
    for (int e = threadIdx.x; e < m; e += blockDim.x)
      if (e != p) {
        a[e * m + p] = -a[e * m + p] * inv;
        a[p * m + e] = a[p * m + e] * inv;
      }
    __syncthreads();
    if (threadIdx.x == 0) a[p * m + p] = inv;
    __syncthreads();
  }
  if (s_fail) {
    for (int e = threadIdx.x; e < m * m; e += blockDim.x) a[e] = 0.0;
    __syncthreads();
    for (int r = threadIdx.x; r < m; r += blockDim.x)
      a[r * m + r] = 1.0 / (double)L.w[(size_t)slot_d(K) * n + cmap[r]];
    if (threadIdx.x == 0) G.st->flags |= 0x40;  // FP_FLAG_COARSE_SINGULAR
  }
}

// e = Minv r on the coarsest level's active nodes (the calling workgroup),
// the FP64 product on the matrix cores.
__device__ __forceinline__ void coarse_solve_block(const TallRec& G) {
  __shared__ double rs[kTallMaxCoarse];
  const TallLevel& L = G.lv[G.L - 1];
  const int m = *G.ncoarse;
  for (size_t i = threadIdx.x; i < (size_t)L.n; i += blockDim.x) L.e[i] = 0.0f;
  for (int q = threadIdx.x; q < m; q += blockDim.x) rs[q] = (double)L.r[G.cmap[q]];
  __syncthreads();
  float* __restrict__ e = L.e;
  const int* __restrict__ cmap = G.cmap;
  gemv_f64_mfma(G.minv, m, rs, [&](int row, double v) { e[cmap[row]] = (float)v; });
}

// The bottom of the V-cycle in one workgroup per grid: levels lb .. L-2 (at
// most kTallBottomNodes nodes each) are smoothed, restricted, solved on the
// coarsest level and prolonged with the same steps and buffer conventions as
// the grid kernels, separated by workgroup barriers instead of launches.
__global__ __launch_bounds__(kTallBottomThreads) void k_lat_bottom(const TallPart* parts) {
  const TallRec& G = *parts[blockIdx.x].g;
  if (G.st->done) return;
  const int lb = G.lb, lc = G.L - 1;
  const int cpre = (kTallCoarseSweeps - 1) & 1;
  for (int l = lb; l < lc; ++l) {
    const int kh = G.reach[2 * l], kv = G.reach[2 * l + 1];
    const size_t n = (size_t)G.lv[l].n;
    for (size_t i = threadIdx.x; i < n; i += blockDim.x) lat_jacobi_node(G, l, 0, 1, kh, kv, i);
    __syncthreads();
    for (int sw = 1, c = 0; sw < kTallCoarseSweeps; ++sw, c ^= 1) {
      for (size_t i = threadIdx.x; i < n; i += blockDim.x) lat_jacobi_node(G, l, c, 0, kh, kv, i);
      __syncthreads();
    }
    for (size_t i = threadIdx.x; i < n; i += blockDim.x) lat_residual_node(G, l, cpre, kh, kv, i);
    __syncthreads();
    for (size_t ci = threadIdx.x; ci < (size_t)G.lv[l + 1].n; ci += blockDim.x) lat_restrict_node(G, l, ci);
    __syncthreads();
  }
  coarse_solve_block(G);
  __syncthreads();
  for (int l = lc - 1; l >= lb; --l) {
    const int kh = G.reach[2 * l], kv = G.reach[2 * l + 1];
    const size_t n = (size_t)G.lv[l].n;
    for (size_t i = threadIdx.x; i < n; i += blockDim.x) lat_prolong_node(G, l, cpre, i);
    __syncthreads();
    for (int sw = 0, c = cpre; sw < kTallCoarseSweeps; ++sw, c ^= 1) {
      for (size_t i = threadIdx.x; i < n; i += blockDim.x) lat_jacobi_node(G, l, c, 0, kh, kv, i);
      __syncthreads();
    }
  }
}

}  // namespace fp

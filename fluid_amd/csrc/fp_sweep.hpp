// fp_sweep.hpp -- one full red-black Gauss-Seidel sweep in a single pass.
//
// The smoother of mgSmoothColour (webgpu-uniform-pressure-multigrid.wgsl.ts:325-344,
// without the LCP clamp) applied as both colour passes of one sweep, written for
// CDNA4: a 256-lane workgroup owns a (TY x TZ) column tile (z fastest) and
// marches along x through a chunk of planes.  Each plane is staged once in LDS
// with a two-cell halo; the first colour is updated on the tile plus a
// one-cell ring from the OLD field (read-only input buffer), the second colour
// on the tile from those fresh values, so the result is exactly one red-black
// sweep no matter how workgroups are scheduled (old and new fields are
// different buffers).  One read of the coefficients, r and e and one write of
// e per cell replace the two passes of separate colour kernels.
//
// Variant (template flag):
//   ZERO     the sweep starts from e = 0 (first pre-smoothing sweep; e_in unread)
#pragma once

#include "fp_common.hpp"
#include "fp_uniform.hpp"

namespace fp {

// Packed per-cell coefficients of one level: the three positive-face
// weights (h^2-scaled, 0 on walls and between air cells) and the inverse
// diagonal (0 for air and for liquid cells without any open face).  The
// negative-face weight of a cell is the positive-face weight of its
// neighbour, so walls and air need no tests in the sweep: they carry zeros.
template <typename T>
struct alignas(4 * sizeof(T)) W4 {
  T wx, wy, wz, id;
};

template <typename T>
__global__ void k_build_weights(int nx, int ny, int nz, float gx, float gy, float gz, const float* __restrict__ phi,
                                W4<T>* __restrict__ w) {
  const size_t n = (size_t)nx * ny * nz;
  const size_t sx = (size_t)ny * nz, sy = (size_t)nz;
  for (size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x; idx < n; idx += (size_t)gridDim.x * blockDim.x) {
    const int k = (int)(idx % (size_t)nz);
    const size_t t = idx / (size_t)nz;
    const int j = (int)(t % (size_t)ny);
    const int i = (int)(t / (size_t)ny);
    const float pc = phi[idx];
    const float wxp = (i + 1 < nx) ? face_weight_fast(pc, phi[idx + sx], gx) : 0.0f;
    const float wyp = (j + 1 < ny) ? face_weight_fast(pc, phi[idx + sy], gy) : 0.0f;
    const float wzp = (k + 1 < nz) ? face_weight_fast(pc, phi[idx + 1], gz) : 0.0f;
    const float wxm = (i > 0) ? face_weight_fast(pc, phi[idx - sx], gx) : 0.0f;
    const float wym = (j > 0) ? face_weight_fast(pc, phi[idx - sy], gy) : 0.0f;
    const float wzm = (k > 0) ? face_weight_fast(pc, phi[idx - 1], gz) : 0.0f;
    const float diag = (wxp + wxm) + (wyp + wym) + (wzp + wzm);
    W4<T> out;
    Store<T>::st(&out.wx, 0, wxp);
    Store<T>::st(&out.wy, 0, wyp);
    Store<T>::st(&out.wz, 0, wzp);
    Store<T>::st(&out.id, 0, (pc < 0.0f && diag > 0.0f) ? 1.0f / diag : 0.0f);
    w[idx] = out;
  }
}

// Tile width along z and height along y of the levels' liquid hulls (the
// 32 x 8 march tiles).
constexpr int kHullTZ = 32, kHullTY = 256 / kHullTZ;

// Workgroups are dispatched round-robin over the 8 XCDs of the device, each
// with its own L2: group g runs on XCD g % 8.  Logical tiles run z fastest,
// then y, then the x chunk.  The x chunks stay in lock step across the device
// and every XCD takes slabs of two tile rows (all z tiles of two y rows) of
// each chunk, slab b going to XCD b % 8: neighbouring tiles in a slab, which
// read each other's halo rows, share an L2, and liquid that fills only part
// of the grid's height still reaches all eight XCDs (the hull-clamped passes
// do work only where liquid is).  A slab is 16 grid rows high: two rows of
// 8-row tiles (slab_rows 2) or one row of 16-row tiles (slab_rows 1).  Grids
// whose tile rows do not split into slabs for 8 XCDs take one contiguous slab
// per XCD (tiles % 8 == 0), else the identity.
__device__ __forceinline__ void tile_of_group(int g, int tiles_z, int tiles_y, int ngroups, int& tzi, int& tyi,
                                              int& cxi, int slab_rows = 2) {
  (void)ngroups;
  const int tiles = tiles_z * tiles_y;
  int l = g;
  if (tiles_y % (8 * slab_rows) == 0) {
    const int S = tiles >> 3;          // tiles per XCD per chunk
    const int Bt = slab_rows * tiles_z;  // tiles per slab
    const int i = g >> 3;
    const int chunk = i / S, r = i - chunk * S;
    const int bx = r / Bt;
    l = chunk * tiles + (bx * 8 + (g & 7)) * Bt + (r - bx * Bt);
  } else if ((tiles & 7) == 0) {
    const int S = tiles >> 3;
    const int i = g >> 3;
    const int chunk = i / S;
    l = chunk * tiles + (g & 7) * S + (i - chunk * S);
  }
  tzi = l % tiles_z;
  const int t = l / tiles_z;
  tyi = t % tiles_y;
  cxi = t / tiles_y;
}

template <typename T>
__device__ __forceinline__ float wf(const T& v) { return Store<T>::ld(&v, 0); }

// Bit-level views for branch-free masking of staged loads.
template <typename T> struct W4Bits;
template <> struct W4Bits<__half> { typedef fp_u32x2 type; };
template <> struct W4Bits<__hip_bfloat16> { typedef fp_u32x2 type; };
template <> struct W4Bits<float> { typedef fp_u32x4 type; };

__device__ __forceinline__ fp_u32x2 bits_unless(fp_u32x2 b, bool keep) { return keep ? b : fp_u32x2(0u); }
__device__ __forceinline__ fp_u32x4 bits_unless(fp_u32x4 b, bool keep) { return keep ? b : fp_u32x4(0u); }

template <typename T>
__device__ __forceinline__ T value_unless(T v, bool keep) {
  if constexpr (sizeof(T) == 2) {
    unsigned short b = keep ? *reinterpret_cast<unsigned short*>(&v) : (unsigned short)0;
    return *reinterpret_cast<T*>(&b);
  } else {
    return keep ? v : T(0);
  }
}

template <typename T>
__device__ __forceinline__ T to_store(float v) {
  T t;
  Store<T>::st(&t, 0, v);
  return t;
}

// p[i] = a, p[i+1] = b with one store (i even; p aligned to two elements).
template <typename T>
__device__ __forceinline__ void gst_pair(T* p, size_t i, float a, float b) {
  struct alignas(2 * sizeof(T)) Two { T lo, hi; };
  stg(reinterpret_cast<Two*>(p), i >> 1, Two{to_store<T>(a), to_store<T>(b)});
}

// One red-black sweep.  TZ x TY = 256 lanes; TR: storage of the right-hand
// side, T: of e, of the packed coefficients and of the LDS staging.
//
// LDS per plane slot: coefficients on rows/columns -2..T of the tile, the old
// correction on -2..T+1, the scaled right-hand side and the first-colour
// values on -1..T.  Four slots of coefficients, r and first-colour values,
// three of the old correction.  Per plane step L:
//   commit plane L | A | prefetch L+D, colour 1 at R = L-1 | B | colour 2 at L-2
//
// The plane loop is unrolled by 4 from x0-2, so the four-slot indices and the
// register set of a plane (D sets of raw prefetched bits, D | 4) are
// compile-time constants; the old-correction slots rotate through three scalar
// offsets.  Masks, scaling and conversions happen at the commit, so a plane's
// loads stay in flight for D plane steps, and planes outside the grid are not
// loaded at all (the commit stages zeros for them).  Colour-1 lanes cover the
// first-colour cells of the ring-extended tile (CY rows of CZ/2), colour-2
// lanes the second-colour cells of the tile (TY rows of TZ/2).
template <int TZ, int D, typename TR, typename T, bool ZERO>
__device__ __forceinline__ void sweep3_body(const Part* parts, int np, int vb, int level, int which_in, int first) {
  static_assert(D == 1 || D == 2 || D == 4, "the prefetch depth must divide the unroll of 4");
  FP_PART_AT(parts, np, vb);
  const State* st = cload(&Gp->st);
  if (cload(&st->done)) return;
  const LevelRec lvl = cload(&Gp->lv[level]);
  const bool top = level == 0;
  const W4<T>* __restrict__ wts = (const W4<T>*)lvl.w;
  const TR* __restrict__ r = (const TR*)(top ? (const void*)cload(&Gp->r) : lvl.r);
  const float rmul = (top ? 1.0f : lvl.Sr) / lvl.Se;
  const int dyn = top;
  const T* __restrict__ ein = (const T*)(which_in ? lvl.e2 : lvl.e);
  T* __restrict__ eout = (T*)(which_in ? lvl.e : lvl.e2);
  const int chunk = lvl.sw.chunk, tiles_z = lvl.sw.tz, tiles_y = lvl.sw.ty;
  constexpr int TY = 256 / TZ;
  constexpr int EY = TY + 4, EZ = TZ + 4;    // old correction: -2 .. T+1
  // coefficients on the old correction's window: their liquid flags mask it
  constexpr int WY = EY, WZ = EZ;
  constexpr int CY = TY + 2, CZ = TZ + 2;    // r, first-colour values: -1 .. T
  constexpr int WN = WY * WZ, EN = EY * EZ, CN = CY * CZ;
  constexpr int NW = (WN + 255) / 256, NE = ZERO ? 0 : (EN + 255) / 256, NR = (CN + 255) / 256;
  constexpr int HF = CZ / 2, HS = TZ / 2;
  static_assert(CY * HF <= 256 && TY * HS <= 256, "lane maps exceed the workgroup");
  using WB = typename W4Bits<T>::type;
  __shared__ WB s_w[4][WN];
  __shared__ T s_e[ZERO ? 1 : 3 * EN];
  __shared__ T s_r[4][CN];
  __shared__ T s_c[4][CN];

  const float rm = dyn ? rmul * cload(&st->rscale) : rmul;
  const int nx = lvl.d[0], ny = lvl.d[1], nz = lvl.d[2];
  int tzi, tyi, cxi;
  tile_of_group(blk, tiles_z, tiles_y, nblk, tzi, tyi, cxi, TY >= 16 ? 1 : 2);
  const int z0 = tzi * TZ, y0 = tyi * TY;
  int x0 = cxi * chunk;
  int x1 = min(x0 + chunk, nx);
  const int tid = threadIdx.x;
  const long long plane = (long long)ny * nz;
  const int second = 1 - first;
  // even rows: both colours of a plane leave in one pass of paired stores
  // (colour 2 writes each (z even, z odd) pair); odd nz keeps single stores
  const bool pairs = (nz & 1) == 0;
  static_assert(kHullTZ % TZ == 0 && TY % kHullTY == 0, "a sweep tile lies in one column of hull tiles");
  {
    // the level's 32 x 8 tiles carry liquid hulls (k_tile_hull): planes of
    // the chunk outside the hull are all air in this tile.  On coarse levels
    // they sweep to zero (the prolongation reads coarse air as zero); the
    // finest level's correction is read only inside the hulls or masked by
    // the liquid flags (these sweeps' staged windows, the marches), so its
    // planes outside the hull are left as they are.
    if (lvl.hull) {
      // the union of the hulls of the march tiles this tile lies in
      const int hz = z0 / kHullTZ, hy0 = y0 / kHullTY, hy1 = min((y0 + TY - 1) / kHullTY, lvl.mg.ty - 1);
      int hlo = nx, hhi = -1;
      for (int hy = hy0; hy <= hy1; ++hy) {
        const int* hp = lvl.hull + 2 * (hy * lvl.mg.tz + hz);
        hlo = min(hlo, cload(hp));
        hhi = max(hhi, cload(hp + 1));
      }
      const int xs = max(x0, hlo), xe = min(x1, hhi + 1);
      if (!top && (xs > x0 || xe < x1)) {
        const int zy = tid / (TZ / 2), zk = tid - (tid / (TZ / 2)) * (TZ / 2);
        const int gy = y0 + zy, gz = z0 + 2 * zk;
        if (tid < TY * (TZ / 2) && gy < ny && gz < nz)
          for (int x = x0; x < x1; ++x) {
            if (x >= xs && x < xe) continue;
            const size_t o = (size_t)((long long)x * plane) + (size_t)gy * nz + gz;
            if (pairs) {
              gst_pair(eout, o, 0.0f, 0.0f);
            } else {
              Store<T>::gst(eout, o, 0.0f);
              if (gz + 1 < nz) Store<T>::gst(eout, o + 1, 0.0f);
            }
          }
      }
      x0 = xs;
      x1 = xe;
      if (x0 >= x1) return;
    }
  }

  // in-plane offsets of the staged cells (-1: outside the grid)
  int ow[NW], oe[NE > 0 ? NE : 1], orr[NR];
#pragma unroll
  for (int k = 0; k < NW; ++k) {
    const int q = tid + 256 * k;
    const int yy = y0 + q / WZ - 2, zz = z0 + q % WZ - 2;
    ow[k] = (q < WN && yy >= 0 && yy < ny && zz >= 0 && zz < nz) ? yy * nz + zz : -1;
  }
#pragma unroll
  for (int k = 0; k < NE; ++k) {
    const int q = tid + 256 * k;
    const int yy = y0 + q / EZ - 2, zz = z0 + q % EZ - 2;
    oe[k] = (q < EN && yy >= 0 && yy < ny && zz >= 0 && zz < nz) ? yy * nz + zz : -1;
  }
#pragma unroll
  for (int k = 0; k < NR; ++k) {
    const int q = tid + 256 * k;
    const int yy = y0 + q / CZ - 1, zz = z0 + q % CZ - 1;
    orr[k] = (q < CN && yy >= 0 && yy < ny && zz >= 0 && zz < nz) ? yy * nz + zz : -1;
  }
  // D register sets of raw prefetched planes
  WB pw[D][NW];
  T pe[D][NE > 0 ? NE : 1];
  TR pr[D][NR];
  int lf = x0 - 2;                 // next plane to fetch
  long long fo = (long long)lf * plane;
  auto fetch = [&](int d) {
    if (lf >= 0 && lf < nx) {
      const W4<T>* wp = wts + fo;
      const TR* rp = r + fo;
#pragma unroll
      for (int k = 0; k < NW; ++k) pw[d][k] = ldg(reinterpret_cast<const WB*>(wp), (size_t)max(ow[k], 0));
      if constexpr (!ZERO) {
        const T* ep = ein + fo;
#pragma unroll
        for (int k = 0; k < NE; ++k) pe[d][k] = ldg(ep, (size_t)max(oe[k], 0));
      }
#pragma unroll
      for (int k = 0; k < NR; ++k) pr[d][k] = ldg(rp, (size_t)max(orr[k], 0));
    }
    ++lf;
    fo += plane;
  };

  // colour-1 lane: a first-colour cell of the ring-extended tile; its z
  // column alternates with the plane parity (index 0: even steps from x0-2)
  const bool c1 = tid < CY * HF;
  const int c1y = tid / HF, c1k = tid - (tid / HF) * HF;
  const int c1gy = y0 + c1y - 1;
  int c1lz[2];
  bool c1st[2];
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const int R = x0 - 3 + h;  // parity of R = L - 1 on steps of parity h
    c1lz[h] = 2 * c1k + ((first + R + c1gy + 1) & 1);
    const int gz = z0 + c1lz[h] - 1;
    c1st[h] = c1 && c1y >= 1 && c1y <= TY && c1lz[h] >= 1 && c1lz[h] <= TZ && c1gy < ny && gz < nz;
  }
  // colour-2 lane: a second-colour cell of the tile
  const bool c2 = tid < TY * HS;
  const int c2y = tid / HS, c2k = tid - (tid / HS) * HS;
  const int c2gy = y0 + c2y;
  int c2lz[2];
  bool c2st[2];
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const int B = x0 - 4 + h;  // parity of B = L - 2 on steps of parity h
    c2lz[h] = 2 * c2k + ((second + B + c2gy) & 1);
    c2st[h] = c2 && c2gy < ny && z0 + c2lz[h] < nz;
  }

#pragma unroll
  for (int d = 0; d < D; ++d) fetch(d);

  int e0 = 2 * EN, e1 = EN, e2 = 0;   // old-correction slots of planes L, L-1, L-2
  long long oR = (long long)(x0 - 3) * plane;  // offset of plane R = L - 1
  float wxm2 = 0.0f;                   // colour 2: -x coefficient of plane B-1, read one step early

  auto step = [&](int L, int d) {
    {
      const int t = e2;
      e2 = e1;
      e1 = e0;
      e0 = t;
    }
    const int h = d & 1;
    const int s0 = d, s1 = (d + 3) & 3, s2 = (d + 2) & 3, s3 = (d + 1) & 3;  // slots of L, L-1, L-2, L-3
    const int rs = d % D;
    const bool pin = (L >= 0 && L < nx);
    // ---- commit plane L ----
#pragma unroll
    for (int k = 0; k < NW; ++k) {
      const int q = tid + 256 * k;
      if (k + 1 < NW || q < WN) s_w[s0][q] = bits_unless(pw[rs][k], pin && ow[k] >= 0);
    }
    if constexpr (!ZERO) {
      // air cells stage a zero correction whatever the buffer holds there
      static_assert(WN == EN && NW == NE, "coefficients and correction share the staged window");
#pragma unroll
      for (int k = 0; k < NE; ++k) {
        const int q = tid + 256 * k;
        const W4<T> wq = __builtin_bit_cast(W4<T>, pw[rs][k]);
        if (k + 1 < NE || q < EN) s_e[e0 + q] = value_unless(pe[rs][k], pin && oe[k] >= 0 && wf(wq.id) > 0.0f);
      }
    }
#pragma unroll
    for (int k = 0; k < NR; ++k) {
      const int q = tid + 256 * k;
      const float v = f32_rounded(Store<TR>::ld(&pr[rs][k], 0) * rm);
      if (k + 1 < NR || q < CN) s_r[s0][q] = to_store<T>((pin && orr[k] >= 0) ? v : 0.0f);
    }
    lds_barrier();  // A
    if (L + D <= x1 + 1) fetch(rs);

    const int R = L - 1, B = L - 2;
    float wxm = 0.0f;
    if (L >= x0) {
      // ---- colour 1 at plane R ----
      const W4<T>* w0 = reinterpret_cast<const W4<T>*>(s_w[s1]);
      const W4<T>* wm = reinterpret_cast<const W4<T>*>(s_w[s2]);
      if (c1) {
        const int lz = c1lz[h];
        const int wi = (c1y + 1) * WZ + lz + 1;
        const int ci = c1y * CZ + lz;
        // all LDS reads first, then the arithmetic: one wait instead of a
        // read-wait-FMA chain per neighbour (the scheduler otherwise
        // interleaves them to shorten live ranges)
        const W4<T> wc = w0[wi];
        const T rv = s_r[s1][ci];
        T ev[6], wn[3];
        if constexpr (!ZERO) {
          const int ei = (c1y + 1) * EZ + lz + 1;
          const T* eo = s_e + e1;
          ev[0] = s_e[e0 + ei];
          ev[1] = s_e[e2 + ei];
          ev[2] = eo[ei + EZ];
          ev[3] = eo[ei - EZ];
          ev[4] = eo[ei + 1];
          ev[5] = eo[ei - 1];
          wn[0] = wm[wi].wx;
          wn[1] = w0[wi - WZ].wy;
          wn[2] = w0[wi - 1].wz;
        }
        __builtin_amdgcn_sched_barrier(0);
        const float idg = wf(wc.id);
        float acc = wf(rv);
        if constexpr (!ZERO) {
          acc += wf(wc.wx) * wf(ev[0]);
          acc += wf(wn[0]) * wf(ev[1]);
          acc += wf(wc.wy) * wf(ev[2]);
          acc += wf(wn[1]) * wf(ev[3]);
          acc += wf(wc.wz) * wf(ev[4]);
          acc += wf(wn[2]) * wf(ev[5]);
        }
        const float val = idg > 0.0f ? f32_rounded(acc * idg) : 0.0f;
        s_c[s1][ci] = to_store<T>(val);
        if (!pairs && R >= x0 && R < x1 && c1st[h])
          Store<T>::gst(eout + oR, (size_t)(c1gy * nz + z0 + lz - 1), val);
      }
      // colour-2 lanes read the -x coefficient of the next step's plane B - 1
      // (plane L - 2) now: its slot is refilled at the next commit
      if (c2) wxm = wf(wm[(c2y + 2) * WZ + c2lz[h ^ 1] + 2].wx);
    }
    lds_barrier();  // B
    if (L >= x0 + 2 && B < x1 && c2) {
      // ---- colour 2 at plane B ----
      const int lz = c2lz[h];
      const int wi = (c2y + 2) * WZ + lz + 2;
      const int ci = (c2y + 1) * CZ + lz + 1;
      const W4<T>* w0 = reinterpret_cast<const W4<T>*>(s_w[s2]);
      const T* cc = s_c[s2];
      const W4<T> wc = w0[wi];
      const T rv = s_r[s2][ci];
      const T c0 = s_c[s1][ci], c1_ = s_c[s3][ci], c2_ = cc[ci + CZ], c3 = cc[ci - CZ], c4 = cc[ci + 1],
              c5 = cc[ci - 1];
      const T wy_ = w0[wi - WZ].wy, wz_ = w0[wi - 1].wz;
      __builtin_amdgcn_sched_barrier(0);
      const float idg = wf(wc.id);
      float acc = wf(rv);
      acc += wf(wc.wx) * wf(c0);
      acc += wxm2 * wf(c1_);
      acc += wf(wc.wy) * wf(c2_);
      acc += wf(wy_) * wf(c3);
      acc += wf(wc.wz) * wf(c4);
      acc += wf(wz_) * wf(c5);
      const float val = idg > 0.0f ? f32_rounded(acc * idg) : 0.0f;
      if (c2st[h]) {
        if (pairs) {
          // the cell pair (z even, z odd) of this row in one store: the
          // first-colour partner is the staged value beside this cell
          const float pv = (lz & 1) ? wf(c5) : wf(c4);
          gst_pair(eout + (oR - plane), (size_t)(c2gy * nz + z0 + (lz & ~1)), (lz & 1) ? pv : val,
                   (lz & 1) ? val : pv);
        } else {
          Store<T>::gst(eout + (oR - plane), (size_t)(c2gy * nz + z0 + lz), val);
        }
      }
    }
    wxm2 = wxm;
    oR += plane;
  };

  for (int L0 = x0 - 2; L0 <= x1 + 1; L0 += 4) {
#pragma unroll
    for (int d = 0; d < 4; ++d)
      if (L0 + d <= x1 + 1) step(L0 + d, d);
  }
}

// One launch per sweep: workgroup b is the table's workgroup b.
template <int TZ, int D, typename TR, typename T, bool ZERO>
__global__ __launch_bounds__(256) void k_sweep3(const Part* parts, int np, int level, int which_in, int first) {
  sweep3_body<TZ, D, TR, T, ZERO>(parts, np, (int)blockIdx.x, level, which_in, first);
}

// The same sweep as a persistent launch: kPersistGroups workgroups (or
// fewer) walk the table's `total` workgroups with a stride of the grid size,
// a multiple of 8, so every workgroup keeps the XCD its table slots assume.
template <int TZ, int D, typename TR, typename T, bool ZERO>
__global__ __launch_bounds__(256) void k_sweep3_persistent(const Part* parts, int np, int total, int level,
                                                           int which_in, int first) {
  for (int vb = (int)blockIdx.x; vb < total; vb += (int)gridDim.x) {
    sweep3_body<TZ, D, TR, T, ZERO>(parts, np, vb, level, which_in, first);
    __syncthreads();  // the next table slot reuses the LDS planes
  }
}

}  // namespace fp

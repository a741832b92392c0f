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
// Variants (template flags):
//   ZERO     the sweep starts from e = 0 (first pre-smoothing sweep; e_in unread)
//   PROLONG  e_in is corrected on load by the trilinear prolongation of the
//            coarse correction (first post-smoothing sweep)
#pragma once

#include "fp_common.hpp"

namespace fp {

struct SweepGeom {
  int nx, ny, nz;
  float gx, gy, gz;
};

struct ProlongArgs {
  const void* ec;      // coarse correction (storage T; zero on coarse air cells), or nullptr
  int cnx, cny, cnz;
  int sx, sy, sz;      // strides fine/coarse per axis (1 or 2)
  float fac;           // Se_c / Se_f
};

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

// Workgroups are dispatched round-robin over the 8 XCDs of the device, each
// with its own L2: group g runs on XCD g % 8.  Remap so that every XCD walks a
// contiguous run of tiles (z fastest, then y, then the x chunk): neighbouring
// tiles, which read each other's halo rows, then share one L2.
__device__ __forceinline__ void tile_of_group(int g, int tiles_z, int tiles_y, int ngroups, int& tzi, int& tyi,
                                              int& cxi) {
  const int per = ngroups >> 3;
  const int l = (g < (per << 3)) ? (g & 7) * per + (g >> 3) : g;
  tzi = l % tiles_z;
  const int t = l / tiles_z;
  tyi = t % tiles_y;
  cxi = t / tiles_y;
}

template <typename T>
__device__ __forceinline__ float wf(const T& v) { return Store<T>::ld(&v, 0); }

// Bit-level views for branch-free masking of staged loads.
template <typename T> struct W4Bits;
template <> struct W4Bits<__half> { using type = uint2; };
template <> struct W4Bits<__hip_bfloat16> { using type = uint2; };
template <> struct W4Bits<float> { using type = uint4; };

__device__ __forceinline__ uint2 bits_unless(uint2 b, bool keep) { return keep ? b : make_uint2(0u, 0u); }
__device__ __forceinline__ uint4 bits_unless(uint4 b, bool keep) { return keep ? b : make_uint4(0u, 0u, 0u, 0u); }

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

// Taps of the cell-centred linear prolongation along one axis, as positions
// in a staged coarse window starting at coarse index c0: fine f of stride s
// reads coarse f>>1 with 3/4 and its neighbour on f's side with 1/4, or
// coarse f>>1 alone at the domain edge; s = 1 copies.
__device__ __forceinline__ void window_taps(int f, int s, int nc, int c0, int& a, float& wa, int& b, float& wb) {
  if (s == 1) {
    a = f - c0;
    wa = 1.0f;
    b = a;
    wb = 0.0f;
    return;
  }
  const int I = f >> 1;
  const int J = (f & 1) ? I + 1 : I - 1;
  const bool two = (J >= 0 && J < nc);
  a = I - c0;
  wa = two ? 0.75f : 1.0f;
  b = two ? J - c0 : a;
  wb = two ? 0.25f : 0.0f;
}

// One red-black sweep.  TZ x TY = 256 lanes; TR: storage of the right-hand
// side, T: of e, of the packed coefficients and of the LDS staging.  LDS holds
// four planes of coefficients (rows/columns -2..T of the tile; -2..T+1 with
// PROLONG), three of the old correction (-2..T+1), four of the scaled
// right-hand side and three of the first-colour values (-1..T).  Per plane
// step L two barriers order the work:
//   commit plane L | A | prefetch L+1, colour 1 at R = L-1 | B | colour 2 at L-2
// (PROLONG adds the coarse correction to plane L's second-colour liquid cells
// between A and a third barrier).  The prefetch holds raw loaded bits: masks,
// scaling and conversions happen at the commit, so the loads of plane L+1
// stay in flight across the colour passes of the current planes.  Colour-1
// lanes cover the first-colour cells of the ring-extended tile (CY rows of
// CZ/2), colour-2 lanes the second-colour cells of the tile (TY rows of TZ/2).
template <int TZ, typename TR, typename T, bool PROLONG, bool ZERO>
__global__ __launch_bounds__(256) void k_sweep3(SweepGeom G, const W4<T>* __restrict__ wts,
                                                const TR* __restrict__ r, float rmul, int dyn,
                                                const T* __restrict__ ein, T* __restrict__ eout, int first,
                                                ProlongArgs P, int chunk, int tiles_z, int tiles_y, int ngroups,
                                                const State* st) {
  if (st->done) return;
  constexpr int TY = 256 / TZ;
  constexpr int EY = TY + 4, EZ = TZ + 4;                    // old correction: -2 .. T+1
  constexpr int WY = PROLONG ? EY : TY + 3;                  // coefficients: -2 .. T (T+1)
  constexpr int WZ = PROLONG ? EZ : TZ + 3;
  constexpr int CY = TY + 2, CZ = TZ + 2;                    // r, first-colour values: -1 .. T
  constexpr int WN = WY * WZ, EN = EY * EZ, CN = CY * CZ;
  constexpr int PN = PROLONG ? 2 * EN : 0;                   // two coarse planes, window <= EY x EZ
  constexpr int NW = (WN + 255) / 256, NE = ZERO ? 0 : (EN + 255) / 256, NR = (CN + 255) / 256;
  constexpr int NP = (PN + 255) / 256;
  constexpr int HF = CZ / 2, HS = TZ / 2;
  static_assert(CY * HF <= 256 && TY * HS <= 256, "lane maps exceed the workgroup");
  using WB = typename W4Bits<T>::type;
  __shared__ WB s_w[4][WN];
  __shared__ T s_e[3][EN];
  __shared__ T s_r[4][CN];
  __shared__ T s_c[3][CN];
  __shared__ T s_p[PROLONG ? PN : 1];

  const float rm = dyn ? rmul * st->rscale : rmul;
  const int nx = G.nx, ny = G.ny, nz = G.nz;
  int tzi, tyi, cxi;
  tile_of_group(blockIdx.x, tiles_z, tiles_y, ngroups, tzi, tyi, cxi);
  const int z0 = tzi * TZ, y0 = tyi * TY, x0 = cxi * chunk;
  const int x1 = min(x0 + chunk, nx);
  const int tid = threadIdx.x;
  const size_t plane = (size_t)ny * nz;
  const int second = 1 - first;

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
  // coarse window (PROLONG): rows/columns from cy0/cz0, planes pa(L), pa(L)+1
  const int cy0 = PROLONG ? (P.sy == 2 ? (y0 >> 1) - 2 : y0 - 2) : 0;
  const int cz0 = PROLONG ? (P.sz == 2 ? (z0 >> 1) - 2 : z0 - 2) : 0;
  const size_t cplane = PROLONG ? (size_t)P.cny * P.cnz : 0;
  int op[NP > 0 ? NP : 1], opl[NP > 0 ? NP : 1];
#pragma unroll
  for (int k = 0; k < NP; ++k) {
    const int q = tid + 256 * k;
    const int qq = q % EN;
    const int yy = cy0 + qq / EZ, zz = cz0 + qq % EZ;
    opl[k] = q / EN;
    op[k] = (q < PN && yy >= 0 && yy < P.cny && zz >= 0 && zz < P.cnz) ? yy * P.cnz + zz : -1;
  }
  auto coarse_plane0 = [&](int L) { return P.sx == 2 ? ((L - 1) >> 1) : L; };

  WB pw[NW];
  T pe[NE > 0 ? NE : 1];
  TR pr[NR];
  T pp[NP > 0 ? NP : 1];
  auto fetch = [&](int L) {
    const size_t base = (size_t)min(max(L, 0), nx - 1) * plane;
#pragma unroll
    for (int k = 0; k < NW; ++k) pw[k] = *reinterpret_cast<const WB*>(wts + base + max(ow[k], 0));
#pragma unroll
    for (int k = 0; k < NE; ++k) pe[k] = ein[base + max(oe[k], 0)];
#pragma unroll
    for (int k = 0; k < NR; ++k) pr[k] = r[base + max(orr[k], 0)];
    if constexpr (PROLONG) {
      const int pa = coarse_plane0(L);
      const T* ec = (const T*)P.ec;
#pragma unroll
      for (int k = 0; k < NP; ++k) {
        const int c = min(max(pa + opl[k], 0), P.cnx - 1);
        pp[k] = ec[(size_t)c * cplane + max(op[k], 0)];
      }
    }
  };

  // colour-1 lane: a first-colour cell of the ring-extended tile
  const bool c1 = tid < CY * HF;
  const int c1y = tid / HF, c1k = tid - (tid / HF) * HF;
  const int c1gy = y0 + c1y - 1;
  // colour-2 lane: a second-colour cell of the tile
  const bool c2 = tid < TY * HS;
  const int c2y = tid / HS, c2k = tid - (tid / HS) * HS;
  const int c2gy = y0 + c2y;

  fetch(x0 - 2);
  for (int L = x0 - 2; L <= x1 + 1; ++L) {
    const int m4 = (L + 4) & 3, m3 = (L + 3) % 3;
    const bool pin = (L >= 0 && L < nx);
    // ---- commit plane L ----
#pragma unroll
    for (int k = 0; k < NW; ++k) {
      const int q = tid + 256 * k;
      if (k + 1 < NW || q < WN) s_w[m4][q] = bits_unless(pw[k], pin && ow[k] >= 0);
    }
#pragma unroll
    for (int k = 0; k < (EN + 255) / 256; ++k) {
      const int q = tid + 256 * k;
      if (k + 1 < (EN + 255) / 256 || q < EN) {
        if constexpr (ZERO) s_e[m3][q] = to_store<T>(0.0f);
        else s_e[m3][q] = value_unless(pe[k], pin && oe[k] >= 0);
      }
    }
#pragma unroll
    for (int k = 0; k < NR; ++k) {
      const int q = tid + 256 * k;
      const float v = Store<TR>::ld(&pr[k], 0) * rm;
      if (k + 1 < NR || q < CN) s_r[m4][q] = to_store<T>((pin && orr[k] >= 0) ? v : 0.0f);
    }
    if constexpr (PROLONG) {
      const int pa = coarse_plane0(L);
#pragma unroll
      for (int k = 0; k < NP; ++k) {
        const int q = tid + 256 * k;
        const int c = pa + opl[k];
        if (k + 1 < NP || q < PN) s_p[q] = value_unless(pp[k], c >= 0 && c < P.cnx && op[k] >= 0);
      }
    }
    lds_barrier();  // A
    if (L + 1 <= x1 + 1) fetch(L + 1);

    if constexpr (PROLONG) {
      // e(L) += fac * P ec at second-colour liquid cells (coarse air holds e = 0)
      if (pin) {
        int xa, xb;
        float wxa, wxb;
        window_taps(L, P.sx, P.cnx, coarse_plane0(L), xa, wxa, xb, wxb);
        for (int q = tid; q < EN; q += 256) {
          const int ey = q / EZ, ez = q - (q / EZ) * EZ;
          const int gy = y0 + ey - 2, gz = z0 + ez - 2;
          const W4<T> wc = *reinterpret_cast<const W4<T>*>(&s_w[m4][q]);
          if (gy >= 0 && gy < ny && gz >= 0 && gz < nz && ((L + gy + gz) & 1) == second && wf(wc.id) > 0.0f) {
            int ya, yb, za, zb;
            float wya, wyb, wza, wzb;
            window_taps(gy, P.sy, P.cny, cy0, ya, wya, yb, wyb);
            window_taps(gz, P.sz, P.cnz, cz0, za, wza, zb, wzb);
            const T* pa_ = s_p + xa * EN;
            const T* pb_ = s_p + xb * EN;
            const float acc =
                wxa * (wya * (wza * wf(pa_[ya * EZ + za]) + wzb * wf(pa_[ya * EZ + zb])) +
                       wyb * (wza * wf(pa_[yb * EZ + za]) + wzb * wf(pa_[yb * EZ + zb]))) +
                wxb * (wya * (wza * wf(pb_[ya * EZ + za]) + wzb * wf(pb_[ya * EZ + zb])) +
                       wyb * (wza * wf(pb_[yb * EZ + za]) + wzb * wf(pb_[yb * EZ + zb])));
            s_e[m3][q] = to_store<T>(wf(s_e[m3][q]) + P.fac * acc);
          }
        }
      }
      lds_barrier();  // A2
    }

    const int B = L - 2;
    float wxm2 = 0.0f;
    int wi2 = 0, ci2 = 0, gz2 = 0;
    if (L >= x0) {
      // ---- colour 1 at plane R = L - 1 ----
      const int R = L - 1;
      const int sR = (R + 4) & 3, sRm = (R + 3) & 3;
      const int eR = (R + 3) % 3, eRm = (R + 2) % 3, eRp = m3;
      const W4<T>* w0 = reinterpret_cast<const W4<T>*>(s_w[sR]);
      const W4<T>* wm = reinterpret_cast<const W4<T>*>(s_w[sRm]);
      if (c1) {
        const int lz = 2 * c1k + ((first + R + c1gy + 1) & 1);
        const int gz = z0 + lz - 1;
        const int wi = (c1y + 1) * WZ + lz + 1;
        const int ei = (c1y + 1) * EZ + lz + 1;
        const W4<T> wc = w0[wi];
        const float idg = wf(wc.id);
        const T* e0 = s_e[eR];
        float acc = wf(s_r[sR][c1y * CZ + lz]);
        acc += wf(wc.wx) * wf(s_e[eRp][ei]);
        acc += wf(wm[wi].wx) * wf(s_e[eRm][ei]);
        acc += wf(wc.wy) * wf(e0[ei + EZ]);
        acc += wf(w0[wi - WZ].wy) * wf(e0[ei - EZ]);
        acc += wf(wc.wz) * wf(e0[ei + 1]);
        acc += wf(w0[wi - 1].wz) * wf(e0[ei - 1]);
        const float val = idg > 0.0f ? acc * idg : 0.0f;
        s_c[eR][c1y * CZ + lz] = to_store<T>(val);
        if (R >= x0 && R < x1 && c1y >= 1 && c1y <= TY && lz >= 1 && lz <= TZ && c1gy < ny && gz < nz)
          Store<T>::st(eout, (size_t)R * plane + (size_t)c1gy * nz + gz, val);
      }
      // colour-2 lanes read the -x coefficient of plane B - 1 now: its slot
      // is refilled at the next commit, which no barrier separates from colour 2
      if (c2) {
        const int lz = 2 * c2k + ((second + B + c2gy) & 1);
        gz2 = z0 + lz;
        wi2 = (c2y + 2) * WZ + lz + 2;
        ci2 = (c2y + 1) * CZ + lz + 1;
        wxm2 = wf(reinterpret_cast<const W4<T>*>(s_w[(B + 3) & 3])[wi2].wx);
      }
    }
    lds_barrier();  // B
    if (L >= x0 + 2 && B < x1 && c2) {
      // ---- colour 2 at plane B = L - 2 ----
      const int sB = (B + 4) & 3;
      const int cB = (B + 3) % 3, cBm = (B + 2) % 3, cBp = (B + 4) % 3;
      const W4<T>* w0 = reinterpret_cast<const W4<T>*>(s_w[sB]);
      const W4<T> wc = w0[wi2];
      const float idg = wf(wc.id);
      const T* c0 = s_c[cB];
      float acc = wf(s_r[sB][ci2]);
      acc += wf(wc.wx) * wf(s_c[cBp][ci2]);
      acc += wxm2 * wf(s_c[cBm][ci2]);
      acc += wf(wc.wy) * wf(c0[ci2 + CZ]);
      acc += wf(w0[wi2 - WZ].wy) * wf(c0[ci2 - CZ]);
      acc += wf(wc.wz) * wf(c0[ci2 + 1]);
      acc += wf(w0[wi2 - 1].wz) * wf(c0[ci2 - 1]);
      const float val = idg > 0.0f ? acc * idg : 0.0f;
      if (c2gy < ny && gz2 < nz) Store<T>::st(eout, (size_t)B * plane + (size_t)c2gy * nz + gz2, val);
    }
  }
}

}  // namespace fp

// fp_sweep.hpp -- one full red-black Gauss-Seidel sweep in a single pass.
//
// The smoother of mgSmoothColour (webgpu-uniform-pressure-multigrid.wgsl.ts:325-344,
// without the LCP clamp) applied as both colour passes of one sweep, written for
// CDNA4: a 256-lane workgroup owns a (TY x TZ) column tile (z fastest, one
// 64-lane wave per TZ=64 row) and marches along x through a chunk of planes.
// Each plane is staged once in LDS with a two-cell halo; the first colour is
// updated on the tile plus a one-cell ring from the OLD field (read-only input
// buffer), the second colour on the tile from those fresh values, so the
// result is exactly one red-black sweep no matter how blocks are scheduled
// (old and new fields are different buffers).  One read of e, phi and r and
// one write of e per cell replace the two passes of separate colour kernels.
//
// Optional fusions:
//   zero_init  the sweep starts from e = 0 (first pre-smoothing sweep; e_in unread)
//   prolong    e_in is corrected on load by the trilinear prolongation of the
//              coarse correction (first post-smoothing sweep)
#pragma once

#include "fp_common.hpp"

namespace fp {

template <typename T>
__device__ __forceinline__ float sw_ld(const T* p, size_t i) { return Store<T>::ld(p, i); }

struct SweepGeom {
  int nx, ny, nz;
  float gx, gy, gz;
};

struct ProlongArgs {
  const void* ec;      // coarse correction (storage T), or nullptr
  const float* phic;   // coarse level set (FP32)
  int cnx, cny, cnz;
  int sx, sy, sz;      // strides fine/coarse per axis (1 or 2)
  float fac;           // Se_c / Se_f
};

__device__ __forceinline__ int sw_ptaps(int f, int s, int nc, int* c, float* w) {
  if (s == 1) { c[0] = f; w[0] = 1.0f; return 1; }
  const int I = f >> 1;
  const int J = (f & 1) ? I + 1 : I - 1;
  if (J < 0 || J >= nc) { c[0] = I; w[0] = 1.0f; return 1; }
  c[0] = I; w[0] = 0.75f; c[1] = J; w[1] = 0.25f;
  return 2;
}

template <typename T>
__device__ __forceinline__ float sw_prolong(const ProlongArgs& P, int i, int j, int k) {
  int ci[2], cj[2], ck[2];
  float wi[2], wj[2], wk[2];
  const int ni = sw_ptaps(i, P.sx, P.cnx, ci, wi);
  const int nj = sw_ptaps(j, P.sy, P.cny, cj, wj);
  const int nk = sw_ptaps(k, P.sz, P.cnz, ck, wk);
  const T* ec = (const T*)P.ec;
  float acc = 0.0f;
  for (int a = 0; a < ni; ++a)
    for (int b = 0; b < nj; ++b) {
      const float wab = wi[a] * wj[b];
      const size_t base = ((size_t)ci[a] * P.cny + cj[b]) * P.cnz;
      for (int q = 0; q < nk; ++q) {
        const size_t m = base + ck[q];
        if (P.phic[m] < 0.0f) acc += wab * wk[q] * sw_ld(ec, m);
      }
    }
  return P.fac * acc;
}

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

template <typename T>
__device__ __forceinline__ float wf(const T& v) { return Store<T>::ld(&v, 0); }

// TZ x TY = 256 lanes.  TR: storage of the right-hand side, T: of e and of
// the packed coefficients.  Lanes are mapped colour-compactly: the first
// colour covers only its cells of the tile plus a one-cell ring (TY+2 rows of
// TZ/2+1), the second colour only its cells of the tile (TY rows of TZ/2).
// Old values are read only at second-colour cells, so only those are loaded
// (and corrected by the prolongation when it is fused).
template <int TZ, typename TR, typename T>
__global__ __launch_bounds__(256) void k_sweep2(SweepGeom G, const W4<T>* __restrict__ wts,
                                                const TR* __restrict__ r, float rmul, int dyn,
                                                const T* __restrict__ ein, T* __restrict__ eout, int first,
                                                int zero_init, ProlongArgs P, int chunk, const State* st) {
  if (st->done) return;
  constexpr int TY = 256 / TZ;
  constexpr int RZ2 = TZ + 4, RY2 = TY + 4;  // two-cell halo (old field, coefficients)
  constexpr int RZ1 = TZ + 2, RY1 = TY + 2;  // one-cell halo (first colour, r)
  constexpr int HF = RZ1 / 2;                // first-colour cells per ring row
  constexpr int HS = TZ / 2;                 // second-colour cells per tile row
  __shared__ W4<T> s_w[4][RY2][RZ2];
  __shared__ float s_e[4][RY2][RZ2];
  __shared__ float s_r[4][RY1][RZ1];
  __shared__ float s_c[3][RY1][RZ1];

  const float rm = dyn ? rmul * st->rscale : rmul;
  const int nx = G.nx, ny = G.ny, nz = G.nz;
  const int z0 = blockIdx.x * TZ, y0 = blockIdx.y * TY;
  const int x0 = blockIdx.z * chunk;
  const int x1 = min(x0 + chunk, nx);
  const int tid = threadIdx.x;
  const size_t sxs = (size_t)ny * nz;
  const int second = 1 - first;
  const bool use_prolong = P.ec != nullptr;

  constexpr int NE = (RY2 * RZ2 + 255) / 256;
  constexpr int NR = (RY1 * RZ1 + 255) / 256;
  W4<T> pw[NE];
  float pe[NE], pr[NR];
  auto fetch = [&](int L) {
    const bool pin = (L >= 0 && L < nx);
#pragma unroll
    for (int k = 0; k < NE; ++k) {
      const int q = tid + 256 * k;
      float ev = 0.0f;
      W4<T> wv;
      Store<T>::st(&wv.wx, 0, 0.0f);
      Store<T>::st(&wv.wy, 0, 0.0f);
      Store<T>::st(&wv.wz, 0, 0.0f);
      Store<T>::st(&wv.id, 0, 0.0f);
      if (q < RY2 * RZ2) {
        const int ly = q / RZ2, lz = q - ly * RZ2;
        const int gy = y0 + ly - 2, gz = z0 + lz - 2;
        if (pin && gy >= 0 && gy < ny && gz >= 0 && gz < nz) {
          const size_t idx = (size_t)L * sxs + (size_t)gy * nz + gz;
          wv = wts[idx];
          if (!zero_init && ((L + gy + gz) & 1) == second && wf(wv.id) > 0.0f) {
            ev = sw_ld(ein, idx);
            if (use_prolong) ev += sw_prolong<T>(P, L, gy, gz);
          }
        }
      }
      pe[k] = ev;
      pw[k] = wv;
    }
#pragma unroll
    for (int k = 0; k < NR; ++k) {
      const int q = tid + 256 * k;
      float rv = 0.0f;
      if (q < RY1 * RZ1) {
        const int ly = q / RZ1, lz = q - ly * RZ1;
        const int gy = y0 + ly - 1, gz = z0 + lz - 1;
        if (pin && gy >= 0 && gy < ny && gz >= 0 && gz < nz) rv = sw_ld(r, (size_t)L * sxs + (size_t)gy * nz + gz) * rm;
      }
      pr[k] = rv;
    }
  };
  auto commit = [&](int L) {
    const int slot = (L + 8) & 3;
#pragma unroll
    for (int k = 0; k < NE; ++k) {
      const int q = tid + 256 * k;
      if (q < RY2 * RZ2) {
        const int ly = q / RZ2, lz = q - ly * RZ2;
        s_e[slot][ly][lz] = pe[k];
        s_w[slot][ly][lz] = pw[k];
      }
    }
#pragma unroll
    for (int k = 0; k < NR; ++k) {
      const int q = tid + 256 * k;
      if (q < RY1 * RZ1) {
        const int ly = q / RZ1, lz = q - ly * RZ1;
        s_r[slot][ly][lz] = pr[k];
      }
    }
  };

  fetch(x0 - 2);
  for (int L = x0 - 2; L <= x1 + 1; ++L) {
    // ---- plane L (prefetched) into LDS; start the loads of plane L + 1 ----
    commit(L);
    lds_barrier();
    if (L + 1 <= x1 + 1) fetch(L + 1);

    // ---- first colour at plane R = L - 1 on the tile plus a one-cell ring ----
    const int R = L - 1;
    if (R >= x0 - 1 && R <= x1 && (R < 0 || R >= nx)) {
      // a plane outside the domain: its first-colour values read as zero
      // (their weights are zero, but stale LDS bits could be NaN)
      const int cslot = (R + 9) % 3;
      for (int q = tid; q < RY1 * RZ1; q += 256) (&s_c[cslot][0][0])[q] = 0.0f;
    }
    if (R >= x0 - 1 && R <= x1 && R >= 0 && R < nx && tid < RY1 * HF) {
      const int sR = (R + 8) & 3, sRm = (R + 7) & 3, sRp = (R + 9) & 3;
      const int cslot = (R + 9) % 3;
      const int ly = tid / HF, k = tid - ly * HF;
      const int gy = y0 + ly - 1;
      const int lz = 2 * k + ((first + R + gy + 1) & 1);
      const int gz = z0 + lz - 1;
      const int ay = ly + 1, az = lz + 1;  // position in the two-halo arrays
      const W4<T> wc = s_w[sR][ay][az];
      const float idg = wf(wc.id);
      float val = 0.0f;
      if (idg > 0.0f) {
        float acc = s_r[sR][ly][lz];
        acc += wf(wc.wx) * s_e[sRp][ay][az];
        acc += wf(s_w[sRm][ay][az].wx) * s_e[sRm][ay][az];
        acc += wf(wc.wy) * s_e[sR][ay + 1][az];
        acc += wf(s_w[sR][ay - 1][az].wy) * s_e[sR][ay - 1][az];
        acc += wf(wc.wz) * s_e[sR][ay][az + 1];
        acc += wf(s_w[sR][ay][az - 1].wz) * s_e[sR][ay][az - 1];
        val = acc * idg;
      }
      s_c[cslot][ly][lz] = val;
      // first-colour cells of the tile are final: write them now
      if (R >= x0 && R < x1 && ly >= 1 && ly <= TY && lz >= 1 && lz <= TZ && gy < ny && gz < nz)
        Store<T>::st(eout, (size_t)R * sxs + (size_t)gy * nz + gz, val);
    }
    lds_barrier();

    // ---- second colour at plane B = L - 2 on the tile ----
    const int B = L - 2;
    if (B >= x0 && B < x1 && tid < TY * HS) {
      const int ly = tid / HS, k = tid - ly * HS;
      const int gy = y0 + ly;
      const int lz = 2 * k + ((second + B + gy) & 1);
      const int gz = z0 + lz;
      if (gy < ny && gz < nz) {
        const int sB = (B + 8) & 3, sBm = (B + 7) & 3;
        const int cB = (B + 9) % 3, cBm = (B + 8) % 3, cBp = (B + 10) % 3;
        const int ay = ly + 2, az = lz + 2, cy = ly + 1, cz = lz + 1;
        const W4<T> wc = s_w[sB][ay][az];
        const float idg = wf(wc.id);
        float val = 0.0f;
        if (idg > 0.0f) {
          float acc = s_r[sB][cy][cz];
          acc += wf(wc.wx) * s_c[cBp][cy][cz];
          acc += wf(s_w[sBm][ay][az].wx) * s_c[cBm][cy][cz];
          acc += wf(wc.wy) * s_c[cB][cy + 1][cz];
          acc += wf(s_w[sB][ay - 1][az].wy) * s_c[cB][cy - 1][cz];
          acc += wf(wc.wz) * s_c[cB][cy][cz + 1];
          acc += wf(s_w[sB][ay][az - 1].wz) * s_c[cB][cy][cz - 1];
          val = acc * idg;
        }
        Store<T>::st(eout, (size_t)B * sxs + (size_t)gy * nz + gz, val);
      }
    }
    lds_barrier();
  }
}

}  // namespace fp

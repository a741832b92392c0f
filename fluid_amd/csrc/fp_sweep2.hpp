// fp_sweep2.hpp -- two red-black Gauss-Seidel sweeps of the finest level in
// one pass over HBM (temporal blocking along the x march).
//
// The smoother is fp_sweep.hpp's (mgSmoothColour,
// webgpu-uniform-pressure-multigrid.wgsl.ts:325-344, without the LCP clamp),
// and so is the arithmetic of every cell update, term for term, so two
// sweeps here give the bits of two single-sweep launches.  A 256-lane
// workgroup owns a (TY x TZ) column tile and marches along x.  Each plane of
// coefficients, old correction and right-hand side is staged once in LDS
// with a four-cell halo; the four colour passes of the two sweeps run one
// plane behind each other:
//
//   step L: commit plane L | A | pass 0 at L-1 | B | pass 1 at L-2 | C |
//           pass 2 at L-3 | D | pass 3 at L-4 (stores the tile's plane L-4)
//
// Pass p updates its colour ((first + p) & 1 of x + y + z) on the tile plus a
// ring of 3 - p cells, from the outputs of pass p - 1 (pass 0: the old
// correction, or zero), so the tile's result needs no other workgroup's
// values.  Coefficients and the old correction are read once per cell (plus
// halo), the right-hand side once, the result written once: about half the
// HBM traffic of two sweep launches.  Only the finest level runs these: its
// planes outside the liquid hulls are left as they are (readers mask air).
//
// Opt-in (fp_tuning.sweep_pairs): on MI355X a pair takes 0.358 ms at 384^3
// against 2 x 0.146 ms for two single-sweep launches.  The pair does about
// 30 % more cell updates per sweep (the rings of passes 0-2) at half the
// single sweep's occupancy (48.8 KB of LDS: 3 workgroups per CU against 7),
// and the single sweep is bound by its LDS passes, not by HBM, so halving
// the HBM traffic does not pay for them.
//
// Variant (template flag):
//   ZERO     the first sweep starts from e = 0 (pre-smoothing; e_in unread)
#pragma once

#include "fp_sweep.hpp"

namespace fp {

template <int P>
struct PassNo {
  static constexpr int value = P;
};

template <int TZ, int D, typename TR, typename T, bool ZERO>
__device__ __forceinline__ void sweep2_body(const Part* parts, int np, int vb, int which_in, int first) {
  FP_PART_AT(parts, np, vb);
  const State* st = cload(&Gp->st);
  if (cload(&st->done)) return;
  const LevelRec lvl = cload(&Gp->lv[0]);
  const W4<T>* __restrict__ wts = (const W4<T>*)lvl.w;
  const TR* __restrict__ r = (const TR*)cload(&Gp->r);
  const T* __restrict__ ein = (const T*)(which_in ? lvl.e2 : lvl.e);
  T* __restrict__ eout = (T*)(which_in ? lvl.e : lvl.e2);
  const int chunk = lvl.sw.chunk, tiles_z = lvl.sw.tz, tiles_y = lvl.sw.ty;
  constexpr int TY = 256 / TZ;
  constexpr int WZ = TZ + 8, WN = (TY + 8) * WZ;  // coefficients, old correction: -4 .. T+3
  constexpr int RZ = TZ + 6, RN = (TY + 6) * RZ;  // right-hand side, pass 0: -3 .. T+2
  constexpr int Z1 = TZ + 4, P1N = (TY + 4) * Z1;  // pass 1: -2 .. T+1
  constexpr int Z2 = TZ + 2, P2N = (TY + 2) * Z2;  // pass 2: -1 .. T
  constexpr int NW = (WN + 255) / 256, NE = ZERO ? 0 : NW, NR = (RN + 255) / 256;
  // Plane slots.  Coefficients: planes L .. L-5 (pass 3 reads its -x
  // coefficient one step early, so the commit of L+1 may take the slot of
  // L-5); right-hand side: L .. L-5 (pass 3 reads L-4 beside the next
  // commit); pass outputs and the old correction: three planes each.
  constexpr int SW = 6, SR = 6;
  using WB = typename W4Bits<T>::type;
  __shared__ WB s_w[SW][WN];
  __shared__ T s_e[ZERO ? 1 : 3][ZERO ? 1 : WN];
  __shared__ T s_r[SR][RN];
  __shared__ T s_p0[3][RN];
  __shared__ T s_p1[3][P1N];
  __shared__ T s_p2[3][P2N];

  const float rm = (1.0f / lvl.Se) * cload(&st->rscale);  // as fp_sweep.hpp forms it
  const int nx = lvl.d[0], ny = lvl.d[1], nz = lvl.d[2];
  int tzi, tyi, cxi;
  tile_of_group(blk, tiles_z, tiles_y, nblk, tzi, tyi, cxi, TY >= 16 ? 1 : 2);
  const int z0 = tzi * TZ, y0 = tyi * TY;
  int x0 = cxi * chunk;
  int x1 = min(x0 + chunk, nx);
  const int tid = threadIdx.x;
  const long long plane = (long long)ny * nz;
  const bool pairs = (nz & 1) == 0;
  static_assert(kHullTZ % TZ == 0 && TY % kHullTY == 0, "a sweep tile lies in one column of hull tiles");
  if (lvl.hull) {
    const int hz = z0 / kHullTZ, hy0 = y0 / kHullTY, hy1 = min((y0 + TY - 1) / kHullTY, lvl.mg.ty - 1);
    int hlo = nx, hhi = -1;
    for (int hy = hy0; hy <= hy1; ++hy) {
      const int* hp = lvl.hull + 2 * (hy * lvl.mg.tz + hz);
      hlo = min(hlo, cload(hp));
      hhi = max(hhi, cload(hp + 1));
    }
    x0 = max(x0, hlo);
    x1 = min(x1, hhi + 1);
  }
  if (x0 >= x1) return;

  // in-plane offsets of the staged cells (-1: outside the grid)
  int ow[NW], orr[NR];
#pragma unroll
  for (int k = 0; k < NW; ++k) {
    const int q = tid + 256 * k;
    const int yy = y0 + q / WZ - 4, zz = z0 + q % WZ - 4;
    ow[k] = (q < WN && yy >= 0 && yy < ny && zz >= 0 && zz < nz) ? yy * nz + zz : -1;
  }
#pragma unroll
  for (int k = 0; k < NR; ++k) {
    const int q = tid + 256 * k;
    const int yy = y0 + q / RZ - 3, zz = z0 + q % RZ - 3;
    orr[k] = (q < RN && yy >= 0 && yy < ny && zz >= 0 && zz < nz) ? yy * nz + zz : -1;
  }
  // raw bits of the D planes in flight (set d: planes x0 - 4 + d + D j).
  // Depths 2-4 fit (the occupancy is set by LDS) but measure slower: the
  // kernel is bound by its LDS passes and barriers, not by load latency.
  static_assert(D >= 1 && D <= 4, "prefetch depth");
  WB pw[D][NW];
  T pe[D][NE > 0 ? NE : 1];
  TR pr[D][NR];
  auto fetch = [&](int x, int d) {
    if (x >= 0 && x < nx) {
      const long long fo = (long long)x * plane;
      const W4<T>* wp = wts + fo;
      const TR* rp = r + fo;
#pragma unroll
      for (int k = 0; k < NW; ++k) pw[d][k] = ldg(reinterpret_cast<const WB*>(wp), (size_t)max(ow[k], 0));
      if constexpr (!ZERO) {
        const T* ep = ein + fo;
#pragma unroll
        for (int k = 0; k < NE; ++k) pe[d][k] = ldg(ep, (size_t)max(ow[k], 0));
      }
#pragma unroll
      for (int k = 0; k < NR; ++k) pr[d][k] = ldg(rp, (size_t)max(orr[k], 0));
    }
  };
  // plane slots: planes from x0 - 4 on, offset to keep the modulus positive
  auto wslot = [&](int x) { return (x - x0 + 12) % SW; };
  auto rslot = [&](int x) { return (x - x0 + 12) % SR; };
  auto tslot = [&](int x) { return (x - x0 + 12) % 3; };

  // pass 3's lanes: one second-of-pair cell of the tile each (TY rows of TZ/2)
  constexpr int H3 = TZ / 2;
  const bool l3 = tid < TY * H3;
  const int py3 = tid / H3, hz3 = tid - (tid / H3) * H3;
  float wxm3 = 0.0f;  // -x coefficient of pass 3's cell at its plane (read a step early)

  // One colour pass at plane X on the tile plus a ring of R = 3 - P cells.
  // prev: the previous pass's output slots of planes X+1, X, X-1 (ring
  // R + 1, rows of PZ + 2), or none (pass 0 from zero); out: this pass's slot
  // of plane X (passes 0-2).  Pass 3 stores the tile's cell pairs instead.
  auto pass = [&](auto PC, int X, const T* pxp, const T* pxc, const T* pxm, T* out) {
    constexpr int P = decltype(PC)::value;
    constexpr int R = 3 - P;
    constexpr int PZ = TZ + 2 * R, HZ = PZ / 2, CNT = (TY + 2 * R) * HZ;
    constexpr int QZ = PZ + 2;  // rows of prev
    constexpr bool FROM_ZERO = ZERO && P == 0;
    const int colour = (first + P) & 1;
    const W4<T>* w0 = reinterpret_cast<const W4<T>*>(s_w[wslot(X)]);
    const W4<T>* wm = reinterpret_cast<const W4<T>*>(s_w[wslot(X - 1)]);
    const T* rr = s_r[rslot(X)];
#pragma unroll
    for (int k = 0; k < (CNT + 255) / 256; ++k) {
      const int q = tid + 256 * k;
      if (P == 3 ? !l3 : q >= CNT) break;
      const int py = P == 3 ? py3 : q / HZ;
      const int hz = P == 3 ? hz3 : q - (q / HZ) * HZ;
      const int gy = y0 - R + py;
      const int pz = 2 * hz + ((colour + X + gy + z0 + R) & 1);
      const int wi = (py + 4 - R) * WZ + (pz + 4 - R);
      const int ri = (py + 3 - R) * RZ + (pz + 3 - R);
      const W4<T> wc = w0[wi];
      const T rv = rr[ri];
      T ev[6], wn[3];
      if constexpr (!FROM_ZERO) {
        const int pi = (py + 1) * QZ + (pz + 1);
        ev[0] = pxp[pi];
        ev[1] = pxm[pi];
        ev[2] = pxc[pi + QZ];
        ev[3] = pxc[pi - QZ];
        ev[4] = pxc[pi + 1];
        ev[5] = pxc[pi - 1];
        if constexpr (P != 3) wn[0] = wm[wi].wx;
        wn[1] = w0[wi - WZ].wy;
        wn[2] = w0[wi - 1].wz;
      }
      __builtin_amdgcn_sched_barrier(0);
      const float idg = wf(wc.id);
      float acc = wf(rv);
      if constexpr (!FROM_ZERO) {
        acc += wf(wc.wx) * wf(ev[0]);
        acc += (P == 3 ? wxm3 : wf(wn[0])) * wf(ev[1]);
        acc += wf(wc.wy) * wf(ev[2]);
        acc += wf(wn[1]) * wf(ev[3]);
        acc += wf(wc.wz) * wf(ev[4]);
        acc += wf(wn[2]) * wf(ev[5]);
      }
      const float val = idg > 0.0f ? f32_rounded(acc * idg) : 0.0f;
      if constexpr (P != 3) {
        out[py * PZ + pz] = to_store<T>(val);
      } else {
        // the tile's cell pair (z even, z odd): this pass's cell and pass 2's
        // partner beside it
        const int gz = z0 + pz;
        const float pv = wf(pxc[(py + 1) * QZ + ((pz ^ 1) + 1)]);
        T* eo = eout + (long long)X * plane;
        if (pairs) {
          if (gy < ny && gz < nz)
            gst_pair(eo, (size_t)(gy * nz + (gz & ~1)), (gz & 1) ? pv : val, (gz & 1) ? val : pv);
        } else if (gy < ny) {
          if (gz < nz) Store<T>::gst(eo, (size_t)(gy * nz + gz), val);
          if ((gz ^ 1) < nz) Store<T>::gst(eo, (size_t)(gy * nz + (gz ^ 1)), pv);
        }
      }
    }
  };

  const int xend = x1 + 3;  // last step
#pragma unroll
  for (int d = 0; d < D; ++d) fetch(x0 - 4 + d, d);
  auto step = [&](int L, int d) {
    const bool pin = L >= 0 && L < nx;
    // ---- commit plane L ----
    {
      WB* sw = s_w[wslot(L)];
#pragma unroll
      for (int k = 0; k < NW; ++k) {
        const int q = tid + 256 * k;
        if (k + 1 < NW || q < WN) sw[q] = bits_unless(pw[d][k], pin && ow[k] >= 0);
      }
      if constexpr (!ZERO) {
        // air cells stage a zero correction whatever the buffer holds there
        T* se = s_e[tslot(L)];
#pragma unroll
        for (int k = 0; k < NE; ++k) {
          const int q = tid + 256 * k;
          const W4<T> wq = __builtin_bit_cast(W4<T>, pw[d][k]);
          if (k + 1 < NE || q < WN) se[q] = value_unless(pe[d][k], pin && ow[k] >= 0 && wf(wq.id) > 0.0f);
        }
      }
      T* sr = s_r[rslot(L)];
#pragma unroll
      for (int k = 0; k < NR; ++k) {
        const int q = tid + 256 * k;
        const float v = f32_rounded(Store<TR>::ld(&pr[d][k], 0) * rm);
        if (k + 1 < NR || q < RN) sr[q] = to_store<T>((pin && orr[k] >= 0) ? v : 0.0f);
      }
    }
    lds_barrier();  // A
    if (L + D <= xend) fetch(L + D, d);
    // pass p covers planes x0 - (3 - p) .. x1 - 1 + (3 - p)
    const int X0 = L - 1, X1 = L - 2, X2 = L - 3, X3 = L - 4;
    if (X0 >= x0 - 3 && X0 <= x1 + 2) {
      if constexpr (ZERO)
        pass(PassNo<0>{}, X0, nullptr, nullptr, nullptr, s_p0[tslot(X0)]);
      else
        pass(PassNo<0>{}, X0, s_e[tslot(X0 + 1)], s_e[tslot(X0)], s_e[tslot(X0 - 1)], s_p0[tslot(X0)]);
    }
    lds_barrier();  // B
    if (X1 >= x0 - 2 && X1 <= x1 + 1)
      pass(PassNo<1>{}, X1, s_p0[tslot(X1 + 1)], s_p0[tslot(X1)], s_p0[tslot(X1 - 1)], s_p1[tslot(X1)]);
    lds_barrier();  // C
    if (X2 >= x0 - 1 && X2 <= x1)
      pass(PassNo<2>{}, X2, s_p1[tslot(X2 + 1)], s_p1[tslot(X2)], s_p1[tslot(X2 - 1)], s_p2[tslot(X2)]);
    lds_barrier();  // D
    if (X3 >= x0 && X3 < x1)
      pass(PassNo<3>{}, X3, s_p2[tslot(X3 + 1)], s_p2[tslot(X3)], s_p2[tslot(X3 - 1)], nullptr);
    // pass 3's -x coefficient for the next step: this plane's +x face weight
    // of the cell the lane takes at plane X3 + 1 (the other parity)
    if (l3 && X3 + 1 >= x0 && X3 + 1 < x1) {
      const int colour = (first + 3) & 1;
      const int gy = y0 + py3;
      const int pz = 2 * hz3 + ((colour + X3 + 1 + gy + z0) & 1);
      wxm3 = wf(reinterpret_cast<const W4<T>*>(s_w[wslot(X3)])[(py3 + 4) * WZ + (pz + 4)].wx);
    }
  };
  for (int L0 = x0 - 4; L0 <= xend; L0 += D) {
#pragma unroll
    for (int d = 0; d < D; ++d)
      if (L0 + d <= xend) step(L0 + d, d);
  }
}

// One launch per pair of sweeps: workgroup b is the table's workgroup b.
#ifndef FP_PAIR_DEPTH
#define FP_PAIR_DEPTH 1
#endif
constexpr int kPairDepth = FP_PAIR_DEPTH;  // planes of loads in flight per workgroup

template <int TZ, typename TR, typename T, bool ZERO>
__global__ __launch_bounds__(256) void k_sweep2(const Part* parts, int np, int which_in, int first) {
  sweep2_body<TZ, kPairDepth, TR, T, ZERO>(parts, np, (int)blockIdx.x, which_in, first);
}

template <int TZ, typename TR, typename T, bool ZERO>
__global__ __launch_bounds__(256) void k_sweep2_persistent(const Part* parts, int np, int total, int which_in,
                                                           int first) {
  for (int vb = (int)blockIdx.x; vb < total; vb += (int)gridDim.x) {
    sweep2_body<TZ, kPairDepth, TR, T, ZERO>(parts, np, vb, which_in, first);
    __syncthreads();  // the next table slot reuses the LDS planes
  }
}

}  // namespace fp

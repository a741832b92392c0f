// fp_uniform.hip -- uniform MAC-grid pressure projection on gfx950, batched.
//
// One projection = divergence/rhs -> FP64 iterative refinement around an
// FP32 multigrid-preconditioned CG (multigrid vectors in FP32, FP16 or BF16)
// -> pressure-gradient subtract.  All arrays are C order, z fastest.  Every
// kernel runs over a part table (fp_uniform.hpp), so one launch per stage
// serves a whole batch of independent grids of different sizes.
//
// Reference algebra restated here (the kernels are written for CDNA4, not
// translated from WGSL):
//   rhs ............ mgBuildFinestRhs + divergenceAt
//                    (webgpu-uniform-pressure-multigrid.wgsl.ts:152-167,
//                     webgpu-uniform-reference.wgsl.ts:1010-1027)
//   A p ............ mgApply (multigrid.wgsl.ts:127-133) with the Neumann-wall
//                    rows of sparse-atlas-composite-projection.ts:12-13,1171-1192
//   coarse phi ..... mgDownsampleTopology (multigrid.wgsl.ts:169-195)
//   smoother ....... mgSmoothColour without the LCP clamp (multigrid.wgsl.ts:325-344)
//   transfers ...... cell-centred trilinear prolongation (multigrid.wgsl.ts:217-241)
//                    and its transpose as restriction (keeps M symmetric)
//   CG control ..... WebGPUOctreePipelinedMGPCG.encode (webgpu-octree-pipelined-mgpcg.ts:973-1166):
//                    device-side alpha/beta, merged reductions, early exit
//   subtract ....... project (webgpu-uniform-reference.wgsl.ts:1052-1099)
#include "fp_uniform.hpp"
#include "fp_sweep.hpp"
#include "fp_sweep2.hpp"
#include "fp_march.hpp"
#include "fp_state.hpp"

#include <algorithm>
#include <type_traits>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

namespace fp {

// --------------------------------------------------------------------------
// index helpers
// --------------------------------------------------------------------------
struct Grid3 {
  int nx, ny, nz;
  __device__ __forceinline__ size_t n() const { return (size_t)nx * ny * nz; }
  __device__ __forceinline__ void ijk(size_t idx, int& i, int& j, int& k) const {
    k = (int)(idx % (size_t)nz);
    const size_t t = idx / (size_t)nz;
    j = (int)(t % (size_t)ny);
    i = (int)(t / (size_t)ny);
  }
};

struct LevelArgs {
  Grid3 g;
  float gx, gy, gz;  // per-axis coefficient factors (h0/h_l,a)^2
  const float* phi;
};

__device__ __forceinline__ Grid3 grid_of(const LevelRec& l) { return Grid3{l.d[0], l.d[1], l.d[2]}; }
__device__ __forceinline__ LevelArgs largs_of(const LevelRec& l) {
  return LevelArgs{grid_of(l), l.g[0], l.g[1], l.g[2], l.phi};
}

// 4-wide loads of the storage types (arrays are 256-byte aligned).
template <typename T> struct V4;
template <> struct V4<float> {
  static __device__ __forceinline__ float4 ld(const float* p, size_t i4) { return reinterpret_cast<const float4*>(p)[i4]; }
};
template <> struct V4<__half> {
  static __device__ __forceinline__ float4 ld(const __half* p, size_t i4) {
    const uint2 u = reinterpret_cast<const uint2*>(p)[i4];
    __half2 a, b;
    __builtin_memcpy(&a, &u.x, 4);
    __builtin_memcpy(&b, &u.y, 4);
    const float2 fa = __half22float2(a), fb = __half22float2(b);
    return make_float4(fa.x, fa.y, fb.x, fb.y);
  }
};
template <> struct V4<__hip_bfloat16> {
  static __device__ __forceinline__ float4 ld(const __hip_bfloat16* p, size_t i4) {
    const uint2 u = reinterpret_cast<const uint2*>(p)[i4];
    return make_float4(__uint_as_float(u.x << 16), __uint_as_float(u.x & 0xffff0000u), __uint_as_float(u.y << 16),
                       __uint_as_float(u.y & 0xffff0000u));
  }
};

template <typename T>
__device__ __forceinline__ float ldT(const T* p, size_t i) { return Store<T>::ld(p, i); }
template <typename T>
__device__ __forceinline__ void stT(T* p, size_t i, float v) { Store<T>::st(p, i, v); }

// which: 0 x (FP64), 1 e, 2 d[0].  skip_converged: leave grids whose
// refinement has converged.
// which: 0 x (unless warm), 1 e, 2 d[0]; over the finest level's liquid hulls
// (every reader of these vectors stays on liquid cells or inside the hulls).
__global__ __launch_bounds__(256) void k_zero(const Part* parts, int np, int which, int skip_converged) {
  FP_PART(parts, np);
  const State* st = cload(&Gp->st);
  if (skip_converged && cload(&st->outer_done)) return;
  if (which == 0 && cload(&Gp->warm)) return;
  const LevelRec l0 = cload(&Gp->lv[0]);
  const TileSpan sp = tile_span(l0, blk, nblk);
  if (!sp.inside) return;
  const size_t plane = (size_t)l0.d[1] * l0.d[2];
  if (which == 0) {
    double* __restrict__ x = cload(&Gp->x);
    for (int px = sp.x0; px < sp.x1; ++px) stg(x, (size_t)px * plane + sp.col, 0.0);
    return;
  }
  float* __restrict__ p = which == 1 ? cload(&Gp->e) : cload(&Gp->d[0]);
  for (int px = sp.x0; px < sp.x1; ++px) stg(p, (size_t)px * plane + sp.col, 0.0f);
}

// --------------------------------------------------------------------------
// rhs: b = -(rho h/dt) sum(u_out - u_in) + (rho h^2/dt) q on liquid cells.
// Also ||b||^2 and max |div u| (1/s) before the projection.
// --------------------------------------------------------------------------
__global__ void k_rhs(const Part* parts, int np) {
  FP_PART(parts, np);
  const LevelRec& l0 = G.lv[0];
  const Grid3 g = grid_of(l0);
  const float* __restrict__ phi = l0.phi;
  const float* __restrict__ u = G.in_u;
  const float* __restrict__ v = G.in_v;
  const float* __restrict__ w = G.in_w;
  const float* __restrict__ q = G.in_q;
  double* __restrict__ b = G.b;
  const double s_flux = G.s_flux, s_q = G.s_q, inv_h = G.inv_h;
  double acc[3] = {0.0, 0.0, 0.0};
  float dmax = 0.0f;
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  FP_STRIDE(idx, g.n()) {
    int i, j, k;
    g.ijk(idx, i, j, k);
    double val = 0.0;
    if (phi[idx] < 0.0f) {
      acc[1] += 1.0;
      const double up = (i + 1 < nx) ? (double)u[((size_t)(i + 1) * ny + j) * nz + k] : 0.0;
      const double um = (i > 0) ? (double)u[((size_t)i * ny + j) * nz + k] : 0.0;
      const double vp = (j + 1 < ny) ? (double)v[((size_t)i * (ny + 1) + j + 1) * nz + k] : 0.0;
      const double vm = (j > 0) ? (double)v[((size_t)i * (ny + 1) + j) * nz + k] : 0.0;
      const double wp = (k + 1 < nz) ? (double)w[((size_t)i * ny + j) * (nz + 1) + k + 1] : 0.0;
      const double wm = (k > 0) ? (double)w[((size_t)i * ny + j) * (nz + 1) + k] : 0.0;
      const double flux = (up - um) + (vp - vm) + (wp - wm);
      const double qq = q ? (double)q[idx] : 0.0;
      val = -s_flux * flux + s_q * qq;
      dmax = fmaxf(dmax, (float)fabs(flux * inv_h - qq));
      acc[0] += val * val;
      acc[2] += val;
    }
    b[idx] = val;
  }
  block_partials_max<3>(acc, dmax, G.partials, blk, nblk);
}

// x += e; e = 0 (folds one inner solve into the FP64 iterate).  A broken CG
// recurrence leaves the FP64 iterate as it was (fail closed).
__global__ __launch_bounds__(256) void k_accumulate(const Part* parts, int np) {
  FP_PART(parts, np);
  const State* st = cload(&Gp->st);
  if (cload(&st->outer_done)) return;
  const bool keep = !(cload(&st->flags) & 0x7);
  double* __restrict__ x = cload(&Gp->x);
  float* __restrict__ e = cload(&Gp->e);
  const LevelRec l0 = cload(&Gp->lv[0]);
  const TileSpan sp = tile_span(l0, blk, nblk);
  const size_t plane = (size_t)l0.d[1] * l0.d[2];
  if (!sp.inside) return;
  for (int px = sp.x0; px < sp.x1; ++px) {
    const size_t i = (size_t)px * plane + sp.col;
    if (keep) stg(x, i, ldg(x, i) + (double)ldg(e, i));
    stg(e, i, 0.0f);
  }
}

// --------------------------------------------------------------------------
// multigrid setup: coarse level set (CM11a Eq. 15-16 with C preserved levels)
// --------------------------------------------------------------------------
// FP32 sum of eight corner terms in the pairing of mgD4Sum8
// (lib/methods/uniform/webgpu-uniform-pressure-multigrid.wgsl.ts:50-54); the
// additions are rounded one by one (no contraction), so the CPU restatement
// reproduces every bit.
__device__ __forceinline__ float d4_sum8(const float* t) {
  const float y0 = __fadd_rn(__fadd_rn(t[0], t[5]), __fadd_rn(t[1], t[4]));
  const float y1 = __fadd_rn(__fadd_rn(t[2], t[7]), __fadd_rn(t[3], t[6]));
  return __fadd_rn(y0, y1);
}

// Coarse level set of mgDownsampleTopology (…multigrid.wgsl.ts:169-195): the
// mean of the eight children, or the mean of the non-negative ones when their
// signs are mixed and the destination is one of the C finest coarse levels.
__global__ void k_downsample_phi(const Part* parts, int np, int level, int preserve) {
  FP_PART(parts, np);
  const LevelRec& fl = G.lv[level - 1];
  const LevelRec& cl = G.lv[level];
  const Grid3 f = grid_of(fl), c = grid_of(cl);
  const float* __restrict__ pf = fl.phi;
  float* __restrict__ pc = cl.phi;
  const int sx = cl.s[0], sy = cl.s[1], sz = cl.s[2];
  const int use_positive = level <= preserve;
  FP_STRIDE(idx, c.n()) {
    int I, J, K;
    c.ijk(idx, I, J, K);
    // the eight corner children (bit 0: x, 1: y, 2: z); an axis that did not
    // coarsen (stride 1) and the last cell of an odd axis repeat one child
    float v[8];
#pragma unroll
    for (int o = 0; o < 8; ++o) {
      const int i = min(I * sx + (o & 1) * (sx - 1), f.nx - 1);
      const int j = min(J * sy + ((o >> 1) & 1) * (sy - 1), f.ny - 1);
      const int k = min(K * sz + ((o >> 2) & 1) * (sz - 1), f.nz - 1);
      v[o] = pf[((size_t)i * f.ny + j) * f.nz + k];
    }
    float ps[8], pn[8], ng[8];
#pragma unroll
    for (int o = 0; o < 8; ++o) {
      ps[o] = v[o] >= 0.0f ? v[o] : 0.0f;
      pn[o] = v[o] >= 0.0f ? 1.0f : 0.0f;
      ng[o] = v[o] >= 0.0f ? 0.0f : 1.0f;
    }
    const float sum = d4_sum8(v), pos = d4_sum8(ps), npos = d4_sum8(pn), nneg = d4_sum8(ng);
    const bool mixed = npos > 0.0f && nneg > 0.0f;
    pc[idx] = (mixed && use_positive) ? __fdiv_rn(pos, fmaxf(npos, 1.0f)) : __fdiv_rn(sum, 8.0f);
  }
}

// Dense inverse of the coarsest level's liquid block (one workgroup, FP64
// Gauss-Jordan in place; SPD so no pivoting).  A component of the block with
// no air face (an all-liquid level) is singular, its last pivot vanishes to
// rounding: that pivot gets its row's original diagonal added, which inverts
// A + a_pp e_p e_p^T instead (SPD, exact on the range of A).  A negative or
// non-finite pivot falls back to the diagonal (Jacobi) map so the
// preconditioner stays a fixed SPD map.
constexpr int kCoarseMax = 512;
__device__ void coarse_setup(LevelArgs L, int* __restrict__ cmap, int* __restrict__ ncount,
                             double* __restrict__ a, int max_n, State* st) {
  __shared__ int s_n;
  __shared__ int s_fail;
  __shared__ double s_diag[kCoarseMax];
  const Grid3 g = L.g;
  const size_t N = g.n();
  if (threadIdx.x == 0) {
    int n = 0;
    for (size_t idx = 0; idx < N; ++idx) {
      const float pc = L.phi[idx];
      if (!(pc < 0.0f)) continue;
      int i, j, k;
      g.ijk(idx, i, j, k);
      float diag = 0.0f;
      if (i > 0) diag += face_weight(pc, L.phi[idx - (size_t)g.ny * g.nz], L.gx);
      if (i + 1 < g.nx) diag += face_weight(pc, L.phi[idx + (size_t)g.ny * g.nz], L.gx);
      if (j > 0) diag += face_weight(pc, L.phi[idx - g.nz], L.gy);
      if (j + 1 < g.ny) diag += face_weight(pc, L.phi[idx + g.nz], L.gy);
      if (k > 0) diag += face_weight(pc, L.phi[idx - 1], L.gz);
      if (k + 1 < g.nz) diag += face_weight(pc, L.phi[idx + 1], L.gz);
      if (diag > 0.0f && n < max_n) cmap[n++] = (int)idx;
    }
    s_n = n;
    s_fail = 0;
    *ncount = n;
  }
  __syncthreads();
  const int n = s_n;
  // assemble
  for (int e = threadIdx.x; e < n * n; e += blockDim.x) {
    const int r = e / n, c = e % n;
    const size_t ir = (size_t)cmap[r], ic = (size_t)cmap[c];
    int i, j, k, i2, j2, k2;
    g.ijk(ir, i, j, k);
    g.ijk(ic, i2, j2, k2);
    const float pr = L.phi[ir];
    double val = 0.0;
    if (r == c) {
      if (i > 0) val += face_weight(pr, L.phi[ir - (size_t)g.ny * g.nz], L.gx);
      if (i + 1 < g.nx) val += face_weight(pr, L.phi[ir + (size_t)g.ny * g.nz], L.gx);
      if (j > 0) val += face_weight(pr, L.phi[ir - g.nz], L.gy);
      if (j + 1 < g.ny) val += face_weight(pr, L.phi[ir + g.nz], L.gy);
      if (k > 0) val += face_weight(pr, L.phi[ir - 1], L.gz);
      if (k + 1 < g.nz) val += face_weight(pr, L.phi[ir + 1], L.gz);
    } else {
      const int di = abs(i - i2), dj = abs(j - j2), dk = abs(k - k2);
      if (di + dj + dk == 1) {
        const float gg = di ? L.gx : (dj ? L.gy : L.gz);
        val = -(double)face_weight(pr, L.phi[ic], gg);
      }
    }
    a[e] = val;
    if (r == c) s_diag[r] = val;
  }
  __syncthreads();
  for (int p = 0; p < n; ++p) {
    double piv = a[p * n + p];
    if (piv <= 1e-9 * s_diag[p] && piv > -1e-9 * s_diag[p]) piv += s_diag[p];  // singular component
    if (!(piv > 0.0)) {
      if (threadIdx.x == 0) s_fail = 1;
    }
    __syncthreads();
    if (s_fail) break;
    // row p scaled, then eliminate column p from every other row
    const double inv = 1.0 / piv;
    for (int e = threadIdx.x; e < n * n; e += blockDim.x) {
      const int r = e / n, c = e % n;
      if (r == p || c == p) continue;
      a[e] -= a[r * n + p] * a[p * n + c] * inv;
    }
    __syncthreads();
    for (int e = threadIdx.x; e < n; e += blockDim.x) {
      if (e != p) {
        a[e * n + p] = -a[e * n + p] * inv;
        a[p * n + e] = a[p * n + e] * inv;
      }
    }
    __syncthreads();
    if (threadIdx.x == 0) a[p * n + p] = inv;
    __syncthreads();
  }
  if (s_fail) {
    // Jacobi fallback
    for (int e = threadIdx.x; e < n * n; e += blockDim.x) a[e] = 0.0;
    __syncthreads();
    for (int r = threadIdx.x; r < n; r += blockDim.x) {
      const size_t ir = (size_t)cmap[r];
      int i, j, k;
      g.ijk(ir, i, j, k);
      const float pr = L.phi[ir];
      double d = 0.0;
      if (i > 0) d += face_weight(pr, L.phi[ir - (size_t)g.ny * g.nz], L.gx);
      if (i + 1 < g.nx) d += face_weight(pr, L.phi[ir + (size_t)g.ny * g.nz], L.gx);
      if (j > 0) d += face_weight(pr, L.phi[ir - g.nz], L.gy);
      if (j + 1 < g.ny) d += face_weight(pr, L.phi[ir + g.nz], L.gy);
      if (k > 0) d += face_weight(pr, L.phi[ir - 1], L.gz);
      if (k + 1 < g.nz) d += face_weight(pr, L.phi[ir + 1], L.gz);
      a[r * n + r] = 1.0 / d;
    }
    if (threadIdx.x == 0) st->flags |= 0x40;  // FP_FLAG_COARSE_SINGULAR
  }
}

__global__ void k_coarse_setup(const Part* parts) {
  const GridRec& G = *parts[blockIdx.x].g;
  coarse_setup(largs_of(G.lv[G.L - 1]), G.cmap, G.ncoarse, G.minv, G.max_coarse, G.st);
}

// --------------------------------------------------------------------------
// V-cycle kernels.  Stored value = true value / scale; TR is the storage of
// the right-hand side (FP32 at the finest level, where it is the CG residual
// scaled by the device-side rscale), T of everything else.  `level` is the
// same for every grid of a launch; the finest level reads the CG residual.
// --------------------------------------------------------------------------
template <typename TR>
__device__ __forceinline__ const TR* level_rhs(const GridRec& G, int level) {
  return (const TR*)(level == 0 ? (const void*)G.r : G.lv[level].r);
}

// One colour of red-black Gauss-Seidel (colour = (i+j+k)&1) at cell idx of a
// level too small for the fused sweep.  zero_init: the first pass from a zero
// guess, which also clears the other colour.  rm: stored-rhs to stored-e scale.
template <typename TR, typename T>
__device__ __forceinline__ void smooth_cell(const GridRec& G, int level, int colour, int zero_init, float rm,
                                            size_t idx) {
  const LevelRec& lv = G.lv[level];
  const LevelArgs L = largs_of(lv);
  const TR* __restrict__ r = level_rhs<TR>(G, level);
  T* __restrict__ e = (T*)lv.e;
  const Grid3 g = L.g;
  const size_t sx = (size_t)g.ny * g.nz, sy = (size_t)g.nz;
  int i, j, k;
  g.ijk(idx, i, j, k);
  // every load first (neighbours past the edge clamped to the cell itself),
  // then the arithmetic: one round trip per pass
  const bool has[6] = {i > 0, i + 1 < g.nx, j > 0, j + 1 < g.ny, k > 0, k + 1 < g.nz};
  const ptrdiff_t off[6] = {-(ptrdiff_t)sx, (ptrdiff_t)sx, -(ptrdiff_t)sy, (ptrdiff_t)sy, -1, 1};
  const float gw[3] = {L.gx, L.gy, L.gz};
  const float pc = L.phi[idx];
  const float rv = ldT(r, idx);
  float pn[6], en[6];
#pragma unroll
  for (int q = 0; q < 6; ++q) {
    const size_t m = has[q] ? (size_t)((ptrdiff_t)idx + off[q]) : idx;
    pn[q] = L.phi[m];
    en[q] = ldT(e, m);
  }
  if (((i + j + k) & 1) != colour) {
    if (zero_init) stT(e, idx, 0.0f);
    return;
  }
  if (!(pc < 0.0f)) {
    stT(e, idx, 0.0f);
    return;
  }
  float diag = 0.0f, acc = rv * rm;
#pragma unroll
  for (int q = 0; q < 6; ++q) {
    const float wf = has[q] ? face_weight(pc, pn[q], gw[q >> 1]) : 0.0f;
    diag += wf;
    acc += (!zero_init && has[q] && pn[q] < 0.0f) ? wf * en[q] : 0.0f;
  }
  stT(e, idx, diag > 0.0f ? acc / diag : 0.0f);
}

template <typename TR, typename T>
__global__ void k_smooth(const Part* parts, int np, int level, int colour, int zero_init) {
  FP_PART(parts, np);
  const State* st = G.st;
  if (st->done) return;
  const LevelRec& lv = G.lv[level];
  const float rmul = (level == 0 ? 1.0f : lv.Sr) / lv.Se;
  const float rm = level == 0 ? rmul * st->rscale : rmul;
  FP_STRIDE(idx, (size_t)lv.n) smooth_cell<TR, T>(G, level, colour, zero_init, rm, idx);
}

// t = r - A e at cell idx of a level too small for the fused march (stored
// with St); rm: stored-rhs to stored-t scale.
template <typename TR, typename T>
__device__ __forceinline__ void residual_cell(const GridRec& G, int level, float rm, size_t idx) {
  const LevelRec& lv = G.lv[level];
  const LevelArgs L = largs_of(lv);
  const TR* __restrict__ r = level_rhs<TR>(G, level);
  const T* __restrict__ e = (const T*)lv.e;
  T* __restrict__ t = (T*)lv.t;
  const float emul = lv.Se / lv.St;
  const Grid3 g = L.g;
  const size_t sx = (size_t)g.ny * g.nz, sy = (size_t)g.nz;
  int i, j, k;
  g.ijk(idx, i, j, k);
  const bool has[6] = {i > 0, i + 1 < g.nx, j > 0, j + 1 < g.ny, k > 0, k + 1 < g.nz};
  const ptrdiff_t off[6] = {-(ptrdiff_t)sx, (ptrdiff_t)sx, -(ptrdiff_t)sy, (ptrdiff_t)sy, -1, 1};
  const float gw[3] = {L.gx, L.gy, L.gz};
  const float pc = L.phi[idx];
  const float ec = ldT(e, idx), rv = ldT(r, idx);
  float pn[6], en[6];
#pragma unroll
  for (int q = 0; q < 6; ++q) {
    const size_t m = has[q] ? (size_t)((ptrdiff_t)idx + off[q]) : idx;
    pn[q] = L.phi[m];
    en[q] = ldT(e, m);
  }
  float res = 0.0f;
  if (pc < 0.0f) {
    float ap = 0.0f;
#pragma unroll
    for (int q = 0; q < 6; ++q) {
      const float wf = has[q] ? face_weight(pc, pn[q], gw[q >> 1]) : 0.0f;
      ap += wf * (ec - (pn[q] < 0.0f ? en[q] : 0.0f));
    }
    res = rv * rm - emul * ap;
  }
  stT(t, idx, res);
}

template <typename TR, typename T>
__global__ void k_residual(const Part* parts, int np, int level) {
  FP_PART(parts, np);
  const State* st = G.st;
  if (st->done) return;
  const LevelRec& lv = G.lv[level];
  const float rmul = (level == 0 ? 1.0f : lv.Sr) / lv.St;
  const float rm = level == 0 ? rmul * st->rscale : rmul;
  FP_STRIDE(idx, (size_t)lv.n) residual_cell<TR, T>(G, level, rm, idx);
}

// The restriction at one coarse cell (the gather form of k_restrict_march, for
// the one-workgroup bottom of the cycle): per axis the fine taps 2I-1 .. 2I+2
// with weights 1/4 3/4 3/4 1/4 (1 at a domain edge), or the cell itself on an
// axis that is not coarsened.
__device__ __forceinline__ int restrict_taps(int I, int s, int nf, int nc, int* f, float* w) {
  if (s == 1) {
    f[0] = I;
    w[0] = 1.0f;
    return 1;
  }
  int c = 0;
  if (I >= 1) { f[c] = 2 * I - 1; w[c] = 0.25f; ++c; }
  f[c] = 2 * I; w[c] = I >= 1 ? 0.75f : 1.0f; ++c;
  if (2 * I + 1 < nf) { f[c] = 2 * I + 1; w[c] = I + 1 >= nc ? 1.0f : 0.75f; ++c; }
  if (2 * I + 2 < nf) { f[c] = 2 * I + 2; w[c] = 0.25f; ++c; }
  return c;
}

template <typename T>
__device__ __forceinline__ void restrict_cell(const GridRec& G, int level, size_t ci) {
  const LevelRec& fl = G.lv[level];
  const LevelRec& cl = G.lv[level + 1];
  const Grid3 f = grid_of(fl), c = grid_of(cl);
  const int sx = cl.s[0], sy = cl.s[1], sz = cl.s[2];
  const int k2 = (sx == 2) + (sy == 2) + (sz == 2);
  const float fac = (1.0f / (float)(1 << k2)) * fl.St / cl.Sr;
  const T* __restrict__ tf = (const T*)fl.t;
  float acc = 0.0f;
  if (cl.phi[ci] < 0.0f) {
    int I, J, K;
    c.ijk(ci, I, J, K);
    int fi[4], fj[4], fk[4];
    float wi[4], wj[4], wk[4];
    const int ni = restrict_taps(I, sx, f.nx, c.nx, fi, wi);
    const int nj = restrict_taps(J, sy, f.ny, c.ny, fj, wj);
    const int nk = restrict_taps(K, sz, f.nz, c.nz, fk, wk);
    for (int a = 0; a < ni; ++a)
      for (int b = 0; b < nj; ++b) {
        const float wab = wi[a] * wj[b];
        const size_t base = ((size_t)fi[a] * f.ny + fj[b]) * f.nz;
        for (int q = 0; q < nk; ++q) acc += wab * wk[q] * ldT(tf, base + fk[q]);
      }
    acc *= fac;
  }
  stT((T*)cl.r, ci, acc);
}

// r_{l+1} = R t_l on the coarse level's liquid cells (R = P^T / 2^k, the
// transpose of the cell-centred trilinear prolongation: per axis the fine
// taps 2I-1 .. 2I+2 with weights 1/4 3/4 3/4 1/4, 1 at a domain edge), as a
// column-tile x-march over the coarse level: a 256-lane workgroup owns a 32 x 8 (z x y) tile of coarse columns (the coarse
// level's march geometry) and walks its chunk of coarse planes.  Each fine
// plane the chunk touches is staged once in LDS (the tile's fine footprint
// plus the one-cell ring of the 4-tap stencil), filtered in y and z per lane,
// and added with its x weight into the two coarse planes it feeds; a coarse
// plane is written once its last fine plane has been seen.  Fine values are
// read once per tile instead of once per coarse tap.  Every reader of r uses
// it on liquid cells only, so on levels with liquid hulls the march covers the
// coarse tile's hull and leaves the other planes unwritten.
template <typename T>
__global__ __launch_bounds__(256) void k_restrict_march(const Part* parts, int np, int level) {
  FP_PART(parts, np);
  if (cload(&cload(&Gp->st)->done)) return;
  const LevelRec fl = cload(&Gp->lv[level]);
  const LevelRec cl = cload(&Gp->lv[level + 1]);
  constexpr int TZ = kMarchTZ, TY = 256 / TZ;
  constexpr int RY = 2 * TY + 2, RZ = 2 * TZ + 2;
  constexpr int NC = (RY * RZ + 255) / 256;
  __shared__ T s[2][RY * RZ];
  const int fnx = fl.d[0], fny = fl.d[1], fnz = fl.d[2];
  const int cnx = cl.d[0], cny = cl.d[1], cnz = cl.d[2];
  const int sx = cl.s[0], sy = cl.s[1], sz = cl.s[2];
  const int k2 = (sx == 2) + (sy == 2) + (sz == 2);
  const float fac = (1.0f / (float)(1 << k2)) * fl.St / cl.Sr;
  const T* __restrict__ tf = (const T*)fl.t;
  T* __restrict__ rc = (T*)cl.r;
  const float* __restrict__ phic = cl.phi;
  int tzi, tyi, cxi;
  tile_of_group(blk, cl.rg.tz, cl.rg.ty, nblk, tzi, tyi, cxi);
  const int cz0 = tzi * TZ, cy0 = tyi * TY;
  int X0 = cxi * cl.rg.chunk, X1 = min(X0 + cl.rg.chunk, cnx);
  if (cl.hull) {
    // r is read on the coarse level's liquid cells only: march over the
    // coarse tile's liquid hull
    const int* hp = cl.hull + 2 * (tyi * cl.mg.tz + tzi);
    X0 = max(X0, cload(hp));
    X1 = min(X1, cload(hp + 1) + 1);
    if (X0 >= X1) return;
  }
  const int tid = threadIdx.x;
  // staged fine window: rows fy0 .. fy0+ry-1, columns fz0 .. fz0+rz-1
  const int fy0 = sy == 2 ? 2 * cy0 - 1 : cy0, fz0 = sz == 2 ? 2 * cz0 - 1 : cz0;
  const int ry = sy == 2 ? RY : TY, rz = sz == 2 ? RZ : TZ;
  // off: the staged element's offset in a fine plane (-1: outside the
  // grid); [hlo, hhi]: the planes of the liquid hull of its fine march tile
  // (the residual march leaves t unwritten outside it, where it is zero)
  int off[NC], lds[NC], hlo[NC], hhi[NC];
#pragma unroll
  for (int k = 0; k < NC; ++k) {
    const int q = tid + 256 * k;
    const int a = q / rz, b = q - (q / rz) * rz;
    const int yy = fy0 + a, zz = fz0 + b;
    lds[k] = q < ry * rz ? a * RZ + b : -1;
    off[k] = (q < ry * rz && yy >= 0 && yy < fny && zz >= 0 && zz < fnz) ? yy * fnz + zz : -1;
    hlo[k] = 0;
    hhi[k] = fnx - 1;
    if (fl.hull && off[k] >= 0) {
      // per-lane values: vector loads (cload would make them wave-uniform)
      const size_t h = 2 * (size_t)((yy / kHullTY) * fl.mg.tz + zz / kHullTZ);
      hlo[k] = ldg((const int*)fl.hull, h);
      hhi[k] = ldg((const int*)fl.hull, h + 1);
    }
  }
  // this lane's coarse column and its 4 + 4 in-plane taps (local rows /
  // columns of the window; absent taps weigh 0)
  const int ly = tid / TZ, lz = tid - (tid / TZ) * TZ;
  const int J = cy0 + ly, K = cz0 + lz;
  int ty_[4], tz_[4];
  float wy[4], wz[4];
  auto taps = [&](int I, int s_, int nf, int nc, int l, int* t, float* w) {
    if (s_ == 1) {
      t[0] = t[1] = t[2] = t[3] = l;
      w[0] = 1.0f;
      w[1] = w[2] = w[3] = 0.0f;
      return;
    }
    for (int q = 0; q < 4; ++q) t[q] = 2 * l + q;  // fine 2I-1+q
    w[0] = I >= 1 ? 0.25f : 0.0f;
    w[1] = I >= 1 ? 0.75f : 1.0f;
    w[2] = 2 * I + 1 < nf ? (I + 1 >= nc ? 1.0f : 0.75f) : 0.0f;
    w[3] = 2 * I + 2 < nf ? 0.25f : 0.0f;
  };
  taps(J, sy, fny, cny, ly, ty_, wy);
  taps(K, sz, fnz, cnz, lz, tz_, wz);
  const bool mine = J < cny && K < cnz;
  const long long fplane = (long long)fny * fnz, cplane = (long long)cny * cnz;
  const size_t cidx = (size_t)J * cnz + K;

  const int p0 = sx == 2 ? max(2 * X0 - 1, 0) : X0;
  const int p1 = sx == 2 ? min(2 * X1, fnx - 1) : X1 - 1;  // last fine plane, inclusive
  T pre[NC];
  auto fetch = [&](int p) {
    const T* src = tf + (long long)p * fplane;
#pragma unroll
    for (int k = 0; k < NC; ++k)
      pre[k] = (off[k] >= 0 && p >= hlo[k] && p <= hhi[k]) ? ldg(src, (size_t)off[k]) : to_store<T>(0.0f);
  };
  fetch(p0);
  int cur = X0;
  float acc0 = 0.0f, acc1 = 0.0f;
  // level set of the coarse cell being accumulated, loaded when its plane
  // starts rather than when it is written
  auto coarse_phi = [&](int c) { return (mine && c < X1) ? ldg(phic, (size_t)((long long)c * cplane) + cidx) : 1.0f; };
  float phc = coarse_phi(cur);
  for (int p = p0; p <= p1; ++p) {
    T* sl = s[p & 1];
#pragma unroll
    for (int k = 0; k < NC; ++k)
      if (lds[k] >= 0) sl[lds[k]] = pre[k];
    lds_barrier();
    if (p < p1) fetch(p + 1);
    float v = 0.0f;
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      const T* row = sl + ty_[b] * RZ;
      const float rowsum = wz[0] * Store<T>::ld(row, tz_[0]) + wz[1] * Store<T>::ld(row, tz_[1]) +
                           wz[2] * Store<T>::ld(row, tz_[2]) + wz[3] * Store<T>::ld(row, tz_[3]);
      v += wy[b] * rowsum;
    }
    // x weights: plane p feeds coarse p>>1 (and its neighbour on p's side)
    if (sx == 1) {
      acc0 += v;
    } else {
      const int I = p >> 1, Jx = (p & 1) ? I + 1 : I - 1;
      const bool two = Jx >= 0 && Jx < cnx;
      const float wa = two ? 0.75f : 1.0f;
      if (I == cur) acc0 += wa * v;
      else if (I == cur + 1) acc1 += wa * v;
      if (two) {
        if (Jx == cur) acc0 += 0.25f * v;
        else if (Jx == cur + 1) acc1 += 0.25f * v;
      }
    }
    // coarse planes whose last fine plane has been staged
    while (cur < X1 && p >= (sx == 2 ? min(2 * cur + 2, fnx - 1) : cur)) {
      if (mine) {
        const size_t o = (size_t)((long long)cur * cplane) + cidx;
        Store<T>::gst(rc, o, phc < 0.0f ? acc0 * fac : 0.0f);
      }
      acc0 = acc1;
      acc1 = 0.0f;
      ++cur;
      phc = coarse_phi(cur);
    }
    // no second barrier: the next plane goes to the other slot, whose last
    // reads came before this plane's barrier
  }
}

// e_l += P e_{l+1} on the fine level's liquid cells; which: the fine buffer.
template <typename T>
__device__ __forceinline__ void prolong_cell(const GridRec& G, int level, int which, size_t idx) {
  const LevelRec& fl = G.lv[level];
  const LevelRec& cl = G.lv[level + 1];
  if (!(fl.phi[idx] < 0.0f)) return;
  const Grid3 f = grid_of(fl), c = grid_of(cl);
  T* __restrict__ ef = (T*)(which ? fl.e2 : fl.e);
  const T* __restrict__ ec = (const T*)(cl.efinal ? cl.e2 : cl.e);
  const float* __restrict__ phic = cl.phi;
  const int sx = cl.s[0], sy = cl.s[1], sz = cl.s[2];
  const float fac = cl.Se / fl.Se;
  int i, j, k;
  f.ijk(idx, i, j, k);
  int ci[2], cj[2], ck[2];
  float wi[2], wj[2], wk[2];
  const int ni = ptaps(i, sx, c.nx, ci, wi);
  const int nj = ptaps(j, sy, c.ny, cj, wj);
  const int nk = ptaps(k, sz, c.nz, ck, wk);
  float acc = 0.0f;
  for (int a = 0; a < ni; ++a)
    for (int b = 0; b < nj; ++b) {
      const float wab = wi[a] * wj[b];
      const size_t base = ((size_t)ci[a] * c.ny + cj[b]) * c.nz;
      for (int q = 0; q < nk; ++q) {
        const size_t m = base + ck[q];
        if (phic[m] < 0.0f) acc += wab * wk[q] * ldT(ec, m);
      }
    }
  stT(ef, idx, ldT(ef, idx) + fac * acc);
}

// e_l += P e_{l+1} on a fused level's liquid cells, over the liquid hulls of
// its march tiles (the other planes are air and keep e = 0): each lane walks its
// cell column along x, and every load of four planes (level set, correction,
// eight coarse taps each, coarse air holding e = 0) is issued before the
// arithmetic.  which: the fine buffer read; which_out: the buffer written
// (the same one: updated in place).
template <typename T>
__global__ __launch_bounds__(256) void k_prolong_march(const Part* parts, int np, int level, int which,
                                                       int which_out) {
  FP_PART(parts, np);
  if (cload(&cload(&Gp->st)->done)) return;
  const LevelRec fl = cload(&Gp->lv[level]);
  const LevelRec cl = cload(&Gp->lv[level + 1]);
  const TileSpan sp = tile_span(fl, blk, nblk);
  if (!sp.inside || sp.x0 >= sp.x1) return;
  const T* __restrict__ ef = (const T*)(which ? fl.e2 : fl.e);
  T* __restrict__ eo = (T*)(which_out ? fl.e2 : fl.e);
  const T* __restrict__ ec = (const T*)(cl.efinal ? cl.e2 : cl.e);
  const float* __restrict__ phi = fl.phi;
  const float fac = cl.Se / fl.Se;
  const int nz = fl.d[2];
  const int gy = (int)(sp.col / (size_t)nz), gz = (int)(sp.col - (size_t)gy * nz);
  const int cny = cl.d[1], cnz = cl.d[2];
  // two taps per axis (a missing second tap repeats the first at weight 0)
  auto taps2 = [](int f, int s_, int nc, int* c, float* w) {
    const int n = ptaps(f, s_, nc, c, w);
    if (n == 1) {
      c[1] = c[0];
      w[1] = 0.0f;
    }
  };
  int cy[2], cz[2];
  float wy[2], wz[2];
  taps2(gy, cl.s[1], cny, cy, wy);
  taps2(gz, cl.s[2], cnz, cz, wz);
  const size_t fplane = (size_t)fl.d[1] * nz, cplane = (size_t)cny * cnz;
  const size_t o00 = (size_t)cy[0] * cnz + cz[0], o01 = (size_t)cy[0] * cnz + cz[1];
  const size_t o10 = (size_t)cy[1] * cnz + cz[0], o11 = (size_t)cy[1] * cnz + cz[1];
  constexpr int U = 4;
  for (int x = sp.x0; x < sp.x1; x += U) {
    float ph[U], ev[U], acc[U];
    T cv[U][8];
    float wx[U][2];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int xx = min(x + u, sp.x1 - 1);
      const size_t i = (size_t)xx * fplane + sp.col;
      ph[u] = ldg(phi, i);
      ev[u] = Store<T>::gld(ef, i);
      int cx[2];
      taps2(xx, cl.s[0], cl.d[0], cx, wx[u]);
      const T* c0 = ec + (size_t)cx[0] * cplane;
      const T* c1 = ec + (size_t)cx[1] * cplane;
      cv[u][0] = ldg(c0, o00);
      cv[u][1] = ldg(c0, o01);
      cv[u][2] = ldg(c0, o10);
      cv[u][3] = ldg(c0, o11);
      cv[u][4] = ldg(c1, o00);
      cv[u][5] = ldg(c1, o01);
      cv[u][6] = ldg(c1, o10);
      cv[u][7] = ldg(c1, o11);
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const float a0 = wy[0] * (wz[0] * wf(cv[u][0]) + wz[1] * wf(cv[u][1])) +
                       wy[1] * (wz[0] * wf(cv[u][2]) + wz[1] * wf(cv[u][3]));
      const float a1 = wy[0] * (wz[0] * wf(cv[u][4]) + wz[1] * wf(cv[u][5])) +
                       wy[1] * (wz[0] * wf(cv[u][6]) + wz[1] * wf(cv[u][7]));
      acc[u] = wx[u][0] * a0 + wx[u][1] * a1;
    }
#pragma unroll
    for (int u = 0; u < U; ++u)
      if (x + u < sp.x1 && ph[u] < 0.0f) Store<T>::gst(eo, (size_t)(x + u) * fplane + sp.col, ev[u] + fac * acc[u]);
  }
}

template <typename T>
__global__ void k_prolong_add(const Part* parts, int np, int level, int which) {
  FP_PART(parts, np);
  if (G.st->done) return;
  FP_STRIDE(idx, (size_t)G.lv[level].n) prolong_cell<T>(G, level, which, idx);
}

// Clears a level's correction buffer (pre-smoothing with zero sweeps).
template <typename T>
__global__ void k_zero_level(const Part* parts, int np, int level) {
  FP_PART(parts, np);
  if (G.st->done) return;
  const LevelRec& lv = G.lv[level];
  T* e = (T*)lv.e;
  FP_STRIDE(i, (size_t)lv.n) stT(e, i, 0.0f);
}

// e = fac * Minv r on each grid's coarsest liquid cells (one workgroup per
// grid).  TR: storage of that level's right-hand side (FP32 when the coarsest
// level is the finest).
template <typename T, typename TR>
__device__ __forceinline__ void coarse_solve_block(const GridRec& G) {
  const State* st = G.st;
  __shared__ double rs[512];
  const int lc = G.L - 1;
  const LevelRec& lv = G.lv[lc];
  const Grid3 g = grid_of(lv);
  const int* __restrict__ cmap = G.cmap;
  const double* __restrict__ minv = G.minv;
  const TR* __restrict__ r = level_rhs<TR>(G, lc);
  T* __restrict__ e = (T*)lv.e;
  const int n = *G.ncoarse;
  const float rm = lc == 0 ? st->rscale : lv.Sr;
  const float fac = 1.0f / lv.Se;
  for (size_t idx = threadIdx.x; idx < g.n(); idx += blockDim.x) stT(e, idx, 0.0f);
  for (int q = threadIdx.x; q < n; q += blockDim.x) rs[q] = (double)(ldT(r, (size_t)cmap[q]) * rm);
  __syncthreads();
  // the FP64 product on the matrix cores
  gemv_f64_mfma(minv, n, rs, [&](int row, double v) { stT(e, (size_t)cmap[row], (float)(v * fac)); });
}

template <typename T, typename TR>
__global__ void k_coarse_solve(const Part* parts) {
  const GridRec& G = *parts[blockIdx.x].g;
  if (G.st->done) return;
  coarse_solve_block<T, TR>(G);
}

// The bottom of the V-cycle in one workgroup per grid: the levels from lb on
// (each under kBottomCells cells, so never fused and never the finest) are
// smoothed, residual-restricted, solved on the coarsest level and prolonged
// with the same per-cell steps as k_smooth / k_residual / k_prolong_add and
// the restriction's weights, separated by workgroup barriers instead of
// launches.
template <typename T>
__global__ __launch_bounds__(kBottomThreads) void k_bottom(const Part* parts, int pre, int post) {
  const GridRec& G = *parts[blockIdx.x].g;
  if (G.st->done) return;
  const int lb = G.lb, lc = G.L - 1;
  for (int l = lb; l < lc; ++l) {
    const LevelRec& lv = G.lv[l];
    const size_t n = (size_t)lv.n;
    const float rs = lv.Sr / lv.Se, rt = lv.Sr / lv.St;
    if (pre == 0) {
      for (size_t i = threadIdx.x; i < n; i += blockDim.x) stT((T*)lv.e, i, 0.0f);
      __syncthreads();
    }
    for (int sw = 0; sw < pre; ++sw)
      for (int colour = 0; colour < 2; ++colour) {
        const int zi = (sw == 0 && colour == 0) ? 1 : 0;
        for (size_t i = threadIdx.x; i < n; i += blockDim.x) smooth_cell<T, T>(G, l, colour, zi, rs, i);
        __syncthreads();
      }
    for (size_t i = threadIdx.x; i < n; i += blockDim.x) residual_cell<T, T>(G, l, rt, i);
    __syncthreads();
    for (size_t ci = threadIdx.x; ci < (size_t)G.lv[l + 1].n; ci += blockDim.x) restrict_cell<T>(G, l, ci);
    __syncthreads();
  }
  coarse_solve_block<T, T>(G);
  __syncthreads();
  for (int l = lc - 1; l >= lb; --l) {
    const LevelRec& lv = G.lv[l];
    const size_t n = (size_t)lv.n;
    const float rs = lv.Sr / lv.Se;
    for (size_t i = threadIdx.x; i < n; i += blockDim.x) prolong_cell<T>(G, l, 0, i);
    __syncthreads();
    for (int sw = 0; sw < post; ++sw)
      for (int cc = 0; cc < 2; ++cc) {
        for (size_t i = threadIdx.x; i < n; i += blockDim.x) smooth_cell<T, T>(G, l, 1 - cc, 0, rs, i);
        __syncthreads();
      }
  }
}

// k_bottom with the bottom levels held in LDS: the level sets of the levels
// from lb on and their correction, right-hand side and residual live in the
// workgroup's dynamic LDS (bottom_lds_bytes), the grid record is copied to
// LDS with those arrays pointing there, and the same per-cell steps run on it
// (they address memory through generic pointers), so every phase is an LDS
// round trip instead of an L2 one.  In: the level sets and level lb's
// restricted right-hand side; out: level lb's correction, the only array of
// these levels read after the bottom.
constexpr size_t kBottomLdsAlign = 16;
__host__ __device__ inline size_t bottom_lds_round(size_t b) { return (b + kBottomLdsAlign - 1) & ~(kBottomLdsAlign - 1); }
template <typename T>
__host__ __device__ inline size_t bottom_lds_bytes_of(const long long* n, int count) {
  size_t b = 0;
  for (int q = 0; q < count; ++q) b += bottom_lds_round(4 * (size_t)n[q]) + 3 * bottom_lds_round(sizeof(T) * (size_t)n[q]);
  return b;
}

template <typename T>
__global__ __launch_bounds__(kBottomThreads) void k_bottom_lds(const Part* parts, int pre, int post) {
  extern __shared__ __align__(16) unsigned char s_bot[];
  __shared__ __align__(16) unsigned char s_rec[sizeof(GridRec)];  // raw: the record's type has initialisers
  GridRec& Gs = *reinterpret_cast<GridRec*>(s_rec);
  const GridRec& Gg = *parts[blockIdx.x].g;
  if (Gg.st->done) return;
  {
    const int* src = reinterpret_cast<const int*>(&Gg);
    int* dst = reinterpret_cast<int*>(&Gs);
    for (int q = threadIdx.x; q < (int)(sizeof(GridRec) / 4); q += blockDim.x) dst[q] = src[q];
  }
  __syncthreads();
  const int lb = Gs.lb, lc = Gs.L - 1;
  if (threadIdx.x == 0) {
    unsigned char* p = s_bot;
    for (int l = lb; l <= lc; ++l) {
      LevelRec& lv = Gs.lv[l];
      const size_t n = (size_t)lv.n;
      lv.phi = reinterpret_cast<float*>(p);
      p += bottom_lds_round(4 * n);
      lv.e = p;
      p += bottom_lds_round(sizeof(T) * n);
      lv.r = p;
      p += bottom_lds_round(sizeof(T) * n);
      lv.t = p;
      p += bottom_lds_round(sizeof(T) * n);
      lv.e2 = lv.e;
    }
  }
  __syncthreads();
  for (int l = lb; l <= lc; ++l) {
    const size_t n = (size_t)Gs.lv[l].n;
    const float* __restrict__ src = Gg.lv[l].phi;
    float* dst = Gs.lv[l].phi;
    for (size_t i = threadIdx.x; i < n; i += blockDim.x) dst[i] = ldg(src, i);
  }
  {
    const size_t n = (size_t)Gs.lv[lb].n;
    const T* __restrict__ src = (const T*)Gg.lv[lb].r;
    T* dst = (T*)Gs.lv[lb].r;
    for (size_t i = threadIdx.x; i < n; i += blockDim.x) dst[i] = ldg(src, i);
  }
  __syncthreads();
  const GridRec& G = Gs;
  for (int l = lb; l < lc; ++l) {
    const LevelRec& lv = G.lv[l];
    const size_t n = (size_t)lv.n;
    const float rs = lv.Sr / lv.Se, rt = lv.Sr / lv.St;
    if (pre == 0) {
      for (size_t i = threadIdx.x; i < n; i += blockDim.x) stT((T*)lv.e, i, 0.0f);
      __syncthreads();
    }
    for (int sw = 0; sw < pre; ++sw)
      for (int colour = 0; colour < 2; ++colour) {
        const int zi = (sw == 0 && colour == 0) ? 1 : 0;
        for (size_t i = threadIdx.x; i < n; i += blockDim.x) smooth_cell<T, T>(G, l, colour, zi, rs, i);
        __syncthreads();
      }
    for (size_t i = threadIdx.x; i < n; i += blockDim.x) residual_cell<T, T>(G, l, rt, i);
    __syncthreads();
    for (size_t ci = threadIdx.x; ci < (size_t)G.lv[l + 1].n; ci += blockDim.x) restrict_cell<T>(G, l, ci);
    __syncthreads();
  }
  coarse_solve_block<T, T>(G);
  __syncthreads();
  for (int l = lc - 1; l >= lb; --l) {
    const LevelRec& lv = G.lv[l];
    const size_t n = (size_t)lv.n;
    const float rs = lv.Sr / lv.Se;
    for (size_t i = threadIdx.x; i < n; i += blockDim.x) prolong_cell<T>(G, l, 0, i);
    __syncthreads();
    for (int sw = 0; sw < post; ++sw)
      for (int cc = 0; cc < 2; ++cc) {
        for (size_t i = threadIdx.x; i < n; i += blockDim.x) smooth_cell<T, T>(G, l, 1 - cc, 0, rs, i);
        __syncthreads();
      }
  }
  {
    const size_t n = (size_t)Gs.lv[lb].n;
    const T* src = (const T*)Gs.lv[lb].e;
    T* __restrict__ dst = (T*)Gg.lv[lb].e;
    for (size_t i = threadIdx.x; i < n; i += blockDim.x) stg(dst, i, src[i]);
  }
}

// Packed coefficients of a fused level.
template <typename T>
__global__ void k_build_weights(const Part* parts, int np, int level) {
  FP_PART(parts, np);
  const LevelRec& lv = G.lv[level];
  const int nx = lv.d[0], ny = lv.d[1], nz = lv.d[2];
  const float gx = lv.g[0], gy = lv.g[1], gz = lv.g[2];
  const float* __restrict__ phi = lv.phi;
  W4<T>* __restrict__ w = (W4<T>*)lv.w;
  const size_t n = (size_t)lv.n;
  const size_t sx = (size_t)ny * nz, sy = (size_t)nz;
  FP_STRIDE(idx, n) {
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

// --------------------------------------------------------------------------
// CG kernels (FP32 vectors, FP64 scalars)
// --------------------------------------------------------------------------
// gamma = r . z (z: the finest level's correction, stored scale Se / rscale).
__device__ __forceinline__ fp_f32x4 ldg4T(const float* p, size_t i4) { return ldg4(p, i4); }
__device__ __forceinline__ fp_f32x4 ldg4T(const __half* p, size_t i4) {
  const fp_u32x2 u = ldg((const fp_u32x2*)p, i4);
  const unsigned ux = u.x, uy = u.y;
  __half2 a, b;
  __builtin_memcpy(&a, &ux, 4);
  __builtin_memcpy(&b, &uy, 4);
  const float2 fa = __half22float2(a), fb = __half22float2(b);
  return fp_f32x4{fa.x, fa.y, fb.x, fb.y};
}
__device__ __forceinline__ fp_f32x4 ldg4T(const __hip_bfloat16* p, size_t i4) {
  const fp_u32x2 u = ldg((const fp_u32x2*)p, i4);
  return fp_f32x4{__uint_as_float(u.x << 16), __uint_as_float(u.x & 0xffff0000u), __uint_as_float(u.y << 16),
                  __uint_as_float(u.y & 0xffff0000u)};
}

// r . z over the finest level's liquid hulls (r and z vanish on air cells).
template <typename T>
__global__ __launch_bounds__(256) void k_cg_rz(const Part* parts, int np) {
  FP_PART(parts, np);
  const State* st = cload(&Gp->st);
  if (cload(&st->done)) return;
  const LevelRec l0 = cload(&Gp->lv[0]);
  const T* __restrict__ z = (const T*)(l0.efinal ? l0.e2 : l0.e);
  const float* __restrict__ r = cload(&Gp->r);
  const float zm = l0.Se / cload(&st->rscale);
  const TileSpan sp = tile_span(l0, blk, nblk);
  const size_t plane = (size_t)l0.d[1] * l0.d[2];
  double acc[1] = {0.0};
  if (sp.inside) {
    int x = sp.x0;
    for (; x + 4 <= sp.x1; x += 4) {
      float rv[4], zv[4];
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const size_t i = (size_t)(x + k) * plane + sp.col;
        rv[k] = ldg(r, i);
        zv[k] = Store<T>::gld(z, i);
      }
#pragma unroll
      for (int k = 0; k < 4; ++k) acc[0] += (double)rv[k] * (double)(zv[k] * zm);
    }
    for (; x < sp.x1; ++x) {
      const size_t i = (size_t)x * plane + sp.col;
      acc[0] += (double)ldg(r, i) * (double)(Store<T>::gld(z, i) * zm);
    }
  }
  block_partials<1>(acc, cload(&Gp->partials), blk, nblk);
}

// e += alpha d ; r -= alpha q ; ||r||^2, max|r| -> convergence and next rscale,
// over the finest level's liquid hulls (d, q and r vanish on air cells, so the
// update is the identity there).  d: the direction this iteration wrote
// (buffer (iter & 1) ^ 1).
__global__ __launch_bounds__(256) void k_cg_update(const Part* parts, int np) {
  FP_PART(parts, np);
  const State* st = cload(&Gp->st);
  if (cload(&st->done)) return;
  const float alpha = (float)cload(&st->alpha);
  const float* __restrict__ d = cload(&Gp->d[(cload(&st->iter) & 1) ^ 1]);
  const float* __restrict__ q = cload(&Gp->q);
  float* __restrict__ e = cload(&Gp->e);
  float* __restrict__ r = cload(&Gp->r);
  const LevelRec l0 = cload(&Gp->lv[0]);
  const TileSpan sp = tile_span(l0, blk, nblk);
  const size_t plane = (size_t)l0.d[1] * l0.d[2];
  double acc[1] = {0.0};
  float rmax = 0.0f;
  if (sp.inside) {
    int x = sp.x0;
    for (; x + 4 <= sp.x1; x += 4) {
      float dv[4], qv[4], ev[4], rv[4];
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const size_t i = (size_t)(x + k) * plane + sp.col;
        dv[k] = ldg(d, i);
        qv[k] = ldg(q, i);
        ev[k] = ldg(e, i);
        rv[k] = ldg(r, i);
      }
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const size_t i = (size_t)(x + k) * plane + sp.col;
        const float en = ev[k] + alpha * dv[k], rn = rv[k] - alpha * qv[k];
        stg(e, i, en);
        stg(r, i, rn);
        acc[0] += (double)rn * (double)rn;
        rmax = fmaxf(rmax, fabsf(rn));
      }
    }
    for (; x < sp.x1; ++x) {
      const size_t i = (size_t)x * plane + sp.col;
      const float en = ldg(e, i) + alpha * ldg(d, i), rn = ldg(r, i) - alpha * ldg(q, i);
      stg(e, i, en);
      stg(r, i, rn);
      acc[0] += (double)rn * (double)rn;
      rmax = fmaxf(rmax, fabsf(rn));
    }
  }
  block_partials_max<1>(acc, rmax, cload(&Gp->partials), blk, nblk);
}

// --------------------------------------------------------------------------
// output: pressure and gradient subtract (FP64 weights and pressures)
// --------------------------------------------------------------------------
__global__ void k_pressure_out(const Part* parts, int np) {
  FP_PART(parts, np);
  float* __restrict__ p = G.out_p;
  if (!p) return;
  const float* __restrict__ phi = G.lv[0].phi;
  const double* __restrict__ x = G.x;
  // an all-liquid grid's pressure is defined up to a constant: its mean is 0
  const State* st = G.st;
  const double xmean = st->nullspace ? st->xsum / (double)st->liquid : 0.0;
  FP_STRIDE(i, (size_t)G.n) p[i] = phi[i] < 0.0f ? (float)(x[i] - xmean) : 0.0f;
}

// axis 0: faces (nx+1, ny, nz); 1: (nx, ny+1, nz); 2: (nx, ny, nz+1).
__global__ void k_subtract(const Part* parts, int np, int axis) {
  FP_PART(parts, np);
  const float* __restrict__ vin = axis == 0 ? G.in_u : (axis == 1 ? G.in_v : G.in_w);
  float* __restrict__ vout = axis == 0 ? G.out_u : (axis == 1 ? G.out_v : G.out_w);
  if (!vout) return;
  const Grid3 g{G.nx, G.ny, G.nz};
  const float* __restrict__ phi = G.lv[0].phi;
  const double* __restrict__ x = G.x;
  const double scale = G.sub_scale;
  const int fx = g.nx + (axis == 0), fy = g.ny + (axis == 1), fz = g.nz + (axis == 2);
  const size_t nf = (size_t)fx * fy * fz;
  FP_STRIDE(f, nf) {
    const int k = (int)(f % (size_t)fz);
    const size_t t = f / (size_t)fz;
    const int j = (int)(t % (size_t)fy);
    const int i = (int)(t / (size_t)fy);
    const int ii = axis == 0 ? i : (axis == 1 ? j : k);
    const int lim = axis == 0 ? g.nx : (axis == 1 ? g.ny : g.nz);
    float out = 0.0f;
    if (ii > 0 && ii < lim) {
      const size_t cp = ((size_t)i * g.ny + j) * g.nz + k;  // cell on the + side
      const size_t cm = cp - (axis == 0 ? (size_t)g.ny * g.nz : (axis == 1 ? (size_t)g.nz : 1));
      const float pp = phi[cp], pm = phi[cm];
      const double wf = face_weight_d(pm, pp);
      if (wf != 0.0) {
        const double xp = pp < 0.0f ? x[cp] : 0.0;
        const double xm = pm < 0.0f ? x[cm] : 0.0;
        out = (float)((double)vin[f] - scale * wf * (xp - xm));
      }
    }
    vout[f] = out;
  }
}

__global__ void k_div_after(const Part* parts, int np) {
  FP_PART(parts, np);
  const float* __restrict__ u = G.out_u;
  const float* __restrict__ v = G.out_v;
  const float* __restrict__ w = G.out_w;
  if (!u || !v || !w) return;
  const float* __restrict__ phi = G.lv[0].phi;
  const float* __restrict__ q = G.in_q;
  const double inv_h = G.inv_h;
  const Grid3 g{G.nx, G.ny, G.nz};
  float dmax = 0.0f;
  const int ny = g.ny, nz = g.nz;
  FP_STRIDE(idx, g.n()) {
    if (!(phi[idx] < 0.0f)) continue;
    int i, j, k;
    g.ijk(idx, i, j, k);
    const double flux = ((double)u[((size_t)(i + 1) * ny + j) * nz + k] - (double)u[((size_t)i * ny + j) * nz + k]) +
                        ((double)v[((size_t)i * (ny + 1) + j + 1) * nz + k] - (double)v[((size_t)i * (ny + 1) + j) * nz + k]) +
                        ((double)w[((size_t)i * ny + j) * (nz + 1) + k + 1] - (double)w[((size_t)i * ny + j) * (nz + 1) + k]);
    dmax = fmaxf(dmax, (float)fabs(flux * inv_h - (q ? (double)q[idx] : 0.0)));
  }
  double none[1] = {0.0};
  block_partials_max<1>(none, dmax, G.partials, blk, nblk);
}

// FP64 operator probe: out = A p (single grid, plain launch).
__global__ void k_apply_d(Grid3 g, const float* __restrict__ phi, const double* __restrict__ p,
                          double* __restrict__ out) {
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  for (size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x; idx < g.n(); idx += (size_t)gridDim.x * blockDim.x) {
    int i, j, k;
    g.ijk(idx, i, j, k);
    const float pc = phi[idx];
    double ap = 0.0;
    if (pc < 0.0f) {
      const double xc = p[idx];
#define ANB(cond, off)                                         \
  if (cond) {                                                  \
    const size_t m = idx + (off);                              \
    const float pn = phi[m];                                   \
    ap += face_weight_d(pc, pn) * (xc - (pn < 0.0f ? p[m] : 0.0)); \
  }
      ANB(i > 0, -(ptrdiff_t)((size_t)ny * nz))
      ANB(i + 1 < nx, (ptrdiff_t)((size_t)ny * nz))
      ANB(j > 0, -(ptrdiff_t)nz)
      ANB(j + 1 < ny, (ptrdiff_t)nz)
      ANB(k > 0, -1)
      ANB(k + 1 < nz, 1)
#undef ANB
    }
    out[idx] = ap;
  }
}

// ==========================================================================
// host side
// ==========================================================================

int plan_uniform_levels(int nx, int ny, int nz, int dims[][3], int strides[][3]) {
  // planUniformCM11aHierarchy (webgpu-uniform-pressure-multigrid.ts:149-169):
  // lockstep halving while the thinnest axis is a power of two, else
  // semi-coarsening of even axes; the haloed coarsest lattice must fit 256
  // cells.  Lattices it rejects (odd edges) are ceil-halved instead.
  auto haloed = [](const int* d) { return (long)(d[0] + 2) * (d[1] + 2) * (d[2] + 2); };
  const int d0[3] = {nx, ny, nz};
  if (nx < 2 || ny < 2 || nz < 2) return -1;
  int count = -1;
  {
    int m = std::min(nx, std::min(ny, nz));
    if ((m & (m - 1)) == 0) {
      int c = 0;
      while ((1 << (c + 1)) <= m) ++c;
      if (c >= 1 && c <= kMaxLevels) {
        const int coarsening = 1 << (c - 1);
        bool ok = (nx % coarsening == 0) && (ny % coarsening == 0) && (nz % coarsening == 0);
        if (ok) {
          for (int l = 0; l < c; ++l)
            for (int a = 0; a < 3; ++a) dims[l][a] = d0[a] >> l;
          if (haloed(dims[c - 1]) <= 256) count = c;
        }
      }
    }
  }
  if (count < 0) {
    int c = 1;
    for (int a = 0; a < 3; ++a) dims[0][a] = d0[a];
    for (;;) {
      int next[3];
      bool same = true;
      for (int a = 0; a < 3; ++a) {
        next[a] = (dims[c - 1][a] % 2 == 0 && dims[c - 1][a] > 2) ? dims[c - 1][a] / 2 : dims[c - 1][a];
        if (next[a] != dims[c - 1][a]) same = false;
      }
      if (same || c >= kMaxLevels) break;
      for (int a = 0; a < 3; ++a) dims[c][a] = next[a];
      ++c;
    }
    if (haloed(dims[c - 1]) <= 256) count = c;
  }
  if (count < 0) {
    int c = 1;
    for (int a = 0; a < 3; ++a) dims[0][a] = d0[a];
    while (haloed(dims[c - 1]) > 256 && c < kMaxLevels) {
      bool same = true;
      for (int a = 0; a < 3; ++a) {
        const int v = dims[c - 1][a] > 2 ? (dims[c - 1][a] + 1) / 2 : dims[c - 1][a];
        if (v != dims[c - 1][a]) same = false;
        dims[c][a] = v;
      }
      if (same) break;
      ++c;
    }
    if (haloed(dims[c - 1]) > 256) return -1;
    count = c;
  }
  for (int l = 0; l < count; ++l)
    for (int a = 0; a < 3; ++a) strides[l][a] = (l == 0 || dims[l][a] == dims[l - 1][a]) ? 1 : 2;
  return count;
}


#ifndef FP_SWEEP_DEPTH
#define FP_SWEEP_DEPTH 1
#endif
constexpr int kSweepDepth = FP_SWEEP_DEPTH;  // planes of loads in flight per workgroup

// Fused sweeps run on levels large enough to fill a 256-lane column tile.
static bool fused_dims(const int* d, long long n) { return d[2] >= 16 && d[1] >= 4 && d[0] >= 4 && n >= 4096; }

// Sweep geometry: tiles of TZ x 256/TZ columns, about FP_SWEEP_GROUPS
// workgroups per sweep of a level, x chunks of min_march_chunk planes or more
// (short on cache-resident coarse levels, where a sweep is bound by the
// latency of its chain of plane steps rather than by HBM).
#ifndef FP_SWEEP_GROUPS
#define FP_SWEEP_GROUPS 8192
#endif
static TileGeo sweep_geometry(const int* d, int TZ, int level, int groups = 0) {
  const int TY = 256 / TZ;
  TileGeo g;
  g.tz = (d[2] + TZ - 1) / TZ;
  g.ty = (d[1] + TY - 1) / TY;
  const long tiles = (long)g.tz * g.ty;
  const long cx_target = std::max<long>(1, (groups > 0 ? groups : FP_SWEEP_GROUPS) / tiles);
  const long min_chunk = min_march_chunk((long long)d[0] * d[1] * d[2], level);
  int chunk = (int)std::max<long>(min_chunk, (d[0] + cx_target - 1) / cx_target);
  chunk = std::min(chunk, d[0]);
  g.chunk = chunk;
  g.ngroups = (int)(tiles * ((d[0] + chunk - 1) / chunk));
  return g;
}

static std::atomic<int> g_next_grid_id{1};

UniformGrid::~UniformGrid() { release(); }

void UniformGrid::release() {
  for (void* p : allocations_) (void)hipFree(p);
  allocations_.clear();
  if (rec_) (void)hipHostFree(rec_);
  rec_ = nullptr;
}

int UniformGrid::alloc(void** p, size_t bytes) {
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

size_t UniformGrid::estimate_bytes(int nx, int ny, int nz) {
  int dims[kMaxLevels][3], strides[kMaxLevels][3];
  const int L = plan_uniform_levels(nx, ny, nz, dims, strides);
  if (L < 1) return 0;
  const size_t n = (size_t)nx * ny * nz;
  size_t total = 0;
  auto add = [&](size_t b) { total += (b + 255) & ~(size_t)255; };
  add(n * 8);  // b
  add(n * 8);  // x
  for (int i = 0; i < 5; ++i) add(n * 4);  // r e q d0 d1
  add((size_t)(nx + 1) * ny * nz * 4);     // input staging u v w q
  add((size_t)nx * (ny + 1) * nz * 4);
  add((size_t)nx * ny * (nz + 1) * 4);
  add(n * 4);
  for (int l = 0; l < L; ++l) {
    const size_t nl = (size_t)dims[l][0] * dims[l][1] * dims[l][2];
    add(nl * 4);      // phi
    add(nl * 4 * 4);  // r e e2 t
    add(nl * 16);     // packed coefficients
  }
  const size_t nc = std::min<size_t>((size_t)dims[L - 1][0] * dims[L - 1][1] * dims[L - 1][2], 512);
  add(nc * 4 + 16);
  add(16);
  add(nc * nc * 8 + 16);
  add(sizeof(State));
  add(4 * (size_t)kMaxBlocks * 8);
  add(sizeof(GridRec));
  return total;
}

int UniformGrid::create(int nx, int ny, int nz, double h) {
  int dims[kMaxLevels][3], strides[kMaxLevels][3];
  const int L = plan_uniform_levels(nx, ny, nz, dims, strides);
  if (L < 1) {
    set_error("no multigrid hierarchy for %dx%dx%d", nx, ny, nz);
    return FP_ERR_LAYOUT;
  }
  HIPCHK(hipHostMalloc((void**)&rec_, sizeof(GridRec), hipHostMallocDefault));
  std::memset(rec_, 0, sizeof(GridRec));
  GridRec& R = *rec_;
  id_ = g_next_grid_id.fetch_add(1);
  R.nx = nx; R.ny = ny; R.nz = nz; R.L = L; R.h = h;
  R.n = (long long)nx * ny * nz;
  R.ns_liquid = R.n;
  R.nu = (long long)(nx + 1) * ny * nz;
  R.nv = (long long)nx * (ny + 1) * nz;
  R.nw = (long long)nx * ny * (nz + 1);
  const size_t n = (size_t)R.n;
  int rc;
  if ((rc = alloc((void**)&R.b, n * 8))) return rc;
  if ((rc = alloc((void**)&R.x, n * 8))) return rc;
  if ((rc = alloc((void**)&R.r, n * 4))) return rc;
  if ((rc = alloc((void**)&R.e, n * 4))) return rc;
  if ((rc = alloc((void**)&R.q, n * 4))) return rc;
  if ((rc = alloc((void**)&R.d[0], n * 4))) return rc;
  if ((rc = alloc((void**)&R.d[1], n * 4))) return rc;
  if ((rc = alloc((void**)&in_u_, (size_t)R.nu * 4))) return rc;
  if ((rc = alloc((void**)&in_v_, (size_t)R.nv * 4))) return rc;
  if ((rc = alloc((void**)&in_w_, (size_t)R.nw * 4))) return rc;
  if ((rc = alloc((void**)&in_q_, n * 4))) return rc;
  for (int l = 0; l < L; ++l) {
    LevelRec& lv = R.lv[l];
    for (int a = 0; a < 3; ++a) {
      lv.d[a] = dims[l][a];
      lv.s[a] = strides[l][a];
      lv.g[a] = l == 0 ? 1.0f : R.lv[l - 1].g[a] / (float)(lv.s[a] * lv.s[a]);
    }
    lv.n = (long long)lv.d[0] * lv.d[1] * lv.d[2];
    const size_t nl = (size_t)lv.n;
    if ((rc = alloc((void**)&lv.phi, nl * 4))) return rc;
    if ((rc = alloc(&lv.r, nl * 4))) return rc;
    if ((rc = alloc(&lv.e, nl * 4))) return rc;
    if ((rc = alloc(&lv.e2, nl * 4))) return rc;
    if ((rc = alloc(&lv.w, nl * 16))) return rc;
    if ((rc = alloc(&lv.t, nl * 4))) return rc;
    lv.Sr = lv.Se = lv.St = 1.0f;
    lv.fused = (l < L - 1) && fused_dims(lv.d, lv.n);
    lv.sweep_tz = lv.d[2] >= 32 ? 32 : 16;
    lv.sw = sweep_geometry(lv.d, lv.sweep_tz, l);
    lv.mg = march_geometry(lv.d[0], lv.d[1], lv.d[2], l);
    // the restriction walks two fine planes per coarse plane, one barrier
    // each, and re-reads two fine planes per chunk: about 4096 workgroups,
    // chunks of 2 to 8 coarse planes (short chains on small levels, few
    // re-reads on large ones)
    lv.rg = lv.mg;
    {
      const long tiles = (long)lv.rg.tz * lv.rg.ty;
      const long c = ((long)lv.d[0] * tiles + 4095) / 4096;
      lv.rg.chunk = (int)std::min<long>(lv.d[0], std::max<long>(2, std::min<long>(8, c)));
    }
    lv.rg.ngroups = lv.rg.tz * lv.rg.ty * ((lv.d[0] + lv.rg.chunk - 1) / lv.rg.chunk);
    lv.hull = nullptr;
    if (lv.fused && (rc = alloc((void**)&lv.hull, (size_t)lv.mg.tz * lv.mg.ty * 2 * sizeof(int)))) return rc;
  }
  R.lb = L - 1;
  for (int l = 1; l < L; ++l)
    if (R.lv[l].n < kBottomCells) {
      R.lb = l;
      break;
    }
  const LevelRec& c = R.lv[L - 1];
  R.max_coarse = (int)std::min<long long>(c.n, kCoarseMax);
  if ((rc = alloc((void**)&R.cmap, (size_t)R.max_coarse * 4 + 16))) return rc;
  if ((rc = alloc((void**)&R.ncoarse, 16))) return rc;
  if ((rc = alloc((void**)&R.minv, (size_t)R.max_coarse * R.max_coarse * 8 + 16))) return rc;
  if ((rc = alloc((void**)&R.st, sizeof(State)))) return rc;
  if ((rc = alloc((void**)&R.partials, 4 * (size_t)kMaxBlocks * 8))) return rc;
  if ((rc = alloc((void**)&dev_rec_, sizeof(GridRec)))) return rc;
  HIPCHK(hipMemset(R.st, 0, sizeof(State)));
  HIPCHK(hipMemset(R.x, 0, n * 8));
  R.inv_h = 1.0 / h;
  R.max_iter = 1;
  HIPCHK(hipMemcpy(dev_rec_, rec_, sizeof(GridRec), hipMemcpyHostToDevice));
  return FP_OK;
}

// pressurePhi (webgpu-uniform-reference.wgsl.ts:215-242) with open fraction
// 1: phi = -(rho' - 1/2) dx from the surface density, in place
__global__ void k_density_to_phi(float* __restrict__ a, size_t n, float h) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    a[i] = -(a[i] - 0.5f) * h;
}

int UniformGrid::bind(hipStream_t s, const fp_inputs* in, fp_outputs* out, const fp_options* opt) {
  GridRec& R = *rec_;
  const size_t n = (size_t)R.n;
  host_io_ = in->memory == FP_MEMORY_HOST;
  // level 0 phi is a private copy, so callers may reuse their buffer; a
  // density input is copied there and converted in place
  const float* phi_src = in->density ? in->density : in->phi;
  if (host_io_) {
    HIPCHK(hipMemcpyAsync(R.lv[0].phi, phi_src, n * 4, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(in_u_, in->u, (size_t)R.nu * 4, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(in_v_, in->v, (size_t)R.nv * 4, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(in_w_, in->w, (size_t)R.nw * 4, hipMemcpyHostToDevice, s));
    if (in->target_divergence)
      HIPCHK(hipMemcpyAsync(in_q_, in->target_divergence, n * 4, hipMemcpyHostToDevice, s));
    R.in_u = in_u_;
    R.in_v = in_v_;
    R.in_w = in_w_;
    R.in_q = in->target_divergence ? in_q_ : nullptr;
    // projected in place on the device copies; FP32 pressure staged in r
    R.out_u = out && out->u ? in_u_ : nullptr;
    R.out_v = out && out->v ? in_v_ : nullptr;
    R.out_w = out && out->w ? in_w_ : nullptr;
    R.out_p = out && out->pressure ? R.r : nullptr;
  } else {
    HIPCHK(hipMemcpyAsync(R.lv[0].phi, phi_src, n * 4, hipMemcpyDeviceToDevice, s));
    R.in_u = in->u;
    R.in_v = in->v;
    R.in_w = in->w;
    R.in_q = in->target_divergence;
    R.out_u = out ? out->u : nullptr;
    R.out_v = out ? out->v : nullptr;
    R.out_w = out ? out->w : nullptr;
    R.out_p = out ? out->pressure : nullptr;
  }
  if (in->density)
    hipLaunchKernelGGL(k_density_to_phi, dim3((unsigned)std::min<size_t>((n + 255) / 256, 65536)), dim3(256), 0, s,
                       R.lv[0].phi, n, (float)R.h);
  const double rho = in->rho, dt = in->dt, h = R.h;
  R.s_flux = rho * h / dt;
  R.s_q = rho * h * h / dt;
  R.inv_h = 1.0 / h;
  R.sub_scale = dt / (rho * h);
  R.tol2 = opt ? opt->relative_tolerance * opt->relative_tolerance : 0.0;
  R.max_iter = opt ? std::max(1, opt->max_iterations) : 1;
  R.warm = opt ? (opt->warm_start ? 1 : 0) : 0;
  R.rscale_bias = opt ? std::ldexp(1.0f, opt->precision_test_shift) : 1.0f;
  phi_staged_ = true;
  return FP_OK;
}

int UniformGrid::finish_outputs(hipStream_t s, const fp_inputs* in, fp_outputs* out) {
  if (!host_io_ || !out) return FP_OK;
  const GridRec& R = *rec_;
  if (out->u) HIPCHK(hipMemcpyAsync(out->u, in_u_, (size_t)R.nu * 4, hipMemcpyDeviceToHost, s));
  if (out->v) HIPCHK(hipMemcpyAsync(out->v, in_v_, (size_t)R.nv * 4, hipMemcpyDeviceToHost, s));
  if (out->w) HIPCHK(hipMemcpyAsync(out->w, in_w_, (size_t)R.nw * 4, hipMemcpyDeviceToHost, s));
  if (out->pressure) HIPCHK(hipMemcpyAsync(out->pressure, R.r, (size_t)R.n * 4, hipMemcpyDeviceToHost, s));
  (void)in;
  return FP_OK;
}

void UniformGrid::set_tuning(int tile_z, int groups, int pairs) {
  GridRec& R = *rec_;
  sweep_pairs_ = pairs;
  for (int l = 0; l < R.L; ++l) {
    LevelRec& lv = R.lv[l];
    const int tz = (tile_z == 32 && lv.d[2] >= 32) ? 32 : (tile_z == 16 ? 16 : (lv.d[2] >= 32 ? 32 : 16));
    lv.sweep_tz = tz;
    lv.sw = sweep_geometry(lv.d, tz, l, groups);
  }
}

void UniformGrid::set_level_scales(int precision, int pre, int post) {
  // FP32: every scale is 1.  FP16/BF16: the finest right-hand side arrives
  // normalised to |r| <= 1 (device rscale); error vectors grow like
  // A^-1 r, i.e. up to ~n^2 cells^2 on smooth modes, so each level stores
  // e / 2^(2 ceil(log2 n_l) - 8) to stay inside the FP16 range.
  GridRec& R = *rec_;
  for (int l = 0; l < R.L; ++l) {
    LevelRec& lv = R.lv[l];
    // fused levels ping-pong e/e2 once per sweep; the coarsest and the
    // unfused levels work in place on e
    lv.efinal = lv.fused ? ((pre + post) & 1) : 0;
    if (precision == FP_PRECISION_FP32) {
      lv.Sr = lv.Se = lv.St = 1.0f;
      continue;
    }
    const int nmax = std::max(lv.d[0], std::max(lv.d[1], lv.d[2]));
    int lg = 0;
    while ((1 << lg) < nmax) ++lg;
    int ee = 2 * lg - 8;
    if (ee < 0) ee = 0;
    lv.Sr = 1.0f;
    lv.St = 1.0f;
    lv.Se = std::ldexp(1.0f, ee);
  }
}

int UniformGrid::upload(hipStream_t s) {
  HIPCHK(hipMemcpyAsync(dev_rec_, rec_, sizeof(GridRec), hipMemcpyHostToDevice, s));
  return FP_OK;
}

// --------------------------------------------------------------------------
// Projector: part tables of a batch and the solve schedule
// --------------------------------------------------------------------------
struct Projector::Plan {
  std::vector<UniformGrid*> grids;
  int precision = 1, pre = 2, post = 2, maxL = 1;
  bool persistent = false;  // sweeps as persistent launches (fp_options.persistent)
  bool pairs = false;       // finest-level sweeps in fused pairs (FP16/BF16 multigrid, every grid opted in)
  std::vector<Part> host;
  Part* arena = nullptr;
  struct Pending {
    PartTable* t;
    size_t off;
  };
  std::vector<Pending> pending;
  PartTable one;        // one workgroup per grid
  PartTable cells;      // finest-level cells, 1 per lane
  PartTable faces[3];   // face arrays
  PartTable march0;     // finest-level column-tile march
  PartTable coarse_top, coarse_deep;  // coarsest level = finest / deeper
  long long bottom_n[kMaxLevels] = {};  // per bottom level (from lb): most cells over coarse_deep's grids
  struct Level {
    PartTable sweep32, sweep16;  // fused grids by sweep tile width
    PartTable res;               // fused grids: residual march
    PartTable fusedcells;        // fused grids: cells of the level
    PartTable smooth;            // unfused grids: cells of the level
    PartTable restrict_;         // grids with a coarser level: coarse cells
    PartTable down;              // grids with this level (l >= 1): its cells
  } lv[kMaxLevels];
  hipGraphExec_t graph = nullptr;
  State* gathered = nullptr;     // device
  State* gathered_host = nullptr;

  static int blocks_for(size_t items, int per_lane) {
    const size_t b = (items + (size_t)FP_BLOCK * per_lane - 1) / ((size_t)FP_BLOCK * per_lane);
    return (int)std::max<size_t>(1, std::min<size_t>(b, 4096));
  }
  // Table of (grid, workgroups); part starts are multiples of 8.
  void add(PartTable& t, const std::vector<std::pair<UniformGrid*, int>>& items) {
    t.n = (int)items.size();
    pending.push_back({&t, host.size()});
    int b0 = 0;
    for (const auto& it : items) {
      host.push_back(Part{b0, it.second, it.first->dev_rec()});
      b0 += (it.second + 7) & ~7;
    }
    t.blocks = b0;
  }
};

Projector::Projector() : noalloc_plan_(new Plan()) {}

Projector::~Projector() {
  delete noalloc_plan_;
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

void Projector::forget(int grid_id) {
  for (auto it = plans_.begin(); it != plans_.end();) {
    bool uses = false;
    // keys: the grid ids, then precision, pre, post, persistent
    for (size_t i = 0; i + 4 < it->first.size(); ++i)
      if (it->first[i] == grid_id) uses = true;
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

// Host part tables of a grid set (dev pointers are set by the upload).
void Projector::fill_tables(Plan* p, const std::vector<UniformGrid*>& grids, int precision, int pre, int post) {
  p->grids = grids;
  p->precision = precision;
  p->pre = pre;
  p->post = post;
  p->pairs = precision != FP_PRECISION_FP32;
  for (UniformGrid* g : grids) p->pairs = p->pairs && g->sweep_pairs() == 1;
  p->host.clear();
  p->pending.clear();
  using Items = std::vector<std::pair<UniformGrid*, int>>;
  Items one, cells, faces[3], march0, ctop, cdeep;
  int maxL = 1;
  for (UniformGrid* g : grids) {
    const GridRec& R = *g->host_rec();
    maxL = std::max(maxL, R.L);
    one.push_back({g, 1});
    cells.push_back({g, Plan::blocks_for((size_t)R.n, 1)});
    faces[0].push_back({g, Plan::blocks_for((size_t)R.nu, 1)});
    faces[1].push_back({g, Plan::blocks_for((size_t)R.nv, 1)});
    faces[2].push_back({g, Plan::blocks_for((size_t)R.nw, 1)});
    march0.push_back({g, R.lv[0].mg.ngroups});
    (R.L == 1 ? ctop : cdeep).push_back({g, 1});
  }
  // LDS of the one-workgroup bottom: sized by the largest bottom of the set
  for (int q = 0; q < kMaxLevels; ++q) p->bottom_n[q] = 0;
  for (UniformGrid* g : grids) {
    const GridRec& R = *g->host_rec();
    if (R.L == 1) continue;
    for (int l = R.lb; l < R.L; ++l) p->bottom_n[l - R.lb] = std::max(p->bottom_n[l - R.lb], R.lv[l].n);
  }
  p->maxL = maxL;
  p->add(p->one, one);
  p->add(p->cells, cells);
  for (int a = 0; a < 3; ++a) p->add(p->faces[a], faces[a]);
  p->add(p->march0, march0);
  p->add(p->coarse_top, ctop);
  p->add(p->coarse_deep, cdeep);
  for (int l = 0; l < maxL; ++l) {
    Items s32, s16, res, fc, sm, rs, dn;
    for (UniformGrid* g : grids) {
      const GridRec& R = *g->host_rec();
      if (l >= R.L) continue;
      const LevelRec& lv = R.lv[l];
      if (l >= 1) dn.push_back({g, Plan::blocks_for((size_t)lv.n, 1)});
      if (l >= R.L - 1) continue;  // the coarsest level is solved directly
      if (l < R.lb) rs.push_back({g, R.lv[l + 1].rg.ngroups});
      if (lv.fused) {
        (lv.sweep_tz == 32 ? s32 : s16).push_back({g, lv.sw.ngroups});
        res.push_back({g, lv.mg.ngroups});
        fc.push_back({g, Plan::blocks_for((size_t)lv.n, 1)});
      } else if (l < R.lb) {
        sm.push_back({g, Plan::blocks_for((size_t)lv.n, 1)});
      }
    }
    Plan::Level& T = p->lv[l];
    p->add(T.sweep32, s32);
    p->add(T.sweep16, s16);
    p->add(T.res, res);
    p->add(T.fusedcells, fc);
    p->add(T.smooth, sm);
    p->add(T.restrict_, rs);
    p->add(T.down, dn);
  }
}

// Copies a plan's parts into dev (capacity checked by the caller) and points
// its tables there.
static void set_table_pointers(Projector::Plan* p, Part* dev) {
  for (const auto& pd : p->pending) pd.t->dev = dev + pd.off;
}

Projector::Plan* Projector::plan_for(const std::vector<UniformGrid*>& grids, const fp_options* opt) {
  std::vector<long long> key;
  for (UniformGrid* g : grids) key.push_back(g->id());
  key.push_back(opt->precision);
  key.push_back(opt->pre_sweeps);
  key.push_back(opt->post_sweeps);
  key.push_back(opt->persistent ? 1 : 0);
  auto it = plans_.find(key);
  if (it != plans_.end()) return it->second;
  Plan* p = new Plan();
  fill_tables(p, grids, opt->precision, opt->pre_sweeps, opt->post_sweeps);
  p->persistent = opt->persistent != 0;
  bool ok = hipMalloc((void**)&p->arena, std::max<size_t>(1, p->host.size()) * sizeof(Part)) == hipSuccess;
  ok = ok && hipMemcpy(p->arena, p->host.data(), p->host.size() * sizeof(Part), hipMemcpyHostToDevice) == hipSuccess;
  ok = ok && hipMalloc((void**)&p->gathered, grids.size() * sizeof(State)) == hipSuccess;
  ok = ok && hipHostMalloc((void**)&p->gathered_host, grids.size() * sizeof(State), hipHostMallocDefault) == hipSuccess;
  if (!ok) {
    set_error("projector plan: device allocation failed");
    if (p->arena) (void)hipFree(p->arena);
    if (p->gathered) (void)hipFree(p->gathered);
    if (p->gathered_host) (void)hipHostFree(p->gathered_host);
    delete p;
    return nullptr;
  }
  set_table_pointers(p, p->arena);
  plans_[key] = p;
  return p;
}

void Projector::timer_begin(hipStream_t s) {
  if ((size_t)(2 * timer_used_ + 2) > timer_events_.size()) {
    hipEvent_t a, b;
    if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) return;
    timer_events_.push_back(a);
    timer_events_.push_back(b);
  }
  (void)hipEventRecord(timer_events_[2 * timer_used_], s);
}

void Projector::timer_end(hipStream_t s) {
  if ((size_t)(2 * timer_used_ + 2) > timer_events_.size()) return;
  (void)hipEventRecord(timer_events_[2 * timer_used_ + 1], s);
  ++timer_used_;
}

#define TIMED(cls, stmt)                        \
  do {                                          \
    const bool _t = (cls) == timed_kernel_;     \
    if (_t) timer_begin(s);                     \
    stmt;                                       \
    if (_t) timer_end(s);                       \
  } while (0)

template <typename TR, typename T, bool ZERO>
static void sweep_launch_persistent(hipStream_t s, const PartTable& t, int tz, int level, int which_in, int first) {
  if (t.n <= 0) return;
  const int grid = std::min(t.blocks, kPersistGroups);
  if (tz == 32)
    hipLaunchKernelGGL((k_sweep3_persistent<32, kSweepDepth, TR, T, ZERO>), dim3(grid), dim3(FP_BLOCK), 0, s, t.dev,
                       t.n, t.blocks, level, which_in, first);
  else
    hipLaunchKernelGGL((k_sweep3_persistent<16, kSweepDepth, TR, T, ZERO>), dim3(grid), dim3(FP_BLOCK), 0, s, t.dev,
                       t.n, t.blocks, level, which_in, first);
}

template <typename TR, typename T>
static void sweep_launch(hipStream_t s, const PartTable& t32, const PartTable& t16, int level, int which_in, int first,
                         bool zero, bool persistent) {
  if (persistent) {
    if (zero) {
      sweep_launch_persistent<TR, T, true>(s, t32, 32, level, which_in, first);
      sweep_launch_persistent<TR, T, true>(s, t16, 16, level, which_in, first);
    } else {
      sweep_launch_persistent<TR, T, false>(s, t32, 32, level, which_in, first);
      sweep_launch_persistent<TR, T, false>(s, t16, 16, level, which_in, first);
    }
    return;
  }
  if (zero) {
    PLAUNCH((k_sweep3<32, kSweepDepth, TR, T, true>), t32, s, level, which_in, first);
    PLAUNCH((k_sweep3<16, kSweepDepth, TR, T, true>), t16, s, level, which_in, first);
  } else {
    PLAUNCH((k_sweep3<32, kSweepDepth, TR, T, false>), t32, s, level, which_in, first);
    PLAUNCH((k_sweep3<16, kSweepDepth, TR, T, false>), t16, s, level, which_in, first);
  }
}

// Two finest-level sweeps in one launch (FP16/BF16 storage only: the FP32
// staging would not leave room for three workgroups per CU).
template <typename T, bool ZERO>
static void sweep2_launch(hipStream_t s, const PartTable& t32, const PartTable& t16, int which_in, int first,
                          bool persistent) {
  if constexpr (sizeof(T) == 2) {
    if (persistent) {
      if (t32.n > 0)
        hipLaunchKernelGGL((k_sweep2_persistent<32, float, T, ZERO>), dim3(std::min(t32.blocks, kPersistGroups)),
                           dim3(FP_BLOCK), 0, s, t32.dev, t32.n, t32.blocks, which_in, first);
      if (t16.n > 0)
        hipLaunchKernelGGL((k_sweep2_persistent<16, float, T, ZERO>), dim3(std::min(t16.blocks, kPersistGroups)),
                           dim3(FP_BLOCK), 0, s, t16.dev, t16.n, t16.blocks, which_in, first);
      return;
    }
    PLAUNCH((k_sweep2<32, float, T, ZERO>), t32, s, which_in, first);
    PLAUNCH((k_sweep2<16, float, T, ZERO>), t16, s, which_in, first);
  } else {
    (void)s, (void)t32, (void)t16, (void)which_in, (void)first, (void)persistent;
  }
}

template <typename T>
int Projector::vcycle_t(hipStream_t s, Plan& p) {
  const int pre = p.pre, post = p.post;
  for (int l = 0; l + 1 < p.maxL; ++l) {
    Plan::Level& T_ = p.lv[l];
    const bool top = l == 0;
    // ---- pre-smoothing ----
    if (pre == 0) {
      PLAUNCH(k_zero_level<T>, T_.fusedcells, s, l);
      PLAUNCH(k_zero_level<T>, T_.smooth, s, l);
    }
    if (top && p.pairs && pre > 0) {
      // fused levels: launches of two sweeps (and one of one for an odd
      // count), each reading one buffer and writing the other; the first,
      // from zero, writes the buffer that leaves the result in pre & 1
      const int launches = pre / 2 + (pre & 1);
      int buf = (pre & 1) ^ ((launches - 1) & 1);
      TIMED(FP_KERNEL_SMOOTH_FINE, (pre >= 2 ? sweep2_launch<T, true>(s, T_.sweep32, T_.sweep16, buf ? 0 : 1, 0,
                                                                       p.persistent)
                                             : sweep_launch<float, T>(s, T_.sweep32, T_.sweep16, l, buf ? 0 : 1, 0,
                                                                      true, p.persistent)));
      for (int left = pre - (pre >= 2 ? 2 : 1); left > 0; left -= 2) {
        TIMED(FP_KERNEL_SMOOTH_FINE, (left >= 2 ? sweep2_launch<T, false>(s, T_.sweep32, T_.sweep16, buf, 0,
                                                                           p.persistent)
                                                : sweep_launch<float, T>(s, T_.sweep32, T_.sweep16, l, buf, 0,
                                                                         false, p.persistent)));
        buf ^= 1;
      }
    }
    for (int sw = 0; sw < pre; ++sw) {
      if (!(top && p.pairs))
        TIMED(top ? FP_KERNEL_SMOOTH_FINE : -1,
              (top ? sweep_launch<float, T>(s, T_.sweep32, T_.sweep16, l, sw & 1, 0, sw == 0, p.persistent)
                   : sweep_launch<T, T>(s, T_.sweep32, T_.sweep16, l, sw & 1, 0, sw == 0, p.persistent)));
      for (int colour = 0; colour < 2; ++colour) {
        const int zi = (sw == 0 && colour == 0) ? 1 : 0;
        if (top) PLAUNCH((k_smooth<float, T>), T_.smooth, s, l, colour, zi);
        else PLAUNCH((k_smooth<T, T>), T_.smooth, s, l, colour, zi);
      }
    }
    // ---- residual, restriction ----
    const MarchLaunch la{l, pre & 1};
    TIMED(top ? FP_KERNEL_RESIDUAL_FINE : -1, {
      if (top) {
        PLAUNCH((k_march<OpResidualW<float, T>>), T_.res, s, 1, la);
        PLAUNCH((k_residual<float, T>), T_.smooth, s, l);
      } else {
        PLAUNCH((k_march<OpResidualW<T, T>>), T_.res, s, 1, la);
        PLAUNCH((k_residual<T, T>), T_.smooth, s, l);
      }
    });
    TIMED(top ? FP_KERNEL_RESTRICT_FINE : -1, PLAUNCH(k_restrict_march<T>, T_.restrict_, s, l));
  }
  FLAUNCH((k_coarse_solve<T, float>), p.coarse_top, s);
  // grids with a coarser level: the levels from lb on and the coarsest solve
  // in one workgroup per grid
  if (p.coarse_deep.n > 0) {
    // LDS-resident bottom when its levels fit beside the static LDS (the
    // grid record, the coarse solve's right-hand side)
    const size_t lds = bottom_lds_bytes_of<T>(p.bottom_n, kMaxLevels);
    if (lds <= kBottomLdsMax)
      hipLaunchKernelGGL((k_bottom_lds<T>), dim3(p.coarse_deep.n), dim3(kBottomThreads), lds, s, p.coarse_deep.dev,
                         pre, post);
    else
      hipLaunchKernelGGL((k_bottom<T>), dim3(p.coarse_deep.n), dim3(kBottomThreads), 0, s, p.coarse_deep.dev, pre,
                         post);
  }
  for (int l = p.maxL - 2; l >= 0; --l) {
    Plan::Level& T_ = p.lv[l];
    const bool top = l == 0;
    // ---- fused levels: prolongation over the liquid hulls, then the sweeps ----
    if (post == 0) {
      PLAUNCH(k_prolong_add<T>, T_.fusedcells, s, l, pre & 1);
    } else if (top && p.pairs) {
      // launches of two sweeps flip the buffer once per pair: the
      // prolongation writes the other buffer when that leaves the result in
      // (pre + post) & 1
      const int launches = post / 2 + (post & 1);
      int buf = pre & 1;
      const int out = (launches & 1) == (post & 1) ? buf : buf ^ 1;
      TIMED(FP_KERNEL_PROLONG_FINE, PLAUNCH(k_prolong_march<T>, T_.res, s, l, buf, out));
      buf = out;
      for (int left = post; left > 0; left -= 2) {
        TIMED(FP_KERNEL_SMOOTH_FINE, (left >= 2 ? sweep2_launch<T, false>(s, T_.sweep32, T_.sweep16, buf, 1,
                                                                           p.persistent)
                                                : sweep_launch<float, T>(s, T_.sweep32, T_.sweep16, l, buf, 1,
                                                                         false, p.persistent)));
        buf ^= 1;
      }
    } else {
      TIMED(top ? FP_KERNEL_PROLONG_FINE : -1, PLAUNCH(k_prolong_march<T>, T_.res, s, l, pre & 1, pre & 1));
      for (int sw = 0; sw < post; ++sw) {
        TIMED(top ? FP_KERNEL_SMOOTH_FINE : -1,
              (top ? sweep_launch<float, T>(s, T_.sweep32, T_.sweep16, l, (pre + sw) & 1, 1, false, p.persistent)
                   : sweep_launch<T, T>(s, T_.sweep32, T_.sweep16, l, (pre + sw) & 1, 1, false, p.persistent)));
      }
    }
    // ---- unfused levels ----
    TIMED(top ? FP_KERNEL_PROLONG_FINE : -1, PLAUNCH(k_prolong_add<T>, T_.smooth, s, l, 0));
    for (int sw = 0; sw < post; ++sw)
      for (int cc = 0; cc < 2; ++cc) {
        const int colour = 1 - cc;
        if (top) PLAUNCH((k_smooth<float, T>), T_.smooth, s, l, colour, 0);
        else PLAUNCH((k_smooth<T, T>), T_.smooth, s, l, colour, 0);
      }
  }
  return FP_OK;
}

int Projector::enqueue_vcycle(hipStream_t s, Plan& p) {
  if (p.precision == FP_PRECISION_MIXED_FP16) return vcycle_t<__half>(s, p);
  if (p.precision == FP_PRECISION_MIXED_BF16) return vcycle_t<__hip_bfloat16>(s, p);
  return vcycle_t<float>(s, p);
}

template <typename T>
int Projector::iteration_t(hipStream_t s, Plan& p) {
  vcycle_t<T>(s, p);
  PLAUNCH(k_cg_rz<T>, p.march0, s);
  FLAUNCH(k_fin_rz<Part>, p.march0, s);
  TIMED(FP_KERNEL_CG_DIRECTION, PLAUNCH(k_march<OpDirection<T>>, p.march0, s, 1, MarchLaunch{0, 0}));
  FLAUNCH(k_fin_dq<Part>, p.march0, s);
  TIMED(FP_KERNEL_CG_UPDATE, PLAUNCH(k_cg_update, p.march0, s));
  FLAUNCH(k_fin_update<Part>, p.march0, s);
  return FP_OK;
}

int Projector::enqueue_iteration(hipStream_t s, Plan& p) {
  if (p.precision == FP_PRECISION_MIXED_FP16) return iteration_t<__half>(s, p);
  if (p.precision == FP_PRECISION_MIXED_BF16) return iteration_t<__hip_bfloat16>(s, p);
  return iteration_t<float>(s, p);
}

int Projector::ensure_graph(hipStream_t s, Plan& p) {
  if (p.graph) return FP_OK;
  // one hipGraph holds two CG + V-cycle iterations (the d ping-pong);
  // kernels of finished grids return at once
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

int Projector::poll(hipStream_t s, Plan& p, int* dev_counts, int* host_counts, int* counts) {
  hipLaunchKernelGGL(k_status<Part>, dim3(1), dim3(FP_BLOCK), 0, s, p.one.dev, p.one.n, dev_counts);
  HIPCHK(hipMemcpyAsync(host_counts, dev_counts, 8, hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  counts[0] = host_counts[0];
  counts[1] = host_counts[1];
  return FP_OK;
}

template <typename T>
static void build_weights_t(hipStream_t s, const Projector* self, int maxL, const PartTable* tabs) {
  (void)self;
  for (int l = 0; l + 1 < maxL; ++l) PLAUNCH(k_build_weights<T>, tabs[l], s, l);
}

int Projector::setup_tables(hipStream_t s, Plan& p, int preserve) {
  for (int l = 1; l < p.maxL; ++l) PLAUNCH(k_downsample_phi, p.lv[l].down, s, l, preserve);
  FLAUNCH(k_coarse_setup, p.one, s);
  PartTable wt[kMaxLevels];
  for (int l = 0; l < p.maxL; ++l) wt[l] = p.lv[l].fusedcells;
  if (p.precision == FP_PRECISION_MIXED_FP16) build_weights_t<__half>(s, this, p.maxL, wt);
  else if (p.precision == FP_PRECISION_MIXED_BF16) build_weights_t<__hip_bfloat16>(s, this, p.maxL, wt);
  else build_weights_t<float>(s, this, p.maxL, wt);
  for (int l = 0; l + 1 < p.maxL; ++l) {
    PLAUNCH(k_hull_reset, p.lv[l].fusedcells, s, l);
    PLAUNCH(k_tile_hull, p.lv[l].res, s, l);
  }
  return FP_OK;
}

// The solve of a prepared plan: set-up, rhs, FP64 refinement passes around
// the CG + V-cycle iterations (replayed as graphs when `graphs`), gradient
// subtract and outputs.  dev_counts / host_counts: 8 bytes each for the
// host's convergence polls.
int Projector::solve(hipStream_t s, Plan& p, const fp_options* opt, bool graphs, int* dev_counts, int* host_counts) {
  FLAUNCH(k_reset_state<Part>, p.one, s);
  setup_tables(s, p, opt->phi_preserve_levels);
  PLAUNCH(k_rhs, p.cells, s);
  FLAUNCH(k_fin_rhs<Part>, p.cells, s);
  PLAUNCH(k_zero, p.march0, s, 0, 0);  // x, except on warm-started grids
  PLAUNCH(k_zero, p.march0, s, 1, 0);
  PLAUNCH(k_zero, p.march0, s, 2, 0);
  const int max_outer = std::max(1, opt->max_outer);
  const int max_iter = std::max(1, opt->max_iterations);
  const int check = std::max(1, opt->check_every);
  int counts[2] = {0, 0};
  for (int pass = 0; pass <= max_outer; ++pass) {
    const int new_pass = pass < max_outer;
    TIMED(FP_KERNEL_TRUE_RESIDUAL,
          PLAUNCH(k_march<OpTrueResidual>, p.march0, s, 2, MarchLaunch{0, pass == 0 ? 1 : 0}));
    FLAUNCH(k_fin_true<Part>, p.march0, s, new_pass);
    int rc = poll(s, p, dev_counts, host_counts, counts);
    if (rc) return rc;
    if (counts[0] == 0 || !new_pass) break;
    if (graphs) {
      rc = ensure_graph(s, p);
      if (rc) return rc;
      const int period = std::max(2, check + (check & 1));
      for (int it = 0; it < max_iter; it += 2) {
        HIPCHK(hipGraphLaunch(p.graph, s));
        if ((it + 2) % period == 0 || it + 2 >= max_iter) {
          if ((rc = poll(s, p, dev_counts, host_counts, counts))) return rc;
          if (counts[1] == 0) break;
        }
      }
    } else {
      for (int it = 0; it < max_iter; ++it) {
        enqueue_iteration(s, p);
        if ((it + 1) % check == 0 || it + 1 == max_iter) {
          if ((rc = poll(s, p, dev_counts, host_counts, counts))) return rc;
          if (counts[1] == 0) break;
        }
      }
    }
    PLAUNCH(k_accumulate, p.march0, s);
    PLAUNCH(k_zero, p.march0, s, 2, 1);
  }
  for (int a = 0; a < 3; ++a) PLAUNCH(k_subtract, p.faces[a], s, a);
  PLAUNCH(k_div_after, p.cells, s);
  FLAUNCH(k_fin_divafter<Part>, p.cells, s);
  PLAUNCH(k_pressure_out, p.cells, s);
  return FP_OK;
}

static void fill_report(const State& S, const GridRec& R, float ms, double timed_ms, int timed_launches,
                        fp_report* rep) {
  std::memset(rep, 0, sizeof(*rep));
  int flags = S.flags;
  if (!S.outer_done || (S.flags & 0x1)) flags |= 0x8;  // FP_FLAG_NONCONVERGENCE
  rep->flags = flags;
  rep->iterations = S.total_iter;
  rep->outer_iterations = S.outer;
  rep->levels = R.L;
  // a non-finite right-hand side (non-finite inputs) reads as NaN
  rep->rhs_norm = std::isfinite(S.bb) ? std::sqrt(std::max(0.0, S.bb)) : std::nan("");
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

// Projections of a grid set; grids whose FP16 multigrid overflowed (a
// non-finite preconditioned residual with a finite right-hand side: the
// FP16 range, not the inputs) are projected again with the BF16 multigrid
// and reported with FP_FLAG_PRECISION_FALLBACK.
int Projector::run(hipStream_t s, const std::vector<UniformGrid*>& grids, const fp_options* opt, fp_report* reps) {
  std::vector<fp_report> local;
  if (!reps) {
    local.resize(grids.size());
    reps = local.data();
  }
  int rc = run_once(s, grids, opt, reps);
  if (rc || opt->precision != FP_PRECISION_MIXED_FP16) return rc;
  std::vector<UniformGrid*> redo;
  std::vector<size_t> at;
  for (size_t i = 0; i < grids.size(); ++i)
    if ((reps[i].flags & 0x1) && std::isfinite(reps[i].rhs_norm)) {
      redo.push_back(grids[i]);
      at.push_back(i);
    }
  if (redo.empty()) return FP_OK;
  fp_options bf = *opt;
  bf.precision = FP_PRECISION_MIXED_BF16;
  std::vector<fp_report> again(redo.size());
  if ((rc = run_once(s, redo, &bf, again.data()))) return rc;
  for (size_t k = 0; k < redo.size(); ++k) {
    const double first_ms = reps[at[k]].device_ms;
    reps[at[k]] = again[k];
    reps[at[k]].flags |= 0x20;  // FP_FLAG_PRECISION_FALLBACK
    reps[at[k]].device_ms += first_ms;
  }
  return FP_OK;
}

int Projector::run_once(hipStream_t s, const std::vector<UniformGrid*>& grids, const fp_options* opt,
                        fp_report* reps) {
  if (grids.empty()) return FP_OK;
  for (UniformGrid* g : grids) {
    g->set_level_scales(opt->precision, opt->pre_sweeps, opt->post_sweeps);
    int rc = g->upload(s);
    if (rc) return rc;
  }
  Plan* pp = plan_for(grids, opt);
  if (!pp) return FP_ERR_OUT_OF_MEMORY;
  Plan& p = *pp;
  if (!dev_counts_) {
    HIPCHK(hipMalloc((void**)&dev_counts_, 16));
    HIPCHK(hipHostMalloc((void**)&host_counts_, 16, hipHostMallocDefault));
  }
  timed_kernel_ = opt->timed_kernel;
  timer_used_ = 0;
  hipEvent_t ev0, ev1;
  HIPCHK(hipEventCreate(&ev0));
  HIPCHK(hipEventCreate(&ev1));
  HIPCHK(hipEventRecord(ev0, s));
  int rc = solve(s, p, opt, opt->graphs && timed_kernel_ == 0, dev_counts_, host_counts_);
  if (rc) return rc;
  HIPCHK(hipEventRecord(ev1, s));
  hipLaunchKernelGGL(k_gather_states<Part>, dim3((unsigned)((grids.size() + 255) / 256)), dim3(256), 0, s, p.one.dev,
                     p.one.n, p.gathered);
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
  if (reps)
    for (size_t i = 0; i < grids.size(); ++i)
      fill_report(p.gathered_host[i], *grids[i]->host_rec(), ms, timed_ms, timed_launches, &reps[i]);
  return FP_OK;
}

// ---- the projection that allocates nothing ---------------------------------
// The device memory fp_project_noalloc needs beyond the grid's own buffers --
// the part tables of its launches, the gathered State and the poll counters --
// lives in this array.  Each device holds its own copy for the life of the
// process; host code never passes `scratch` itself to a kernel or a copy, only
// the address hipGetSymbolAddress returns for the current device.
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
unsigned char* g_scratch_ptr[kMaxDevices] = {};
}  // namespace

// Looks up this device's copy of `scratch` once (with the device current).
int noalloc_scratch_init(int device) {
  if (device < 0 || device >= kMaxDevices) {
    set_error("device %d out of range for the no-allocation scratch", device);
    return FP_ERR_DEVICE;
  }
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_ptr[device]) return FP_OK;
  void* ptr = nullptr;
  hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !ptr) {
    set_error("hipGetSymbolAddress(scratch) on device %d: %s", device, hipGetErrorString(e));
    return FP_ERR_DEVICE;
  }
  g_scratch_ptr[device] = (unsigned char*)ptr;
  return FP_OK;
}

int Projector::run_noalloc(hipStream_t s, int device, UniformGrid* g, const fp_options* opt, fp_report* rep) {
  unsigned char* base = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    base = (device >= 0 && device < kMaxDevices) ? g_scratch_ptr[device] : nullptr;
  }
  if (!base) {
    set_error("fp_project_noalloc: the scratch of device %d was not looked up (context set-up failed)", device);
    return FP_ERR_STATE;
  }
  g->set_level_scales(opt->precision, opt->pre_sweeps, opt->post_sweeps);
  int rc = g->upload(s);
  if (rc) return rc;
  // host tables in a plan kept for this entry point (its vectors keep their
  // capacity), device tables in the scratch: parts, then the State, then
  // the counters
  Plan& p = *noalloc_plan_;
  const std::vector<UniformGrid*> one_grid(1, g);
  fill_tables(&p, one_grid, opt->precision, opt->pre_sweeps, opt->post_sweeps);
  p.persistent = opt->persistent != 0;
  const size_t parts_bytes = p.host.size() * sizeof(Part);
  const size_t state_off = (parts_bytes + 255) & ~(size_t)255;
  const size_t counts_off = state_off + ((sizeof(State) + 255) & ~(size_t)255);
  const size_t need = counts_off + 16;
  if (need > kNoAllocScratchBytes) {
    set_error("fp_project_noalloc: %zu bytes of launch tables exceed the %zu-byte scratch", need,
              (size_t)kNoAllocScratchBytes);
    return FP_ERR_INVALID_ARGUMENT;
  }
  Part* parts = (Part*)base;
  State* gathered = (State*)(base + state_off);
  int* dev_counts = (int*)(base + counts_off);
  HIPCHK(hipMemcpyAsync(parts, p.host.data(), parts_bytes, hipMemcpyHostToDevice, s));
  set_table_pointers(&p, parts);
  timed_kernel_ = 0;
  timer_used_ = 0;
  int host_counts[4] = {0, 0, 0, 0};
  const auto t0 = std::chrono::steady_clock::now();
  rc = solve(s, p, opt, false, dev_counts, host_counts);
  if (rc) return rc;
  hipLaunchKernelGGL(k_gather_states<Part>, dim3(1), dim3(256), 0, s, p.one.dev, p.one.n, gathered);
  State S;
  HIPCHK(hipMemcpyAsync(&S, gathered, sizeof(State), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  const float ms = std::chrono::duration<float, std::milli>(std::chrono::steady_clock::now() - t0).count();
  if (rep) fill_report(S, *g->host_rec(), ms, 0.0, 0, rep);
  return FP_OK;
}


// ---- continuous batching ---------------------------------------------------
static double now_seconds() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

int Projector::serve_round(hipStream_t s, std::vector<Job*>& jobs) {
  struct Group {
    fp_options opt;
    std::vector<Job*> jobs;
  };
  auto key_of = [](const fp_options& o) {
    return ((long long)o.persistent << 40) | ((long long)o.precision << 32) | ((long long)o.pre_sweeps << 24) |
           ((long long)o.post_sweeps << 16) | (long long)(o.phi_preserve_levels & 0xffff);
  };
  std::map<long long, Group> setup, accum, iter, resid[2];
  std::vector<Job*> finish, live;
  for (Job* j : jobs) {
    if (j->stage == Job::DONE) continue;
    live.push_back(j);
    const long long k = key_of(j->opt);
    Group* g = nullptr;
    if (j->stage == Job::SETUP) g = &setup[k];
    else if (j->stage == Job::ACCUMULATE) g = &accum[k];
    else if (j->stage == Job::ITERATE) g = &iter[k];
    else finish.push_back(j);
    if (g) {
      g->opt = j->opt;
      g->jobs.push_back(j);
    }
    if (j->stage == Job::SETUP || j->stage == Job::ACCUMULATE) {
      // an FP64 residual follows in this round; another inner pass may follow it
      Group& r = resid[j->passes < std::max(1, j->opt.max_outer) ? 1 : 0][k];
      r.opt = j->opt;
      r.jobs.push_back(j);
    }
  }
  if (live.empty()) return FP_OK;
  std::vector<std::unique_ptr<Plan>> plans;
  auto make = [&](const std::vector<Job*>& js, const fp_options& o) {
    std::vector<UniformGrid*> gs;
    for (Job* j : js) gs.push_back(j->grid);
    plans.emplace_back(new Plan());
    fill_tables(plans.back().get(), gs, o.precision, o.pre_sweeps, o.post_sweeps);
    plans.back()->persistent = o.persistent != 0;
    return plans.back().get();
  };
  std::vector<std::pair<Plan*, const Group*>> setup_p, accum_p, iter_p, resid_p[2];
  for (auto& kv : setup) setup_p.push_back({make(kv.second.jobs, kv.second.opt), &kv.second});
  for (auto& kv : accum) accum_p.push_back({make(kv.second.jobs, kv.second.opt), &kv.second});
  for (int np = 0; np < 2; ++np)
    for (auto& kv : resid[np]) resid_p[np].push_back({make(kv.second.jobs, kv.second.opt), &kv.second});
  for (auto& kv : iter) iter_p.push_back({make(kv.second.jobs, kv.second.opt), &kv.second});
  fp_options plain;
  std::memset(&plain, 0, sizeof(plain));
  Plan* fin_p = finish.empty() ? nullptr : make(finish, plain);
  Plan* all_p = make(live, plain);
  // tables of the round: one staging copy into the reusable arena (the
  // previous round ended with a stream synchronisation, so it is free)
  size_t total = 0;
  for (auto& p : plans) total += p->host.size();
  if (total > round_arena_cap_) {
    if (round_arena_) (void)hipFree(round_arena_);
    if (round_staging_) (void)hipHostFree(round_staging_);
    round_arena_ = nullptr;
    round_staging_ = nullptr;
    const size_t cap = std::max<size_t>(total * 2, 4096);
    HIPCHK(hipMalloc((void**)&round_arena_, cap * sizeof(Part)));
    HIPCHK(hipHostMalloc((void**)&round_staging_, cap * sizeof(Part), hipHostMallocDefault));
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
    std::memcpy(round_staging_ + off, p->host.data(), p->host.size() * sizeof(Part));
    set_table_pointers(p.get(), round_arena_ + off);
    off += p->host.size();
  }
  HIPCHK(hipMemcpyAsync(round_arena_, round_staging_, std::max<size_t>(total, 1) * sizeof(Part),
                        hipMemcpyHostToDevice, s));
  // ---- set-up: records, level sets, coarse inverses, coefficients, rhs ----
  for (auto& sp : setup_p)
    for (Job* j : sp.second->jobs) {
      j->t_start = now_seconds();
      int rc = j->grid->bind(s, &j->in, &j->out, &j->opt);
      if (rc) return rc;
      j->grid->set_level_scales(j->opt.precision, j->opt.pre_sweeps, j->opt.post_sweeps);
      if ((rc = j->grid->upload(s))) return rc;
    }
  for (auto& sp : setup_p) {
    Plan& p = *sp.first;
    FLAUNCH(k_reset_state<Part>, p.one, s);
    setup_tables(s, p, sp.second->opt.phi_preserve_levels);
    PLAUNCH(k_rhs, p.cells, s);
    FLAUNCH(k_fin_rhs<Part>, p.cells, s);
    PLAUNCH(k_zero, p.march0, s, 0, 0);
    PLAUNCH(k_zero, p.march0, s, 1, 0);
    PLAUNCH(k_zero, p.march0, s, 2, 0);
  }
  // ---- finished inner solves fold into the FP64 iterate ----
  for (auto& ap : accum_p) {
    PLAUNCH(k_accumulate, ap.first->march0, s);
    PLAUNCH(k_zero, ap.first->march0, s, 2, 1);
  }
  // ---- FP64 true residuals (refinement control) ----
  for (int np = 0; np < 2; ++np)
    for (auto& rp : resid_p[np]) {
      PLAUNCH(k_march<OpTrueResidual>, rp.first->march0, s, 2, MarchLaunch{0, 0});
      FLAUNCH(k_fin_true<Part>, rp.first->march0, s, np);
    }
  // ---- one CG + V-cycle iteration of every iterating grid ----
  for (auto& ip : iter_p) enqueue_iteration(s, *ip.first);
  // ---- gradient subtract of converged grids ----
  if (fin_p) {
    for (int a = 0; a < 3; ++a) PLAUNCH(k_subtract, fin_p->faces[a], s, a);
    PLAUNCH(k_div_after, fin_p->cells, s);
    FLAUNCH(k_fin_divafter<Part>, fin_p->cells, s);
    PLAUNCH(k_pressure_out, fin_p->cells, s);
    for (Job* j : finish) {
      int rc = j->grid->finish_outputs(s, &j->in, &j->out);
      if (rc) return rc;
    }
  }
  hipLaunchKernelGGL(k_gather_states<Part>, dim3((unsigned)((live.size() + 255) / 256)), dim3(256), 0, s, all_p->one.dev,
                     all_p->one.n, round_states_);
  HIPCHK(hipMemcpyAsync(round_states_host_, round_states_, live.size() * sizeof(State), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  const double t = now_seconds();
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
        if (S.done && (S.flags & 0x1) && j->opt.precision == FP_PRECISION_MIXED_FP16 && std::isfinite(S.bb)) {
          // the FP16 multigrid overflowed: start again with the BF16 multigrid
          j->opt.precision = FP_PRECISION_MIXED_BF16;
          j->fallback = 1;
          j->passes = 0;
          j->stage = Job::SETUP;
        } else if (S.done) {
          j->stage = Job::ACCUMULATE;
        }
        break;
      case Job::FINISH: {
        const GridRec& R = *j->grid->host_rec();
        fp_report* rep = &j->report;
        std::memset(rep, 0, sizeof(*rep));
        int flags = S.flags;
        if (!S.outer_done || (S.flags & 0x1)) flags |= 0x8;
        if (j->fallback) flags |= 0x20;  // FP_FLAG_PRECISION_FALLBACK
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

// ---- parity probes (single grid) -------------------------------------------
int Projector::probe_rhs(hipStream_t s, UniformGrid* g, double* b_host) {
  fp_options o;
  std::memset(&o, 0, sizeof(o));
  o.precision = FP_PRECISION_FP32;
  Plan* p = plan_for({g}, &o);
  if (!p) return FP_ERR_OUT_OF_MEMORY;
  g->host_rec()->tol2 = 0.0;
  int rc = g->upload(s);
  if (rc) return rc;
  FLAUNCH(k_reset_state<Part>, p->one, s);
  PLAUNCH(k_rhs, p->cells, s);
  FLAUNCH(k_fin_rhs<Part>, p->cells, s);
  HIPCHK(hipMemcpyAsync(b_host, g->host_rec()->b, (size_t)g->host_rec()->n * 8, hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  return FP_OK;
}

int Projector::probe_vcycle(hipStream_t s, UniformGrid* g, const fp_options* opt, float rscale, void* z_out) {
  Plan* p = plan_for({g}, opt);
  if (!p) return FP_ERR_OUT_OF_MEMORY;
  g->set_level_scales(opt->precision, opt->pre_sweeps, opt->post_sweeps);
  int rc = g->upload(s);
  if (rc) return rc;
  FLAUNCH(k_reset_state<Part>, p->one, s);
  setup_tables(s, *p, opt->phi_preserve_levels);
  HIPCHK(hipMemcpyAsync(&g->host_rec()->st->rscale, &rscale, sizeof(float), hipMemcpyHostToDevice, s));
  // the finest level's sweeps leave its air planes outside the liquid hulls
  // untouched: the probe returns zeros there
  const LevelRec& l0 = g->level(0);
  const size_t ebytes = (size_t)l0.n * (opt->precision == FP_PRECISION_FP32 ? 4 : 2);
  HIPCHK(hipMemsetAsync(l0.e, 0, ebytes, s));
  HIPCHK(hipMemsetAsync(l0.e2, 0, ebytes, s));
  HIPCHK(hipStreamSynchronize(s));
  enqueue_vcycle(s, *p);
  const size_t bytes = (size_t)l0.n * (opt->precision == FP_PRECISION_FP32 ? 4 : 2);
  HIPCHK(hipMemcpyAsync(z_out, l0.efinal ? l0.e2 : l0.e, bytes, hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  return FP_OK;
}

int Projector::probe_level_phi(hipStream_t s, UniformGrid* g, int preserve, int level, float* out) {
  fp_options o;
  std::memset(&o, 0, sizeof(o));
  o.precision = FP_PRECISION_FP32;
  Plan* p = plan_for({g}, &o);
  if (!p) return FP_ERR_OUT_OF_MEMORY;
  int rc = g->upload(s);
  if (rc) return rc;
  FLAUNCH(k_reset_state<Part>, p->one, s);
  for (int l = 1; l <= level; ++l) PLAUNCH(k_downsample_phi, p->lv[l].down, s, l, preserve);
  HIPCHK(hipMemcpyAsync(out, g->level(level).phi, (size_t)g->level(level).n * 4, hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  return FP_OK;
}

int probe_apply(hipStream_t s, const UniformGrid* g, const float* phi_host, const double* p_host, double* out_host) {
  const GridRec& R = *const_cast<UniformGrid*>(g)->host_rec();
  const size_t n = (size_t)R.n;
  float* phi = nullptr;
  double *p = nullptr, *o = nullptr;
  HIPCHK(hipMalloc(&phi, n * 4));
  HIPCHK(hipMalloc(&p, n * 8));
  HIPCHK(hipMalloc(&o, n * 8));
  HIPCHK(hipMemcpy(phi, phi_host, n * 4, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(p, p_host, n * 8, hipMemcpyHostToDevice));
  const unsigned blocks = (unsigned)std::min<size_t>(4096, (n + FP_BLOCK - 1) / FP_BLOCK);
  hipLaunchKernelGGL(k_apply_d, dim3(blocks), dim3(FP_BLOCK), 0, s, Grid3{R.nx, R.ny, R.nz}, (const float*)phi,
                     (const double*)p, o);
  HIPCHK(hipStreamSynchronize(s));
  HIPCHK(hipMemcpy(out_host, o, n * 8, hipMemcpyDeviceToHost));
  (void)hipFree(phi);
  (void)hipFree(p);
  (void)hipFree(o);
  return FP_OK;
}

// The 1-D transfer taps along one axis: per fine index its prolongation taps
// (ptaps: two coarse cells, -1 where unused) and per coarse index its
// restriction taps (restrict_taps: four fine cells, -1 where unused).
__global__ void k_probe_taps(int s, int nf, int nc, int* pc, float* pw, int* rf, float* rw) {
  const int t = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (t < nf) {
    int c[2] = {-1, -1};
    float w[2] = {0.0f, 0.0f};
    ptaps(t, s, nc, c, w);
    for (int q = 0; q < 2; ++q) {
      pc[2 * t + q] = c[q];
      pw[2 * t + q] = w[q];
    }
  }
  if (t < nc) {
    int f[4] = {-1, -1, -1, -1};
    float w[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    restrict_taps(t, s, nf, nc, f, w);
    for (int q = 0; q < 4; ++q) {
      rf[4 * t + q] = f[q];
      rw[4 * t + q] = w[q];
    }
  }
}

int probe_taps(hipStream_t s, const UniformGrid* g, int level, int axis, int* pc_host, float* pw_host, int* rf_host,
               float* rw_host) {
  const LevelRec& fl = g->level(level - 1);
  const LevelRec& cl = g->level(level);
  const int nf = fl.d[axis], nc = cl.d[axis], st = cl.s[axis];
  int* pc = nullptr;
  int* rf = nullptr;
  float *pw = nullptr, *rw = nullptr;
  HIPCHK(hipMalloc(&pc, (size_t)nf * 2 * 4));
  HIPCHK(hipMalloc(&pw, (size_t)nf * 2 * 4));
  HIPCHK(hipMalloc(&rf, (size_t)nc * 4 * 4));
  HIPCHK(hipMalloc(&rw, (size_t)nc * 4 * 4));
  const unsigned blocks = (unsigned)((std::max(nf, nc) + FP_BLOCK - 1) / FP_BLOCK);
  hipLaunchKernelGGL(k_probe_taps, dim3(blocks), dim3(FP_BLOCK), 0, s, st, nf, nc, pc, pw, rf, rw);
  HIPCHK(hipStreamSynchronize(s));
  HIPCHK(hipMemcpy(pc_host, pc, (size_t)nf * 2 * 4, hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(pw_host, pw, (size_t)nf * 2 * 4, hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(rf_host, rf, (size_t)nc * 4 * 4, hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(rw_host, rw, (size_t)nc * 4 * 4, hipMemcpyDeviceToHost));
  (void)hipFree(pc);
  (void)hipFree(pw);
  (void)hipFree(rf);
  (void)hipFree(rw);
  return FP_OK;
}

}  // namespace fp

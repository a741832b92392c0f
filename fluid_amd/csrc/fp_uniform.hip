// fp_uniform.hip -- uniform MAC-grid pressure projection on gfx950.
//
// One projection = divergence/rhs -> FP64 iterative refinement around an
// FP32 multigrid-preconditioned CG (multigrid vectors in FP32, FP16 or BF16)
// -> pressure-gradient subtract.  All arrays are C order, z fastest.
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
#include "fp_march.hpp"

#include <algorithm>
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

#define GRID_STRIDE(idx, N) for (size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x; idx < (N); idx += (size_t)gridDim.x * blockDim.x)

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

// --------------------------------------------------------------------------
// generic fills
// --------------------------------------------------------------------------
template <typename T>
__global__ void k_fill(T* p, size_t n, T v) {
  GRID_STRIDE(i, n) p[i] = v;
}
__global__ void k_fill4(float* p, size_t n, float v) {
  const size_t n4 = n / 4;
  GRID_STRIDE(i, n4) reinterpret_cast<float4*>(p)[i] = make_float4(v, v, v, v);
  if (blockIdx.x == 0 && threadIdx.x < (n & 3)) p[n4 * 4 + threadIdx.x] = v;
}

// --------------------------------------------------------------------------
// rhs: b = -(rho h/dt) sum(u_out - u_in) + (rho h^2/dt) q on liquid cells.
// Also ||b||^2 and max |div u| (1/s) before the projection.
// --------------------------------------------------------------------------
__global__ void k_rhs(Grid3 g, const float* __restrict__ phi, const float* __restrict__ u,
                      const float* __restrict__ v, const float* __restrict__ w, const float* __restrict__ q,
                      double s_flux, double s_q, double inv_h, double tol2, double* __restrict__ b, State* st,
                      double* partials) {
  double acc[2] = {0.0, 0.0};
  float dmax = 0.0f;
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  GRID_STRIDE(idx, g.n()) {
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
    }
    b[idx] = val;
  }
  block_partials_max<2>(acc, dmax, partials);
}

__global__ void k_fin_rhs(const double* __restrict__ partials, unsigned G, double tol2, State* st) {
  double out[2];
  float mx;
  sum_partials_max<2>(partials, G, out, mx);
  if (threadIdx.x == 0) {
    st->divmax_before = __float_as_uint(mx);
    st->bb = out[0];
    // The refinement stops at a quarter of the requested relative residual:
    // the max-norm divergence left in the FP32 output then stays below that
    // of an FP64 solve stopped at the requested tolerance.
    st->tol2_bb = 0.0625 * tol2 * out[0];
    st->liquid = (long long)out[1];
  }
}

// Refinement control after the FP64 true residual (OpTrueResidual): decides
// convergence of the refinement and the target of the next inner solve.
__global__ void k_fin_true(const double* __restrict__ partials, unsigned G, State* st) {
  double out[1];
  float m;
  sum_partials_max<1>(partials, G, out, m);
  if (threadIdx.x == 0) {
    const double rr = out[0];
    st->rr64 = rr;
    const bool finite = isfinite(rr);
    st->outer_done = (!finite) || rr <= st->tol2_bb;
    if (!finite) st->flags |= 0x1;  // FP_FLAG_NONFINITE
    // inner solve: a tenth of the refinement target (margin for the max-norm
    // divergence), at most what FP32 CG can be trusted to reduce in one pass.
    st->inner_target2 = fmax(0.01 * st->tol2_bb, 1e-12 * rr);
    st->iter = 0;
    st->done = st->outer_done;
    st->rscale = (m > 0.0f && isfinite(m)) ? exp2f(-ceilf(log2f(m))) : 1.0f;
  }
}

// x += e; e = 0 (folds one inner solve into the FP64 iterate).
__global__ void k_accumulate(size_t n, double* __restrict__ x, float* __restrict__ e) {
  const size_t n4 = n / 4;
  GRID_STRIDE(i, n4) {
    const float4 ev = reinterpret_cast<const float4*>(e)[i];
    double4 xv = reinterpret_cast<const double4*>(x)[i];
    xv.x += ev.x; xv.y += ev.y; xv.z += ev.z; xv.w += ev.w;
    reinterpret_cast<double4*>(x)[i] = xv;
    reinterpret_cast<float4*>(e)[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  }
  if (blockIdx.x == 0 && threadIdx.x < (n & 3)) {
    const size_t i = n4 * 4 + threadIdx.x;
    x[i] += (double)e[i];
    e[i] = 0.0f;
  }
}

// --------------------------------------------------------------------------
// multigrid setup: coarse level set (CM11a Eq. 15-16 with C preserved levels)
// --------------------------------------------------------------------------
__global__ void k_downsample_phi(Grid3 f, const float* __restrict__ pf, Grid3 c, float* __restrict__ pc, int sx,
                                 int sy, int sz, int use_positive) {
  GRID_STRIDE(idx, c.n()) {
    int I, J, K;
    c.ijk(idx, I, J, K);
    float sum = 0.0f, pos = 0.0f;
    int cnt = 0, npos = 0, nneg = 0;
    for (int oi = 0; oi < sx; ++oi) {
      const int i = I * sx + oi;
      if (i >= f.nx) continue;
      for (int oj = 0; oj < sy; ++oj) {
        const int j = J * sy + oj;
        if (j >= f.ny) continue;
        for (int ok = 0; ok < sz; ++ok) {
          const int k = K * sz + ok;
          if (k >= f.nz) continue;
          const float v = pf[((size_t)i * f.ny + j) * f.nz + k];
          sum += v;
          ++cnt;
          if (v >= 0.0f) { pos += v; ++npos; } else ++nneg;
        }
      }
    }
    const bool mixed = npos > 0 && nneg > 0;
    pc[idx] = (mixed && use_positive) ? pos / (float)npos : sum / (float)cnt;
  }
}

// Dense inverse of the coarsest level's liquid block (one workgroup, FP64
// Gauss-Jordan in place; SPD so no pivoting).  A failed pivot falls back to
// the diagonal (Jacobi) map so the preconditioner stays a fixed SPD map.
__global__ void k_coarse_setup(LevelArgs L, int* __restrict__ cmap, int* __restrict__ ncount,
                               double* __restrict__ a, int max_n, State* st) {
  __shared__ int s_n;
  __shared__ int s_fail;
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
  }
  __syncthreads();
  for (int p = 0; p < n; ++p) {
    const double piv = a[p * n + p];
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

// --------------------------------------------------------------------------
// V-cycle kernels.  Stored value = true value / scale; TR is the storage of
// the right-hand side (FP32 at the finest level, where it is the CG residual
// scaled by the device-side rscale), T of everything else.
// --------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ float ldT(const T* p, size_t i) { return Store<T>::ld(p, i); }
template <typename T>
__device__ __forceinline__ void stT(T* p, size_t i, float v) { Store<T>::st(p, i, v); }

// One colour of red-black Gauss-Seidel (colour = (i+j+k)&1).  zero_init:
// the first pass from a zero guess, which also clears the other colour.
template <typename TR, typename T>
__global__ void k_smooth(LevelArgs L, const TR* __restrict__ r, float rmul, int dyn, T* __restrict__ e, int colour,
                         int zero_init, const State* st) {
  if (st->done) return;
  const float rm = dyn ? rmul * st->rscale : rmul;
  const Grid3 g = L.g;
  const size_t sx = (size_t)g.ny * g.nz, sy = (size_t)g.nz;
  GRID_STRIDE(idx, g.n()) {
    int i, j, k;
    g.ijk(idx, i, j, k);
    if (((i + j + k) & 1) != colour) {
      if (zero_init) stT(e, idx, 0.0f);
      continue;
    }
    const float pc = L.phi[idx];
    if (!(pc < 0.0f)) {
      stT(e, idx, 0.0f);
      continue;
    }
    float diag = 0.0f, acc = ldT(r, idx) * rm;
#define SNB(cond, off, gg)                                      \
  if (cond) {                                                   \
    const size_t m = idx + (off);                               \
    const float pn = L.phi[m];                                  \
    const float wf = face_weight(pc, pn, gg);                   \
    diag += wf;                                                 \
    if (!zero_init && pn < 0.0f) acc += wf * ldT(e, m);         \
  }
    SNB(i > 0, -(ptrdiff_t)sx, L.gx)
    SNB(i + 1 < g.nx, (ptrdiff_t)sx, L.gx)
    SNB(j > 0, -(ptrdiff_t)sy, L.gy)
    SNB(j + 1 < g.ny, (ptrdiff_t)sy, L.gy)
    SNB(k > 0, -1, L.gz)
    SNB(k + 1 < g.nz, 1, L.gz)
#undef SNB
    stT(e, idx, diag > 0.0f ? acc / diag : 0.0f);
  }
}

// t = r - A e  (stored with multiplier tmul applied: t_s = (r*rm - emul*A e_s) * tinv)
template <typename TR, typename T>
__global__ void k_residual(LevelArgs L, const TR* __restrict__ r, float rmul, int dyn, const T* __restrict__ e,
                           float emul, float tinv, T* __restrict__ t, const State* st) {
  if (st->done) return;
  const float rm = dyn ? rmul * st->rscale : rmul;
  const Grid3 g = L.g;
  const size_t sx = (size_t)g.ny * g.nz, sy = (size_t)g.nz;
  GRID_STRIDE(idx, g.n()) {
    int i, j, k;
    g.ijk(idx, i, j, k);
    const float pc = L.phi[idx];
    float res = 0.0f;
    if (pc < 0.0f) {
      const float ec = ldT(e, idx);
      float ap = 0.0f;
#define RNB(cond, off, gg)                                          \
  if (cond) {                                                       \
    const size_t m = idx + (off);                                   \
    const float pn = L.phi[m];                                      \
    const float wf = face_weight(pc, pn, gg);                       \
    ap += wf * (ec - (pn < 0.0f ? ldT(e, m) : 0.0f));               \
  }
      RNB(i > 0, -(ptrdiff_t)sx, L.gx)
      RNB(i + 1 < g.nx, (ptrdiff_t)sx, L.gx)
      RNB(j > 0, -(ptrdiff_t)sy, L.gy)
      RNB(j + 1 < g.ny, (ptrdiff_t)sy, L.gy)
      RNB(k > 0, -1, L.gz)
      RNB(k + 1 < g.nz, 1, L.gz)
#undef RNB
      res = (ldT(r, idx) * rm - emul * ap) * tinv;
    }
    stT(t, idx, res);
  }
}

// 1-D taps of the transpose of the prolongation (restriction gather).
__device__ __forceinline__ int rtaps(int I, int s, int nf, int nc, int* f, float* w) {
  if (s == 1) { f[0] = I; w[0] = 1.0f; return 1; }
  int c = 0;
  if (I >= 1) { f[c] = 2 * I - 1; w[c] = 0.25f; ++c; }
  f[c] = 2 * I; w[c] = (I - 1 < 0) ? 1.0f : 0.75f; ++c;
  if (2 * I + 1 < nf) { f[c] = 2 * I + 1; w[c] = (I + 1 >= nc) ? 1.0f : 0.75f; ++c; }
  if (2 * I + 2 < nf) { f[c] = 2 * I + 2; w[c] = 0.25f; ++c; }
  return c;
}
// 1-D taps of the prolongation for fine index f.
__device__ __forceinline__ int ptaps(int f, int s, int nc, int* c, float* w) {
  if (s == 1) { c[0] = f; w[0] = 1.0f; return 1; }
  const int I = f >> 1;
  const int J = (f & 1) ? I + 1 : I - 1;
  if (J < 0 || J >= nc) { c[0] = I; w[0] = 1.0f; return 1; }
  c[0] = I; w[0] = 0.75f; c[1] = J; w[1] = 0.25f;
  return 2;
}

template <typename T>
__global__ void k_restrict(Grid3 f, const T* __restrict__ tf, LevelArgs C, int sx, int sy, int sz, float fac,
                           T* __restrict__ rc, const State* st) {
  if (st->done) return;
  const Grid3 c = C.g;
  GRID_STRIDE(idx, c.n()) {
    int I, J, K;
    c.ijk(idx, I, J, K);
    float acc = 0.0f;
    if (C.phi[idx] < 0.0f) {
      int fi[4], fj[4], fk[4];
      float wi[4], wj[4], wk[4];
      const int ni = rtaps(I, sx, f.nx, c.nx, fi, wi);
      const int nj = rtaps(J, sy, f.ny, c.ny, fj, wj);
      const int nk = rtaps(K, sz, f.nz, c.nz, fk, wk);
      for (int a = 0; a < ni; ++a)
        for (int b = 0; b < nj; ++b) {
          const float wab = wi[a] * wj[b];
          const size_t base = ((size_t)fi[a] * f.ny + fj[b]) * f.nz;
          for (int q = 0; q < nk; ++q) acc += wab * wk[q] * ldT(tf, base + fk[q]);
        }
      acc *= fac;
    }
    stT(rc, idx, acc);
  }
}

template <typename T>
__global__ void k_prolong_add(LevelArgs F, T* __restrict__ ef, Grid3 c, const float* __restrict__ phic,
                              const T* __restrict__ ec, int sx, int sy, int sz, float fac, const State* st) {
  if (st->done) return;
  const Grid3 f = F.g;
  GRID_STRIDE(idx, f.n()) {
    if (!(F.phi[idx] < 0.0f)) continue;
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
}

// e = fac * Minv r on the coarsest liquid cells (one workgroup).
template <typename T, typename TR>
__global__ void k_coarse_solve(Grid3 g, const int* __restrict__ cmap, const int* __restrict__ ncount,
                               const double* __restrict__ minv, const TR* __restrict__ r, float rmul, int dyn,
                               T* __restrict__ e, float fac, const State* st) {
  if (st->done) return;
  __shared__ double rs[512];
  const int n = *ncount;
  const float rm = dyn ? rmul * st->rscale : rmul;
  for (size_t idx = threadIdx.x; idx < g.n(); idx += blockDim.x) stT(e, idx, 0.0f);
  for (int q = threadIdx.x; q < n; q += blockDim.x) rs[q] = (double)(ldT(r, (size_t)cmap[q]) * rm);
  __syncthreads();
  for (int row = threadIdx.x; row < n; row += blockDim.x) {
    double s = 0.0;
    for (int q = 0; q < n; ++q) s += minv[(size_t)row * n + q] * rs[q];
    stT(e, (size_t)cmap[row], (float)(s * fac));
  }
}

// --------------------------------------------------------------------------
// CG kernels (FP32 vectors, FP64 scalars)
// --------------------------------------------------------------------------
// gamma = r . z ; beta = gamma / gamma_prev (Fletcher-Reeves) or the
// Polak-Ribiere form when z_prev is given.
template <typename T>
__global__ void k_cg_rz(size_t n, const float* __restrict__ r, const T* __restrict__ z, float zmul_static,
                        State* st, double* partials) {
  if (st->done) return;
  const float zm = zmul_static / st->rscale;
  double acc[1] = {0.0};
  {
    const size_t n4 = n / 4;
    GRID_STRIDE(i, n4) {
      const float4 rv = reinterpret_cast<const float4*>(r)[i];
      const float4 zv = V4<T>::ld(z, i);
      acc[0] += ((double)rv.x * (double)(zv.x * zm) + (double)rv.y * (double)(zv.y * zm)) +
                ((double)rv.z * (double)(zv.z * zm) + (double)rv.w * (double)(zv.w * zm));
    }
    if (blockIdx.x == 0 && threadIdx.x < (n & 3)) {
      const size_t i = n4 * 4 + threadIdx.x;
      acc[0] += (double)r[i] * (double)(ldT(z, i) * zm);
    }
  }
  block_partials<1>(acc, partials);
}

__global__ void k_fin_rz(const double* __restrict__ partials, unsigned G, State* st) {
  if (st->done) return;
  double out[1];
  sum_partials<1>(partials, G, out);
  if (threadIdx.x == 0) {
    const double rz = out[0];
    if (!(rz > 0.0) || !isfinite(rz)) {
      st->flags |= isfinite(rz) ? 0x4 : 0x1;  // NONPOSITIVE_PRECONDITIONER / NONFINITE
      st->done = 1;
      return;
    }
    st->beta = (st->iter == 0) ? 0.0 : rz / st->rz;
    st->rz = rz;
  }
}

__global__ void k_fin_dq(const double* __restrict__ partials, unsigned G, State* st) {
  if (st->done) return;
  double out[1];
  sum_partials<1>(partials, G, out);
  if (threadIdx.x == 0) {
    const double dq = out[0];
    if (!(dq > 0.0) || !isfinite(dq)) {
      st->flags |= isfinite(dq) ? 0x2 : 0x1;  // NONPOSITIVE_CURVATURE / NONFINITE
      st->done = 1;
      return;
    }
    st->alpha = st->rz / dq;
  }
}

// e += alpha d ; r -= alpha q ; ||r||^2, max|r| -> convergence and next rscale.
__global__ void k_cg_update(size_t n, const float* __restrict__ d, const float* __restrict__ q,
                            float* __restrict__ e, float* __restrict__ r, State* st, double* partials) {
  if (st->done) return;
  const float alpha = (float)st->alpha;
  double acc[1] = {0.0};
  float rmax = 0.0f;
  const size_t n4 = n / 4;
  GRID_STRIDE(i, n4) {
    const float4 dv = reinterpret_cast<const float4*>(d)[i];
    const float4 qv = reinterpret_cast<const float4*>(q)[i];
    float4 ev = reinterpret_cast<const float4*>(e)[i];
    float4 rv = reinterpret_cast<const float4*>(r)[i];
    ev.x += alpha * dv.x; ev.y += alpha * dv.y; ev.z += alpha * dv.z; ev.w += alpha * dv.w;
    rv.x -= alpha * qv.x; rv.y -= alpha * qv.y; rv.z -= alpha * qv.z; rv.w -= alpha * qv.w;
    reinterpret_cast<float4*>(e)[i] = ev;
    reinterpret_cast<float4*>(r)[i] = rv;
    acc[0] += ((double)rv.x * rv.x + (double)rv.y * rv.y) + ((double)rv.z * rv.z + (double)rv.w * rv.w);
    rmax = fmaxf(rmax, fmaxf(fmaxf(fabsf(rv.x), fabsf(rv.y)), fmaxf(fabsf(rv.z), fabsf(rv.w))));
  }
  if (blockIdx.x == 0 && threadIdx.x < (n & 3)) {
    const size_t i = n4 * 4 + threadIdx.x;
    e[i] += alpha * d[i];
    const float rn = r[i] - alpha * q[i];
    r[i] = rn;
    acc[0] += (double)rn * (double)rn;
    rmax = fmaxf(rmax, fabsf(rn));
  }
  block_partials_max<1>(acc, rmax, partials);
}

__global__ void k_fin_update(const double* __restrict__ partials, unsigned G, int max_iter, State* st) {
  if (st->done) return;
  double out[1];
  float m;
  sum_partials_max<1>(partials, G, out, m);
  if (threadIdx.x == 0) {
    const double rr = out[0];
    st->rr = rr;
    st->iter += 1;
    st->total_iter += 1;
    if (!isfinite(rr)) {
      st->flags |= 0x1;
      st->done = 1;
    } else if (rr <= st->inner_target2 || st->iter >= max_iter) {
      st->done = 1;
    }
    st->rscale = (m > 0.0f && isfinite(m)) ? exp2f(-ceilf(log2f(m))) : 1.0f;
  }
}

// --------------------------------------------------------------------------
// output: pressure and gradient subtract (FP64 weights and pressures)
// --------------------------------------------------------------------------
__global__ void k_pressure_out(size_t n, const float* __restrict__ phi, const double* __restrict__ x,
                               float* __restrict__ p) {
  GRID_STRIDE(i, n) p[i] = phi[i] < 0.0f ? (float)x[i] : 0.0f;
}

// axis 0: faces (nx+1, ny, nz); 1: (nx, ny+1, nz); 2: (nx, ny, nz+1).
__global__ void k_subtract(Grid3 g, int axis, const float* __restrict__ phi, const double* __restrict__ x,
                           const float* __restrict__ vin, float* __restrict__ vout, double scale) {
  const int fx = g.nx + (axis == 0), fy = g.ny + (axis == 1), fz = g.nz + (axis == 2);
  const size_t nf = (size_t)fx * fy * fz;
  GRID_STRIDE(f, nf) {
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

__global__ void k_div_after(Grid3 g, const float* __restrict__ phi, const float* __restrict__ u,
                            const float* __restrict__ v, const float* __restrict__ w, const float* __restrict__ q,
                            double inv_h, double* partials) {
  float dmax = 0.0f;
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  GRID_STRIDE(idx, g.n()) {
    if (!(phi[idx] < 0.0f)) continue;
    int i, j, k;
    g.ijk(idx, i, j, k);
    const double flux = ((double)u[((size_t)(i + 1) * ny + j) * nz + k] - (double)u[((size_t)i * ny + j) * nz + k]) +
                        ((double)v[((size_t)i * (ny + 1) + j + 1) * nz + k] - (double)v[((size_t)i * (ny + 1) + j) * nz + k]) +
                        ((double)w[((size_t)i * ny + j) * (nz + 1) + k + 1] - (double)w[((size_t)i * ny + j) * (nz + 1) + k]);
    dmax = fmaxf(dmax, (float)fabs(flux * inv_h - (q ? (double)q[idx] : 0.0)));
  }
  double none[1] = {0.0};
  block_partials_max<1>(none, dmax, partials);
}

__global__ void k_fin_divafter(const double* __restrict__ partials, unsigned G, State* st) {
  double out[1];
  float mx;
  sum_partials_max<1>(partials, G, out, mx);
  if (threadIdx.x == 0) st->divmax_after = __float_as_uint(mx);
}

// FP64 operator probe: out = A p.
__global__ void k_apply_d(Grid3 g, const float* __restrict__ phi, const double* __restrict__ p,
                          double* __restrict__ out) {
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  GRID_STRIDE(idx, g.n()) {
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
static inline unsigned grid_for(size_t n) {
  size_t b = (n + FP_BLOCK - 1) / FP_BLOCK;
  if (b < 1) b = 1;
  if (b > kMaxBlocks) b = kMaxBlocks;
  return (unsigned)b;
}

#define LAUNCH(kern, n, stream, ...) hipLaunchKernelGGL(kern, dim3(grid_for(n)), dim3(FP_BLOCK), 0, stream, __VA_ARGS__)
// LAUNCH with HIP events around it when `cls` is the kernel class being timed.
#define TLAUNCH(cls, kern, n, stream, ...)              \
  do {                                                  \
    const bool _t = (cls) == timed_kernel_;             \
    if (_t) timer_begin(stream);                        \
    LAUNCH(kern, n, stream, __VA_ARGS__);               \
    if (_t) timer_end(stream);                          \
  } while (0)

static LevelArgs largs(const UniformLevel& l) {
  LevelArgs a;
  a.g = Grid3{l.d[0], l.d[1], l.d[2]};
  a.gx = l.g[0];
  a.gy = l.g[1];
  a.gz = l.g[2];
  a.phi = l.phi;
  return a;
}

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

#define HIPCHK(expr)                                                   \
  do {                                                                 \
    hipError_t _e = (expr);                                            \
    if (_e != hipSuccess) {                                            \
      set_error("%s: %s", #expr, hipGetErrorString(_e));               \
      return FP_ERR_DEVICE;                                            \
    }                                                                  \
  } while (0)

UniformSolver::~UniformSolver() { release(); }

void UniformSolver::timer_begin(hipStream_t s) {
  if ((size_t)(2 * timer_used_ + 2) > timer_events_.size()) {
    hipEvent_t a, b;
    if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) return;
    timer_events_.push_back(a);
    timer_events_.push_back(b);
  }
  (void)hipEventRecord(timer_events_[2 * timer_used_], s);
}

void UniformSolver::timer_end(hipStream_t s) {
  if ((size_t)(2 * timer_used_ + 2) > timer_events_.size()) return;
  (void)hipEventRecord(timer_events_[2 * timer_used_ + 1], s);
  ++timer_used_;
}

void UniformSolver::release() {
  if (graph_exec_) (void)hipGraphExecDestroy(graph_exec_);
  graph_exec_ = nullptr;
  for (void* p : allocations_) (void)hipFree(p);
  allocations_.clear();
  for (void* p : pinned_) (void)hipHostFree(p);
  pinned_.clear();
  for (hipEvent_t e : timer_events_) (void)hipEventDestroy(e);
  timer_events_.clear();
}

int UniformSolver::alloc(void** p, size_t bytes) {
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

size_t UniformSolver::estimate_bytes(int nx, int ny, int nz) {
  int dims[kMaxLevels][3], strides[kMaxLevels][3];
  const int L = plan_uniform_levels(nx, ny, nz, dims, strides);
  if (L < 1) return 0;
  const size_t n = (size_t)nx * ny * nz;
  size_t total = 0;
  auto add = [&](size_t b) { total += (b + 255) & ~(size_t)255; };
  add(n * 8);  // b
  add(n * 8);  // x
  for (int i = 0; i < 5; ++i) add(n * 4);  // r e q d0 d1
  add(n * 4 * 3 + 4 * ((size_t)(nx + 1) * ny * nz)); // phi0 copy + u v w copies (upper bound)
  for (int l = 0; l < L; ++l) {
    const size_t nl = (size_t)dims[l][0] * dims[l][1] * dims[l][2];
    add(nl * 4);      // phi
    add(nl * 4 * 4);  // r e e2 t
    add(nl * 16);     // packed coefficients
  }
  add(512 * 512 * 8 + 4096);
  add(sizeof(State) + 8 * kMaxBlocks * 8);
  return total;
}

int UniformSolver::create(int nx, int ny, int nz, double h) {
  nx_ = nx; ny_ = ny; nz_ = nz; h_ = h;
  n_ = (size_t)nx * ny * nz;
  int dims[kMaxLevels][3], strides[kMaxLevels][3];
  L_ = plan_uniform_levels(nx, ny, nz, dims, strides);
  if (L_ < 1) {
    set_error("no multigrid hierarchy for %dx%dx%d", nx, ny, nz);
    return FP_ERR_LAYOUT;
  }
  int rc;
  const size_t nu = (size_t)(nx + 1) * ny * nz, nv = (size_t)nx * (ny + 1) * nz, nw = (size_t)nx * ny * (nz + 1);
  if ((rc = alloc((void**)&b_, n_ * 8))) return rc;
  if ((rc = alloc((void**)&x_, n_ * 8))) return rc;
  if ((rc = alloc((void**)&r_, n_ * 4))) return rc;
  if ((rc = alloc((void**)&e_, n_ * 4))) return rc;
  if ((rc = alloc((void**)&q_, n_ * 4))) return rc;
  if ((rc = alloc((void**)&d_[0], n_ * 4))) return rc;
  if ((rc = alloc((void**)&d_[1], n_ * 4))) return rc;
  if ((rc = alloc((void**)&in_u_, nu * 4))) return rc;
  if ((rc = alloc((void**)&in_v_, nv * 4))) return rc;
  if ((rc = alloc((void**)&in_w_, nw * 4))) return rc;
  if ((rc = alloc((void**)&in_q_, n_ * 4))) return rc;
  for (int l = 0; l < L_; ++l) {
    UniformLevel& lv = lv_[l];
    for (int a = 0; a < 3; ++a) {
      lv.d[a] = dims[l][a];
      lv.s[a] = strides[l][a];
      lv.g[a] = l == 0 ? 1.0f : lv_[l - 1].g[a] / (float)(lv.s[a] * lv.s[a]);
    }
    lv.n = (size_t)lv.d[0] * lv.d[1] * lv.d[2];
    if ((rc = alloc((void**)&lv.phi, lv.n * 4))) return rc;
    if ((rc = alloc(&lv.r, lv.n * 4))) return rc;
    if ((rc = alloc(&lv.e, lv.n * 4))) return rc;
    if ((rc = alloc(&lv.e2, lv.n * 4))) return rc;
    if ((rc = alloc(&lv.w, lv.n * 16))) return rc;
    if ((rc = alloc(&lv.t, lv.n * 4))) return rc;
  }
  const UniformLevel& c = lv_[L_ - 1];
  max_coarse_ = (int)std::min<size_t>(c.n, 512);
  if ((rc = alloc((void**)&cmap_, (size_t)max_coarse_ * 4 + 16))) return rc;
  if ((rc = alloc((void**)&ncoarse_, 16))) return rc;
  if ((rc = alloc((void**)&minv_, (size_t)max_coarse_ * max_coarse_ * 8 + 16))) return rc;
  if ((rc = alloc((void**)&st_, sizeof(State)))) return rc;
  if ((rc = alloc((void**)&partials_, 4 * (size_t)kMaxBlocks * 8))) return rc;
  HIPCHK(hipMemset(st_, 0, sizeof(State)));
  HIPCHK(hipMemset(x_, 0, n_ * 8));
  HIPCHK(hipHostMalloc((void**)&host_state_, sizeof(State), hipHostMallocDefault));
  pinned_.push_back(host_state_);
  HIPCHK(hipHostMalloc((void**)&host_upload_, sizeof(State), hipHostMallocDefault));
  pinned_.push_back(host_upload_);
  return FP_OK;
}

void UniformSolver::set_level_scales(int precision) {
  // FP32: every scale is 1.  FP16/BF16: the finest right-hand side arrives
  // normalised to |r| <= 1 (device rscale); error vectors grow like
  // A^-1 r, i.e. up to ~n^2 cells^2 on smooth modes, so each level stores
  // e / 2^(2 ceil(log2 n_l) - 8) to stay inside the FP16 range.
  for (int l = 0; l < L_; ++l) {
    UniformLevel& lv = lv_[l];
    if (precision == 0) {
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

// Fused sweeps run on levels large enough to fill a 256-lane column tile.
static bool fused_level(const UniformLevel& l) { return l.d[2] >= 16 && l.d[1] >= 4 && l.d[0] >= 4 && l.n >= 4096; }

#ifndef FP_SWEEP_DEPTH
#define FP_SWEEP_DEPTH 1
#endif
constexpr int kSweepDepth = FP_SWEEP_DEPTH;  // planes of loads in flight per workgroup

template <int TZ, typename TR, typename T>
static void launch_sweep_tz(hipStream_t s, const UniformLevel& f, const TR* r, float rmul, int dyn, const T* ein,
                            T* eout, int first, int zero_init, const ProlongArgs& P, const State* st) {
  constexpr int TY = 256 / TZ;
  const W4<T>* w = (const W4<T>*)f.w;
  const SweepGeom G{f.d[0], f.d[1], f.d[2], f.g[0], f.g[1], f.g[2]};
  const int tz = (f.d[2] + TZ - 1) / TZ, ty = (f.d[1] + TY - 1) / TY;
  const long tiles = (long)tz * ty;
  // about two rounds of 8 resident workgroups per CU on 256 CUs, at least 8
  // planes per workgroup (the march carries 4 planes of start-up)
  const long cx_target = std::max<long>(1, 4096 / tiles);
  int chunk = (int)std::max<long>(8, (f.d[0] + cx_target - 1) / cx_target);
  chunk = std::min(chunk, f.d[0]);
  const int cx = (f.d[0] + chunk - 1) / chunk;
  const int ng = (int)(tiles * cx);
  const dim3 grid(ng), block(256);
  if (zero_init)
    hipLaunchKernelGGL((k_sweep3<TZ, kSweepDepth, TR, T, false, true>), grid, block, 0, s, G, w, r, rmul, dyn, ein, eout, first, P,
                       chunk, tz, ty, ng, st);
  else if (P.ec)
    hipLaunchKernelGGL((k_sweep3<TZ, kSweepDepth, TR, T, true, false>), grid, block, 0, s, G, w, r, rmul, dyn, ein, eout, first, P,
                       chunk, tz, ty, ng, st);
  else
    hipLaunchKernelGGL((k_sweep3<TZ, kSweepDepth, TR, T, false, false>), grid, block, 0, s, G, w, r, rmul, dyn, ein, eout, first, P,
                       chunk, tz, ty, ng, st);
}

template <typename TR, typename T>
static void launch_sweep(hipStream_t s, const UniformLevel& f, const TR* r, float rmul, int dyn, const T* ein, T* eout,
                         int first, int zero_init, const ProlongArgs& P, const State* st) {
  if (f.d[2] >= 32) launch_sweep_tz<32, TR, T>(s, f, r, rmul, dyn, ein, eout, first, zero_init, P, st);
  else launch_sweep_tz<16, TR, T>(s, f, r, rmul, dyn, ein, eout, first, zero_init, P, st);
}

void UniformSolver::build_weights(hipStream_t s, int precision) {
  for (int l = 0; l < L_ - 1; ++l) {
    const UniformLevel& f = lv_[l];
    if (!fused_level(f)) continue;
    if (precision == FP_PRECISION_MIXED_FP16)
      LAUNCH(k_build_weights<__half>, f.n, s, f.d[0], f.d[1], f.d[2], f.g[0], f.g[1], f.g[2], (const float*)f.phi,
             (W4<__half>*)f.w);
    else if (precision == FP_PRECISION_MIXED_BF16)
      LAUNCH(k_build_weights<__hip_bfloat16>, f.n, s, f.d[0], f.d[1], f.d[2], f.g[0], f.g[1], f.g[2],
             (const float*)f.phi, (W4<__hip_bfloat16>*)f.w);
    else
      LAUNCH(k_build_weights<float>, f.n, s, f.d[0], f.d[1], f.d[2], f.g[0], f.g[1], f.g[2], (const float*)f.phi,
             (W4<float>*)f.w);
  }
}

template <typename T>
int UniformSolver::vcycle(hipStream_t s, int pre, int post) {
  // level 0's right-hand side is the FP32 CG residual times the device rscale.
  // cur[l] tracks which of the level's two correction buffers holds e.
  int cur[kMaxLevels] = {0};
  const ProlongArgs noP{nullptr, 0, 0, 0, 1, 1, 1, 1.0f};
  for (int l = 0; l < L_ - 1; ++l) {
    UniformLevel& f = lv_[l];
    UniformLevel& c = lv_[l + 1];
    const LevelArgs fa = largs(f);
    const bool top = (l == 0);
    const void* rr = top ? (const void*)r_ : f.r;
    const float rmul = (top ? 1.0f : f.Sr) / f.Se;  // r_true/Se in stored-e units
    const bool fused = fused_level(f);
    if (pre == 0) LAUNCH(k_fill<T>, f.n, s, (T*)f.e, f.n, T(0.0f));
    for (int sw = 0; sw < pre; ++sw) {
      if (fused) {
        T* ein = (T*)(cur[l] ? f.e2 : f.e);
        T* eout = (T*)(cur[l] ? f.e : f.e2);
        const bool t = top && timed_kernel_ == FP_KERNEL_SMOOTH_FINE;
        if (t) timer_begin(s);
        if (top) launch_sweep<float, T>(s, f, (const float*)rr, rmul, 1, ein, eout, 0, sw == 0, noP, st_);
        else launch_sweep<T, T>(s, f, (const T*)rr, rmul, 0, ein, eout, 0, sw == 0, noP, st_);
        if (t) timer_end(s);
        cur[l] ^= 1;
      } else {
        T* e = (T*)(cur[l] ? f.e2 : f.e);
        for (int colour = 0; colour < 2; ++colour) {
          const int zi = (sw == 0 && colour == 0) ? 1 : 0;
          if (top) LAUNCH((k_smooth<float, T>), f.n, s, fa, (const float*)rr, rmul, 1, e, colour, zi, (const State*)st_);
          else LAUNCH((k_smooth<T, T>), f.n, s, fa, (const T*)rr, rmul, 0, e, colour, zi, (const State*)st_);
        }
      }
    }
    T* e = (T*)(cur[l] ? f.e2 : f.e);
    // t = r - A e, stored with scale St
    const float r_t = (top ? 1.0f : f.Sr) / f.St;
    const float e_t = f.Se / f.St;
    if (fused) {
      const bool t = top && timed_kernel_ == FP_KERNEL_RESIDUAL_FINE;
      if (t) timer_begin(s);
      if (top) {
        OpResidualW<float, T> op{f.d[1], f.d[2], (const W4<T>*)f.w, (const float*)rr, (const T*)e, (T*)f.t, r_t, e_t,
                                 1.0f, 1, 0.0f, st_, nullptr};
        launch_march(s, f.d[0], f.d[1], f.d[2], op, st_, 1);
      } else {
        OpResidualW<T, T> op{f.d[1], f.d[2], (const W4<T>*)f.w, (const T*)rr, (const T*)e, (T*)f.t, r_t, e_t, 1.0f, 0,
                             0.0f, st_, nullptr};
        launch_march(s, f.d[0], f.d[1], f.d[2], op, st_, 1);
      }
      if (t) timer_end(s);
    } else if (top) {
      TLAUNCH(FP_KERNEL_RESIDUAL_FINE, (k_residual<float, T>), f.n, s, fa, (const float*)rr, r_t, 1, (const T*)e, e_t, 1.0f, (T*)f.t, (const State*)st_);
    } else {
      LAUNCH((k_residual<T, T>), f.n, s, fa, (const T*)rr, r_t, 0, (const T*)e, e_t, 1.0f, (T*)f.t, (const State*)st_);
    }
    const int k2 = (c.s[0] == 2) + (c.s[1] == 2) + (c.s[2] == 2);
    const float fac = (1.0f / (float)(1 << k2)) * f.St / c.Sr;
    TLAUNCH(top ? FP_KERNEL_RESTRICT_FINE : -1, k_restrict<T>, c.n, s, Grid3{f.d[0], f.d[1], f.d[2]}, (const T*)f.t,
            largs(c), c.s[0], c.s[1], c.s[2], fac, (T*)c.r, (const State*)st_);
  }
  {
    UniformLevel& c = lv_[L_ - 1];
    const bool top = (L_ == 1);
    const float rmul = top ? 1.0f : c.Sr;
    cur[L_ - 1] = 0;
    if (top)
      hipLaunchKernelGGL((k_coarse_solve<T, float>), dim3(1), dim3(FP_BLOCK), 0, s, Grid3{c.d[0], c.d[1], c.d[2]},
                         (const int*)cmap_, (const int*)ncoarse_, (const double*)minv_, (const float*)r_, rmul, 1,
                         (T*)c.e, 1.0f / c.Se, (const State*)st_);
    else
      hipLaunchKernelGGL((k_coarse_solve<T, T>), dim3(1), dim3(FP_BLOCK), 0, s, Grid3{c.d[0], c.d[1], c.d[2]},
                         (const int*)cmap_, (const int*)ncoarse_, (const double*)minv_, (const T*)c.r, rmul, 0,
                         (T*)c.e, 1.0f / c.Se, (const State*)st_);
  }
  for (int l = L_ - 2; l >= 0; --l) {
    UniformLevel& f = lv_[l];
    UniformLevel& c = lv_[l + 1];
    const LevelArgs fa = largs(f);
    const bool top = (l == 0);
    const void* rr = top ? (const void*)r_ : f.r;
    const float rmul = (top ? 1.0f : f.Sr) / f.Se;
    const bool fused = fused_level(f);
    const T* ec = (const T*)(cur[l + 1] ? c.e2 : c.e);
    if (!fused || post == 0) {
      T* e = (T*)(cur[l] ? f.e2 : f.e);
      TLAUNCH(top ? FP_KERNEL_PROLONG_FINE : -1, k_prolong_add<T>, f.n, s, fa, e, Grid3{c.d[0], c.d[1], c.d[2]},
              (const float*)c.phi, ec, c.s[0], c.s[1], c.s[2], c.Se / f.Se, (const State*)st_);
      for (int sw = 0; sw < post; ++sw)
        for (int cc = 0; cc < 2; ++cc) {
          const int colour = 1 - cc;
          if (top) LAUNCH((k_smooth<float, T>), f.n, s, fa, (const float*)rr, rmul, 1, e, colour, 0, (const State*)st_);
          else LAUNCH((k_smooth<T, T>), f.n, s, fa, (const T*)rr, rmul, 0, e, colour, 0, (const State*)st_);
        }
    } else {
      for (int sw = 0; sw < post; ++sw) {
        ProlongArgs P = noP;
        if (sw == 0) P = ProlongArgs{(const void*)ec, c.d[0], c.d[1], c.d[2], c.s[0], c.s[1], c.s[2], c.Se / f.Se};
        T* ein = (T*)(cur[l] ? f.e2 : f.e);
        T* eout = (T*)(cur[l] ? f.e : f.e2);
        const bool t = top && timed_kernel_ == FP_KERNEL_SMOOTH_FINE;
        if (t) timer_begin(s);
        if (top) launch_sweep<float, T>(s, f, (const float*)rr, rmul, 1, ein, eout, 1, 0, P, st_);
        else launch_sweep<T, T>(s, f, (const T*)rr, rmul, 0, ein, eout, 1, 0, P, st_);
        if (t) timer_end(s);
        cur[l] ^= 1;
      }
    }
  }
  zbuf_ = cur[0] ? lv_[0].e2 : lv_[0].e;
  return FP_OK;
}

template <typename T>
int UniformSolver::iteration(hipStream_t s, int pre, int post, int max_iter, int parity) {
  vcycle<T>(s, pre, post);
  const UniformLevel& f = lv_[0];
  const float zmul = f.Se;  // z_true * rscale = e_stored * Se
  const T* z = (const T*)zbuf_;
  LAUNCH(k_cg_rz<T>, n_ / 4 + 1, s, n_, (const float*)r_, z, zmul, st_, partials_);
  hipLaunchKernelGGL(k_fin_rz, dim3(1), dim3(FP_BLOCK), 0, s, (const double*)partials_, grid_for(n_ / 4 + 1), st_);
  {
    const bool t = timed_kernel_ == FP_KERNEL_CG_DIRECTION;
    if (t) timer_begin(s);
    OpDirection<T> op{ny_, nz_, f.phi, z, (const float*)d_[parity], d_[parity ^ 1], q_, zmul, 0.0f, 0.0f,
                      f.g[0], f.g[1], f.g[2], st_, partials_};
    const unsigned G = launch_march(s, nx_, ny_, nz_, op, st_, 1);
    if (t) timer_end(s);
    hipLaunchKernelGGL(k_fin_dq, dim3(1), dim3(FP_BLOCK), 0, s, (const double*)partials_, G, st_);
  }
  TLAUNCH(FP_KERNEL_CG_UPDATE, k_cg_update, n_ / 4 + 1, s, n_, (const float*)d_[parity ^ 1], (const float*)q_, e_, r_,
          st_, partials_);
  hipLaunchKernelGGL(k_fin_update, dim3(1), dim3(FP_BLOCK), 0, s, (const double*)partials_, grid_for(n_ / 4 + 1),
                     max_iter, st_);
  return FP_OK;
}

void UniformSolver::run_iteration(hipStream_t s, int precision, int pre, int post, int max_iter, int parity) {
  if (precision == FP_PRECISION_MIXED_FP16) iteration<__half>(s, pre, post, max_iter, parity);
  else if (precision == FP_PRECISION_MIXED_BF16) iteration<__hip_bfloat16>(s, pre, post, max_iter, parity);
  else iteration<float>(s, pre, post, max_iter, parity);
}

int UniformSolver::ensure_graph(hipStream_t s, int precision, int pre, int post, int max_iter) {
  if (graph_exec_ && graph_key_[0] == precision && graph_key_[1] == pre && graph_key_[2] == post &&
      graph_key_[3] == max_iter)
    return FP_OK;
  if (graph_exec_) {
    (void)hipGraphExecDestroy(graph_exec_);
    graph_exec_ = nullptr;
  }
  hipGraph_t g = nullptr;
  HIPCHK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
  run_iteration(s, precision, pre, post, max_iter, 0);
  run_iteration(s, precision, pre, post, max_iter, 1);
  HIPCHK(hipStreamEndCapture(s, &g));
  hipError_t e = hipGraphInstantiate(&graph_exec_, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (e != hipSuccess) {
    graph_exec_ = nullptr;
    set_error("hipGraphInstantiate: %s", hipGetErrorString(e));
    return FP_ERR_DEVICE;
  }
  graph_key_[0] = precision;
  graph_key_[1] = pre;
  graph_key_[2] = post;
  graph_key_[3] = max_iter;
  return FP_OK;
}

int UniformSolver::setup_levels(hipStream_t s, const float* phi0, int preserve) {
  // level 0 phi is a private copy so callers may reuse their buffer
  HIPCHK(hipMemcpyAsync(lv_[0].phi, phi0, n_ * 4, hipMemcpyDeviceToDevice, s));
  for (int l = 1; l < L_; ++l) {
    const UniformLevel& f = lv_[l - 1];
    UniformLevel& c = lv_[l];
    LAUNCH(k_downsample_phi, c.n, s, Grid3{f.d[0], f.d[1], f.d[2]}, (const float*)f.phi, Grid3{c.d[0], c.d[1], c.d[2]},
           c.phi, c.s[0], c.s[1], c.s[2], (l <= preserve) ? 1 : 0);
  }
  const UniformLevel& c = lv_[L_ - 1];
  hipLaunchKernelGGL(k_coarse_setup, dim3(1), dim3(FP_BLOCK), 0, s, largs(c), cmap_, ncoarse_, minv_, max_coarse_, st_);
  return FP_OK;
}

int UniformSolver::project(hipStream_t s, const fp_inputs* in, fp_outputs* out, const fp_options* opt,
                           fp_report* rep) {
  const size_t nu = (size_t)(nx_ + 1) * ny_ * nz_, nv = (size_t)nx_ * (ny_ + 1) * nz_,
               nw = (size_t)nx_ * ny_ * (nz_ + 1);
  const bool host = in->memory == FP_MEMORY_HOST;
  const float *phi = in->phi, *u = in->u, *v = in->v, *w = in->w, *q = in->target_divergence;
  timed_kernel_ = opt->timed_kernel;
  timer_used_ = 0;
  hipEvent_t ev0, ev1;
  HIPCHK(hipEventCreate(&ev0));
  HIPCHK(hipEventCreate(&ev1));
  HIPCHK(hipEventRecord(ev0, s));
  if (host) {
    // stage host inputs in the grid's resident buffers (level-0 phi doubles as
    // the phi staging buffer)
    HIPCHK(hipMemcpyAsync(lv_[0].phi, phi, n_ * 4, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(in_u_, u, nu * 4, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(in_v_, v, nv * 4, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(in_w_, w, nw * 4, hipMemcpyHostToDevice, s));
    if (q) HIPCHK(hipMemcpyAsync(in_q_, q, n_ * 4, hipMemcpyHostToDevice, s));
    phi = lv_[0].phi;
    u = in_u_; v = in_v_; w = in_w_;
    if (q) q = in_q_;
  }
  // reset scalars (fill kernels rather than memsets so the sequence is graph-safe)
  const double tol = opt->relative_tolerance;
  std::memset(host_upload_, 0, sizeof(State));
  host_upload_->rscale = 1.0f;
  HIPCHK(hipMemcpyAsync(st_, host_upload_, offsetof(State, counters), hipMemcpyHostToDevice, s));
  if (phi != lv_[0].phi) {
    if (setup_levels(s, phi, opt->phi_preserve_levels)) return FP_ERR_DEVICE;
  } else {
    for (int l = 1; l < L_; ++l) {
      const UniformLevel& f = lv_[l - 1];
      UniformLevel& c = lv_[l];
      LAUNCH(k_downsample_phi, c.n, s, Grid3{f.d[0], f.d[1], f.d[2]}, (const float*)f.phi,
             Grid3{c.d[0], c.d[1], c.d[2]}, c.phi, c.s[0], c.s[1], c.s[2], (l <= opt->phi_preserve_levels) ? 1 : 0);
    }
    const UniformLevel& c = lv_[L_ - 1];
    hipLaunchKernelGGL(k_coarse_setup, dim3(1), dim3(FP_BLOCK), 0, s, largs(c), cmap_, ncoarse_, minv_, max_coarse_,
                       st_);
  }
  const float* phi_l0 = lv_[0].phi;
  const Grid3 g{nx_, ny_, nz_};
  const double rho = in->rho, dt = in->dt;
  LAUNCH(k_rhs, n_, s, g, phi_l0, u, v, w, q, rho * h_ / dt, rho * h_ * h_ / dt, 1.0 / h_, tol * tol, b_, st_,
         partials_);
  hipLaunchKernelGGL(k_fin_rhs, dim3(1), dim3(FP_BLOCK), 0, s, (const double*)partials_, grid_for(n_), tol * tol, st_);
  if (!opt->warm_start) LAUNCH(k_fill<double>, n_, s, x_, n_, 0.0);
  LAUNCH(k_fill4, n_ / 4 + 1, s, e_, n_, 0.0f);
  LAUNCH(k_fill4, n_ / 4 + 1, s, d_[0], n_, 0.0f);
  set_level_scales(opt->precision);
  build_weights(s, opt->precision);
  int outer = 0, total = 0;
  int flags = 0;
  const int max_iter = std::max(1, opt->max_iterations);
  const int check = std::max(1, opt->check_every);
  bool converged = false;
  for (outer = 0; outer < std::max(1, opt->max_outer) + 1; ++outer) {
    {
      const bool t = timed_kernel_ == FP_KERNEL_TRUE_RESIDUAL;
      if (t) timer_begin(s);
      OpTrueResidual op{ny_, nz_, phi_l0, (const double*)b_, (const double*)x_, r_, st_, partials_};
      const unsigned G = launch_march(s, nx_, ny_, nz_, op, st_, 0);
      if (t) timer_end(s);
      hipLaunchKernelGGL(k_fin_true, dim3(1), dim3(FP_BLOCK), 0, s, (const double*)partials_, G, st_);
    }
    HIPCHK(hipMemcpyAsync(host_state_, st_, sizeof(State), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    if (host_state_->outer_done) {
      converged = !(host_state_->flags & 0x1);
      break;
    }
    if (outer == std::max(1, opt->max_outer)) break;
    const bool graphs = opt->graphs && timed_kernel_ == 0;
    if (graphs) {
      // one hipGraph holds two CG + V-cycle iterations (the d ping-pong);
      // kernels of finished iterations return at once, so replays past
      // convergence cost only their launch
      int rc = ensure_graph(s, opt->precision, opt->pre_sweeps, opt->post_sweeps, max_iter);
      if (rc) return rc;
      for (int it = 0; it < max_iter; it += 2) {
        HIPCHK(hipGraphLaunch(graph_exec_, s));
        if ((it + 2) % std::max(2, check + (check & 1)) == 0 || it + 2 >= max_iter) {
          HIPCHK(hipMemcpyAsync(host_state_, st_, sizeof(State), hipMemcpyDeviceToHost, s));
          HIPCHK(hipStreamSynchronize(s));
          if (host_state_->done) break;
        }
      }
    } else {
      int parity = 0;
      for (int it = 0; it < max_iter; ++it) {
        run_iteration(s, opt->precision, opt->pre_sweeps, opt->post_sweeps, max_iter, parity);
        parity ^= 1;
        if ((it + 1) % check == 0 || it + 1 == max_iter) {
          HIPCHK(hipMemcpyAsync(host_state_, st_, sizeof(State), hipMemcpyDeviceToHost, s));
          HIPCHK(hipStreamSynchronize(s));
          if (host_state_->done) break;
        }
      }
    }
    if (host_state_->flags & 0x7) {
      // a broken recurrence: keep what the FP64 iterate has (fail closed)
      flags |= host_state_->flags;
      LAUNCH(k_fill4, n_ / 4 + 1, s, e_, n_, 0.0f);
    }
    LAUNCH(k_accumulate, n_ / 4 + 1, s, n_, x_, e_);
    LAUNCH(k_fill4, n_ / 4 + 1, s, d_[0], n_, 0.0f);
  }
  total = host_state_->total_iter;
  flags |= host_state_->flags;
  if (!converged) flags |= 0x8;  // FP_FLAG_NONCONVERGENCE
  // outputs
  float* pout = out->pressure;
  float *uo = out->u, *vo = out->v, *wo = out->w;
  const bool host_out = host;
  float* dev_p = nullptr;
  float *dev_u = nullptr, *dev_v = nullptr, *dev_w = nullptr;
  if (host_out) {
    dev_u = in_u_; dev_v = in_v_; dev_w = in_w_;  // projected in place on the device copy
    dev_p = r_;  // reuse r as the FP32 pressure staging buffer
  } else {
    dev_u = uo; dev_v = vo; dev_w = wo; dev_p = pout;
  }
  const double sc = dt / (rho * h_);
  if (dev_u) LAUNCH(k_subtract, nu, s, g, 0, phi_l0, (const double*)x_, u, dev_u, sc);
  if (dev_v) LAUNCH(k_subtract, nv, s, g, 1, phi_l0, (const double*)x_, v, dev_v, sc);
  if (dev_w) LAUNCH(k_subtract, nw, s, g, 2, phi_l0, (const double*)x_, w, dev_w, sc);
  if (dev_u && dev_v && dev_w) {
    LAUNCH(k_div_after, n_, s, g, phi_l0, dev_u, dev_v, dev_w, q, 1.0 / h_, partials_);
    hipLaunchKernelGGL(k_fin_divafter, dim3(1), dim3(FP_BLOCK), 0, s, (const double*)partials_, grid_for(n_), st_);
  }
  if (pout && dev_p) LAUNCH(k_pressure_out, n_, s, n_, phi_l0, (const double*)x_, dev_p);
  HIPCHK(hipEventRecord(ev1, s));
  if (host_out) {
    if (uo) HIPCHK(hipMemcpyAsync(uo, dev_u, nu * 4, hipMemcpyDeviceToHost, s));
    if (vo) HIPCHK(hipMemcpyAsync(vo, dev_v, nv * 4, hipMemcpyDeviceToHost, s));
    if (wo) HIPCHK(hipMemcpyAsync(wo, dev_w, nw * 4, hipMemcpyDeviceToHost, s));
    if (pout) HIPCHK(hipMemcpyAsync(pout, dev_p, n_ * 4, hipMemcpyDeviceToHost, s));
  }
  HIPCHK(hipMemcpyAsync(host_state_, st_, sizeof(State), hipMemcpyDeviceToHost, s));
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
  if (rep) {
    std::memset(rep, 0, sizeof(*rep));
    rep->flags = flags;
    rep->iterations = total;
    rep->outer_iterations = outer;
    rep->levels = L_;
    rep->rhs_norm = std::sqrt(std::max(0.0, host_state_->bb));
    rep->relative_residual = host_state_->bb > 0 ? std::sqrt(std::max(0.0, host_state_->rr64) / host_state_->bb) : 0.0;
    unsigned a = host_state_->divmax_before, b = host_state_->divmax_after;
    float fa, fb;
    std::memcpy(&fa, &a, 4);
    std::memcpy(&fb, &b, 4);
    rep->max_div_before = fa;
    rep->max_div_after = fb;
    rep->device_ms = ms;
    rep->samples = (int64_t)n_;
    rep->liquid_samples = (int64_t)host_state_->liquid;
    rep->timed_kernel_ms = timed_ms;
    rep->timed_kernel_launches = timed_launches;
  }
  return FP_OK;
}

int UniformSolver::probe_apply(hipStream_t s, const float* phi_host, const double* p_host, double* out_host) {
  float* phi;
  double *p, *o;
  HIPCHK(hipMalloc(&phi, n_ * 4));
  HIPCHK(hipMalloc(&p, n_ * 8));
  HIPCHK(hipMalloc(&o, n_ * 8));
  HIPCHK(hipMemcpy(phi, phi_host, n_ * 4, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(p, p_host, n_ * 8, hipMemcpyHostToDevice));
  LAUNCH(k_apply_d, n_, s, Grid3{nx_, ny_, nz_}, (const float*)phi, (const double*)p, o);
  HIPCHK(hipStreamSynchronize(s));
  HIPCHK(hipMemcpy(out_host, o, n_ * 8, hipMemcpyDeviceToHost));
  (void)hipFree(phi); (void)hipFree(p); (void)hipFree(o);
  return FP_OK;
}

int UniformSolver::probe_rhs(hipStream_t s, const fp_inputs* in, double* b_host) {
  State init;
  std::memset(&init, 0, sizeof(init));
  HIPCHK(hipMemcpy(st_, &init, sizeof(State), hipMemcpyHostToDevice));
  const size_t nu = (size_t)(nx_ + 1) * ny_ * nz_, nv = (size_t)nx_ * (ny_ + 1) * nz_,
               nw = (size_t)nx_ * ny_ * (nz_ + 1);
  HIPCHK(hipMemcpy(lv_[0].phi, in->phi, n_ * 4, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(in_u_, in->u, nu * 4, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(in_v_, in->v, nv * 4, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(in_w_, in->w, nw * 4, hipMemcpyHostToDevice));
  const float* q = nullptr;
  if (in->target_divergence) {
    HIPCHK(hipMemcpy(in_q_, in->target_divergence, n_ * 4, hipMemcpyHostToDevice));
    q = in_q_;
  }
  const double rho = in->rho, dt = in->dt;
  LAUNCH(k_rhs, n_, s, Grid3{nx_, ny_, nz_}, (const float*)lv_[0].phi, (const float*)in_u_, (const float*)in_v_,
         (const float*)in_w_, q, rho * h_ / dt, rho * h_ * h_ / dt, 1.0 / h_, 0.0, b_, st_, partials_);
  hipLaunchKernelGGL(k_fin_rhs, dim3(1), dim3(FP_BLOCK), 0, s, (const double*)partials_, grid_for(n_), 0.0, st_);
  HIPCHK(hipStreamSynchronize(s));
  HIPCHK(hipMemcpy(b_host, b_, n_ * 8, hipMemcpyDeviceToHost));
  return FP_OK;
}

int UniformSolver::probe_vcycle(hipStream_t s, const float* phi_host, const fp_options* opt, const float* r_host,
                                float* z_host) {
  State init;
  std::memset(&init, 0, sizeof(init));
  init.rscale = 1.0f;
  HIPCHK(hipMemcpy(st_, &init, sizeof(State), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(lv_[0].phi, phi_host, n_ * 4, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(r_, r_host, n_ * 4, hipMemcpyHostToDevice));
  for (int l = 1; l < L_; ++l) {
    const UniformLevel& f = lv_[l - 1];
    UniformLevel& c = lv_[l];
    LAUNCH(k_downsample_phi, c.n, s, Grid3{f.d[0], f.d[1], f.d[2]}, (const float*)f.phi, Grid3{c.d[0], c.d[1], c.d[2]},
           c.phi, c.s[0], c.s[1], c.s[2], (l <= opt->phi_preserve_levels) ? 1 : 0);
  }
  {
    const UniformLevel& c = lv_[L_ - 1];
    hipLaunchKernelGGL(k_coarse_setup, dim3(1), dim3(FP_BLOCK), 0, s, largs(c), cmap_, ncoarse_, minv_, max_coarse_,
                       st_);
  }
  set_level_scales(opt->precision);
  build_weights(s, opt->precision);
  // normalise like the CG does: rscale = 2^-ceil(log2 max|r|)
  float m = 0.0f;
  for (size_t i = 0; i < n_; ++i) m = std::max(m, std::fabs(r_host[i]));
  const float rs = m > 0.0f ? std::exp2(-std::ceil(std::log2(m))) : 1.0f;
  HIPCHK(hipMemcpy(&st_->rscale, &rs, sizeof(float), hipMemcpyHostToDevice));
  const UniformLevel& f = lv_[0];
  std::vector<float> tmp(n_);
  if (opt->precision == FP_PRECISION_MIXED_FP16) {
    vcycle<__half>(s, opt->pre_sweeps, opt->post_sweeps);
    std::vector<__half> hz(n_);
    HIPCHK(hipStreamSynchronize(s));
    HIPCHK(hipMemcpy(hz.data(), zbuf_, n_ * 2, hipMemcpyDeviceToHost));
    for (size_t i = 0; i < n_; ++i) z_host[i] = __half2float(hz[i]) * f.Se / rs;
  } else if (opt->precision == FP_PRECISION_MIXED_BF16) {
    vcycle<__hip_bfloat16>(s, opt->pre_sweeps, opt->post_sweeps);
    std::vector<__hip_bfloat16> hz(n_);
    HIPCHK(hipStreamSynchronize(s));
    HIPCHK(hipMemcpy(hz.data(), zbuf_, n_ * 2, hipMemcpyDeviceToHost));
    for (size_t i = 0; i < n_; ++i) z_host[i] = __bfloat162float(hz[i]) * f.Se / rs;
  } else {
    vcycle<float>(s, opt->pre_sweeps, opt->post_sweeps);
    HIPCHK(hipStreamSynchronize(s));
    HIPCHK(hipMemcpy(tmp.data(), zbuf_, n_ * 4, hipMemcpyDeviceToHost));
    for (size_t i = 0; i < n_; ++i) z_host[i] = tmp[i] * f.Se / rs;
  }
  return FP_OK;
}

int UniformSolver::probe_level_phi(hipStream_t s, const float* phi_host, int level, int preserve, float* out_host) {
  if (level < 0 || level >= L_) {
    set_error("level %d out of range [0,%d)", level, L_);
    return FP_ERR_INVALID_ARGUMENT;
  }
  HIPCHK(hipMemcpy(lv_[0].phi, phi_host, n_ * 4, hipMemcpyHostToDevice));
  for (int l = 1; l <= level; ++l) {
    const UniformLevel& f = lv_[l - 1];
    UniformLevel& c = lv_[l];
    LAUNCH(k_downsample_phi, c.n, s, Grid3{f.d[0], f.d[1], f.d[2]}, (const float*)f.phi, Grid3{c.d[0], c.d[1], c.d[2]},
           c.phi, c.s[0], c.s[1], c.s[2], (l <= preserve) ? 1 : 0);
  }
  HIPCHK(hipStreamSynchronize(s));
  HIPCHK(hipMemcpy(out_host, lv_[level].phi, lv_[level].n * 4, hipMemcpyDeviceToHost));
  return FP_OK;
}

}  // namespace fp

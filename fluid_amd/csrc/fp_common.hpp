// fp_common.hpp -- device helpers shared by the projection kernels (gfx950).
#pragma once

#include <hip/hip_bf16.h>
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#define FP_BLOCK 256
#define FP_WAVES (FP_BLOCK / 64)

namespace fp {

// Ghost-fluid fraction, lib/core/cm12-numerics.ts:37-43
// (min(1, max(0.05, |phi_l| / max(|phi_l| + |phi_a|, eps)))), eps as the
// multigrid coefficient bake uses (webgpu-uniform-pressure-multigrid.wgsl.ts:90-92).
__device__ __forceinline__ float ghost_theta(float phi_l, float phi_a) {
  const float den = fmaxf(fabsf(phi_l) + fabsf(phi_a), 1e-9f);
  return fminf(1.0f, fmaxf(0.05f, fabsf(phi_l) / den));
}
__device__ __forceinline__ double ghost_theta_d(double phi_l, double phi_a) {
  const double den = fmax(fabs(phi_l) + fabs(phi_a), 1e-9);
  return fmin(1.0, fmax(0.05, fabs(phi_l) / den));
}

// h^2-scaled face coefficient between cells with level sets pc, pn:
// g between two liquid cells, g/theta across the free surface, 0 between two
// air cells (mgCoefficientRaw, webgpu-uniform-pressure-multigrid.wgsl.ts:93-115).
__device__ __forceinline__ float face_weight(float pc, float pn, float g) {
  const bool lc = pc < 0.0f, ln = pn < 0.0f;
  if (lc && ln) return g;
  if (!(lc || ln)) return 0.0f;
  return lc ? g / ghost_theta(pc, pn) : g / ghost_theta(pn, pc);
}
__device__ __forceinline__ double face_weight_d(float pc, float pn) {
  const bool lc = pc < 0.0f, ln = pn < 0.0f;
  if (lc && ln) return 1.0;
  if (!(lc || ln)) return 0.0;
  return lc ? 1.0 / ghost_theta_d(pc, pn) : 1.0 / ghost_theta_d(pn, pc);
}

// ---- explicit address spaces ------------------------------------------------
// Arrays reached through a grid record are generic pointers to the compiler,
// which would issue FLAT instructions for them; FLAT loads also count against
// the LDS counter, so an LDS-only barrier would wait for an in-flight global
// prefetch.  ldg/stg move the bits of a value through a global (address space
// 1) pointer, and cload reads a record from constant memory (address space 4)
// into scalar registers, readfirstlane keeping every word uniform.
#define FP_AS1 __attribute__((address_space(1)))
typedef unsigned int fp_u32x2 __attribute__((ext_vector_type(2)));
typedef unsigned int fp_u32x4 __attribute__((ext_vector_type(4)));
typedef float fp_f32x4 __attribute__((ext_vector_type(4)));
typedef double fp_f64x4 __attribute__((ext_vector_type(4)));

template <int N> struct RawOf;
template <> struct RawOf<2> { typedef unsigned short type; };
template <> struct RawOf<4> { typedef unsigned int type; };
template <> struct RawOf<8> { typedef fp_u32x2 type; };
template <> struct RawOf<16> { typedef fp_u32x4 type; };

template <typename T>
__device__ __forceinline__ T ldg(const T* p, size_t i) {
  typedef typename RawOf<sizeof(T)>::type R;
  const R r = ((const FP_AS1 R*)p)[i];
  return __builtin_bit_cast(T, r);
}
template <typename T>
__device__ __forceinline__ void stg(T* p, size_t i, T v) {
  typedef typename RawOf<sizeof(T)>::type R;
  ((FP_AS1 R*)p)[i] = __builtin_bit_cast(R, v);
}
__device__ __forceinline__ fp_f32x4 ldg4(const float* p, size_t i4) { return ((const FP_AS1 fp_f32x4*)p)[i4]; }
__device__ __forceinline__ void stg4(float* p, size_t i4, fp_f32x4 v) { ((FP_AS1 fp_f32x4*)p)[i4] = v; }

template <typename T>
__device__ __forceinline__ T cload(const T* p) {
  static_assert(sizeof(T) % 4 == 0, "records are whole dwords");

  typedef const __attribute__((address_space(4))) unsigned CU;
  CU* src = (CU*)(p);
  T out;
  unsigned* dst = reinterpret_cast<unsigned*>(&out);
#pragma unroll
  for (int i = 0; i < (int)(sizeof(T) / 4); ++i) dst[i] = __builtin_amdgcn_readfirstlane(src[i]);
  return out;
}

// Storage conversions of multigrid vectors.
template <typename T> struct Store;
template <> struct Store<float> {
  static __device__ __forceinline__ float ld(const float* p, size_t i) { return p[i]; }
  static __device__ __forceinline__ void st(float* p, size_t i, float v) { p[i] = v; }
  static __device__ __forceinline__ float gld(const float* p, size_t i) { return ldg(p, i); }
  static __device__ __forceinline__ void gst(float* p, size_t i, float v) { stg(p, i, v); }
};
template <> struct Store<__half> {
  static __device__ __forceinline__ float ld(const __half* p, size_t i) { return __half2float(p[i]); }
  static __device__ __forceinline__ void st(__half* p, size_t i, float v) { p[i] = __float2half_rn(v); }
  static __device__ __forceinline__ float gld(const __half* p, size_t i) { return __half2float(ldg(p, i)); }
  static __device__ __forceinline__ void gst(__half* p, size_t i, float v) { stg(p, i, __float2half_rn(v)); }
};
template <> struct Store<__hip_bfloat16> {
  static __device__ __forceinline__ float ld(const __hip_bfloat16* p, size_t i) { return __bfloat162float(p[i]); }
  static __device__ __forceinline__ void st(__hip_bfloat16* p, size_t i, float v) { p[i] = __float2bfloat16(v); }
  static __device__ __forceinline__ float gld(const __hip_bfloat16* p, size_t i) { return __bfloat162float(ldg(p, i)); }
  static __device__ __forceinline__ void gst(__hip_bfloat16* p, size_t i, float v) {
    stg(p, i, __float2bfloat16(v));
  }
};

// Wave (64 lanes) and block sums of doubles.
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  return v;
}
__device__ __forceinline__ float wave_max(float v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_down(v, o, 64));
  return v;
}

// Block-reduce NV doubles; the result is valid in thread 0.
template <int NV>
__device__ __forceinline__ void block_sum(double (&v)[NV]) {
  __shared__ double sh[NV][FP_WAVES];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    double s = wave_sum(v[k]);
    if (lane == 0) sh[k][wid] = s;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) {
      double s = 0.0;
      for (int w = 0; w < FP_WAVES; ++w) s += sh[k][w];
      v[k] = s;
    }
  }
  __syncthreads();
}

__device__ __forceinline__ double atomic_load_d(const double* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Workgroup barrier for LDS hand-offs only: waits for this wave's LDS
// operations and meets the other waves, without draining outstanding global
// loads (so a prefetch issued before it stays in flight).  The memory clobber
// keeps the compiler from moving LDS accesses across it.
__device__ __forceinline__ void lds_barrier() { asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory"); }

// The value of an FP32 expression rounded to FP32 in a register.  The
// compiler may otherwise fold a product and its conversion to FP16/BF16 into
// one mixed-precision FMA that rounds once (v_fma_mixlo_f16), at some sites
// and not at others; the smoothers round every update to FP32 first, so two
// kernels that compute the same sweep store the same bits.
__device__ __forceinline__ float f32_rounded(float x) {
  asm volatile("" : "+v"(x));
  return x;
}

// Two-kernel deterministic reduction.  Producers store one partial per
// workgroup of their grid (no fence: the kernel boundary publishes them); a
// one-workgroup-per-grid finisher sums them in a fixed order.  This keeps
// device-scope fences (an L2 write-back on CDNA's per-XCD caches) out of the
// bandwidth-bound kernels.  blk / nblk: the workgroup's index and count within
// its grid (a batched launch covers several grids).
template <int NV>
__device__ __forceinline__ void block_partials(double (&v)[NV], double* partials, int blk, int nblk) {
  block_sum<NV>(v);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) partials[(size_t)k * nblk + blk] = v[k];
  }
}

template <int NV>
__device__ __forceinline__ void sum_partials(const double* partials, unsigned G, double (&out)[NV]) {
  double s[NV];
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    // four independent chains keep several loads in flight per thread; the
    // order of the additions is fixed, so the sum stays deterministic
    const double* p = partials + (size_t)k * G;
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
    const unsigned step = blockDim.x;
    unsigned b = threadIdx.x;
    for (; b + 3 * step < G; b += 4 * step) {
      a0 += p[b];
      a1 += p[b + step];
      a2 += p[b + 2 * step];
      a3 += p[b + 3 * step];
    }
    for (; b < G; b += step) a0 += p[b];
    s[k] = (a0 + a1) + (a2 + a3);
  }
  block_sum<NV>(s);
#pragma unroll
  for (int k = 0; k < NV; ++k) out[k] = s[k];
}

// Block max of non-negative floats; valid in thread 0.
__device__ __forceinline__ float block_max(float m) {
  __shared__ float shm[FP_WAVES];
  m = wave_max(m);
  if ((threadIdx.x & 63) == 0) shm[threadIdx.x >> 6] = m;
  __syncthreads();
  float r = 0.0f;
  if (threadIdx.x == 0)
    for (int w = 0; w < FP_WAVES; ++w) r = fmaxf(r, shm[w]);
  __syncthreads();
  return r;
}

// As block_partials, plus a max stored after the NV sums (slot NV).  Maxima go
// through partials too: one device-scope atomic per wave on a single address
// serialises across the XCDs.
template <int NV>
__device__ __forceinline__ void block_partials_max(double (&v)[NV], float m, double* partials, int blk, int nblk) {
  const float bm = block_max(m);
  block_sum<NV>(v);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) partials[(size_t)k * nblk + blk] = v[k];
    partials[(size_t)NV * nblk + blk] = (double)bm;
  }
}

template <int NV>
__device__ __forceinline__ void sum_partials_max(const double* partials, unsigned G, double (&out)[NV], float& mx) {
  sum_partials<NV>(partials, G, out);
  float m0 = 0.0f, m1 = 0.0f;
  const double* p = partials + (size_t)NV * G;
  const unsigned step = blockDim.x;
  unsigned b = threadIdx.x;
  for (; b + step < G; b += 2 * step) {
    m0 = fmaxf(m0, (float)p[b]);
    m1 = fmaxf(m1, (float)p[b + step]);
  }
  for (; b < G; b += step) m0 = fmaxf(m0, (float)p[b]);
  mx = block_max(fmaxf(m0, m1));
}

// Fast h^2-scaled face coefficient for the FP32 solve and the multigrid:
// 1/theta = clamp(max(|l| + |a|, eps) / |l|, 1, 20) with a hardware reciprocal
// (one ulp), selected without branches.  The FP64 operator keeps the exact
// quotient of face_weight_d.
__device__ __forceinline__ float face_weight_fast(float pc, float pn, float g) {
  const bool lc = pc < 0.0f, ln = pn < 0.0f;
  const float l = lc ? pc : pn, a = lc ? pn : pc;
  const float al = fabsf(l);
  const float inv = fminf(20.0f, fmaxf(1.0f, fmaxf(al + fabsf(a), 1e-9f) * __builtin_amdgcn_rcpf(al)));
  const float w = (lc && ln) ? g : g * inv;
  return (lc || ln) ? w : 0.0f;
}

// 1-D taps of the cell-centred linear prolongation for fine index f along one
// axis (s: 1 when the axis is not coarsened).
__device__ __forceinline__ int ptaps(int f, int s, int nc, int* c, float* w) {
  if (s == 1) { c[0] = f; w[0] = 1.0f; return 1; }
  const int I = f >> 1;
  const int J = (f & 1) ? I + 1 : I - 1;
  if (J < 0 || J >= nc) { c[0] = I; w[0] = 1.0f; return 1; }
  c[0] = I; w[0] = 0.75f; c[1] = J; w[1] = 0.25f;
  return 2;
}

// Max of non-negative floats through their bit patterns (order preserving).
__device__ __forceinline__ void atomic_max_nonneg(unsigned* p, float v) {
  if (v > 0.0f) atomicMax(p, __float_as_uint(v));
}

// y = M x for the dense row-major FP64 n x n inverse of a coarsest level, on
// the matrix cores (v_mfma_f64_16x16x4f64).  Each wave of the calling
// workgroup takes 16-row blocks of M; per step of 4 columns, lane l supplies
// A = M[row0 + l % 16][k + l / 16] and, in column 0 of B (lanes l % 16 == 0),
// x[k + l / 16] (0 in the other columns), so column 0 of the 16 x 16
// accumulator is the block's slice of M x: lane l holds column l % 16, rows
// l / 16 + 4 i in its registers i = 0..3 (measured on gfx950 with an
// identity A and a B of distinct entries).  out(row, value) receives each row
// once.  The calling workgroup's size is a multiple of 64.
typedef double fp_f64x4 __attribute__((ext_vector_type(4)));
template <class Out>
__device__ __forceinline__ void gemv_f64_mfma(const double* __restrict__ M, int n, const double* __restrict__ x,
                                              Out&& out) {
  const int lane = (int)(threadIdx.x & 63), wave = (int)(threadIdx.x >> 6), nwaves = (int)(blockDim.x >> 6);
  const int r16 = lane & 15, kq = lane >> 4;
  const int nb = (n + 15) >> 4;
  for (int rb = wave; rb < nb; rb += nwaves) {
    fp_f64x4 acc = {0.0, 0.0, 0.0, 0.0};
    const int row = rb * 16 + r16;
    for (int k0 = 0; k0 < n; k0 += 4) {
      const int k = k0 + kq;
      const double a = (row < n && k < n) ? M[(size_t)row * n + k] : 0.0;
      const double b = (r16 == 0 && k < n) ? x[k] : 0.0;
      acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc, 0, 0, 0);
    }
    if (r16 == 0) {
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int r = rb * 16 + kq + 4 * i;
        if (r < n) out(r, acc[i]);
      }
    }
  }
}

}  // namespace fp

// fp_extrap.hip -- velocity extrapolation of a uniform grid on MI355X.
//
// Kernels restate lib/methods/uniform/webgpu-uniform-velocity-extrapolation.wgsl.ts
// (cited per kernel) and its schedule (…extrapolation.ts:312-392).  The
// narrow band runs only where it can change anything: a face whose lattice
// (L1) distance to every source face exceeds 3 never activates, since each
// Godunov step adds at least h/sqrt(3) and the band ends at 2h, so after the
// seed three dilations of the source masks and one compaction give the list
// of cells the passes visit; cells off the list keep their seed state, which
// the seed writes to both ping-pong sides.  Each pass is one launch over the
// list that returns at once when the previous pass left no face active (the
// reference's indirect dispatch chain).  The hierarchy and the seed stream
// the whole grid in XCD-contiguous chunks.  No contraction anywhere: results
// equal oracle/fp_oracle_extrap.c bit for bit.
#pragma clang fp contract(off)

#include "fp_extrap.hpp"
#include "fp_uniform.hpp"

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstring>

namespace fp {

namespace {

constexpr float kInf = 65504.0f;  // DISTANCE_INFINITY (…extrapolation.wgsl.ts:61)
constexpr int kThreads = 256;

// Division by a launch-invariant divisor: floor(n / d) = (n * m) >> s for
// every n < 2^31 with s = 31 + ceil(log2 d), m = ceil(2^s / d) < 2^32
// (Granlund and Montgomery, "Division by invariant integers using
// multiplication", PLDI 1994, Thm 4.2).  Cell indices stay below 2^31.
struct FastDiv {
  unsigned m;
  int s;
};
FastDiv make_div(unsigned d) {
  int l = 0;
  while ((1ull << l) < d) ++l;
  const unsigned long long m = ((1ull << (31 + l)) + d - 1) / d;
  return FastDiv{(unsigned)m, 31 + l};
}
__device__ __forceinline__ unsigned fdiv(unsigned n, FastDiv f) {
  return (unsigned)(((unsigned long long)n * f.m) >> f.s);
}

struct ExGeo {
  int d[3];
  float h;
  int open_top;
  float eps_min, eps_max;  // fimConvergenceEpsilon's clamp (…wgsl.ts:114-119)
  FastDiv dz, dy;          // division by nz and by ny
};

__device__ __forceinline__ unsigned eidx(const ExGeo& g, int x, int y, int z) {
  return ((unsigned)x * g.d[1] + y) * g.d[2] + z;
}
__device__ __forceinline__ bool ein(const int* d, int x, int y, int z) {
  return x >= 0 && y >= 0 && z >= 0 && x < d[0] && y < d[1] && z < d[2];
}
// cell index -> (x, y, z) of a lattice with divisors dz = nz, dy = ny
__device__ __forceinline__ void decode3(unsigned i, int ny, int nz, FastDiv dz, FastDiv dy, int& x, int& y, int& z) {
  const unsigned t = fdiv(i, dz);
  z = (int)(i - t * (unsigned)nz);
  const unsigned xx = fdiv(t, dy);
  y = (int)(t - xx * (unsigned)ny);
  x = (int)xx;
}
__device__ __forceinline__ void edecode(const ExGeo& g, unsigned i, int& x, int& y, int& z) {
  decode3(i, g.d[1], g.d[2], g.dz, g.dy, x, y, z);
}
// positive face of cell (x,y,z) along c in the (n_c + 1)-long face array
__device__ __forceinline__ unsigned face_of(const ExGeo& g, int c, int x, int y, int z) {
  return c == 0 ? ((unsigned)(x + 1) * g.d[1] + y) * g.d[2] + z
       : c == 1 ? ((unsigned)x * (g.d[1] + 1) + y + 1) * g.d[2] + z
                : ((unsigned)x * g.d[1] + y) * (g.d[2] + 1) + z + 1;
}
// XCD-contiguous chunk of a streaming launch: workgroup b runs on XCD b mod 8
__device__ __forceinline__ unsigned chunk_of(unsigned chunks) {
  const unsigned per = (chunks + 7) / 8;
  return (blockIdx.x & 7) * per + (blockIdx.x >> 3);
}
__device__ __forceinline__ float eps_of(const ExGeo& g, float reference) {
  float a = fabsf(reference);
  a = a < g.eps_min ? g.eps_min : (a > g.eps_max ? g.eps_max : a);
  return a * 1.1920929e-7f;
}
__device__ __forceinline__ int block_sum(int v) {
  __shared__ int part[kThreads / 64];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = v;
  __syncthreads();
  int s = 0;
  if (threadIdx.x == 0)
    for (int k = 0; k < kThreads / 64; ++k) s += part[k];
  return s;
}

// ---- classes: open (bits 0-2) and source (bits 3-5) per component ----
// faceOpenFraction (webgpu-uniform-reference.wgsl.ts:511-519) without solids,
// sourceFace (…extrapolation.wgsl.ts:100-104)
// Streaming kernels take kCells cells per thread (kCells x kThreads per
// workgroup, the cells of one pass kThreads apart) and issue every load of
// their cells before the first store: neighbour loads are clamped to the cell
// itself at the domain edge and masked afterwards, so no branch separates
// them.  One cell per thread, or loads behind bounds branches, left these
// kernels waiting on one memory round trip at a time.
constexpr int kCells = 4;

__device__ __forceinline__ unsigned cell_of(unsigned n, int j) {
  const unsigned per_group = kThreads * kCells;
  return chunk_of((n + per_group - 1) / per_group) * per_group + j * kThreads + threadIdx.x;
}

// exclusive prefix of v over the workgroup, and the workgroup total
__device__ __forceinline__ int block_scan(int v, int* total) {
  __shared__ int wsum[kThreads / 64];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  int x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  if (lane == 63) wsum[wid] = x;
  __syncthreads();
  int before = 0, all = 0;
#pragma unroll
  for (int k = 0; k < kThreads / 64; ++k) {
    before += k < wid ? wsum[k] : 0;
    all += wsum[k];
  }
  *total = all;
  return before + x - v;
}

__global__ __launch_bounds__(kThreads) void k_ex_classify(ExGeo g, const float* __restrict__ field, int is_phi,
                                                         uint8_t* __restrict__ cls, uint8_t* __restrict__ cand,
                                                         unsigned n) {
  const unsigned sx = (unsigned)g.d[1] * g.d[2], sy = (unsigned)g.d[2];
  float f[kCells][4];
  int pos[kCells][3];
#pragma unroll
  for (int j = 0; j < kCells; ++j) {
    const unsigned i = min(cell_of(n, j), n - 1);
    int x, y, z;
    edecode(g, i, x, y, z);
    pos[j][0] = x;
    pos[j][1] = y;
    pos[j][2] = z;
    f[j][0] = field[i];
    f[j][1] = field[i + (x + 1 < g.d[0] ? sx : 0u)];
    f[j][2] = field[i + (y + 1 < g.d[1] ? sy : 0u)];
    f[j][3] = field[i + (z + 1 < g.d[2] ? 1u : 0u)];
  }
#pragma unroll
  for (int j = 0; j < kCells; ++j) {
    const unsigned i = cell_of(n, j);
    if (i >= n) break;
    bool wet[4];
    for (int k = 0; k < 4; ++k) wet[k] = is_phi ? f[j][k] < 0.0f : f[j][k] > 0.5f;
    unsigned b = 0;
    for (int c = 0; c < 3; ++c) {
      const bool inside = pos[j][c] + 1 < g.d[c];
      const bool open = inside || (c == 1 && g.open_top);
      if (!open) continue;
      b |= 1u << c;
      if (wet[0] || (inside && wet[1 + c])) b |= 8u << c;
    }
    cls[i] = (uint8_t)b;
    cand[i] = (uint8_t)(b >> 3);
  }
}

// bits 0-2 open(c), 3-5 source(c)
__device__ __forceinline__ unsigned cls_at(const ExGeo& g, const uint8_t* cls, int x, int y, int z) {
  return ein(g.d, x, y, z) ? cls[eidx(g, x, y, z)] : 0u;
}

// the byte of each 6-neighbour (0 outside the domain), loaded without branches
__device__ __forceinline__ void load6(const ExGeo& g, const uint8_t* __restrict__ a, unsigned i, int x, int y, int z,
                                      unsigned* out) {
  const unsigned sx = (unsigned)g.d[1] * g.d[2], sy = (unsigned)g.d[2];
  const bool ok[6] = {x > 0, x + 1 < g.d[0], y > 0, y + 1 < g.d[1], z > 0, z + 1 < g.d[2]};
  const unsigned idx[6] = {ok[0] ? i - sx : i, ok[1] ? i + sx : i, ok[2] ? i - sy : i,
                           ok[3] ? i + sy : i, ok[4] ? i - 1 : i,  ok[5] ? i + 1 : i};
#pragma unroll
  for (int k = 0; k < 6; ++k) out[k] = a[idx[k]];
#pragma unroll
  for (int k = 0; k < 6; ++k) out[k] = ok[k] ? out[k] : 0u;
}

// seedActiveFront (…extrapolation.wgsl.ts:148-170).  Side A everywhere; side
// B only where a pass can read it before writing it: the band list's cells
// and their 6-neighbours, all within one cell of the last candidate
// dilation (near: that mask; a pass writes every list cell of its
// destination side, and reads its source side at list cells and their
// neighbours only).
__global__ __launch_bounds__(kThreads) void k_ex_seed(ExGeo g, const uint8_t* __restrict__ cls,
                                                     const uint8_t* __restrict__ near,
                                                     const float* __restrict__ u, const float* __restrict__ v,
                                                     const float* __restrict__ w, ExState a, ExState b,
                                                     ExCounters* cnt, unsigned n) {
  unsigned me[kCells], nb[kCells][6], nr[kCells];
  float in[kCells][3];
#pragma unroll
  for (int j = 0; j < kCells; ++j) {
    const unsigned i = min(cell_of(n, j), n - 1);
    int x, y, z;
    edecode(g, i, x, y, z);
    me[j] = cls[i];
    load6(g, cls, i, x, y, z, nb[j]);
    unsigned m6[6];
    load6(g, near, i, x, y, z, m6);
    nr[j] = near[i] | m6[0] | m6[1] | m6[2] | m6[3] | m6[4] | m6[5];
    // the positive face of every cell exists in the face arrays
    in[j][0] = u[face_of(g, 0, x, y, z)];
    in[j][1] = v[face_of(g, 1, x, y, z)];
    in[j][2] = w[face_of(g, 2, x, y, z)];
  }
  int active = 0;
#pragma unroll
  for (int j = 0; j < kCells; ++j) {
    const unsigned i = cell_of(n, j);
    if (i >= n) break;
    unsigned srcnb = 0;  // per component: a 6-neighbour is a source
    for (int k = 0; k < 6; ++k) srcnb |= nb[j][k];
    srcnb >>= 3;
    unsigned flags = 0;
    for (int c = 0; c < 3; ++c) {
      float val = 0.0f, dist = kInf;
      if (me[j] >> c & 1) {
        if (me[j] >> (3 + c) & 1) {
          val = in[j][c];
          dist = 0.0f;
          flags |= 1u << c;
        } else if (srcnb >> c & 1) {
          flags |= 8u << c;
          active += 1;
        }
      }
      a.val[c][i] = val;
      a.dist[c][i] = dist;
      if (nr[j]) {
        b.val[c][i] = val;
        b.dist[c][i] = dist;
      }
    }
    a.flags[i] = (uint8_t)flags;
    if (nr[j]) b.flags[i] = (uint8_t)flags;
  }
  const int s = block_sum(active);
  if (threadIdx.x == 0 && s) atomicAdd(&cnt->active[0], s);
}

// one 6-neighbour OR of the candidate bits; the last one appends the cells
// with an open non-source candidate face to the band list (one atomic per
// workgroup) and also stores its mask (the seed's near set)
__global__ __launch_bounds__(kThreads) void k_ex_dilate(ExGeo g, const uint8_t* __restrict__ cls,
                                                       const uint8_t* __restrict__ in, uint8_t* __restrict__ out,
                                                       int last, int* __restrict__ list, ExCounters* cnt,
                                                       unsigned n) {
  unsigned m[kCells], me[kCells];
#pragma unroll
  for (int j = 0; j < kCells; ++j) {
    const unsigned i = min(cell_of(n, j), n - 1);
    int x, y, z;
    edecode(g, i, x, y, z);
    unsigned nb[6];
    load6(g, in, i, x, y, z, nb);
    m[j] = in[i] | nb[0] | nb[1] | nb[2] | nb[3] | nb[4] | nb[5];
    me[j] = last ? cls[i] : 0u;
  }
#pragma unroll
  for (int j = 0; j < kCells; ++j) {
    const unsigned i = cell_of(n, j);
    if (i < n) out[i] = (uint8_t)m[j];
  }
  if (!last) return;
  unsigned takes = 0;
#pragma unroll
  for (int j = 0; j < kCells; ++j)
    if (cell_of(n, j) < n && (m[j] & (me[j] & 7u) & ~(me[j] >> 3)) != 0) takes |= 1u << j;
  int total;
  int at = block_scan(__popc(takes), &total);
  __shared__ int base;
  if (threadIdx.x == 0) base = total ? atomicAdd(&cnt->list_count, total) : 0;
  __syncthreads();
  at += base;
#pragma unroll
  for (int j = 0; j < kCells; ++j)
    if (takes >> j & 1) list[at++] = (int)cell_of(n, j);
}

// ---- the band pass (updateActiveFront and helpers, …extrapolation.wgsl.ts:172-331) ----
// Neighbour state during a pass (the pass's source side).  Every accessor
// loads at an index clamped into the domain and selects afterwards, so the
// loads of a stencil are issued together instead of one branch at a time.
struct Nb {
  const ExGeo& g;
  const uint8_t* cls;
  const ExState& s;
  __device__ unsigned at(int x, int y, int z, bool* inside) const {
    *inside = ein(g.d, x, y, z);
    return eidx(g, min(max(x, 0), g.d[0] - 1), min(max(y, 0), g.d[1] - 1), min(max(z, 0), g.d[2] - 1));
  }
  __device__ bool open(int x, int y, int z, int c) const {
    bool in;
    const unsigned i = at(x, y, z, &in);
    return in && (cls[i] >> c & 1);
  }
  __device__ bool source(int x, int y, int z, int c) const {
    bool in;
    const unsigned i = at(x, y, z, &in);
    return in && (cls[i] >> (3 + c) & 1);
  }
  __device__ float dist(int x, int y, int z, int c) const {
    bool in;
    const unsigned i = at(x, y, z, &in);
    const unsigned cl = cls[i];
    const float d = s.dist[c][i];
    return in && (cl >> c & 1) ? d : kInf;
  }
  __device__ float value(int x, int y, int z, int c) const {
    bool in;
    const unsigned i = at(x, y, z, &in);
    const unsigned cl = cls[i];
    const unsigned fl = s.flags[i];
    const float v = s.val[c][i];
    return in && (cl >> c & 1) && (fl >> c & 1) ? v : 0.0f;
  }
  __device__ unsigned flags(int x, int y, int z) const {
    bool in;
    const unsigned i = at(x, y, z, &in);
    const unsigned fl = s.flags[i];
    return in ? fl : 0u;
  }
};

__device__ __forceinline__ void step(int a, int side, int& x, int& y, int& z) {
  if (a == 0) x += side;
  else if (a == 1) y += side;
  else z += side;
}

// godunovDistance (…extrapolation.wgsl.ts:188-215), cubic cells
__device__ float godunov(const Nb& nb, int x, int y, int z, int c) {
  const float h = nb.g.h;
  float m[3];
  for (int a = 0; a < 3; ++a) {
    int lx = x, ly = y, lz = z, hx = x, hy = y, hz = z;
    step(a, -1, lx, ly, lz);
    step(a, 1, hx, hy, hz);
    const float dl = nb.dist(lx, ly, lz, c), dh = nb.dist(hx, hy, hz, c);
    m[a] = dl < dh ? dl : dh;
  }
  for (int i = 0; i < 2; ++i)
    for (int j = i + 1; j < 3; ++j)
      if (m[j] < m[i]) {
        const float t = m[i];
        m[i] = m[j];
        m[j] = t;
      }
  if (m[0] >= 0.5f * kInf) return kInf;
  float root = m[0] + h;
  for (int count = 2; count <= 3; ++count) {
    if (m[count - 1] >= 0.5f * kInf || root <= m[count - 1]) break;
    float A = 0.0f, B = 0.0f, C = -1.0f;
    for (int k = 0; k < count; ++k) {
      const float inv = 1.0f / (h * h);
      A += inv;
      B += m[k] * inv;
      C += m[k] * m[k] * inv;
    }
    const float disc = B * B - A * C;
    root = (B + sqrtf(disc > 0.0f ? disc : 0.0f)) / A;
  }
  return root;
}

// upwindExtensionValue (…extrapolation.wgsl.ts:221-248)
__device__ float upwind(const Nb& nb, int x, int y, int z, int c, float solved) {
  const float h = nb.g.h;
  float wt[3], ws[3];
  const float eps = eps_of(nb.g, solved);
  for (int a = 0; a < 3; ++a) {
    wt[a] = 0.0f;
    ws[a] = 0.0f;
    int lx = x, ly = y, lz = z, hx = x, hy = y, hz = z;
    step(a, -1, lx, ly, lz);
    step(a, 1, hx, hy, hz);
    const float dl = nb.dist(lx, ly, lz, c), dh = nb.dist(hx, hy, hz, c);
    const float mn = dl < dh ? dl : dh;
    if (mn >= solved - eps) continue;
    float mv = 0.0f, mc = 0.0f;
    if (fabsf(dl - mn) <= eps) {
      mv += nb.value(lx, ly, lz, c);
      mc += 1.0f;
    }
    if (fabsf(dh - mn) <= eps) {
      mv += nb.value(hx, hy, hz, c);
      mc += 1.0f;
    }
    if (mc <= 0.0f) continue;
    const float weight = (solved - mn) / (h * h);
    wt[a] = weight * mv / mc;
    ws[a] = weight;
  }
  const float value = (wt[0] + wt[2]) + wt[1], sum = (ws[0] + ws[2]) + ws[1];
  return sum > 0.0f ? value / sum : 0.0f;
}

// activeNodeConverges (…extrapolation.wgsl.ts:253-265)
__device__ bool converges(const Nb& nb, int x, int y, int z, int c, float* updated) {
  if (!nb.open(x, y, z, c) || nb.source(x, y, z, c)) return false;
  if (!(nb.flags(x, y, z) >> (3 + c) & 1)) return false;
  const float old = nb.dist(x, y, z, c);
  if (old >= 0.5f * kInf) return false;
  const float gd = godunov(nb, x, y, z, c);
  const float upd = old < gd ? old : gd;
  if (upd > 2.0f * nb.g.h) return false;
  *updated = upd;
  return fabsf(upd - old) <= eps_of(nb.g, old);
}

// activatedByConvergedUpwindNeighbor (…extrapolation.wgsl.ts:266-283); only
// called when a neighbour is active, the first condition of activeNodeConverges
__device__ bool activated(const Nb& nb, int x, int y, int z, int c, float own) {
  const float cand = godunov(nb, x, y, z, c);
  if (cand > 2.0f * nb.g.h) return false;
  const float eps = eps_of(nb.g, own);
  for (int a = 0; a < 3; ++a)
    for (int side = -1; side <= 1; side += 2) {
      int qx = x, qy = y, qz = z;
      step(a, side, qx, qy, qz);
      float un;
      if (converges(nb, qx, qy, qz, c, &un) && own > un + eps) return true;
    }
  return false;
}

__global__ __launch_bounds__(kThreads) void k_ex_pass(ExGeo g, const uint8_t* __restrict__ cls,
                                                     const int* __restrict__ list, int count, ExState src,
                                                     ExState dst, ExCounters* cnt, int pass) {
  if (cnt->active[pass - 1] == 0) return;  // the previous pass left the front empty
  if (blockIdx.x == 0 && threadIdx.x == 0) cnt->latest = pass & 1;
  const int k = blockIdx.x * kThreads + threadIdx.x;
  int active = 0;
  if (k < count) {
    const unsigned i = (unsigned)list[k];
    int x, y, z;
    edecode(g, i, x, y, z);
    const Nb nb{g, cls, src};
    const unsigned me = cls[i];
    unsigned flags = src.flags[i];
    // active bits of the 6 neighbours per component: an inactive face with no
    // active neighbour cannot be activated
    unsigned nbact = 0;
    for (int a = 0; a < 3; ++a)
      for (int side = -1; side <= 1; side += 2) {
        int qx = x, qy = y, qz = z;
        step(a, side, qx, qy, qz);
        nbact |= nb.flags(qx, qy, qz);
      }
    nbact >>= 3;
    for (int c = 0; c < 3; ++c) {
      float val = src.val[c][i], dist = src.dist[c][i];
      if ((me >> c & 1) && !(me >> (3 + c) & 1)) {
        if (flags >> (3 + c) & 1) {
          const float old = dist;
          const float eps = eps_of(g, old);
          const float cand = godunov(nb, x, y, z, c);
          const float upd = old < cand ? old : cand;
          if (upd <= 2.0f * g.h) {
            dist = upd;
            val = upwind(nb, x, y, z, c, upd);
            flags |= 1u << c;
          }
          const bool conv = old < 0.5f * kInf && fabsf(upd - old) <= eps;
          if (!conv && upd <= 2.0f * g.h) flags |= 8u << c;
          else flags &= ~(8u << c);
        } else if ((nbact >> c & 1) && activated(nb, x, y, z, c, dist)) {
          flags |= 8u << c;  // JRW07: a newly added point is updated from the next pass
        }
      }
      dst.val[c][i] = val;
      dst.dist[c][i] = dist;
    }
    dst.flags[i] = (uint8_t)flags;
    active = __popc(flags >> 3 & 7u);
  }
  const int s = block_sum(active);
  if (threadIdx.x == 0 && s) atomicAdd(&cnt->active[pass], s);
}

// ---- hierarchy (…extrapolation.wgsl.ts:385-488) ----
struct ExSrc {
  int d[3];
  const float* val[3];
  const uint8_t* known;
};

// hierarchyComponentSample: renormalised trilinear sample of known values,
// the component axis face-centred; returns the weight sum
__device__ float hier_sample(const ExSrc& S, const int* T, const int* p, int c, float* value) {
  float pos[3];
  for (int a = 0; a < 3; ++a) pos[a] = ((float)p[a] + 0.5f) * (float)S.d[a] / (float)T[a] - 0.5f;
  pos[c] = (float)(p[c] + 1) * (float)S.d[c] / (float)T[c] - 1.0f;
  int lower[3];
  float fr[3];
  for (int a = 0; a < 3; ++a) {
    const float f = floorf(pos[a]);
    lower[a] = (int)f;
    fr[a] = pos[a] - f;
  }
  float cx[8], cw[8];
#pragma unroll
  for (int corner = 0; corner < 8; ++corner) {
    const int off[3] = {corner & 1, corner >> 1 & 1, corner >> 2 & 1};
    int q[3];
    float wgt[3];
    for (int a = 0; a < 3; ++a) {
      q[a] = min(max(lower[a] + off[a], 0), S.d[a] - 1);
      wgt[a] = off[a] ? fr[a] : 1.0f - fr[a];
    }
    const float weight = wgt[0] * wgt[1] * wgt[2];
    const unsigned i = ((unsigned)q[0] * S.d[1] + q[1]) * S.d[2] + q[2];
    // load every corner (the clamped index is always valid) and select:
    // branch-free, so the eight corners' loads are in flight together
    const unsigned kn = S.known[i];
    const float val = S.val[c][i];
    const bool use = weight > 0.0f && (kn >> c & 1);
    cx[corner] = use ? weight * val : 0.0f;
    cw[corner] = use ? weight : 0.0f;
  }
  const float xs = ((cx[0] + cx[5]) + (cx[1] + cx[4])) + ((cx[2] + cx[7]) + (cx[3] + cx[6]));
  const float wsum = ((cw[0] + cw[5]) + (cw[1] + cw[4])) + ((cw[2] + cw[7]) + (cw[3] + cw[6]));
  *value = wsum > 0.0f ? xs / wsum : 0.0f;
  return wsum;
}

__device__ __forceinline__ void known_at(const ExSrc& S, int x, int y, int z, int c, float* v, float* w) {
  *v = 0.0f;
  *w = 0.0f;
  if (!ein(S.d, x, y, z)) return;
  const unsigned i = ((unsigned)x * S.d[1] + y) * S.d[2] + z;
  if (S.known[i] >> c & 1) {
    *v = S.val[c][i];
    *w = 1.0f;
  }
}

// hierarchyCorrespondingCellSample (…wgsl.ts:430-447)
__device__ float corresponding_sample(const ExSrc& S, const int* p, int c, float* value) {
  const int lx = 2 * p[0], ly = 2 * p[1], lz = 2 * p[2];
  float sx = 0.0f, sw = 0.0f;
  for (int oy = 0; oy <= 1; ++oy) {
    float v00, w00, v11, w11, v10, w10, v01, w01;
    known_at(S, lx, ly + oy, lz, c, &v00, &w00);
    known_at(S, lx + 1, ly + oy, lz + 1, c, &v11, &w11);
    known_at(S, lx + 1, ly + oy, lz, c, &v10, &w10);
    known_at(S, lx, ly + oy, lz + 1, c, &v01, &w01);
    sx += (v00 + v11) + (v10 + v01);
    sw += (w00 + w11) + (w10 + w01);
  }
  *value = sw > 0.0f ? sx / sw : 0.0f;
  return sw;
}

// resolveConvergedFront (…wgsl.ts:359-369): the band cells' final state,
// on the side the last pass wrote, back into side A, which then holds the
// band result of every cell
__global__ __launch_bounds__(kThreads) void k_ex_resolve(const int* __restrict__ list, int count, ExState a,
                                                        ExState b, const ExCounters* cnt) {
  const int k = blockIdx.x * kThreads + threadIdx.x;
  if (k >= count || cnt->latest == 0) return;
  const unsigned i = (unsigned)list[k];
  for (int c = 0; c < 3; ++c) {
    a.val[c][i] = b.val[c][i];
    a.dist[c][i] = b.dist[c][i];
  }
  a.flags[i] = b.flags[i];
}

// restrictKnownVelocity (…wgsl.ts:450-471); level 1 samples the band result
__global__ __launch_bounds__(kThreads) void k_ex_restrict(ExSrc src, ExLevel tgt, FastDiv tz, FastDiv ty) {
  const unsigned n = (unsigned)tgt.n;
  const unsigned i = chunk_of((n + kThreads - 1) / kThreads) * kThreads + threadIdx.x;
  if (i >= n) return;
  int p[3];
  decode3(i, tgt.d[1], tgt.d[2], tz, ty, p[0], p[1], p[2]);
  unsigned known = 0;
  for (int c = 0; c < 3; ++c) {
    float v;
    float w = hier_sample(src, tgt.d, p, c, &v);
    if (w <= 0.0f && c == 1) w = corresponding_sample(src, p, c, &v);
    tgt.val[c][i] = w > 0.0f ? v : 0.0f;
    if (w > 0.0f) known |= 1u << c;
  }
  tgt.known[i] = (uint8_t)known;
}

// prolongUnknownVelocity (…wgsl.ts:473-488) into a level's up arrays
__global__ __launch_bounds__(kThreads) void k_ex_prolong(ExSrc src, ExLevel tgt, FastDiv tz, FastDiv ty) {
  const unsigned n = (unsigned)tgt.n;
  const unsigned i = chunk_of((n + kThreads - 1) / kThreads) * kThreads + threadIdx.x;
  if (i >= n) return;
  int p[3];
  decode3(i, tgt.d[1], tgt.d[2], tz, ty, p[0], p[1], p[2]);
  unsigned known = tgt.known[i];
  for (int c = 0; c < 3; ++c) {
    float out = tgt.val[c][i];
    if (!(known >> c & 1)) {
      float v;
      if (hier_sample(src, tgt.d, p, c, &v) > 0.0f) {
        out = v;
        known |= 1u << c;
      }
    }
    tgt.up_val[c][i] = out;
  }
  tgt.up_known[i] = (uint8_t)known;
}

// the last prolongation onto the band result, packed into the face arrays
// (packTransportShell, …wgsl.ts:491-512): known values on open faces, 0 else
__global__ __launch_bounds__(kThreads) void k_ex_finish(ExGeo g, ExSrc src, int have_src, ExState r,
                                                       const uint8_t* __restrict__ cls, ExCounters* cnt,
                                                       float* __restrict__ uo, float* __restrict__ vo,
                                                       float* __restrict__ wo, unsigned n) {
  int unknown = 0, band = 0;
  unsigned mes[kCells], knowns[kCells];
  float owns[kCells][3];
#pragma unroll
  for (int j = 0; j < kCells; ++j) {
    const unsigned i = min(cell_of(n, j), n - 1);
    mes[j] = cls[i];
    knowns[j] = r.flags[i] & 7u;
    owns[j][0] = r.val[0][i];
    owns[j][1] = r.val[1][i];
    owns[j][2] = r.val[2][i];
  }
#pragma unroll
  for (int j = 0; j < kCells; ++j) {
    const unsigned i = cell_of(n, j);
    if (i >= n) break;
    int x, y, z;
    edecode(g, i, x, y, z);
    const unsigned me = mes[j];
    const unsigned known = knowns[j];
    const float* own = owns[j];
    band += __popc(known);
    const int p[3] = {x, y, z};
    float* out[3] = {uo, vo, wo};
    for (int c = 0; c < 3; ++c) {
      float val = own[c];
      bool k = known >> c & 1;
      if (!k && have_src) {
        float v;
        if (hier_sample(src, g.d, p, c, &v) > 0.0f) {
          val = v;
          k = true;
        }
      }
      const bool open = me >> c & 1;
      if (open && !k) unknown += 1;
      out[c][face_of(g, c, x, y, z)] = open && k ? val : 0.0f;
    }
    // the negative domain faces are walls
    if (x == 0) uo[(unsigned)y * g.d[2] + z] = 0.0f;
    if (y == 0) vo[(unsigned)x * (g.d[1] + 1) * g.d[2] + z] = 0.0f;
    if (z == 0) wo[((unsigned)x * g.d[1] + y) * (g.d[2] + 1)] = 0.0f;
  }
  const int su = block_sum(unknown);
  __syncthreads();
  const int sb = block_sum(band);
  if (threadIdx.x == 0) {
    if (su) atomicAdd(&cnt->unknown, (unsigned long long)su);
    if (sb) atomicAdd(&cnt->band_known, (unsigned long long)sb);
  }
}

// workgroups of a streaming launch over n < 2^31 items, cells per thread k
int chunked_grid(long long n, int k = 1) {
  const long long chunks = (n + (long long)kThreads * k - 1) / ((long long)kThreads * k);
  return (int)(8 * ((chunks + 7) / 8));
}

}  // namespace

Extrapolator::~Extrapolator() {
  for (void* p : allocations_) (void)hipFree(p);
}

int Extrapolator::alloc(void** p, size_t bytes) {
  bytes = (bytes + 255) & ~(size_t)255;
  if (hipMalloc(p, bytes) != hipSuccess) {
    set_error("velocity extrapolation: hipMalloc(%zu) failed", bytes);
    return FP_ERR_OUT_OF_MEMORY;
  }
  allocations_.push_back(*p);
  bytes_ += bytes;
  return FP_OK;
}

static void level_dims(int nx, int ny, int nz, std::vector<std::array<int, 3>>* out) {
  int d[3] = {nx, ny, nz};
  while (std::min(d[0], std::min(d[1], d[2])) > 1) {
    for (int a = 0; a < 3; ++a) d[a] = (d[a] + 1) / 2;
    out->push_back({d[0], d[1], d[2]});
  }
}

size_t Extrapolator::estimate_bytes(int nx, int ny, int nz) {
  auto round = [](size_t b) { return (b + 255) & ~(size_t)255; };
  const size_t n = (size_t)nx * ny * nz;
  size_t total = 2 * 6 * round(n * 4) + 2 * round(n) + round(n) + 2 * round(n) + round(n * 4) + round(sizeof(ExCounters));
  total += round(n * 4) + round(((size_t)(nx + 1) * ny * nz + (size_t)nx * (ny + 1) * nz + (size_t)nx * ny * (nz + 1)) * 4);
  std::vector<std::array<int, 3>> dims;
  level_dims(nx, ny, nz, &dims);
  for (const auto& d : dims) {
    const size_t m = (size_t)d[0] * d[1] * d[2];
    total += 6 * round(m * 4) + 2 * round(m);
  }
  return total;
}

int Extrapolator::create(int nx, int ny, int nz, double h) {
  d_[0] = nx;
  d_[1] = ny;
  d_[2] = nz;
  n_ = (long long)nx * ny * nz;
  h_ = (float)h;
  const size_t n = (size_t)n_;
  int rc;
  for (int k = 0; k < 2; ++k) {
    for (int c = 0; c < 3; ++c) {
      if ((rc = alloc((void**)&st_[k].val[c], n * 4))) return rc;
      if ((rc = alloc((void**)&st_[k].dist[c], n * 4))) return rc;
    }
    if ((rc = alloc((void**)&st_[k].flags, n))) return rc;
  }
  if ((rc = alloc((void**)&cls_, n))) return rc;
  if ((rc = alloc((void**)&cand_[0], n))) return rc;
  if ((rc = alloc((void**)&cand_[1], n))) return rc;
  if ((rc = alloc((void**)&list_, n * 4))) return rc;
  if ((rc = alloc((void**)&counters_, sizeof(ExCounters)))) return rc;
  const size_t faces = (size_t)(nx + 1) * ny * nz + (size_t)nx * (ny + 1) * nz + (size_t)nx * ny * (nz + 1);
  if ((rc = alloc((void**)&in_, n * 4 + faces * 4))) return rc;
  std::vector<std::array<int, 3>> dims;
  level_dims(nx, ny, nz, &dims);
  for (const auto& d : dims) {
    ExLevel L;
    std::memset(&L, 0, sizeof(L));
    L.d[0] = d[0];
    L.d[1] = d[1];
    L.d[2] = d[2];
    L.n = (long long)d[0] * d[1] * d[2];
    for (int c = 0; c < 3; ++c) {
      if ((rc = alloc((void**)&L.val[c], (size_t)L.n * 4))) return rc;
      if ((rc = alloc((void**)&L.up_val[c], (size_t)L.n * 4))) return rc;
    }
    if ((rc = alloc((void**)&L.known, (size_t)L.n))) return rc;
    if ((rc = alloc((void**)&L.up_known, (size_t)L.n))) return rc;
    levels_.push_back(L);
  }
  return FP_OK;
}

int Extrapolator::run(hipStream_t s, const fp_extrapolate_args* a, fp_extrapolate_report* rep, bool same_band) {
  const auto t0 = std::chrono::steady_clock::now();
  const int nx = d_[0], ny = d_[1], nz = d_[2];
  const long long n = n_;
  const size_t nu = (size_t)(nx + 1) * ny * nz, nv = (size_t)nx * (ny + 1) * nz, nw = (size_t)nx * ny * (nz + 1);
  if ((long long)(nx + 1) * ny * nz >= (1ll << 31) || (long long)nx * (ny + 1) * nz >= (1ll << 31) ||
      (long long)nx * ny * (nz + 1) >= (1ll << 31)) {
    set_error("fp_extrapolate_velocity: grids of 2^31 or more faces per component are not supported");
    return FP_ERR_UNSUPPORTED;
  }
  const float* field = a->phi ? a->phi : a->density;
  const float *u = a->u, *v = a->v, *w = a->w;
  // kernels write device memory only: host outputs go through the input
  // staging (free once the seed has read the inputs) and are copied back
  float *uo = a->u_out, *vo = a->v_out, *wo = a->w_out;
  if (a->memory == FP_MEMORY_HOST) {
    float* f = in_;
    float* hu = f + n;
    float* hv = hu + nu;
    float* hw = hv + nv;
    HIPCHK(hipMemcpyAsync(f, field, (size_t)n * 4, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(hu, u, nu * 4, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(hv, v, nv * 4, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(hw, w, nw * 4, hipMemcpyHostToDevice, s));
    field = f;
    u = hu;
    v = hv;
    w = hw;
    uo = hu;
    vo = hv;
    wo = hw;
  }
  ExGeo g;
  g.d[0] = nx;
  g.d[1] = ny;
  g.d[2] = nz;
  g.h = h_;
  g.open_top = a->open_top ? 1 : 0;
  g.dz = make_div((unsigned)nz);
  g.dy = make_div((unsigned)ny);
  g.eps_min = 2.0f * h_;
  {
    const float dx = (float)nx * h_, dy = (float)ny * h_, dz = (float)nz * h_;
    g.eps_max = std::sqrt((dx * dx + dy * dy) + dz * dz);
  }
  int ceiling = std::min(std::max(nx, std::max(ny, nz)), 16);
  if (a->max_passes > 0) ceiling = std::min(a->max_passes, 19);
  const dim3 grid(chunked_grid(n, kCells)), block(kThreads);
  HIPCHK(hipMemsetAsync(counters_, 0, sizeof(ExCounters), s));
  const unsigned un = (unsigned)n;
  same_band = same_band && band_count_ >= 0;
  if (!same_band) {
    hipLaunchKernelGGL(k_ex_classify, grid, block, 0, s, g, field, a->phi ? 1 : 0, cls_, cand_[0], un);
    hipLaunchKernelGGL(k_ex_dilate, grid, block, 0, s, g, (const uint8_t*)cls_, (const uint8_t*)cand_[0], cand_[1],
                       0, list_, counters_, un);
    hipLaunchKernelGGL(k_ex_dilate, grid, block, 0, s, g, (const uint8_t*)cls_, (const uint8_t*)cand_[1], cand_[0],
                       0, list_, counters_, un);
    hipLaunchKernelGGL(k_ex_dilate, grid, block, 0, s, g, (const uint8_t*)cls_, (const uint8_t*)cand_[0], cand_[1],
                       1, list_, counters_, un);
  }
  hipLaunchKernelGGL(k_ex_seed, grid, block, 0, s, g, (const uint8_t*)cls_, (const uint8_t*)cand_[1], u, v, w,
                     st_[0], st_[1], counters_, un);
  ExCounters host;
  if (!same_band) {
    HIPCHK(hipMemcpyAsync(&host, counters_, sizeof(ExCounters), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    band_count_ = host.list_count;
  }
  const int count = band_count_;
  if (count > 0) {
    for (int pass = 1; pass <= ceiling; ++pass)
      hipLaunchKernelGGL(k_ex_pass, dim3((count + kThreads - 1) / kThreads), block, 0, s, g, (const uint8_t*)cls_,
                         (const int*)list_, count, st_[(pass - 1) & 1], st_[pass & 1], counters_, pass);
    hipLaunchKernelGGL(k_ex_resolve, dim3((count + kThreads - 1) / kThreads), block, 0, s, (const int*)list_, count,
                       st_[0], st_[1], (const ExCounters*)counters_);
  }
  // hierarchy: restrict down, prolong up, the last prolongation fused with the packing
  const int L = (int)levels_.size();
  const ExSrc none{{0, 0, 0}, {nullptr, nullptr, nullptr}, nullptr};
  for (int l = 0; l < L; ++l) {
    const ExLevel& T = levels_[l];
    ExSrc src{{nx, ny, nz}, {st_[0].val[0], st_[0].val[1], st_[0].val[2]}, st_[0].flags};
    if (l > 0) {
      const ExLevel& S = levels_[l - 1];
      src = ExSrc{{S.d[0], S.d[1], S.d[2]}, {S.val[0], S.val[1], S.val[2]}, S.known};
    }
    hipLaunchKernelGGL(k_ex_restrict, dim3(chunked_grid(T.n)), block, 0, s, src, T, make_div((unsigned)T.d[2]),
                       make_div((unsigned)T.d[1]));
  }
  ExSrc coarser = none;
  if (L > 0) {
    const ExLevel& C = levels_[L - 1];
    coarser = ExSrc{{C.d[0], C.d[1], C.d[2]}, {C.val[0], C.val[1], C.val[2]}, C.known};
  }
  for (int l = L - 2; l >= 0; --l) {
    const ExLevel& T = levels_[l];
    hipLaunchKernelGGL(k_ex_prolong, dim3(chunked_grid(T.n)), block, 0, s, coarser, T, make_div((unsigned)T.d[2]),
                       make_div((unsigned)T.d[1]));
    coarser = ExSrc{{T.d[0], T.d[1], T.d[2]}, {T.up_val[0], T.up_val[1], T.up_val[2]}, T.up_known};
  }
  hipLaunchKernelGGL(k_ex_finish, grid, block, 0, s, g, coarser, L > 0 ? 1 : 0, st_[0], (const uint8_t*)cls_,
                     counters_, uo, vo, wo, un);
  if (a->memory == FP_MEMORY_HOST) {
    HIPCHK(hipMemcpyAsync(a->u_out, uo, nu * 4, hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(a->v_out, vo, nv * 4, hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(a->w_out, wo, nw * 4, hipMemcpyDeviceToHost, s));
  }
  HIPCHK(hipMemcpyAsync(&host, counters_, sizeof(ExCounters), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  if (rep) {
    std::memset(rep, 0, sizeof(*rep));
    int passes = 0;
    for (int k = 1; k <= ceiling; ++k)
      if (host.active[k - 1] > 0) passes = k;
    rep->passes = count > 0 ? passes : 0;
    rep->active_left = rep->passes > 0 ? host.active[rep->passes] : host.active[0];
    rep->levels = L;
    rep->band_faces = (int64_t)host.band_known;
    rep->unknown_faces = (int64_t)host.unknown;
    rep->device_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  }
  return FP_OK;
}

}  // namespace fp

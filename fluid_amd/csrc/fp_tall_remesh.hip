// fp_tall_remesh.hip -- re-laying a restricted tall-cell grid out around its
// current surface, on the device (fp_tall_remesh).
//
// Per column (one thread): the lowest and highest fine rows where the liquid
// (phi < 0) meets air along the column, with phi read from the current
// layout -- linear in the row inside the tall cell (the reconstruction of
// representedInitialPhiVolumeCells, lib/core/tall-cell-grid.ts:220-238), the
// band samples above it, air beyond the band -- and a surface cell being one
// whose wetness differs from the row below or above (isInitialSurfaceCell,
// tall-cell-grid.ts:263-279); the base from them (chooseTallCellBase,
// tall-cell-grid.ts:61-82; a dry column takes the highest base, :378-386);
// then Jacobi passes of base <= min(neighbour) + D until no base changes
// (limitNeighboringTallCellBases, :84-106, run to convergence as the layout
// builder's nx + nz passes do, :395).  The quantities move to the new layout
// as tallCells.txt:240-252 states: band cells copy the old value at their
// fine row (read linearly from the old tall cell where it covered the row),
// the new tall cell's bottom and top samples are the least-squares line
// through the old values of its rows; vertical face slots likewise (the tall
// cell's interior faces share the mean), horizontal ones refit their affine
// part over the rows both columns' tall cells share.  A column (pair) whose
// base(s) did not change is copied as is.  FP64 sums in row order without
// contraction, so a CPU restatement gives the same bits.
#pragma clang fp contract(off)

#include "fp_tall.hpp"

#include <algorithm>
#include <cstring>
#include <vector>

namespace fp {

namespace {

struct RemeshGeo {
  int nx, ny, nz, B, P, S, D;
  double air;  // phi of the rows beyond the band (5 h, buildInitialPhi's limit)
};

// phi of fine row y of column c in a layout with base b
__device__ __forceinline__ double column_phi(const RemeshGeo& g, const float* __restrict__ phi, int x, int z, int b,
                                             int y) {
  if (y < b) {
    const double bottom = phi[((size_t)x * g.P + 0) * g.nz + z];
    const double top = phi[((size_t)x * g.P + 1) * g.nz + z];
    const double t = fmin(1.0, fmax(0.0, (double)y / (double)max(b - 1, 1)));
    return bottom + (top - bottom) * t;
  }
  const int py = 2 + y - b;
  return py < g.P ? (double)phi[((size_t)x * g.P + py) * g.nz + z] : g.air;
}

// v at fine face f (between rows f-1 and f) of a column with base b: slot 0
// inside the tall cell, slot f - b + 1 from the tall top up
__device__ __forceinline__ double column_v(const RemeshGeo& g, const float* __restrict__ v, int x, int z, int b,
                                           int f) {
  if (f <= 0 || f >= g.ny) return 0.0;
  if (f < b) return v[((size_t)x * g.P + 0) * g.nz + z];
  const int s = f - b + 1;
  return s < g.P ? (double)v[((size_t)x * g.P + s) * g.nz + z] : 0.0;
}

// horizontal slots of one face: the affine part over rows 0..m-1 (mean, slope
// about (m-1)/2), then one slot per row from m
__device__ __forceinline__ double face_value(const RemeshGeo& g, const float* __restrict__ a, size_t base_index,
                                             size_t stride, int m, int y) {
  if (y < m) {
    const double mean = a[base_index];
    const double slope = a[base_index + stride];
    return mean + slope * ((double)y - 0.5 * (double)(m - 1));
  }
  const int s = 2 + y - m;
  return s < g.S ? (double)a[base_index + (size_t)s * stride] : 0.0;
}

// chooseTallCellBase (tall-cell-grid.ts:61-82) with Math.round
__device__ __forceinline__ int choose_base(int lowest, int highest, int fine_ny, int B, int liquid_halo,
                                           int air_halo) {
  const int maximum_base = max(0, fine_ny - B);
  const int lower = highest + 1 + air_halo - B;
  const int upper = lowest + 1 - liquid_halo;
  const int desired = lower <= upper ? (int)floor((double)(lower + upper) / 2.0 + 0.5) : lower;
  const int clamped = max(0, min(maximum_base, desired));
  return maximum_base >= 2 ? max(2, clamped) : 0;
}

__global__ void k_remesh_bases(RemeshGeo g, const float* __restrict__ phi, const int* __restrict__ bases,
                               int* __restrict__ raw, int liquid_halo, int air_halo) {
  const int ncol = g.nx * g.nz;
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < ncol; c += gridDim.x * blockDim.x) {
    const int x = c / g.nz, z = c - (c / g.nz) * g.nz;
    const int b = bases[c];
    int lowest = g.ny, highest = -1;
    bool below = false, here = column_phi(g, phi, x, z, b, 0) < 0.0;
    for (int y = 0; y < g.ny; ++y) {
      const bool above = y + 1 < g.ny && column_phi(g, phi, x, z, b, y + 1) < 0.0;
      const bool under = y == 0 ? here : below;
      if (here != under || here != above) {
        lowest = min(lowest, y);
        highest = max(highest, y);
      }
      below = here;
      here = above;
    }
    const int maximum_base = max(0, g.ny - g.B);
    const int lh = min(liquid_halo, g.B), ah = min(air_halo, g.B);
    const int surface = highest >= 0 ? choose_base(lowest, highest, g.ny, g.B, lh, ah)
                                     : (maximum_base <= 3 ? min(2, maximum_base) : maximum_base);
    raw[c] = maximum_base >= 2 ? max(2, surface) : 0;
  }
}

// one Jacobi pass of base <= min(4-neighbour) + D; *changed set when a base drops
__global__ void k_remesh_limit(RemeshGeo g, const int* __restrict__ cur, int* __restrict__ next, int* changed) {
  const int ncol = g.nx * g.nz;
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < ncol; c += gridDim.x * blockDim.x) {
    const int x = c / g.nz, z = c - (c / g.nz) * g.nz;
    int up = cur[c];
    if (x > 0) up = min(up, cur[c - g.nz] + g.D);
    if (x + 1 < g.nx) up = min(up, cur[c + g.nz] + g.D);
    if (z > 0) up = min(up, cur[c - 1] + g.D);
    if (z + 1 < g.nz) up = min(up, cur[c + 1] + g.D);
    next[c] = up;
    if (up != cur[c]) atomicOr(changed, 1);
  }
}

// new samples and vertical slots of one column
__global__ void k_remesh_column(RemeshGeo g, const int* __restrict__ ob, const int* __restrict__ nb,
                                const float* __restrict__ phi, const float* __restrict__ v, float* __restrict__ phi_o,
                                float* __restrict__ v_o, unsigned long long* changed) {
  const int ncol = g.nx * g.nz;
  unsigned long long moved = 0;
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < ncol; c += gridDim.x * blockDim.x) {
    const int x = c / g.nz, z = c - (c / g.nz) * g.nz;
    const int bo = ob[c], bn = nb[c];
    const size_t i0 = (size_t)x * g.P * g.nz + z;
    if (bo == bn) {
      for (int py = 0; py < g.P; ++py) {
        phi_o[i0 + (size_t)py * g.nz] = phi[i0 + (size_t)py * g.nz];
        v_o[i0 + (size_t)py * g.nz] = v[i0 + (size_t)py * g.nz];
      }
      continue;
    }
    ++moved;
    // least-squares line through rows 0..bn-1, in t = y - (bn - 1)/2
    const double mid = 0.5 * (double)(bn - 1);
    double sf = 0.0, stf = 0.0, stt = 0.0;
    for (int y = 0; y < bn; ++y) {
      const double f = column_phi(g, phi, x, z, bo, y);
      const double t = (double)y - mid;
      sf += f;
      stf += t * f;
      stt += t * t;
    }
    const double mean = sf / (double)bn, slope = stt > 0.0 ? stf / stt : 0.0;
    phi_o[i0] = (float)(mean - slope * mid);
    phi_o[i0 + (size_t)g.nz] = (float)(mean + slope * mid);
    for (int py = 2; py < g.P; ++py) {
      const int y = bn + py - 2;
      phi_o[i0 + (size_t)py * g.nz] = y < g.ny ? (float)column_phi(g, phi, x, z, bo, y) : (float)g.air;
    }
    // vertical faces: the tall cell's interior faces 1..bn-1 share their mean
    double sv = 0.0;
    for (int f = 1; f < bn; ++f) sv += column_v(g, v, x, z, bo, f);
    v_o[i0] = (float)(sv / (double)max(bn - 1, 1));
    for (int s = 1; s < g.P; ++s) v_o[i0 + (size_t)s * g.nz] = (float)column_v(g, v, x, z, bo, bn + s - 1);
  }
  if (moved) atomicAdd(changed, moved);
}

// new horizontal slots of the faces between columns (axis 0: x faces, 1: z faces)
__global__ void k_remesh_faces(RemeshGeo g, int axis, const int* __restrict__ ob, const int* __restrict__ nb,
                               const float* __restrict__ a, float* __restrict__ a_o) {
  // u: (nx+1, S, nz); w: (nx, S, nz+1)
  const int fx = axis == 0 ? g.nx + 1 : g.nx, fz = axis == 0 ? g.nz : g.nz + 1;
  const int nf = fx * fz;
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < nf; k += gridDim.x * blockDim.x) {
    const int x = k / fz, z = k - (k / fz) * fz;
    const size_t stride = (size_t)fz;
    const size_t i0 = (size_t)x * g.S * fz + z;
    const bool wall = axis == 0 ? (x == 0 || x == g.nx) : (z == 0 || z == g.nz);
    int ca = 0, cn = 0;
    if (!wall) {
      ca = axis == 0 ? (x - 1) * g.nz + z : x * g.nz + z - 1;
      cn = axis == 0 ? x * g.nz + z : x * g.nz + z;
    }
    if (wall || (ob[ca] == nb[ca] && ob[cn] == nb[cn])) {
      for (int s = 0; s < g.S; ++s) a_o[i0 + (size_t)s * stride] = a[i0 + (size_t)s * stride];
      continue;
    }
    const int mo = min(ob[ca], ob[cn]), mn = min(nb[ca], nb[cn]);
    const int top = min(g.ny, max(nb[ca], nb[cn]) + g.B);
    const double mid = 0.5 * (double)(mn - 1);
    double sf = 0.0, stf = 0.0, stt = 0.0;
    for (int y = 0; y < mn; ++y) {
      const double f = face_value(g, a, i0, stride, mo, y);
      const double t = (double)y - mid;
      sf += f;
      stf += t * f;
      stt += t * t;
    }
    a_o[i0] = (float)(sf / (double)mn);
    a_o[i0 + stride] = (float)(stt > 0.0 ? stf / stt : 0.0);
    for (int s = 2; s < g.S; ++s) {
      const int y = mn + s - 2;
      a_o[i0 + (size_t)s * stride] = y < top ? (float)face_value(g, a, i0, stride, mo, y) : 0.0f;
    }
  }
}

int grid_of(long long n) { return (int)std::min<long long>(std::max<long long>((n + 255) / 256, 1), 4096); }

}  // namespace

size_t TallGrid::remesh_bytes(int nx, int nz, int band, int max_delta) {
  const size_t P = (size_t)band + 2, S = P + (size_t)max_delta;
  const size_t state = (size_t)nx * P * nz * 2 + (size_t)(nx + 1) * S * nz + (size_t)nx * S * (nz + 1);
  auto round = [](size_t b) { return (b + 255) & ~(size_t)255; };
  return 2 * round(state * 4) + round((size_t)nx * nz * 8 + 64);
}

int TallGrid::remesh(hipStream_t s, const fp_tall_remesh_args* a, int64_t* changed_columns) {
  TallRec& R = *rec_;
  if (R.ny - R.band < 2) {
    set_error("fp_tall_remesh: fine_ny %d - band %d < 2 leaves no tall cell to re-lay", R.ny, R.band);
    return FP_ERR_LAYOUT;
  }
  const long long ncol = (long long)R.nx * R.nz;
  const size_t state = (size_t)(R.n + R.nu + R.nv + R.nw);
  if (!rm_in_) {
    int rc = alloc((void**)&rm_in_, state * 4);
    if (!rc) rc = alloc((void**)&rm_out_, state * 4);
    if (!rc) rc = alloc((void**)&rm_b_, (size_t)ncol * 8 + 64);
    if (rc) return rc;
  }
  const size_t bytes[4] = {(size_t)R.n * 4, (size_t)R.nu * 4, (size_t)R.nv * 4, (size_t)R.nw * 4};
  const void* in[4] = {a->phi, a->u, a->v, a->w};
  void* out[4] = {a->phi_out, a->u_out, a->v_out, a->w_out};
  float* stage_in[4];
  float* stage_out[4];
  stage_in[0] = rm_in_;
  stage_out[0] = rm_out_;
  for (int k = 1; k < 4; ++k) {
    stage_in[k] = stage_in[k - 1] + bytes[k - 1] / 4;
    stage_out[k] = stage_out[k - 1] + bytes[k - 1] / 4;
  }
  const float* src[4];
  float* dst[4];
  if (a->memory == FP_MEMORY_HOST) {
    for (int k = 0; k < 4; ++k) {
      HIPCHK(hipMemcpyAsync(stage_in[k], in[k], bytes[k], hipMemcpyHostToDevice, s));
      src[k] = stage_in[k];
      dst[k] = stage_out[k];
    }
  } else {
    // device arrays are read in place; an output is written directly unless
    // it overlaps an input that a later kernel still reads
    for (int k = 0; k < 4; ++k) {
      src[k] = (const float*)in[k];
      bool overlaps = false;
      for (int j = 0; j < 4 && out[k]; ++j) {
        const char *o0 = (const char*)out[k], *i0 = (const char*)in[j];
        overlaps = overlaps || (o0 < i0 + bytes[j] && i0 < o0 + bytes[k]);
      }
      dst[k] = out[k] && !overlaps ? (float*)out[k] : stage_out[k];
    }
  }
  const float *ip = src[0], *iu = src[1], *iv = src[2], *iw = src[3];
  float *op = dst[0], *ou = dst[1], *ov = dst[2], *ow = dst[3];
  int* raw = rm_b_;
  int* tmp = rm_b_ + ncol;
  int* flag = (int*)(rm_b_ + 2 * ncol);
  unsigned long long* moved = (unsigned long long*)(rm_b_ + 2 * ncol + 8);
  const RemeshGeo g{R.nx, R.ny, R.nz, R.band, R.P, R.S, R.D, 5.0 * R.h};
  // tallCellSettings.high halos (tall-cell-grid.ts:57)
  const int lh = a->liquid_halo > 0 ? a->liquid_halo : 16, ah = a->air_halo > 0 ? a->air_halo : 8;
  const int blocks = grid_of(ncol);
  hipLaunchKernelGGL(k_remesh_bases, dim3(blocks), dim3(256), 0, s, g, (const float*)ip, R.bases, raw, lh, ah);
  // the neighbour limit to convergence, checked every 8 passes
  int* cur = raw;
  int* nxt = tmp;
  for (long long pass = 0; pass <= (long long)R.nx + R.nz; pass += 8) {
    HIPCHK(hipMemsetAsync(flag, 0, 4, s));
    for (int k = 0; k < 8; ++k) {
      hipLaunchKernelGGL(k_remesh_limit, dim3(blocks), dim3(256), 0, s, g, (const int*)cur, nxt, flag);
      std::swap(cur, nxt);
    }
    int h_flag = 0;
    HIPCHK(hipMemcpyAsync(&h_flag, flag, 4, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    if (!h_flag) break;
  }
  const int* nb = cur;
  HIPCHK(hipMemsetAsync(moved, 0, 8, s));
  hipLaunchKernelGGL(k_remesh_column, dim3(blocks), dim3(256), 0, s, g, R.bases, nb, (const float*)ip,
                     (const float*)iv, op, ov, moved);
  hipLaunchKernelGGL(k_remesh_faces, dim3(grid_of((long long)(R.nx + 1) * R.nz)), dim3(256), 0, s, g, 0, R.bases, nb,
                     (const float*)iu, ou);
  hipLaunchKernelGGL(k_remesh_faces, dim3(grid_of((long long)R.nx * (R.nz + 1))), dim3(256), 0, s, g, 1, R.bases, nb,
                     (const float*)iw, ow);
  // the grid takes the new layout; its resident pressure belonged to the old one
  HIPCHK(hipMemcpyAsync(const_cast<int*>(R.bases), nb, (size_t)ncol * 4, hipMemcpyDeviceToDevice, s));
  HIPCHK(hipMemsetAsync(R.x, 0, (size_t)R.n * 8, s));
  const hipMemcpyKind kout = a->memory == FP_MEMORY_HOST ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice;
  for (int k = 0; k < 4; ++k)
    if (out[k] && dst[k] != out[k]) HIPCHK(hipMemcpyAsync(out[k], dst[k], bytes[k], kout, s));
  std::vector<int> hb((size_t)ncol);
  unsigned long long h_moved = 0;
  HIPCHK(hipMemcpyAsync(hb.data(), nb, (size_t)ncol * 4, hipMemcpyDeviceToHost, s));
  HIPCHK(hipMemcpyAsync(&h_moved, moved, 8, hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  if (a->column_bases_out) std::memcpy(a->column_bases_out, hb.data(), (size_t)ncol * 4);
  // every column's band reaching the lid leaves the pressure unanchored
  R.ns_liquid = 0;
  for (long long i = 0; i < ncol && R.ns_liquid >= 0; ++i)
    R.ns_liquid = hb[(size_t)i] + R.band >= R.ny ? R.ns_liquid + 2 + (R.ny - hb[(size_t)i]) : -1;
  // the state machine reads ns_liquid from the device record
  int rc = upload(s);
  if (rc) return rc;
  HIPCHK(hipStreamSynchronize(s));
  if (changed_columns) *changed_columns = (int64_t)h_moved;
  return FP_OK;
}

}  // namespace fp

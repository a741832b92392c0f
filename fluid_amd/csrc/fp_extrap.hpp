// fp_extrap.hpp -- velocity extrapolation of a uniform grid (host side).
//
// The step after the projection in the reference's uniform solver
// (lib/methods/uniform/webgpu-uniform-velocity-extrapolation{,.wgsl}.ts): every
// velocity component is extended from the faces next to liquid into the air,
// first by Jeong-Ross-Whitaker passes of the Godunov distance clipped to a
// two-cell band with the first-order upwind extension of CM11b Eq. 7, then
// into the far field by a restrict / prolong hierarchy of renormalised
// trilinear interpolation of known values.  FP32 in the reference's operation
// order without contraction: results equal oracle/fp_oracle_extrap.c bit for
// bit.
#pragma once

#include "fluid_projection.h"

#include <hip/hip_runtime.h>

#include <vector>

namespace fp {

// One ping-pong side of the narrow-band state (per cell, z fastest).
struct ExState {
  float* val[3];
  float* dist[3];
  uint8_t* flags;  // bits 0-2 known(c), 3-5 active(c)
};

// One level of the far-field hierarchy.
struct ExLevel {
  int d[3];
  long long n;
  float* val[3];
  uint8_t* known;
  float* up_val[3];
  uint8_t* up_known;
};

struct ExCounters {
  int active[20];                  // [0] after the seed, [k] after pass k
  int latest;                      // side holding the band result
  int list_count;                  // cells of the band list
  unsigned long long unknown;      // open faces left unknown
  unsigned long long band_known;   // faces known after the band
};

class Extrapolator {
 public:
  ~Extrapolator();
  static size_t estimate_bytes(int nx, int ny, int nz);
  int create(int nx, int ny, int nz, double h);
  // One extrapolation of a->phi / a->density and a->u, a->v, a->w into the
  // outputs; synchronous on s.
  // same_band: the field (phi / density) and open_top are those of this
  // object's previous run, whose classes, near mask and band list are reused
  // (a second extrapolation over the same surface, as a MacCormack step makes)
  int run(hipStream_t s, const fp_extrapolate_args* a, fp_extrapolate_report* rep, bool same_band = false);
  size_t bytes() const { return bytes_; }

 private:
  int alloc(void** p, size_t bytes);
  std::vector<void*> allocations_;
  std::vector<ExLevel> levels_;
  size_t bytes_ = 0;
  int d_[3] = {0, 0, 0};
  long long n_ = 0;
  float h_ = 0.0f;
  ExState st_[2];
  uint8_t* cls_ = nullptr;
  uint8_t* cand_[2] = {nullptr, nullptr};
  int band_count_ = -1;  // band list length of the previous run (-1: none)
  int* list_ = nullptr;
  float* in_ = nullptr;  // host-memory staging: phi or density, u, v, w
  ExCounters* counters_ = nullptr;
};

}  // namespace fp

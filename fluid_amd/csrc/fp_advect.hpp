// fp_advect.hpp -- velocity advection and body forces of a uniform grid (host side).
//
// The step before the projection in the reference's uniform solver
// (lib/methods/uniform/webgpu-uniform-reference.wgsl.ts:872-1004): the
// velocity is transported along RK2 characteristics through its extrapolated
// field, semi-Lagrangian or bounded MacCormack, then gravity and viscosity
// act on it.  FP32 in the reference's operation order without contraction:
// results equal oracle/fp_oracle_advect.c bit for bit.
#pragma once

#include "fluid_projection.h"

#include <hip/hip_runtime.h>

#include <vector>

namespace fp {

class Extrapolator;

class Advector {
 public:
  ~Advector();
  static size_t estimate_bytes(int nx, int ny, int nz);
  int create(int nx, int ny, int nz, double h);
  // One advection of a->u, a->v, a->w into the outputs; ex extrapolates the
  // MacCormack prediction (and the current field when no transport is
  // given).  Synchronous on s.
  int run(hipStream_t s, Extrapolator* ex, const fp_advect_args* a, fp_advect_report* rep);
  size_t bytes() const { return bytes_; }

 private:
  int alloc(void** p, size_t bytes);
  int stage_host(size_t floats);
  std::vector<void*> allocations_;
  size_t bytes_ = 0;
  int d_[3] = {0, 0, 0};
  float h_ = 0.0f;
  size_t nf_[3] = {0, 0, 0};
  float* transport_ = nullptr;  // the current field's extrapolation, when the caller gives none
  float* pred_ = nullptr;       // MacCormack prediction
  float* rev_ = nullptr;        // MacCormack reverse advection
  float* pred_ext_ = nullptr;   // the prediction's extrapolation
  float* out_ = nullptr;        // output staging (host memory or outputs overlapping inputs)
  float* host_ = nullptr;       // input staging for host memory
  size_t host_floats_ = 0;
};

}  // namespace fp

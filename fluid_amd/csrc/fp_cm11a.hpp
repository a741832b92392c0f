// fp_cm11a.hpp -- the CM11a separating-wall pressure solve on MI355X (host side).
//
// A second uniform-grid solver behind fp_project (fp_options.solver =
// FP_SOLVER_CM11A): the reference's own uniform solver, a linear
// complementarity problem on halo-padded levels solved by a fixed multigrid
// schedule (lib/methods/uniform/webgpu-uniform-pressure-multigrid.ts:488-576,
// …multigrid.wgsl.ts).  Liquid next to a box wall couples to a halo unknown
// with p >= 0, so the liquid may pull away from the wall instead of being held
// by a negative pressure.  Arithmetic is FP32 in the reference's operation
// order (D4-ordered sums, a double-single coarsest solve), without
// contraction, so results equal the CPU restatement bit for bit.
#pragma once

#include "fluid_projection.h"

#include <hip/hip_runtime.h>

#include <vector>

namespace fp {

class UniformGrid;

// One halo-padded level on the device (z fastest: (x * D1 + y) * D2 + z).
struct CmLevel {
  int D0, D1, D2, n;
  float h0, h1, h2;
  float* phi_raw;  // coarsened level set (the raw pyramid)
  float* phi;      // continued one cell into the walls
  float4* topo;    // cell dual volume, +x/+y/+z face dual volumes
  float4* coef;    // +x/+y/+z coefficients V / (h^2 theta), liquid flag
  float* p;
  float* rhs0;     // V-cycle right-hand side
  float* rhs1;     // full-cycle correction right-hand side
  float* res;      // residual scratch
  float* pmin;     // p_min of the level
};

struct CmStats {
  int coarse_sweeps;
  int coarse_failed;
  int coarse_solves;
  unsigned coarse_residual_bits;
  unsigned lcp_residual_bits;  // finest projected residual max (1/s), float bits
  unsigned div_before_bits, div_after_bits;
  int liquid;  // liquid cells of the finest level
};

class Cm11a {
 public:
  ~Cm11a();
  // Device bytes the solver of an (nx, ny, nz) grid allocates.
  static size_t estimate_bytes(int nx, int ny, int nz);
  int create(int nx, int ny, int nz, double h);
  // One projection of the grid's bound inputs (UniformGrid::bind) into its
  // bound outputs; synchronous on s.
  int run(hipStream_t s, UniformGrid* g, const fp_options* opt, double dt, double rho, fp_report* rep);
  size_t bytes() const { return bytes_; }

 private:
  int alloc(void** p, size_t bytes);
  void vcycle(hipStream_t s, int l, const float* rhs, const float* pmin0);
  void full_cycle(hipStream_t s, const float* pmin0);
  void coarse_solve(hipStream_t s, const float* rhs, const float* pmin);
  std::vector<CmLevel> lv_;
  std::vector<void*> allocations_;
  size_t bytes_ = 0;
  int nx_ = 0, ny_ = 0, nz_ = 0;
  double h_ = 0.0;
  float* b_ = nullptr;      // finest right-hand side
  float* pmin0_ = nullptr;  // finest p_min
  float* pshift_ = nullptr; // finest p_min - p during a full cycle
  float* backup_ = nullptr;
  double* partials_ = nullptr;
  CmStats* stats_ = nullptr;
  // schedule of the current run
  int pre_ = 6, post_ = 6, cap_ = 4096, constraint_levels_ = 3, open_top_ = 0;
  float dt_ = 0.0f, rho_ = 0.0f, tol_ = 1e-4f;
};

}  // namespace fp

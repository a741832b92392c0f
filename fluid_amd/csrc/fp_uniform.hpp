// fp_uniform.hpp -- uniform-grid projection solver (host class + device state).
#pragma once

#include "../../include/fluid_projection.h"
#include "fp_common.hpp"

#include <vector>

namespace fp {

constexpr int kMaxLevels = 16;
constexpr unsigned kMaxBlocks = 16384;

void set_error(const char* fmt, ...);

// Device-resident scalars of one projection (the GPU-side counterpart of the
// reference's convergence / error words, webgpu-octree-pipelined-mgpcg.ts:66-73).
struct State {
  double bb;             // ||b||^2
  double tol2_bb;        // (tol ||b||)^2
  double rr64;           // ||b - A x||^2 of the FP64 iterate
  double inner_target2;  // inner FP32 solve target ||r||^2
  double rz;             // r . z of the current iteration
  double alpha, beta;
  double rr;             // inner ||r||^2
  long long liquid;      // liquid samples
  int done;              // inner solve finished
  int outer_done;        // FP64 tolerance met
  int flags;             // FP_FLAG_*
  int iter;              // inner iterations this pass
  int total_iter;
  float rscale;          // power-of-two normalisation of the V-cycle input
  unsigned rmax_bits;    // max |r| accumulator (float bits)
  unsigned divmax_before, divmax_after;
  unsigned counters[8];  // grid-reduction tickets
};

struct UniformLevel {
  int d[3] = {0, 0, 0};
  int s[3] = {1, 1, 1};
  float g[3] = {1, 1, 1};
  size_t n = 0;
  float* phi = nullptr;
  void* r = nullptr;  // restricted right-hand side (storage T)
  void* e = nullptr;  // correction (storage T); level 0: the preconditioned residual z
  void* e2 = nullptr; // second correction buffer (fused sweeps ping-pong)
  void* w = nullptr;  // packed coefficients W4<T> (fused levels)
  void* t = nullptr;  // residual scratch (storage T)
  float Sr = 1, Se = 1, St = 1;  // stored = true / S
};

int plan_uniform_levels(int nx, int ny, int nz, int dims[][3], int strides[][3]);

class UniformSolver {
 public:
  ~UniformSolver();
  static size_t estimate_bytes(int nx, int ny, int nz);
  int create(int nx, int ny, int nz, double h);
  int project(hipStream_t s, const fp_inputs* in, fp_outputs* out, const fp_options* opt, fp_report* rep);
  int probe_apply(hipStream_t s, const float* phi, const double* p, double* out);
  int probe_rhs(hipStream_t s, const fp_inputs* in, double* b);
  int probe_vcycle(hipStream_t s, const float* phi, const fp_options* opt, const float* r, float* z);
  int probe_level_phi(hipStream_t s, const float* phi, int level, int preserve, float* out);

  int levels() const { return L_; }
  const UniformLevel& level(int l) const { return lv_[l]; }
  size_t bytes() const { return bytes_; }
  size_t cells() const { return n_; }

 private:
  int alloc(void** p, size_t bytes);
  void release();
  void set_level_scales(int precision);
  void build_weights(hipStream_t s, int precision);
  int setup_levels(hipStream_t s, const float* phi0, int preserve);
  template <typename T> int vcycle(hipStream_t s, int pre, int post);
  template <typename T> int iteration(hipStream_t s, int pre, int post, int max_iter, int parity);
  void run_iteration(hipStream_t s, int precision, int pre, int post, int max_iter, int parity);
  int ensure_graph(hipStream_t s, int precision, int pre, int post, int max_iter);
  hipGraphExec_t graph_exec_ = nullptr;
  int graph_key_[4] = {-1, -1, -1, -1};

  int nx_ = 0, ny_ = 0, nz_ = 0;
  double h_ = 0;
  size_t n_ = 0;
  int L_ = 0;
  UniformLevel lv_[kMaxLevels];
  double* b_ = nullptr;   // FP64 right-hand side (h^2-scaled)
  double* x_ = nullptr;   // FP64 pressure iterate (resident: warm start)
  float* r_ = nullptr;    // FP32 CG residual
  float* e_ = nullptr;    // FP32 inner solution
  float* q_ = nullptr;    // A d
  float* d_[2] = {nullptr, nullptr};
  float *in_u_ = nullptr, *in_v_ = nullptr, *in_w_ = nullptr, *in_q_ = nullptr;  // host-input staging
  int* cmap_ = nullptr;
  int* ncoarse_ = nullptr;
  double* minv_ = nullptr;
  int max_coarse_ = 0;
  State* st_ = nullptr;
  double* partials_ = nullptr;
  State* host_state_ = nullptr;
  State* host_upload_ = nullptr;
  void timer_begin(hipStream_t s);
  void timer_end(hipStream_t s);
  void* zbuf_ = nullptr;  // level-0 buffer holding z after the last V-cycle
  int timed_kernel_ = 0;
  int timer_used_ = 0;
  std::vector<hipEvent_t> timer_events_;
  std::vector<void*> allocations_;
  std::vector<void*> pinned_;
  size_t bytes_ = 0;
};

}  // namespace fp

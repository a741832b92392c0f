// fp_tall.hpp -- restricted tall-cell projection: per-grid device record, the
// grid resource holder and the (batched) projector.
//
// Layout (C order, z fastest): column bases (nx, nz); samples (nx, P, nz) with
// P = band + 2 -- packed row 0 the tall bottom (fine row 0), 1 the tall top
// (fine row base-1), 2+k the band cell base+k; face slots u (nx+1, S, nz),
// v (nx, P, nz), w (nx, S, nz+1), S = band + 2 + max_delta.  The operator is
// the variational reduction E^T A_u E of the uniform ghost-fluid operator on
// the expanded fine grid (pressure linear inside a tall cell, air samples at
// p = 0); the kernels evaluate it per sample in closed form from per-face
// weight moments, so no fine grid is ever materialised.
#pragma once

#include "fp_uniform.hpp"

namespace fp {

constexpr int kTallMaxLevels = 14;
constexpr int kTallMaxCoarse = 64;  // nodes of the coarsest lattice (dense FP64 inverse)
constexpr int kTallMaxK = 8;        // banded coupling reach of the coarse levels (max neighbour delta 11)
constexpr long long kTallBottomNodes = 1024;  // levels this small run inside the one-workgroup bottom of the V-cycle
constexpr int kTallSlotsMax = 5 * kTallMaxK + 3;

// Per-column FP64 face data: the weight moments sum w, sum w t, sum w t^2
// (t = row - (m-1)/2) over the rows 0..m-1 shared with the +x and the +z
// neighbour, and the summed weight of the base-1 faces inside the tall cell.
struct TallCol {
  double mx[3];
  double mz[3];
  double wint;
  double pad;
};

// One lattice level of the multigrid (levels >= 1; level 0 is the sample
// space).  Nodes (nx, ny, nz): rows 0/1 keep the tall bottom/top layers, rows
// >= 2 aggregate the band.  w holds the banded Galerkin couplings
// (fp_tall_galerkin.hpp, slot-major), dinv the scaled l1-Jacobi weight (0 on
// inactive nodes).
struct TallLevel {
  int nx, ny, nz;
  int cy;              // band rows of the finer level per band row of this one (1 or 2)
  long long n;
  float* w;
  float* dinv;
  float* r;
  float* e;
  float* e2;
  float* t;
};

struct TallRec {
  int nx, ny, nz;      // columns and fine rows
  int band, P, S, D;   // band rows, packed rows, face slots, max neighbour base delta
  int L;               // levels including the sample level
  int K;               // banded coupling reach of the coarse levels
  int lb;              // first lattice level of the V-cycle's one-workgroup bottom (L-1: the coarsest only)
  int zfinal;          // V-cycle buffer (0: z, 1: z2) holding M^-1 r
  long long n, nu, nv, nw, ncol;
  long long ns_liquid;  // unknowns at which no air face anchors the pressure: every in-domain
                        // sample, when every column's band reaches the lid (else -1: never)
  double h;
  const int* bases;    // (nx, nz)
  float* phi;          // samples (private copy)
  TallCol* cols;       // (nx, nz)
  int* top;            // (nx, nz): highest packed band row holding liquid (1: none)
  fp_f32x4* blk;       // per column: inverse (bottom, top) block bb, bt, tt and 0
  float* dinv;         // per sample: inverse diagonal of band unknowns
  double* b;           // FP64 right-hand side
  double* x;           // FP64 pressure (resident: warm start)
  float* r;            // FP32 CG residual
  float* e;            // FP32 inner solution
  float* q;
  float* d[2];
  float* z;            // V-cycle correction at the sample level (ping-pong z / z2)
  float* z2;
  float* t;            // sample-level residual
  TallLevel lv[kTallMaxLevels];
  int* reach;          // per lattice level l: [2l] the largest row offset of a nonzero x/z
                       // coupling, [2l+1] of a same-column one (k_tall_l1diag)
  int* cmap;
  int* ncoarse;
  double* minv;
  int max_coarse;
  int max_iter;
  State* st;
  double* partials;
  const float* in_u;
  const float* in_v;
  const float* in_w;
  float* out_u;
  float* out_v;
  float* out_w;
  float* out_p;
  double s_flux, inv_h, tol2, sub_scale;
  float omega;         // block-Jacobi damping at the sample level
  int warm;
  float rscale_bias;   // 1 (the tall-cell multigrid is FP32)
};

using TallPart = PartOf<TallRec>;
using TallTable = PartTableOf<TallPart>;

// Lattice plan of a layout: level dims (level 0 = samples) and band rows.
int plan_tall_levels(int nx, int band, int nz, int dims[][3], int cy[]);

class TallGrid {
 public:
  ~TallGrid();
  static size_t estimate_bytes(int nx, int ny, int nz, int band, int max_delta);
  int create(int nx, int ny, int nz, int band, int max_delta, double h, const int* bases_host);
  int bind(hipStream_t s, const fp_inputs* in, fp_outputs* out, const fp_options* opt);
  int upload(hipStream_t s);
  int finish_outputs(hipStream_t s, const fp_inputs* in, fp_outputs* out);
  // Re-lay the grid out around the surface of a->phi and move the quantities
  // to the new layout (fp_tall_remesh.hip); synchronous on s.
  int remesh(hipStream_t s, const fp_tall_remesh_args* a, int64_t* changed_columns);
  // Device bytes the first remesh adds (its staging buffers).
  static size_t remesh_bytes(int nx, int nz, int band, int max_delta);
  bool remesh_ready() const { return rm_in_ != nullptr; }
  int levels() const { return rec_->L; }
  const TallLevel& level(int l) const { return rec_->lv[l]; }
  size_t bytes() const { return bytes_; }
  TallRec* host_rec() { return rec_; }
  const TallRec* dev_rec() const { return dev_rec_; }
  int id() const { return id_; }

 private:
  int alloc(void** p, size_t bytes);
  void release();
  TallRec* rec_ = nullptr;
  TallRec* dev_rec_ = nullptr;
  float *in_u_ = nullptr, *in_v_ = nullptr, *in_w_ = nullptr;
  float *rm_in_ = nullptr, *rm_out_ = nullptr;  // remesh staging: phi, u, v, w in and out
  int* rm_b_ = nullptr;                         // remesh bases (two buffers), flag, counter
  bool host_io_ = false;
  int id_ = 0;
  std::vector<void*> allocations_;
  size_t bytes_ = 0;
};

class TallProjector {
 public:
  struct Plan;
  ~TallProjector();
  int run(hipStream_t s, const std::vector<TallGrid*>& grids, const fp_options* opt, fp_report* reps);
  void forget(int grid_id);
  // Parity probes on one grid (host arrays): out = A p in FP64; b; the layout
  // maps (fine row and unknown flag per sample).
  int probe_apply(hipStream_t s, TallGrid* g, const float* phi, const double* p, double* out);
  int probe_rhs(hipStream_t s, TallGrid* g, const fp_inputs* in, double* b);
  int probe_layout(hipStream_t s, TallGrid* g, const float* phi, int* rows, unsigned char* unknown);
  // One round of continuous batching over tall-cell jobs (the stages of
  // Projector::serve_round).
  int serve_round(hipStream_t s, std::vector<Job*>& jobs);

 private:
  void fill_tables(Plan* p, const std::vector<TallGrid*>& grids, int pre, int post);
  Plan* plan_for(const std::vector<TallGrid*>& grids, const fp_options* opt);
  int setup(hipStream_t s, Plan& p);
  int enqueue_vcycle(hipStream_t s, Plan& p);
  int enqueue_iteration(hipStream_t s, Plan& p);
  int ensure_graph(hipStream_t s, Plan& p);
  int poll(hipStream_t s, Plan& p, int* counts);
  void timer_begin(hipStream_t s);
  void timer_end(hipStream_t s);
  std::map<std::vector<long long>, Plan*> plans_;
  int timed_kernel_ = 0;
  int timer_used_ = 0;
  std::vector<hipEvent_t> timer_events_;
  int* host_counts_ = nullptr;
  int* dev_counts_ = nullptr;
  TallPart* round_arena_ = nullptr;
  size_t round_arena_cap_ = 0;
  TallPart* round_staging_ = nullptr;
  State* round_states_ = nullptr;
  State* round_states_host_ = nullptr;
  size_t round_states_cap_ = 0;
};

}  // namespace fp

// fp_uniform.hpp -- uniform-grid projection: per-grid device records, the
// grid resource holder and the (batched) projector.
//
// Every kernel of the solve is launched over a table of parts: part i covers
// workgroups [block0_i, block0_i + nblocks_i) of the launch and names the
// grid record it works on.  One launch per solver stage therefore serves any
// number of independent grids of different sizes (a single projection is a
// table of one part); each grid keeps its own block decomposition, scalars and
// reductions, so its result does not depend on what it is batched with.
#pragma once

#include "../../include/fluid_projection.h"
#include "fp_common.hpp"

#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

namespace fp {

constexpr int kMaxLevels = 16;
constexpr unsigned kMaxBlocks = 16384;
// Dynamic LDS of the LDS-resident V-cycle bottom (k_bottom_lds): with its
// static LDS (grid record, coarse right-hand side) within 64 KB.
constexpr size_t kBottomLdsMax = 48 * 1024;
constexpr long long kBottomCells = 4096;  // levels below this size run in the one-workgroup bottom of the V-cycle
constexpr int kBottomThreads = 1024;

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
  int done;              // inner solve finished (or nothing left to do)
  int outer_done;        // FP64 tolerance met
  int flags;             // FP_FLAG_*
  int iter;              // inner iterations this pass
  int total_iter;
  int outer;             // refinement passes run
  float rscale;          // power-of-two normalisation of the V-cycle input
  unsigned divmax_before, divmax_after;
  int nullspace;         // every sample is liquid: A is singular (constants)
  double bmean;          // mean of b removed on such a grid (0 otherwise)
  double xsum;           // sum of the iterate over the liquid samples
};

// Launch geometry of the column-tile kernels on one level.
struct TileGeo {
  int tz = 0, ty = 0, chunk = 0, ngroups = 0;
};

struct LevelRec {
  int d[3];
  int s[3];           // strides fine/coarse per axis from the previous level
  float g[3];         // per-axis coefficient factors (h0/h_l,a)^2
  long long n;
  float* phi;
  void* r;            // restricted right-hand side (storage T)
  void* e;            // correction (storage T); level 0: the preconditioned residual z
  void* e2;           // second correction buffer (fused sweeps ping-pong)
  void* w;            // packed coefficients W4<T> (fused levels)
  void* t;            // residual scratch (storage T)
  float Sr, Se, St;   // stored = true / S
  int fused;          // sweeps and residual run as column-tile marches
  int sweep_tz;       // sweep tile width along z (32 or 16)
  int efinal;         // buffer (0: e, 1: e2) holding e after the post-smoothing
  TileGeo sw;         // sweep geometry
  TileGeo mg;         // march geometry (tile 32 x 8)
  TileGeo rg;         // restriction into this level: the march tiles, short x chunks
  int* hull;          // per march tile: first and last plane holding liquid
                      // (fused levels, k_tile_hull; null otherwise)
};

struct GridRec {
  int nx, ny, nz, L;
  int lb;             // first level of the V-cycle's one-workgroup bottom (levels >= 1 under kBottomCells)
  long long n, nu, nv, nw;
  long long ns_liquid;  // liquid samples at which no air face anchors the pressure (all of them)
  double h;
  LevelRec lv[kMaxLevels];
  double* b;          // FP64 right-hand side (h^2-scaled)
  double* x;          // FP64 pressure iterate (resident: warm start)
  float* r;           // FP32 CG residual
  float* e;           // FP32 inner solution
  float* q;           // A d
  float* d[2];
  int* cmap;
  int* ncoarse;
  double* minv;
  int max_coarse;
  int max_iter;
  State* st;
  double* partials;
  // per-projection inputs / outputs (device) and constants
  const float* in_u;
  const float* in_v;
  const float* in_w;
  const float* in_q;
  float* out_u;
  float* out_v;
  float* out_w;
  float* out_p;
  double s_flux, s_q, inv_h, tol2, sub_scale;
  int warm;           // keep x from the previous projection (warm start)
  float rscale_bias;  // 2^precision_test_shift (1 in production)
};

// One grid's slice of a part-table launch (R: the grid record type).
template <class R>
struct PartOf {
  int block0;
  int nblocks;
  const R* g;
};
using Part = PartOf<GridRec>;

// Workgroups of a part-table launch: the part covering blockIdx.x, the grid it
// names and the workgroup's index within that grid.  Part starts are multiples
// of 8, so a grid's workgroup b runs on XCD b % 8 as tile_of_group assumes;
// workgroups in the padding return at once.
template <class P>
__device__ __forceinline__ int part_index(const P* parts, int n, int b = -1) {
  if (b < 0) b = (int)blockIdx.x;
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (cload(&parts[mid].block0) <= b) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

#define FP_PART(parts, np)                                \
  const auto fp_part_ = cload((parts) + part_index(parts, np)); \
  const auto* Gp = fp_part_.g;                            \
  const auto& G = *Gp;                                    \
  const int blk = (int)blockIdx.x - fp_part_.block0; \
  const int nblk = fp_part_.nblocks;                 \
  if (blk >= nblk) return;

// FP_PART for virtual workgroup vb of a persistent launch, whose workgroups
// walk the table's workgroups with a stride of the (multiple of 8) grid size.
#define FP_PART_AT(parts, np, vb)                                   \
  const auto fp_part_ = cload((parts) + part_index(parts, np, vb)); \
  const auto* Gp = fp_part_.g;                                      \
  const auto& G = *Gp;                                              \
  const int blk = (vb) - fp_part_.block0;                           \
  const int nblk = fp_part_.nblocks;                                \
  if (blk >= nblk) return;

// Workgroups of a persistent sweep launch (a multiple of 8: each keeps its XCD).
constexpr int kPersistGroups = 2048;

// Grid-stride loop over one grid's workgroups.
#define FP_STRIDE(idx, N) \
  for (size_t idx = (size_t)blk * blockDim.x + threadIdx.x; idx < (size_t)(N); idx += (size_t)nblk * blockDim.x)

int plan_uniform_levels(int nx, int ny, int nz, int dims[][3], int strides[][3]);

// Device scratch of fp_project_noalloc: the part tables of one grid's
// launches (at most 8 + 7 kMaxLevels parts of 16 bytes), its gathered State
// and the poll counters, each 256-byte aligned.
constexpr size_t kNoAllocScratchBytes = 4096;
// Looks up the device's copy of the scratch (at context set-up).
int noalloc_scratch_init(int device);

// ---- host helpers -----------------------------------------------------------
#define HIPCHK(expr)                                                   \
  do {                                                                 \
    hipError_t _e = (expr);                                            \
    if (_e != hipSuccess) {                                            \
      set_error("%s: %s", #expr, hipGetErrorString(_e));               \
      return FP_ERR_DEVICE;                                            \
    }                                                                  \
  } while (0)

// Launch helpers: a part-table launch skips empty tables; a finisher runs one
// workgroup per part of its producer's table.
#define PLAUNCH(kern, tab, stream, ...)                                                               \
  do {                                                                                                \
    if ((tab).n > 0)                                                                                  \
      hipLaunchKernelGGL(kern, dim3((tab).blocks), dim3(FP_BLOCK), 0, stream, (tab).dev, (tab).n, ##__VA_ARGS__); \
  } while (0)
#define FLAUNCH(kern, tab, stream, ...)                                                               \
  do {                                                                                                \
    if ((tab).n > 0) hipLaunchKernelGGL(kern, dim3((tab).n), dim3(FP_BLOCK), 0, stream, (tab).dev, ##__VA_ARGS__); \
  } while (0)


// Resources of one grid: the multigrid hierarchy, CG vectors and its record.
class UniformGrid {
 public:
  ~UniformGrid();
  static size_t estimate_bytes(int nx, int ny, int nz);
  int create(int nx, int ny, int nz, double h);

  // Per-projection set-up of the record: stage host inputs (async on s) or
  // take device pointers, the output targets and the scalars.
  int bind(hipStream_t s, const fp_inputs* in, fp_outputs* out, const fp_options* opt);
  void set_level_scales(int precision, int pre, int post);
  // Sweep tile width along z (0, 16 or 32) and workgroups per sweep (0 or
  // 256..65536) of the fused levels; 0 restores the automatic choice.
  // pairs 1: pairs of finest-level sweeps run as one launch when the
  // multigrid stores FP16/BF16; 0: one launch per sweep.
  void set_tuning(int tile_z, int groups, int pairs = 0);
  int sweep_pairs() const { return sweep_pairs_; }
  int upload(hipStream_t s);  // record -> device (async)
  int finish_outputs(hipStream_t s, const fp_inputs* in, fp_outputs* out);

  int levels() const { return rec_->L; }
  const LevelRec& level(int l) const { return rec_->lv[l]; }
  size_t bytes() const { return bytes_; }
  size_t cells() const { return (size_t)rec_->n; }
  GridRec* host_rec() { return rec_; }
  const GridRec* dev_rec() const { return dev_rec_; }
  State* dev_state() const { return rec_->st; }
  bool phi_is_staged() const { return phi_staged_; }
  int id() const { return id_; }

 private:
  int alloc(void** p, size_t bytes);
  void release();
  GridRec* rec_ = nullptr;        // pinned host copy
  GridRec* dev_rec_ = nullptr;    // device copy
  float *in_u_ = nullptr, *in_v_ = nullptr, *in_w_ = nullptr, *in_q_ = nullptr;  // host-input staging
  bool host_io_ = false;
  bool phi_staged_ = false;
  int sweep_pairs_ = 0;
  int id_ = 0;
  std::vector<void*> allocations_;
  size_t bytes_ = 0;
};

class TallGrid;
class TallProjector;

// One asynchronous projection request (fp_submit) and its progress; exactly
// one of grid / tall is set.
struct Job {
  enum Stage { SETUP, ITERATE, ACCUMULATE, FINISH, DONE };
  uint64_t ticket = 0;
  UniformGrid* grid = nullptr;
  TallGrid* tall = nullptr;
  fp_inputs in{};
  fp_outputs out{};
  fp_options opt{};
  int stage = SETUP;
  int passes = 0;     // FP64 true residuals evaluated
  int fallback = 0;   // re-solved with the BF16 multigrid after an FP16 overflow
  int rc = FP_OK;
  fp_report report{};
  double t_submit = 0, t_start = 0, t_done = 0;  // host seconds
  hipEvent_t ready = nullptr;  // recorded on the caller's stream at submit (fp_inputs.stream)
};

// Device part table of one stage.
template <class P>
struct PartTableOf {
  const P* dev = nullptr;
  int n = 0;        // parts (grids)
  int blocks = 0;   // workgroups of the launch
};
using PartTable = PartTableOf<Part>;

// Runs projections of a set of bound grids: one launch per solver stage
// covers every grid of the set.  Plans (part tables and the replayed
// iteration graph) are cached per grid set and options.
class Projector {
 public:
  struct Plan;
  Projector();
  ~Projector();
  int run(hipStream_t s, const std::vector<UniformGrid*>& grids, const fp_options* opt, fp_report* reps);
  int run_once(hipStream_t s, const std::vector<UniformGrid*>& grids, const fp_options* opt, fp_report* reps);
  // One grid's projection with every launch table, state and counter in the
  // device's static scratch (no device or pinned allocation, no graph).
  int run_noalloc(hipStream_t s, int device, UniformGrid* g, const fp_options* opt, fp_report* rep);
  // Drops cached plans that name the grid (before the grid is destroyed).
  void forget(int grid_id);
  // Parity probes on one grid.
  int probe_rhs(hipStream_t s, UniformGrid* g, double* b_host);
  int probe_vcycle(hipStream_t s, UniformGrid* g, const fp_options* opt, float rscale, void* z_out);
  int probe_level_phi(hipStream_t s, UniformGrid* g, int preserve, int level, float* out);

  // One round of continuous batching: every job advances one stage (set-up +
  // FP64 residual, one CG + V-cycle iteration, accumulation + FP64 residual,
  // or the gradient subtract); each stage group is one batched launch set.
  // Jobs whose stage became DONE carry their report.
  int serve_round(hipStream_t s, std::vector<Job*>& jobs);
  void fill_tables(Plan* p, const std::vector<UniformGrid*>& grids, int precision, int pre, int post);

 private:
  Plan* plan_for(const std::vector<UniformGrid*>& grids, const fp_options* opt);
  int setup_tables(hipStream_t s, Plan& p, int preserve);
  template <typename T> int vcycle_t(hipStream_t s, Plan& p);
  template <typename T> int iteration_t(hipStream_t s, Plan& p);
  int enqueue_vcycle(hipStream_t s, Plan& p);
  int enqueue_iteration(hipStream_t s, Plan& p);
  int ensure_graph(hipStream_t s, Plan& p);
  int solve(hipStream_t s, Plan& p, const fp_options* opt, bool graphs, int* dev_counts, int* host_counts);
  int poll(hipStream_t s, Plan& p, int* dev_counts, int* host_counts, int* counts);
  void timer_begin(hipStream_t s);
  void timer_end(hipStream_t s);
  std::map<std::vector<long long>, Plan*> plans_;
  Plan* noalloc_plan_ = nullptr;  // host tables of run_noalloc
  int timed_kernel_ = 0;
  int timer_used_ = 0;
  std::vector<hipEvent_t> timer_events_;
  int* host_counts_ = nullptr;
  int* dev_counts_ = nullptr;
  // serving: per-round tables in one reusable arena, states gathered per round
  Part* round_arena_ = nullptr;
  size_t round_arena_cap_ = 0;
  Part* round_staging_ = nullptr;
  State* round_states_ = nullptr;
  State* round_states_host_ = nullptr;
  size_t round_states_cap_ = 0;
};

// The serving worker of a context: fp_submit queues jobs, a host thread runs
// serve_round while jobs are active, fp_poll waits for a ticket.  GPU work of
// the worker and of synchronous fp_project calls is serialised by gpu_mutex.
class Server {
 public:
  Server(hipStream_t s, int device, Projector* proj, TallProjector* tall, std::mutex* gpu_mutex);
  ~Server();
  // g: a UniformGrid or a TallGrid (the other NULL)
  int submit(UniformGrid* g, TallGrid* t, const fp_inputs* in, const fp_outputs* out, const fp_options* opt,
             uint64_t* ticket);
  // FP_OK with the report when done; 1 while pending (after waiting up to timeout_ms).
  int poll(uint64_t ticket, int timeout_ms, fp_report* rep, double* latency_ms);
  int jobs_on(const void* grid);
  int pending();

 private:
  void loop();
  int jobs_on_locked(const void* grid);
  hipStream_t stream_;
  int device_;
  Projector* proj_;
  TallProjector* tall_;
  std::mutex* gpu_mutex_;
  std::mutex mu_;
  std::condition_variable work_cv_, done_cv_;
  std::deque<std::unique_ptr<Job>> incoming_;
  std::vector<std::unique_ptr<Job>> active_;
  std::map<uint64_t, std::unique_ptr<Job>> done_;
  uint64_t next_ticket_ = 1;
  bool stop_ = false;
  std::thread thread_;
};

// FP64 operator probe out = A p on the grid's finest level (host arrays).
int probe_apply(hipStream_t s, const UniformGrid* g, const float* phi_host, const double* p_host, double* out_host);
// Transfer taps between level-1 (fine) and level (coarse) along one axis.
int probe_taps(hipStream_t s, const UniformGrid* g, int level, int axis, int* pc_host, float* pw_host, int* rf_host,
               float* rw_host);

}  // namespace fp

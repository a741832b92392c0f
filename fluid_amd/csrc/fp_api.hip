// fp_api.hip -- the C ABI (include/fluid_projection.h) over the HIP solvers.
#include "fp_advect.hpp"
#include "fp_cm11a.hpp"
#include "fp_extrap.hpp"
#include "fp_tall.hpp"
#include "fp_uniform.hpp"

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <new>
#include <mutex>
#include <vector>

namespace fp {
static thread_local char g_error[1024] = "";
void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_error, sizeof(g_error), fmt, ap);
  va_end(ap);
}
}  // namespace fp

using fp::set_error;

struct fp_ctx {
  int device = 0;
  uint64_t budget = 0;
  uint64_t in_use = 0;
  hipStream_t stream = nullptr;
  hipEvent_t order = nullptr;    // orders the library stream after a caller's stream
  int grids = 0;
  fp::Projector* proj = nullptr;
  fp::TallProjector* tall = nullptr;
  fp::Server* server = nullptr;  // created by the first fp_submit
  std::mutex gpu;                // serialises synchronous calls with the serving worker
};

struct fp_grid {
  fp_ctx* ctx = nullptr;
  fp_grid_desc desc{};
  fp::UniformGrid* uniform = nullptr;  // FP_LAYOUT_UNIFORM
  fp::TallGrid* tall = nullptr;        // FP_LAYOUT_TALL_CELL
  fp::Cm11a* cm11a = nullptr;          // the CM11a solver's levels, from its first projection
  fp::Extrapolator* extrap = nullptr;  // velocity extrapolation state, from its first call
  fp::Advector* advect = nullptr;      // velocity advection state, from its first call
  uint64_t bytes = 0;
};

// Tall-cell max neighbour base delta: fp_grid_desc.max_delta, 0 meaning the
// maximumNeighborDelta of the high preset (lib/core/tall-cell-grid.ts:55).
static int tall_delta(const fp_grid_desc* d) { return d->max_delta > 0 ? d->max_delta : 4; }

#define CHECK_ARG(cond, ...)              \
  do {                                    \
    if (!(cond)) {                        \
      set_error(__VA_ARGS__);             \
      return FP_ERR_INVALID_ARGUMENT;     \
    }                                     \
  } while (0)

extern "C" {

const char* fp_last_error(void) { return fp::g_error; }

int fp_abi_version(void) { return FP_ABI_VERSION; }

int fp_ctx_create(int32_t device, uint64_t hbm_budget_bytes, fp_ctx** out) {
  CHECK_ARG(out != nullptr, "fp_ctx_create: out is NULL");
  *out = nullptr;
  int count = 0;
  hipError_t e = hipGetDeviceCount(&count);
  if (e != hipSuccess || count == 0) {
    set_error("fp_ctx_create: no HIP device (%s)", hipGetErrorString(e));
    return FP_ERR_DEVICE;
  }
  CHECK_ARG(device >= 0 && device < count, "fp_ctx_create: device %d not in [0,%d)", device, count);
  e = hipSetDevice(device);
  if (e != hipSuccess) {
    set_error("hipSetDevice(%d): %s", device, hipGetErrorString(e));
    return FP_ERR_DEVICE;
  }
  // this device's copy of the no-allocation scratch, looked up once
  int rc = fp::noalloc_scratch_init(device);
  if (rc) return rc;
  fp_ctx* ctx = new (std::nothrow) fp_ctx();
  if (!ctx) return FP_ERR_OUT_OF_MEMORY;
  ctx->device = device;
  ctx->budget = hbm_budget_bytes;
  ctx->proj = new (std::nothrow) fp::Projector();
  ctx->tall = new (std::nothrow) fp::TallProjector();
  if (!ctx->proj || !ctx->tall) {
    delete ctx->proj;
    delete ctx->tall;
    delete ctx;
    return FP_ERR_OUT_OF_MEMORY;
  }
  // default priority stream (no priority streams on this pool)
  e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&ctx->order, hipEventDisableTiming);
  if (e != hipSuccess) {
    set_error("hipStreamCreate: %s", hipGetErrorString(e));
    delete ctx->proj;
    delete ctx->tall;
    delete ctx;
    return FP_ERR_DEVICE;
  }
  *out = ctx;
  return FP_OK;
}

int fp_ctx_destroy(fp_ctx* ctx) {
  if (!ctx) return FP_OK;
  if (ctx->grids != 0) {
    set_error("fp_ctx_destroy: %d grids still alive", ctx->grids);
    return FP_ERR_STATE;
  }
  (void)hipSetDevice(ctx->device);
  delete ctx->server;
  ctx->server = nullptr;
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  delete ctx->proj;
  delete ctx->tall;
  if (ctx->order) (void)hipEventDestroy(ctx->order);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return FP_OK;
}

// A grid with a submitted projection that has not been polled to completion.
static bool grid_busy(const fp_grid* g) {
  return g->ctx->server && g->ctx->server->jobs_on(g->uniform ? (const void*)g->uniform : (const void*)g->tall) > 0;
}

#define CHECK_IDLE(g, what)                                                                      \
  do {                                                                                           \
    if (grid_busy(g)) {                                                                          \
      set_error("%s: the grid has a submitted projection in flight (fp_poll it first)", what);   \
      return FP_ERR_STATE;                                                                       \
    }                                                                                            \
  } while (0)

int fp_ctx_bytes_in_use(const fp_ctx* ctx, uint64_t* bytes) {
  CHECK_ARG(ctx && bytes, "fp_ctx_bytes_in_use: NULL argument");
  *bytes = ctx->in_use;
  return FP_OK;
}

int fp_default_options(fp_options* o) {
  CHECK_ARG(o != nullptr, "fp_default_options: NULL");
  std::memset(o, 0, sizeof(*o));
  o->relative_tolerance = 1e-6;
  o->max_iterations = 100;
  o->max_outer = 4;
  o->precision = FP_PRECISION_MIXED_FP16;
  o->pre_sweeps = 2;
  o->post_sweeps = 2;
  o->warm_start = 0;
  o->phi_preserve_levels = 2;
  o->check_every = 1;
  o->timed_kernel = FP_KERNEL_NONE;
  o->graphs = 1;
  return FP_OK;
}

static int validate_desc(const fp_grid_desc* d) {
  CHECK_ARG(d != nullptr, "grid description is NULL");
  CHECK_ARG(d->layout == FP_LAYOUT_UNIFORM || d->layout == FP_LAYOUT_TALL_CELL, "unknown layout %d", d->layout);
  CHECK_ARG(d->nx >= 2 && d->ny >= 2 && d->nz >= 2, "grid dims must be >= 2 (got %d x %d x %d)", d->nx, d->ny, d->nz);
  CHECK_ARG((uint64_t)d->nx * d->ny * d->nz < (1ull << 34), "grid too large");
  CHECK_ARG(d->h > 0.0, "cell size must be positive");
  if (d->layout == FP_LAYOUT_TALL_CELL) {
    CHECK_ARG(d->column_bases != nullptr, "tall-cell grid: column_bases is NULL");
    CHECK_ARG(d->band >= 1 && d->band <= 1024, "tall-cell band rows must be in [1,1024] (got %d)", d->band);
    CHECK_ARG(d->max_delta >= 0 && tall_delta(d) <= d->band, "tall-cell max neighbour delta %d must be in [1, band]",
              tall_delta(d));
  }
  return FP_OK;
}

int fp_grid_bytes(const fp_grid_desc* desc, int32_t precision, uint64_t* bytes) {
  (void)precision;
  int rc = validate_desc(desc);
  if (rc) return rc;
  CHECK_ARG(bytes != nullptr, "bytes is NULL");
  const size_t b = desc->layout == FP_LAYOUT_TALL_CELL
                       ? fp::TallGrid::estimate_bytes(desc->nx, desc->ny, desc->nz, desc->band, tall_delta(desc))
                       : fp::UniformGrid::estimate_bytes(desc->nx, desc->ny, desc->nz);
  if (b == 0) {
    set_error("no multigrid hierarchy for %d x %d x %d", desc->nx, desc->ny, desc->nz);
    return FP_ERR_LAYOUT;
  }
  *bytes = b;
  return FP_OK;
}

int fp_grid_create(fp_ctx* ctx, const fp_grid_desc* desc, fp_grid** out) {
  CHECK_ARG(ctx && out, "fp_grid_create: NULL argument");
  *out = nullptr;
  int rc = validate_desc(desc);
  if (rc) return rc;
  uint64_t need = 0;
  rc = fp_grid_bytes(desc, 0, &need);
  if (rc) return rc;
  if (ctx->budget && ctx->in_use + need > ctx->budget) {
    set_error("fp_grid_create: %llu bytes would exceed the HBM budget (%llu of %llu in use)",
              (unsigned long long)need, (unsigned long long)ctx->in_use, (unsigned long long)ctx->budget);
    return FP_ERR_BUDGET;
  }
  (void)hipSetDevice(ctx->device);
  fp_grid* g = new (std::nothrow) fp_grid();
  if (!g) return FP_ERR_OUT_OF_MEMORY;
  g->ctx = ctx;
  g->desc = *desc;
  g->desc.column_bases = nullptr;
  if (desc->layout == FP_LAYOUT_TALL_CELL) {
    g->tall = new (std::nothrow) fp::TallGrid();
    if (!g->tall) { delete g; return FP_ERR_OUT_OF_MEMORY; }
    rc = g->tall->create(desc->nx, desc->ny, desc->nz, desc->band, tall_delta(desc), desc->h, desc->column_bases);
    if (rc) {
      delete g->tall;
      delete g;
      return rc;
    }
    g->bytes = g->tall->bytes();
  } else {
    g->uniform = new (std::nothrow) fp::UniformGrid();
    if (!g->uniform) { delete g; return FP_ERR_OUT_OF_MEMORY; }
    rc = g->uniform->create(desc->nx, desc->ny, desc->nz, desc->h);
    if (rc) {
      delete g->uniform;
      delete g;
      return rc;
    }
    g->bytes = g->uniform->bytes();
  }
  ctx->in_use += g->bytes;
  ctx->grids += 1;
  *out = g;
  return FP_OK;
}

int fp_grid_destroy(fp_grid* g) {
  if (!g) return FP_OK;
  CHECK_IDLE(g, "fp_grid_destroy");
  std::lock_guard<std::mutex> lock(g->ctx->gpu);
  (void)hipSetDevice(g->ctx->device);
  (void)hipStreamSynchronize(g->ctx->stream);
  g->ctx->in_use -= g->bytes;
  g->ctx->grids -= 1;
  if (g->uniform) g->ctx->proj->forget(g->uniform->id());
  if (g->tall) g->ctx->tall->forget(g->tall->id());
  if (g->cm11a) g->ctx->in_use -= g->cm11a->bytes();
  if (g->extrap) g->ctx->in_use -= g->extrap->bytes();
  if (g->advect) g->ctx->in_use -= g->advect->bytes();
  delete g->cm11a;
  delete g->extrap;
  delete g->advect;
  delete g->uniform;
  delete g->tall;
  delete g;
  return FP_OK;
}

int fp_grid_levels(const fp_grid* g, int32_t max_levels, int32_t* dims, int32_t* strides, int32_t* levels) {
  CHECK_ARG(g && levels, "fp_grid_levels: NULL argument");
  if (g->tall) {
    // sample level, then the lattice levels; strides: the aggregation factors
    const fp::TallRec& R = *g->tall->host_rec();
    *levels = R.L;
    for (int l = 0; l < R.L && l < max_levels; ++l) {
      const fp::TallLevel& lv = R.lv[l];
      const int d3[3] = {lv.nx, lv.ny, lv.nz};
      const int s3[3] = {l == 0 ? 1 : 2, l == 0 ? 1 : lv.cy, l == 0 ? 1 : 2};
      for (int a = 0; a < 3; ++a) {
        if (dims) dims[l * 3 + a] = d3[a];
        if (strides) strides[l * 3 + a] = s3[a];
      }
    }
    return FP_OK;
  }
  const int L = g->uniform->levels();
  *levels = L;
  for (int l = 0; l < L && l < max_levels; ++l)
    for (int a = 0; a < 3; ++a) {
      if (dims) dims[l * 3 + a] = g->uniform->level(l).d[a];
      if (strides) strides[l * 3 + a] = g->uniform->level(l).s[a];
    }
  return FP_OK;
}

int fp_grid_set_tuning(fp_grid* g, const fp_tuning* t) {
  CHECK_ARG(g && t, "fp_grid_set_tuning: NULL argument");
  if (!g->uniform) {
    set_error("fp_grid_set_tuning: uniform grids only");
    return FP_ERR_UNSUPPORTED;
  }
  CHECK_ARG(t->sweep_tile_z == 0 || t->sweep_tile_z == 16 || t->sweep_tile_z == 32,
            "sweep_tile_z must be 0, 16 or 32 (got %d)", t->sweep_tile_z);
  CHECK_ARG(t->sweep_tile_z != 32 || g->desc.nz >= 32, "32-wide sweep tiles need nz >= 32 (nz = %d)", g->desc.nz);
  CHECK_ARG(t->sweep_groups == 0 || (t->sweep_groups >= 256 && t->sweep_groups <= 65536),
            "sweep_groups must be 0 or in [256, 65536] (got %d)", t->sweep_groups);
  CHECK_ARG(t->sweep_pairs == 0 || t->sweep_pairs == 1, "sweep_pairs must be 0 or 1 (got %d)", t->sweep_pairs);
  CHECK_IDLE(g, "fp_grid_set_tuning");
  std::lock_guard<std::mutex> lock(g->ctx->gpu);
  (void)hipSetDevice(g->ctx->device);
  (void)hipStreamSynchronize(g->ctx->stream);
  g->ctx->proj->forget(g->uniform->id());  // cached plans hold the old launch shapes
  g->uniform->set_tuning(t->sweep_tile_z, t->sweep_groups, t->sweep_pairs);
  return FP_OK;
}

int fp_grid_face_counts(const fp_grid* g, int64_t* nu, int64_t* nv, int64_t* nw) {
  CHECK_ARG(g && nu && nv && nw, "fp_grid_face_counts: NULL argument");
  const fp_grid_desc& d = g->desc;
  if (g->tall) {
    const int P = d.band + 2, S = d.band + 2 + tall_delta(&d);
    *nu = (int64_t)(d.nx + 1) * S * d.nz;
    *nv = (int64_t)d.nx * P * d.nz;
    *nw = (int64_t)d.nx * S * (d.nz + 1);
    return FP_OK;
  }
  *nu = (int64_t)(d.nx + 1) * d.ny * d.nz;
  *nv = (int64_t)d.nx * (d.ny + 1) * d.nz;
  *nw = (int64_t)d.nx * d.ny * (d.nz + 1);
  return FP_OK;
}

static int validate_inputs(const fp_inputs* in, const fp_outputs* out, const fp_options* o) {
  CHECK_ARG(in && out && o, "fp_project: NULL argument");
  CHECK_ARG(in->memory == FP_MEMORY_HOST || in->memory == FP_MEMORY_DEVICE, "unknown memory kind %d", in->memory);
  CHECK_ARG((in->phi || in->density) && in->u && in->v && in->w, "fp_project: phi (or density)/u/v/w must be given");
  CHECK_ARG(in->dt > 0.0 && in->rho > 0.0, "fp_project: dt and rho must be positive");
  CHECK_ARG(o->relative_tolerance > 0.0 && o->relative_tolerance < 1.0, "relative_tolerance must be in (0,1)");
  CHECK_ARG(o->max_iterations >= 1 && o->max_outer >= 1, "iteration limits must be positive");
  CHECK_ARG(o->precision >= 0 && o->precision <= 2, "unknown precision %d", o->precision);
  CHECK_ARG(o->pre_sweeps >= 0 && o->post_sweeps >= 0 && o->pre_sweeps <= 16 && o->post_sweeps <= 16,
            "sweep counts must be in [0,16]");
  CHECK_ARG(o->pre_sweeps == o->post_sweeps, "pre and post sweeps must match (symmetric preconditioner)");
  CHECK_ARG(o->precision_test_shift >= 0 && o->precision_test_shift <= 64, "precision_test_shift must be in [0,64]");
  CHECK_ARG(o->solver == FP_SOLVER_MGPCG || o->solver == FP_SOLVER_CM11A, "unknown solver %d", o->solver);
  CHECK_ARG(o->open_top == 0 || o->open_top == 1, "open_top must be 0 or 1 (got %d)", o->open_top);
  CHECK_ARG(o->full_cycles >= 0 && o->full_cycles <= 64, "full_cycles must be in [0,64]");
  CHECK_ARG(o->v_cycles >= 0 && o->v_cycles <= 256, "v_cycles must be in [0,256]");
  CHECK_ARG(o->lcp_sweeps >= 0 && o->lcp_sweeps <= 64, "lcp_sweeps must be in [0,64]");
  if (o->solver == FP_SOLVER_MGPCG && o->open_top) {
    set_error("an open top needs FP_SOLVER_CM11A (the MGPCG system has closed walls)");
    return FP_ERR_UNSUPPORTED;
  }
  return FP_OK;
}

// One CM11a projection of a uniform grid (fp_project with FP_SOLVER_CM11A).
static int project_cm11a(fp_grid* g, const fp_inputs* in, fp_outputs* out, const fp_options* o, fp_report* rep) {
  fp_ctx* ctx = g->ctx;
  if (!g->uniform) {
    set_error("FP_SOLVER_CM11A: uniform grids only");
    return FP_ERR_UNSUPPORTED;
  }
  std::lock_guard<std::mutex> lock(ctx->gpu);
  (void)hipSetDevice(ctx->device);
  if (!g->cm11a) {
    const fp_grid_desc& d = g->desc;
    const uint64_t need = fp::Cm11a::estimate_bytes(d.nx, d.ny, d.nz);
    if (ctx->budget && ctx->in_use + need > ctx->budget) {
      set_error("FP_SOLVER_CM11A: %llu more bytes exceed the context's HBM budget (%llu of %llu in use)",
                (unsigned long long)need, (unsigned long long)ctx->in_use, (unsigned long long)ctx->budget);
      return FP_ERR_BUDGET;
    }
    fp::Cm11a* c = new (std::nothrow) fp::Cm11a();
    if (!c) return FP_ERR_OUT_OF_MEMORY;
    int rc = c->create(d.nx, d.ny, d.nz, d.h);
    if (rc) {
      delete c;
      return rc;
    }
    g->cm11a = c;
    ctx->in_use += c->bytes();
  }
  if (in->stream_order) {
    hipError_t e = hipEventRecord(ctx->order, (hipStream_t)in->stream);
    if (e == hipSuccess) e = hipStreamWaitEvent(ctx->stream, ctx->order, 0);
    if (e != hipSuccess) {
      set_error("fp_project: ordering after the caller's stream: %s", hipGetErrorString(e));
      return FP_ERR_DEVICE;
    }
  }
  int rc = g->uniform->bind(ctx->stream, in, out, o);
  if (rc) return rc;
  rc = g->cm11a->run(ctx->stream, g->uniform, o, in->dt, in->rho, rep);
  if (rc) return rc;
  rc = g->uniform->finish_outputs(ctx->stream, in, out);
  if (rc) return rc;
  hipError_t e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) {
    set_error("fp_project: %s", hipGetErrorString(e));
    return FP_ERR_DEVICE;
  }
  return FP_OK;
}

int fp_project(fp_grid* g, const fp_inputs* in, fp_outputs* out, const fp_options* o, fp_report* rep) {
  CHECK_ARG(g != nullptr, "fp_project: grid is NULL");
  return fp_project_batch(&g, 1, in, out, o, rep);
}

int fp_project_batch(fp_grid* const* grids, int32_t count, const fp_inputs* in, fp_outputs* out,
                     const fp_options* o, fp_report* reps) {
  CHECK_ARG(grids != nullptr && count >= 1, "fp_project_batch: no grids");
  CHECK_ARG(in != nullptr && out != nullptr, "fp_project_batch: NULL inputs/outputs");
  fp_ctx* ctx = grids[0] ? grids[0]->ctx : nullptr;
  CHECK_ARG(ctx != nullptr, "fp_project_batch: grid 0 is NULL");
  if (o && o->solver == FP_SOLVER_CM11A) {
    if (count != 1) {
      set_error("FP_SOLVER_CM11A: one grid per call (fp_project)");
      return FP_ERR_UNSUPPORTED;
    }
    CHECK_IDLE(grids[0], "fp_project");
    int rc = validate_inputs(&in[0], &out[0], o);
    if (rc) return rc;
    return project_cm11a(grids[0], &in[0], &out[0], o, reps);
  }
  std::vector<fp::UniformGrid*> list;
  std::vector<fp::TallGrid*> tlist;
  std::vector<int> ui, ti;  // batch positions of the uniform / tall-cell grids
  for (int i = 0; i < count; ++i) {
    CHECK_ARG(grids[i] != nullptr, "fp_project_batch: grid %d is NULL", i);
    CHECK_ARG(grids[i]->ctx == ctx, "fp_project_batch: grid %d belongs to another context", i);
    for (int j = 0; j < i; ++j) CHECK_ARG(grids[j] != grids[i], "fp_project_batch: grid %d repeated", i);
    CHECK_IDLE(grids[i], "fp_project");
    int rc = validate_inputs(&in[i], &out[i], o);
    if (rc) return rc;
    if (grids[i]->tall) {
      if (in[i].density) {
        set_error("fp_project: density inputs are not supported on tall-cell grids (pass phi)");
        return FP_ERR_UNSUPPORTED;
      }
      if (in[i].target_divergence) {
        set_error("fp_project: target_divergence is not supported on tall-cell grids");
        return FP_ERR_UNSUPPORTED;
      }
      tlist.push_back(grids[i]->tall);
      ti.push_back(i);
    } else {
      list.push_back(grids[i]->uniform);
      ui.push_back(i);
    }
  }
  std::lock_guard<std::mutex> lock(ctx->gpu);
  (void)hipSetDevice(ctx->device);
  // device work after what the callers queued on their streams
  for (int i = 0; i < count; ++i) {
    if (!in[i].stream_order) continue;
    hipError_t e = hipEventRecord(ctx->order, (hipStream_t)in[i].stream);
    if (e == hipSuccess) e = hipStreamWaitEvent(ctx->stream, ctx->order, 0);
    if (e != hipSuccess) {
      set_error("fp_project: ordering after the caller's stream: %s", hipGetErrorString(e));
      return FP_ERR_DEVICE;
    }
  }
  for (size_t k = 0; k < list.size(); ++k) {
    int rc = list[k]->bind(ctx->stream, &in[ui[k]], &out[ui[k]], o);
    if (rc) return rc;
  }
  for (size_t k = 0; k < tlist.size(); ++k) {
    int rc = tlist[k]->bind(ctx->stream, &in[ti[k]], &out[ti[k]], o);
    if (rc) return rc;
  }
  std::vector<fp_report> ureps(list.size()), treps(tlist.size());
  int rc = ctx->proj->run(ctx->stream, list, o, reps ? ureps.data() : nullptr);
  if (rc) return rc;
  rc = ctx->tall->run(ctx->stream, tlist, o, reps ? treps.data() : nullptr);
  if (rc) return rc;
  if (reps) {
    for (size_t k = 0; k < list.size(); ++k) reps[ui[k]] = ureps[k];
    for (size_t k = 0; k < tlist.size(); ++k) reps[ti[k]] = treps[k];
  }
  for (size_t k = 0; k < list.size(); ++k) {
    rc = list[k]->finish_outputs(ctx->stream, &in[ui[k]], &out[ui[k]]);
    if (rc) return rc;
  }
  for (size_t k = 0; k < tlist.size(); ++k) {
    rc = tlist[k]->finish_outputs(ctx->stream, &in[ti[k]], &out[ti[k]]);
    if (rc) return rc;
  }
  hipError_t e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) {
    set_error("fp_project: %s", hipGetErrorString(e));
    return FP_ERR_DEVICE;
  }
  return FP_OK;
}

int fp_project_noalloc(fp_grid* g, const fp_inputs* in, fp_outputs* out, const fp_options* o, fp_report* rep) {
  CHECK_ARG(g != nullptr, "fp_project_noalloc: grid is NULL");
  CHECK_ARG(g->uniform != nullptr, "fp_project_noalloc: uniform grids only");
  CHECK_IDLE(g, "fp_project_noalloc");
  int rc = validate_inputs(in, out, o);
  if (rc) return rc;
  if (o->solver != FP_SOLVER_MGPCG) {
    set_error("fp_project_noalloc: FP_SOLVER_MGPCG only");
    return FP_ERR_UNSUPPORTED;
  }
  fp_ctx* ctx = g->ctx;
  std::lock_guard<std::mutex> lock(ctx->gpu);
  (void)hipSetDevice(ctx->device);
  if (in->stream_order) {
    hipError_t e = hipEventRecord(ctx->order, (hipStream_t)in->stream);
    if (e == hipSuccess) e = hipStreamWaitEvent(ctx->stream, ctx->order, 0);
    if (e != hipSuccess) {
      set_error("fp_project_noalloc: ordering after the caller's stream: %s", hipGetErrorString(e));
      return FP_ERR_DEVICE;
    }
  }
  rc = g->uniform->bind(ctx->stream, in, out, o);
  if (rc) return rc;
  rc = ctx->proj->run_noalloc(ctx->stream, ctx->device, g->uniform, o, rep);
  if (rc) return rc;
  rc = g->uniform->finish_outputs(ctx->stream, in, out);
  if (rc) return rc;
  hipError_t e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) {
    set_error("fp_project_noalloc: %s", hipGetErrorString(e));
    return FP_ERR_DEVICE;
  }
  return FP_OK;
}

int fp_submit(fp_grid* g, const fp_inputs* in, fp_outputs* out, const fp_options* o, uint64_t* ticket) {
  CHECK_ARG(g != nullptr && ticket != nullptr, "fp_submit: NULL argument");
  int rc = validate_inputs(in, out, o);
  if (rc) return rc;
  if (o->solver != FP_SOLVER_MGPCG) {
    set_error("fp_submit: FP_SOLVER_MGPCG only (CM11a projections run through fp_project)");
    return FP_ERR_UNSUPPORTED;
  }
  fp_ctx* ctx = g->ctx;
  if (g->tall && in->density) {
    set_error("fp_submit: density inputs are not supported on tall-cell grids (pass phi)");
    return FP_ERR_UNSUPPORTED;
  }
  if (g->tall && in->target_divergence) {
    set_error("fp_submit: target_divergence is not supported on tall-cell grids");
    return FP_ERR_UNSUPPORTED;
  }
  if (!ctx->server) {
    ctx->server = new (std::nothrow) fp::Server(ctx->stream, ctx->device, ctx->proj, ctx->tall, &ctx->gpu);
    if (!ctx->server) return FP_ERR_OUT_OF_MEMORY;
  }
  return ctx->server->submit(g->uniform, g->tall, in, out, o, ticket);  // FP_ERR_STATE on a busy grid
}

int fp_poll(fp_ctx* ctx, uint64_t ticket, int32_t timeout_ms, fp_report* report, double* latency_ms) {
  CHECK_ARG(ctx != nullptr, "fp_poll: context is NULL");
  if (!ctx->server) {
    set_error("fp_poll: nothing was submitted on this context");
    return FP_ERR_INVALID_ARGUMENT;
  }
  return ctx->server->poll(ticket, timeout_ms, report, latency_ms);
}

int fp_probe_apply(fp_grid* g, const float* phi, const double* p, double* out) {
  CHECK_ARG(g && phi && p && out, "fp_probe_apply: NULL argument");
  CHECK_IDLE(g, "fp_probe_apply");
  std::lock_guard<std::mutex> lock(g->ctx->gpu);
  (void)hipSetDevice(g->ctx->device);
  if (g->tall) return g->ctx->tall->probe_apply(g->ctx->stream, g->tall, phi, p, out);
  return fp::probe_apply(g->ctx->stream, g->uniform, phi, p, out);
}

int fp_probe_rhs(fp_grid* g, const fp_inputs* in, double* b) {
  CHECK_ARG(g && in && b && in->phi && in->u && in->v && in->w, "fp_probe_rhs: NULL argument");
  CHECK_ARG(in->memory == FP_MEMORY_HOST, "fp_probe_rhs takes host arrays");
  CHECK_IDLE(g, "fp_probe_rhs");
  std::lock_guard<std::mutex> lock(g->ctx->gpu);
  (void)hipSetDevice(g->ctx->device);
  if (g->tall) return g->ctx->tall->probe_rhs(g->ctx->stream, g->tall, in, b);
  int rc = g->uniform->bind(g->ctx->stream, in, nullptr, nullptr);
  if (rc) return rc;
  return g->ctx->proj->probe_rhs(g->ctx->stream, g->uniform, b);
}

int fp_probe_vcycle(fp_grid* g, const float* phi, const fp_options* o, const float* r, float* z) {
  CHECK_ARG(g && phi && o && r && z, "fp_probe_vcycle: NULL argument");
  CHECK_ARG(g->uniform != nullptr, "fp_probe_vcycle: uniform grids only");
  CHECK_IDLE(g, "fp_probe_vcycle");
  CHECK_ARG(o->precision >= 0 && o->precision <= 2, "unknown precision %d", o->precision);
  CHECK_ARG(o->pre_sweeps == o->post_sweeps && o->pre_sweeps >= 0 && o->pre_sweeps <= 16,
            "pre and post sweeps must match and lie in [0,16]");
  std::lock_guard<std::mutex> lock(g->ctx->gpu);
  (void)hipSetDevice(g->ctx->device);
  fp::GridRec& R = *g->uniform->host_rec();
  const size_t n = (size_t)R.n;
  hipError_t e = hipMemcpy(R.lv[0].phi, phi, n * 4, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(R.r, r, n * 4, hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    set_error("fp_probe_vcycle: %s", hipGetErrorString(e));
    return FP_ERR_DEVICE;
  }
  // normalise like the CG does: rscale = 2^-ceil(log2 max|r|)
  float m = 0.0f;
  for (size_t i = 0; i < n; ++i) m = std::max(m, std::fabs(r[i]));
  const float rs = m > 0.0f ? std::exp2(-std::ceil(std::log2(m))) : 1.0f;
  std::vector<uint32_t> raw(n);
  int rc = g->ctx->proj->probe_vcycle(g->ctx->stream, g->uniform, o, rs, raw.data());
  if (rc) return rc;
  const float se = R.lv[0].Se;
  if (o->precision == FP_PRECISION_MIXED_FP16) {
    const __half* h = reinterpret_cast<const __half*>(raw.data());
    for (size_t i = 0; i < n; ++i) z[i] = __half2float(h[i]) * se / rs;
  } else if (o->precision == FP_PRECISION_MIXED_BF16) {
    const __hip_bfloat16* h = reinterpret_cast<const __hip_bfloat16*>(raw.data());
    for (size_t i = 0; i < n; ++i) z[i] = __bfloat162float(h[i]) * se / rs;
  } else {
    const float* f = reinterpret_cast<const float*>(raw.data());
    for (size_t i = 0; i < n; ++i) z[i] = f[i] * se / rs;
  }
  return FP_OK;
}

int fp_probe_level_phi(fp_grid* g, const float* phi, int32_t level, int32_t preserve, float* out) {
  CHECK_ARG(g && phi && out, "fp_probe_level_phi: NULL argument");
  CHECK_ARG(g->uniform != nullptr, "fp_probe_level_phi: uniform grids only");
  CHECK_IDLE(g, "fp_probe_level_phi");
  CHECK_ARG(level >= 0 && level < g->uniform->levels(), "level %d out of range [0,%d)", level, g->uniform->levels());
  std::lock_guard<std::mutex> lock(g->ctx->gpu);
  (void)hipSetDevice(g->ctx->device);
  const fp::GridRec& R = *g->uniform->host_rec();
  hipError_t e = hipMemcpy(R.lv[0].phi, phi, (size_t)R.n * 4, hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    set_error("fp_probe_level_phi: %s", hipGetErrorString(e));
    return FP_ERR_DEVICE;
  }
  return g->ctx->proj->probe_level_phi(g->ctx->stream, g->uniform, preserve, level, out);
}

int fp_probe_taps(fp_grid* g, int32_t level, int32_t axis, int32_t* prolong_coarse, float* prolong_weight,
                  int32_t* restrict_fine, float* restrict_weight) {
  CHECK_ARG(g && prolong_coarse && prolong_weight && restrict_fine && restrict_weight, "fp_probe_taps: NULL argument");
  CHECK_ARG(g->uniform != nullptr, "fp_probe_taps: uniform grids only");
  CHECK_ARG(level >= 1 && level < g->uniform->levels(), "level %d out of range [1,%d)", level, g->uniform->levels());
  CHECK_ARG(axis >= 0 && axis < 3, "axis %d out of range [0,3)", axis);
  std::lock_guard<std::mutex> lock(g->ctx->gpu);
  (void)hipSetDevice(g->ctx->device);
  return fp::probe_taps(g->ctx->stream, g->uniform, level, axis, prolong_coarse, prolong_weight, restrict_fine,
                        restrict_weight);
}

// The grid's extrapolation state, created on first use within the context's budget.
static int ensure_extrapolator(fp_grid* g, const char* what) {
  if (g->extrap) return FP_OK;
  fp_ctx* ctx = g->ctx;
  const fp_grid_desc& d = g->desc;
  const uint64_t need = fp::Extrapolator::estimate_bytes(d.nx, d.ny, d.nz);
  if (ctx->budget && ctx->in_use + need > ctx->budget) {
    set_error("%s: %llu more bytes exceed the context's HBM budget (%llu of %llu in use)", what,
              (unsigned long long)need, (unsigned long long)ctx->in_use, (unsigned long long)ctx->budget);
    return FP_ERR_BUDGET;
  }
  fp::Extrapolator* e = new (std::nothrow) fp::Extrapolator();
  if (!e) return FP_ERR_OUT_OF_MEMORY;
  int rc = e->create(d.nx, d.ny, d.nz, d.h);
  if (rc) {
    delete e;
    return rc;
  }
  g->extrap = e;
  ctx->in_use += e->bytes();
  return FP_OK;
}

int fp_extrapolate_velocity(fp_grid* g, const fp_extrapolate_args* a, fp_extrapolate_report* rep) {
  CHECK_ARG(g && a, "fp_extrapolate_velocity: NULL argument");
  CHECK_ARG(g->uniform != nullptr, "fp_extrapolate_velocity: uniform grids only");
  CHECK_ARG(a->memory == FP_MEMORY_HOST || a->memory == FP_MEMORY_DEVICE,
            "fp_extrapolate_velocity: memory must be 0 or 1");
  CHECK_ARG(a->phi || a->density, "fp_extrapolate_velocity: phi or density is required");
  CHECK_ARG(a->u && a->v && a->w && a->u_out && a->v_out && a->w_out,
            "fp_extrapolate_velocity: u, v, w and their outputs are required");
  CHECK_ARG(a->max_passes >= 0 && a->max_passes <= 16, "fp_extrapolate_velocity: max_passes must be in [0,16]");
  CHECK_IDLE(g, "fp_extrapolate_velocity");
  fp_ctx* ctx = g->ctx;
  std::lock_guard<std::mutex> lock(ctx->gpu);
  (void)hipSetDevice(ctx->device);
  int rc = ensure_extrapolator(g, "fp_extrapolate_velocity");
  if (rc) return rc;
  return g->extrap->run(ctx->stream, a, rep);
}

int fp_advect_velocity(fp_grid* g, const fp_advect_args* a, fp_advect_report* rep) {
  CHECK_ARG(g && a, "fp_advect_velocity: NULL argument");
  CHECK_ARG(g->uniform != nullptr, "fp_advect_velocity: uniform grids only");
  CHECK_ARG(a->memory == FP_MEMORY_HOST || a->memory == FP_MEMORY_DEVICE, "fp_advect_velocity: memory must be 0 or 1");
  CHECK_ARG(a->scheme == FP_ADVECT_SEMI_LAGRANGIAN || a->scheme == FP_ADVECT_MACCORMACK,
            "fp_advect_velocity: unknown scheme %d", a->scheme);
  CHECK_ARG(a->phi || a->density, "fp_advect_velocity: phi or density is required");
  CHECK_ARG(a->u && a->v && a->w && a->u_out && a->v_out && a->w_out,
            "fp_advect_velocity: u, v, w and their outputs are required");
  CHECK_ARG((a->tu && a->tv && a->tw) || (!a->tu && !a->tv && !a->tw),
            "fp_advect_velocity: give all of tu, tv, tw or none");
  CHECK_ARG(std::isfinite(a->dt) && a->dt > 0.0 && std::isfinite(a->gravity) && std::isfinite(a->viscosity) &&
                a->viscosity >= 0.0,
            "fp_advect_velocity: dt must be positive, gravity and viscosity finite, viscosity >= 0");
  CHECK_IDLE(g, "fp_advect_velocity");
  fp_ctx* ctx = g->ctx;
  std::lock_guard<std::mutex> lock(ctx->gpu);
  (void)hipSetDevice(ctx->device);
  int rc = ensure_extrapolator(g, "fp_advect_velocity");
  if (rc) return rc;
  if (!g->advect) {
    const fp_grid_desc& d = g->desc;
    const uint64_t need = fp::Advector::estimate_bytes(d.nx, d.ny, d.nz);
    if (ctx->budget && ctx->in_use + need > ctx->budget) {
      set_error("fp_advect_velocity: %llu more bytes exceed the context's HBM budget (%llu of %llu in use)",
                (unsigned long long)need, (unsigned long long)ctx->in_use, (unsigned long long)ctx->budget);
      return FP_ERR_BUDGET;
    }
    fp::Advector* v = new (std::nothrow) fp::Advector();
    if (!v) return FP_ERR_OUT_OF_MEMORY;
    rc = v->create(d.nx, d.ny, d.nz, d.h);
    if (rc) {
      delete v;
      return rc;
    }
    g->advect = v;
    ctx->in_use += v->bytes();
  }
  const size_t before = g->advect->bytes();
  rc = g->advect->run(ctx->stream, g->extrap, a, rep);
  ctx->in_use += g->advect->bytes() - before;  // host staging, on the first host-memory call
  return rc;
}

int fp_tall_remesh(fp_grid* g, const fp_tall_remesh_args* a, int64_t* changed_columns) {
  CHECK_ARG(g && a, "fp_tall_remesh: NULL argument");
  CHECK_ARG(g->tall != nullptr, "fp_tall_remesh: tall-cell grids only");
  CHECK_ARG(a->memory == FP_MEMORY_HOST || a->memory == FP_MEMORY_DEVICE, "fp_tall_remesh: memory must be 0 or 1");
  CHECK_ARG(a->phi && a->u && a->v && a->w, "fp_tall_remesh: phi, u, v and w are required");
  CHECK_ARG(a->liquid_halo >= 0 && a->air_halo >= 0, "fp_tall_remesh: halos must be >= 0");
  CHECK_IDLE(g, "fp_tall_remesh");
  fp_ctx* ctx = g->ctx;
  std::lock_guard<std::mutex> lock(ctx->gpu);
  (void)hipSetDevice(ctx->device);
  if (!g->tall->remesh_ready()) {
    const fp_grid_desc& d = g->desc;
    const uint64_t need = fp::TallGrid::remesh_bytes(d.nx, d.nz, d.band, tall_delta(&d));
    if (ctx->budget && ctx->in_use + need > ctx->budget) {
      set_error("fp_tall_remesh: %llu more bytes exceed the context's HBM budget (%llu of %llu in use)",
                (unsigned long long)need, (unsigned long long)ctx->in_use, (unsigned long long)ctx->budget);
      return FP_ERR_BUDGET;
    }
  }
  const size_t before = g->tall->bytes();
  const int rc = g->tall->remesh(ctx->stream, a, changed_columns);
  const size_t grown = g->tall->bytes() - before;
  ctx->in_use += grown;
  g->bytes += grown;
  // plans hold per-grid state of the old layout
  ctx->tall->forget(g->tall->id());
  return rc;
}

int fp_probe_tall_layout(fp_grid* g, const float* phi, int32_t* rows, uint8_t* unknown) {
  CHECK_ARG(g && phi && rows && unknown, "fp_probe_tall_layout: NULL argument");
  CHECK_ARG(g->tall != nullptr, "fp_probe_tall_layout: tall-cell grids only");
  CHECK_IDLE(g, "fp_probe_tall_layout");
  std::lock_guard<std::mutex> lock(g->ctx->gpu);
  (void)hipSetDevice(g->ctx->device);
  return g->ctx->tall->probe_layout(g->ctx->stream, g->tall, phi, rows, unknown);
}

}  // extern "C"

// fp_api.hip -- the C ABI (include/fluid_projection.h) over the HIP solvers.
#include "fp_uniform.hpp"

#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <new>

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
  int grids = 0;
};

struct fp_grid {
  fp_ctx* ctx = nullptr;
  fp_grid_desc desc{};
  fp::UniformSolver* uniform = nullptr;
  uint64_t bytes = 0;
};

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
  fp_ctx* ctx = new (std::nothrow) fp_ctx();
  if (!ctx) return FP_ERR_OUT_OF_MEMORY;
  ctx->device = device;
  ctx->budget = hbm_budget_bytes;
  // default priority stream (no priority streams on this pool)
  e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    set_error("hipStreamCreate: %s", hipGetErrorString(e));
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
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return FP_OK;
}

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
    set_error("tall-cell layout not built into this library");
    return FP_ERR_UNSUPPORTED;
  }
  return FP_OK;
}

int fp_grid_bytes(const fp_grid_desc* desc, int32_t precision, uint64_t* bytes) {
  (void)precision;
  int rc = validate_desc(desc);
  if (rc) return rc;
  CHECK_ARG(bytes != nullptr, "bytes is NULL");
  const size_t b = fp::UniformSolver::estimate_bytes(desc->nx, desc->ny, desc->nz);
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
  g->uniform = new (std::nothrow) fp::UniformSolver();
  if (!g->uniform) { delete g; return FP_ERR_OUT_OF_MEMORY; }
  rc = g->uniform->create(desc->nx, desc->ny, desc->nz, desc->h);
  if (rc) {
    delete g->uniform;
    delete g;
    return rc;
  }
  g->bytes = g->uniform->bytes();
  ctx->in_use += g->bytes;
  ctx->grids += 1;
  *out = g;
  return FP_OK;
}

int fp_grid_destroy(fp_grid* g) {
  if (!g) return FP_OK;
  (void)hipSetDevice(g->ctx->device);
  (void)hipStreamSynchronize(g->ctx->stream);
  g->ctx->in_use -= g->bytes;
  g->ctx->grids -= 1;
  delete g->uniform;
  delete g;
  return FP_OK;
}

int fp_grid_levels(const fp_grid* g, int32_t max_levels, int32_t* dims, int32_t* strides, int32_t* levels) {
  CHECK_ARG(g && levels, "fp_grid_levels: NULL argument");
  const int L = g->uniform->levels();
  *levels = L;
  for (int l = 0; l < L && l < max_levels; ++l)
    for (int a = 0; a < 3; ++a) {
      if (dims) dims[l * 3 + a] = g->uniform->level(l).d[a];
      if (strides) strides[l * 3 + a] = g->uniform->level(l).s[a];
    }
  return FP_OK;
}

int fp_grid_face_counts(const fp_grid* g, int64_t* nu, int64_t* nv, int64_t* nw) {
  CHECK_ARG(g && nu && nv && nw, "fp_grid_face_counts: NULL argument");
  const fp_grid_desc& d = g->desc;
  *nu = (int64_t)(d.nx + 1) * d.ny * d.nz;
  *nv = (int64_t)d.nx * (d.ny + 1) * d.nz;
  *nw = (int64_t)d.nx * d.ny * (d.nz + 1);
  return FP_OK;
}

static int validate_inputs(const fp_inputs* in, const fp_outputs* out, const fp_options* o) {
  CHECK_ARG(in && out && o, "fp_project: NULL argument");
  CHECK_ARG(in->memory == FP_MEMORY_HOST || in->memory == FP_MEMORY_DEVICE, "unknown memory kind %d", in->memory);
  CHECK_ARG(in->phi && in->u && in->v && in->w, "fp_project: phi/u/v/w must be given");
  CHECK_ARG(in->dt > 0.0 && in->rho > 0.0, "fp_project: dt and rho must be positive");
  CHECK_ARG(o->relative_tolerance > 0.0 && o->relative_tolerance < 1.0, "relative_tolerance must be in (0,1)");
  CHECK_ARG(o->max_iterations >= 1 && o->max_outer >= 1, "iteration limits must be positive");
  CHECK_ARG(o->precision >= 0 && o->precision <= 2, "unknown precision %d", o->precision);
  CHECK_ARG(o->pre_sweeps >= 0 && o->post_sweeps >= 0 && o->pre_sweeps <= 16 && o->post_sweeps <= 16,
            "sweep counts must be in [0,16]");
  CHECK_ARG(o->pre_sweeps == o->post_sweeps, "pre and post sweeps must match (symmetric preconditioner)");
  return FP_OK;
}

int fp_project(fp_grid* g, const fp_inputs* in, fp_outputs* out, const fp_options* o, fp_report* rep) {
  CHECK_ARG(g != nullptr, "fp_project: grid is NULL");
  int rc = validate_inputs(in, out, o);
  if (rc) return rc;
  (void)hipSetDevice(g->ctx->device);
  return g->uniform->project(g->ctx->stream, in, out, o, rep);
}

int fp_probe_apply(fp_grid* g, const float* phi, const double* p, double* out) {
  CHECK_ARG(g && phi && p && out, "fp_probe_apply: NULL argument");
  (void)hipSetDevice(g->ctx->device);
  return g->uniform->probe_apply(g->ctx->stream, phi, p, out);
}

int fp_probe_rhs(fp_grid* g, const fp_inputs* in, double* b) {
  CHECK_ARG(g && in && b && in->phi && in->u && in->v && in->w, "fp_probe_rhs: NULL argument");
  CHECK_ARG(in->memory == FP_MEMORY_HOST, "fp_probe_rhs takes host arrays");
  (void)hipSetDevice(g->ctx->device);
  return g->uniform->probe_rhs(g->ctx->stream, in, b);
}

int fp_probe_vcycle(fp_grid* g, const float* phi, const fp_options* o, const float* r, float* z) {
  CHECK_ARG(g && phi && o && r && z, "fp_probe_vcycle: NULL argument");
  (void)hipSetDevice(g->ctx->device);
  return g->uniform->probe_vcycle(g->ctx->stream, phi, o, r, z);
}

int fp_probe_level_phi(fp_grid* g, const float* phi, int32_t level, int32_t preserve, float* out) {
  CHECK_ARG(g && phi && out, "fp_probe_level_phi: NULL argument");
  (void)hipSetDevice(g->ctx->device);
  return g->uniform->probe_level_phi(g->ctx->stream, phi, level, preserve, out);
}

}  // extern "C"

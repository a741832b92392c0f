// fp_addon.cc -- thin N-API binding of the C ABI (include/fluid_projection.h).
//
// What the reference's TypeScript host would load in place of its WebGPU
// projection stage (WebGPUUniformPressureMultigrid.encode + `project`,
// lib/methods/uniform/webgpu-uniform-pressure-multigrid.ts:327-364).  Every
// call maps one-to-one onto a C entry point; `project` runs as napi_async_work
// and resolves a Promise, so the JS thread never blocks on the GPU (the
// reference's readStats is asynchronous the same way,
// lib/methods/uniform/webgpu-uniform-reference.ts:1358-1421).
#include <node_api.h>

#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "../../include/fluid_projection.h"

namespace {

#define NAPI_OK(call)                                                     \
  do {                                                                    \
    napi_status _s = (call);                                              \
    if (_s != napi_ok) {                                                  \
      napi_throw_error(env, nullptr, "N-API call failed: " #call);        \
      return nullptr;                                                     \
    }                                                                     \
  } while (0)

napi_value throw_fp(napi_env env, int rc) {
  char code[32];
  snprintf(code, sizeof(code), "FP_ERR_%d", -rc);
  std::string msg = fp_last_error();
  // contract violations throw RangeError like the reference's constructor
  if (rc == FP_ERR_INVALID_ARGUMENT || rc == FP_ERR_LAYOUT) napi_throw_range_error(env, code, msg.c_str());
  else napi_throw_error(env, code, msg.c_str());
  return nullptr;
}

bool get_named(napi_env env, napi_value obj, const char* name, napi_value* out) {
  bool has = false;
  if (napi_has_named_property(env, obj, name, &has) != napi_ok || !has) return false;
  if (napi_get_named_property(env, obj, name, out) != napi_ok) return false;
  napi_valuetype t;
  napi_typeof(env, *out, &t);
  return t != napi_undefined && t != napi_null;
}

double get_number(napi_env env, napi_value obj, const char* name, double dflt) {
  napi_value v;
  if (!get_named(env, obj, name, &v)) return dflt;
  double d = dflt;
  napi_get_value_double(env, v, &d);
  return d;
}

// Float32Array -> pointer + element count (borrowed; the JS object is kept alive by a reference)
bool get_f32(napi_env env, napi_value v, float** data, size_t* len) {
  bool is = false;
  napi_is_typedarray(env, v, &is);
  if (!is) return false;
  napi_typedarray_type type;
  napi_value ab;
  size_t off;
  void* p;
  if (napi_get_typedarray_info(env, v, &type, len, &p, &ab, &off) != napi_ok) return false;
  if (type != napi_float32_array) return false;
  *data = (float*)p;
  return true;
}

void* unwrap_external(napi_env env, napi_value v) {
  void* p = nullptr;
  napi_get_value_external(env, v, &p);
  return p;
}

// Per grid (all on the JS thread): its context, and the projections queued
// on it that have not completed.  destroyGrid refuses a grid with pending
// work, whose buffers a worker thread may still be using.
struct GridInfo {
  fp_ctx* ctx = nullptr;
  int pending = 0;
};
std::map<fp_grid*, GridInfo> g_grids;

GridInfo* grid_info(fp_grid* g) {
  auto it = g_grids.find(g);
  return it == g_grids.end() ? nullptr : &it->second;
}

// ---- context ----------------------------------------------------------------
napi_value create_context(napi_env env, napi_callback_info info) {
  size_t argc = 2;
  napi_value argv[2];
  NAPI_OK(napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr));
  int32_t device = 0;
  double budget = 0;
  if (argc > 0) napi_get_value_int32(env, argv[0], &device);
  if (argc > 1) napi_get_value_double(env, argv[1], &budget);
  fp_ctx* ctx = nullptr;
  int rc = fp_ctx_create(device, (uint64_t)budget, &ctx);
  if (rc) return throw_fp(env, rc);
  napi_value ext;
  NAPI_OK(napi_create_external(env, ctx, nullptr, nullptr, &ext));
  return ext;
}

napi_value destroy_context(napi_env env, napi_callback_info info) {
  size_t argc = 1;
  napi_value argv[1];
  NAPI_OK(napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr));
  if (argc < 1) return nullptr;
  int rc = fp_ctx_destroy((fp_ctx*)unwrap_external(env, argv[0]));
  if (rc) return throw_fp(env, rc);
  return nullptr;
}

// ---- grid -----------------------------------------------------------------------
napi_value create_grid(napi_env env, napi_callback_info info) {
  size_t argc = 2;
  napi_value argv[2];
  NAPI_OK(napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr));
  if (argc < 2) {
    napi_throw_type_error(env, nullptr, "createGrid(context, {nx, ny, nz, h, layout?, band?, columnBases?})");
    return nullptr;
  }
  fp_ctx* ctx = (fp_ctx*)unwrap_external(env, argv[0]);
  fp_grid_desc d;
  std::memset(&d, 0, sizeof(d));
  d.layout = (int32_t)get_number(env, argv[1], "layout", FP_LAYOUT_UNIFORM);
  d.nx = (int32_t)get_number(env, argv[1], "nx", 0);
  d.ny = (int32_t)get_number(env, argv[1], "ny", 0);
  d.nz = (int32_t)get_number(env, argv[1], "nz", 0);
  d.band = (int32_t)get_number(env, argv[1], "band", 0);
  d.max_delta = (int32_t)get_number(env, argv[1], "maxDelta", 0);
  d.h = get_number(env, argv[1], "h", 0.0);
  std::vector<int32_t> bases;
  napi_value cb;
  if (get_named(env, argv[1], "columnBases", &cb)) {
    // an Int32Array of nx * nz tall-cell heights (the reference stores them as
    // float, so a Float32Array of integral values is accepted and converted)
    bool is_ta = false;
    napi_is_typedarray(env, cb, &is_ta);
    napi_typedarray_type type = napi_int8_array;
    size_t len = 0, off = 0;
    void* p = nullptr;
    napi_value ab;
    if (is_ta) napi_get_typedarray_info(env, cb, &type, &len, &p, &ab, &off);
    if (!is_ta || (type != napi_int32_array && type != napi_float32_array)) {
      napi_throw_type_error(env, nullptr, "columnBases must be an Int32Array (or Float32Array)");
      return nullptr;
    }
    if (d.nx <= 0 || d.nz <= 0 || len != (size_t)d.nx * (size_t)d.nz) {
      napi_throw_range_error(env, nullptr, "columnBases length must be nx * nz");
      return nullptr;
    }
    bases.resize(len);
    for (size_t i = 0; i < len; ++i) {
      const double b = type == napi_int32_array ? (double)((int32_t*)p)[i] : (double)((float*)p)[i];
      if (b != (double)(int32_t)b) {
        napi_throw_range_error(env, nullptr, "columnBases must hold integers");
        return nullptr;
      }
      bases[i] = (int32_t)b;
    }
    d.column_bases = bases.data();
  }
  fp_grid* g = nullptr;
  int rc = fp_grid_create(ctx, &d, &g);
  if (rc) return throw_fp(env, rc);
  g_grids[g].ctx = ctx;
  napi_value ext;
  NAPI_OK(napi_create_external(env, g, nullptr, nullptr, &ext));
  return ext;
}

napi_value destroy_grid(napi_env env, napi_callback_info info) {
  size_t argc = 1;
  napi_value argv[1];
  NAPI_OK(napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr));
  if (argc < 1) return nullptr;
  fp_grid* g = (fp_grid*)unwrap_external(env, argv[0]);
  GridInfo* gi = grid_info(g);
  if (gi && gi->pending > 0) {
    napi_throw_error(env, "FP_ERR_6", "destroyGrid: the grid has projections in flight (await them first)");
    return nullptr;
  }
  int rc = fp_grid_destroy(g);
  if (rc) return throw_fp(env, rc);
  g_grids.erase(g);
  return nullptr;
}

napi_value bytes_in_use(napi_env env, napi_callback_info info) {
  size_t argc = 1;
  napi_value argv[1];
  NAPI_OK(napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr));
  uint64_t b = 0;
  int rc = fp_ctx_bytes_in_use(argc > 0 ? (const fp_ctx*)unwrap_external(env, argv[0]) : nullptr, &b);
  if (rc) return throw_fp(env, rc);
  napi_value v;
  NAPI_OK(napi_create_double(env, (double)b, &v));
  return v;
}

napi_value grid_levels(napi_env env, napi_callback_info info) {
  size_t argc = 1;
  napi_value argv[1];
  NAPI_OK(napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr));
  fp_grid* g = (fp_grid*)unwrap_external(env, argv[0]);
  int32_t dims[48], strides[48], L = 0;
  int rc = fp_grid_levels(g, 16, dims, strides, &L);
  if (rc) return throw_fp(env, rc);
  napi_value arr;
  NAPI_OK(napi_create_array_with_length(env, L, &arr));
  for (int l = 0; l < L; ++l) {
    napi_value row;
    NAPI_OK(napi_create_array_with_length(env, 3, &row));
    for (int a = 0; a < 3; ++a) {
      napi_value v;
      NAPI_OK(napi_create_int32(env, dims[l * 3 + a], &v));
      NAPI_OK(napi_set_element(env, row, a, v));
    }
    NAPI_OK(napi_set_element(env, arr, l, row));
  }
  return arr;
}

// ---- project / submit / batch (async) ----------------------------------------------
// One grid's part of a job: its C records and the JS result object.
struct GridIO {
  fp_grid* grid = nullptr;
  fp_inputs in{};
  fp_outputs out{};
  fp_report rep{};
  napi_ref result = nullptr;   // {pressure, u, v, w}
};

struct ProjectJob {
  enum Kind { PROJECT, SUBMIT, BATCH } kind = PROJECT;
  napi_async_work work = nullptr;
  napi_deferred deferred = nullptr;
  std::vector<napi_ref> keep;  // input/output typed arrays stay alive
  std::vector<GridIO> io;
  fp_options opt{};
  fp_ctx* ctx = nullptr;       // SUBMIT: the context polled
  uint64_t ticket = 0;
  double latency_ms = 0.0;
  int rc = 0;
  std::string error;
};

void project_execute(napi_env, void* data) {
  ProjectJob* j = (ProjectJob*)data;
  if (j->kind == ProjectJob::PROJECT) {
    j->rc = fp_project(j->io[0].grid, &j->io[0].in, &j->io[0].out, &j->opt, &j->io[0].rep);
  } else if (j->kind == ProjectJob::SUBMIT) {
    j->rc = fp_poll(j->ctx, j->ticket, -1, &j->io[0].rep, &j->latency_ms);
  } else {
    std::vector<fp_grid*> grids;
    std::vector<fp_inputs> in;
    std::vector<fp_outputs> out;
    std::vector<fp_report> reps(j->io.size());
    for (const GridIO& g : j->io) {
      grids.push_back(g.grid);
      in.push_back(g.in);
      out.push_back(g.out);
    }
    j->rc = fp_project_batch(grids.data(), (int32_t)grids.size(), in.data(), out.data(), &j->opt, reps.data());
    for (size_t i = 0; i < reps.size(); ++i) j->io[i].rep = reps[i];
  }
  if (j->rc) j->error = fp_last_error();
}

napi_value report_object(napi_env env, const fp_report& r) {
  napi_value o, v;
  napi_create_object(env, &o);
  auto num = [&](const char* k, double x) {
    napi_create_double(env, x, &v);
    napi_set_named_property(env, o, k, v);
  };
  num("flags", r.flags);
  num("iterations", r.iterations);
  num("outerIterations", r.outer_iterations);
  num("levels", r.levels);
  num("relativeResidual", r.relative_residual);
  num("rhsNorm", r.rhs_norm);
  num("maxDivergenceBefore", r.max_div_before);
  num("maxDivergenceAfter", r.max_div_after);
  num("deviceMs", r.device_ms);
  num("liquidSamples", (double)r.liquid_samples);
  num("samples", (double)r.samples);
  num("lcpResidual", r.lcp_residual);
  num("coarseSweeps", r.coarse_sweeps);
  num("coarseFailures", r.coarse_failures);
  return o;
}

void project_complete(napi_env env, napi_status, void* data) {
  ProjectJob* j = (ProjectJob*)data;
  for (const GridIO& g : j->io) {
    GridInfo* gi = grid_info(g.grid);
    if (gi) gi->pending -= 1;
  }
  if (j->rc) {
    napi_value err, msg, code;
    napi_create_string_utf8(env, j->error.c_str(), NAPI_AUTO_LENGTH, &msg);
    char c[32];
    snprintf(c, sizeof(c), "FP_ERR_%d", -j->rc);
    napi_create_string_utf8(env, c, NAPI_AUTO_LENGTH, &code);
    napi_create_error(env, code, msg, &err);
    napi_reject_deferred(env, j->deferred, err);
  } else {
    std::vector<napi_value> res(j->io.size());
    for (size_t i = 0; i < j->io.size(); ++i) {
      napi_get_reference_value(env, j->io[i].result, &res[i]);
      napi_value rep = report_object(env, j->io[i].rep);
      if (j->kind == ProjectJob::SUBMIT) {
        napi_value lat;
        napi_create_double(env, j->latency_ms, &lat);
        napi_set_named_property(env, rep, "latencyMs", lat);
      }
      napi_set_named_property(env, res[i], "report", rep);
    }
    if (j->kind == ProjectJob::BATCH) {
      napi_value arr;
      napi_create_array_with_length(env, res.size(), &arr);
      for (size_t i = 0; i < res.size(); ++i) napi_set_element(env, arr, (uint32_t)i, res[i]);
      napi_resolve_deferred(env, j->deferred, arr);
    } else {
      napi_resolve_deferred(env, j->deferred, res[0]);
    }
  }
  for (napi_ref r : j->keep) napi_delete_reference(env, r);
  for (const GridIO& g : j->io)
    if (g.result) napi_delete_reference(env, g.result);
  napi_delete_async_work(env, j->work);
  delete j;
}

// Parses one grid's {phi | density, u, v, w, targetDivergence?, dt, rho} and
// {pressure, u, v, w} into g; false after throwing a JS error.
bool parse_io(napi_env env, ProjectJob* j, GridIO& g, napi_value inputs, napi_value outputs) {
  // array lengths the grid reads and writes: level-0 samples and the three face arrays
  int32_t dims[3 * 16], levels = 0;
  int64_t nf[3] = {0, 0, 0};
  if (!g.grid || fp_grid_levels(g.grid, 16, dims, nullptr, &levels) != FP_OK ||
      fp_grid_face_counts(g.grid, &nf[0], &nf[1], &nf[2]) != FP_OK) {
    napi_throw_type_error(env, nullptr, "not a live projection grid");
    return false;
  }
  const size_t cells = (size_t)dims[0] * dims[1] * dims[2];
  const size_t want_in[6] = {cells, (size_t)nf[0], (size_t)nf[1], (size_t)nf[2], cells, cells};
  const size_t want_out[4] = {cells, (size_t)nf[0], (size_t)nf[1], (size_t)nf[2]};
  g.in.memory = FP_MEMORY_HOST;
  const char* in_names[6] = {"phi", "u", "v", "w", "targetDivergence", "density"};
  const float** in_ptrs[6] = {&g.in.phi, &g.in.u, &g.in.v, &g.in.w, &g.in.target_divergence, &g.in.density};
  for (int k = 0; k < 6; ++k) {
    napi_value v;
    if (!get_named(env, inputs, in_names[k], &v)) continue;
    float* p;
    size_t len;
    if (!get_f32(env, v, &p, &len)) {
      napi_throw_type_error(env, nullptr, "inputs must be Float32Array");
      return false;
    }
    if (len != want_in[k]) {
      napi_throw_range_error(env, nullptr, "input array length does not match the grid");
      return false;
    }
    *in_ptrs[k] = p;
    napi_ref r;
    napi_create_reference(env, v, 1, &r);
    j->keep.push_back(r);
  }
  g.in.dt = get_number(env, inputs, "dt", 0.0);
  g.in.rho = get_number(env, inputs, "rho", 998.2);
  const char* out_names[4] = {"pressure", "u", "v", "w"};
  float** out_ptrs[4] = {&g.out.pressure, &g.out.u, &g.out.v, &g.out.w};
  for (int k = 0; k < 4; ++k) {
    napi_value v;
    if (!get_named(env, outputs, out_names[k], &v)) continue;
    float* p;
    size_t len;
    if (!get_f32(env, v, &p, &len)) {
      napi_throw_type_error(env, nullptr, "outputs must be Float32Array");
      return false;
    }
    if (len != want_out[k]) {
      napi_throw_range_error(env, nullptr, "output array length does not match the grid");
      return false;
    }
    *out_ptrs[k] = p;
  }
  napi_create_reference(env, outputs, 1, &g.result);
  return true;
}

void parse_options(napi_env env, napi_value o, fp_options* opt) {
  fp_default_options(opt);
  napi_valuetype t;
  napi_typeof(env, o, &t);
  if (t != napi_object) return;
  opt->relative_tolerance = get_number(env, o, "relativeTolerance", opt->relative_tolerance);
  opt->max_iterations = (int32_t)get_number(env, o, "maxIterations", opt->max_iterations);
  opt->max_outer = (int32_t)get_number(env, o, "maxOuter", opt->max_outer);
  opt->precision = (int32_t)get_number(env, o, "precision", opt->precision);
  opt->pre_sweeps = (int32_t)get_number(env, o, "preSweeps", opt->pre_sweeps);
  opt->post_sweeps = (int32_t)get_number(env, o, "postSweeps", opt->post_sweeps);
  opt->warm_start = (int32_t)get_number(env, o, "warmStart", opt->warm_start);
  opt->solver = (int32_t)get_number(env, o, "solver", opt->solver);
  opt->open_top = (int32_t)get_number(env, o, "openTop", opt->open_top);
  opt->full_cycles = (int32_t)get_number(env, o, "fullCycles", opt->full_cycles);
  opt->v_cycles = (int32_t)get_number(env, o, "vCycles", opt->v_cycles);
  opt->lcp_sweeps = (int32_t)get_number(env, o, "lcpSweeps", opt->lcp_sweeps);
}

void discard(napi_env env, ProjectJob* j) {
  for (napi_ref r : j->keep) napi_delete_reference(env, r);
  for (const GridIO& g : j->io)
    if (g.result) napi_delete_reference(env, g.result);
  delete j;
}

napi_value queue(napi_env env, ProjectJob* j) {
  for (const GridIO& g : j->io) {
    GridInfo* gi = grid_info(g.grid);
    if (gi) gi->pending += 1;
  }
  napi_value promise, name;
  NAPI_OK(napi_create_promise(env, &j->deferred, &promise));
  NAPI_OK(napi_create_string_utf8(env, "fluidProjection", NAPI_AUTO_LENGTH, &name));
  NAPI_OK(napi_create_async_work(env, nullptr, name, project_execute, project_complete, j, &j->work));
  NAPI_OK(napi_queue_async_work(env, j->work));
  return promise;
}

// project(grid, {phi, u, v, w, targetDivergence?, dt, rho}, {pressure, u, v, w}, options?) -> Promise
// submit(...same...) -> Promise (fp_submit now, fp_poll on a worker thread)
napi_value project_or_submit(napi_env env, napi_callback_info info, bool submit) {
  size_t argc = 4;
  napi_value argv[4];
  NAPI_OK(napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr));
  if (argc < 3) {
    napi_throw_type_error(env, nullptr, "project(grid, inputs, outputs, options?)");
    return nullptr;
  }
  ProjectJob* j = new ProjectJob();
  j->kind = submit ? ProjectJob::SUBMIT : ProjectJob::PROJECT;
  j->io.resize(1);
  j->io[0].grid = (fp_grid*)unwrap_external(env, argv[0]);
  if (!parse_io(env, j, j->io[0], argv[1], argv[2])) {
    discard(env, j);
    return nullptr;
  }
  if (argc > 3) parse_options(env, argv[3], &j->opt);
  else fp_default_options(&j->opt);
  if (submit) {
    GridInfo* gi = grid_info(j->io[0].grid);
    j->ctx = gi ? gi->ctx : nullptr;
    int rc = fp_submit(j->io[0].grid, &j->io[0].in, &j->io[0].out, &j->opt, &j->ticket);
    if (rc) {
      discard(env, j);
      return throw_fp(env, rc);
    }
  }
  return queue(env, j);
}

napi_value project(napi_env env, napi_callback_info info) { return project_or_submit(env, info, false); }
napi_value submit(napi_env env, napi_callback_info info) { return project_or_submit(env, info, true); }

// projectBatch([grid...], [inputs...], [outputs...], options?) -> Promise<[result...]>
napi_value project_batch(napi_env env, napi_callback_info info) {
  size_t argc = 4;
  napi_value argv[4];
  NAPI_OK(napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr));
  uint32_t n = 0, ni = 0, no = 0;
  if (argc < 3 || napi_get_array_length(env, argv[0], &n) != napi_ok ||
      napi_get_array_length(env, argv[1], &ni) != napi_ok || napi_get_array_length(env, argv[2], &no) != napi_ok ||
      n == 0 || ni != n || no != n) {
    napi_throw_type_error(env, nullptr, "projectBatch([grid...], [inputs...], [outputs...], options?) of equal lengths");
    return nullptr;
  }
  ProjectJob* j = new ProjectJob();
  j->kind = ProjectJob::BATCH;
  j->io.resize(n);
  for (uint32_t i = 0; i < n; ++i) {
    napi_value g, in, out;
    napi_get_element(env, argv[0], i, &g);
    napi_get_element(env, argv[1], i, &in);
    napi_get_element(env, argv[2], i, &out);
    j->io[i].grid = (fp_grid*)unwrap_external(env, g);
    if (!parse_io(env, j, j->io[i], in, out)) {
      discard(env, j);
      return nullptr;
    }
  }
  fp_default_options(&j->opt);
  if (argc > 3) parse_options(env, argv[3], &j->opt);
  return queue(env, j);
}

// remeshTallGrid(grid, {phi, u, v, w}, {phi, u, v, w}?, {liquidHalo, airHalo}?)
//   -> {changedColumns, columnBases: Int32Array}
// fp_tall_remesh, synchronous (about a millisecond on a 512^2-column grid);
// without an output object the input arrays are updated in place.
napi_value remesh_tall_grid(napi_env env, napi_callback_info info) {
  size_t argc = 4;
  napi_value argv[4];
  NAPI_OK(napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr));
  if (argc < 2) {
    napi_throw_type_error(env, nullptr, "remeshTallGrid(grid, {phi, u, v, w}, outputs?, {liquidHalo, airHalo}?)");
    return nullptr;
  }
  fp_grid* g = (fp_grid*)unwrap_external(env, argv[0]);
  GridInfo* gi = g ? grid_info(g) : nullptr;
  int32_t dims[3 * 16], levels = 0;
  int64_t nf[3] = {0, 0, 0};
  if (!gi || fp_grid_levels(g, 16, dims, nullptr, &levels) != FP_OK ||
      fp_grid_face_counts(g, &nf[0], &nf[1], &nf[2]) != FP_OK) {
    napi_throw_type_error(env, nullptr, "not a live projection grid");
    return nullptr;
  }
  if (gi->pending > 0) {
    napi_throw_error(env, "FP_ERR_6", "remeshTallGrid: the grid has a projection in flight (await it first)");
    return nullptr;
  }
  const size_t samples = (size_t)dims[0] * dims[1] * dims[2];
  const size_t want[4] = {samples, (size_t)nf[0], (size_t)nf[1], (size_t)nf[2]};
  const char* names[4] = {"phi", "u", "v", "w"};
  fp_tall_remesh_args a;
  std::memset(&a, 0, sizeof(a));
  a.memory = FP_MEMORY_HOST;
  const float** in_ptrs[4] = {&a.phi, &a.u, &a.v, &a.w};
  float** out_ptrs[4] = {&a.phi_out, &a.u_out, &a.v_out, &a.w_out};
  bool has_out = false;
  if (argc > 2) {
    napi_valuetype t;
    napi_typeof(env, argv[2], &t);
    has_out = t == napi_object;
  }
  for (int k = 0; k < 4; ++k) {
    napi_value v;
    float* p;
    size_t len;
    if (!get_named(env, argv[1], names[k], &v) || !get_f32(env, v, &p, &len) || len != want[k]) {
      napi_throw_range_error(env, nullptr, "remeshTallGrid: phi, u, v, w must be Float32Arrays of the grid's sizes");
      return nullptr;
    }
    *in_ptrs[k] = p;
    *out_ptrs[k] = p;
    if (has_out && get_named(env, argv[2], names[k], &v)) {
      if (!get_f32(env, v, &p, &len) || len != want[k]) {
        napi_throw_range_error(env, nullptr, "remeshTallGrid: outputs must be Float32Arrays of the grid's sizes");
        return nullptr;
      }
      *out_ptrs[k] = p;
    }
  }
  if (argc > 3) {
    a.liquid_halo = (int32_t)get_number(env, argv[3], "liquidHalo", 0);
    a.air_halo = (int32_t)get_number(env, argv[3], "airHalo", 0);
  }
  const size_t columns = (size_t)dims[0] * dims[2];
  napi_value ab, bases, res, changed;
  void* bp = nullptr;
  NAPI_OK(napi_create_arraybuffer(env, columns * 4, &bp, &ab));
  NAPI_OK(napi_create_typedarray(env, napi_int32_array, columns, ab, 0, &bases));
  a.column_bases_out = (int32_t*)bp;
  int64_t moved = 0;
  int rc = fp_tall_remesh(g, &a, &moved);
  if (rc) return throw_fp(env, rc);
  NAPI_OK(napi_create_object(env, &res));
  NAPI_OK(napi_create_double(env, (double)moved, &changed));
  NAPI_OK(napi_set_named_property(env, res, "changedColumns", changed));
  NAPI_OK(napi_set_named_property(env, res, "columnBases", bases));
  return res;
}

// extrapolateVelocity(grid, {phi | density, u, v, w}, {u, v, w}?, {openTop, maxPasses}?)
//   -> {passes, activeLeft, levels, bandFaces, unknownFaces, deviceMs}
// fp_extrapolate_velocity, synchronous; without an output object u, v, w are
// updated in place.
napi_value extrapolate_velocity(napi_env env, napi_callback_info info) {
  size_t argc = 4;
  napi_value argv[4];
  NAPI_OK(napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr));
  if (argc < 2) {
    napi_throw_type_error(env, nullptr, "extrapolateVelocity(grid, {phi|density, u, v, w}, outputs?, options?)");
    return nullptr;
  }
  fp_grid* g = (fp_grid*)unwrap_external(env, argv[0]);
  GridInfo* gi = g ? grid_info(g) : nullptr;
  int32_t dims[3 * 16], levels = 0;
  int64_t nf[3] = {0, 0, 0};
  if (!gi || fp_grid_levels(g, 16, dims, nullptr, &levels) != FP_OK ||
      fp_grid_face_counts(g, &nf[0], &nf[1], &nf[2]) != FP_OK) {
    napi_throw_type_error(env, nullptr, "not a live projection grid");
    return nullptr;
  }
  if (gi->pending > 0) {
    napi_throw_error(env, "FP_ERR_6", "extrapolateVelocity: the grid has a projection in flight (await it first)");
    return nullptr;
  }
  const size_t cells = (size_t)dims[0] * dims[1] * dims[2];
  fp_extrapolate_args a;
  std::memset(&a, 0, sizeof(a));
  a.memory = FP_MEMORY_HOST;
  napi_value val;
  float* p;
  size_t len;
  const bool has_phi = get_named(env, argv[1], "phi", &val);
  if (!has_phi && !get_named(env, argv[1], "density", &val)) {
    napi_throw_type_error(env, nullptr, "extrapolateVelocity: phi or density is required");
    return nullptr;
  }
  if (!get_f32(env, val, &p, &len) || len != cells) {
    napi_throw_range_error(env, nullptr, "extrapolateVelocity: phi/density must be a Float32Array of the grid's cells");
    return nullptr;
  }
  if (has_phi) a.phi = p;
  else a.density = p;
  const char* names[3] = {"u", "v", "w"};
  const float** in_ptrs[3] = {&a.u, &a.v, &a.w};
  float** out_ptrs[3] = {&a.u_out, &a.v_out, &a.w_out};
  bool has_out = false;
  if (argc > 2) {
    napi_valuetype t;
    napi_typeof(env, argv[2], &t);
    has_out = t == napi_object;
  }
  for (int k = 0; k < 3; ++k) {
    if (!get_named(env, argv[1], names[k], &val) || !get_f32(env, val, &p, &len) || len != (size_t)nf[k]) {
      napi_throw_range_error(env, nullptr, "extrapolateVelocity: u, v, w must be Float32Arrays of the grid's faces");
      return nullptr;
    }
    *in_ptrs[k] = p;
    *out_ptrs[k] = p;
    if (has_out && get_named(env, argv[2], names[k], &val)) {
      if (!get_f32(env, val, &p, &len) || len != (size_t)nf[k]) {
        napi_throw_range_error(env, nullptr, "extrapolateVelocity: outputs must be Float32Arrays of the grid's faces");
        return nullptr;
      }
      *out_ptrs[k] = p;
    }
  }
  if (argc > 3) {
    a.open_top = (int32_t)get_number(env, argv[3], "openTop", 0);
    a.max_passes = (int32_t)get_number(env, argv[3], "maxPasses", 0);
  }
  fp_extrapolate_report r;
  int rc = fp_extrapolate_velocity(g, &a, &r);
  if (rc) return throw_fp(env, rc);
  napi_value res, num;
  NAPI_OK(napi_create_object(env, &res));
  const struct {
    const char* name;
    double value;
  } fields[] = {{"passes", (double)r.passes},        {"activeLeft", (double)r.active_left},
                {"levels", (double)r.levels},        {"bandFaces", (double)r.band_faces},
                {"unknownFaces", (double)r.unknown_faces}, {"deviceMs", r.device_ms}};
  for (const auto& f : fields) {
    NAPI_OK(napi_create_double(env, f.value, &num));
    NAPI_OK(napi_set_named_property(env, res, f.name, num));
  }
  return res;
}

// advectVelocity(grid, {phi | density, u, v, w, transport?: {u, v, w}}, {u, v, w}?,
//                {dt, scheme, gravity, viscosity, openTop, noSlip})
//   -> {deviceMs}
// fp_advect_velocity, synchronous; without an output object u, v, w are
// updated in place.  scheme: 0 semi-Lagrangian, 1 bounded MacCormack.
napi_value advect_velocity(napi_env env, napi_callback_info info) {
  size_t argc = 4;
  napi_value argv[4];
  NAPI_OK(napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr));
  if (argc < 4) {
    napi_throw_type_error(env, nullptr, "advectVelocity(grid, {phi|density, u, v, w, transport?}, outputs|null, options)");
    return nullptr;
  }
  fp_grid* g = (fp_grid*)unwrap_external(env, argv[0]);
  GridInfo* gi = g ? grid_info(g) : nullptr;
  int32_t dims[3 * 16], levels = 0;
  int64_t nf[3] = {0, 0, 0};
  if (!gi || fp_grid_levels(g, 16, dims, nullptr, &levels) != FP_OK ||
      fp_grid_face_counts(g, &nf[0], &nf[1], &nf[2]) != FP_OK) {
    napi_throw_type_error(env, nullptr, "not a live projection grid");
    return nullptr;
  }
  if (gi->pending > 0) {
    napi_throw_error(env, "FP_ERR_6", "advectVelocity: the grid has a projection in flight (await it first)");
    return nullptr;
  }
  const size_t cells = (size_t)dims[0] * dims[1] * dims[2];
  fp_advect_args a;
  std::memset(&a, 0, sizeof(a));
  a.memory = FP_MEMORY_HOST;
  napi_value val, tr;
  float* p;
  size_t len;
  const bool has_phi = get_named(env, argv[1], "phi", &val);
  if (!has_phi && !get_named(env, argv[1], "density", &val)) {
    napi_throw_type_error(env, nullptr, "advectVelocity: phi or density is required");
    return nullptr;
  }
  if (!get_f32(env, val, &p, &len) || len != cells) {
    napi_throw_range_error(env, nullptr, "advectVelocity: phi/density must be a Float32Array of the grid's cells");
    return nullptr;
  }
  if (has_phi) a.phi = p;
  else a.density = p;
  const char* names[3] = {"u", "v", "w"};
  const float** in_ptrs[3] = {&a.u, &a.v, &a.w};
  const float** tr_ptrs[3] = {&a.tu, &a.tv, &a.tw};
  float** out_ptrs[3] = {&a.u_out, &a.v_out, &a.w_out};
  napi_valuetype t;
  napi_typeof(env, argv[2], &t);
  const bool has_out = t == napi_object;
  const bool has_tr = get_named(env, argv[1], "transport", &tr);
  for (int k = 0; k < 3; ++k) {
    if (!get_named(env, argv[1], names[k], &val) || !get_f32(env, val, &p, &len) || len != (size_t)nf[k]) {
      napi_throw_range_error(env, nullptr, "advectVelocity: u, v, w must be Float32Arrays of the grid's faces");
      return nullptr;
    }
    *in_ptrs[k] = p;
    *out_ptrs[k] = p;
    if (has_tr) {
      if (!get_named(env, tr, names[k], &val) || !get_f32(env, val, &p, &len) || len != (size_t)nf[k]) {
        napi_throw_range_error(env, nullptr, "advectVelocity: transport u, v, w must be Float32Arrays of the grid's faces");
        return nullptr;
      }
      *tr_ptrs[k] = p;
    }
    if (has_out && get_named(env, argv[2], names[k], &val)) {
      if (!get_f32(env, val, &p, &len) || len != (size_t)nf[k]) {
        napi_throw_range_error(env, nullptr, "advectVelocity: outputs must be Float32Arrays of the grid's faces");
        return nullptr;
      }
      *out_ptrs[k] = p;
    }
  }
  a.dt = (float)get_number(env, argv[3], "dt", 0.0);
  a.scheme = (int32_t)get_number(env, argv[3], "scheme", FP_ADVECT_SEMI_LAGRANGIAN);
  a.gravity = (float)get_number(env, argv[3], "gravity", -9.80665);
  a.viscosity = (float)get_number(env, argv[3], "viscosity", 0.0);
  a.open_top = (int32_t)get_number(env, argv[3], "openTop", 0);
  a.no_slip = (int32_t)get_number(env, argv[3], "noSlip", 0);
  fp_advect_report r;
  int rc = fp_advect_velocity(g, &a, &r);
  if (rc) return throw_fp(env, rc);
  napi_value res, num;
  NAPI_OK(napi_create_object(env, &res));
  NAPI_OK(napi_create_double(env, r.device_ms, &num));
  NAPI_OK(napi_set_named_property(env, res, "deviceMs", num));
  return res;
}

napi_value last_error(napi_env env, napi_callback_info) {
  napi_value s;
  NAPI_OK(napi_create_string_utf8(env, fp_last_error(), NAPI_AUTO_LENGTH, &s));
  return s;
}

napi_value abi_version(napi_env env, napi_callback_info) {
  napi_value v;
  NAPI_OK(napi_create_int32(env, fp_abi_version(), &v));
  return v;
}

napi_value init(napi_env env, napi_value exports) {
  struct {
    const char* name;
    napi_callback cb;
  } fns[] = {{"createContext", create_context}, {"destroyContext", destroy_context}, {"createGrid", create_grid},
             {"destroyGrid", destroy_grid},     {"gridLevels", grid_levels},         {"project", project},
             {"submit", submit},                {"projectBatch", project_batch},     {"bytesInUse", bytes_in_use},
             {"lastError", last_error},         {"abiVersion", abi_version},         {"remeshTallGrid", remesh_tall_grid},
             {"extrapolateVelocity", extrapolate_velocity}, {"advectVelocity", advect_velocity}};
  for (auto& f : fns) {
    napi_value fn;
    NAPI_OK(napi_create_function(env, f.name, NAPI_AUTO_LENGTH, f.cb, nullptr, &fn));
    NAPI_OK(napi_set_named_property(env, exports, f.name, fn));
  }
  return exports;
}

}  // namespace

NAPI_MODULE(NODE_GYP_MODULE_NAME, init)

// fp_state.hpp -- per-grid solver state kernels shared by the uniform and the
// tall-cell projections.  P is a part type (PartOf<Record>); a record has
// st (State*), partials, max_iter, tol2 and out_u/out_v/out_w.  Finishers run
// one workgroup per part of their producer's table and reduce its partials in
// a fixed order.
#pragma once

#include "fp_uniform.hpp"

namespace fp {

// Refinement targets: the FP64 refinement stops at kOuterFactor times the
// requested relative residual; each inner FP32 solve at kInnerFactor of that
// target, or after reducing its residual by kInnerCap (what FP32 CG can be
// trusted to reduce in one pass).
#ifndef FP_OUTER_FACTOR
#define FP_OUTER_FACTOR 0.25
#endif
#ifndef FP_INNER_FACTOR
#define FP_INNER_FACTOR 0.5
#endif
#ifndef FP_INNER_CAP
#define FP_INNER_CAP 1e-6
#endif
constexpr double kOuterFactor = FP_OUTER_FACTOR;
constexpr double kInnerFactor = FP_INNER_FACTOR;
constexpr double kInnerCap = FP_INNER_CAP;

// --------------------------------------------------------------------------
// per-projection state and vector resets
// --------------------------------------------------------------------------
// One workgroup per grid: a fresh State (rscale 1).
template <class P>
__global__ void k_reset_state(const P* parts) {
  if (threadIdx.x != 0) return;
  State* st = parts[blockIdx.x].g->st;
  State z;
  __builtin_memset(&z, 0, sizeof(z));
  z.rscale = 1.0f;
  *st = z;
}

// Pending work of a batch: counts[0] grids whose refinement has not
// converged, counts[1] grids whose inner solve is still running.
template <class P>
__global__ void k_status(const P* parts, int np, int* counts) {
  __shared__ int s_outer, s_inner;
  if (threadIdx.x == 0) {
    s_outer = 0;
    s_inner = 0;
  }
  __syncthreads();
  int o = 0, in = 0;
  for (int i = threadIdx.x; i < np; i += blockDim.x) {
    const State* st = parts[i].g->st;
    o += st->outer_done ? 0 : 1;
    in += st->done ? 0 : 1;
  }
  atomicAdd(&s_outer, o);
  atomicAdd(&s_inner, in);
  __syncthreads();
  if (threadIdx.x == 0) {
    counts[0] = s_outer;
    counts[1] = s_inner;
  }
}

// The States of a batch gathered in part order (one copy to the host).
template <class P>
__global__ void k_gather_states(const P* parts, int np, State* out) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < np; i += gridDim.x * blockDim.x) out[i] = *parts[i].g->st;
}

// finishers run one workgroup per part of their producer's table
template <class P>
__global__ void k_fin_rhs(const P* parts) {
  const P p = parts[blockIdx.x];
  const auto& G = *p.g;
  State* st = G.st;
  double out[3];
  float mx;
  sum_partials_max<3>(G.partials, p.nblocks, out, mx);
  if (threadIdx.x == 0) {
    st->divmax_before = __float_as_uint(mx);
    const long long liquid = (long long)out[1];
    st->liquid = liquid;
    double bb = out[0];
    // Every sample liquid (a tall-cell grid: and every band up to the lid):
    // no free-surface (Dirichlet) face anchors the pressure, A is singular
    // with the constants as its null space
    // (sparse-atlas-composite-projection.ts:1086-1146,1194-1211).  Walls are
    // the only other boundary, and a liquid component that is not the whole
    // domain has a neighbour outside it, which is air, so this is the only
    // unanchored component a grid can have.  The mean of b is removed (the
    // system becomes consistent) and the pressure is returned with zero mean.
    if (liquid > 0 && liquid == G.ns_liquid) {
      const double mean = out[2] / (double)liquid;
      st->nullspace = 1;
      st->bmean = mean;
      st->flags |= 0x10;  // FP_FLAG_UNANCHORED_COMPONENT
      bb = fmax(0.0, bb - (double)liquid * mean * mean);
    }
    st->bb = bb;
    // The refinement stops at a quarter of the requested relative residual:
    // the max-norm divergence left in the FP32 output then stays below that
    // of an FP64 solve stopped at the requested tolerance.
    st->tol2_bb = kOuterFactor * kOuterFactor * G.tol2 * bb;
  }
}

// Refinement control after the FP64 true residual (OpTrueResidual): decides
// convergence of the refinement and the target of the next inner solve.
// new_pass: an inner solve follows for grids that have not converged.
template <class P>
__global__ void k_fin_true(const P* parts, int new_pass) {
  const P p = parts[blockIdx.x];
  State* st = p.g->st;
  if (st->outer_done) return;
  double out[2];
  float m;
  sum_partials_max<2>(p.g->partials, p.nblocks, out, m);
  if (threadIdx.x == 0) {
    const double rr = out[0];
    st->rr64 = rr;
    st->xsum = out[1];
    const bool finite = isfinite(rr);
    st->outer_done = (!finite) || rr <= st->tol2_bb;
    if (!finite) st->flags |= 0x1;  // FP_FLAG_NONFINITE
    // inner solve: a tenth of the refinement target (margin for the max-norm
    // divergence), at most what FP32 CG can be trusted to reduce in one pass.
    st->inner_target2 = fmax(kInnerFactor * kInnerFactor * st->tol2_bb, kInnerCap * kInnerCap * rr);
    st->iter = 0;
    st->done = st->outer_done || !new_pass;
    if (!st->outer_done && new_pass) st->outer += 1;
    st->rscale = p.g->rscale_bias * ((m > 0.0f && isfinite(m)) ? exp2f(-ceilf(log2f(m))) : 1.0f);
  }
}

template <class P>
__global__ void k_fin_rz(const P* parts) {
  const P p = parts[blockIdx.x];
  State* st = p.g->st;
  if (st->done) return;
  double out[1];
  sum_partials<1>(p.g->partials, p.nblocks, out);
  if (threadIdx.x == 0) {
    const double rz = out[0];
    if (!(rz > 0.0) || !isfinite(rz)) {
      st->flags |= isfinite(rz) ? 0x4 : 0x1;  // NONPOSITIVE_PRECONDITIONER / NONFINITE
      st->done = 1;
      return;
    }
    st->beta = (st->iter == 0) ? 0.0 : rz / st->rz;
    st->rz = rz;
  }
}

template <class P>
__global__ void k_fin_dq(const P* parts) {
  const P p = parts[blockIdx.x];
  State* st = p.g->st;
  if (st->done) return;
  double out[1];
  sum_partials<1>(p.g->partials, p.nblocks, out);
  if (threadIdx.x == 0) {
    const double dq = out[0];
    if (!(dq > 0.0) || !isfinite(dq)) {
      st->flags |= isfinite(dq) ? 0x2 : 0x1;  // NONPOSITIVE_CURVATURE / NONFINITE
      st->done = 1;
      return;
    }
    st->alpha = st->rz / dq;
  }
}

template <class P>
__global__ void k_fin_update(const P* parts) {
  const P p = parts[blockIdx.x];
  State* st = p.g->st;
  if (st->done) return;
  double out[1];
  float m;
  sum_partials_max<1>(p.g->partials, p.nblocks, out, m);
  if (threadIdx.x == 0) {
    const double rr = out[0];
    st->rr = rr;
    st->iter += 1;
    st->total_iter += 1;
    if (!isfinite(rr)) {
      st->flags |= 0x1;
      st->done = 1;
    } else if (rr <= st->inner_target2 || st->iter >= p.g->max_iter) {
      st->done = 1;
    }
    st->rscale = p.g->rscale_bias * ((m > 0.0f && isfinite(m)) ? exp2f(-ceilf(log2f(m))) : 1.0f);
  }
}

template <class P>
__global__ void k_fin_divafter(const P* parts) {
  const P p = parts[blockIdx.x];
  if (!p.g->out_u || !p.g->out_v || !p.g->out_w) return;
  double out[1];
  float mx;
  sum_partials_max<1>(p.g->partials, p.nblocks, out, mx);
  if (threadIdx.x == 0) p.g->st->divmax_after = __float_as_uint(mx);
}

}  // namespace fp

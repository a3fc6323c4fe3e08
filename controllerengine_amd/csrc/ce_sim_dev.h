// Generic (interpreted) closed-loop simulation kernel, reference src/discrete_solves.cu:12-79,
// templated on the arithmetic type (double: FP64 chain; float: mixed chain) and on the observer /
// plant dimensions.  One candidate per lane; the state lives in VGPRs; plant and observer matrices
// and the decoded instruction chains are wave-uniform (scalar loads).  Used when runtime
// specialisation is disabled and for batches of uneven problems.
//
// The float form keeps the reference's compensated state: each state is a hi/lo pair, the lo parts
// enter K x, C x and the increments through dot products of their own, and a step adds its
// increment pair with TwoSum followed by Fast2Sum (src/main.cu:129-155 hladdf,
// src/discrete_solves.cu:35-77).  The instruction chains read the hi parts only, as the reference's
// do.  The FP64 form needs no pairs.
#pragma once

#include <hip/hip_runtime.h>

#include "ce_stage.h"

namespace ce {

template <typename R>
struct RealOps;
template <>
struct RealOps<double> {
  __device__ static double fma(double a, double b, double c) { return __builtin_fma(a, b, c); }
  __device__ static double fmin(double a, double b) { return __builtin_fmin(a, b); }
  __device__ static double fmax(double a, double b) { return __builtin_fmax(a, b); }
  __device__ static double inf() { return __builtin_inf(); }
  __device__ static double eps6() { return EPS6; }
  __device__ static double aval(const ProblemDev* P, const Ins& in, int which) { return which ? in.bval : in.aval; }
  __device__ static const double* sai(const ProblemDev* P) { return P->sai; }
  __device__ static const double* sb(const ProblemDev* P) { return P->sb; }
  __device__ static const double* sc(const ProblemDev* P) { return P->sc; }
  __device__ static double sd(const ProblemDev* P) { return P->sd; }
  __device__ static const double* dai(const ProblemDev* P) { return P->dai; }
  __device__ static const double* dbu(const ProblemDev* P) { return P->dbu; }
  __device__ static const double* dbw(const ProblemDev* P) { return P->dbw; }
  __device__ static double amp(const ProblemDev* P) { return P->ref_amp; }
};
template <>
struct RealOps<float> {
  __device__ static float fma(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
  __device__ static float fmin(float a, float b) { return __builtin_fminf(a, b); }
  __device__ static float fmax(float a, float b) { return __builtin_fmaxf(a, b); }
  __device__ static float inf() { return __builtin_inff(); }
  __device__ static float eps6() { return (float)EPS6; }
  __device__ static float aval(const ProblemDev* P, const Ins& in, int which) {
    return which ? (float)in.bval : (float)in.aval;
  }
  __device__ static const float* sai(const ProblemDev* P) { return P->fsai; }
  __device__ static const float* sb(const ProblemDev* P) { return P->fsb; }
  __device__ static const float* sc(const ProblemDev* P) { return P->fsc; }
  __device__ static float sd(const ProblemDev* P) { return P->fsd; }
  __device__ static const float* dai(const ProblemDev* P) { return P->fdai; }
  __device__ static const float* dbu(const ProblemDev* P) { return P->fdbu; }
  __device__ static const float* dbw(const ProblemDev* P) { return P->fdbw; }
  __device__ static float amp(const ProblemDev* P) { return P->fref_amp; }
};

// (HI, LO) += (h, l): TwoSum of the hi parts, the lo parts and the TwoSum error added, then Fast2Sum
// (reference src/main.cu:129-155).  Every operation is a separate rounding (no contraction).
__device__ __forceinline__ void hl_add(float& HI, float& LO, float h, float l) {
  const float sh = HI + h;
  const float b0 = sh - HI;
  const float sl = (HI - (sh - b0)) + (h - b0);
  const float v = (LO + l) + sl;
  HI = sh + v;
  LO = v - (HI - sh);
}

// Instruction chain (reference src/main.cu:260-291).  ops/count/col are wave-uniform.
template <typename R, int DIM, int MAXCOUNT, bool HAS_SEED>
__device__ __forceinline__ R interp_dev(const Ins* __restrict__ ops, int count, const R (&src)[DIM], R seed) {
  using O = RealOps<R>;
  R result = R(0);
  bool started = false;
#pragma unroll
  for (int r = 0; r < MAXCOUNT; r++) {
    if (r >= count) break;
    const Ins in = ops[r];
    if (in.op == 0) continue;
    R pos = src[0];
#pragma unroll
    for (int k = 1; k < DIM; k++)
      if (in.col == k) pos = src[k];
    R prev;
    if (started) prev = result;
    else if (HAS_SEED) prev = (seed == seed) ? seed : pos;
    else prev = pos;
    const R av = (R)in.aval, bv = (R)in.bval;
    const R a = in.akind == K_CONST ? av : (in.akind == K_STATE ? pos : prev);
    const R b = in.bkind == K_CONST ? bv : (in.bkind == K_STATE ? pos : prev);
    switch (in.op) {
      case 1: result = a + b; break;
      case 2: result = a - b; break;
      case 3: result = a * b; break;
      case 4: result = a / b; break;
      case 5: result = O::fmin(O::fmax(prev, a), b); break;
      case 6: result = (R)((R(0) < a) - (a < R(0))); break;
      default: break;
    }
    started = true;
  }
  return result;
}

// Closed-loop simulation over the compacted candidate list.  Stores the fit window of y
// (SoA: ybuf[(t - fit_start) * n + slot]) and the steady-state mean and overshoot of the INFO row
// (reference src/modalsetup.cu:52-62).  max_t fl(y_t yss) equals fl(max_t y_t * yss) for yss > 0 and
// fl(min_t y_t * yss) for yss < 0, since rounding is monotonic.
template <typename R, int NS, int ND>
__device__ __forceinline__ void k_sim_body(const ProblemDev* __restrict__ P, const StageBufs& B, int bx, int nbx, int by,
                                           int bz) {
  using O = RealOps<R>;
  R* __restrict__ ybuf_base = (R*)(sizeof(R) == 8 ? (void*)B.ybuf : (void*)B.ybuf32);
  const int idx = bx * 64 + threadIdx.x;
  if (idx >= *B.count) return;
  const int c = B.list[idx];
  const bool unstable = (B.flags[c] & F_UNSTABLE) != 0;
  const int n = B.n;
  const float* g = B.gains + (size_t)c * P->cuda_w;
  R K[NS], L[NS];
#pragma unroll
  for (int i = 0; i < NS; i++) { K[i] = (R)g[i]; L[i] = (R)g[NS + i]; }
  const R N = (R)g[2 * NS];
  constexpr bool HL = sizeof(R) == 4;  // compensated float state
  R xh[NS], xp[ND], xhl[NS], xpl[ND];
#pragma unroll
  for (int i = 0; i < NS; i++) xh[i] = xhl[i] = R(0);
#pragma unroll
  for (int i = 0; i < ND; i++) xp[i] = xpl[i] = R(0);
  const int T = P->T, f0 = P->fit_start_id, f1 = P->fit_start_id + P->fit_count;
  const int s0 = P->steady_start_id, k0 = P->ref_step_id, nw = P->nw;
  const int dc_n = P->dc_n, u_n = P->u_n;
  const R amp = O::amp(P), sd = O::sd(P);
  const R* sai = O::sai(P);
  const R* sb = O::sb(P);
  const R* scv = O::sc(P);
  const R* dai = O::dai(P);
  const R* dbu = O::dbu(P);
  const R* dbw = O::dbw(P);
  R ysum = R(0), ymax = -O::inf(), ymin = O::inf();
  R* ybuf = ybuf_base + idx;
  float* dy = (B.diag && B.d_y) ? B.d_y + (size_t)c * T : nullptr;
  for (int t = 0; t < T; t++) {
    const R y = interp_dev<R, ND, CE_MAX_DC_INSTRUCTIONS, false>(P->dc_ops, dc_n, xp, R(0));
    if (dy) dy[t] = (float)y;
    if (t >= f0 && t < f1) ybuf[(size_t)(t - f0) * n] = y;
    if (t >= s0) ysum = ysum + y;
    ymax = O::fmax(ymax, y);
    ymin = O::fmin(ymin, y);
    R kx = R(0);
#pragma unroll
    for (int k = 0; k < NS; k++) kx = O::fma(K[k], xh[k], kx);
    if (HL) {
#pragma unroll
      for (int k = 0; k < NS; k++) kx = O::fma(K[k], xhl[k], kx);
    }
    const R r = (t < k0) ? R(0) : amp;
    const R u_raw = O::fma(N, r, -kx);
    const R u = interp_dev<R, NS, CE_MAX_U_INSTRUCTIONS, true>(P->u_ops, u_n, xh, u_raw);
    R cx = R(0);
#pragma unroll
    for (int k = 0; k < NS; k++) cx = O::fma(scv[k], xh[k], cx);
    if (HL) {
#pragma unroll
      for (int k = 0; k < NS; k++) cx = O::fma(scv[k], xhl[k], cx);
    }
    const R yerr = y - O::fma(sd, u, cx);
    R dh[NS], dl[NS];
#pragma unroll
    for (int i = 0; i < NS; i++) {
      R acc = O::fma(L[i], yerr, sb[i] * u);
#pragma unroll
      for (int k = 0; k < NS; k++) acc = O::fma(sai[i * NS + k], xh[k], acc);
      dh[i] = acc;
      R lo = R(0);
      if (HL) {
#pragma unroll
        for (int k = 0; k < NS; k++) lo = O::fma(sai[i * NS + k], xhl[k], lo);
      }
      dl[i] = lo;
    }
#pragma unroll
    for (int i = 0; i < NS; i++) {
      if constexpr (HL) hl_add(xh[i], xhl[i], dh[i], dl[i]);
      else xh[i] = xh[i] + dh[i];
    }
    R nls[ND + 1];
#pragma unroll
    for (int i = 0; i < ND; i++) nls[i] = xp[i];
    nls[ND] = y;
    R W[MAXW];
#pragma unroll
    for (int q = 0; q < MAXW; q++) {
      W[q] = R(0);
      if (q < nw) W[q] = interp_dev<R, ND + 1, CE_MAX_NL_INSTRUCTIONS, false>(P->nl_ops[q], P->nl_n[q], nls, R(0));
    }
    R ph[ND], pl[ND];
#pragma unroll
    for (int i = 0; i < ND; i++) {
      R base = dbu[i] * u;
#pragma unroll
      for (int q = 0; q < MAXW; q++)
        if (q < nw) base = O::fma(dbw[i * nw + q], W[q], base);
      R acc = base;
#pragma unroll
      for (int k = 0; k < ND; k++) acc = O::fma(dai[i * ND + k], xp[k], acc);
      ph[i] = acc;
      R lo = R(0);
      if (HL) {
#pragma unroll
        for (int k = 0; k < ND; k++) lo = O::fma(dai[i * ND + k], xpl[k], lo);
      }
      pl[i] = lo;
    }
#pragma unroll
    for (int i = 0; i < ND; i++) {
      if constexpr (HL) hl_add(xp[i], xpl[i], ph[i], pl[i]);
      else xp[i] = xp[i] + ph[i];
    }
  }
  if (unstable) return;
  const R yss = ysum / (R)(T - s0);
  const R pk = O::fmax(R(0), ((yss < R(0)) ? ymin : ymax) * yss);
  double* fr = B.info + (size_t)c * P->info_w + P->info_freqresp;
  fr[2] = (double)(pk / O::fmax(yss * yss, O::eps6()) - R(1));
  fr[3] = (double)yss;
}

// The simulation is one long serial recurrence per candidate and holds only one wave on the SIMDs
// it uses; its waves take issue priority over the waves of kernels that run beside it on another
// stream, which fill the remaining issue slots.
template <typename R, int NS, int ND>
__global__ __launch_bounds__(64) void k_sim(const ProblemDev* __restrict__ P, StageBufs B) {
  __builtin_amdgcn_s_setprio(3);
  k_sim_body<R, NS, ND>(P, B, blockIdx.x, gridDim.x, 0, 0);
}

template <typename R, int NS, int ND>
__global__ __launch_bounds__(64) void k_sim_batch(BatchTab T) {
  __builtin_amdgcn_s_setprio(3);
  for (int b = blockIdx.x; b < T.total; b += gridDim.x) {
    const BatchSlot q = batch_slot(T, b);
    k_sim_body<R, NS, ND>(T.probs[q.p], T.bufs[q.p], q.lb, q.nb, 0, 0);
  }
}

template <typename R, int NS, int ND>
static void launch_sim_t(const ProblemDev* P, const StageBufs& B, R* ybuf, hipStream_t st) {
  (void)ybuf;
  hipLaunchKernelGGL((k_sim<R, NS, ND>), dim3((B.n + 63) / 64), dim3(64), 0, st, P, B);
  kmark(sizeof(R) == 8 ? "k_sim_fp64" : "k_sim_f32", st);
}

template <typename R, int NS>
static bool launch_sim_ns(int nd, const ProblemDev* P, const StageBufs& B, R* ybuf, hipStream_t st) {
  switch (nd) {
    case 1: launch_sim_t<R, NS, 1>(P, B, ybuf, st); return true;
    case 2: launch_sim_t<R, NS, 2>(P, B, ybuf, st); return true;
    case 3: launch_sim_t<R, NS, 3>(P, B, ybuf, st); return true;
    case 4: launch_sim_t<R, NS, 4>(P, B, ybuf, st); return true;
    case 5: launch_sim_t<R, NS, 5>(P, B, ybuf, st); return true;
    case 6: launch_sim_t<R, NS, 6>(P, B, ybuf, st); return true;
    case 7: launch_sim_t<R, NS, 7>(P, B, ybuf, st); return true;
    case 8: launch_sim_t<R, NS, 8>(P, B, ybuf, st); return true;
  }
  return false;
}

template <typename R>
static bool launch_sim(int ns, int nd, const ProblemDev* P, const StageBufs& B, R* ybuf, hipStream_t st) {
  switch (ns) {
    case 1: return launch_sim_ns<R, 1>(nd, P, B, ybuf, st);
    case 2: return launch_sim_ns<R, 2>(nd, P, B, ybuf, st);
    case 3: return launch_sim_ns<R, 3>(nd, P, B, ybuf, st);
    case 4: return launch_sim_ns<R, 4>(nd, P, B, ybuf, st);
    case 5: return launch_sim_ns<R, 5>(nd, P, B, ybuf, st);
    case 6: return launch_sim_ns<R, 6>(nd, P, B, ybuf, st);
    case 7: return launch_sim_ns<R, 7>(nd, P, B, ybuf, st);
    case 8: return launch_sim_ns<R, 8>(nd, P, B, ybuf, st);
  }
  return false;
}

}  // namespace ce

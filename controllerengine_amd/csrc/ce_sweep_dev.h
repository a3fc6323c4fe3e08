// One point of the margin sweep (reference src/allmargin.cu:100-117): G(z) = K^T (zI - A1)^-1 L by
// elimination with partial pivoting in registers (rows swapped by selects, right-hand side
// eliminated with the same multipliers, column-oriented back substitution), loop = P(z) G(z), its
// magnitude and argument, and b(w) of the Vinnicombe margin.  Templated on the arithmetic type: the
// double instance is the FP64 sweep and the mixed chain's refinement, the float instance the mixed
// sweep.
#pragma once

#include <hip/hip_runtime.h>

#include "ce_stage.h"

namespace ce {

template <typename Tp>
__device__ __forceinline__ const Tp* tab(const ProblemDev* P, int64_t off) {
  return reinterpret_cast<const Tp*>(reinterpret_cast<const uint8_t*>(P) + off);
}

__device__ __forceinline__ bool margin_skip(const ProblemDev* P, const StageBufs& B, int c) {
  const int f = B.flags[c];
  if (f & (F_NOMINAL_SKIPPED | F_NOMINAL_UNSTABLE)) return true;
  return B.info[(size_t)c * P->info_w + P->info_scoring] >= P->unstable_err;
}

template <typename R>
struct SweepMath;
template <>
struct SweepMath<double> {
  __device__ static double hyp(double a, double b) { return cedm::hypot(a, b); }
  __device__ static double arg(double y, double x) { return cedm::atan2(y, x); }
  __device__ static double sqrt(double a) { return cedm::sqrt_(a); }
  __device__ static bool fin(double a) { return cedm::isfinite_(a); }
  __device__ static double fma(double a, double b, double c) { return __builtin_fma(a, b, c); }
};
template <>
struct SweepMath<float> {
  __device__ static float hyp(float a, float b) { return hypotf(a, b); }
  __device__ static float arg(float y, float x) { return atan2f(y, x); }
  __device__ static float sqrt(float a) { return __builtin_sqrtf(a); }
  __device__ static bool fin(float a) { return __builtin_isfinite(a); }
  __device__ static float fma(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
};

// oa: A1 (NS x NS, row-major, double); gg: the candidate's gains row [K, L, N].  Returns false when
// the solution of the shifted system is not finite.
template <int NS, typename R>
__device__ __forceinline__ bool sweep_point(const ProblemDev* __restrict__ P, const double* __restrict__ oa,
                                            const float* __restrict__ gg, int w, R& m, R& raw, R& b) {
  using M = SweepMath<R>;
  const R zr = (R)tab<double>(P, P->off_zr)[w], zi = (R)tab<double>(P, P->off_zi)[w];
  R Mr[NS][NS], Mi[NS][NS], vr[NS], vi[NS];
#pragma unroll
  for (int i = 0; i < NS; i++) {
#pragma unroll
    for (int j = 0; j < NS; j++) {
      Mr[i][j] = (i == j ? zr : R(0)) - (R)oa[i * NS + j];
      Mi[i][j] = (i == j ? zi : R(0));
    }
    vr[i] = (R)gg[NS + i];
    vi[i] = R(0);
  }
#pragma unroll
  for (int k = 0; k < NS; k++) {
    int p = k;
    R best = Mr[k][k] * Mr[k][k] + Mi[k][k] * Mi[k][k];
#pragma unroll
    for (int i = k + 1; i < NS; i++) {
      const R s = Mr[i][k] * Mr[i][k] + Mi[i][k] * Mi[i][k];
      if (s > best) { best = s; p = i; }
    }
    if (best != R(0)) {
#pragma unroll
      for (int i = k + 1; i < NS; i++) {
        const bool sw = (p == i);
#pragma unroll
        for (int j = k; j < NS; j++) {
          const R ar = Mr[k][j], ai = Mi[k][j], br = Mr[i][j], bi = Mi[i][j];
          Mr[k][j] = sw ? br : ar; Mi[k][j] = sw ? bi : ai;
          Mr[i][j] = sw ? ar : br; Mi[i][j] = sw ? ai : bi;
        }
        const R ar = vr[k], ai = vi[k], br = vr[i], bi = vi[i];
        vr[k] = sw ? br : ar; vi[k] = sw ? bi : ai;
        vr[i] = sw ? ar : br; vi[i] = sw ? ai : bi;
      }
      const R pr = Mr[k][k], pim = Mi[k][k];
      const R dd = pr * pr + pim * pim;
#pragma unroll
      for (int i = k + 1; i < NS; i++) {
        const R ar = Mr[i][k], ai = Mi[i][k];
        Mr[i][k] = (ar * pr + ai * pim) / dd;
        Mi[i][k] = (ai * pr - ar * pim) / dd;
      }
    }
#pragma unroll
    for (int i = k + 1; i < NS; i++) {
      const R lr = Mr[i][k], li = Mi[i][k];
#pragma unroll
      for (int j = k + 1; j < NS; j++) {
        const R qr = lr * Mr[k][j] - li * Mi[k][j];
        const R qi = lr * Mi[k][j] + li * Mr[k][j];
        Mr[i][j] = Mr[i][j] - qr;
        Mi[i][j] = Mi[i][j] - qi;
      }
      const R qr = lr * vr[k] - li * vi[k];
      const R qi = lr * vi[k] + li * vr[k];
      vr[i] = vr[i] - qr;
      vi[i] = vi[i] - qi;
    }
  }
  R xr[NS], xi[NS];
  bool fin = true;
#pragma unroll
  for (int i = NS - 1; i >= 0; i--) {
    const R pr = Mr[i][i], pim = Mi[i][i];
    const R dd = pr * pr + pim * pim;
    xr[i] = (vr[i] * pr + vi[i] * pim) / dd;
    xi[i] = (vi[i] * pr - vr[i] * pim) / dd;
#pragma unroll
    for (int j = 0; j < i; j++) {
      const R qr = Mr[j][i] * xr[i] - Mi[j][i] * xi[i];
      const R qi = Mr[j][i] * xi[i] + Mi[j][i] * xr[i];
      vr[j] = vr[j] - qr;
      vi[j] = vi[j] - qi;
    }
  }
#pragma unroll
  for (int i = 0; i < NS; i++) fin = fin && M::fin(xr[i]) && M::fin(xi[i]);
  if (!fin) return false;
  R gr = R(0), gi = R(0);
#pragma unroll
  for (int i = 0; i < NS; i++) {
    const R k = (R)gg[i];
    gr = M::fma(k, xr[i], gr);
    gi = M::fma(k, xi[i], gi);
  }
  const R pr = (R)tab<double>(P, P->off_pr)[w], pim = (R)tab<double>(P, P->off_pi)[w];
  const R rr = pr * gr - pim * gi, ri = pr * gi + pim * gr;
  m = M::hyp(rr, ri);
  raw = M::arg(ri, rr);
  const R alpha = (R)P->alpha;
  const R Ps = alpha * (R)tab<double>(P, P->off_pabs)[w];
  const R Cs = M::hyp(gr, gi) / alpha;
  const R num = M::hyp(R(1) + rr, ri);
  b = num / (M::sqrt(R(1) + Ps * Ps) * M::sqrt(R(1) + Cs * Cs));
  return true;
}

}  // namespace ce

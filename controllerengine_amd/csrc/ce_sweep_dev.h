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

// Flags-only form, for sweeps that run beside the modal fit: the score the fit writes may still
// turn the candidate into a skipped one, which the crossing scan decides after both.
__device__ __forceinline__ bool margin_skip_flags(const StageBufs& B, int c) {
  return (B.flags[c] & (F_NOMINAL_SKIPPED | F_NOMINAL_UNSTABLE)) != 0;
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

template <typename R>
__device__ __forceinline__ void sweep_tail(const ProblemDev* __restrict__ P, int w, R gr, R gi, R& m, R& raw, R& b);

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
  sweep_tail<R>(P, w, gr, gi, m, raw, b);
  return true;
}

// |loop|, arg(loop) and b(w) from G(z_w) (reference src/allmargin.cu:104-117).
template <typename R>
__device__ __forceinline__ void sweep_tail(const ProblemDev* __restrict__ P, int w, R gr, R gi, R& m, R& raw, R& b) {
  using M = SweepMath<R>;
  const R pr = (R)tab<double>(P, P->off_pr)[w], pim = (R)tab<double>(P, P->off_pi)[w];
  const R rr = pr * gr - pim * gi, ri = pr * gi + pim * gr;
  m = M::hyp(rr, ri);
  raw = M::arg(ri, rr);
  const R alpha = (R)P->alpha;
  const R Ps = alpha * (R)tab<double>(P, P->off_pabs)[w];
  const R Cs = M::hyp(gr, gi) / alpha;
  const R num = M::hyp(R(1) + rr, ri);
  b = num / (M::sqrt(R(1) + Ps * Ps) * M::sqrt(R(1) + Cs * Cs));
}

// Float sweep on the Hessenberg form (mixed chain only; the FP64 chain keeps the elimination
// above for parity).  A1 = Q H Q^T by Householder reflectors, so
//   G(z) = K^T (zI - A1)^-1 L = (Q^T K)^T (zI - H)^-1 (Q^T L),
// and each frequency costs an O(NS^2) solve of the Hessenberg system (partial pivoting between
// the two candidate rows of each column) instead of an O(NS^3) dense elimination.  The pivots are
// inverted with the hardware reciprocal (1 ulp): these float values only locate the points that
// are recomputed in FP64, with guard bands far wider than the rounding.
template <int NS>
__device__ __forceinline__ void hess_prep(const double* __restrict__ oa, const float* __restrict__ gg,
                                          float (&H)[NS][NS], float (&kq)[NS], float (&lq)[NS]) {
#pragma unroll
  for (int i = 0; i < NS; i++) {
#pragma unroll
    for (int j = 0; j < NS; j++) H[i][j] = (float)oa[i * NS + j];
    kq[i] = gg[i];
    lq[i] = gg[NS + i];
  }
#pragma unroll
  for (int c = 0; c + 2 < NS; c++) {
    float sig = 0.0f;
#pragma unroll
    for (int i = c + 2; i < NS; i++) sig = __builtin_fmaf(H[i][c], H[i][c], sig);
    if (sig == 0.0f) continue;
    const float alpha = H[c + 1][c];
    const float nrm = __builtin_sqrtf(__builtin_fmaf(alpha, alpha, sig));
    const float beta = alpha >= 0.0f ? -nrm : nrm;
    float v[NS];
#pragma unroll
    for (int i = 0; i < NS; i++) v[i] = (i == c + 1) ? alpha - beta : (i > c + 1 ? H[i][c] : 0.0f);
    const float vtv = __builtin_fmaf(v[c + 1], v[c + 1], sig);
    const float tau = 2.0f / vtv;
    // H <- P H, then H <- H P, with P = I - tau v v^T (v zero above row c+1)
#pragma unroll
    for (int j = 0; j < NS; j++) {
      float sdot = 0.0f;
#pragma unroll
      for (int i = c + 1; i < NS; i++) sdot = __builtin_fmaf(v[i], H[i][j], sdot);
      const float f = tau * sdot;
#pragma unroll
      for (int i = c + 1; i < NS; i++) H[i][j] = __builtin_fmaf(-f, v[i], H[i][j]);
    }
#pragma unroll
    for (int i = 0; i < NS; i++) {
      float sdot = 0.0f;
#pragma unroll
      for (int j = c + 1; j < NS; j++) sdot = __builtin_fmaf(H[i][j], v[j], sdot);
      const float f = tau * sdot;
#pragma unroll
      for (int j = c + 1; j < NS; j++) H[i][j] = __builtin_fmaf(-f, v[j], H[i][j]);
    }
    float sk = 0.0f, sl = 0.0f;
#pragma unroll
    for (int i = c + 1; i < NS; i++) {
      sk = __builtin_fmaf(v[i], kq[i], sk);
      sl = __builtin_fmaf(v[i], lq[i], sl);
    }
    const float fk = tau * sk, fl = tau * sl;
#pragma unroll
    for (int i = c + 1; i < NS; i++) {
      kq[i] = __builtin_fmaf(-fk, v[i], kq[i]);
      lq[i] = __builtin_fmaf(-fl, v[i], lq[i]);
    }
  }
}

template <int NS>
__device__ __forceinline__ bool hess_point(const ProblemDev* __restrict__ P, const float (&H)[NS][NS],
                                           const float (&kq)[NS], const float (&lq)[NS], int w, float& m, float& raw,
                                           float& b) {
  const float zr = (float)tab<double>(P, P->off_zr)[w], zi = (float)tab<double>(P, P->off_zi)[w];
  // M = zI - H, upper Hessenberg; only entries with j >= i - 1 are used
  float Mr[NS][NS], Mi[NS][NS], vr[NS], vi[NS];
#pragma unroll
  for (int i = 0; i < NS; i++) {
#pragma unroll
    for (int j = 0; j < NS; j++) {
      Mr[i][j] = (i == j ? zr : 0.0f) - H[i][j];
      Mi[i][j] = (i == j ? zi : 0.0f);
    }
    vr[i] = lq[i];
    vi[i] = 0.0f;
  }
#pragma unroll
  for (int k = 0; k + 1 < NS; k++) {
    const float a = Mr[k][k] * Mr[k][k] + Mi[k][k] * Mi[k][k];
    const float c = Mr[k + 1][k] * Mr[k + 1][k] + Mi[k + 1][k] * Mi[k + 1][k];
    const bool sw = c > a;
#pragma unroll
    for (int j = k; j < NS; j++) {
      const float ar = Mr[k][j], ai = Mi[k][j], br = Mr[k + 1][j], bi = Mi[k + 1][j];
      Mr[k][j] = sw ? br : ar;
      Mi[k][j] = sw ? bi : ai;
      Mr[k + 1][j] = sw ? ar : br;
      Mi[k + 1][j] = sw ? ai : bi;
    }
    {
      const float ar = vr[k], ai = vi[k], br = vr[k + 1], bi = vi[k + 1];
      vr[k] = sw ? br : ar;
      vi[k] = sw ? bi : ai;
      vr[k + 1] = sw ? ar : br;
      vi[k + 1] = sw ? ai : bi;
    }
    const float pr = Mr[k][k], pi = Mi[k][k];
    const float inv = __builtin_amdgcn_rcpf(pr * pr + pi * pi);
    const float lr = (Mr[k + 1][k] * pr + Mi[k + 1][k] * pi) * inv;
    const float li = (Mi[k + 1][k] * pr - Mr[k + 1][k] * pi) * inv;
#pragma unroll
    for (int j = k + 1; j < NS; j++) {
      Mr[k + 1][j] = Mr[k + 1][j] - (lr * Mr[k][j] - li * Mi[k][j]);
      Mi[k + 1][j] = Mi[k + 1][j] - (lr * Mi[k][j] + li * Mr[k][j]);
    }
    const float qr = lr * vr[k] - li * vi[k], qi = lr * vi[k] + li * vr[k];
    vr[k + 1] = vr[k + 1] - qr;
    vi[k + 1] = vi[k + 1] - qi;
  }
  float xr[NS], xi[NS];
  bool fin = true;
#pragma unroll
  for (int i = NS - 1; i >= 0; i--) {
    float sr = vr[i], si = vi[i];
#pragma unroll
    for (int j = i + 1; j < NS; j++) {
      sr = sr - (Mr[i][j] * xr[j] - Mi[i][j] * xi[j]);
      si = si - (Mr[i][j] * xi[j] + Mi[i][j] * xr[j]);
    }
    const float pr = Mr[i][i], pi = Mi[i][i];
    const float inv = __builtin_amdgcn_rcpf(pr * pr + pi * pi);
    xr[i] = (sr * pr + si * pi) * inv;
    xi[i] = (si * pr - sr * pi) * inv;
    fin = fin && __builtin_isfinite(xr[i]) && __builtin_isfinite(xi[i]);
  }
  if (!fin) return false;
  float gr = 0.0f, gi = 0.0f;
#pragma unroll
  for (int i = 0; i < NS; i++) {
    gr = __builtin_fmaf(kq[i], xr[i], gr);
    gi = __builtin_fmaf(kq[i], xi[i], gi);
  }
  sweep_tail<float>(P, w, gr, gi, m, raw, b);
  return true;
}

}  // namespace ce

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

// |loop|, arg(loop) and b(w) from G(z_w) and the plant response P(z_w) = pr + i pim, |P| = pabs
// (reference src/allmargin.cu:104-117).
template <typename R>
__device__ __forceinline__ void sweep_tail_vals(R pr, R pim, R pabs, R alpha, R gr, R gi, R& m, R& raw, R& b) {
  using M = SweepMath<R>;
  const R rr = pr * gr - pim * gi, ri = pr * gi + pim * gr;
  m = M::hyp(rr, ri);
  raw = M::arg(ri, rr);
  const R Ps = alpha * pabs;
  const R Cs = M::hyp(gr, gi) / alpha;
  const R num = M::hyp(R(1) + rr, ri);
  b = num / (M::sqrt(R(1) + Ps * Ps) * M::sqrt(R(1) + Cs * Cs));
}
template <typename R>
__device__ __forceinline__ void sweep_tail(const ProblemDev* __restrict__ P, int w, R gr, R gi, R& m, R& raw, R& b) {
  sweep_tail_vals<R>((R)tab<double>(P, P->off_pr)[w], (R)tab<double>(P, P->off_pi)[w],
                     (R)tab<double>(P, P->off_pabs)[w], (R)P->alpha, gr, gi, m, raw, b);
}

// |loop|, arg(loop) and b(w) in float for the mixed sweep's locator: the formulas of
// sweep_tail_vals with each modulus as sqrt(x^2 + y^2) (no rescaling: a value beyond the float
// range comes out infinite, and such a point is recomputed in FP64) and the quotient by the
// hardware reciprocal; a few float ulps, within the locator's error allowance.
__device__ __forceinline__ void locator_tail(float pr, float pim, float pabs, float alpha, float gr, float gi,
                                             float& m, float& raw, float& b) {
  const float rr = __builtin_fmaf(pr, gr, -(pim * gi)), ri = __builtin_fmaf(pr, gi, pim * gr);
  m = __builtin_sqrtf(__builtin_fmaf(rr, rr, ri * ri));
  raw = atan2f(ri, rr);
  const float Ps = alpha * pabs;
  const float ia = __builtin_amdgcn_rcpf(alpha);
  const float Cs2 = __builtin_fmaf(gr, gr, gi * gi) * (ia * ia);
  const float a1 = 1.0f + rr;
  const float num = __builtin_sqrtf(__builtin_fmaf(a1, a1, ri * ri));
  b = num * __builtin_amdgcn_rcpf(__builtin_sqrtf(__builtin_fmaf(Ps, Ps, 1.0f)) * __builtin_sqrtf(1.0f + Cs2));
}

// Sweep of the mixed chain on the Hessenberg form (the FP64 chain keeps the dense elimination
// above for parity), in delta form: A1 - I = Q H Q^T by Householder reflectors, so
//   G(z) = K^T (zI - A1)^-1 L = (Q^T K)^T ((z - 1) I - H)^-1 (Q^T L),
// and each frequency costs an O(NS^2) solve of the Hessenberg system instead of an O(NS^3) dense
// elimination.  The solve is FP64: the sampled plants put the eigenvalues of A1 within O(ts) of
// 1, where (z - 1) I - H at the low frequencies is conditioned far beyond what float carries
// (relative errors of 1e-2 in a float solve).  The loop value is then rounded to float and the
// tail (|loop|, arg, b) evaluated in float; these values only locate the points that are
// recomputed with the dense FP64 elimination.
// The Householder reduction of A1 - I (delta form: its diagonal is 1 - O(ts) for a sampled plant)
// on the LDS block sh = [H (row-major), Q^T K, Q^T L] of one workgroup (one wave).  Each reflector
// is formed by every lane from the same LDS values; then lane j updates column j of H (H <- P H)
// and lanes NS, NS + 1 update Q^T K and Q^T L; then lane i updates row i (H <- H P).  O(NS^2) work
// per reflector and lane, and no matrix in registers.
template <int NS>
__device__ __forceinline__ void hess_prep_lds(const double* __restrict__ oa, const float* __restrict__ gg,
                                              double* sh, double& hnorm) {
  const int lane = threadIdx.x;
  double* H = sh;
  double* kq = sh + NS * NS;
  double* lq = sh + NS * NS + NS;
  for (int e = lane; e < NS * NS; e += 64) H[e] = oa[e] - ((e / NS) == (e % NS) ? 1.0 : 0.0);
  if (lane < NS) {
    kq[lane] = (double)gg[lane];
    lq[lane] = (double)gg[NS + lane];
  }
  __syncthreads();
#pragma unroll
  for (int c = 0; c + 2 < NS; c++) {
    double sig = 0.0;
#pragma unroll
    for (int i = c + 2; i < NS; i++) sig = __builtin_fma(H[i * NS + c], H[i * NS + c], sig);
    if (sig == 0.0) continue;  // the same LDS values in every lane: uniform
    const double alpha = H[(c + 1) * NS + c];
    const double nrm = __builtin_sqrt(__builtin_fma(alpha, alpha, sig));
    const double beta = alpha >= 0.0 ? -nrm : nrm;
    double v[NS];
#pragma unroll
    for (int i = 0; i < NS; i++) v[i] = (i == c + 1) ? alpha - beta : (i > c + 1 ? H[i * NS + c] : 0.0);
    const double vtv = __builtin_fma(v[c + 1], v[c + 1], sig);
    const double tau = 2.0 / vtv;
    if (lane < NS + 2) {
      // lane j < NS: column j of H; lane NS: Q^T K; lane NS + 1: Q^T L
      double* col = lane < NS ? H + lane : (lane == NS ? kq : lq);
      const int stride = lane < NS ? NS : 1;
      double sdot = 0.0;
#pragma unroll
      for (int i = c + 1; i < NS; i++) sdot = __builtin_fma(v[i], col[i * stride], sdot);
      const double f = tau * sdot;
#pragma unroll
      for (int i = c + 1; i < NS; i++) col[i * stride] = __builtin_fma(-f, v[i], col[i * stride]);
    }
    __syncthreads();
    if (lane < NS) {
      double* row = H + lane * NS;
      double sdot = 0.0;
#pragma unroll
      for (int j = c + 1; j < NS; j++) sdot = __builtin_fma(row[j], v[j], sdot);
      const double f = tau * sdot;
#pragma unroll
      for (int j = c + 1; j < NS; j++) row[j] = __builtin_fma(-f, v[j], row[j]);
    }
    __syncthreads();
  }
  double hf = 0.0;
#pragma unroll
  for (int e = 0; e < NS * NS; e++) hf = __builtin_fma(H[e], H[e], hf);
  hnorm = __builtin_sqrt(hf);
}

// One frequency: G = kq^T M^-1 lq with M = (z - 1) I - H upper Hessenberg, by Gaussian elimination
// with partial pivoting between the two candidate rows of each column, M = P L U.  With
// w = L^-1 P^T lq (the eliminated right-hand side) and y = U^-T kq (forward substitution, run
// row by row as the rows of U are produced), G = kq^T U^-1 w = y^T w, so no row of U is kept:
// the live state is the carried row, the partly reduced kq and the running sum.
// err: relative error of the float loop value (and, in radians, of its phase): the FP64 estimate
//   SWEEP_ERR_SAFETY eps_d (||M|| / min |pivot|) ||y|| ||w|| / |G|
// plus SWEEP_ERR_FLOOR for the float rounding of G and the float tail; 4 when unknown.
// The Hessenberg data come from a wave-uniform LDS block sh = [H (row-major), Q^T K, Q^T L];
// (zmr, zi) = z - 1, and (pr, pim, pabs) the plant response at the point.
template <int NS>
__device__ __forceinline__ bool hess_point(const double* __restrict__ sh, double hnorm, double zmr, double zi,
                                           float pr, float pim, float pabs, float alpha, float& m, float& raw,
                                           float& b, float& err) {
  const double* H0 = sh;
  const double* kq = sh + NS * NS;
  const double* lq = sh + NS * NS + NS;
#define H(i, j) H0[(i) * NS + (j)]
  // carried row (row k of the elimination before its pivot choice), its right-hand side, and
  // kq reduced by the rows of U produced so far
  double cr[NS], ci[NS], yr[NS], yi[NS];
#pragma unroll
  for (int j = 0; j < NS; j++) {
    cr[j] = (j == 0 ? zmr : 0.0) - H(0, j);
    ci[j] = (j == 0 ? zi : 0.0);
    yr[j] = kq[j];
    yi[j] = 0.0;
  }
  double vr = lq[0], vi = 0.0;
  double sr = 0.0, si = 0.0, y2 = 0.0, w2 = 0.0, pmin2 = __builtin_inf();
  // (the products are written as explicit fused multiply-adds: these values locate points and
  // carry their own error estimate, and the file is built without contraction)
#define FMA __builtin_fma
#pragma unroll
  for (int k = 0; k < NS; k++) {
    // row k + 1 of M: -H[k+1][j], plus z - 1 on the diagonal; its column-k entry is real
    const double nk = (k + 1 < NS) ? -H(k + 1 < NS ? k + 1 : k, k) : 0.0;
    const bool sw = (k + 1 < NS) && nk * nk > FMA(cr[k], cr[k], ci[k] * ci[k]);
    const double ukr = sw ? nk : cr[k], uki = sw ? 0.0 : ci[k];
    const double okr = sw ? cr[k] : nk, oki = sw ? ci[k] : 0.0;
    const double nv = (k + 1 < NS) ? lq[k + 1 < NS ? k + 1 : k] : 0.0;
    const double wr = sw ? nv : vr, wi = sw ? 0.0 : vi;
    const double ovr = sw ? vr : nv, ovi = sw ? vi : 0.0;
    const double p2 = FMA(ukr, ukr, uki * uki);
    pmin2 = __builtin_fmin(pmin2, p2);
    // 1 / p2 by the hardware reciprocal and two Newton steps (full double precision; a zero,
    // infinite or NaN pivot gives NaN here, and the point is then reported not finite)
    double inv = __builtin_amdgcn_rcp(p2);
    inv = FMA(inv, FMA(-p2, inv, 1.0), inv);
    inv = FMA(inv, FMA(-p2, inv, 1.0), inv);
    // l = o_k / u_k, y_k = kq'_k / u_k
    const double lr = FMA(okr, ukr, oki * uki) * inv, li = FMA(oki, ukr, -(okr * uki)) * inv;
    const double ykr = FMA(yr[k], ukr, yi[k] * uki) * inv, yki = FMA(yi[k], ukr, -(yr[k] * uki)) * inv;
#pragma unroll
    for (int j = k + 1; j < NS; j++) {
      const double njr = (j == k + 1 ? zmr : 0.0) - H(k + 1 < NS ? k + 1 : k, j), nji = (j == k + 1 ? zi : 0.0);
      const double ujr = sw ? njr : cr[j], uji = sw ? nji : ci[j];
      const double ojr = sw ? cr[j] : njr, oji = sw ? ci[j] : nji;
      // carried row o - l u, and kq' - u y_k
      cr[j] = FMA(-lr, ujr, FMA(li, uji, ojr));
      ci[j] = FMA(-lr, uji, FMA(-li, ujr, oji));
      yr[j] = FMA(-ujr, ykr, FMA(uji, yki, yr[j]));
      yi[j] = FMA(-ujr, yki, FMA(-uji, ykr, yi[j]));
    }
    vr = FMA(-lr, wr, FMA(li, wi, ovr));
    vi = FMA(-lr, wi, FMA(-li, wr, ovi));
    sr = FMA(ykr, wr, FMA(-yki, wi, sr));
    si = FMA(ykr, wi, FMA(yki, wr, si));
    y2 = FMA(ykr, ykr, FMA(yki, yki, y2));
    w2 = FMA(wr, wr, FMA(wi, wi, w2));
  }
  const double g2 = FMA(sr, sr, si * si);
#undef FMA
  if (!(pmin2 > 0.0) || !__builtin_isfinite(g2)) return false;
  // the estimate needs a digit or two: float, with approximate reciprocals.  ||M|| <=
  // sqrt(NS) |z - 1| + ||H||_F.  An overflow or a NaN gives 4 (every guard then holds).
  const float mn = __builtin_sqrtf((float)NS) * __builtin_sqrtf((float)__builtin_fma(zmr, zmr, zi * zi)) + (float)hnorm;
  const float ratio = (float)((y2 * w2) * __builtin_amdgcn_rcp(pmin2 * g2));
  const float e = SWEEP_ERR_SAFETY * 1.1102230246251565e-16f * mn * __builtin_sqrtf(ratio);
  err = (e <= 1.0f) ? e + SWEEP_ERR_FLOOR : 4.0f;
  locator_tail(pr, pim, pabs, alpha, (float)sr, (float)si, m, raw, b);
  return true;
#undef H
}

}  // namespace ce

// Normal-equation solve and behaviour score of the modal fit, reference src/modalsolve.cu:150-203:
// column scaling by sqrt(diag), LDL^T with symmetric pivoting (Eigen LDLT<Lower> restated; zero
// pivots solve to zero), per-mode energy on the unscaled Gram, energy-weighted wn and zeta of the
// oscillatory modes, hinge errors.  Shared by the FP64 fit and the mixed closed-form fit; runs on
// one lane per candidate over LDS-resident data.
#pragma once

#include <float.h>

#include "ce_stage.h"

namespace ce {

__device__ __forceinline__ int tri(int i, int j) { return i * (i + 1) / 2 + j; }

// G2: packed lower triangle of the 2nt x 2nt Gram (overwritten), atb: right-hand side; dg, sd, cs,
// x: scratch vectors of 2nt.  u: unique modes; modes read from B.mode_re/mode_im of candidate c.
__device__ __noinline__ void solve_and_score(const ProblemDev* __restrict__ P, const StageBufs& B, int c, int u,
                                             double* row, double* G2, const double* atb, double* dg, double* sd,
                                             double* cs, double* x) {
  const int nt = P->nt, nc2 = 2 * nt;
  const int n = B.n;
  double* sc = row + P->info_scoring;
  for (int i = 0; i < nc2; i++) {
    dg[i] = G2[tri(i, i)];
    sd[i] = (i + 1 < nc2) ? G2[tri(i + 1, i)] : 0.0;
    const double s = cedm::sqrt_(dg[i]);
    cs[i] = (s > EPS6) ? s : 1.0;
  }
  // scaled system D^-1 G D^-1 (lower triangle) and right-hand side D^-1 b
  for (int i = 0; i < nc2; i++) {
    const double inv_i = 1.0 / cs[i];
    for (int j = 0; j <= i; j++) {
      const double inv_j = 1.0 / cs[j];
      G2[tri(i, j)] = (inv_i * G2[tri(i, j)]) * inv_j;
    }
    x[i] = atb[i] / cs[i];
  }
  // LDL^T with symmetric pivoting on the largest remaining diagonal (lower storage)
  int trans[MAXNC];
  bool all_zero = false;
  const int N2 = nc2;
  for (int k = 0; k < N2; k++) {
    int idx = k;
    double best = cedm::fabs_(G2[tri(k, k)]);
    for (int i = k + 1; i < N2; i++) {
      const double a = cedm::fabs_(G2[tri(i, i)]);
      if (a > best) { best = a; idx = i; }
    }
    trans[k] = idx;
    if (k != idx) {
      for (int j = 0; j < k; j++) { const double t = G2[tri(k, j)]; G2[tri(k, j)] = G2[tri(idx, j)]; G2[tri(idx, j)] = t; }
      for (int i = idx + 1; i < N2; i++) { const double t = G2[tri(i, k)]; G2[tri(i, k)] = G2[tri(i, idx)]; G2[tri(i, idx)] = t; }
      { const double t = G2[tri(k, k)]; G2[tri(k, k)] = G2[tri(idx, idx)]; G2[tri(idx, idx)] = t; }
      for (int i = k + 1; i < idx; i++) { const double t = G2[tri(i, k)]; G2[tri(i, k)] = G2[tri(idx, i)]; G2[tri(idx, i)] = t; }
    }
    if (k > 0) {
      double tmp[MAXNC];
      for (int j = 0; j < k; j++) tmp[j] = G2[tri(j, j)] * G2[tri(k, j)];
      double acc = 0.0;
      for (int j = 0; j < k; j++) acc = cedm::fma_(G2[tri(k, j)], tmp[j], acc);
      G2[tri(k, k)] = G2[tri(k, k)] - acc;
      for (int i = k + 1; i < N2; i++) {
        double a = 0.0;
        for (int j = 0; j < k; j++) a = cedm::fma_(G2[tri(i, j)], tmp[j], a);
        G2[tri(i, k)] = G2[tri(i, k)] - a;
      }
    }
    const double akk = G2[tri(k, k)];
    const bool valid = cedm::fabs_(akk) > 0.0;
    if (k == 0 && !valid) { all_zero = true; break; }
    if (valid)
      for (int i = k + 1; i < N2; i++) G2[tri(i, k)] = G2[tri(i, k)] / akk;
  }
  if (all_zero) {
    for (int i = 0; i < N2; i++) x[i] = 0.0;
  } else {
    for (int k = 0; k < N2; k++) { const double t = x[k]; x[k] = x[trans[k]]; x[trans[k]] = t; }
    for (int i = 0; i < N2; i++) {
      double acc = 0.0;
      for (int j = 0; j < i; j++) acc = cedm::fma_(G2[tri(i, j)], x[j], acc);
      x[i] = x[i] - acc;
    }
    for (int i = 0; i < N2; i++) {
      const double d = G2[tri(i, i)];
      x[i] = (cedm::fabs_(d) > DBL_MIN) ? x[i] / d : 0.0;
    }
    for (int i = N2 - 1; i >= 0; i--) {
      double acc = 0.0;
      for (int j = i + 1; j < N2; j++) acc = cedm::fma_(G2[tri(j, i)], x[j], acc);
      x[i] = x[i] - acc;
    }
    for (int k = N2 - 1; k >= 0; k--) { const double t = x[k]; x[k] = x[trans[k]]; x[trans[k]] = t; }
  }
  for (int i = 0; i < N2; i++) x[i] = x[i] / cs[i];

  // per-mode energy, energy-weighted wn and zeta of the oscillatory modes (src/modalsolve.cu:168-190)
  double* md = row + P->info_modal;
  double* fr = row + P->info_freqresp;
  double esum = 0.0, lw = 0.0, zw = 0.0;
  for (int e = 0; e < u; e++) {
    const int id = 2 * e;
    const double c0 = x[id], c1 = x[id + 1];
    const double g00 = dg[id], g01 = sd[id], g11 = dg[id + 1];
    const double energy = ((c0 * c0) * g00 + ((2.0 * c0) * c1) * g01) + (c1 * c1) * g11;
    const double re_e = B.mode_re[(size_t)e * n + c], im_e = B.mode_im[(size_t)e * n + c];
    const double wn_e = cedm::fmax_(cedm::sqrt_(cedm::fma_(re_e, re_e, im_e * im_e)), EPS6);
    md[id] = wn_e;
    md[id + 1] = -re_e / wn_e;
    md[2 * nt + e] = energy;
    if (im_e < EPS6) continue;
    esum = esum + energy;
    lw = lw + energy * cedm::log(wn_e);
    zw = zw + energy * (-re_e / wn_e);
  }
  md[3 * nt] = (double)u;
  double wn, zeta;
  if (esum <= EPS6) { wn = EPS6; zeta = 0.0; }
  else { wn = cedm::exp(lw / esum); zeta = zw / esum; }
  const double dmp = P->damping, d2 = dmp * dmp;
  const double inner = ((4.0 * d2) * d2 - 4.0 * d2) + 2.0;
  const double twn = ((2.0 * PI_D) * P->bandwidth) / cedm::sqrt_((1.0 - 2.0 * d2) + cedm::sqrt_(inner));
  const double lr = cedm::log(wn / twn);
  double beh = lr * lr + ((zeta - dmp) * (zeta - dmp) / dmp) / dmp;
  const double os = P->overshoot, sse = P->sse;
  const double err_os = cedm::fmax_(0.0, fr[2] - os) / os;
  beh = beh + err_os * err_os;
  const double err_ss = cedm::fmax_(0.0, cedm::fabs_(fr[3] / P->ref_amp - 1.0) - sse) / sse;
  beh = beh + err_ss * err_ss;
  sc[0] = cedm::isfinite_(beh) ? beh : P->eig_fail_err;
  sc[1] = sc[0];
  fr[0] = wn;
  fr[1] = zeta;
}

}  // namespace ce

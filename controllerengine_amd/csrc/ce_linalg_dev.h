// Per-lane dense linear algebra for the gfx950 FP64 stages: the real-Schur eigenvalue solver
// (Householder Hessenberg + Francis double-shift QR) and the pivoted LDL^T solve.
//
// Each routine follows the step order of the reference's Eigen calls (EigenSolver<double> at
// reference src/eigensolve.cu:11 and src/allmargin.cu:72; LDLT at src/modalsolve.cu:67) with the
// operation order fixed (fma chains for inner products, no contraction elsewhere) so that the
// device result equals the CPU restatement bit for bit.
#pragma once

#include <float.h>

#include "ce_common.h"

namespace ce {

constexpr int TS = MAXT;  // row stride of the eigen work matrix

struct HHv {
  double e0, e1, tau, beta;
};

__device__ __forceinline__ HHv make_hh(double c0, double v1, double v2, int len) {
  HHv h;
  double tail = cedm::fma_(v1, v1, 0.0);
  if (len == 3) tail = cedm::fma_(v2, v2, tail);
  if (tail <= DBL_MIN) {
    h.tau = 0.0;
    h.beta = c0;
    h.e0 = 0.0;
    h.e1 = 0.0;
  } else {
    double beta = cedm::sqrt_(cedm::fma_(c0, c0, tail));
    if (c0 >= 0.0) beta = -beta;
    const double den = c0 - beta;
    h.e0 = v1 / den;
    h.e1 = (len == 3) ? v2 / den : 0.0;
    h.tau = (beta - c0) / beta;
    h.beta = beta;
  }
  return h;
}

// Householder with a 1- or 2-entry essential part applied from the left to rows r0..r0+rows-1,
// columns c0..c1-1 (rows = 2 or 3).
__device__ __forceinline__ void hh_left_small(double* T, int r0, int c0, int c1, int rows, const HHv& h) {
  if (h.tau == 0.0) return;
  for (int j = c0; j < c1; j++) {
    double tmp = cedm::fma_(h.e0, T[(r0 + 1) * TS + j], 0.0);
    if (rows == 3) tmp = cedm::fma_(h.e1, T[(r0 + 2) * TS + j], tmp);
    tmp = tmp + T[r0 * TS + j];
    T[r0 * TS + j] = T[r0 * TS + j] - h.tau * tmp;
    T[(r0 + 1) * TS + j] = T[(r0 + 1) * TS + j] - (h.tau * h.e0) * tmp;
    if (rows == 3) T[(r0 + 2) * TS + j] = T[(r0 + 2) * TS + j] - (h.tau * h.e1) * tmp;
  }
}

__device__ __forceinline__ void hh_right_small(double* T, int r1, int c0, int cols, const HHv& h) {
  if (h.tau == 0.0) return;
  for (int i = 0; i < r1; i++) {
    double* row = T + i * TS + c0;
    double tmp = cedm::fma_(row[1], h.e0, 0.0);
    if (cols == 3) tmp = cedm::fma_(row[2], h.e1, tmp);
    tmp = tmp + row[0];
    row[0] = row[0] - h.tau * tmp;
    const double tt = h.tau * tmp;
    row[1] = row[1] - tt * h.e0;
    if (cols == 3) row[2] = row[2] - tt * h.e1;
  }
}

// Eigenvalues of the n x n matrix in T (row stride TS, destroyed).  0 = success, 1 = no
// convergence within 40 n iterations, 2 = non-finite eigenvalue.
__device__ __noinline__ int eigvals_dev(double* T, int n, double* wr, double* wi) {
  double scale = 0.0;
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++) {
      const double a = cedm::fabs_(T[i * TS + j]);
      scale = (scale < a) ? a : scale;
    }
  if (scale < DBL_MIN) {
    for (int i = 0; i < n; i++) { wr[i] = 0.0; wi[i] = 0.0; }
    return 0;
  }
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++) T[i * TS + j] = T[i * TS + j] / scale;

  // Hessenberg reduction, column by column
  for (int i = 0; i + 1 < n; i++) {
    const int rem = n - i - 1;
    double tail = 0.0;
    for (int k = 1; k < rem; k++) tail = cedm::fma_(T[(i + 1 + k) * TS + i], T[(i + 1 + k) * TS + i], tail);
    const double c0 = T[(i + 1) * TS + i];
    double tau, beta;
    double ess[MAXT];
    if (tail <= DBL_MIN) {
      tau = 0.0;
      beta = c0;
      for (int k = 0; k + 1 < rem; k++) ess[k] = 0.0;
    } else {
      beta = cedm::sqrt_(cedm::fma_(c0, c0, tail));
      if (c0 >= 0.0) beta = -beta;
      const double den = c0 - beta;
      for (int k = 1; k < rem; k++) ess[k - 1] = T[(i + 1 + k) * TS + i] / den;
      tau = (beta - c0) / beta;
    }
    T[(i + 1) * TS + i] = beta;
    for (int k = 1; k < rem; k++) T[(i + 1 + k) * TS + i] = 0.0;
    if (rem == 1) {
      T[(i + 1) * TS + i + 1] = T[(i + 1) * TS + i + 1] * (1.0 - tau);
      for (int r = 0; r < n; r++) T[r * TS + i + 1] = T[r * TS + i + 1] * (1.0 - tau);
    } else if (tau != 0.0) {
      // left: rows i+1..n-1, columns i+1..n-1
      for (int j = i + 1; j < n; j++) {
        double tmp = 0.0;
        for (int k = 0; k < rem - 1; k++) tmp = cedm::fma_(ess[k], T[(i + 2 + k) * TS + j], tmp);
        tmp = tmp + T[(i + 1) * TS + j];
        T[(i + 1) * TS + j] = T[(i + 1) * TS + j] - tau * tmp;
        for (int k = 0; k < rem - 1; k++) {
          const double te = tau * ess[k];
          T[(i + 2 + k) * TS + j] = T[(i + 2 + k) * TS + j] - te * tmp;
        }
      }
      // right: rows 0..n-1, columns i+1..n-1
      for (int r = 0; r < n; r++) {
        double* row = T + r * TS + i + 1;
        double tmp = 0.0;
        for (int k = 0; k < rem - 1; k++) tmp = cedm::fma_(row[1 + k], ess[k], tmp);
        tmp = tmp + row[0];
        row[0] = row[0] - tau * tmp;
        const double tt = tau * tmp;
        for (int k = 0; k < rem - 1; k++) row[1 + k] = row[1 + k] - tt * ess[k];
      }
    }
  }

  const double eps = DBL_EPSILON;
  int iu = n - 1, iter = 0, total_iter = 0;
  const int max_iters = 40 * n;
  double exshift = 0.0;
  double norm = 0.0;
  for (int j = 0; j < n; j++) {
    const int lim = (n < j + 2) ? n : j + 2;
    for (int i = 0; i < lim; i++) norm = norm + cedm::fabs_(T[i * TS + j]);
  }
  double consider_zero = norm * (eps * eps);
  consider_zero = (consider_zero < DBL_MIN) ? DBL_MIN : consider_zero;
  if (norm != 0.0) {
    while (iu >= 0) {
      int il = iu;
      while (il > 0) {
        double s = cedm::fabs_(T[(il - 1) * TS + il - 1]) + cedm::fabs_(T[il * TS + il]);
        s = s * eps;
        s = (s < consider_zero) ? consider_zero : s;
        if (cedm::fabs_(T[il * TS + il - 1]) <= s) break;
        il--;
      }
      if (il == iu) {
        T[iu * TS + iu] = T[iu * TS + iu] + exshift;
        if (iu > 0) T[iu * TS + iu - 1] = 0.0;
        iu--;
        iter = 0;
      } else if (il == iu - 1) {
        const double p = 0.5 * (T[(iu - 1) * TS + iu - 1] - T[iu * TS + iu]);
        const double q = p * p + T[iu * TS + iu - 1] * T[(iu - 1) * TS + iu];
        T[iu * TS + iu] = T[iu * TS + iu] + exshift;
        T[(iu - 1) * TS + iu - 1] = T[(iu - 1) * TS + iu - 1] + exshift;
        if (q >= 0.0) {
          const double z = cedm::sqrt_(cedm::fabs_(q));
          const double gp = (p >= 0.0) ? p + z : p - z;
          const double gq = T[iu * TS + iu - 1];
          double c, s;
          if (gq == 0.0) {
            c = gp < 0.0 ? -1.0 : 1.0;
            s = 0.0;
          } else if (gp == 0.0) {
            c = 0.0;
            s = gq < 0.0 ? 1.0 : -1.0;
          } else if (cedm::fabs_(gp) > cedm::fabs_(gq)) {
            const double t = gq / gp;
            double u = cedm::sqrt_(1.0 + t * t);
            if (gp < 0.0) u = -u;
            c = 1.0 / u;
            s = -t * c;
          } else {
            const double t = gp / gq;
            double u = cedm::sqrt_(1.0 + t * t);
            if (gq < 0.0) u = -u;
            s = -1.0 / u;
            c = -t * s;
          }
          for (int j = iu - 1; j < n; j++) {
            const double x = T[(iu - 1) * TS + j], y = T[iu * TS + j];
            T[(iu - 1) * TS + j] = c * x - s * y;
            T[iu * TS + j] = s * x + c * y;
          }
          for (int i = 0; i <= iu; i++) {
            const double x = T[i * TS + iu - 1], y = T[i * TS + iu];
            T[i * TS + iu - 1] = c * x - s * y;
            T[i * TS + iu] = s * x + c * y;
          }
          T[iu * TS + iu - 1] = 0.0;
        }
        if (iu > 1) T[(iu - 1) * TS + iu - 2] = 0.0;
        iu -= 2;
        iter = 0;
      } else {
        double sh0 = T[iu * TS + iu];
        double sh1 = T[(iu - 1) * TS + iu - 1];
        double sh2 = T[iu * TS + iu - 1] * T[(iu - 1) * TS + iu];
        if (iter == 10) {
          exshift = exshift + sh0;
          for (int i = 0; i <= iu; i++) T[i * TS + i] = T[i * TS + i] - sh0;
          const double s = cedm::fabs_(T[iu * TS + iu - 1]) + cedm::fabs_(T[(iu - 1) * TS + iu - 2]);
          sh0 = 0.75 * s;
          sh1 = 0.75 * s;
          sh2 = -0.4375 * s * s;
        }
        if (iter == 30) {
          double s = (sh1 - sh0) / 2.0;
          s = s * s + sh2;
          if (s > 0.0) {
            s = cedm::sqrt_(s);
            if (sh1 < sh0) s = -s;
            s = s + (sh1 - sh0) / 2.0;
            s = sh0 - sh2 / s;
            exshift = exshift + s;
            for (int i = 0; i <= iu; i++) T[i * TS + i] = T[i * TS + i] - s;
            sh0 = sh1 = sh2 = 0.964;
          }
        }
        iter++;
        total_iter++;
        if (total_iter > max_iters) return 1;
        double v0 = 0.0, v1 = 0.0, v2 = 0.0;
        int im;
        for (im = iu - 2; im >= il; --im) {
          const double tmm = T[im * TS + im];
          const double r = sh0 - tmm;
          const double s = sh1 - tmm;
          v0 = (r * s - sh2) / T[(im + 1) * TS + im] + T[im * TS + im + 1];
          v1 = T[(im + 1) * TS + im + 1] - tmm - r - s;
          v2 = T[(im + 2) * TS + im + 1];
          if (im == il) break;
          const double lhs = T[im * TS + im - 1] * (cedm::fabs_(v1) + cedm::fabs_(v2));
          const double rhs = v0 * (cedm::fabs_(T[(im - 1) * TS + im - 1]) + cedm::fabs_(tmm) +
                                   cedm::fabs_(T[(im + 1) * TS + im + 1]));
          if (cedm::fabs_(lhs) < eps * rhs) break;
        }
        for (int k = im; k <= iu - 2; k++) {
          const bool first = (k == im);
          double w0, w1, w2;
          if (first) { w0 = v0; w1 = v1; w2 = v2; }
          else { w0 = T[k * TS + k - 1]; w1 = T[(k + 1) * TS + k - 1]; w2 = T[(k + 2) * TS + k - 1]; }
          const HHv h = make_hh(w0, w1, w2, 3);
          if (h.beta != 0.0) {
            if (first && k > il) T[k * TS + k - 1] = -T[k * TS + k - 1];
            else if (!first) T[k * TS + k - 1] = h.beta;
            hh_left_small(T, k, k, n, 3, h);
            const int r1 = ((iu < k + 3) ? iu : k + 3) + 1;
            hh_right_small(T, r1, k, 3, h);
          }
        }
        {
          const HHv h = make_hh(T[(iu - 1) * TS + iu - 2], T[iu * TS + iu - 2], 0.0, 2);
          if (h.beta != 0.0) {
            T[(iu - 1) * TS + iu - 2] = h.beta;
            hh_left_small(T, iu - 1, iu - 1, n, 2, h);
            hh_right_small(T, iu + 1, iu - 1, 2, h);
          }
        }
        for (int i = im + 2; i <= iu; i++) {
          T[i * TS + i - 2] = 0.0;
          if (i > im + 2) T[i * TS + i - 3] = 0.0;
        }
      }
    }
  }
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++) T[i * TS + j] = T[i * TS + j] * scale;
  int i = 0;
  while (i < n) {
    if (i == n - 1 || T[(i + 1) * TS + i] == 0.0) {
      wr[i] = T[i * TS + i];
      wi[i] = 0.0;
      if (!cedm::isfinite_(wr[i])) return 2;
      i++;
    } else {
      const double p = 0.5 * (T[i * TS + i] - T[(i + 1) * TS + i + 1]);
      double t0 = T[(i + 1) * TS + i], t1 = T[i * TS + i + 1];
      const double ap = cedm::fabs_(p), a0 = cedm::fabs_(t0), a1 = cedm::fabs_(t1);
      const double m01 = (a0 < a1) ? a1 : a0;
      const double maxval = (ap < m01) ? m01 : ap;
      t0 = t0 / maxval;
      t1 = t1 / maxval;
      const double p0 = p / maxval;
      const double z = maxval * cedm::sqrt_(cedm::fabs_(p0 * p0 + t0 * t1));
      wr[i] = T[(i + 1) * TS + i + 1] + p;
      wi[i] = z;
      wr[i + 1] = T[(i + 1) * TS + i + 1] + p;
      wi[i + 1] = -z;
      if (!(cedm::isfinite_(wr[i]) && cedm::isfinite_(z))) return 2;
      i += 2;
    }
  }
  return 0;
}

}  // namespace ce

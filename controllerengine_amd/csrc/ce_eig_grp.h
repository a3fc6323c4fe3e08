// Lane-group real-Schur eigenvalue solver for gfx950 (FP64).
//
// One candidate's matrix lives in LDS and is worked on by a group of G lanes (G = 4, 8 or 16, at
// least the matrix order).  The step order is the one of the per-lane solver (Eigen's
// EigenSolver<double>: scaling, Householder Hessenberg reduction, Francis double-shift QR with
// exceptional shifts, read-out; reference call sites src/eigensolve.cu:11 and src/allmargin.cu:72):
//
//  * every scalar that steers control flow (Householder coefficients, shifts, deflation tests,
//    iteration counters) is computed by every lane of the group from the same LDS values, so the
//    lanes of a group stay convergent;
//  * every row / column update is split over lanes by its independent index (a left reflector
//    updates column j in lane j, a right reflector row i in lane i), and each element sees exactly
//    the operations, in the same order, as in the per-lane solver.
//
// Only eigenvalues are computed, so the QR phase leaves out what the sequential solver does for
// the Schur vectors' sake but never reads again: left reflectors stop at column iu (the solver's
// active window ends there and iu never grows back, so columns beyond it are read by no later
// step), and the 2 x 2 standardisation rotates only its own block.  Right reflectors keep every
// row from 0: the window's top il can move back up (its deflation test is relative to a diagonal
// entry the iteration changes), after which rows above the old window are read again.
//
// The result is therefore bit-identical to the sequential solver.  Groups of one wave may diverge
// from each other; a group never waits on another.  Phases that read what other lanes of the group
// wrote are separated by grp_sync(): a wavefront-scope fence, which keeps the compiler from moving
// LDS accesses across it.  The lanes of a group are lanes of one wave, and a wave's memory
// operations are seen by all of its lanes in issue order, so the fence needs no wait instruction
// (the AMDGPU memory model emits none at wavefront scope).  Kernels using it run one wave per
// workgroup.
//
// Sign-flip events.  For D = diag(+-1), every step of the solver maps the iterates of D A D onto D
// times those of A times D, operation for operation (negation commutes with rounding), with two
// exceptions: a Householder vector whose leading entry is zero while its tail is not (the sign of
// beta then does not follow the vector), and the Francis start test |lhs| < eps rhs, whose rhs is
// signed (it carries the sign of v0) and changes sign where D does.  With sign_block = k >= 0
// (D = diag(I_k, -I)), *event is set when either happens; when it is not, the solve of D A D would
// return the same eigenvalues (and the same status) bit for bit.
#pragma once

#include <float.h>

#include "ce_common.h"

namespace ce {

__device__ __forceinline__ void grp_sync() { __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront"); }

// Maxima over the G lanes of a group (lanes gi G .. gi G + G - 1 of the wave, which run the same
// control flow): exact and independent of the combination order.
// Groups of 4, 8 and 16 lanes are quads, half rows and rows of the wave: the partners come by
// DPP (lane ^ 1 and lane ^ 2 by quad permutes, then the mirrored lane within 8 and within 16),
// each step joining two disjoint halves; other sizes pair lanes through the LDS crossbar.
template <int CTRL>
__device__ __forceinline__ int max_i_dpp(int v) {
  const int o = __builtin_amdgcn_mov_dpp(v, CTRL, 0xF, 0xF, true);
  return v > o ? v : o;
}
template <int G>
__device__ __forceinline__ int grp_max_i(int v) {
  if constexpr (G == 4 || G == 8 || G == 16) {
    v = max_i_dpp<0xB1>(v);                       // quad_perm [1,0,3,2]
    v = max_i_dpp<0x4E>(v);                       // quad_perm [2,3,0,1]
    if constexpr (G >= 8) v = max_i_dpp<0x141>(v);  // row_half_mirror
    if constexpr (G == 16) v = max_i_dpp<0x140>(v); // row_mirror
  } else {
#pragma unroll
    for (int d = 1; d < G; d <<= 1) {
      const int o = __shfl_xor(v, d);
      v = v > o ? v : o;
    }
  }
  return v;
}
// v never NaN
template <int G>
__device__ __forceinline__ double grp_max_d(double v) {
#pragma unroll
  for (int d = 1; d < G; d <<= 1) {
    const double o = __shfl_xor(v, d);
    v = (v < o) ? o : v;
  }
  return v;
}

// The value v holds in lane SRC of the caller's group, in every lane of the group (all lanes of
// the group active).  Groups of 16 lanes are the rows of the wave and groups of 4 its quads, so a
// DPP row broadcast / quad permute moves the two halves without LDS; other group sizes permute
// through the LDS crossbar.
template <int G, int SRC>
__device__ __forceinline__ double grp_bcast(double v) {
  static_assert(SRC < G, "source lane outside the group");
  if constexpr (G == 16 || G == 4) {
    constexpr int ctrl = (G == 16) ? (0x150 + SRC) : (SRC * 0x55);  // row_newbcast:SRC / quad_perm [SRC x4]
    const long long b = __builtin_bit_cast(long long, v);
    const int lo = __builtin_amdgcn_mov_dpp((int)b, ctrl, 0xF, 0xF, true);
    const int hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), ctrl, 0xF, 0xF, true);
    return __builtin_bit_cast(double, ((long long)hi << 32) | (unsigned int)lo);
  } else {
    return __shfl(v, (int)(threadIdx.x & ~(unsigned)(G - 1)) + SRC);
  }
}

// IEEE square root.  For x in [2^-700, 2^700] this is the sequence the compiler emits for sqrt on
// gfx950 (hardware reciprocal square root, then Goldschmidt and two Newton corrections), without
// its range scaling and special-value selects, which leave such x unchanged; other x take the full
// sqrt.  The same bits as sqrt for every x (tests/test_gpu_eig_math.py).
__device__ __forceinline__ double grp_sqrt(double x) {
  if (x >= 0x1p-700 && x <= 0x1p700) {
    const double y = __builtin_amdgcn_rsq(x);
    double g = x * y;
    double h = y * 0.5;
    const double r = __builtin_fma(-h, g, 0.5);
    g = __builtin_fma(g, r, g);
    h = __builtin_fma(h, r, h);
    double d = __builtin_fma(-g, g, x);
    g = __builtin_fma(d, h, g);
    d = __builtin_fma(-g, g, x);
    return __builtin_fma(d, h, g);
  }
  return cedm::sqrt_(x);
}

struct GHh {
  double e0, e1, tau, beta;
};

// Householder reflector of (c0, v1[, v2]) (Eigen's makeHouseholderInPlace).  Every lane of the
// group evaluates the norm and beta; the quotients e0 = v1 / den, e1 = v2 / den and
// tau = (beta - c0) / beta = -(den / beta) (beta - c0 rounds to exactly -den) are IEEE divisions on
// lanes 0, 1 and 2 of the group, broadcast to the others: the roundings of evaluating all three in
// every lane, at a third of the divisions.  Without branches: when the tail is negligible (tau = 0,
// beta = c0) the quotients are evaluated and ignored, since callers apply the reflector only when
// tau != 0.
template <int G>
__device__ __forceinline__ GHh grp_make_hh(double c0, double v1, double v2, int len, int l, int* event) {
  static_assert(G >= 3, "the quotients need three lanes");
  GHh h;
  double tail = cedm::fma_(v1, v1, 0.0);
  if (len == 3) tail = cedm::fma_(v2, v2, tail);
  if (event && c0 == 0.0 && tail > DBL_MIN) *event = 1;
  const bool flat = tail <= DBL_MIN;
  double beta = grp_sqrt(cedm::fma_(c0, c0, tail));
  if (c0 >= 0.0) beta = -beta;
  const double den = c0 - beta;
  const double num = (l == 0) ? v1 : ((l == 1) ? v2 : den);
  const double q = num / ((l == 2) ? beta : den);
  h.e0 = grp_bcast<G, 0>(q);
  h.e1 = (len == 3) ? grp_bcast<G, 1>(q) : 0.0;
  h.tau = flat ? 0.0 : -grp_bcast<G, 2>(q);
  h.beta = flat ? c0 : beta;
  return h;
}

// Eigenvalues of the n x n matrix T (LDS, row stride S = G R + 1, destroyed) into WR/WI (LDS, n
// entries), n <= G R.  Lane l of the group owns the rows / columns l, l + G, ..., l + (R-1) G.
// V: LDS scratch of 3 G R doubles (WR and WI may lie inside it: they are written only at the
// read-out).  Returns (in every lane) 0 = success, 1 = no convergence
// within 40 n iterations, 2 = non-finite eigenvalue.
#define CE_LANES(j) _Pragma("unroll") for (int q_ = 0, j = l; q_ < R; q_++, j += G)
template <int G, int R = 1>
__device__ __forceinline__ int eig_grp(double* T, double* V, double* WR, double* WI, int n, int l, int sign_block = -1,
                                       int* event = nullptr) {
  constexpr int N = G * R, S = N + 1;
  // largest |a_ij| (NaN entries never win, as in the sequential fold): column maxima per lane,
  // then the group maximum
  double scale = 0.0;
  CE_LANES(j) if (j < n)
    for (int i = 0; i < n; i++) {
      const double a = cedm::fabs_(T[i * S + j]);
      scale = (scale < a) ? a : scale;
    }
  scale = grp_max_d<G>(scale);
  grp_sync();
  if (scale < DBL_MIN) {
    CE_LANES(j) if (j < n) { WR[j] = 0.0; WI[j] = 0.0; }
    grp_sync();
    return 0;
  }
  CE_LANES(j) if (j < n)
    for (int i = 0; i < n; i++) T[i * S + j] = T[i * S + j] / scale;
  grp_sync();

  // Hessenberg reduction, column by column
  for (int i = 0; i + 1 < n; i++) {
    const int rem = n - i - 1;
    double tail = 0.0;
    for (int k = 1; k < rem; k++) tail = cedm::fma_(T[(i + 1 + k) * S + i], T[(i + 1 + k) * S + i], tail);
    const double c0 = T[(i + 1) * S + i];
    if (event && c0 == 0.0 && tail > DBL_MIN) *event = 1;
    double tau, beta, den = 1.0;
    const bool flat = tail <= DBL_MIN;
    if (flat) {
      tau = 0.0;
      beta = c0;
    } else {
      beta = grp_sqrt(cedm::fma_(c0, c0, tail));
      if (c0 >= 0.0) beta = -beta;
      den = c0 - beta;
      tau = (beta - c0) / beta;
    }
    grp_sync();
    CE_LANES(k) if (k >= 1 && k < rem) {
      V[k - 1] = flat ? 0.0 : T[(i + 1 + k) * S + i] / den;
      T[(i + 1 + k) * S + i] = 0.0;
    }
    if (l == 0) T[(i + 1) * S + i] = beta;
    grp_sync();
    if (rem == 1) {
      CE_LANES(j) if (j < n) {
        if (j == i + 1) T[(i + 1) * S + i + 1] = T[(i + 1) * S + i + 1] * (1.0 - tau);
        T[j * S + i + 1] = T[j * S + i + 1] * (1.0 - tau);
      }
      grp_sync();
    } else if (tau != 0.0) {
      CE_LANES(j) if (j > i && j < n) {  // left: rows i+1..n-1 of column j
        double tmp = 0.0;
        for (int k = 0; k < rem - 1; k++) tmp = cedm::fma_(V[k], T[(i + 2 + k) * S + j], tmp);
        tmp = tmp + T[(i + 1) * S + j];
        T[(i + 1) * S + j] = T[(i + 1) * S + j] - tau * tmp;
        for (int k = 0; k < rem - 1; k++) {
          const double te = tau * V[k];
          T[(i + 2 + k) * S + j] = T[(i + 2 + k) * S + j] - te * tmp;
        }
      }
      grp_sync();
      CE_LANES(r) if (r < n) {  // right: columns i+1..n-1 of row r
        double* row = T + r * S + i + 1;
        double tmp = 0.0;
        for (int k = 0; k < rem - 1; k++) tmp = cedm::fma_(row[1 + k], V[k], tmp);
        tmp = tmp + row[0];
        row[0] = row[0] - tau * tmp;
        const double tt = tau * tmp;
        for (int k = 0; k < rem - 1; k++) row[1 + k] = row[1 + k] - tt * V[k];
      }
      grp_sync();
    }
  }

  const double eps = DBL_EPSILON;
  int iu = n - 1, iter = 0, total_iter = 0;
  const int max_iters = 40 * n;
  double exshift = 0.0;
  double norm = 0.0;
  for (int j = 0; j < n; j++) {
    const int lim = (n < j + 2) ? n : j + 2;
    for (int i = 0; i < lim; i++) norm = norm + cedm::fabs_(T[i * S + j]);
  }
  double consider_zero = norm * (eps * eps);
  consider_zero = (consider_zero < DBL_MIN) ? DBL_MIN : consider_zero;
  if (norm != 0.0) {
    while (iu >= 0) {
      // Deflations run in a loop of their own until the window needs an iteration, so every pass
      // of the outer loop is one Francis iteration in every group that is not finished: groups of
      // a wave that deflate at different passes do not wait through each other's iterations.
      int il;
      for (;;) {
        // findSmallSubdiagEntry: the largest il <= iu whose subdiagonal entry is negligible (0 when
        // none is).  Every lane tests its own rows; the group maximum is the row the downward scan
        // stops at.
        {
          int best = 0;
          CE_LANES(r) if (r >= 1 && r <= iu) {
            double s = cedm::fabs_(T[(r - 1) * S + r - 1]) + cedm::fabs_(T[r * S + r]);
            s = s * eps;
            s = (s < consider_zero) ? consider_zero : s;
            if (cedm::fabs_(T[r * S + r - 1]) <= s && r > best) best = r;
          }
          il = grp_max_i<G>(best);
        }
        if (il == iu) {
          grp_sync();
          if (l == 0) {
            T[iu * S + iu] = T[iu * S + iu] + exshift;
            if (iu > 0) T[iu * S + iu - 1] = 0.0;
          }
          grp_sync();
          iu--;
          iter = 0;
        } else if (il == iu - 1) {
          const double p = 0.5 * (T[(iu - 1) * S + iu - 1] - T[iu * S + iu]);
          const double q = p * p + T[iu * S + iu - 1] * T[(iu - 1) * S + iu];
          const double gq = T[iu * S + iu - 1];
          grp_sync();
          if (l == 0) {
            T[iu * S + iu] = T[iu * S + iu] + exshift;
            T[(iu - 1) * S + iu - 1] = T[(iu - 1) * S + iu - 1] + exshift;
          }
          grp_sync();
          if (q >= 0.0) {
            const double z = grp_sqrt(cedm::fabs_(q));
            const double gp = (p >= 0.0) ? p + z : p - z;
            double c, s;
            if (gq == 0.0) {
              c = gp < 0.0 ? -1.0 : 1.0;
              s = 0.0;
            } else if (gp == 0.0) {
              c = 0.0;
              s = gq < 0.0 ? 1.0 : -1.0;
            } else if (cedm::fabs_(gp) > cedm::fabs_(gq)) {
              const double t = gq / gp;
              double u = grp_sqrt(1.0 + t * t);
              if (gp < 0.0) u = -u;
              c = 1.0 / u;
              s = -t * c;
            } else {
              const double t = gp / gq;
              double u = grp_sqrt(1.0 + t * t);
              if (gq < 0.0) u = -u;
              s = -1.0 / u;
              c = -t * s;
            }
            CE_LANES(j) if (j >= iu - 1 && j <= iu) {
              const double x = T[(iu - 1) * S + j], y = T[iu * S + j];
              T[(iu - 1) * S + j] = c * x - s * y;
              T[iu * S + j] = s * x + c * y;
            }
            grp_sync();
            CE_LANES(r) if (r >= iu - 1 && r <= iu) {
              const double x = T[r * S + iu - 1], y = T[r * S + iu];
              T[r * S + iu - 1] = c * x - s * y;
              T[r * S + iu] = s * x + c * y;
            }
            grp_sync();
            if (l == 0) T[iu * S + iu - 1] = 0.0;
          }
          if (l == 0 && iu > 1) T[(iu - 1) * S + iu - 2] = 0.0;
          grp_sync();
          iu -= 2;
          iter = 0;
        } else {
          break;
        }
        if (iu < 0) break;
      }
      if (iu < 0) break;
      {
        double sh0 = T[iu * S + iu];
        double sh1 = T[(iu - 1) * S + iu - 1];
        double sh2 = T[iu * S + iu - 1] * T[(iu - 1) * S + iu];
        if (iter == 10) {
          exshift = exshift + sh0;
          const double s = cedm::fabs_(T[iu * S + iu - 1]) + cedm::fabs_(T[(iu - 1) * S + iu - 2]);
          grp_sync();
          CE_LANES(j) if (j <= iu) T[j * S + j] = T[j * S + j] - sh0;
          grp_sync();
          sh0 = 0.75 * s;
          sh1 = 0.75 * s;
          sh2 = -0.4375 * s * s;
        }
        if (iter == 30) {
          double s = (sh1 - sh0) / 2.0;
          s = s * s + sh2;
          if (s > 0.0) {
            s = grp_sqrt(s);
            if (sh1 < sh0) s = -s;
            s = s + (sh1 - sh0) / 2.0;
            s = sh0 - sh2 / s;
            exshift = exshift + s;
            grp_sync();
            CE_LANES(j) if (j <= iu) T[j * S + j] = T[j * S + j] - s;
            grp_sync();
            sh0 = sh1 = sh2 = 0.964;
          }
        }
        iter++;
        total_iter++;
        if (total_iter > max_iters) return 1;
        // initFrancisQRStep: the downward scan from iu - 2 stops at the first row whose start test
        // passes, else at il.  Every lane tests its own rows in (il, iu - 2]; the group maximum of
        // the passing rows is where the scan stops.  The scan tests the rows from iu - 2 down to that
        // row, so a sign-flip event counts only there.
        int im;
        {
          int best = il, evm = -1;
          CE_LANES(m) if (m >= il + 1 && m <= iu - 2) {
            const double tmm = T[m * S + m];
            const double r = sh0 - tmm;
            const double s = sh1 - tmm;
            const double a0 = (r * s - sh2) / T[(m + 1) * S + m] + T[m * S + m + 1];
            const double a1 = T[(m + 1) * S + m + 1] - tmm - r - s;
            const double a2 = T[(m + 2) * S + m + 1];
            const double lhs = T[m * S + m - 1] * (cedm::fabs_(a1) + cedm::fabs_(a2));
            const double rhs =
                a0 * (cedm::fabs_(T[(m - 1) * S + m - 1]) + cedm::fabs_(tmm) + cedm::fabs_(T[(m + 1) * S + m + 1]));
            const bool brk = cedm::fabs_(lhs) < eps * rhs;
            if (brk && m > best) best = m;
            if (event && m == sign_block - 1 && brk != (cedm::fabs_(lhs) < eps * -rhs)) evm = m;
          }
          im = grp_max_i<G>(best);
          if (event && grp_max_i<G>(evm) >= im) *event = 1;
        }
        double v0, v1, v2;
        {
          const double tmm = T[im * S + im];
          const double r = sh0 - tmm;
          const double s = sh1 - tmm;
          v0 = (r * s - sh2) / T[(im + 1) * S + im] + T[im * S + im + 1];
          v1 = T[(im + 1) * S + im + 1] - tmm - r - s;
          v2 = T[(im + 2) * S + im + 1];
        }
        double w0 = v0, w1 = v1, w2 = v2;  // the vector of step k: column k - 1 below row k - 1 after the first
        for (int k = im; k <= iu - 2; k++) {
          const bool first = (k == im);
          // Rows k..k+2 of the lane's window columns, loaded before the reflector is formed (nothing
          // writes them before the left update).
          double a0[R], a1[R], a2[R];
          CE_LANES(j) if (j >= k && j <= iu) {
            a0[q_] = T[k * S + j];
            a1[q_] = T[(k + 1) * S + j];
            a2[q_] = T[(k + 2) * S + j];
          }
          const GHh h = grp_make_hh<G>(w0, w1, w2, 3, l, event);
          if (h.beta != 0.0) {
            grp_sync();
            if (l == 0) {
              if (first && k > il) T[k * S + k - 1] = -T[k * S + k - 1];
              else if (!first) T[k * S + k - 1] = h.beta;
            }
            if (h.tau != 0.0) {
              const double te0 = h.tau * h.e0, te1 = h.tau * h.e1;
              CE_LANES(j) if (j >= k && j <= iu) {  // left: rows k..k+2 of column j
                double tmp = cedm::fma_(h.e0, a1[q_], 0.0);
                tmp = cedm::fma_(h.e1, a2[q_], tmp);
                tmp = tmp + a0[q_];
                T[k * S + j] = a0[q_] - h.tau * tmp;
                T[(k + 1) * S + j] = a1[q_] - te0 * tmp;
                T[(k + 2) * S + j] = a2[q_] - te1 * tmp;
              }
              grp_sync();
              const int r1 = ((iu < k + 3) ? iu : k + 3) + 1;
              CE_LANES(r) if (r < r1) {  // right: columns k..k+2 of row r
                double* row = T + r * S + k;
                double tmp = cedm::fma_(row[1], h.e0, 0.0);
                tmp = cedm::fma_(row[2], h.e1, tmp);
                tmp = tmp + row[0];
                row[0] = row[0] - h.tau * tmp;
                const double tt = h.tau * tmp;
                row[1] = row[1] - tt * h.e0;
                row[2] = row[2] - tt * h.e1;
              }
            }
            grp_sync();
          }
          if (k < iu - 2) {
            w0 = T[(k + 1) * S + k];
            w1 = T[(k + 2) * S + k];
            w2 = T[(k + 3) * S + k];
          }
        }
        {
          const GHh h = grp_make_hh<G>(T[(iu - 1) * S + iu - 2], T[iu * S + iu - 2], 0.0, 2, l, event);
          if (h.beta != 0.0) {
            grp_sync();
            if (l == 0) T[(iu - 1) * S + iu - 2] = h.beta;
            if (h.tau != 0.0) {
              CE_LANES(j) if (j >= iu - 1 && j <= iu) {
                double tmp = cedm::fma_(h.e0, T[iu * S + j], 0.0);
                tmp = tmp + T[(iu - 1) * S + j];
                T[(iu - 1) * S + j] = T[(iu - 1) * S + j] - h.tau * tmp;
                T[iu * S + j] = T[iu * S + j] - (h.tau * h.e0) * tmp;
              }
              grp_sync();
              CE_LANES(r) if (r <= iu) {
                double* row = T + r * S + iu - 1;
                double tmp = cedm::fma_(row[1], h.e0, 0.0);
                tmp = tmp + row[0];
                row[0] = row[0] - h.tau * tmp;
                const double tt = h.tau * tmp;
                row[1] = row[1] - tt * h.e0;
              }
            }
            grp_sync();
          }
        }
        CE_LANES(r) if (r >= im + 2 && r <= iu) {
          T[r * S + r - 2] = 0.0;
          if (r > im + 2) T[r * S + r - 3] = 0.0;
        }
        grp_sync();
      }
    }
  }
  grp_sync();
  CE_LANES(j) if (j < n)  // the read-out uses the diagonal, sub- and superdiagonal
    for (int i = (j > 0 ? j - 1 : 0); i <= j + 1 && i < n; i++) T[i * S + j] = T[i * S + j] * scale;
  grp_sync();
  int rc = 0;
  int i = 0;
  while (i < n) {
    if (i == n - 1 || T[(i + 1) * S + i] == 0.0) {
      const double w = T[i * S + i];
      if (l == 0) { WR[i] = w; WI[i] = 0.0; }
      if (!cedm::isfinite_(w)) { rc = 2; break; }
      i++;
    } else {
      const double p = 0.5 * (T[i * S + i] - T[(i + 1) * S + i + 1]);
      double t0 = T[(i + 1) * S + i], t1 = T[i * S + i + 1];
      const double ap = cedm::fabs_(p), a0 = cedm::fabs_(t0), a1 = cedm::fabs_(t1);
      const double m01 = (a0 < a1) ? a1 : a0;
      const double maxval = (ap < m01) ? m01 : ap;
      t0 = t0 / maxval;
      t1 = t1 / maxval;
      const double p0 = p / maxval;
      const double z = maxval * grp_sqrt(cedm::fabs_(p0 * p0 + t0 * t1));
      const double w = T[(i + 1) * S + i + 1] + p;
      if (l == 0) { WR[i] = w; WI[i] = z; WR[i + 1] = w; WI[i + 1] = -z; }
      if (!(cedm::isfinite_(w) && cedm::isfinite_(z))) { rc = 2; break; }
      i += 2;
    }
  }
  grp_sync();
  return rc;
}

#undef CE_LANES

}  // namespace ce

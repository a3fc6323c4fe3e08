// CPU oracle for the ControllerEngine candidate-evaluation chain -- TEST INFRASTRUCTURE ONLY.
//
// This file restates, in plain sequential C++ (double precision, OpenMP over candidates), the
// reference pipeline of ishanbhatia999/ControllerEngine:
//   derive()                src/config.cu:40-78, make_engine_params src/config.cu:93-104
//   decode_instructions     src/main.cu:317-337 (+ NL y-slot rule src/main.cu:377)
//   interpreter             src/main.cu:99-105, 260-291
//   discrete_solve<Luen.>   src/discrete_solves.cu:12-79   (state in double, no hi/lo pairs)
//   discrete_solve<Kalman>  src/discrete_solves.cu:83-185  (ora_simulate_kalman; gains row 2n+2)
//   eigensetup              src/eigensetup.cu:1-57
//   eigensolve              src/eigensolve.cu:1-19         (Eigen EigenSolver<double>, restated)
//   modalsetup              src/modalsetup.cu:1-97
//   modalsolve              src/modalsolve.cu:1-108        (Eigen LDLT, restated, in double)
//   allmargin               src/allmargin.cu:1-200         (Eigen PartialPivLU<complex>, restated, in double)
//   orchestrate_sims        src/main.cu:402-582 (one round)
// and, apart (namespace r32, ora_evaluate_ref32), the same chain in the reference's own precision
// map: float simulation with hi/lo state, float eigensetup, double eigensolve stored as float,
// float modal fit and LDLT, complex<float> margin sweep.
//
// Third-party algorithms restated (Eigen, version unpinned: lib/eigen is an empty submodule):
//   EigenSolver/RealSchur: Householder Hessenberg reduction, Francis double-shift QR with the
//     EISPACK-hqr deflation test and exceptional shifts at iterations 10 and 30, 2x2 block split,
//     eigenvalue read-out top-down with complex pairs as (p+iz, p-iz).
//   LDLT<Lower>: symmetric pivoting on the largest remaining diagonal, zero-pivot rows solved as 0.
//   PartialPivLU + solve: unblocked elimination with row pivoting on the largest modulus, then the
//     column-oriented triangular solves Eigen uses for n <= 16 (panel width 16).
//
// Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg load this library, and
// only as the checker or the reported CPU baseline; the product never calls it.
//
// Arithmetic conventions (the GPU FP64 path follows the same text, so results agree bit for bit):
// dot products, matrix-vector products, Gram sums and polynomials are explicit fma chains in the
// stated index order; every other formula is evaluated operation by operation as written, with no
// contraction (-ffp-contract=off).  Elementary functions come from ce_detmath.h.
#include "../controllerengine_amd/csrc/ce_detmath.h"
#include "../include/ControllerEngine.h"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#if defined(_OPENMP)
#include <omp.h>
#endif

#define ORA_EXPORT extern "C" __attribute__((visibility("default")))

namespace {

constexpr float PI_F = 0x1.921fb6p+1f;                    // src/main.cu:41
constexpr double EPS30 = (double)0x1.a484cp-100f;         // src/modalsetup.cu:20, src/allmargin.cu:84,179
constexpr double EPS6 = (double)0x1.0c6f7ap-20f;          // src/modalsetup.cu:35, modalsolve.cu:158,182,189
constexpr double BIG10 = (double)0x1.2a05f2p+33f;         // src/modalsetup.cu:29-30
constexpr double FLTMAX = (double)0x1.fffffep127f;        // src/allmargin.cu:95
constexpr double PI_D = cedm::PI_HI;
constexpr double TWO_PI_D = 2.0 * cedm::PI_HI;
constexpr double RAD2DEG = 180.0 / cedm::PI_HI;
constexpr double DEG2RAD = cedm::PI_HI / 180.0;

void set_err(char* err, int errlen, const std::string& msg) {
  if (err && errlen > 0) {
    std::snprintf(err, (size_t)errlen, "%s", msg.c_str());
  }
}

// ---------------------------------------------------------------------------------------------
// derive(): src/config.cu:40-78 and make_engine_params src/config.cu:93-104, float32 as written.
struct Derived {
  int n_threads, nt, cuda_w, info_w, info_scoring, info_freqresp, info_allmargin, info_modal;
  int fit_start_id, steady_start_id, fit_count;
  float ts, freq_max, fit_t0, fit_t1, crossover;
};

bool derive(const ce_config& c, Derived& d, std::string& err) {
  d.n_threads = (c.n_cores / 128) * 1536;                                   // config.cu:41
  d.ts = c.duration / (float)c.time;                                        // config.cu:42
  d.freq_max = PI_F / d.ts;                                                 // config.cu:43
  d.nt = c.statespace_dim + c.discrete_dim;                                 // config.cu:44
  d.cuda_w = 2 * c.statespace_dim + 1;                                      // config.cu:51
  d.info_scoring = c.input_size;                                            // config.cu:52-56
  d.info_freqresp = d.info_scoring + 3;
  d.info_allmargin = d.info_freqresp + 4;
  d.info_modal = d.info_allmargin + 5;
  d.info_w = d.info_modal + 3 * d.nt + 1;
  d.fit_t0 = c.fit_start * c.duration;                                      // config.cu:62
  d.fit_t1 = (c.fit_end * (float)c.time - 1.0f) * d.ts;                     // config.cu:63
  const float damping2 = c.damping * c.damping;                             // config.cu:65
  const float damping4 = damping2 * damping2;
  const float wn = 2.0f * PI_F * c.bandwidth /
                   std::sqrt((1.0f - 2.0f * damping2) + std::sqrt(4.0f * damping4 - 4.0f * damping2 + 2.0f));
  d.crossover = c.crossover_coeff * wn * std::sqrt(std::sqrt(4.0f * damping4 + 1.0f) - 2.0f * damping2);
  d.fit_start_id = (int)(c.fit_start * (float)c.time);                      // config.cu:99
  d.steady_start_id = (int)(c.steady_start * (float)c.time);                // config.cu:100
  // modalsetup.cu:75-76: the fit loop visits every int t with (float)t < fit_len
  const float fit_len = (c.fit_end - c.fit_start) * (float)c.time;
  int cnt = 0;
  if (fit_len > 0.0f) {
    cnt = (int)std::ceil(fit_len);
    while (cnt > 0 && !((float)(cnt - 1) < fit_len)) cnt--;
    while ((float)cnt < fit_len) cnt++;
  }
  d.fit_count = cnt;
  if (d.fit_start_id + d.fit_count > c.time) {
    // The reference reads past the trajectory row here (modalsetup.cu:76-84); reject instead.
    err = "fit window [" + std::to_string(d.fit_start_id) + ", " + std::to_string(d.fit_start_id + d.fit_count) +
          ") runs past time = " + std::to_string(c.time);
    return false;
  }
  return true;
}

// ---------------------------------------------------------------------------------------------
// Instruction decode (src/main.cu:317-337) and interpreter (src/main.cu:260-291).
enum { K_CONST = 0, K_STATE = 1, K_PREV = 2 };
struct Ins {
  int op, col, akind, bkind;
  double aval, bval;
};

bool decode(const int* info, int dim, int count, const float* base, int n_params, Ins* out, std::string& err) {
  for (int r = 0; r < count; r++) {
    const int* row = info + (size_t)r * dim;
    int packed = 0, col = 0;
    for (int c = 0; c < dim; c++)
      if (row[c] != 0) { packed = row[c]; col = c; }
    out[r] = Ins{0, 0, K_CONST, K_CONST, 0.0, 0.0};
    if (packed == 0) continue;
    if (packed < 0) { err = "instruction entry " + std::to_string(packed) + " is negative"; return false; }
    const int d1 = (packed / 36) % 36, d2 = packed % 36;
    auto kind = [](int d) { return d == 0 ? K_STATE : (d == 35 ? K_PREV : K_CONST); };
    double v[2] = {0.0, 0.0};
    const int ds[2] = {d1, d2};
    for (int q = 0; q < 2; q++) {
      const int d = ds[q];
      if (d == 0 || d == 35) continue;
      if (d - 1 >= n_params) {
        err = "instruction digit " + std::to_string(d) + " needs BASE_PARAMS[" + std::to_string(d - 1) +
              "] but n_params is " + std::to_string(n_params);
        return false;
      }
      v[q] = (double)base[d - 1];
    }
    out[r] = Ins{packed / 1296, col, kind(d1), kind(d2), v[0], v[1]};
  }
  return true;
}

// interpreter<DIM,COUNT,HAS_SEED>: the chain's running value starts from the seed (U chain, unless
// the seed is NaN) or the state under the first non-SKIP instruction's column.
double interp(const Ins* ins, int count, const double* src, bool has_seed, double seed) {
  double result = 0.0;
  bool started = false;
  for (int r = 0; r < count; r++) {
    const Ins& in = ins[r];
    if (in.op == 0) continue;
    const double pos = src[in.col];
    const double prev = started ? result : ((has_seed && !std::isnan(seed)) ? seed : pos);
    const double a = in.akind == K_CONST ? in.aval : (in.akind == K_STATE ? pos : prev);
    const double b = in.bkind == K_CONST ? in.bval : (in.bkind == K_STATE ? pos : prev);
    switch (in.op) {
      case 1: result = a + b; break;
      case 2: result = a - b; break;
      case 3: result = a * b; break;
      case 4: result = a / b; break;
      case 5: result = cedm::fmin_(cedm::fmax_(prev, a), b); break;
      case 6: result = (double)((0.0 < a) - (a < 0.0)); break;
      default: break;
    }
    started = true;
  }
  return result;
}

// ---------------------------------------------------------------------------------------------
// Eigenvalues of a real n x n matrix (row-major, overwritten): Eigen EigenSolver<double> without
// eigenvectors, restated.  Returns 0 on success, 1 on no convergence, 2 on non-finite eigenvalue.
struct HH {  // Householder vector of a 2- or 3-element segment
  double ess[2];
  double tau, beta;
};

// makeHouseholder: v = [c0; tail].  *event (when not NULL) is set when c0 is zero and the tail is
// not: the one case in which the reflector of a sign-flipped vector is not the sign-flipped reflector.
void make_householder(const double* v, int len, HH& h, int* event) {
  double tail_sq = 0.0;
  for (int k = 1; k < len; k++) tail_sq = cedm::fma_(v[k], v[k], tail_sq);
  const double c0 = v[0];
  if (event && c0 == 0.0 && tail_sq > DBL_MIN) *event = 1;
  if (tail_sq <= DBL_MIN) {
    h.tau = 0.0;
    h.beta = c0;
    h.ess[0] = h.ess[1] = 0.0;
  } else {
    double beta = cedm::sqrt_(cedm::fma_(c0, c0, tail_sq));
    if (c0 >= 0.0) beta = -beta;
    const double den = c0 - beta;
    for (int k = 1; k < len; k++) h.ess[k - 1] = v[k] / den;
    for (int k = len; k < 3; k++) h.ess[k - 1] = 0.0;
    h.tau = (beta - c0) / beta;
    h.beta = beta;
  }
}

// applyHouseholderOnTheLeft on rows r0..r0+rows-1, columns c0..c0+cols-1 of A (stride n)
void hh_left(double* A, int n, int r0, int c0, int rows, int cols, const double* ess, double tau) {
  if (rows == 1) {
    for (int j = 0; j < cols; j++) A[r0 * n + c0 + j] = A[r0 * n + c0 + j] * (1.0 - tau);
    return;
  }
  if (tau == 0.0) return;
  for (int j = 0; j < cols; j++) {
    double tmp = 0.0;
    for (int k = 0; k < rows - 1; k++) tmp = cedm::fma_(ess[k], A[(r0 + 1 + k) * n + c0 + j], tmp);
    tmp = tmp + A[r0 * n + c0 + j];
    A[r0 * n + c0 + j] = A[r0 * n + c0 + j] - tau * tmp;
    for (int k = 0; k < rows - 1; k++) {
      const double te = tau * ess[k];
      A[(r0 + 1 + k) * n + c0 + j] = A[(r0 + 1 + k) * n + c0 + j] - te * tmp;
    }
  }
}

// applyHouseholderOnTheRight on rows r0..r0+rows-1, columns c0..c0+cols-1 of A (stride n)
void hh_right(double* A, int n, int r0, int c0, int rows, int cols, const double* ess, double tau) {
  if (cols == 1) {
    for (int i = 0; i < rows; i++) A[(r0 + i) * n + c0] = A[(r0 + i) * n + c0] * (1.0 - tau);
    return;
  }
  if (tau == 0.0) return;
  for (int i = 0; i < rows; i++) {
    double* row = A + (r0 + i) * n + c0;
    double tmp = 0.0;
    for (int k = 0; k < cols - 1; k++) tmp = cedm::fma_(row[1 + k], ess[k], tmp);
    tmp = tmp + row[0];
    row[0] = row[0] - tau * tmp;
    const double tt = tau * tmp;
    for (int k = 0; k < cols - 1; k++) row[1 + k] = row[1 + k] - tt * ess[k];
  }
}

// sign_block / event (test instrumentation, both optional): with D = diag(I_{sign_block}, -I), the
// solve of D A D takes the same steps and returns the same bits as the solve of A unless *event is
// set: a Householder vector with a zero leading entry and a nonzero tail, or a Francis start test
// at row sign_block - 1 whose outcome depends on the sign of its right-hand side.
int eigvals(double* A, int n, double* wr, double* wi, int sign_block = -1, int* event = nullptr) {
  // RealSchur::compute: scale by the largest |a_ij|
  double scale = 0.0;
  for (int i = 0; i < n * n; i++) scale = std::max(scale, cedm::fabs_(A[i]));
  if (scale < DBL_MIN) {
    for (int i = 0; i < n; i++) { wr[i] = 0.0; wi[i] = 0.0; }
    return 0;
  }
  for (int i = 0; i < n * n; i++) A[i] = A[i] / scale;

  // HessenbergDecomposition: column i, Householder on A[i+1..n-1][i]
  for (int i = 0; i + 1 < n; i++) {
    const int rem = n - i - 1;
    double v[CE_MAX_TOTAL_DIM];
    for (int k = 0; k < rem; k++) v[k] = A[(i + 1 + k) * n + i];
    // makeHouseholderInPlace on a vector of length rem (essential part has rem-1 entries)
    double tail_sq = 0.0;
    for (int k = 1; k < rem; k++) tail_sq = cedm::fma_(v[k], v[k], tail_sq);
    const double c0 = v[0];
    double tau, beta, ess[CE_MAX_TOTAL_DIM];
    if (event && c0 == 0.0 && tail_sq > DBL_MIN) *event = 1;
    if (tail_sq <= DBL_MIN) {
      tau = 0.0;
      beta = c0;
      for (int k = 0; k + 1 < rem; k++) ess[k] = 0.0;
    } else {
      beta = cedm::sqrt_(cedm::fma_(c0, c0, tail_sq));
      if (c0 >= 0.0) beta = -beta;
      const double den = c0 - beta;
      for (int k = 1; k < rem; k++) ess[k - 1] = v[k] / den;
      tau = (beta - c0) / beta;
    }
    A[(i + 1) * n + i] = beta;
    for (int k = 1; k < rem; k++) A[(i + 1 + k) * n + i] = 0.0;  // matrixH() keeps only the Hessenberg part
    // A = H A H' on the trailing block
    if (rem == 1) {
      // applyHouseholderOnTheLeft/Right of a 1-row/1-column block scales by (1 - tau)
      hh_left(A, n, i + 1, i + 1, 1, rem, ess, tau);
      hh_right(A, n, 0, i + 1, n, 1, ess, tau);
    } else {
      hh_left(A, n, i + 1, i + 1, rem, rem, ess, tau);
      hh_right(A, n, 0, i + 1, n, rem, ess, tau);
    }
  }

  // RealSchur::computeFromHessenberg
  double* T = A;
  const double eps = DBL_EPSILON;
  int iu = n - 1, iter = 0, total_iter = 0;
  const int max_iters = 40 * n;
  double exshift = 0.0;
  double norm = 0.0;
  for (int j = 0; j < n; j++) {
    const int lim = std::min(n, j + 2);
    for (int i = 0; i < lim; i++) norm = norm + cedm::fabs_(T[i * n + j]);
  }
  const double consider_zero = std::max(norm * (eps * eps), DBL_MIN);
  bool converged = true;
  if (norm != 0.0) {
    while (iu >= 0) {
      // findSmallSubdiagEntry
      int il = iu;
      while (il > 0) {
        double s = cedm::fabs_(T[(il - 1) * n + il - 1]) + cedm::fabs_(T[il * n + il]);
        s = std::max(s * eps, consider_zero);
        if (cedm::fabs_(T[il * n + il - 1]) <= s) break;
        il--;
      }
      if (il == iu) {  // one root
        T[iu * n + iu] = T[iu * n + iu] + exshift;
        if (iu > 0) T[iu * n + iu - 1] = 0.0;
        iu--;
        iter = 0;
      } else if (il == iu - 1) {  // two roots: splitOffTwoRows
        const double p = 0.5 * (T[(iu - 1) * n + iu - 1] - T[iu * n + iu]);
        const double q = p * p + T[iu * n + iu - 1] * T[(iu - 1) * n + iu];
        T[iu * n + iu] = T[iu * n + iu] + exshift;
        T[(iu - 1) * n + iu - 1] = T[(iu - 1) * n + iu - 1] + exshift;
        if (q >= 0.0) {
          const double z = cedm::sqrt_(cedm::fabs_(q));
          const double gp = (p >= 0.0) ? p + z : p - z;
          const double gq = T[iu * n + iu - 1];
          // JacobiRotation::makeGivens(gp, gq)
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
          // rows iu-1, iu of columns iu-1..n-1, rotation (c, -s): x' = c x - s y, y' = s x + c y
          for (int j = iu - 1; j < n; j++) {
            const double x = T[(iu - 1) * n + j], y = T[iu * n + j];
            T[(iu - 1) * n + j] = c * x - s * y;
            T[iu * n + j] = s * x + c * y;
          }
          // columns iu-1, iu of rows 0..iu, same rotation
          for (int i = 0; i <= iu; i++) {
            const double x = T[i * n + iu - 1], y = T[i * n + iu];
            T[i * n + iu - 1] = c * x - s * y;
            T[i * n + iu] = s * x + c * y;
          }
          T[iu * n + iu - 1] = 0.0;
        }
        if (iu > 1) T[(iu - 1) * n + iu - 2] = 0.0;
        iu -= 2;
        iter = 0;
      } else {
        // computeShift
        double sh[3];
        sh[0] = T[iu * n + iu];
        sh[1] = T[(iu - 1) * n + iu - 1];
        sh[2] = T[iu * n + iu - 1] * T[(iu - 1) * n + iu];
        if (iter == 10) {  // Wilkinson's ad hoc shift
          exshift = exshift + sh[0];
          for (int i = 0; i <= iu; i++) T[i * n + i] = T[i * n + i] - sh[0];
          const double s = cedm::fabs_(T[iu * n + iu - 1]) + cedm::fabs_(T[(iu - 1) * n + iu - 2]);
          sh[0] = 0.75 * s;
          sh[1] = 0.75 * s;
          sh[2] = -0.4375 * s * s;
        }
        if (iter == 30) {  // MATLAB's ad hoc shift
          double s = (sh[1] - sh[0]) / 2.0;
          s = s * s + sh[2];
          if (s > 0.0) {
            s = cedm::sqrt_(s);
            if (sh[1] < sh[0]) s = -s;
            s = s + (sh[1] - sh[0]) / 2.0;
            s = sh[0] - sh[2] / s;
            exshift = exshift + s;
            for (int i = 0; i <= iu; i++) T[i * n + i] = T[i * n + i] - s;
            sh[0] = sh[1] = sh[2] = 0.964;
          }
        }
        iter++;
        total_iter++;
        if (total_iter > max_iters) { converged = false; break; }
        // initFrancisQRStep
        double v[3];
        int im;
        for (im = iu - 2; im >= il; --im) {
          const double tmm = T[im * n + im];
          const double r = sh[0] - tmm;
          const double s = sh[1] - tmm;
          v[0] = (r * s - sh[2]) / T[(im + 1) * n + im] + T[im * n + im + 1];
          v[1] = T[(im + 1) * n + im + 1] - tmm - r - s;
          v[2] = T[(im + 2) * n + im + 1];
          if (im == il) break;
          const double lhs = T[im * n + im - 1] * (cedm::fabs_(v[1]) + cedm::fabs_(v[2]));
          const double rhs = v[0] * (cedm::fabs_(T[(im - 1) * n + im - 1]) + cedm::fabs_(tmm) +
                                     cedm::fabs_(T[(im + 1) * n + im + 1]));
          if (event && im == sign_block - 1 && (cedm::fabs_(lhs) < eps * rhs) != (cedm::fabs_(lhs) < eps * -rhs))
            *event = 1;
          if (cedm::fabs_(lhs) < eps * rhs) break;
        }
        // performFrancisQRStep
        for (int k = im; k <= iu - 2; k++) {
          const bool first = (k == im);
          double w[3];
          if (first) {
            w[0] = v[0]; w[1] = v[1]; w[2] = v[2];
          } else {
            w[0] = T[k * n + k - 1]; w[1] = T[(k + 1) * n + k - 1]; w[2] = T[(k + 2) * n + k - 1];
          }
          HH h;
          make_householder(w, 3, h, event);
          if (h.beta != 0.0) {
            if (first && k > il) T[k * n + k - 1] = -T[k * n + k - 1];
            else if (!first) T[k * n + k - 1] = h.beta;
            hh_left(T, n, k, k, 3, n - k, h.ess, h.tau);
            hh_right(T, n, 0, k, std::min(iu, k + 3) + 1, 3, h.ess, h.tau);
          }
        }
        {
          double w[2] = {T[(iu - 1) * n + iu - 2], T[iu * n + iu - 2]};
          HH h;
          make_householder(w, 2, h, event);
          if (h.beta != 0.0) {
            T[(iu - 1) * n + iu - 2] = h.beta;
            hh_left(T, n, iu - 1, iu - 1, 2, n - iu + 1, h.ess, h.tau);
            hh_right(T, n, 0, iu - 1, iu + 1, 2, h.ess, h.tau);
          }
        }
        for (int i = im + 2; i <= iu; i++) {
          T[i * n + i - 2] = 0.0;
          if (i > im + 2) T[i * n + i - 3] = 0.0;
        }
      }
    }
  }
  if (!converged) return 1;
  for (int i = 0; i < n * n; i++) T[i] = T[i] * scale;
  // EigenSolver::compute: read eigenvalues off the quasi-triangular T
  int i = 0;
  while (i < n) {
    if (i == n - 1 || T[(i + 1) * n + i] == 0.0) {
      wr[i] = T[i * n + i];
      wi[i] = 0.0;
      if (!cedm::isfinite_(wr[i])) return 2;
      i++;
    } else {
      const double p = 0.5 * (T[i * n + i] - T[(i + 1) * n + i + 1]);
      double t0 = T[(i + 1) * n + i], t1 = T[i * n + i + 1];
      const double maxval = std::max(cedm::fabs_(p), std::max(cedm::fabs_(t0), cedm::fabs_(t1)));
      t0 = t0 / maxval;
      t1 = t1 / maxval;
      const double p0 = p / maxval;
      const double z = maxval * cedm::sqrt_(cedm::fabs_(p0 * p0 + t0 * t1));
      wr[i] = T[(i + 1) * n + i + 1] + p;
      wi[i] = z;
      wr[i + 1] = T[(i + 1) * n + i + 1] + p;
      wi[i + 1] = -z;
      if (!(cedm::isfinite_(wr[i]) && cedm::isfinite_(z))) return 2;
      i += 2;
    }
  }
  return 0;
}

// ---------------------------------------------------------------------------------------------
// LDLT<Lower> with symmetric pivoting, restated; solves in place.  M is n x n row-major, full
// symmetric (only the lower triangle is read).  R = double (the plain chain) or float (the
// reference's own modalsolve.cu:73-74, MatF).
inline double fmaR(double a, double b, double c) { return cedm::fma_(a, b, c); }
inline float fmaR(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
template <typename R>
R min_normal();
template <>
double min_normal<double>() { return DBL_MIN; }
template <>
float min_normal<float>() { return FLT_MIN; }

template <typename R>
void ldlt_solve_t(R* M, int n, R* x) {
  int trans[2 * CE_MAX_TOTAL_DIM];
  R temp[2 * CE_MAX_TOTAL_DIM];
  bool all_zero = false;
  for (int k = 0; k < n; k++) {
    int idx = k;
    double best = cedm::fabs_(M[k * n + k]);
    for (int i = k + 1; i < n; i++) {
      const double a = cedm::fabs_(M[i * n + i]);
      if (a > best) { best = a; idx = i; }
    }
    trans[k] = idx;
    if (k != idx) {
      for (int j = 0; j < k; j++) std::swap(M[k * n + j], M[idx * n + j]);
      for (int i = idx + 1; i < n; i++) std::swap(M[i * n + k], M[i * n + idx]);
      std::swap(M[k * n + k], M[idx * n + idx]);
      for (int i = k + 1; i < idx; i++) {
        const double tmp = M[i * n + k];
        M[i * n + k] = M[idx * n + i];
        M[idx * n + i] = tmp;
      }
    }
    const int rs = n - k - 1;
    if (k > 0) {
      for (int j = 0; j < k; j++) temp[j] = M[j * n + j] * M[k * n + j];
      R acc = R(0);
      for (int j = 0; j < k; j++) acc = fmaR(M[k * n + j], temp[j], acc);
      M[k * n + k] = M[k * n + k] - acc;
      for (int i = k + 1; i < n; i++) {
        R a = R(0);
        for (int j = 0; j < k; j++) a = fmaR(M[i * n + j], temp[j], a);
        M[i * n + k] = M[i * n + k] - a;
      }
    }
    const R akk = M[k * n + k];
    const bool valid = cedm::fabs_(akk) > 0.0;
    if (k == 0 && !valid) {
      all_zero = true;
      break;
    }
    if (rs > 0 && valid)
      for (int i = k + 1; i < n; i++) M[i * n + k] = M[i * n + k] / akk;
  }
  if (all_zero) {
    for (int i = 0; i < n; i++) x[i] = R(0);
    return;
  }
  for (int k = 0; k < n; k++) std::swap(x[k], x[trans[k]]);
  for (int i = 0; i < n; i++) {
    R acc = R(0);
    for (int j = 0; j < i; j++) acc = fmaR(M[i * n + j], x[j], acc);
    x[i] = x[i] - acc;
  }
  for (int i = 0; i < n; i++) {
    const R d = M[i * n + i];
    if (cedm::fabs_(d) > min_normal<R>()) x[i] = x[i] / d;
    else x[i] = R(0);
  }
  for (int i = n - 1; i >= 0; i--) {
    R acc = R(0);
    for (int j = i + 1; j < n; j++) acc = fmaR(M[j * n + i], x[j], acc);
    x[i] = x[i] - acc;
  }
  for (int k = n - 1; k >= 0; k--) std::swap(x[k], x[trans[k]]);
}
void ldlt_solve(double* M, int n, double* x) { ldlt_solve_t<double>(M, n, x); }

// ---------------------------------------------------------------------------------------------
// Complex partial-pivot LU solve of (z I - A) x = b with A real n x n (row-major) and b real.
// Returns false when x is not finite (Eigen's allFinite check in allmargin.cu:27,105).
template <typename R>
struct CxT {
  R r, i;
};
using Cx = CxT<double>;
template <typename R>
inline CxT<R> cmul(CxT<R> a, CxT<R> b) { return CxT<R>{a.r * b.r - a.i * b.i, a.r * b.i + a.i * b.r}; }
template <typename R>
inline CxT<R> cdiv(CxT<R> a, CxT<R> b) {
  const R d = b.r * b.r + b.i * b.i;
  return CxT<R>{(a.r * b.r + a.i * b.i) / d, (a.i * b.r - a.r * b.i) / d};
}

template <typename R>
bool shifted_solve_t(const R* A, int n, CxT<R> z, const R* b, CxT<R>* x) {
  // Gaussian elimination with partial pivoting on [zI - A | b]: row k of the pivot search is the
  // first row of largest squared modulus; the right-hand side is eliminated with the same
  // multipliers in the same order as a column-oriented forward substitution, and the back
  // substitution is column-oriented (x_i = v_i / U_ii, then v_j -= U_ji x_i for j < i).
  using C = CxT<R>;
  C M[CE_MAX_STATESPACE_DIM][CE_MAX_STATESPACE_DIM];
  C v[CE_MAX_STATESPACE_DIM];
  for (int i = 0; i < n; i++) {
    for (int j = 0; j < n; j++) M[i][j] = C{(i == j ? z.r : R(0)) - A[i * n + j], (i == j ? z.i : R(0))};
    v[i] = C{b[i], R(0)};
  }
  for (int k = 0; k < n; k++) {
    int p = k;
    R best = M[k][k].r * M[k][k].r + M[k][k].i * M[k][k].i;
    for (int i = k + 1; i < n; i++) {
      const R s = M[i][k].r * M[i][k].r + M[i][k].i * M[i][k].i;
      if (s > best) { best = s; p = i; }
    }
    if (best != R(0)) {
      if (p != k) {
        for (int j = k; j < n; j++) std::swap(M[k][j], M[p][j]);
        std::swap(v[k], v[p]);
      }
      for (int i = k + 1; i < n; i++) M[i][k] = cdiv(M[i][k], M[k][k]);
    }
    for (int i = k + 1; i < n; i++) {
      for (int j = k + 1; j < n; j++) {
        const C pr = cmul(M[i][k], M[k][j]);
        M[i][j] = C{M[i][j].r - pr.r, M[i][j].i - pr.i};
      }
      const C pv = cmul(M[i][k], v[k]);
      v[i] = C{v[i].r - pv.r, v[i].i - pv.i};
    }
  }
  for (int i = n - 1; i >= 0; i--) {
    x[i] = cdiv(v[i], M[i][i]);
    for (int j = 0; j < i; j++) {
      const C pr = cmul(M[j][i], x[i]);
      v[j] = C{v[j].r - pr.r, v[j].i - pr.i};
    }
  }
  bool fin = true;
  for (int i = 0; i < n; i++) fin = fin && cedm::isfinite_((double)x[i].r) && cedm::isfinite_((double)x[i].i);
  return fin;
}
bool shifted_solve(const double* A, int n, Cx z, const double* b, Cx* x) { return shifted_solve_t<double>(A, n, z, b, x); }

// real-weighted dot sum_i w_i x_i of a complex vector
inline Cx rdot(const double* w, const Cx* x, int n) {
  Cx acc{0.0, 0.0};
  for (int i = 0; i < n; i++) acc = Cx{cedm::fma_(w[i], x[i].r, acc.r), cedm::fma_(w[i], x[i].i, acc.i)};
  return acc;
}

// ---------------------------------------------------------------------------------------------
struct Problem {
  ce_config cfg;
  Derived d;
  int ns, nd, nt, nw;
  // observer model (STATESPACE_*) and plant (DISCRETE_*), promoted to double
  double sa[64], sai[64], sb[8], sc[8], sd;
  double da[64], dai[64], dbu[8], dbw[32], dc[8];
  Ins dc_ops[CE_MAX_DC_INSTRUCTIONS], u_ops[CE_MAX_U_INSTRUCTIONS];
  Ins nl_ops[CE_MAX_NONLIN_DIM][CE_MAX_NL_INSTRUCTIONS];
  int nl_count[CE_MAX_NONLIN_DIM];
  std::vector<double> ref;                 // reference step, main.cu:411-419
  std::vector<Cx> z;                       // frequency grid points e^{j theta}
  std::vector<double> theta;               // arg(z)
  // allmargin plant (batch: observer model; serial: true plant with D = 0), allmargin.cu:11-15
  double pa[64], pb[8], pc[8], pd;
  std::vector<Cx> presp;
  std::vector<char> pok;
  double alpha;
  // the same plant response and scaling in complex<float> (the reference's own arithmetic,
  // allmargin.cu:21-41), for the ref-FP32 chain
  std::vector<CxT<float>> presp32;
  std::vector<char> pok32;
  float alpha32 = 0.0f;
  bool alpha32_ok = false;
};

bool setup(const ce_config& c, const ce_plant& pl, Problem& P, std::string& err) {
  P.cfg = c;
  if (!derive(c, P.d, err)) return false;
  P.ns = c.statespace_dim; P.nd = c.discrete_dim; P.nt = P.ns + P.nd; P.nw = c.nonlin_dim;
  const int ns = P.ns, nd = P.nd;
  for (int i = 0; i < ns * ns; i++) { P.sa[i] = pl.ss_a[i]; P.sai[i] = pl.ss_a_i[i]; }
  for (int i = 0; i < ns; i++) { P.sb[i] = pl.ss_b[i]; P.sc[i] = pl.ss_c[i]; }
  P.sd = pl.ss_d;
  for (int i = 0; i < nd * nd; i++) { P.da[i] = pl.d_a[i]; P.dai[i] = pl.d_a_i[i]; }
  for (int i = 0; i < nd; i++) { P.dbu[i] = pl.d_bu[i]; P.dc[i] = pl.d_c[i]; }
  for (int i = 0; i < nd * P.nw; i++) P.dbw[i] = pl.d_bw[i];
  if (!decode(pl.dc_info, nd, c.dc_instructions, pl.base_params, c.n_params, P.dc_ops, err)) return false;
  if (!decode(pl.u_info, ns, c.u_instructions, pl.base_params, c.n_params, P.u_ops, err)) return false;
  for (int i = 0, row = 0; i < P.nw; row += c.nl_instructions[i], i++) {
    if (!decode(pl.nl_info + (size_t)row * (nd + 1), nd + 1, c.nl_instructions[i], pl.base_params, c.n_params,
                P.nl_ops[i], err))
      return false;
    P.nl_count[i] = c.nl_instructions[i];
  }
  // reference vector (main.cu:414): double compare against ref_step_time*time
  P.ref.resize(c.time);
  for (int t = 0; t < c.time; t++) P.ref[t] = ((double)t < c.ref_step_time * (double)c.time) ? 0.0 : (double)c.ref_amplitude;
  // frequency grid (main.cu:423-427), float32 as written: increment = (logf(fmax) - logf(fmin)) /
  // (F - 1), theta_i = expf(logf(fmin) + increment * i) * ts, z_i = (cosf theta_i, sinf theta_i);
  // promoted to double.  theta_w = arg(z_w) taken in float (allmargin.cu:111).
  const int F = c.freq;
  const float inc = (cedm::logf_(P.d.freq_max) - cedm::logf_(c.freq_min)) / (float)(F - 1);
  P.z.resize(F);
  P.theta.resize(F);
  for (int w = 0; w < F; w++) {
    const float th = cedm::expf_(cedm::logf_(c.freq_min) + inc * (float)w) * P.d.ts;
    float s, co;
    cedm::sincosf_(th, s, co);
    P.z[w] = Cx{(double)co, (double)s};
    P.theta[w] = (double)cedm::atan2f_(s, co);
  }
  // margins plant
  if (c.serial) {
    for (int i = 0; i < ns * ns; i++) P.pa[i] = P.da[i];
    for (int i = 0; i < ns; i++) { P.pb[i] = P.dbu[i]; P.pc[i] = P.dc[i]; }
    P.pd = 0.0;
  } else {
    for (int i = 0; i < ns * ns; i++) P.pa[i] = P.sa[i];
    for (int i = 0; i < ns; i++) { P.pb[i] = P.sb[i]; P.pc[i] = P.sc[i]; }
    P.pd = P.sd;
  }
  P.presp.resize(F);
  P.pok.resize(F);
  for (int w = 0; w < F; w++) {
    Cx x[CE_MAX_STATESPACE_DIM];
    const bool ok = shifted_solve(P.pa, ns, P.z[w], P.pb, x);
    P.pok[w] = ok;
    if (ok) {
      Cx g = rdot(P.pc, x, ns);
      P.presp[w] = Cx{g.r + P.pd, g.i};
    } else {
      P.presp[w] = Cx{0.0, 0.0};
    }
  }
  {  // alpha: allmargin.cu:32-41; zc = std::polar(1.0f, crossover * ts) in float
    float s, co;
    cedm::sincosf_(P.d.crossover * P.d.ts, s, co);
    Cx x[CE_MAX_STATESPACE_DIM];
    const bool ok = shifted_solve(P.pa, ns, Cx{(double)co, (double)s}, P.pb, x);
    Cx g = rdot(P.pc, x, ns);
    g.r = g.r + P.pd;
    const double pmag = cedm::hypot(g.r, g.i);
    if (!ok || !(pmag > 0.0) || !cedm::isfinite_(pmag)) {
      char buf[160];
      std::snprintf(buf, sizeof buf, "allmargin: plant has no finite nonzero gain at crossover = %g rad/s",
                    (double)P.d.crossover);
      err = buf;
      return false;
    }
    P.alpha = 1.0 / pmag;
  }
  {  // float forms (allmargin.cu:21-41 as written: complex<float> LU solves, std::abs in float)
    float pa[64], pb[8], pc[8];
    for (int i = 0; i < ns * ns; i++) pa[i] = (float)P.pa[i];
    for (int i = 0; i < ns; i++) { pb[i] = (float)P.pb[i]; pc[i] = (float)P.pc[i]; }
    const float pd = (float)P.pd;
    auto resp32 = [&](CxT<float> z, CxT<float>& g) {
      CxT<float> x[CE_MAX_STATESPACE_DIM];
      const bool ok = shifted_solve_t<float>(pa, ns, z, pb, x);
      CxT<float> acc{0.0f, 0.0f};
      for (int i = 0; i < ns; i++) acc = CxT<float>{acc.r + pc[i] * x[i].r, acc.i + pc[i] * x[i].i};
      g = CxT<float>{acc.r + pd, acc.i};
      return ok;
    };
    P.presp32.resize(F);
    P.pok32.resize(F);
    for (int w = 0; w < F; w++) {
      CxT<float> g;
      P.pok32[w] = resp32(CxT<float>{(float)P.z[w].r, (float)P.z[w].i}, g);
      P.presp32[w] = P.pok32[w] ? g : CxT<float>{0.0f, 0.0f};
    }
    float s, co;
    cedm::sincosf_(P.d.crossover * P.d.ts, s, co);
    CxT<float> g;
    const bool ok = resp32(CxT<float>{co, s}, g);
    const float pmag = (float)cedm::hypot((double)g.r, (double)g.i);
    P.alpha32_ok = ok && pmag > 0.0f && cedm::isfinite_((double)pmag);
    P.alpha32 = P.alpha32_ok ? 1.0f / pmag : 0.0f;
  }
  return true;
}

struct Diag {
  double* y;      // n x time
  double* lina;   // n x nt*nt
  double* eig;    // n x 2nt
  double* ata;    // n x (2nt)^2
  double* atb;    // n x 2nt
};

// One candidate through the six stages.  info: info_w doubles, read and written in place.
void run_candidate(const Problem& P, const float* gains, double* info, const Diag* dg, size_t cand,
                   std::vector<double>& y) {
  const ce_config& c = P.cfg;
  const Derived& D = P.d;
  const int ns = P.ns, nd = P.nd, nt = P.nt, T = c.time;
  double K[8], L[8];
  for (int i = 0; i < ns; i++) { K[i] = gains[i]; L[i] = gains[ns + i]; }
  const double N = gains[2 * ns];

  // ---- discrete_solve<Luenberger>: src/discrete_solves.cu:12-79
  {
    double xh[8] = {0}, xp[8] = {0};
    double nls[CE_MAX_NL_DIM];
    for (int t = 0; t < T; t++) {
      const double yv = interp(P.dc_ops, c.dc_instructions, xp, false, 0.0);
      y[t] = yv;
      double kx = 0.0;
      for (int k = 0; k < ns; k++) kx = cedm::fma_(K[k], xh[k], kx);
      const double u_raw = cedm::fma_(N, P.ref[t], -kx);
      const double u = interp(P.u_ops, c.u_instructions, xh, true, u_raw);
      double cx = 0.0;
      for (int k = 0; k < ns; k++) cx = cedm::fma_(P.sc[k], xh[k], cx);
      const double yhat = cedm::fma_(P.sd, u, cx);
      const double yerr = yv - yhat;
      double dh[8];
      for (int i = 0; i < ns; i++) {
        double acc = cedm::fma_(L[i], yerr, P.sb[i] * u);
        for (int k = 0; k < ns; k++) acc = cedm::fma_(P.sai[i * ns + k], xh[k], acc);
        dh[i] = acc;
      }
      for (int i = 0; i < ns; i++) xh[i] = xh[i] + dh[i];
      for (int i = 0; i < nd; i++) nls[i] = xp[i];
      nls[nd] = yv;
      double W[4];
      for (int q = 0; q < P.nw; q++) W[q] = interp(P.nl_ops[q], P.nl_count[q], nls, false, 0.0);
      double ph[8];
      for (int i = 0; i < nd; i++) {
        double base = P.dbu[i] * u;
        for (int q = 0; q < P.nw; q++) base = cedm::fma_(P.dbw[i * P.nw + q], W[q], base);
        double acc = base;
        for (int k = 0; k < nd; k++) acc = cedm::fma_(P.dai[i * nd + k], xp[k], acc);
        ph[i] = acc;
      }
      for (int i = 0; i < nd; i++) xp[i] = xp[i] + ph[i];
    }
    if (dg && dg->y) std::memcpy(dg->y + cand * (size_t)T, y.data(), sizeof(double) * T);
  }

  // ---- eigensetup: src/eigensetup.cu:1-57
  double lina[CE_MAX_TOTAL_DIM * CE_MAX_TOTAL_DIM];
  for (int i = 0; i < ns; i++) {
    for (int j = 0; j < ns; j++) {
      const double bk = P.sb[i] * K[j];
      const double cdk = P.sc[j] - P.sd * K[j];
      const double lcd = L[i] * cdk;
      lina[i * nt + j] = (P.sa[i * ns + j] - bk) - lcd;
    }
    for (int j = 0; j < nd; j++) lina[i * nt + ns + j] = L[i] * P.dc[j];
  }
  for (int i = 0; i < nd; i++) {
    for (int j = 0; j < ns; j++) lina[(ns + i) * nt + j] = -(P.dbu[i] * K[j]);
    for (int j = 0; j < nd; j++) lina[(ns + i) * nt + ns + j] = P.da[i * nd + j];
  }
  if (dg && dg->lina) std::memcpy(dg->lina + cand * (size_t)nt * nt, lina, sizeof(double) * nt * nt);

  // ---- eigensolve: src/eigensolve.cu:1-19 (failure leaves NaN; the reference reads unset memory)
  double wr[CE_MAX_TOTAL_DIM], wi[CE_MAX_TOTAL_DIM];
  {
    double A[CE_MAX_TOTAL_DIM * CE_MAX_TOTAL_DIM];
    std::memcpy(A, lina, sizeof(double) * nt * nt);
    if (eigvals(A, nt, wr, wi) != 0)
      for (int i = 0; i < nt; i++) { wr[i] = cedm::qnan(); wi[i] = cedm::qnan(); }
  }
  if (dg && dg->eig)
    for (int i = 0; i < nt; i++) { dg->eig[cand * 2 * nt + 2 * i] = wr[i]; dg->eig[cand * 2 * nt + 2 * i + 1] = wi[i]; }

  double* sc = info + D.info_scoring;
  double* fr = info + D.info_freqresp;
  double* am = info + D.info_allmargin;
  double* md = info + D.info_modal;
  const double ts = D.ts;

  // continuous-time modes (modalsetup.cu:8-22, modalsolve.cu:114-128)
  double Re[CE_MAX_TOTAL_DIM], Im[CE_MAX_TOTAL_DIM];
  double max_mag = 0.0;
  for (int i = 0; i < nt; i++) {
    const double mag = cedm::hypot(wr[i], wi[i]);
    max_mag = cedm::fmax_(max_mag, mag);
    Re[i] = cedm::log(cedm::fmax_(mag, EPS30)) / ts;
    Im[i] = cedm::atan2(wi[i], wr[i]) / ts;
  }
  // adjacent-duplicate (conjugate) removal, modalsetup.cu:29-47
  double mr[2 * CE_MAX_TOTAL_DIM];
  int u = 0;
  {
    double pre = BIG10, pim = BIG10;
    for (int i = 0; i < nt; i++) {
      const bool dup = cedm::fabs_(Re[i] - pre) < EPS6 && cedm::fabs_(cedm::fabs_(Im[i]) - cedm::fabs_(pim)) < EPS6;
      pre = Re[i];
      pim = Im[i];
      if (!dup) { mr[2 * u] = Re[i]; mr[2 * u + 1] = cedm::fabs_(Im[i]); u++; }
    }
  }
  const int nc2 = 2 * nt;
  double AtA[4 * CE_MAX_TOTAL_DIM * CE_MAX_TOTAL_DIM];
  double Atb[2 * CE_MAX_TOTAL_DIM];
  bool have_gram = false;

  // ---- modalsetup: src/modalsetup.cu:1-97
  if (max_mag > 1.0) {
    sc[0] = (double)c.unstable_err + cedm::log(max_mag) / ts;
    sc[1] = sc[0];
  } else {
    double shift[CE_MAX_TOTAL_DIM];
    for (int e = 0; e < u; e++) shift[e] = cedm::fmax_(mr[2 * e] * (double)D.fit_t0, mr[2 * e] * (double)D.fit_t1);
    double yss = 0.0;
    for (int t = D.steady_start_id; t < T; t++) yss = yss + y[t];
    yss = yss / (double)(T - D.steady_start_id);
    double peak = 0.0;
    for (int t = 0; t < T; t++) peak = cedm::fmax_(peak, y[t] * yss);
    fr[2] = peak / cedm::fmax_(yss * yss, EPS6) - 1.0;
    fr[3] = yss;
    for (int i = 0; i < nc2 * nc2; i++) AtA[i] = 0.0;
    for (int i = 0; i < nc2; i++) Atb[i] = 0.0;
    const int nc = 2 * u;
    double cf[2 * CE_MAX_TOTAL_DIM];
    for (int t = 0; t < D.fit_count; t++) {
      const double tval = (double)(D.fit_start_id + t) * ts;
      for (int e = 0; e < u; e++) {
        const double decay = cedm::exp(cedm::fma_(mr[2 * e], tval, -shift[e]));
        double s, co;
        cedm::sincos(mr[2 * e + 1] * tval, s, co);
        cf[2 * e] = decay * co;
        cf[2 * e + 1] = decay * s;
      }
      const double yv = y[D.fit_start_id + t] - yss;
      for (int i = 0; i < nc; i++) {
        for (int j = 0; j <= i; j++) AtA[i * nc2 + j] = cedm::fma_(cf[i], cf[j], AtA[i * nc2 + j]);
        Atb[i] = cedm::fma_(cf[i], yv, Atb[i]);
      }
    }
    have_gram = true;
  }
  if (dg && dg->ata && have_gram) std::memcpy(dg->ata + cand * (size_t)nc2 * nc2, AtA, sizeof(double) * nc2 * nc2);
  if (dg && dg->atb && have_gram) std::memcpy(dg->atb + cand * (size_t)nc2, Atb, sizeof(double) * nc2);

  // ---- modalsolve: src/modalsolve.cu:1-108
  if (sc[0] == 0.0) {
    if (max_mag > 1.0) {
      sc[0] = (double)c.unstable_err + cedm::log(max_mag) / ts;
      sc[1] = sc[0];
    } else {
      double cs[2 * CE_MAX_TOTAL_DIM];
      for (int i = 0; i < nc2; i++) {
        const double s = cedm::sqrt_(AtA[i * nc2 + i]);
        cs[i] = (s > EPS6) ? s : 1.0;
      }
      double S[4 * CE_MAX_TOTAL_DIM * CE_MAX_TOTAL_DIM];
      double xs[2 * CE_MAX_TOTAL_DIM];
      for (int i = 0; i < nc2; i++) {
        const double inv_i = 1.0 / cs[i];
        for (int j = 0; j < nc2; j++) {
          const double inv_j = 1.0 / cs[j];
          const double m = (i >= j) ? AtA[i * nc2 + j] : AtA[j * nc2 + i];
          S[i * nc2 + j] = (inv_i * m) * inv_j;
        }
        xs[i] = Atb[i] / cs[i];
      }
      ldlt_solve(S, nc2, xs);
      double coef[2 * CE_MAX_TOTAL_DIM];
      for (int i = 0; i < nc2; i++) coef[i] = xs[i] / cs[i];
      double esum = 0.0, lw = 0.0, zw = 0.0;
      for (int e = 0; e < u; e++) {
        const int id = 2 * e;
        const double c0 = coef[id], c1 = coef[id + 1];
        const double g00 = AtA[id * nc2 + id], g01 = AtA[(id + 1) * nc2 + id], g11 = AtA[(id + 1) * nc2 + id + 1];
        const double energy = ((c0 * c0) * g00 + ((2.0 * c0) * c1) * g01) + (c1 * c1) * g11;
        const double re_e = mr[id], im_e = mr[id + 1];
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
      const double dmp = c.damping, d2 = dmp * dmp;
      const double inner = ((4.0 * d2) * d2 - 4.0 * d2) + 2.0;
      const double twn = ((2.0 * PI_D) * (double)c.bandwidth) / cedm::sqrt_((1.0 - 2.0 * d2) + cedm::sqrt_(inner));
      const double lr = cedm::log(wn / twn);
      double beh = lr * lr + ((zeta - dmp) * (zeta - dmp) / dmp) / dmp;
      const double os = c.overshoot, sse = c.steady_state_error;
      const double err_os = cedm::fmax_(0.0, fr[2] - os) / os;
      beh = beh + err_os * err_os;
      const double err_ss = cedm::fmax_(0.0, cedm::fabs_(fr[3] / (double)c.ref_amplitude - 1.0) - sse) / sse;
      beh = beh + err_ss * err_ss;
      sc[0] = cedm::isfinite_(beh) ? beh : (double)c.eig_fail_err;
      sc[1] = sc[0];
      fr[0] = wn;
      fr[1] = zeta;
    }
  }

  // ---- allmargin: src/allmargin.cu:43-181
  const double UE = c.unstable_err, EF = c.eig_fail_err;
  if (sc[0] >= UE) {
    sc[1] = sc[0]; sc[2] = sc[0];
    am[0] = am[1] = am[2] = am[3] = am[4] = 0.0;
    return;
  }
  double oa[64];
  for (int i = 0; i < ns; i++)
    for (int j = 0; j < ns; j++) oa[i * ns + j] = lina[i * nt + j];
  bool nominal_stable;
  {
    const int m = 2 * ns;
    double Acl[4 * 64];
    for (int i = 0; i < ns; i++)
      for (int j = 0; j < ns; j++) {
        Acl[i * m + j] = oa[i * ns + j] - P.pd * (L[i] * K[j]);
        Acl[i * m + ns + j] = -(L[i] * P.pc[j]);
        Acl[(ns + i) * m + j] = P.pb[i] * K[j];
        Acl[(ns + i) * m + ns + j] = P.pa[i * ns + j];
      }
    double er[16], ei[16];
    nominal_stable = eigvals(Acl, m, er, ei) == 0;
    if (nominal_stable)
      for (int i = 0; i < m; i++)
        if (cedm::hypot(er[i], ei[i]) >= 1.0) { nominal_stable = false; break; }
  }
  if (!nominal_stable) {
    sc[1] = sc[0];
    sc[2] = UE;
    am[0] = am[1] = am[2] = am[3] = am[4] = 0.0;
    const double total = cedm::log((sc[0] + UE) + EPS30);
    sc[0] = cedm::isfinite_(total) ? total : EF;
    return;
  }
  double PM = FLTMAX, GM = FLTMAX, DM = FLTMAX, B = 1.0;
  bool crossed = false, ok = true;
  double prevMag = 0.0, prevPhase = 0.0, prevRaw = 0.0, prevTheta = 0.0;
  for (int w = 0; w < c.freq; w++) {
    if (!P.pok[w]) { ok = false; break; }
    Cx x[CE_MAX_STATESPACE_DIM];
    if (!shifted_solve(oa, ns, P.z[w], L, x)) { ok = false; break; }
    const Cx G = rdot(K, x, ns);
    const Cx resp = cmul(P.presp[w], G);
    const double m = cedm::hypot(resp.r, resp.i);
    const double raw = cedm::atan2(resp.i, resp.r);
    const double theta_w = P.theta[w];
    {
      const double Ps = P.alpha * cedm::hypot(P.presp[w].r, P.presp[w].i);
      const double Cs = cedm::hypot(G.r, G.i) / P.alpha;
      const double num = cedm::hypot(1.0 + resp.r, resp.i);
      B = cedm::fmin_(B, num / (cedm::sqrt_(1.0 + Ps * Ps) * cedm::sqrt_(1.0 + Cs * Cs)));
    }
    if (w > 0) {
      double delta = raw - prevRaw;
      if (delta > PI_D) delta = delta - TWO_PI_D;
      else if (delta < -PI_D) delta = delta + TWO_PI_D;
      const double ph = prevPhase + delta;
      if ((prevMag - 1.0) * (m - 1.0) <= 0.0 && m != prevMag) {
        const double frac = (1.0 - prevMag) / (m - prevMag);
        const double phase_c = prevPhase + frac * (ph - prevPhase);
        crossed = true;
        double pm_deg = 180.0 + phase_c * RAD2DEG;
        pm_deg = pm_deg - 360.0 * cedm::floor_((pm_deg + 180.0) / 360.0);
        PM = cedm::fmin_(PM, pm_deg);
        const double theta_c = prevTheta + frac * (theta_w - prevTheta);
        const double pm_rad = pm_deg * DEG2RAD;
        if (theta_c > EPS6 && pm_rad >= 0.0) DM = cedm::fmin_(DM, ((pm_rad / theta_c) * ts) * 1000.0);
      }
      {
        const double a = prevPhase + PI_D, b = ph + PI_D;
        const double ka = cedm::floor_(a / TWO_PI_D), kb = cedm::floor_(b / TWO_PI_D);
        if (ka != kb && b != a) {
          const double target = TWO_PI_D * (b > a ? kb : ka);
          const double frac = (target - a) / (b - a);
          const double mag_p = prevMag + frac * (m - prevMag);
          GM = cedm::fmin_(GM, -20.0 * cedm::log10(mag_p));
        }
      }
      prevPhase = ph;
    } else {
      prevPhase = raw;
    }
    prevMag = m;
    prevRaw = raw;
    prevTheta = theta_w;
  }
  if (!ok) {
    sc[1] = sc[0]; sc[2] = EF; sc[0] = EF;
    am[0] = 0.0; am[1] = 0.0; am[2] = 0.0; am[3] = 1.0; am[4] = 0.0;
    return;
  }
  const double tgm = c.target_gm, tpm = c.target_pm, tb = c.target_b;
  const double err_gm = cedm::fmax_(0.0, tgm - GM) / tgm;
  const double err_pm = crossed ? cedm::fmax_(0.0, tpm - PM) / tpm : 0.0;
  const double err_b = cedm::fmax_(0.0, tb - B) / tb;
  const double margin = (err_gm * err_gm + err_pm * err_pm) + err_b * err_b;
  sc[1] = sc[0];
  sc[2] = margin;
  am[0] = GM; am[1] = PM; am[2] = DM; am[3] = 1.0; am[4] = B;
  const double total = cedm::log((sc[0] + margin) + EPS30);
  sc[0] = cedm::isfinite_(total) ? total : EF;
}

// =============================================================================================
// ref-FP32 chain: the reference's own precision map, restated for the drift statement (how far
// the reference's arithmetic lies from the plain FP64 chain) and as the bit-exact check of the
// product's float simulation.
//   discrete_solve<Luen.>  src/discrete_solves.cu:12-79   float, hi/lo state (hladdf), fmaf chains
//   eigensetup             src/eigensetup.cu:1-57         float products and differences
//   eigensolve             src/eigensolve.cu:1-19         double solve of the float matrix, stored float
//   modalsetup             src/modalsetup.cu:1-97         float modes, statistics and Gram sums
//   modalsolve             src/modalsolve.cu:1-108        float scaling and LDLT (MatF)
//   allmargin              src/allmargin.cu:1-200         complex<float> LU sweep; double nominal solve
// The device stages (simulation, eigensetup, modalsetup) use fma where the reference's source
// calls __fmaf_rn or a product feeds a sum of the same expression; the host stages (modalsolve,
// allmargin) evaluate as written.  Float elementary functions are the double ones of ce_detmath.h
// rounded to float, and Eigen's complex reductions are taken in index order, so this is the
// reference's precision map rather than its bits (nvcc's logf / Eigen's packet order differ).
namespace r32 {
constexpr float PIF = 0x1.921fb6p+1f;          // src/main.cu:41
constexpr float EPS30F = 0x1.a484cp-100f;
constexpr float EPS6F = 0x1.0c6f7ap-20f;
constexpr float BIG10F = 0x1.2a05f2p+33f;
constexpr float FLTMAXF = 0x1.fffffep127f;
inline float fmaf_(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
inline float fminf_(float a, float b) { return __builtin_fminf(a, b); }
inline float fmaxf_(float a, float b) { return __builtin_fmaxf(a, b); }
inline float fabsf_(float a) { return __builtin_fabsf(a); }
inline float sqrtf_(float a) { return __builtin_sqrtf(a); }
inline float floorf_(float a) { return __builtin_floorf(a); }
inline float hypotf_(float a, float b) { return (float)cedm::hypot((double)a, (double)b); }
inline float log10f_(float a) { return (float)cedm::log10((double)a); }
inline bool finite(float a) { return cedm::isfinite_((double)a); }

// the interpreter in float (src/main.cu:260-291)
float interp(const Ins* ins, int count, const float* src, bool has_seed, float seed) {
  float result = 0.0f;
  bool started = false;
  for (int r = 0; r < count; r++) {
    const Ins& in = ins[r];
    if (in.op == 0) continue;
    const float pos = src[in.col];
    const float prev = started ? result : ((has_seed && seed == seed) ? seed : pos);
    const float a = in.akind == K_CONST ? (float)in.aval : (in.akind == K_STATE ? pos : prev);
    const float b = in.bkind == K_CONST ? (float)in.bval : (in.bkind == K_STATE ? pos : prev);
    switch (in.op) {
      case 1: result = a + b; break;
      case 2: result = a - b; break;
      case 3: result = a * b; break;
      case 4: result = a / b; break;
      case 5: result = fminf_(fmaxf_(prev, a), b); break;
      case 6: result = (float)((0.0f < a) - (a < 0.0f)); break;
      default: break;
    }
    started = true;
  }
  return result;
}

// hladdf (src/main.cu:129-155): TwoSum of the hi parts, lo parts plus the error, Fast2Sum
inline void hladd(float& HI, float& LO, float h, float l) {
  const float sh = HI + h;
  const float b0 = sh - HI;
  const float sl = (HI - (sh - b0)) + (h - b0);
  const float v = (LO + l) + sl;
  HI = sh + v;
  LO = v - (HI - sh);
}

// discrete_solve<Luenberger> in float (src/discrete_solves.cu:12-79)
void simulate(const Problem& P, const float* gains, float* y) {
  const ce_config& c = P.cfg;
  const int ns = P.ns, nd = P.nd, nw = P.nw, T = c.time;
  float K[8], L[8], sai[64], sb[8], scv[8], dai[64], dbu[8], dbw[32];
  for (int i = 0; i < ns; i++) { K[i] = gains[i]; L[i] = gains[ns + i]; }
  const float N = gains[2 * ns];
  for (int i = 0; i < ns * ns; i++) sai[i] = (float)P.sai[i];
  for (int i = 0; i < ns; i++) { sb[i] = (float)P.sb[i]; scv[i] = (float)P.sc[i]; }
  const float sd = (float)P.sd;
  for (int i = 0; i < nd * nd; i++) dai[i] = (float)P.dai[i];
  for (int i = 0; i < nd; i++) dbu[i] = (float)P.dbu[i];
  for (int i = 0; i < nd * nw; i++) dbw[i] = (float)P.dbw[i];
  float xh[8] = {0}, xhl[8] = {0}, xp[8] = {0}, xpl[8] = {0};
  for (int t = 0; t < T; t++) {
    const float yv = interp(P.dc_ops, c.dc_instructions, xp, false, 0.0f);
    y[t] = yv;
    float kx = 0.0f;
    for (int k = 0; k < ns; k++) kx = fmaf_(K[k], xh[k], kx);
    for (int k = 0; k < ns; k++) kx = fmaf_(K[k], xhl[k], kx);
    const float u_raw = fmaf_(N, (float)P.ref[t], -kx);
    const float u = interp(P.u_ops, c.u_instructions, xh, true, u_raw);
    float cx = 0.0f;
    for (int k = 0; k < ns; k++) cx = fmaf_(scv[k], xh[k], cx);
    for (int k = 0; k < ns; k++) cx = fmaf_(scv[k], xhl[k], cx);
    const float yerr = yv - fmaf_(sd, u, cx);
    float dh[8], dl[8];
    for (int i = 0; i < ns; i++) {
      float acc = fmaf_(L[i], yerr, sb[i] * u);
      for (int k = 0; k < ns; k++) acc = fmaf_(sai[i * ns + k], xh[k], acc);
      dh[i] = acc;
      float lo = 0.0f;
      for (int k = 0; k < ns; k++) lo = fmaf_(sai[i * ns + k], xhl[k], lo);
      dl[i] = lo;
    }
    for (int i = 0; i < ns; i++) hladd(xh[i], xhl[i], dh[i], dl[i]);
    float nls[CE_MAX_NL_DIM];
    for (int i = 0; i < nd; i++) nls[i] = xp[i];
    nls[nd] = yv;
    float W[4];
    for (int q = 0; q < nw; q++) W[q] = interp(P.nl_ops[q], P.nl_count[q], nls, false, 0.0f);
    float ph[8], pl[8];
    for (int i = 0; i < nd; i++) {
      float base = dbu[i] * u;
      for (int q = 0; q < nw; q++) base = fmaf_(dbw[i * nw + q], W[q], base);
      float acc = base;
      for (int k = 0; k < nd; k++) acc = fmaf_(dai[i * nd + k], xp[k], acc);
      ph[i] = acc;
      float lo = 0.0f;
      for (int k = 0; k < nd; k++) lo = fmaf_(dai[i * nd + k], xpl[k], lo);
      pl[i] = lo;
    }
    for (int i = 0; i < nd; i++) hladd(xp[i], xpl[i], ph[i], pl[i]);
  }
}

// continuous-time modes of the float eigenvalues (modalsetup.cu:8-22, modalsolve.cu:21-33)
float modes(const float* wr, const float* wi, int nt, float ts, float* Re, float* Im) {
  float max_mag = 0.0f;
  for (int i = 0; i < nt; i++) {
    const float mag = hypotf_(wr[i], wi[i]);
    max_mag = fmaxf_(max_mag, mag);
    Re[i] = cedm::logf_(fmaxf_(mag, EPS30F)) / ts;
    Im[i] = cedm::atan2f_(wi[i], wr[i]) / ts;
  }
  return max_mag;
}

// One candidate; info: info_w floats in place; y: time floats (the trajectory).
void candidate(const Problem& P, const float* gains, float* info, float* y) {
  const ce_config& c = P.cfg;
  const Derived& D = P.d;
  const int ns = P.ns, nd = P.nd, nt = P.nt, T = c.time;
  float K[8], L[8];
  for (int i = 0; i < ns; i++) { K[i] = gains[i]; L[i] = gains[ns + i]; }
  simulate(P, gains, y);

  // eigensetup (float): A - B K - L (C - D K), L C_d, -B_u K, A_d
  float lina[CE_MAX_TOTAL_DIM * CE_MAX_TOTAL_DIM];
  for (int i = 0; i < ns; i++) {
    for (int j = 0; j < ns; j++) {
      const float bk = (float)P.sb[i] * K[j];
      const float cdk = (float)P.sc[j] - (float)P.sd * K[j];
      const float lcd = L[i] * cdk;
      lina[i * nt + j] = ((float)P.sa[i * ns + j] - bk) - lcd;
    }
    for (int j = 0; j < nd; j++) lina[i * nt + ns + j] = L[i] * (float)P.dc[j];
  }
  for (int i = 0; i < nd; i++) {
    for (int j = 0; j < ns; j++) lina[(ns + i) * nt + j] = -((float)P.dbu[i] * K[j]);
    for (int j = 0; j < nd; j++) lina[(ns + i) * nt + ns + j] = (float)P.da[i * nd + j];
  }
  // eigensolve: double solve of the float matrix, eigenvalues stored as float
  float wr[CE_MAX_TOTAL_DIM], wi[CE_MAX_TOTAL_DIM];
  {
    double A[CE_MAX_TOTAL_DIM * CE_MAX_TOTAL_DIM], er[CE_MAX_TOTAL_DIM], ei[CE_MAX_TOTAL_DIM];
    for (int i = 0; i < nt * nt; i++) A[i] = lina[i];
    const bool ok = eigvals(A, nt, er, ei) == 0;
    for (int i = 0; i < nt; i++) {
      wr[i] = ok ? (float)er[i] : (float)cedm::qnan();
      wi[i] = ok ? (float)ei[i] : (float)cedm::qnan();
    }
  }
  float* sc = info + D.info_scoring;
  float* fr = info + D.info_freqresp;
  float* am = info + D.info_allmargin;
  float* md = info + D.info_modal;
  const float ts = D.ts;
  float Re[CE_MAX_TOTAL_DIM], Im[CE_MAX_TOTAL_DIM];
  const float max_mag = modes(wr, wi, nt, ts, Re, Im);
  float mr[2 * CE_MAX_TOTAL_DIM];
  int u = 0;
  {
    float pre = BIG10F, pim = BIG10F;
    for (int i = 0; i < nt; i++) {
      const bool dup = fabsf_(Re[i] - pre) < EPS6F && fabsf_(fabsf_(Im[i]) - fabsf_(pim)) < EPS6F;
      pre = Re[i];
      pim = Im[i];
      if (!dup) { mr[2 * u] = Re[i]; mr[2 * u + 1] = fabsf_(Im[i]); u++; }
    }
  }
  const int nc2 = 2 * nt;
  float AtA[4 * CE_MAX_TOTAL_DIM * CE_MAX_TOTAL_DIM] = {0};
  float Atb[2 * CE_MAX_TOTAL_DIM] = {0};

  // modalsetup (float)
  if (max_mag > 1.0f) {
    sc[0] = c.unstable_err + cedm::logf_(max_mag) / ts;
    sc[1] = sc[0];
  } else {
    float shift[CE_MAX_TOTAL_DIM];
    for (int e = 0; e < u; e++) shift[e] = fmaxf_(mr[2 * e] * D.fit_t0, mr[2 * e] * D.fit_t1);
    float yss = 0.0f;
    for (int t = D.steady_start_id; t < T; t++) yss = yss + y[t];
    yss = yss / (float)(T - D.steady_start_id);
    float peak = 0.0f;
    for (int t = 0; t < T; t++) peak = fmaxf_(peak, y[t] * yss);
    fr[2] = peak / fmaxf_(yss * yss, EPS6F) - 1.0f;
    fr[3] = yss;
    const int nc = 2 * u;
    float cf[2 * CE_MAX_TOTAL_DIM];
    for (int t = 0; t < D.fit_count; t++) {
      const float tval = (float)(D.fit_start_id + t) * ts;
      for (int e = 0; e < u; e++) {
        const float decay = cedm::expf_(fmaf_(mr[2 * e], tval, -shift[e]));
        float s, co;
        cedm::sincosf_(mr[2 * e + 1] * tval, s, co);
        cf[2 * e] = decay * co;
        cf[2 * e + 1] = decay * s;
      }
      const float yv = y[D.fit_start_id + t] - yss;
      for (int i = 0; i < nc; i++) {
        for (int j = 0; j <= i; j++) AtA[i * nc2 + j] = fmaf_(cf[i], cf[j], AtA[i * nc2 + j]);
        Atb[i] = fmaf_(cf[i], yv, Atb[i]);
      }
    }
  }

  // modalsolve (float, host: as written)
  if (sc[0] == 0.0f) {
    float cs[2 * CE_MAX_TOTAL_DIM], S[4 * CE_MAX_TOTAL_DIM * CE_MAX_TOTAL_DIM], xs[2 * CE_MAX_TOTAL_DIM];
    for (int i = 0; i < nc2; i++) {
      const float s = sqrtf_(AtA[i * nc2 + i]);
      cs[i] = (s > EPS6F) ? s : 1.0f;
    }
    for (int i = 0; i < nc2; i++) {
      const float inv_i = 1.0f / cs[i];
      for (int j = 0; j < nc2; j++) {
        const float inv_j = 1.0f / cs[j];
        const float m = (i >= j) ? AtA[i * nc2 + j] : AtA[j * nc2 + i];
        S[i * nc2 + j] = (inv_i * m) * inv_j;
      }
      xs[i] = Atb[i] / cs[i];
    }
    ldlt_solve_t<float>(S, nc2, xs);
    float coef[2 * CE_MAX_TOTAL_DIM];
    for (int i = 0; i < nc2; i++) coef[i] = xs[i] / cs[i];
    float esum = 0.0f, lw = 0.0f, zw = 0.0f;
    for (int e = 0; e < u; e++) {
      const int id = 2 * e;
      const float c0 = coef[id], c1 = coef[id + 1];
      const float g00 = AtA[id * nc2 + id], g01 = AtA[(id + 1) * nc2 + id], g11 = AtA[(id + 1) * nc2 + id + 1];
      const float energy = ((c0 * c0) * g00 + ((2.0f * c0) * c1) * g01) + (c1 * c1) * g11;
      const float re_e = mr[id], im_e = mr[id + 1];
      const float wn_e = fmaxf_(sqrtf_(re_e * re_e + im_e * im_e), EPS6F);
      md[id] = wn_e;
      md[id + 1] = -re_e / wn_e;
      md[2 * nt + e] = energy;
      if (im_e < EPS6F) continue;
      esum = esum + energy;
      lw = lw + energy * cedm::logf_(wn_e);
      zw = zw + energy * (-re_e / wn_e);
    }
    md[3 * nt] = (float)u;
    float wn, zeta;
    if (esum <= EPS6F) { wn = EPS6F; zeta = 0.0f; }
    else { wn = cedm::expf_(lw / esum); zeta = zw / esum; }
    const float dmp = c.damping, d2 = dmp * dmp;
    const float twn = ((2.0f * PIF) * c.bandwidth) / sqrtf_((1.0f - 2.0f * d2) + sqrtf_(((4.0f * d2) * d2 - 4.0f * d2) + 2.0f));
    const float lr = cedm::logf_(wn / twn);
    float beh = lr * lr + ((zeta - dmp) * (zeta - dmp) / dmp) / dmp;
    const float err_os = fmaxf_(0.0f, fr[2] - c.overshoot) / c.overshoot;
    beh = beh + err_os * err_os;
    const float err_ss = fmaxf_(0.0f, fabsf_(fr[3] / c.ref_amplitude - 1.0f) - c.steady_state_error) / c.steady_state_error;
    beh = beh + err_ss * err_ss;
    sc[0] = finite(beh) ? beh : c.eig_fail_err;
    sc[1] = sc[0];
    fr[0] = wn;
    fr[1] = zeta;
  }

  // allmargin (complex<float> sweep; double nominal eigensolve of the float closed loop)
  const float UE = c.unstable_err, EF = c.eig_fail_err;
  if (sc[0] >= UE) {
    sc[1] = sc[0]; sc[2] = sc[0];
    am[0] = am[1] = am[2] = am[3] = am[4] = 0.0f;
    return;
  }
  float oa[64], pa[64], pb[8], pc[8];
  for (int i = 0; i < ns; i++)
    for (int j = 0; j < ns; j++) oa[i * ns + j] = lina[i * nt + j];
  for (int i = 0; i < ns * ns; i++) pa[i] = (float)P.pa[i];
  for (int i = 0; i < ns; i++) { pb[i] = (float)P.pb[i]; pc[i] = (float)P.pc[i]; }
  const float pd = (float)P.pd;
  bool nominal_stable;
  {
    const int m = 2 * ns;
    double Acl[4 * 64], er[16], ei[16];
    for (int i = 0; i < ns; i++)
      for (int j = 0; j < ns; j++) {
        Acl[i * m + j] = oa[i * ns + j] - pd * (L[i] * K[j]);
        Acl[i * m + ns + j] = -(L[i] * pc[j]);
        Acl[(ns + i) * m + j] = pb[i] * K[j];
        Acl[(ns + i) * m + ns + j] = pa[i * ns + j];
      }
    nominal_stable = eigvals(Acl, m, er, ei) == 0;
    if (nominal_stable)
      for (int i = 0; i < m; i++)
        if (cedm::hypot(er[i], ei[i]) >= 1.0) { nominal_stable = false; break; }
  }
  if (!nominal_stable) {
    sc[1] = sc[0];
    sc[2] = UE;
    am[0] = am[1] = am[2] = am[3] = am[4] = 0.0f;
    const float total = cedm::logf_((sc[0] + UE) + EPS30F);
    sc[0] = finite(total) ? total : EF;
    return;
  }
  using C = CxT<float>;
  const float alpha = P.alpha32;
  float PM = FLTMAXF, GM = FLTMAXF, DM = FLTMAXF, B = 1.0f;
  bool crossed = false, ok = true;
  float prevMag = 0.0f, prevPhase = 0.0f, prevRaw = 0.0f, prevTheta = 0.0f;
  for (int w = 0; w < c.freq; w++) {
    if (!P.pok32[w]) { ok = false; break; }
    C x[CE_MAX_STATESPACE_DIM];
    if (!shifted_solve_t<float>(oa, ns, C{(float)P.z[w].r, (float)P.z[w].i}, L, x)) { ok = false; break; }
    C G{0.0f, 0.0f};
    for (int i = 0; i < ns; i++) G = C{G.r + K[i] * x[i].r, G.i + K[i] * x[i].i};
    const C pr = P.presp32[w];
    const C resp = cmul(pr, G);
    const float m = hypotf_(resp.r, resp.i);
    const float raw = cedm::atan2f_(resp.i, resp.r);
    const float theta_w = (float)P.theta[w];
    {
      const float Ps = alpha * hypotf_(pr.r, pr.i), Cs = hypotf_(G.r, G.i) / alpha;
      B = fminf_(B, hypotf_(1.0f + resp.r, resp.i) / (sqrtf_(1.0f + Ps * Ps) * sqrtf_(1.0f + Cs * Cs)));
    }
    if (w > 0) {
      float delta = raw - prevRaw;
      if (delta > PIF) delta = delta - 2.0f * PIF;
      else if (delta < -PIF) delta = delta + 2.0f * PIF;
      const float ph = prevPhase + delta;
      if ((prevMag - 1.0f) * (m - 1.0f) <= 0.0f && m != prevMag) {
        const float frac = (1.0f - prevMag) / (m - prevMag);
        const float phase_c = prevPhase + frac * (ph - prevPhase);
        crossed = true;
        float pm_deg = 180.0f + phase_c * (180.0f / PIF);
        pm_deg = pm_deg - 360.0f * floorf_((pm_deg + 180.0f) / 360.0f);
        PM = fminf_(PM, pm_deg);
        const float theta_c = prevTheta + frac * (theta_w - prevTheta);
        const float pm_rad = pm_deg * (PIF / 180.0f);
        if (theta_c > EPS6F && pm_rad >= 0.0f) DM = fminf_(DM, ((pm_rad / theta_c) * ts) * 1000.0f);
      }
      {
        const float a = prevPhase + PIF, b = ph + PIF;
        const float ka = floorf_(a / (2.0f * PIF)), kb = floorf_(b / (2.0f * PIF));
        if (ka != kb && b != a) {
          const float target = (2.0f * PIF) * (b > a ? kb : ka);
          const float frac = (target - a) / (b - a);
          const float mag_p = prevMag + frac * (m - prevMag);
          GM = fminf_(GM, -20.0f * log10f_(mag_p));
        }
      }
      prevPhase = ph;
    } else {
      prevPhase = raw;
    }
    prevMag = m;
    prevRaw = raw;
    prevTheta = theta_w;
  }
  if (!ok) {
    sc[1] = sc[0]; sc[2] = EF; sc[0] = EF;
    am[0] = 0.0f; am[1] = 0.0f; am[2] = 0.0f; am[3] = 1.0f; am[4] = 0.0f;
    return;
  }
  const float err_gm = fmaxf_(0.0f, c.target_gm - GM) / c.target_gm;
  const float err_pm = crossed ? fmaxf_(0.0f, c.target_pm - PM) / c.target_pm : 0.0f;
  const float err_b = fmaxf_(0.0f, c.target_b - B) / c.target_b;
  const float margin = (err_gm * err_gm + err_pm * err_pm) + err_b * err_b;
  sc[1] = sc[0];
  sc[2] = margin;
  am[0] = GM; am[1] = PM; am[2] = DM; am[3] = 1.0f; am[4] = B;
  const float total = cedm::logf_((sc[0] + margin) + EPS30F);
  sc[0] = finite(total) ? total : EF;
}
}  // namespace r32

// discrete_solve<Kalman> (src/discrete_solves.cu:83-185), never launched by the reference, restated in
// double: gains row [K (ns), N, R, Q (ns)] of 2 ns + 2 floats (the reference sizes the row 2 ns + 1,
// src/config.cu:51, and reads one float past it).  Per step: y = DC(xp); u = U(N r - K xh); predict
// xh += (A - I) xh + B u; P = A P A' + diag(Q); gain KG = P C' / (C P C' + R); update
// xh += KG (y - C xh); P = (I - KG C) P (I - KG C)' + R KG KG'; then the plant step of the
// Luenberger simulator.  The reference's covariance update is written in place through restrict
// pointers (ABATf(IKHP, P, IKHP)); this is the product it intends.  Its fast reciprocal
// (__fdividef) is an IEEE division here.  y_out: n x time.
static void kalman_candidate(const Problem& P, const float* gains, double* y) {
  const ce_config& c = P.cfg;
  const int ns = P.ns, nd = P.nd, T = c.time;
  double K[8], Q[8];
  for (int i = 0; i < ns; i++) { K[i] = gains[i]; Q[i] = gains[ns + 2 + i]; }
  const double N = gains[ns], R = gains[ns + 1];
  double xh[8] = {0}, xp[8] = {0}, Pm[64] = {0};
  double nls[CE_MAX_NL_DIM];
  for (int t = 0; t < T; t++) {
    const double yv = interp(P.dc_ops, c.dc_instructions, xp, false, 0.0);
    y[t] = yv;
    double kx = 0.0;
    for (int k = 0; k < ns; k++) kx = cedm::fma_(K[k], xh[k], kx);
    const double u_raw = cedm::fma_(N, P.ref[t], -kx);
    const double u = interp(P.u_ops, c.u_instructions, xh, true, u_raw);
    double dh[8];
    for (int i = 0; i < ns; i++) {
      double acc = P.sb[i] * u;
      for (int k = 0; k < ns; k++) acc = cedm::fma_(P.sai[i * ns + k], xh[k], acc);
      dh[i] = acc;
    }
    for (int i = 0; i < ns; i++) xh[i] = xh[i] + dh[i];
    // P = A P A' + diag(Q)
    double C0[64], Pn[64];
    for (int i = 0; i < ns; i++)
      for (int j = 0; j < ns; j++) {
        double a = 0.0;
        for (int k = 0; k < ns; k++) a = cedm::fma_(P.sa[i * ns + k], Pm[k * ns + j], a);
        C0[i * ns + j] = a;
      }
    for (int i = 0; i < ns; i++)
      for (int j = 0; j < ns; j++) {
        double a = 0.0;
        for (int k = 0; k < ns; k++) a = cedm::fma_(C0[i * ns + k], P.sa[j * ns + k], a);
        Pn[i * ns + j] = a;
      }
    for (int i = 0; i < ns; i++) Pn[i * ns + i] = Pn[i * ns + i] + Q[i];
    // KG = P C' / (C P C' + R)
    double pht[8], cp[8];
    for (int i = 0; i < ns; i++) {
      double a = 0.0;
      for (int k = 0; k < ns; k++) a = cedm::fma_(Pn[i * ns + k], P.sc[k], a);
      pht[i] = a;
    }
    for (int j = 0; j < ns; j++) {
      double a = 0.0;
      for (int k = 0; k < ns; k++) a = cedm::fma_(P.sc[k], Pn[k * ns + j], a);
      cp[j] = a;
    }
    double hphr = 0.0;
    for (int k = 0; k < ns; k++) hphr = cedm::fma_(cp[k], P.sc[k], hphr);
    hphr = 1.0 / (hphr + R);
    double kg[8];
    for (int i = 0; i < ns; i++) kg[i] = pht[i] * hphr;
    // xh += KG (y - C xh)
    double dz = -yv;
    for (int k = 0; k < ns; k++) dz = cedm::fma_(P.sc[k], xh[k], dz);
    dz = -dz;
    for (int i = 0; i < ns; i++) xh[i] = xh[i] + kg[i] * dz;
    // P = (I - KG C) P (I - KG C)' + R KG KG'
    double ikh[64];
    for (int i = 0; i < ns; i++)
      for (int j = 0; j < ns; j++) ikh[i * ns + j] = (i == j ? 1.0 : 0.0) - kg[i] * P.sc[j];
    for (int i = 0; i < ns; i++)
      for (int j = 0; j < ns; j++) {
        double a = 0.0;
        for (int k = 0; k < ns; k++) a = cedm::fma_(ikh[i * ns + k], Pn[k * ns + j], a);
        C0[i * ns + j] = a;
      }
    for (int i = 0; i < ns; i++)
      for (int j = 0; j < ns; j++) {
        double a = 0.0;
        for (int k = 0; k < ns; k++) a = cedm::fma_(C0[i * ns + k], ikh[j * ns + k], a);
        Pm[i * ns + j] = a + (kg[i] * R) * kg[j];
      }
    // plant step, as the Luenberger simulator
    for (int i = 0; i < nd; i++) nls[i] = xp[i];
    nls[nd] = yv;
    double W[4];
    for (int q = 0; q < P.nw; q++) W[q] = interp(P.nl_ops[q], P.nl_count[q], nls, false, 0.0);
    double ph[8];
    for (int i = 0; i < nd; i++) {
      double base = P.dbu[i] * u;
      for (int q = 0; q < P.nw; q++) base = cedm::fma_(P.dbw[i * P.nw + q], W[q], base);
      double acc = base;
      for (int k = 0; k < nd; k++) acc = cedm::fma_(P.dai[i * nd + k], xp[k], acc);
      ph[i] = acc;
    }
    for (int i = 0; i < nd; i++) xp[i] = xp[i] + ph[i];
  }
}

}  // namespace

// =============================================================================================
// Exported entry points (ctypes)

ORA_EXPORT int ora_derive(const ce_config* cfg, ce_derived* out, char* err, int errlen) {
  Derived d;
  std::string e;
  const bool ok = derive(*cfg, d, e);
  out->n_threads = d.n_threads; out->total_dim = d.nt; out->cuda_mx_width = d.cuda_w; out->info_mx_width = d.info_w;
  out->info_scoring = d.info_scoring; out->info_freqresp = d.info_freqresp; out->info_allmargin = d.info_allmargin;
  out->info_modal = d.info_modal; out->fit_start_id = d.fit_start_id; out->steady_start_id = d.steady_start_id;
  out->fit_count = d.fit_count; out->ts = d.ts; out->freq_max = d.freq_max; out->fit_t0 = d.fit_t0; out->fit_t1 = d.fit_t1;
  out->crossover = d.crossover;
  if (!ok) { set_err(err, errlen, e); return 1; }
  return 0;
}

// decoded instruction as 4 ints + 2 doubles per row: op, col, akind, bkind, aval, bval
ORA_EXPORT int ora_decode(const int* info, int dim, int count, const float* base, int n_params, int* iout,
                          double* vout, char* err, int errlen) {
  std::vector<Ins> out(count > 0 ? count : 1);
  std::string e;
  if (!decode(info, dim, count, base, n_params, out.data(), e)) { set_err(err, errlen, e); return 1; }
  for (int r = 0; r < count; r++) {
    iout[4 * r] = out[r].op; iout[4 * r + 1] = out[r].col; iout[4 * r + 2] = out[r].akind; iout[4 * r + 3] = out[r].bkind;
    vout[2 * r] = out[r].aval; vout[2 * r + 1] = out[r].bval;
  }
  return 0;
}

ORA_EXPORT double ora_interp(const int* iops, const double* vops, int count, const double* src, int has_seed, double seed) {
  std::vector<Ins> ins(count > 0 ? count : 1);
  for (int r = 0; r < count; r++)
    ins[r] = Ins{iops[4 * r], iops[4 * r + 1], iops[4 * r + 2], iops[4 * r + 3], vops[2 * r], vops[2 * r + 1]};
  return interp(ins.data(), count, src, has_seed != 0, seed);
}

ORA_EXPORT int ora_eigvals(const double* a, int n, double* wr, double* wi) {
  std::vector<double> A(a, a + (size_t)n * n);
  return eigvals(A.data(), n, wr, wi);
}

// As ora_eigvals, with the sign-flip events of the solve (see eigvals) for D = diag(I_sign_block, -I).
ORA_EXPORT int ora_eigvals_events(const double* a, int n, int sign_block, double* wr, double* wi, int* event) {
  std::vector<double> A(a, a + (size_t)n * n);
  *event = 0;
  return eigvals(A.data(), n, wr, wi, sign_block, event);
}

ORA_EXPORT void ora_ldlt_solve(const double* m, int n, const double* b, double* x) {
  std::vector<double> M(m, m + (size_t)n * n);
  for (int i = 0; i < n; i++) x[i] = b[i];
  ldlt_solve(M.data(), n, x);
}

ORA_EXPORT int ora_shifted_solve(const double* a, int n, double zr, double zi, const double* b, double* xr, double* xi) {
  Cx x[CE_MAX_STATESPACE_DIM];
  const bool ok = shifted_solve(a, n, Cx{zr, zi}, b, x);
  for (int i = 0; i < n; i++) { xr[i] = x[i].r; xi[i] = x[i].i; }
  return ok ? 1 : 0;
}

ORA_EXPORT int ora_detmath(int fn, const double* x, const double* y, double* out, int n) {
  for (int i = 0; i < n; i++) {
    switch (fn) {
      case 0: out[i] = cedm::exp(x[i]); break;
      case 1: out[i] = cedm::log(x[i]); break;
      case 2: out[i] = cedm::sin(x[i]); break;
      case 3: out[i] = cedm::cos(x[i]); break;
      case 4: out[i] = cedm::atan2(x[i], y[i]); break;
      case 5: out[i] = cedm::hypot(x[i], y[i]); break;
      case 6: out[i] = cedm::log10(x[i]); break;
      default: return 1;
    }
  }
  return 0;
}

// Frequency grid (theta_w, z_w) and plant response of a problem: exposed for tests.
ORA_EXPORT int ora_grid(const ce_config* cfg, const ce_plant* plant, double* zr, double* zi, double* theta,
                        double* pr, double* pi, int* pok, double* alpha, char* err, int errlen) {
  Problem P;
  std::string e;
  if (!setup(*cfg, *plant, P, e)) { set_err(err, errlen, e); return 1; }
  for (int w = 0; w < cfg->freq; w++) {
    zr[w] = P.z[w].r; zi[w] = P.z[w].i; theta[w] = P.theta[w];
    pr[w] = P.presp[w].r; pi[w] = P.presp[w].i; pok[w] = P.pok[w];
  }
  *alpha = P.alpha;
  return 0;
}

// Full FP64 chain for n candidates of one round.
//   gains: n x (2ns+1) floats; info_in: n x info_w floats; info_out (float, may be NULL),
//   info64 (double, may be NULL).  Diagnostics (double, may be NULL): y n x time, lina n x nt^2,
//   eig n x 2nt, ata n x (2nt)^2, atb n x 2nt.  threads <= 0 uses the OpenMP default.
ORA_EXPORT int ora_evaluate(const ce_config* cfg, const ce_plant* plant, const float* gains, const float* info_in,
                            float* info_out, double* info64, int n, int threads, double* dy, double* dlina,
                            double* deig, double* data, double* datb, char* err, int errlen) {
  Problem P;
  std::string e;
  if (!setup(*cfg, *plant, P, e)) { set_err(err, errlen, e); return 1; }
  const int iw = P.d.info_w, gw = P.d.cuda_w;
  Diag dg{dy, dlina, deig, data, datb};
  const Diag* dgp = (dy || dlina || deig || data || datb) ? &dg : nullptr;
#if defined(_OPENMP)
  if (threads > 0) omp_set_num_threads(threads);
#endif
#pragma omp parallel
  {
    std::vector<double> y(cfg->time);
    std::vector<double> row(iw);
#pragma omp for schedule(dynamic, 8)
    for (int t = 0; t < n; t++) {
      for (int k = 0; k < iw; k++) row[k] = info_in[(size_t)t * iw + k];
      run_candidate(P, gains + (size_t)t * gw, row.data(), dgp, (size_t)t, y);
      if (info64)
        for (int k = 0; k < iw; k++) info64[(size_t)t * iw + k] = row[k];
      if (info_out)
        for (int k = 0; k < iw; k++) info_out[(size_t)t * iw + k] = (float)row[k];
    }
  }
  return 0;
}

// The float forms of the fit solve and the margin sweep's complex solve (ref-FP32 chain).
ORA_EXPORT void ora_ldlt_solve32(const float* m, int n, const float* b, float* x) {
  float M[4 * CE_MAX_TOTAL_DIM * CE_MAX_TOTAL_DIM];
  std::memcpy(M, m, sizeof(float) * n * n);
  std::memcpy(x, b, sizeof(float) * n);
  ldlt_solve_t<float>(M, n, x);
}

ORA_EXPORT int ora_shifted_solve32(const float* a, int n, float zr, float zi, const float* b, float* xr, float* xi) {
  CxT<float> x[CE_MAX_STATESPACE_DIM];
  const bool ok = shifted_solve_t<float>(a, n, CxT<float>{zr, zi}, b, x);
  for (int i = 0; i < n; i++) { xr[i] = x[i].r; xi[i] = x[i].i; }
  return ok ? 1 : 0;
}

// ref-FP32 chain (the reference's precision map) for n candidates of one round: float INFO rows
// (info_out, n x info_w) and, when y_out is not NULL, the float trajectories (n x time).
ORA_EXPORT int ora_evaluate_ref32(const ce_config* cfg, const ce_plant* plant, const float* gains, const float* info_in,
                                  float* info_out, float* y_out, int n, int threads, char* err, int errlen) {
  Problem P;
  std::string e;
  if (!setup(*cfg, *plant, P, e)) { set_err(err, errlen, e); return 1; }
  if (!P.alpha32_ok) {
    set_err(err, errlen, "allmargin: plant has no finite nonzero float gain at the crossover");
    return 1;
  }
  const int iw = P.d.info_w, gw = P.d.cuda_w;
#if defined(_OPENMP)
  if (threads > 0) omp_set_num_threads(threads);
#endif
#pragma omp parallel
  {
    std::vector<float> y(cfg->time);
#pragma omp for schedule(dynamic, 8)
    for (int t = 0; t < n; t++) {
      float* row = info_out + (size_t)t * iw;
      for (int k = 0; k < iw; k++) row[k] = info_in[(size_t)t * iw + k];
      r32::candidate(P, gains + (size_t)t * gw, row, y.data());
      if (y_out) std::memcpy(y_out + (size_t)t * cfg->time, y.data(), sizeof(float) * cfg->time);
    }
  }
  return 0;
}

ORA_EXPORT int ora_simulate_kalman(const ce_config* cfg, const ce_plant* plant, const float* gains, int n, double* y_out,
                                   char* err, int errlen) {
  Problem P;
  std::string e;
  if (!setup(*cfg, *plant, P, e)) { set_err(err, errlen, e); return 1; }
  const int gw = 2 * P.ns + 2;
#pragma omp parallel for schedule(dynamic, 8)
  for (int t = 0; t < n; t++) kalman_candidate(P, gains + (size_t)t * gw, y_out + (size_t)t * cfg->time);
  return 0;
}

ORA_EXPORT int ora_max_threads(void) {
#if defined(_OPENMP)
  return omp_get_max_threads();
#else
  return 1;
#endif
}

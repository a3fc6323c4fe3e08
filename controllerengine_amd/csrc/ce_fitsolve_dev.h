// Normal-equation solve and behaviour score of the modal fit, reference src/modalsolve.cu:150-203:
// column scaling by sqrt(diag), LDL^T with symmetric pivoting (Eigen LDLT<Lower> restated; zero
// pivots solve to zero), per-mode energy on the unscaled Gram, energy-weighted wn and zeta of the
// oscillatory modes, hinge errors.  Shared by the FP64 fit and the mixed closed-form fit.
#pragma once

#include <float.h>

#include "ce_stage.h"

namespace ce {

__device__ __forceinline__ int tri(int i, int j) { return i * (i + 1) / 2 + j; }

// Group-parallel form: G lanes of one candidate (l = lane in group) work on LDS-resident data.
// Element-wise steps are split over lanes by their independent index; every element sees the
// same operations in the same order as in a sequential solve (the pivot search keeps the
// sequential scan's choice, and the triangular substitutions run whole on every lane, on a
// register copy of the vector), so the result does not depend on G.  The lanes of the group must
// be convergent; phases are separated by a wavefront-scope fence (the group's lanes are lanes of
// one wave: no wait is needed, only the compiler's program order).  NCP: the padded order (a
// compile-time bound on n2).
// G2: packed lower triangle of the 2nt x 2nt Gram (overwritten), atb: right-hand side; dg, sd, cs,
// x, tmp: LDS vectors of 2nt; trans: LDS ints of 2nt.  u: unique modes of candidate c.
// n2: order of the system solved, 2nt, or 2u when the rows and columns of the modes u..nt-1 are
// exactly zero (then they never take part: no pivot is chosen among them before the last nonzero
// one, their eliminations and substitutions leave zeros, and the leading 2u x 2u solve returns the
// same values).
__device__ __forceinline__ void fs_sync() { __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront"); }

// Group reduction to the largest v, the smallest i among equal v (v never NaN): every lane of
// the group ends with the result.  The selection is a total order, so any pairing gives the same
// winner.  For 16-lane groups (the rows of the wave) the partner comes by DPP: lane ^ 1 and
// lane ^ 2 (quad permutes), then the mirrored lane within 8 and within 16 (row_half_mirror,
// row_mirror), each step joining two disjoint halves; other sizes pair lanes through the LDS
// crossbar.
template <int CTRL>
__device__ __forceinline__ void argmax_first_dpp(double& v, int& i) {
  const long long b = __builtin_bit_cast(long long, v);
  const int lo = __builtin_amdgcn_mov_dpp((int)b, CTRL, 0xF, 0xF, true);
  const int hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), CTRL, 0xF, 0xF, true);
  const double ov = __builtin_bit_cast(double, ((long long)hi << 32) | (unsigned int)lo);
  const int oi = __builtin_amdgcn_mov_dpp(i, CTRL, 0xF, 0xF, true);
  if (ov > v || (ov == v && oi < i)) { v = ov; i = oi; }
}
template <int G>
__device__ __forceinline__ void grp_argmax_first(double& v, int& i) {
  if constexpr (G == 16) {
    argmax_first_dpp<0xB1>(v, i);   // quad_perm [1,0,3,2]
    argmax_first_dpp<0x4E>(v, i);   // quad_perm [2,3,0,1]
    argmax_first_dpp<0x141>(v, i);  // row_half_mirror
    argmax_first_dpp<0x140>(v, i);  // row_mirror
  } else {
#pragma unroll
    for (int d = 1; d < G; d <<= 1) {
      const double ov = __shfl_xor(v, d);
      const int oi = __shfl_xor(i, d);
      if (ov > v || (ov == v && oi < i)) { v = ov; i = oi; }
    }
  }
}

// The value v holds in lane 0 of the caller's group, in every lane of the group (DPP row
// broadcast for 16-lane groups, the LDS crossbar otherwise).
template <int G>
__device__ __forceinline__ double grp_lane0(double v) {
  if constexpr (G == 16) {
    const long long b = __builtin_bit_cast(long long, v);
    const int lo = __builtin_amdgcn_mov_dpp((int)b, 0x150, 0xF, 0xF, true);  // row_newbcast:0
    const int hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), 0x150, 0xF, 0xF, true);
    return __builtin_bit_cast(double, ((long long)hi << 32) | (unsigned int)lo);
  } else {
    return __shfl(v, (int)(threadIdx.x & ~(unsigned)(G - 1)));
  }
}

template <int G, int NCP>
__device__ __forceinline__ void solve_and_score_grp(const ProblemDev* __restrict__ P, const StageBufs& B, int c, int u,
                                                    double* row, double* G2, const double* atb, double* dg,
                                                    double* sd, double* cs, double* x, double* tmp, int* trans,
                                                    int l, int n2) {
  static_assert(NCP <= 2 * G, "at most two rows of the system per lane");
  const int nt = P->nt, N2 = n2;
  const int n = B.n;
  double* sc = row + P->info_scoring;
  for (int i = l; i < N2; i += G) {
    dg[i] = G2[tri(i, i)];
    sd[i] = (i + 1 < N2) ? G2[tri(i + 1, i)] : 0.0;
    const double s = cedm::sqrt_(dg[i]);
    cs[i] = (s > EPS6) ? s : 1.0;
  }
  fs_sync();
  // scaled system D^-1 G D^-1 (lower triangle) and right-hand side D^-1 b
  for (int i = l; i < N2; i += G) {
    const double inv_i = 1.0 / cs[i];
    for (int j = 0; j <= i; j++) {
      const double inv_j = 1.0 / cs[j];
      G2[tri(i, j)] = (inv_i * G2[tri(i, j)]) * inv_j;
    }
    x[i] = atb[i] / cs[i];
  }
  fs_sync();
  // LDL^T with symmetric pivoting on the largest remaining diagonal (lower storage)
  bool all_zero = false;
  for (int k = 0; k < N2; k++) {
    // Lane l owns rows k + l and k + l + G of the step (NCP <= 2G: every row from k on).
    const int r0 = k + l, r1 = r0 + G;
    const bool h0 = r0 < N2, h1 = r1 < N2;
    // pivot: the first i > k whose |diag| exceeds every earlier candidate and |G2[k][k]|, else k
    // (the sequential scan with a strict comparison: NaN never wins, and a NaN at k keeps k).
    // Each lane takes the first maximum of its own rows k + 1 + l, k + 1 + l + G, then the group
    // keeps the largest value and, among equal values, the smallest row.  The diagonal entries are
    // loaded together (addresses clamped into the matrix, values past its end dropped).
    int idx = k;
    {
      const int i0 = r0 + 1, i1 = r1 + 1;
      const double bk = cedm::fabs_(G2[tri(k, k)]);
      const double e0 = cedm::fabs_(G2[tri(i0 < N2 ? i0 : k, i0 < N2 ? i0 : k)]);
      const double e1 = cedm::fabs_(G2[tri(i1 < N2 ? i1 : k, i1 < N2 ? i1 : k)]);
      double lbest = -1.0;
      int lidx = N2;
      if (i0 < N2 && e0 > lbest) { lbest = e0; lidx = i0; }
      if (i1 < N2 && e1 > lbest) { lbest = e1; lidx = i1; }
      grp_argmax_first<G>(lbest, lidx);
      if (lidx < N2 && lbest > bk) idx = lidx;
    }
    fs_sync();
    if (l == 0) trans[k] = idx;
    if (k != idx) {
      // symmetric swap of rows / columns k and idx.  The four parts touch disjoint entries (row k
      // and row idx left of k; column k and column idx below idx; the two diagonal entries; column k
      // and row idx between k and idx), each entry on one lane: every load is issued before the
      // first store.
      const int ja = l, jb = l + G;
      const int pa = idx + 1 + l, pb = pa + G;
      const int qa = k + 1 + l, qb = qa + G;
      const bool hja = ja < k, hjb = jb < k, hpa = pa < N2, hpb = pb < N2, hqa = qa < idx, hqb = qb < idx;
      double s1a = 0.0, s2a = 0.0, s1b = 0.0, s2b = 0.0, u1a = 0.0, u2a = 0.0, u1b = 0.0, u2b = 0.0;
      double w1a = 0.0, w2a = 0.0, w1b = 0.0, w2b = 0.0, dk = 0.0, di = 0.0;
      if (hja) { s1a = G2[tri(k, ja)]; s2a = G2[tri(idx, ja)]; }
      if (hjb) { s1b = G2[tri(k, jb)]; s2b = G2[tri(idx, jb)]; }
      if (hpa) { u1a = G2[tri(pa, k)]; u2a = G2[tri(pa, idx)]; }
      if (hpb) { u1b = G2[tri(pb, k)]; u2b = G2[tri(pb, idx)]; }
      if (hqa) { w1a = G2[tri(qa, k)]; w2a = G2[tri(idx, qa)]; }
      if (hqb) { w1b = G2[tri(qb, k)]; w2b = G2[tri(idx, qb)]; }
      if (l == 0) { dk = G2[tri(k, k)]; di = G2[tri(idx, idx)]; }
      if (hja) { G2[tri(k, ja)] = s2a; G2[tri(idx, ja)] = s1a; }
      if (hjb) { G2[tri(k, jb)] = s2b; G2[tri(idx, jb)] = s1b; }
      if (hpa) { G2[tri(pa, k)] = u2a; G2[tri(pa, idx)] = u1a; }
      if (hpb) { G2[tri(pb, k)] = u2b; G2[tri(pb, idx)] = u1b; }
      if (hqa) { G2[tri(qa, k)] = w2a; G2[tri(idx, qa)] = w1a; }
      if (hqb) { G2[tri(qb, k)] = w2b; G2[tri(idx, qb)] = w1b; }
      if (l == 0) { G2[tri(k, k)] = di; G2[tri(idx, idx)] = dk; }
      fs_sync();
    }
    // column k: rows r0 and r1 (row k, the diagonal, on lane 0), each the sequential dot product
    // over j < k from 0.0 then one subtraction; the two rows' chains are interleaved and their
    // operands loaded four terms at a time (a lane without a second row reads row k and drops the
    // result).  The new column stays in registers: the pivot comes from lane 0 by DPP, and the
    // division below the diagonal follows before the one store.
    double c0 = G2[tri(h0 ? r0 : k, k)], c1 = G2[tri(h1 ? r1 : k, k)];
    if (k > 0) {
      for (int j = l; j < k; j += G) tmp[j] = G2[tri(j, j)] * G2[tri(k, j)];
      fs_sync();
      const double* p0 = G2 + tri(h0 ? r0 : k, 0);
      const double* p1 = G2 + tri(h1 ? r1 : k, 0);
      double a0 = 0.0, a1 = 0.0;
      int j = 0;
      for (; j + 4 <= k; j += 4) {
        const double t0 = tmp[j], t1 = tmp[j + 1], t2 = tmp[j + 2], t3 = tmp[j + 3];
        const double x0 = p0[j], x1 = p0[j + 1], x2 = p0[j + 2], x3 = p0[j + 3];
        const double y0 = p1[j], y1 = p1[j + 1], y2 = p1[j + 2], y3 = p1[j + 3];
        a0 = cedm::fma_(x0, t0, a0);
        a1 = cedm::fma_(y0, t0, a1);
        a0 = cedm::fma_(x1, t1, a0);
        a1 = cedm::fma_(y1, t1, a1);
        a0 = cedm::fma_(x2, t2, a0);
        a1 = cedm::fma_(y2, t2, a1);
        a0 = cedm::fma_(x3, t3, a0);
        a1 = cedm::fma_(y3, t3, a1);
      }
      if (j < k) {  // the last one to three terms, loaded together
        const int m = k - j;
        const int j1 = (m > 1) ? j + 1 : j, j2 = (m > 2) ? j + 2 : j;
        const double t0 = tmp[j], t1 = tmp[j1], t2 = tmp[j2];
        const double x0 = p0[j], x1 = p0[j1], x2 = p0[j2];
        const double y0 = p1[j], y1 = p1[j1], y2 = p1[j2];
        a0 = cedm::fma_(x0, t0, a0);
        a1 = cedm::fma_(y0, t0, a1);
        if (m > 1) {
          a0 = cedm::fma_(x1, t1, a0);
          a1 = cedm::fma_(y1, t1, a1);
        }
        if (m > 2) {
          a0 = cedm::fma_(x2, t2, a0);
          a1 = cedm::fma_(y2, t2, a1);
        }
      }
      c0 = c0 - a0;
      c1 = c1 - a1;
    }
    const double akk = grp_lane0<G>(c0);
    const bool valid = cedm::fabs_(akk) > 0.0;
    if (k == 0 && !valid) { all_zero = true; break; }
    if (h0) G2[tri(r0, k)] = (valid && l > 0) ? c0 / akk : c0;
    if (h1) G2[tri(r1, k)] = valid ? c1 / akk : c1;
    fs_sync();
  }
  fs_sync();
  if (all_zero) {
    for (int i = l; i < N2; i += G) x[i] = 0.0 / cs[i];
  } else {
    // Substitutions: P^T L D L^T P x = b, in the sequential solve's order of operations.  The
    // swaps k = 0..N2-1 of the right-hand side leave at position i the entry that tracing i back
    // through them reaches; each lane traces its own positions.
    for (int i = l; i < N2; i += G) {
      int p = i;
      for (int k = N2 - 1; k >= 0; k--) {
        const int t = trans[k];
        p = (p == k) ? t : ((p == t) ? k : p);
      }
      tmp[i] = x[p];
    }
    fs_sync();
    // Every lane of the group runs both triangular solves whole, on a register copy of the vector
    // (loops unrolled to the padded order NCP; rows from N2 on are masked off).  Each row is the
    // sequential dot product, fma over j in increasing order from 0.0, then one subtraction; the
    // lanes only remove the round trips through LDS.  Entries from N2 on hold -0.0 and the Gram
    // rows from N2 on hold +0.0, so a padded term fma(+0, -0, acc) = acc exactly, signed zeros
    // and NaN included.
    double xr[NCP];
#pragma unroll
    for (int i = 0; i < NCP; i++) xr[i] = (i < N2) ? tmp[i] : -0.0;
    // L y = P b (unit lower; row 0 is y0 = b0 - 0.0 = b0)
#pragma unroll
    for (int i = 1; i < NCP; i++) {
      if (i < N2) {
        double acc = 0.0;
#pragma unroll
        for (int j = 0; j < i; j++) acc = cedm::fma_(G2[tri(i, j)], xr[j], acc);
        xr[i] = xr[i] - acc;
      }
    }
    fs_sync();
    // D z = y: one division per entry, on the lane that owns the entry
#pragma unroll
    for (int i = 0; i < NCP; i++)
      if (i % G == l && i < N2) tmp[i] = xr[i];
    fs_sync();
    for (int i = l; i < N2; i += G) {
      const double d = G2[tri(i, i)];
      tmp[i] = (cedm::fabs_(d) > DBL_MIN) ? tmp[i] / d : 0.0;
    }
    fs_sync();
#pragma unroll
    for (int i = 0; i < NCP; i++) xr[i] = (i < N2) ? tmp[i] : -0.0;
    // L^T w = z (row N2 - 1 is w = z - 0.0 = z)
#pragma unroll
    for (int i = NCP - 2; i >= 0; i--) {
      if (i + 1 < N2) {
        double acc = 0.0;
#pragma unroll
        for (int j = i + 1; j < NCP; j++) acc = cedm::fma_(G2[tri(j, i)], xr[j], acc);
        xr[i] = xr[i] - acc;
      }
    }
    fs_sync();
#pragma unroll
    for (int i = 0; i < NCP; i++)
      if (i % G == l && i < N2) tmp[i] = xr[i];
    fs_sync();
    // x = P^T w (the swaps k = N2-1..0), then the column scaling
    for (int i = l; i < N2; i += G) {
      int p = i;
      for (int k = 0; k < N2; k++) {
        const int t = trans[k];
        p = (p == k) ? t : ((p == t) ? k : p);
      }
      x[i] = tmp[p] / cs[i];
    }
  }
  fs_sync();

  // per-mode energy, energy-weighted wn and zeta of the oscillatory modes (src/modalsolve.cu:168-190);
  // per-mode terms on lane e, the sums over modes in mode order on lane 0
  double* md = row + P->info_modal;
  double* fr = row + P->info_freqresp;
  static_assert(G >= MAXT, "one lane per mode");
  double* t_en = tmp;           // energy
  double* t_lw = dg;            // energy * log(wn_e)
  double* t_zw = sd;            // energy * zeta_e
  int* t_osc = trans;           // oscillatory mode: counted in the sums
  const int e = l;
  double energy = 0.0, lwe = 0.0, zwe = 0.0;
  bool osc = false;
  if (e < u) {
    const int id = 2 * e;
    const double c0 = x[id], c1 = x[id + 1];
    const double g00 = dg[id], g01 = sd[id], g11 = dg[id + 1];
    energy = ((c0 * c0) * g00 + ((2.0 * c0) * c1) * g01) + (c1 * c1) * g11;
    const double re_e = B.mode_re[(size_t)e * n + c], im_e = B.mode_im[(size_t)e * n + c];
    const double wn_e = cedm::fmax_(cedm::sqrt_(cedm::fma_(re_e, re_e, im_e * im_e)), EPS6);
    md[id] = wn_e;
    md[id + 1] = -re_e / wn_e;
    md[2 * nt + e] = energy;
    osc = !(im_e < EPS6);
    if (osc) {
      lwe = energy * cedm::log(wn_e);
      zwe = energy * (-re_e / wn_e);
    }
  }
  fs_sync();
  if (e < u) {
    t_en[e] = energy;
    t_lw[e] = lwe;
    t_zw[e] = zwe;
    t_osc[e] = osc ? 1 : 0;
  }
  fs_sync();
  if (l != 0) return;
  double esum = 0.0, lw = 0.0, zw = 0.0;
  for (int q = 0; q < u; q++) {
    if (!t_osc[q]) continue;
    esum = esum + t_en[q];
    lw = lw + t_lw[q];
    zw = zw + t_zw[q];
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

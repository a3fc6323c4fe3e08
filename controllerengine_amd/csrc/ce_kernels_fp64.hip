// FP64 stage kernels for gfx950.
//
// Stage map (reference file -> kernel):
//   eigensetup + eigensolve + modal prelude   src/eigensetup.cu, src/eigensolve.cu, src/modalsetup.cu:8-50
//        -> k_eig_fp64        one candidate per lane; LIN_A built in private memory, real-Schur QR.
//   discrete_solve<Luenberger> + yss/overshoot   src/discrete_solves.cu:12-79, src/modalsetup.cu:52-62
//        -> k_sim_fp64<NS,ND> one candidate per lane; state in VGPRs, plant/observer matrices read as
//                             wave-uniform scalar loads; only the fit window of y is stored, SoA.
//   modalsetup Gram + modalsolve              src/modalsetup.cu:64-96, src/modalsolve.cu
//        -> k_fit_fp64<NB>    G lanes per candidate own 8x4 / 4x4 tiles of the normal equations;
//                             basis values staged in LDS; LDL^T solve and behaviour score in LDS.
//   allmargin sweep                           src/allmargin.cu:61-159
//        -> k_sweep_fp64<NS>  one frequency per lane, one candidate per workgroup; register LU.
//        -> k_scan_fp64       one candidate per lane: phase unwrap, crossings, score (sequential in w).
#include <hip/hip_runtime.h>

#include "ce_linalg_dev.h"
#include "ce_stage.h"

namespace ce {

template <typename Tp>
__device__ __forceinline__ const Tp* tab(const ProblemDev* P, int64_t off) {
  return reinterpret_cast<const Tp*>(reinterpret_cast<const uint8_t*>(P) + off);
}

// ------------------------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_info_init(const float* __restrict__ in, double* __restrict__ out, size_t count) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) out[i] = (double)in[i];
}

__global__ __launch_bounds__(256) void k_info_final(const double* __restrict__ in, float* __restrict__ out, size_t count) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) out[i] = (float)in[i];
}

// Compaction: the candidates whose trajectory is needed (linearly stable ones, or all of them when
// the trajectory dump is requested) get consecutive slots; one atomic per wave.
__global__ __launch_bounds__(64) void k_compact(StageBufs B, int all) {
  const int c = blockIdx.x * 64 + threadIdx.x;
  const bool pred = c < B.n && (all || (B.flags[c] & F_UNSTABLE) == 0);
  const unsigned long long ballot = __ballot(pred);
  const int cnt = __popcll(ballot);
  int base = 0;
  if (threadIdx.x == __ffsll((long long)ballot) - 1 && cnt > 0) base = atomicAdd(B.count, cnt);
  base = __shfl(base, __ffsll((long long)ballot) - 1);
  if (pred) {
    const unsigned long long below = ballot & ((1ull << threadIdx.x) - 1ull);
    B.list[base + __popcll(below)] = c;
  }
}

// ------------------------------------------------------------------------------------------------
// Eigen stage: LIN_A (reference src/eigensetup.cu:16-56), its eigenvalues, the continuous-time
// modes with the adjacent-duplicate rule (src/modalsetup.cu:8-47), the unstable flag
// (src/modalsetup.cu:26) and the nominal closed loop of the margins stage (src/allmargin.cu:61-79).
__global__ __launch_bounds__(64) void k_eig_fp64(const ProblemDev* __restrict__ P, StageBufs B) {
  const int c = blockIdx.x * 64 + threadIdx.x;
  if (c >= B.n) return;
  const int ns = P->ns, nd = P->nd, nt = P->nt;
  const float* g = B.gains + (size_t)c * P->cuda_w;
  double K[MAXS], L[MAXS];
  for (int i = 0; i < ns; i++) { K[i] = g[i]; L[i] = g[ns + i]; }

  double A[MAXT * TS];
  for (int i = 0; i < ns; i++) {
    for (int j = 0; j < ns; j++) {
      const double bk = P->sb[i] * K[j];
      const double cdk = P->sc[j] - P->sd * K[j];
      const double lcd = L[i] * cdk;
      A[i * TS + j] = (P->sa[i * ns + j] - bk) - lcd;
    }
    for (int j = 0; j < nd; j++) A[i * TS + ns + j] = L[i] * P->dc[j];
  }
  for (int i = 0; i < nd; i++) {
    for (int j = 0; j < ns; j++) A[(ns + i) * TS + j] = -(P->dbu[i] * K[j]);
    for (int j = 0; j < nd; j++) A[(ns + i) * TS + ns + j] = P->da[i * nd + j];
  }
  double* oa = B.open_a + (size_t)c * ns * ns;
  for (int i = 0; i < ns; i++)
    for (int j = 0; j < ns; j++) oa[i * ns + j] = A[i * TS + j];
  if (B.diag && B.d_lina) {
    float* dl = B.d_lina + (size_t)c * nt * nt;
    for (int i = 0; i < nt; i++)
      for (int j = 0; j < nt; j++) dl[i * nt + j] = (float)A[i * TS + j];
  }

  double wr[MAXT], wi[MAXT];
  if (eigvals_dev(A, nt, wr, wi) != 0)
    for (int i = 0; i < nt; i++) { wr[i] = cedm::qnan(); wi[i] = cedm::qnan(); }
  if (B.diag && B.d_eig) {
    float* de = B.d_eig + (size_t)c * 2 * nt;
    for (int i = 0; i < nt; i++) { de[2 * i] = (float)wr[i]; de[2 * i + 1] = (float)wi[i]; }
  }

  const double ts = P->ts;
  double max_mag = 0.0;
  double pre = BIG10, pim = BIG10;
  int u = 0;
  const int n = B.n;
  for (int i = 0; i < nt; i++) {
    const double mag = cedm::hypot(wr[i], wi[i]);
    max_mag = cedm::fmax_(max_mag, mag);
    const double re = cedm::log(cedm::fmax_(mag, EPS30)) / ts;
    const double im = cedm::atan2(wi[i], wr[i]) / ts;
    const bool dup = cedm::fabs_(re - pre) < EPS6 && cedm::fabs_(cedm::fabs_(im) - cedm::fabs_(pim)) < EPS6;
    pre = re;
    pim = im;
    if (!dup) {
      B.mode_re[(size_t)u * n + c] = re;
      B.mode_im[(size_t)u * n + c] = cedm::fabs_(im);
      u++;
    }
  }
  B.uniq[c] = u;
  double* row = B.info + (size_t)c * P->info_w;
  double* sc = row + P->info_scoring;
  int flags = 0;
  if (max_mag > 1.0) {
    sc[0] = P->unstable_err + cedm::log(max_mag) / ts;
    sc[1] = sc[0];
    flags |= F_UNSTABLE;
  }
  if ((flags & F_UNSTABLE) || sc[0] >= P->unstable_err) {
    flags |= F_NOMINAL_SKIPPED;
  } else {
    // nominal closed loop [[A1 - D L K^T, -L Pc^T], [Pb K^T, Pa]] (src/allmargin.cu:63-67)
    const int m = 2 * ns;
    for (int i = 0; i < ns; i++)
      for (int j = 0; j < ns; j++) {
        A[i * TS + j] = oa[i * ns + j] - P->pd * (L[i] * K[j]);
        A[i * TS + ns + j] = -(L[i] * P->pc[j]);
        A[(ns + i) * TS + j] = P->pb[i] * K[j];
        A[(ns + i) * TS + ns + j] = P->pa[i * ns + j];
      }
    bool stable = eigvals_dev(A, m, wr, wi) == 0;
    if (stable)
      for (int i = 0; i < m; i++)
        if (cedm::hypot(wr[i], wi[i]) >= 1.0) { stable = false; break; }
    if (!stable) flags |= F_NOMINAL_UNSTABLE;
  }
  B.flags[c] = flags;
}

// ------------------------------------------------------------------------------------------------
// Instruction chain (reference src/main.cu:260-291).  ops/count/col are wave-uniform.
template <int DIM, int MAXCOUNT, bool HAS_SEED>
__device__ __forceinline__ double interp_dev(const Ins* __restrict__ ops, int count, const double (&src)[DIM], double seed) {
  double result = 0.0;
  bool started = false;
#pragma unroll
  for (int r = 0; r < MAXCOUNT; r++) {
    if (r >= count) break;
    const Ins in = ops[r];
    if (in.op == 0) continue;
    double pos = src[0];
#pragma unroll
    for (int k = 1; k < DIM; k++)
      if (in.col == k) pos = src[k];
    double prev;
    if (started) prev = result;
    else if (HAS_SEED) prev = (seed == seed) ? seed : pos;
    else prev = pos;
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

// Closed-loop simulation, reference src/discrete_solves.cu:12-79, in double without hi/lo pairs.
template <int NS, int ND>
__global__ __launch_bounds__(64) void k_sim_fp64(const ProblemDev* __restrict__ P, StageBufs B) {
  const int idx = blockIdx.x * 64 + threadIdx.x;
  if (idx >= *B.count) return;
  const int c = B.list[idx];
  const bool unstable = (B.flags[c] & F_UNSTABLE) != 0;
  const int n = B.n;
  const float* g = B.gains + (size_t)c * P->cuda_w;
  double K[NS], L[NS];
#pragma unroll
  for (int i = 0; i < NS; i++) { K[i] = g[i]; L[i] = g[NS + i]; }
  const double N = g[2 * NS];
  double xh[NS], xp[ND];
#pragma unroll
  for (int i = 0; i < NS; i++) xh[i] = 0.0;
#pragma unroll
  for (int i = 0; i < ND; i++) xp[i] = 0.0;
  const int T = P->T, f0 = P->fit_start_id, f1 = P->fit_start_id + P->fit_count;
  const int s0 = P->steady_start_id, k0 = P->ref_step_id, nw = P->nw;
  const int dc_n = P->dc_n, u_n = P->u_n;
  const double amp = P->ref_amp, sd = P->sd;
  double ysum = 0.0, ymax = -cedm::inf(), ymin = cedm::inf();
  double* ybuf = B.ybuf + idx;
  float* dy = (B.diag && B.d_y) ? B.d_y + (size_t)c * T : nullptr;
  for (int t = 0; t < T; t++) {
    const double y = interp_dev<ND, CE_MAX_DC_INSTRUCTIONS, false>(P->dc_ops, dc_n, xp, 0.0);
    if (dy) dy[t] = (float)y;
    if (t >= f0 && t < f1) ybuf[(size_t)(t - f0) * n] = y;
    if (t >= s0) ysum = ysum + y;
    ymax = cedm::fmax_(ymax, y);
    ymin = cedm::fmin_(ymin, y);
    double kx = 0.0;
#pragma unroll
    for (int k = 0; k < NS; k++) kx = cedm::fma_(K[k], xh[k], kx);
    const double r = (t < k0) ? 0.0 : amp;
    const double u_raw = cedm::fma_(N, r, -kx);
    const double u = interp_dev<NS, CE_MAX_U_INSTRUCTIONS, true>(P->u_ops, u_n, xh, u_raw);
    double cx = 0.0;
#pragma unroll
    for (int k = 0; k < NS; k++) cx = cedm::fma_(P->sc[k], xh[k], cx);
    const double yerr = y - cedm::fma_(sd, u, cx);
    double dh[NS];
#pragma unroll
    for (int i = 0; i < NS; i++) {
      double acc = cedm::fma_(L[i], yerr, P->sb[i] * u);
#pragma unroll
      for (int k = 0; k < NS; k++) acc = cedm::fma_(P->sai[i * NS + k], xh[k], acc);
      dh[i] = acc;
    }
#pragma unroll
    for (int i = 0; i < NS; i++) xh[i] = xh[i] + dh[i];
    double nls[ND + 1];
#pragma unroll
    for (int i = 0; i < ND; i++) nls[i] = xp[i];
    nls[ND] = y;
    double W[MAXW];
#pragma unroll
    for (int q = 0; q < MAXW; q++) {
      W[q] = 0.0;
      if (q < nw) W[q] = interp_dev<ND + 1, CE_MAX_NL_INSTRUCTIONS, false>(P->nl_ops[q], P->nl_n[q], nls, 0.0);
    }
    double ph[ND];
#pragma unroll
    for (int i = 0; i < ND; i++) {
      double base = P->dbu[i] * u;
#pragma unroll
      for (int q = 0; q < MAXW; q++)
        if (q < nw) base = cedm::fma_(P->dbw[i * nw + q], W[q], base);
      double acc = base;
#pragma unroll
      for (int k = 0; k < ND; k++) acc = cedm::fma_(P->dai[i * ND + k], xp[k], acc);
      ph[i] = acc;
    }
#pragma unroll
    for (int i = 0; i < ND; i++) xp[i] = xp[i] + ph[i];
  }
  if (unstable) return;
  // steady-state mean and overshoot (src/modalsetup.cu:52-62).  max_t fl(y_t * yss) equals
  // fl(max_t y_t * yss) for yss > 0 and fl(min_t y_t * yss) for yss < 0 (rounding is monotonic).
  const double yss = ysum / (double)(T - s0);
  const double pk = cedm::fmax_(0.0, ((yss < 0.0) ? ymin : ymax) * yss);
  double* fr = B.info + (size_t)c * P->info_w + P->info_freqresp;
  fr[2] = pk / cedm::fmax_(yss * yss, EPS6) - 1.0;
  fr[3] = yss;
}

// ------------------------------------------------------------------------------------------------
// Modal fit: Gram accumulation (src/modalsetup.cu:64-96) + LDLT solve and behaviour score
// (src/modalsolve.cu:150-203).  NB = number of 8-column blocks of the 2nt-wide normal equations.
//
// Lane roles inside a candidate's group of G lanes: lane g < NB owns the diagonal block g (its
// lower triangle, as an 8x4 rectangle plus a 4x4 square) and the right-hand side of those 8 rows;
// the remaining lanes own the two 8x4 halves of one off-diagonal block each.  Every Gram entry is
// accumulated by exactly one lane, in time order, as fma(phi_i, phi_j, acc).
template <int NB>
struct FitShape {
  static constexpr int G = (NB == 1) ? 1 : ((NB == 2) ? 4 : 16);
  static constexpr int CPW = 64 / G;  // candidates per workgroup
  static constexpr int NCP = 8 * NB;  // padded basis width
  static constexpr int NTRI = NCP * (NCP + 1) / 2;
  static constexpr int TB = 8;        // samples staged per LDS round
};

__device__ __forceinline__ int tri(int i, int j) { return i * (i + 1) / 2 + j; }

template <int NB>
__global__ __launch_bounds__(64) void k_fit_fp64(const ProblemDev* __restrict__ P, StageBufs B) {
  using S = FitShape<NB>;
  constexpr int G = S::G, CPW = S::CPW, NCP = S::NCP, TB = S::TB;
  __shared__ double phi[CPW][TB][NCP];
  __shared__ double gram[CPW][S::NTRI];
  __shared__ double vec[CPW][5][NCP];  // atb, diag, subdiag, col scale, solution

  const int lane = threadIdx.x;
  const int slot = lane / G, g = lane % G;
  const int idx = blockIdx.x * CPW + slot;
  const int n = B.n;
  const int nt = P->nt, nc2 = 2 * nt;
  bool active = false;
  int u = 0;
  int c = 0;
  double* row = nullptr;
  if (idx < *B.count) {
    c = B.list[idx];
    row = B.info + (size_t)c * P->info_w;
    const bool stable = (B.flags[c] & F_UNSTABLE) == 0;
    active = stable && (row[P->info_scoring] == 0.0 || (B.diag && (B.d_ata || B.d_atb)));
    u = B.uniq[c];
  }
  // unit owned by this lane
  int rb = 0, cb = 0, coff = 0;
  bool diag_unit = false, idle = false;
  if (g < NB) {
    diag_unit = true;
    rb = g;
    cb = g;
  } else {
    int p = (g - NB) >> 1;
    coff = ((g - NB) & 1) * 4;
    rb = 1;
    while (p >= rb) { p -= rb; rb++; }
    cb = p;
    idle = rb >= NB;
  }
  // this candidate's modes (u <= nt), read by the lanes that evaluate them
  const double ts = P->ts, t0 = P->fit_t0, t1 = P->fit_t1;
  double mre[(MAXT + G - 1) / G], mim[(MAXT + G - 1) / G], msh[(MAXT + G - 1) / G];
  constexpr int MPL = (MAXT + G - 1) / G;  // modes per lane
#pragma unroll
  for (int q = 0; q < MPL; q++) {
    const int e = g + q * G;
    mre[q] = 0.0; mim[q] = 0.0; msh[q] = 0.0;
    if (active && e < u) {
      mre[q] = B.mode_re[(size_t)e * n + c];
      mim[q] = B.mode_im[(size_t)e * n + c];
      msh[q] = cedm::fmax_(mre[q] * t0, mre[q] * t1);
    }
  }
  double R1[8][4], R2[4][4], AB[8];
#pragma unroll
  for (int i = 0; i < 8; i++) {
    AB[i] = 0.0;
#pragma unroll
    for (int j = 0; j < 4; j++) R1[i][j] = 0.0;
  }
#pragma unroll
  for (int i = 0; i < 4; i++)
#pragma unroll
    for (int j = 0; j < 4; j++) R2[i][j] = 0.0;

  const double yss = active ? row[P->info_freqresp + 3] : 0.0;
  const int f0 = P->fit_start_id, cnt = P->fit_count;
  const double* ycol = B.ybuf + (active ? idx : 0);
  const int rbase = rb * 8, cbase = cb * 8 + coff;
  for (int tb = 0; tb < cnt; tb += TB) {
    // stage basis values of TB samples
#pragma unroll
    for (int s = 0; s < TB; s++) {
      const int t = tb + s;
      const double tval = (double)(f0 + t) * ts;
#pragma unroll
      for (int q = 0; q < MPL; q++) {
        const int e = g + q * G;
        if (e < NCP / 2) {
          double cv = 0.0, sv = 0.0;
          if (active && e < u && t < cnt) {
            const double decay = cedm::exp(cedm::fma_(mre[q], tval, -msh[q]));
            double sn, cs;
            cedm::sincos(mim[q] * tval, sn, cs);
            cv = decay * cs;
            sv = decay * sn;
          }
          phi[slot][s][2 * e] = cv;
          phi[slot][s][2 * e + 1] = sv;
        }
      }
    }
    __syncthreads();
    if (active && !idle) {
#pragma unroll 1
      for (int s = 0; s < TB; s++) {
        const int t = tb + s;
        if (t >= cnt) break;
        double pr[8], pc[4], pc2[4];
#pragma unroll
        for (int i = 0; i < 8; i++) pr[i] = phi[slot][s][rbase + i];
#pragma unroll
        for (int j = 0; j < 4; j++) pc[j] = phi[slot][s][cbase + j];
#pragma unroll
        for (int j = 0; j < 4; j++) pc2[j] = phi[slot][s][cb * 8 + 4 + j];
#pragma unroll
        for (int i = 0; i < 8; i++)
#pragma unroll
          for (int j = 0; j < 4; j++) R1[i][j] = cedm::fma_(pr[i], pc[j], R1[i][j]);
        if (diag_unit) {
#pragma unroll
          for (int i = 0; i < 4; i++)
#pragma unroll
            for (int j = 0; j < 4; j++) R2[i][j] = cedm::fma_(pr[4 + i], pc2[j], R2[i][j]);
          const double yv = ycol[(size_t)t * n] - yss;
#pragma unroll
          for (int i = 0; i < 8; i++) AB[i] = cedm::fma_(pr[i], yv, AB[i]);
        }
      }
    }
    __syncthreads();
  }

  // scatter the lower triangle into LDS (padded rows/cols beyond 2nt are dropped)
  if (active && !idle) {
#pragma unroll
    for (int i = 0; i < 8; i++)
#pragma unroll
      for (int j = 0; j < 4; j++) {
        const int gi = rbase + i, gj = cbase + j;
        if (gi >= gj && gi < nc2) gram[slot][tri(gi, gj)] = R1[i][j];
      }
    if (diag_unit) {
#pragma unroll
      for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) {
          const int gi = rbase + 4 + i, gj = cb * 8 + 4 + j;
          if (gi >= gj && gi < nc2) gram[slot][tri(gi, gj)] = R2[i][j];
        }
#pragma unroll
      for (int i = 0; i < 8; i++)
        if (rbase + i < nc2) vec[slot][0][rbase + i] = AB[i];
    }
  }
  __syncthreads();
  if (!active || g != 0) return;

  double* G2 = gram[slot];
  double* atb = vec[slot][0];
  double* dg = vec[slot][1];
  double* sd = vec[slot][2];
  double* cs = vec[slot][3];
  double* x = vec[slot][4];
  if (B.diag && B.d_ata) {
    float* da = B.d_ata + (size_t)c * nc2 * nc2;
    for (int i = 0; i < nc2; i++)
      for (int j = 0; j < nc2; j++) da[i * nc2 + j] = (j <= i) ? (float)G2[tri(i, j)] : 0.0f;
  }
  if (B.diag && B.d_atb) {
    float* db = B.d_atb + (size_t)c * nc2;
    for (int i = 0; i < nc2; i++) db[i] = (float)atb[i];
  }
  double* sc = row + P->info_scoring;
  if (sc[0] != 0.0) return;  // modalsolve skips flagged candidates (src/modalsolve.cu:112)

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

// ------------------------------------------------------------------------------------------------
// Margin sweep, one frequency per lane (src/allmargin.cu:100-117): G(z) = K^T (zI - A1)^-1 L by
// elimination with partial pivoting in registers; stores |P G| and arg(P G); min b(w) by atomics.
__device__ __forceinline__ bool margin_skip(const ProblemDev* P, const StageBufs& B, int c) {
  const int f = B.flags[c];
  if (f & (F_NOMINAL_SKIPPED | F_NOMINAL_UNSTABLE)) return true;
  return B.info[(size_t)c * P->info_w + P->info_scoring] >= P->unstable_err;
}

template <int NS>
__global__ __launch_bounds__(64) void k_sweep_fp64(const ProblemDev* __restrict__ P, StageBufs B) {
  const int c = blockIdx.x;
  const int w = blockIdx.y * 64 + threadIdx.x;
  if (margin_skip(P, B, c)) return;
  const int F = P->F;
  if (w >= F) return;
  if (!tab<int32_t>(P, P->off_pok)[w]) { atomicOr(&B.sweep_fail[c], 1); return; }
  const double* oa = B.open_a + (size_t)c * NS * NS;
  const float* gg = B.gains + (size_t)c * P->cuda_w;
  const double zr = tab<double>(P, P->off_zr)[w], zi = tab<double>(P, P->off_zi)[w];
  double Mr[NS][NS], Mi[NS][NS], vr[NS], vi[NS];
#pragma unroll
  for (int i = 0; i < NS; i++) {
#pragma unroll
    for (int j = 0; j < NS; j++) {
      Mr[i][j] = (i == j ? zr : 0.0) - oa[i * NS + j];
      Mi[i][j] = (i == j ? zi : 0.0);
    }
    vr[i] = (double)gg[NS + i];
    vi[i] = 0.0;
  }
#pragma unroll
  for (int k = 0; k < NS; k++) {
    int p = k;
    double best = Mr[k][k] * Mr[k][k] + Mi[k][k] * Mi[k][k];
#pragma unroll
    for (int i = k + 1; i < NS; i++) {
      const double s = Mr[i][k] * Mr[i][k] + Mi[i][k] * Mi[i][k];
      if (s > best) { best = s; p = i; }
    }
    if (best != 0.0) {
      // row swap k <-> p on columns k.. and the right-hand side, by selects
#pragma unroll
      for (int i = k + 1; i < NS; i++) {
        const bool sw = (p == i);
#pragma unroll
        for (int j = k; j < NS; j++) {
          const double ar = Mr[k][j], ai = Mi[k][j], br = Mr[i][j], bi = Mi[i][j];
          Mr[k][j] = sw ? br : ar; Mi[k][j] = sw ? bi : ai;
          Mr[i][j] = sw ? ar : br; Mi[i][j] = sw ? ai : bi;
        }
        const double ar = vr[k], ai = vi[k], br = vr[i], bi = vi[i];
        vr[k] = sw ? br : ar; vi[k] = sw ? bi : ai;
        vr[i] = sw ? ar : br; vi[i] = sw ? ai : bi;
      }
      const double pr = Mr[k][k], pim = Mi[k][k];
      const double dd = pr * pr + pim * pim;
#pragma unroll
      for (int i = k + 1; i < NS; i++) {
        const double ar = Mr[i][k], ai = Mi[i][k];
        Mr[i][k] = (ar * pr + ai * pim) / dd;
        Mi[i][k] = (ai * pr - ar * pim) / dd;
      }
    }
#pragma unroll
    for (int i = k + 1; i < NS; i++) {
      const double lr = Mr[i][k], li = Mi[i][k];
#pragma unroll
      for (int j = k + 1; j < NS; j++) {
        const double qr = lr * Mr[k][j] - li * Mi[k][j];
        const double qi = lr * Mi[k][j] + li * Mr[k][j];
        Mr[i][j] = Mr[i][j] - qr;
        Mi[i][j] = Mi[i][j] - qi;
      }
      const double qr = lr * vr[k] - li * vi[k];
      const double qi = lr * vi[k] + li * vr[k];
      vr[i] = vr[i] - qr;
      vi[i] = vi[i] - qi;
    }
  }
  double xr[NS], xi[NS];
  bool fin = true;
#pragma unroll
  for (int i = NS - 1; i >= 0; i--) {
    const double pr = Mr[i][i], pim = Mi[i][i];
    const double dd = pr * pr + pim * pim;
    xr[i] = (vr[i] * pr + vi[i] * pim) / dd;
    xi[i] = (vi[i] * pr - vr[i] * pim) / dd;
#pragma unroll
    for (int j = 0; j < i; j++) {
      const double qr = Mr[j][i] * xr[i] - Mi[j][i] * xi[i];
      const double qi = Mr[j][i] * xi[i] + Mi[j][i] * xr[i];
      vr[j] = vr[j] - qr;
      vi[j] = vi[j] - qi;
    }
  }
#pragma unroll
  for (int i = 0; i < NS; i++) fin = fin && cedm::isfinite_(xr[i]) && cedm::isfinite_(xi[i]);
  if (!fin) { atomicOr(&B.sweep_fail[c], 1); return; }
  double gr = 0.0, gi = 0.0;
#pragma unroll
  for (int i = 0; i < NS; i++) {
    const double k = (double)gg[i];
    gr = cedm::fma_(k, xr[i], gr);
    gi = cedm::fma_(k, xi[i], gi);
  }
  const double pr = tab<double>(P, P->off_pr)[w], pim = tab<double>(P, P->off_pi)[w];
  const double rr = pr * gr - pim * gi, ri = pr * gi + pim * gr;
  const double m = cedm::hypot(rr, ri);
  const double raw = cedm::atan2(ri, rr);
  const double Ps = P->alpha * tab<double>(P, P->off_pabs)[w];
  const double Cs = cedm::hypot(gr, gi) / P->alpha;
  const double num = cedm::hypot(1.0 + rr, ri);
  const double b = num / (cedm::sqrt_(1.0 + Ps * Ps) * cedm::sqrt_(1.0 + Cs * Cs));
  double* out = B.mag_raw + ((size_t)c * F + w) * 2;
  out[0] = m;
  out[1] = raw;
  if (b == b) atomicMin(&B.bmin[c], (unsigned long long)__double_as_longlong(b));
}

// Sequential part of the sweep (src/allmargin.cu:119-180), one candidate per lane.
__global__ __launch_bounds__(64) void k_scan_fp64(const ProblemDev* __restrict__ P, StageBufs B) {
  const int c = blockIdx.x * 64 + threadIdx.x;
  if (c >= B.n) return;
  double* row = B.info + (size_t)c * P->info_w;
  double* sc = row + P->info_scoring;
  double* am = row + P->info_allmargin;
  const double UE = P->unstable_err, EF = P->eig_fail_err;
  const int flags = B.flags[c];
  if (sc[0] >= UE) {
    sc[1] = sc[0]; sc[2] = sc[0];
    am[0] = am[1] = am[2] = am[3] = am[4] = 0.0;
    return;
  }
  if (flags & F_NOMINAL_UNSTABLE) {
    sc[1] = sc[0];
    sc[2] = UE;
    am[0] = am[1] = am[2] = am[3] = am[4] = 0.0;
    const double total = cedm::log((sc[0] + UE) + EPS30);
    sc[0] = cedm::isfinite_(total) ? total : EF;
    return;
  }
  if (B.sweep_fail[c]) {
    sc[1] = sc[0]; sc[2] = EF; sc[0] = EF;
    am[0] = 0.0; am[1] = 0.0; am[2] = 0.0; am[3] = 1.0; am[4] = 0.0;
    return;
  }
  const int F = P->F;
  const double ts = P->ts;
  const double* th = tab<double>(P, P->off_theta);
  const double* mr = B.mag_raw + (size_t)c * F * 2;
  double PM = FLTMAX, GM = FLTMAX, DM = FLTMAX;
  bool crossed = false;
  double prevMag = 0.0, prevPhase = 0.0, prevRaw = 0.0, prevTheta = 0.0;
  for (int w = 0; w < F; w++) {
    const double m = mr[2 * w], raw = mr[2 * w + 1], theta_w = th[w];
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
      const double a = prevPhase + PI_D, b = ph + PI_D;
      const double ka = cedm::floor_(a / TWO_PI_D), kb = cedm::floor_(b / TWO_PI_D);
      if (ka != kb && b != a) {
        const double target = TWO_PI_D * (b > a ? kb : ka);
        const double frac = (target - a) / (b - a);
        const double mag_p = prevMag + frac * (m - prevMag);
        GM = cedm::fmin_(GM, -20.0 * cedm::log10(mag_p));
      }
      prevPhase = ph;
    } else {
      prevPhase = raw;
    }
    prevMag = m;
    prevRaw = raw;
    prevTheta = theta_w;
  }
  const double Bpc = cedm::fmin_(1.0, __longlong_as_double((long long)B.bmin[c]));
  const double err_gm = cedm::fmax_(0.0, P->target_gm - GM) / P->target_gm;
  const double err_pm = crossed ? cedm::fmax_(0.0, P->target_pm - PM) / P->target_pm : 0.0;
  const double err_b = cedm::fmax_(0.0, P->target_b - Bpc) / P->target_b;
  const double margin = (err_gm * err_gm + err_pm * err_pm) + err_b * err_b;
  sc[1] = sc[0];
  sc[2] = margin;
  am[0] = GM; am[1] = PM; am[2] = DM; am[3] = 1.0; am[4] = Bpc;
  const double total = cedm::log((sc[0] + margin) + EPS30);
  sc[0] = cedm::isfinite_(total) ? total : EF;
}

// ------------------------------------------------------------------------------------------------
// Launch helpers (host side of this translation unit)

using SimFn = void (*)(const ProblemDev*, StageBufs);

template <int NS, int ND>
static void launch_sim_t(const ProblemDev* P, const StageBufs& B, hipStream_t st) {
  hipLaunchKernelGGL((k_sim_fp64<NS, ND>), dim3((B.n + 63) / 64), dim3(64), 0, st, P, B);
}

template <int NS>
static bool launch_sim_ns(int nd, const ProblemDev* P, const StageBufs& B, hipStream_t st) {
  switch (nd) {
    case 1: launch_sim_t<NS, 1>(P, B, st); return true;
    case 2: launch_sim_t<NS, 2>(P, B, st); return true;
    case 3: launch_sim_t<NS, 3>(P, B, st); return true;
    case 4: launch_sim_t<NS, 4>(P, B, st); return true;
    case 5: launch_sim_t<NS, 5>(P, B, st); return true;
    case 6: launch_sim_t<NS, 6>(P, B, st); return true;
    case 7: launch_sim_t<NS, 7>(P, B, st); return true;
    case 8: launch_sim_t<NS, 8>(P, B, st); return true;
  }
  return false;
}

bool launch_sim_fp64(int ns, int nd, const ProblemDev* P, const StageBufs& B, hipStream_t st) {
  switch (ns) {
    case 1: return launch_sim_ns<1>(nd, P, B, st);
    case 2: return launch_sim_ns<2>(nd, P, B, st);
    case 3: return launch_sim_ns<3>(nd, P, B, st);
    case 4: return launch_sim_ns<4>(nd, P, B, st);
    case 5: return launch_sim_ns<5>(nd, P, B, st);
    case 6: return launch_sim_ns<6>(nd, P, B, st);
    case 7: return launch_sim_ns<7>(nd, P, B, st);
    case 8: return launch_sim_ns<8>(nd, P, B, st);
  }
  return false;
}

void launch_eig_fp64(const ProblemDev* P, const StageBufs& B, hipStream_t st) {
  hipLaunchKernelGGL(k_eig_fp64, dim3((B.n + 63) / 64), dim3(64), 0, st, P, B);
}

void launch_compact(const StageBufs& B, int all, hipStream_t st) {
  hipLaunchKernelGGL(k_compact, dim3((B.n + 63) / 64), dim3(64), 0, st, B, all);
}

bool launch_fit_fp64(int nt, const ProblemDev* P, const StageBufs& B, hipStream_t st) {
  const int nb = (2 * nt + 7) / 8;
  switch (nb) {
    case 1: hipLaunchKernelGGL(k_fit_fp64<1>, dim3((B.n + FitShape<1>::CPW - 1) / FitShape<1>::CPW), dim3(64), 0, st, P, B); return true;
    case 2: hipLaunchKernelGGL(k_fit_fp64<2>, dim3((B.n + FitShape<2>::CPW - 1) / FitShape<2>::CPW), dim3(64), 0, st, P, B); return true;
    case 3: hipLaunchKernelGGL(k_fit_fp64<3>, dim3((B.n + FitShape<3>::CPW - 1) / FitShape<3>::CPW), dim3(64), 0, st, P, B); return true;
    case 4: hipLaunchKernelGGL(k_fit_fp64<4>, dim3((B.n + FitShape<4>::CPW - 1) / FitShape<4>::CPW), dim3(64), 0, st, P, B); return true;
  }
  return false;
}

bool launch_sweep_fp64(int ns, int F, const ProblemDev* P, const StageBufs& B, hipStream_t st) {
  const dim3 grid(B.n, (F + 63) / 64);
  switch (ns) {
    case 1: hipLaunchKernelGGL(k_sweep_fp64<1>, grid, dim3(64), 0, st, P, B); break;
    case 2: hipLaunchKernelGGL(k_sweep_fp64<2>, grid, dim3(64), 0, st, P, B); break;
    case 3: hipLaunchKernelGGL(k_sweep_fp64<3>, grid, dim3(64), 0, st, P, B); break;
    case 4: hipLaunchKernelGGL(k_sweep_fp64<4>, grid, dim3(64), 0, st, P, B); break;
    case 5: hipLaunchKernelGGL(k_sweep_fp64<5>, grid, dim3(64), 0, st, P, B); break;
    case 6: hipLaunchKernelGGL(k_sweep_fp64<6>, grid, dim3(64), 0, st, P, B); break;
    case 7: hipLaunchKernelGGL(k_sweep_fp64<7>, grid, dim3(64), 0, st, P, B); break;
    case 8: hipLaunchKernelGGL(k_sweep_fp64<8>, grid, dim3(64), 0, st, P, B); break;
    default: return false;
  }
  hipLaunchKernelGGL(k_scan_fp64, dim3((B.n + 63) / 64), dim3(64), 0, st, P, B);
  return true;
}

void launch_info_init(const float* in, double* out, size_t count, hipStream_t st) {
  hipLaunchKernelGGL(k_info_init, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, st, in, out, count);
}
void launch_info_final(const double* in, float* out, size_t count, hipStream_t st) {
  hipLaunchKernelGGL(k_info_final, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, st, in, out, count);
}

}  // namespace ce

// Kalman-filter closed-loop simulation (reference src/discrete_solves.cu:83-185, the
// discrete_solve<ControllerType::Kalman> kernel, which the reference never launches).
//
// Gains row: [K (ns), N, R, Q (ns)], 2 ns + 2 floats.  The reference sizes every gains row 2 ns + 1
// (src/config.cu:51) and so reads the last Q entry from the next candidate's row; here the row is
// 2 ns + 2 wide.
//
// Per step, one candidate per lane, state and covariance in registers:
//   y = DC(xp); u = U(N r - K xh)
//   predict   xh += (A - I) xh + B u;  P = A P A' + diag(Q)
//   gain      KG = P C' / (C P C' + R)
//   update    xh += KG (y - C xh);  P = (I - KG C) P (I - KG C)' + R KG KG'
//   plant     as the Luenberger simulator (NL chains, BU, BW)
// The reference writes the covariance product in place through restrict pointers
// (ABATf(IKHP, P, IKHP)); this is the product it intends.  Its reciprocal is the fast __fdividef;
// here it is an IEEE division.
//
// FP64 form: double, the oracle's operation order (rows bit-identical to it).  Float form: the
// reference's arithmetic, the state as hi/lo pairs (TwoSum / Fast2Sum, the update's product split
// exactly with an FMA, src/main.cu:129-155).  Loops run over the problem's dimensions inside
// kernels instantiated for dimension capacities 2, 4 and 8.
#include <hip/hip_runtime.h>

#include "ce_sim_dev.h"
#include "ce_stage.h"

namespace ce {

template <typename R, int NS, int ND>
__global__ __launch_bounds__(64) void k_sim_kalman(const ProblemDev* __restrict__ P, const float* __restrict__ gains,
                                                    int n, R* __restrict__ yout) {
  using O = RealOps<R>;
  constexpr bool HL = sizeof(R) == 4;
  const int c = blockIdx.x * 64 + threadIdx.x;
  if (c >= n) return;
  const int ns = P->ns, nd = P->nd, nw = P->nw, T = P->T, k0 = P->ref_step_id;
  const float* g = gains + (size_t)c * (2 * ns + 2);
  R K[NS], Q[NS];
#pragma unroll
  for (int i = 0; i < NS; i++) {
    K[i] = i < ns ? (R)g[i] : R(0);
    Q[i] = i < ns ? (R)g[ns + 2 + i] : R(0);
  }
  const R N = (R)g[ns], Rn = (R)g[ns + 1];
  const R amp = O::amp(P);
  const R* sai = O::sai(P);
  const R* sb = O::sb(P);
  const R* scv = O::sc(P);
  const R* dai = O::dai(P);
  const R* dbu = O::dbu(P);
  const R* dbw = O::dbw(P);
  // the covariance products use A itself (STATESPACE_A)
  R sa[NS * NS];
#pragma unroll
  for (int i = 0; i < NS; i++)
#pragma unroll
    for (int j = 0; j < NS; j++) sa[i * NS + j] = (i < ns && j < ns) ? (R)(HL ? P->fsa[i * ns + j] : P->sa[i * ns + j]) : R(0);
  R xh[NS], xhl[NS], xp[ND], xpl[ND], Pm[NS * NS];
#pragma unroll
  for (int i = 0; i < NS; i++) xh[i] = xhl[i] = R(0);
#pragma unroll
  for (int i = 0; i < ND; i++) xp[i] = xpl[i] = R(0);
#pragma unroll
  for (int i = 0; i < NS * NS; i++) Pm[i] = R(0);
  R* yrow = yout + (size_t)c * T;
  for (int t = 0; t < T; t++) {
    const R y = interp_dev<R, ND, CE_MAX_DC_INSTRUCTIONS, false>(P->dc_ops, P->dc_n, xp, R(0));
    yrow[t] = y;
    R kx = R(0);
#pragma unroll
    for (int k = 0; k < NS; k++)
      if (k < ns) kx = O::fma(K[k], xh[k], kx);
    if (HL) {
#pragma unroll
      for (int k = 0; k < NS; k++)
        if (k < ns) kx = O::fma(K[k], xhl[k], kx);
    }
    const R r = (t < k0) ? R(0) : amp;
    const R u_raw = O::fma(N, r, -kx);
    const R u = interp_dev<R, NS, CE_MAX_U_INSTRUCTIONS, true>(P->u_ops, P->u_n, xh, u_raw);
    // predict
    R dh[NS], dl[NS];
#pragma unroll
    for (int i = 0; i < NS; i++) {
      R acc = R(0), lo = R(0);
      if (i < ns) {
        acc = sb[i] * u;
#pragma unroll
        for (int k = 0; k < NS; k++)
          if (k < ns) acc = O::fma(sai[i * ns + k], xh[k], acc);
        if (HL) {
#pragma unroll
          for (int k = 0; k < NS; k++)
            if (k < ns) lo = O::fma(sai[i * ns + k], xhl[k], lo);
        }
      }
      dh[i] = acc;
      dl[i] = lo;
    }
#pragma unroll
    for (int i = 0; i < NS; i++) {
      if (i < ns) {
        if constexpr (HL) hl_add(xh[i], xhl[i], dh[i], dl[i]);
        else xh[i] = xh[i] + dh[i];
      }
    }
    // P = A P A' + diag(Q)
    R C0[NS * NS], Pn[NS * NS];
#pragma unroll
    for (int i = 0; i < NS; i++)
#pragma unroll
      for (int j = 0; j < NS; j++) {
        R a = R(0);
#pragma unroll
        for (int k = 0; k < NS; k++)
          if (k < ns) a = O::fma(sa[i * NS + k], Pm[k * NS + j], a);
        C0[i * NS + j] = a;
      }
#pragma unroll
    for (int i = 0; i < NS; i++)
#pragma unroll
      for (int j = 0; j < NS; j++) {
        R a = R(0);
#pragma unroll
        for (int k = 0; k < NS; k++)
          if (k < ns) a = O::fma(C0[i * NS + k], sa[j * NS + k], a);
        Pn[i * NS + j] = a;
      }
#pragma unroll
    for (int i = 0; i < NS; i++)
      if (i < ns) Pn[i * NS + i] = Pn[i * NS + i] + Q[i];
    // KG = P C' / (C P C' + R)
    R pht[NS], cp[NS];
#pragma unroll
    for (int i = 0; i < NS; i++) {
      R a = R(0), b = R(0);
#pragma unroll
      for (int k = 0; k < NS; k++)
        if (k < ns) {
          a = O::fma(Pn[i * NS + k], scv[k < ns ? k : 0], a);
          b = O::fma(scv[k < ns ? k : 0], Pn[k * NS + i], b);
        }
      pht[i] = a;
      cp[i] = b;
    }
    R hphr = R(0);
#pragma unroll
    for (int k = 0; k < NS; k++)
      if (k < ns) hphr = O::fma(cp[k], scv[k < ns ? k : 0], hphr);
    hphr = R(1) / (hphr + Rn);
    R kg[NS];
#pragma unroll
    for (int i = 0; i < NS; i++) kg[i] = pht[i] * hphr;
    // xh += KG (y - C xh)
    R dz = -y;
#pragma unroll
    for (int k = 0; k < NS; k++)
      if (k < ns) dz = O::fma(scv[k < ns ? k : 0], xh[k], dz);
    dz = -dz;
    if constexpr (HL) {
      R dzl = R(0);
#pragma unroll
      for (int k = 0; k < NS; k++)
        if (k < ns) dzl = O::fma(scv[k < ns ? k : 0], xhl[k], dzl);
      dzl = -dzl;
#pragma unroll
      for (int i = 0; i < NS; i++)
        if (i < ns) {
          const R s = kg[i] * dz;
          R e = O::fma(kg[i], dz, -s);
          e = O::fma(kg[i], dzl, e);
          hl_add(xh[i], xhl[i], s, e);
        }
    } else {
#pragma unroll
      for (int i = 0; i < NS; i++)
        if (i < ns) xh[i] = xh[i] + kg[i] * dz;
    }
    // P = (I - KG C) P (I - KG C)' + R KG KG'
    R ikh[NS * NS];
#pragma unroll
    for (int i = 0; i < NS; i++)
#pragma unroll
      for (int j = 0; j < NS; j++) ikh[i * NS + j] = (i == j ? R(1) : R(0)) - kg[i] * (j < ns ? scv[j] : R(0));
#pragma unroll
    for (int i = 0; i < NS; i++)
#pragma unroll
      for (int j = 0; j < NS; j++) {
        R a = R(0);
#pragma unroll
        for (int k = 0; k < NS; k++)
          if (k < ns) a = O::fma(ikh[i * NS + k], Pn[k * NS + j], a);
        C0[i * NS + j] = a;
      }
#pragma unroll
    for (int i = 0; i < NS; i++)
#pragma unroll
      for (int j = 0; j < NS; j++) {
        R a = R(0);
#pragma unroll
        for (int k = 0; k < NS; k++)
          if (k < ns) a = O::fma(C0[i * NS + k], ikh[j * NS + k], a);
        Pm[i * NS + j] = (i < ns && j < ns) ? a + (kg[i] * Rn) * kg[j] : R(0);
      }
    // plant step (the Luenberger simulator's)
    R nls[ND + 1];
#pragma unroll
    for (int i = 0; i < ND; i++) nls[i] = xp[i];
    nls[ND] = y;
    // the NL chains read column nd as y: move it into the slot the chains index
    R src[ND + 1];
#pragma unroll
    for (int i = 0; i <= ND; i++) src[i] = (i < nd) ? nls[i] : y;
    R W[MAXW];
#pragma unroll
    for (int q = 0; q < MAXW; q++) {
      W[q] = R(0);
      if (q < nw) W[q] = interp_dev<R, ND + 1, CE_MAX_NL_INSTRUCTIONS, false>(P->nl_ops[q], P->nl_n[q], src, R(0));
    }
    R ph[ND], pl[ND];
#pragma unroll
    for (int i = 0; i < ND; i++) {
      R base = R(0), lo = R(0);
      if (i < nd) {
        base = dbu[i] * u;
#pragma unroll
        for (int q = 0; q < MAXW; q++)
          if (q < nw) base = O::fma(dbw[i * nw + q], W[q], base);
#pragma unroll
        for (int k = 0; k < ND; k++)
          if (k < nd) base = O::fma(dai[i * nd + k], xp[k], base);
        if (HL) {
#pragma unroll
          for (int k = 0; k < ND; k++)
            if (k < nd) lo = O::fma(dai[i * nd + k], xpl[k], lo);
        }
      }
      ph[i] = base;
      pl[i] = lo;
    }
#pragma unroll
    for (int i = 0; i < ND; i++) {
      if (i < nd) {
        if constexpr (HL) hl_add(xp[i], xpl[i], ph[i], pl[i]);
        else xp[i] = xp[i] + ph[i];
      }
    }
  }
}

template <typename R, int NS>
static bool launch_kalman_ns(int nd, const ProblemDev* P, const float* gains, int n, R* y, hipStream_t st) {
  const dim3 grid((n + 63) / 64), block(64);
  if (nd <= 2) hipLaunchKernelGGL((k_sim_kalman<R, NS, 2>), grid, block, 0, st, P, gains, n, y);
  else if (nd <= 4) hipLaunchKernelGGL((k_sim_kalman<R, NS, 4>), grid, block, 0, st, P, gains, n, y);
  else if (nd <= 8) hipLaunchKernelGGL((k_sim_kalman<R, NS, 8>), grid, block, 0, st, P, gains, n, y);
  else return false;
  kmark(sizeof(R) == 8 ? "k_sim_kalman_fp64" : "k_sim_kalman_f32", st);
  return true;
}

template <typename R>
static bool launch_kalman_t(int ns, int nd, const ProblemDev* P, const float* gains, int n, R* y, hipStream_t st) {
  if (ns <= 2) return launch_kalman_ns<R, 2>(nd, P, gains, n, y, st);
  if (ns <= 4) return launch_kalman_ns<R, 4>(nd, P, gains, n, y, st);
  if (ns <= 8) return launch_kalman_ns<R, 8>(nd, P, gains, n, y, st);
  return false;
}

bool launch_sim_kalman(int ns, int nd, bool mixed, const ProblemDev* P, const float* gains, int n, double* y64,
                       float* y32, hipStream_t st) {
  return mixed ? launch_kalman_t<float>(ns, nd, P, gains, n, y32, st)
               : launch_kalman_t<double>(ns, nd, P, gains, n, y64, st);
}

}  // namespace ce

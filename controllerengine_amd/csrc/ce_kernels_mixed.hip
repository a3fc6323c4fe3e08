// Mixed-precision stage kernels for gfx950.
//
//   simulation     float state (generic: k_sim<float,NS,ND>, or the specialised kernel); the fit
//                  window of y is stored as float, SoA.
//   modal fit      k_fitcf: the normal equations of the damped-sinusoid fit (reference
//                  src/modalsetup.cu:64-91) are formed in double without a pass over time for the
//                  Gram: each entry sum_t e^{(s_e+s_f) tau} trig(w_e tau) trig(w_f tau) is a geometric
//                  series, summed in closed form from per-mode expm1 values (no cancellation for
//                  slow modes).  The
//                  right-hand side sum_t phi(t)(y(t) - yss) runs over the stored trajectory with the
//                  basis advanced by a double complex recurrence.  Then the same LDL^T solve and
//                  score as the FP64 chain (src/modalsolve.cu).
//   margin sweep   k_sweep_mark32: one wave per candidate, every frequency located by an O(ns^2)
//                  FP64 solve on the Hessenberg form with float |loop|, arg and b(w), each with an
//                  error estimate; the points whose values show (or come within an error-widened
//                  guard band of) a gain or phase crossing, a non-finite value, or the smallest
//                  b(w) and its band are appended to a dense list.  k_refine_full64: the listed
//                  points are recomputed with the FP64 sweep point (the FP64 chain's arithmetic),
//                  and candidates with more than RMAX such points in full.  k_scan32: the
//                  reference's unwrap / crossing / interpolation scan (src/allmargin.cu:119-180) in
//                  double over the refined values, one wave per candidate (chunk-wise phase prefix,
//                  wave minima).
#include <hip/hip_runtime.h>

#include <string>

#include "ce_fitsolve_dev.h"
#include "ce_sim_dev.h"
#include "ce_stage.h"
#include "ce_sweep_dev.h"

namespace ce {

constexpr float GUARD_MAG = 1e-3f;   // |m - 1| below this: refine in FP64
constexpr float GUARD_PHASE = 1e-3f; // phase within this of an odd multiple of pi, or a jump within
                                     // this of pi: refine in FP64
constexpr float BAND_SPREAD = 1e-6f; // b band over a flat minimum: refine only its smallest point

// ------------------------------------------------------------------------------------------------
struct Cplx {
  double r, i;
};
__device__ __forceinline__ Cplx cmul(Cplx a, Cplx b) { return Cplx{a.r * b.r - a.i * b.i, a.r * b.i + a.i * b.r}; }
__device__ __forceinline__ Cplx cdiv(Cplx a, Cplx b) {
  const double d = b.r * b.r + b.i * b.i;
  return Cplx{(a.r * b.r + a.i * b.i) / d, (a.i * b.r - a.r * b.i) / d};
}
// e^{a} - 1 for complex a, accurate near 0
__device__ __forceinline__ Cplx cexpm1(double ar, double ai) {
  double s, c;
  cedm::sincos(ai, s, c);
  double sh, ch;
  cedm::sincos(0.5 * ai, sh, ch);
  const double em = cedm::expm1(ar);
  return Cplx{em * c - 2.0 * sh * sh, (em + 1.0) * s};
}

// Right-hand side of the fit, sum_t phi_e(t) (y(t) - yss), by Horner's rule in double:
// sum_{t in part} (y_t - yss) z_e^(t - t_part) with z_e = exp(s_e ts), over the part's samples in
// reverse.  Lane = list slot (coalesced reads of the SoA trajectory), 8 modes per lane
// (blockIdx.y selects modes 8y..8y+7), blockIdx.z = time part.
constexpr int ATB_MODES = 8;

// M modes of one lane's candidate over samples [tb, te), newest first; samples are loaded in
// groups, one group ahead of the Horner steps.  Each mode's recurrence is the same operation
// sequence for any M, so the sums do not depend on the mode count the wave runs with.
template <int M>
__device__ __forceinline__ void atb_run(const ProblemDev* __restrict__ P, const StageBufs& B, int c, int idx, int u,
                                        int e0, int tb, int te, int bz, double yss) {
  const int n = B.n;
  const double ts = P->ts;
  double zr[M], zi[M], ar[M], ai[M];
#pragma unroll
  for (int k = 0; k < M; k++) {
    zr[k] = 0.0;
    zi[k] = 0.0;
    ar[k] = 0.0;
    ai[k] = 0.0;
    if (e0 + k < u) {
      const double re = B.mode_re[(size_t)(e0 + k) * n + c], im = B.mode_im[(size_t)(e0 + k) * n + c];
      const double ms = cedm::exp(re * ts);
      double s, co;
      cedm::sincos(im * ts, s, co);
      zr[k] = ms * co;
      zi[k] = ms * s;
    }
  }
  const float* ycol = B.ybuf32 + idx;
  auto step = [&](float yf) {
    const double y = (double)yf - yss;
#pragma unroll
    for (int k = 0; k < M; k++) {
      const double nr = __builtin_fma(ar[k], zr[k], __builtin_fma(-ai[k], zi[k], y));
      const double ni = __builtin_fma(ar[k], zi[k], ai[k] * zr[k]);
      ar[k] = nr;
      ai[k] = ni;
    }
  };
  // groups of 8 samples; the next group's loads are issued before the current group's steps
  constexpr int GS = 8;
  int t = te - 1;
  float cur[GS], nxt[GS];
  if (t - (GS - 1) >= tb) {
#pragma unroll
    for (int j = 0; j < GS; j++) nxt[j] = ycol[(size_t)(t - j) * n];
  }
  for (; t - (GS - 1) >= tb; t -= GS) {
#pragma unroll
    for (int j = 0; j < GS; j++) cur[j] = nxt[j];
    if (t - GS - (GS - 1) >= tb) {
#pragma unroll
      for (int j = 0; j < GS; j++) nxt[j] = ycol[(size_t)(t - GS - j) * n];
    }
#pragma unroll
    for (int j = 0; j < GS; j++) step(cur[j]);
  }
  for (; t >= tb; t--) step(ycol[(size_t)t * n]);
  double2* out = B.atbp + ((size_t)bz * P->nt + e0) * n + idx;
#pragma unroll
  for (int k = 0; k < M; k++)
    if (e0 + k < u) out[(size_t)k * n] = make_double2(ar[k], ai[k]);
}

// list slot idx belongs to the part of the list this launch covers (split simulation)
__device__ __forceinline__ bool in_sim_part(const StageBufs& B, int idx, int live) {
  if (B.sim_part == 0) return true;
  const int lim = sim_split_slots(live, B.sim_cap);
  return (B.sim_part == 1) ? idx < lim : idx >= lim;
}

__device__ __forceinline__ void k_atb_body(const ProblemDev* __restrict__ P, const StageBufs& B, int bx, int nbx, int by, int bz) {
  const int idx = bx * 64 + threadIdx.x;
  const int live = *B.count;
  if (idx >= live || !in_sim_part(B, idx, live)) return;
  const int c = B.list[idx];
  const double* row = B.info + (size_t)c * P->info_w;
  if (B.flags[c] & F_UNSTABLE) return;
  if (!(row[P->info_scoring] == 0.0 || (B.diag && (B.d_ata || B.d_atb)))) return;
  const int u = B.uniq[c];
  const int e0 = by * ATB_MODES;
  if (e0 >= u) return;
  const int N = P->fit_count;
  const int tb = bz * B.atb_chunk;
  const int te = (tb + B.atb_chunk < N) ? tb + B.atb_chunk : N;
  const double yss = row[P->info_freqresp + 3];
  // the wave runs with the largest mode count among its active lanes (2, 4 or 8)
  const int um = u - e0;
  if (__ballot(um > 4) != 0ull) atb_run<8>(P, B, c, idx, u, e0, tb, te, bz, yss);
  else if (__ballot(um > 2) != 0ull) atb_run<4>(P, B, c, idx, u, e0, tb, te, bz, yss);
  else atb_run<2>(P, B, c, idx, u, e0, tb, te, bz, yss);
}

__global__ __launch_bounds__(64) void k_atb(const ProblemDev* __restrict__ P, StageBufs B) {
  k_atb_body(P, B, blockIdx.x, gridDim.x, blockIdx.y, blockIdx.z);
}

// local block -> (candidate block, mode block, time part) of the problem's own grid
template <bool LOOP>
__global__ __launch_bounds__(64) void k_atb_batch(BatchTab T) {
  batch_blocks<LOOP>(T, [&](const BatchSlot& q) {
    const ProblemDev* P = T.probs[q.p];
    const StageBufs& B = T.bufs[q.p];
    const int nbx = (B.n + 63) / 64, nby = (P->nt + ATB_MODES - 1) / ATB_MODES;
    const int r = q.lb / nbx;
    k_atb_body(P, B, q.lb % nbx, nbx, r % nby, r / nby);
  });
}


template <int NB>
struct CfShape {
  static constexpr int G = 16;
  static constexpr int CPW = 4;
  static constexpr int NCP = 8 * NB;
  static constexpr int NTRI = NCP * (NCP + 1) / 2;
};

// Gram matrix in closed form, right-hand side from the partial sums of k_atb, then the LDL^T
// solve and score.  G = 16 lanes per candidate, 4 candidates per single-wave workgroup.
template <int NB>
__device__ __forceinline__ void k_fitcf_body(const ProblemDev* __restrict__ P, const StageBufs& B, int bx, int nbx, int by, int bz) {
  using S = CfShape<NB>;
  constexpr int G = S::G, CPW = S::CPW, NTRI = S::NTRI, NCP = S::NCP;
  __shared__ double gram[CPW][NTRI];
  __shared__ double vec[CPW][6][NCP];
  __shared__ int itr[CPW][NCP];
  const int lane = threadIdx.x;
  const int slot = lane / G, g = lane % G;
  const int idx = bx * CPW + slot;
  const int n = B.n;
  const int nt = P->nt, nc2 = 2 * nt;
  bool active = false;
  int u = 0, c = 0;
  double* row = nullptr;
  const int live = *B.count;
  if (idx < live && in_sim_part(B, idx, live)) {
    c = B.list[idx];
    row = B.info + (size_t)c * P->info_w;
    const bool stable = (B.flags[c] & F_UNSTABLE) == 0;
    active = stable && (row[P->info_scoring] == 0.0 || (B.diag && (B.d_ata || B.d_atb)));
    u = B.uniq[c];
  }
  for (int k = g; k < NTRI; k += G) gram[slot][k] = 0.0;
  for (int k = g; k < NCP; k += G) vec[slot][0][k] = 0.0;
  __syncthreads();
  const double ts = P->ts, t0 = P->fit_t0, t1 = P->fit_t1;
  const int f0 = P->fit_start_id, N = P->fit_count;
  const double tau0 = (double)f0 * ts;
  // per-mode factors of the closed form (modes of this candidate in vec[1..4], which the solve
  // below uses as scratch only after the Gram is complete): PM(q, e), q = E1 re/im, EN re/im,
  // W re/im, x1, xN
  constexpr int NM = NCP / 2;
  double* pm = &vec[slot][1][0];
  auto PM = [&](int q, int e) -> double& { return pm[q * NM + e]; };
  if (active && g < u) {
    // Atb_e = sum_parts phi_e(t_part) * partial_e,part (complex)
    const int e = g;
    const double re = B.mode_re[(size_t)e * n + c], im = B.mode_im[(size_t)e * n + c];
    const double sh = cedm::fmax_(re * t0, re * t1);
    // basis value at each part's first sample: the window start's value advanced by
    // z_e^chunk = exp((re + i im) chunk ts) per part
    const double mag = cedm::exp(re * tau0 - sh);
    double sw0, cw0;
    cedm::sincos(im * tau0, sw0, cw0);
    double sr = 0.0, si = 0.0;
    {
      double br = mag * cw0, bi = mag * sw0;
      const double tch = (double)B.atb_chunk * ts;
      const double mw = cedm::exp(re * tch);
      double sw, cw;
      cedm::sincos(im * tch, sw, cw);
      const double wr = mw * cw, wi = mw * sw;
      for (int q = 0; q < B.atb_parts; q++) {
        const double2 pv = B.atbp[((size_t)q * nt + e) * n + idx];
        sr = sr + (br * pv.x - bi * pv.y);
        si = si + (br * pv.y + bi * pv.x);
        const double nr = br * wr - bi * wi;
        bi = br * wi + bi * wr;
        br = nr;
      }
    }
    vec[slot][0][2 * e] = sr;
    vec[slot][0][2 * e + 1] = si;
    const double ar = re * ts, ai = im * ts;
    const Cplx e1 = cexpm1(ar, ai), en = cexpm1((double)N * ar, (double)N * ai);
    PM(0, e) = e1.r;
    PM(1, e) = e1.i;
    PM(2, e) = en.r;
    PM(3, e) = en.i;
    PM(4, e) = mag * cw0;
    PM(5, e) = mag * sw0;
    PM(6, e) = cedm::expm1(2.0 * ar);
    PM(7, e) = cedm::expm1(2.0 * ((double)N * ar));
  }
  __syncthreads();
  if (active) {
    // Gram in closed form: mode pairs (e >= f) spread over the group's lanes.  Each entry pairs
    // sums of e^{(z_e + z_f) t} and e^{(z_e + conj z_f) t}, geometric series whose ratio
    // expm1(N z) / expm1(z) is formed from the per-mode expm1 values:
    // expm1(a + b) = A + B + A B with A = expm1(a), B = expm1(b) (no cancellation: the real parts
    // of stable modes share their sign), and for e = f the difference exponent is the real 2 s_e ts.
    const int npairs = u * (u + 1) / 2;
    for (int p = g; p < npairs; p += G) {
      int ea = 0;
      while ((ea + 1) * (ea + 2) / 2 <= p) ea++;
      const int fb = p - ea * (ea + 1) / 2;
      const double se = B.mode_re[(size_t)ea * n + c], we = B.mode_im[(size_t)ea * n + c];
      const double sf = B.mode_re[(size_t)fb * n + c], wf = B.mode_im[(size_t)fb * n + c];
      const Cplx E1e{PM(0, ea), PM(1, ea)}, ENe{PM(2, ea), PM(3, ea)}, We{PM(4, ea), PM(5, ea)};
      const Cplx E1f{PM(0, fb), PM(1, fb)}, ENf{PM(2, fb), PM(3, fb)}, Wf{PM(4, fb), PM(5, fb)};
      const Cplx E1c{E1f.r, -E1f.i}, ENc{ENf.r, -ENf.i}, Wc{Wf.r, -Wf.i};
      auto xm1 = [](Cplx a, Cplx b) {  // expm1(x + y) from A = expm1(x), B = expm1(y)
        const Cplx ab = cmul(a, b);
        return Cplx{(a.r + b.r) + ab.r, (a.i + b.i) + ab.i};
      };
      Cplx rd, rs;
      if (se + sf == 0.0 && we - wf == 0.0) rd = Cplx{(double)N, 0.0};
      else if (ea == fb) rd = Cplx{PM(7, ea) / PM(6, ea), 0.0};
      else rd = cdiv(xm1(ENe, ENc), xm1(E1e, E1c));
      if (se + sf == 0.0 && we + wf == 0.0) rs = Cplx{(double)N, 0.0};
      else rs = cdiv(xm1(ENe, ENf), xm1(E1e, E1f));
      const Cplx sdif = cmul(cmul(We, Wc), rd);
      const Cplx ssum = cmul(cmul(We, Wf), rs);
      const double cc = 0.5 * (sdif.r + ssum.r), ss = 0.5 * (sdif.r - ssum.r);
      const double cs = 0.5 * (ssum.i - sdif.i), sc = 0.5 * (ssum.i + sdif.i);
      gram[slot][tri(2 * ea, 2 * fb)] = cc;
      gram[slot][tri(2 * ea + 1, 2 * fb + 1)] = ss;
      gram[slot][tri(2 * ea + 1, 2 * fb)] = sc;
      if (ea != fb) gram[slot][tri(2 * ea, 2 * fb + 1)] = cs;
    }
  }
  __syncthreads();
  if (!active) return;
  double* G2 = gram[slot];
  double* atb = vec[slot][0];
  if (g == 0 && B.diag && B.d_ata) {
    float* da = B.d_ata + (size_t)c * nc2 * nc2;
    for (int i = 0; i < nc2; i++)
      for (int j = 0; j < nc2; j++) da[i * nc2 + j] = (j <= i) ? (float)G2[tri(i, j)] : 0.0f;
  }
  if (g == 0 && B.diag && B.d_atb) {
    float* db = B.d_atb + (size_t)c * nc2;
    for (int i = 0; i < nc2; i++) db[i] = (float)atb[i];
  }
  if (row[P->info_scoring] != 0.0) return;
  solve_and_score_grp<G, NCP>(P, B, c, u, row, G2, atb, vec[slot][1], vec[slot][2], vec[slot][3], vec[slot][4],
                         vec[slot][5], itr[slot], g, 2 * u);
}

template <int NB>
__global__ __launch_bounds__(64) void k_fitcf(const ProblemDev* __restrict__ P, StageBufs B) {
  k_fitcf_body<NB>(P, B, blockIdx.x, gridDim.x, blockIdx.y, blockIdx.z);
}

template <int NB, bool LOOP>
__global__ __launch_bounds__(64) void k_fitcf_batch(BatchTab T) {
  batch_blocks<LOOP>(T, [&](const BatchSlot& q) { k_fitcf_body<NB>(T.probs[q.p], T.bufs[q.p], q.lb, q.nb, 0, 0); });
}


// ------------------------------------------------------------------------------------------------
// Margin sweep, mixed chain.  One wave per candidate; lane l owns the contiguous frequency chunk
// [l*KM, l*KM + KM), KM = ceil(F / 64) <= KMAX.  Chunk-wise scans combine in-lane sequential
// passes with a wave prefix over the lanes' chunk totals.
constexpr int KMAX = 8;  // F <= 512

__device__ __forceinline__ bool near_odd_pi(float ph, float tol) {
  const float x = ph + (float)PI_D;
  const float k = __builtin_rintf(x / (float)TWO_PI_D);
  return __builtin_fabsf(x - k * (float)TWO_PI_D) < tol;
}

template <typename V>
__device__ __forceinline__ V wave_excl_sum(V v, int lane) {
  V incl = v;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const V u = __shfl_up(incl, d);
    if (lane >= d) incl = incl + u;
  }
  return incl - v;
}

// Candidate c has too many points to refine: append it to the list of candidates whose every
// frequency is recomputed in FP64 (one entry per candidate).
__device__ __forceinline__ void sweep_fallback(const StageBufs& B, int c, int lane) {
  if (lane == 0) {
    const int pos = atomicAdd(B.fcount, 1);
    B.flist[pos] = c;
    atomicOr(&B.flags[c], F_SWEEP_FP64);
  }
}

// Located sweep of every frequency, then the points whose values show (or come within a guard
// band of) a gain or phase crossing, a non-finite value, or the smallest b(w) (and those within a
// band of it) are marked for FP64 refinement and appended to the dense refinement list.  The guard
// bands are widened by each point's error estimate.  A candidate with more than RMAX marked
// points has its whole sweep recomputed in FP64 instead.
template <int NS>
__device__ __forceinline__ void k_sweep_mark32_body(const ProblemDev* __restrict__ P, const StageBufs& B, int bx, int nbx, int by, int bz) {
  const int c = bx;
  if (margin_skip_flags(B, c)) return;
  const int F = P->F, lane = threadIdx.x;
  const int KM = (F + 63) / 64;
  const int w0 = lane * KM;
  const double* oa = B.open_a + (size_t)c * NS * NS;
  const float* gg = B.gains + (size_t)c * P->cuda_w;
  bool failp = false;
  // the located values in the frequency tables' lane-interleaved order (point k of this lane at
  // slot k * 64 + lane, one 16-byte store per point; the 64 lanes' stores of one k are consecutive)
  float4* mr = reinterpret_cast<float4*>(B.mr32) + (size_t)c * KM * 64 + lane;
  const int nk = (w0 < F) ? ((F - w0 < KM) ? F - w0 : KM) : 0;  // points of this lane
  // The Hessenberg form is the same for every lane: it is reduced and kept in LDS (H, Q^T K,
  // Q^T L) and read by broadcast, which leaves the registers to the point solves.
  __shared__ double sh[NS * NS + 2 * NS];
  double hnorm;
  __syncthreads();  // a batched launch reuses sh for its next candidate
  hess_prep_lds<NS>(oa, gg, sh, hnorm);
  // point k of this lane: slot k * 64 + lane of the lane-interleaved table
  const SweepPointIl* spt = tab<SweepPointIl>(P, P->off_sweep_il) + lane;
  const float alpha = (float)P->alpha;
  // the sweep points go to memory and are read back below, so that nothing else is live across
  // the register-heavy elimination; each point's table values are loaded one point ahead
  SweepPointIl nq{};
  if (nk > 0) nq = spt[0];
#pragma unroll 1
  for (int k = 0; k < nk; k++) {
    // keeps the LDS reads of the Hessenberg data inside the loop (not hoisted into registers)
    __asm__ volatile("" ::: "memory");
    const double zr = nq.zm1, zi = nq.zi;
    const float pr = nq.pr, pim = nq.pi, pa = nq.pabs;
    const int ok = nq.ok;
    if (k + 1 < nk) nq = spt[(k + 1) * 64];
    float mm = __builtin_nanf(""), rr = __builtin_nanf(""), b2 = __builtin_nanf(""), e2 = 1.0f;
    if (!ok) {
      failp = true;
    } else {
      float m1, r1, bq, e1;
      if (hess_point<NS>(sh, hnorm, zr, zi, pr, pim, pa, alpha, m1, r1, bq, e1)) { mm = m1; rr = r1; b2 = bq; e2 = e1; }
    }
    mr[k * 64] = make_float4(mm, rr, b2, e2);
  }
  if (__ballot(failp)) {
    if (lane == 0) B.sweep_fail[c] = 1;
    return;
  }
  float m[KMAX], raw[KMAX], bb[KMAX], er[KMAX];
#pragma unroll
  for (int k = 0; k < KMAX; k++) {
    m[k] = raw[k] = bb[k] = __builtin_nanf("");
    er[k] = 1.0f;
    if (k < nk) {
      const float4 v = mr[k * 64];
      m[k] = v.x;
      raw[k] = v.y;
      bb[k] = v.z;
      er[k] = v.w;
    }
  }
  const float PIf = (float)PI_D, TPf = (float)TWO_PI_D;
  // unwrapped float phase: ph[w] = raw[0] + sum_{v=1..w} wrap(raw[v] - raw[v-1])
  float lastRaw = raw[0], lastM = m[0], lastEr = er[0];
#pragma unroll
  for (int k = 1; k < KMAX; k++)
    if (k < nk) { lastRaw = raw[k]; lastM = m[k]; lastEr = er[k]; }
  const float prevRaw0 = __shfl_up(lastRaw, 1);
  const float prevM0 = __shfl_up(lastM, 1);
  const float prevEr0 = __shfl_up(lastEr, 1);
  float dl[KMAX];
  float loc = 0.0f;
#pragma unroll
  for (int k = 0; k < KMAX; k++) {
    dl[k] = 0.0f;
    const int w = w0 + k;
    if (k < nk && w > 0) {
      float d = raw[k] - (k == 0 ? prevRaw0 : raw[k > 0 ? k - 1 : 0]);
      if (d > PIf) d -= TPf;
      else if (d < -PIf) d += TPf;
      dl[k] = d;
      loc = loc + d;
    }
  }
  const float off = wave_excl_sum(loc, lane) + __shfl(raw[0], 0);
  float ph[KMAX];
  float run = off;
#pragma unroll
  for (int k = 0; k < KMAX; k++) {
    run = run + dl[k];
    ph[k] = run;
  }
  float lastPh = ph[0];
#pragma unroll
  for (int k = 1; k < KMAX; k++)
    if (k < nk) lastPh = ph[k];
  const float prevPh0 = __shfl_up(lastPh, 1);
  // crossing marks: interval (w-1, w) flagged -> mark both ends.  Each guard is widened by the
  // error estimate of the points involved (relative in |loop|, radians in the phase).
  unsigned cross = 0;
  bool send_prev = false;
#pragma unroll
  for (int k = 0; k < KMAX; k++) {
    if (k >= nk) continue;
    const int w = w0 + k;
    const bool fin = __builtin_isfinite(m[k]) && __builtin_isfinite(raw[k]);
    if (w == 0) {
      if (!fin) cross |= 1u << k;
      continue;
    }
    const float pm_ = (k == 0) ? prevM0 : m[k > 0 ? k - 1 : 0];
    const float pr_ = (k == 0) ? prevRaw0 : raw[k > 0 ? k - 1 : 0];
    const float pp_ = (k == 0) ? prevPh0 : ph[k > 0 ? k - 1 : 0];
    const float pe_ = (k == 0) ? prevEr0 : er[k > 0 ? k - 1 : 0];
    const float delta = raw[k] - pr_;
    const bool amb = __builtin_fabsf(__builtin_fabsf(delta) - PIf) < GUARD_PHASE + pe_ + er[k];
    const bool gcross = (pm_ - 1.0f) * (m[k] - 1.0f) <= 0.0f;
    const bool gnear = __builtin_fabsf(pm_ - 1.0f) < GUARD_MAG + pe_ * pm_ ||
                       __builtin_fabsf(m[k] - 1.0f) < GUARD_MAG + er[k] * m[k];
    const float ka = __builtin_floorf((pp_ + PIf) / TPf), kb = __builtin_floorf((ph[k] + PIf) / TPf);
    const bool pcross = ka != kb;
    const bool pnear = near_odd_pi(pp_, GUARD_PHASE + pe_) || near_odd_pi(ph[k], GUARD_PHASE + er[k]);
    if (gcross || gnear || pcross || pnear || amb || !fin) {
      cross |= 1u << k;
      if (k == 0) send_prev = true;
      else cross |= 1u << (k - 1);
    }
  }
  const bool recv = __shfl_down(send_prev ? 1 : 0, 1) != 0 && lane < 63;
  if (recv && nk > 0) cross |= 1u << (nk - 1);
  // smallest b (first index on ties; NaN never wins) and the band around it
  float bv = __builtin_inff();
  int bi = 0x7fffffff;
#pragma unroll
  for (int k = 0; k < KMAX; k++)
    if (k < nk && bb[k] < bv) { bv = bb[k]; bi = w0 + k; }
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) {
    const float ov = __shfl_xor(bv, d);
    const int oi = __shfl_xor(bi, d);
    if (ov < bv || (ov == bv && oi < bi)) { bv = ov; bi = oi; }
  }
  // band: every point whose lower bound on b reaches the smallest upper bound.  With the loop
  // value L = P G off by a relative e, |1 + L| moves by at most e |L| and each square root of the
  // denominator by a relative e, and the denominator is at least 1 + |L| (Cauchy-Schwarz), so
  // |db| <= e (b + |L| / (1 + |L|)).  A point outside the band cannot hold the smallest b, and its
  // float value cannot undercut it; the band (with the smallest point) is refined in FP64, so the
  // minimum is the FP64 one.
  float eb[KMAX];
  float hi = __builtin_inff(), lo = __builtin_inff();
#pragma unroll
  for (int k = 0; k < KMAX; k++) {
    eb[k] = er[k] * (bb[k] + m[k] / (1.0f + m[k]));
    if (k < nk && bb[k] + eb[k] < hi) hi = bb[k] + eb[k];
    if (k < nk && bb[k] - eb[k] < lo) lo = bb[k] - eb[k];
  }
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) {
    hi = __builtin_fminf(hi, __shfl_xor(hi, d));
    lo = __builtin_fminf(lo, __shfl_xor(lo, d));
  }
  unsigned band = 0;
  if (bi != 0x7fffffff) {
    // margin for the rounding of b itself from L (a few float ulps)
    const float lim = hi * (1.0f + 1e-5f);
#pragma unroll
    for (int k = 0; k < KMAX; k++)
      if (k < nk && bb[k] - eb[k] <= lim) band |= 1u << k;
    if (bi >= w0 && bi < w0 + nk) cross |= 1u << (bi - w0);
  }
  // More than RMAX marks: a band over a flat minimum (all its points within BAND_SPREAD of one
  // another, so any of them is the minimum to within that) is dropped but for the smallest float
  // point; otherwise the whole sweep is recomputed in FP64.
  unsigned keep = cross | band;
  int nkeep = __popc(keep);
  int pre = wave_excl_sum(nkeep, lane);
  int tot = __shfl(pre + nkeep, 63);
  if (tot > RMAX && hi - lo <= BAND_SPREAD) {
    keep = cross;
    nkeep = __popc(keep);
    pre = wave_excl_sum(nkeep, lane);
    tot = __shfl(pre + nkeep, 63);
  }
  if (tot > RMAX) {
    sweep_fallback(B, c, lane);
    return;
  }
  // refinement list (one atomic per wave) and the per-point state
  int base = 0;
  if (lane == 0 && tot > 0) base = atomicAdd(B.rcount, tot);
  base = __shfl(base, 0);
  int8_t* rf = B.refined + (size_t)c * F;
  int r = 0;
#pragma unroll
  for (int k = 0; k < KMAX; k++) {
    if (k >= nk) continue;
    const bool mk = (keep >> k) & 1u;
    rf[w0 + k] = mk ? 1 : 0;
    if (mk) B.rlist[base + pre + (r++)] = make_int2(c, w0 + k);
  }
}

template <int NS>
// at least 3 waves per SIMD: the order-8 point solve otherwise holds 182 VGPRs (2 waves), with
// its LDS reads of the Hessenberg data issued far ahead
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(3))) void k_sweep_mark32(const ProblemDev* __restrict__ P, StageBufs B) {
  k_sweep_mark32_body<NS>(P, B, blockIdx.x, gridDim.x, blockIdx.y, blockIdx.z);
}

template <int NS, bool LOOP>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(3))) void k_sweep_mark32_batch(BatchTab T) {
  batch_blocks<LOOP>(T, [&](const BatchSlot& q) { k_sweep_mark32_body<NS>(T.probs[q.p], T.bufs[q.p], q.lb, q.nb, 0, 0); });
}


template <int NS>
__device__ __forceinline__ void k_refine64_body(const ProblemDev* __restrict__ P, const StageBufs& B, int bx, int nbx, int by, int bz) {
  const int total = *B.rcount;
  const int F = P->F;
  for (int i = bx * 64 + threadIdx.x; i < total; i += nbx * 64) {
    const int2 e = B.rlist[i];
    const int c = e.x, w = e.y;
    double m, raw, b;
    const bool ok = sweep_point<NS, double>(P, B.open_a + (size_t)c * NS * NS, B.gains + (size_t)c * P->cuda_w, w, m, raw, b);
    double* out = B.mag_raw + ((size_t)c * F + w) * 3;
    out[0] = m;
    out[1] = raw;
    out[2] = b;
    B.refined[(size_t)c * F + w] = ok ? 2 : 3;
  }
}

// Every frequency of the listed candidates in FP64 (the FP64 chain's sweep point), one wave per
// candidate, lanes over frequencies.
template <int NS>
__device__ __forceinline__ void k_full64_body(const ProblemDev* __restrict__ P, const StageBufs& B, int bx, int nbx, int by, int bz) {
  const int total = *B.fcount;
  const int F = P->F;
  for (int i = bx; i < total; i += nbx) {
    const int c = B.flist[i];
    const double* oa = B.open_a + (size_t)c * NS * NS;
    const float* gg = B.gains + (size_t)c * P->cuda_w;
    for (int w = threadIdx.x; w < F; w += 64) {
      double m, raw, b;
      const bool ok = sweep_point<NS, double>(P, oa, gg, w, m, raw, b);
      double* out = B.mag_raw + ((size_t)c * F + w) * 3;
      out[0] = m;
      out[1] = raw;
      out[2] = b;
      B.refined[(size_t)c * F + w] = ok ? 2 : 3;
    }
  }
}

template <int NS, bool LOOP>
__global__ __launch_bounds__(64) void k_full64_batch(BatchTab T) {
  batch_blocks<LOOP>(T, [&](const BatchSlot& q) { k_full64_body<NS>(T.probs[q.p], T.bufs[q.p], q.lb, q.nb, 0, 0); });
}

// The point refinement (first REFINE_BLOCKS workgroups) and the whole-candidate FP64 sweeps (the
// rest) in one launch: they write disjoint candidates' points.
constexpr int REFINE_BLOCKS = 1024, FULL64_BLOCKS = 256;
template <int NS>
__global__ __launch_bounds__(64) void k_refine_full64(const ProblemDev* __restrict__ P, StageBufs B) {
  const int bx = blockIdx.x;
  if (bx < REFINE_BLOCKS) k_refine64_body<NS>(P, B, bx, REFINE_BLOCKS, 0, 0);
  else k_full64_body<NS>(P, B, bx - REFINE_BLOCKS, FULL64_BLOCKS, 0, 0);
}

template <int NS, bool LOOP>
__global__ __launch_bounds__(64) void k_refine64_batch(BatchTab T) {
  batch_blocks<LOOP>(T, [&](const BatchSlot& q) { k_refine64_body<NS>(T.probs[q.p], T.bufs[q.p], q.lb, q.nb, 0, 0); });
}


// Margin error and total score from the scan's minima (src/allmargin.cu:182-196).
__device__ __forceinline__ void margin_score(const ProblemDev* __restrict__ P, double* sc, double* am, double GM,
                                             double PM, double DM, double Bmin, bool crossed) {
  const double EF = P->eig_fail_err;
  const double err_gm = cedm::fmax_(0.0, P->target_gm - GM) / P->target_gm;
  const double err_pm = crossed ? cedm::fmax_(0.0, P->target_pm - PM) / P->target_pm : 0.0;
  const double err_b = cedm::fmax_(0.0, P->target_b - Bmin) / P->target_b;
  const double margin = (err_gm * err_gm + err_pm * err_pm) + err_b * err_b;
  sc[1] = sc[0];
  sc[2] = margin;
  am[0] = GM; am[1] = PM; am[2] = DM; am[3] = 1.0; am[4] = Bmin;
  const double total = cedm::log((sc[0] + margin) + EPS30);
  sc[0] = cedm::isfinite_(total) ? total : EF;
}

// Crossing scan (src/allmargin.cu:119-180) over the float values with the FP64 refinements in
// place, one wave per candidate: each interval's contributions are independent once the phase is
// unwrapped (chunk-wise prefix in double); the minima and flags reduce over the wave.
// SPLIT: the scan runs beside the fit (second stream), so the behaviour score is not final yet: it
// only stores its minima in B.mres, for k_margin_combine_final to apply after the fit.
template <bool SPLIT>
__device__ __forceinline__ void k_scan32_body(const ProblemDev* __restrict__ P, const StageBufs& B, int bx, int nbx, int by, int bz) {
  const int c = bx;
  const int lane = threadIdx.x;
  double* row = B.info + (size_t)c * P->info_w;
  double* sc = row + P->info_scoring;
  double* am = row + P->info_allmargin;
  const double UE = P->unstable_err, EF = P->eig_fail_err;
  const int flags = B.flags[c];
  if (SPLIT && margin_skip_flags(B, c)) return;  // not swept; the combine step writes its row
  if (!SPLIT && sc[0] >= UE) {
    if (lane == 0) {
      sc[1] = sc[0]; sc[2] = sc[0];
      am[0] = am[1] = am[2] = am[3] = am[4] = 0.0;
    }
    return;
  }
  if (!SPLIT && (flags & F_NOMINAL_UNSTABLE)) {
    if (lane == 0) {
      sc[1] = sc[0];
      sc[2] = UE;
      am[0] = am[1] = am[2] = am[3] = am[4] = 0.0;
      const double total = cedm::log((sc[0] + UE) + EPS30);
      sc[0] = cedm::isfinite_(total) ? total : EF;
    }
    return;
  }
  const int F = P->F;
  const int KM = (F + 63) / 64;
  const int w0 = lane * KM;
  const int nk = (w0 < F) ? ((F - w0 < KM) ? F - w0 : KM) : 0;
  bool fail = B.sweep_fail[c] != 0;
  double m[KMAX], raw[KMAX], bb[KMAX];
  bool failp = false;
  if (!fail) {
    const float4* mr = reinterpret_cast<const float4*>(B.mr32) + (size_t)c * KM * 64 + lane;
    const double* m64 = B.mag_raw + (size_t)c * F * 3;
    const int8_t* rf = B.refined + (size_t)c * F;
    // the lane's float values and refinement marks are loaded together, ahead of their use
    float fv[KMAX][3];
    int rv[KMAX];
#pragma unroll
    for (int k = 0; k < KMAX; k++) {
      rv[k] = 0;
      fv[k][0] = fv[k][1] = fv[k][2] = 0.0f;
      if (k < nk) {
        const int w = w0 + k;
        rv[k] = rf[w];
        const float4 v = mr[k * 64];
        fv[k][0] = v.x;
        fv[k][1] = v.y;
        fv[k][2] = v.z;
      }
    }
#pragma unroll
    for (int k = 0; k < KMAX; k++) {
      m[k] = raw[k] = bb[k] = 0.0;
      if (k < nk) { m[k] = (double)fv[k][0]; raw[k] = (double)fv[k][1]; bb[k] = (double)fv[k][2]; }
    }
    // the FP64 values of the refined points: every load is issued before any is used, so a wave
    // waits once for all of them
#pragma unroll
    for (int k = 0; k < KMAX; k++) {
      if (k < nk && rv[k] == 2) {
        const int w = w0 + k;
        m[k] = m64[3 * w];
        raw[k] = m64[3 * w + 1];
        bb[k] = m64[3 * w + 2];
      }
    }
#pragma unroll
    for (int k = 0; k < KMAX; k++) {
      if (k >= nk) continue;
      if (rv[k] == 3) failp = true;
      if (!(cedm::isfinite_(m[k]) && cedm::isfinite_(raw[k]))) failp = true;
    }
  }
  if (fail || __ballot(failp)) {
    if (lane == 0) {
      if (SPLIT) {
        B.mres[(size_t)c * 6 + 5] = 1.0;
      } else {
        sc[1] = sc[0]; sc[2] = EF; sc[0] = EF;
        am[0] = 0.0; am[1] = 0.0; am[2] = 0.0; am[3] = 1.0; am[4] = 0.0;
      }
    }
    return;
  }
  const double ts = P->ts;
  // grid angles, lane-interleaved: theta of point k of this lane at slot k * 64 + lane
  const double* thi = tab<double>(P, P->off_theta_il);
  double Bmin = 1.0;
  double lastRaw = raw[0], lastM = m[0];
#pragma unroll
  for (int k = 0; k < KMAX; k++)
    if (k < nk) {
      Bmin = cedm::fmin_(Bmin, bb[k]);
      lastRaw = raw[k];
      lastM = m[k];
    }
  const double prevRaw0 = __shfl_up(lastRaw, 1);
  const double prevM0 = __shfl_up(lastM, 1);
  double dl[KMAX];
  double loc = 0.0;
#pragma unroll
  for (int k = 0; k < KMAX; k++) {
    dl[k] = 0.0;
    const int w = w0 + k;
    if (k < nk && w > 0) {
      double d = raw[k] - (k == 0 ? prevRaw0 : raw[k > 0 ? k - 1 : 0]);
      if (d > PI_D) d = d - TWO_PI_D;
      else if (d < -PI_D) d = d + TWO_PI_D;
      dl[k] = d;
      loc = loc + d;
    }
  }
  const double off = wave_excl_sum(loc, lane) + __shfl(raw[0], 0);
  double ph[KMAX];
  double run = off;
#pragma unroll
  for (int k = 0; k < KMAX; k++) {
    run = run + dl[k];
    ph[k] = run;
  }
  double lastPh = ph[0];
#pragma unroll
  for (int k = 1; k < KMAX; k++)
    if (k < nk) lastPh = ph[k];
  const double prevPh0 = __shfl_up(lastPh, 1);
  double PM = FLTMAX, GM = FLTMAX, DM = FLTMAX;
  bool crossed = false;
#pragma unroll
  for (int k = 0; k < KMAX; k++) {
    if (k >= nk) continue;
    const int w = w0 + k;
    if (w == 0) continue;
    const double prevMag = (k == 0) ? prevM0 : m[k > 0 ? k - 1 : 0];
    const double prevPhase = (k == 0) ? prevPh0 : ph[k > 0 ? k - 1 : 0];
    // theta_{w-1}: the previous point of this lane, or (k = 0, lane > 0) the last of the lane below
    const double prevTheta = (k == 0) ? thi[(KM - 1) * 64 + lane - 1] : thi[(k - 1) * 64 + lane];
    const double theta_w = thi[k * 64 + lane];
    const double mk = m[k], phk = ph[k];
    if ((prevMag - 1.0) * (mk - 1.0) <= 0.0 && mk != prevMag) {
      const double frac = (1.0 - prevMag) / (mk - prevMag);
      const double phase_c = prevPhase + frac * (phk - prevPhase);
      crossed = true;
      double pm_deg = 180.0 + phase_c * RAD2DEG;
      pm_deg = pm_deg - 360.0 * cedm::floor_((pm_deg + 180.0) / 360.0);
      PM = cedm::fmin_(PM, pm_deg);
      const double theta_c = prevTheta + frac * (theta_w - prevTheta);
      const double pm_rad = pm_deg * DEG2RAD;
      if (theta_c > EPS6 && pm_rad >= 0.0) DM = cedm::fmin_(DM, ((pm_rad / theta_c) * ts) * 1000.0);
    }
    const double a = prevPhase + PI_D, b2 = phk + PI_D;
    const double ka = cedm::floor_(a / TWO_PI_D), kb = cedm::floor_(b2 / TWO_PI_D);
    if (ka != kb && b2 != a) {
      const double target = TWO_PI_D * (b2 > a ? kb : ka);
      const double frac = (target - a) / (b2 - a);
      const double mag_p = prevMag + frac * (mk - prevMag);
      GM = cedm::fmin_(GM, -20.0 * cedm::log10(mag_p));
    }
  }
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) {
    PM = cedm::fmin_(PM, __shfl_xor(PM, d));
    GM = cedm::fmin_(GM, __shfl_xor(GM, d));
    DM = cedm::fmin_(DM, __shfl_xor(DM, d));
    Bmin = cedm::fmin_(Bmin, __shfl_xor(Bmin, d));
  }
  crossed = __ballot(crossed) != 0ull;
  if (lane != 0) return;
  if (SPLIT) {
    double* r = B.mres + (size_t)c * 6;
    r[0] = GM; r[1] = PM; r[2] = DM; r[3] = Bmin; r[4] = crossed ? 1.0 : 0.0; r[5] = 0.0;
  } else {
    margin_score(P, sc, am, GM, PM, DM, Bmin, crossed);
  }
}

// After the fit: the row updates of the crossing scan from its stored minima (lane per candidate);
// the same branches, in the same order, as the unsplit scan.
__device__ __forceinline__ void margin_combine_row(const ProblemDev* __restrict__ P, const StageBufs& B, int c) {
  double* row = B.info + (size_t)c * P->info_w;
  double* sc = row + P->info_scoring;
  double* am = row + P->info_allmargin;
  const double UE = P->unstable_err, EF = P->eig_fail_err;
  if (sc[0] >= UE) {
    sc[1] = sc[0]; sc[2] = sc[0];
    am[0] = am[1] = am[2] = am[3] = am[4] = 0.0;
    return;
  }
  if (B.flags[c] & F_NOMINAL_UNSTABLE) {
    sc[1] = sc[0];
    sc[2] = UE;
    am[0] = am[1] = am[2] = am[3] = am[4] = 0.0;
    const double total = cedm::log((sc[0] + UE) + EPS30);
    sc[0] = cedm::isfinite_(total) ? total : EF;
    return;
  }
  const double* r = B.mres + (size_t)c * 6;
  if (r[5] != 0.0) {
    sc[1] = sc[0]; sc[2] = EF; sc[0] = EF;
    am[0] = 0.0; am[1] = 0.0; am[2] = 0.0; am[3] = 1.0; am[4] = 0.0;
    return;
  }
  margin_score(P, sc, am, r[0], r[1], r[2], r[3], r[4] != 0.0);
}

// The row updates, then the float copy of the workgroup's rows (the INFO rows the caller gets):
// one launch at the end of the two-stream chain.
__global__ __launch_bounds__(256) void k_margin_combine_final(const ProblemDev* __restrict__ P, StageBufs B,
                                                              float* __restrict__ out) {
  const int c0 = blockIdx.x * 256;
  const int c = c0 + threadIdx.x;
  if (c < B.n) margin_combine_row(P, B, c);
  __syncthreads();
  const size_t w = (size_t)P->info_w;
  const size_t e0 = (size_t)c0 * w, e1 = (size_t)(c0 + 256 < B.n ? c0 + 256 : B.n) * w;
  for (size_t i = e0 + threadIdx.x; i < e1; i += 256) out[i] = (float)B.info[i];
}

__global__ __launch_bounds__(64) void k_scan32(const ProblemDev* __restrict__ P, StageBufs B) {
  k_scan32_body<false>(P, B, blockIdx.x, gridDim.x, blockIdx.y, blockIdx.z);
}

__global__ __launch_bounds__(64) void k_scan32_split(const ProblemDev* __restrict__ P, StageBufs B) {
  k_scan32_body<true>(P, B, blockIdx.x, gridDim.x, blockIdx.y, blockIdx.z);
}

template <bool LOOP>
__global__ __launch_bounds__(64) void k_scan32_batch(BatchTab T) {
  batch_blocks<LOOP>(T, [&](const BatchSlot& q) { k_scan32_body<false>(T.probs[q.p], T.bufs[q.p], q.lb, q.nb, 0, 0); });
}


// ------------------------------------------------------------------------------------------------
bool launch_sim_mixed(int ns, int nd, const ProblemDev* P, const StageBufs& B, hipStream_t st) {
  return launch_sim<float>(ns, nd, P, B, B.ybuf32, st);
}

bool launch_fit_mixed(int nt, const ProblemDev* P, const StageBufs& B, hipStream_t st) {
  const int nb = (2 * nt + 7) / 8;
  hipLaunchKernelGGL(k_atb, dim3((B.n + 63) / 64, (nt + ATB_MODES - 1) / ATB_MODES, B.atb_parts), dim3(64), 0, st, P, B);
  kmark("k_atb", st);
  const dim3 grid((B.n + 3) / 4);
  switch (nb) {
    case 1: hipLaunchKernelGGL(k_fitcf<1>, grid, dim3(64), 0, st, P, B); kmark("k_fitcf", st); return true;
    case 2: hipLaunchKernelGGL(k_fitcf<2>, grid, dim3(64), 0, st, P, B); kmark("k_fitcf", st); return true;
    case 3: hipLaunchKernelGGL(k_fitcf<3>, grid, dim3(64), 0, st, P, B); kmark("k_fitcf", st); return true;
    case 4: hipLaunchKernelGGL(k_fitcf<4>, grid, dim3(64), 0, st, P, B); kmark("k_fitcf", st); return true;
  }
  return false;
}

template <int NS>
static void launch_mark_refine_t(const ProblemDev* P, const StageBufs& B, hipStream_t st) {
  hipLaunchKernelGGL(k_sweep_mark32<NS>, dim3(B.n), dim3(64), 0, st, P, B);
  kmark("k_sweep_mark32", st);
  hipLaunchKernelGGL(k_refine_full64<NS>, dim3(REFINE_BLOCKS + FULL64_BLOCKS), dim3(64), 0, st, P, B);
  kmark("k_refine_full64", st);
}

// located sweep + marking + FP64 refinement (independent of the simulation and the fit)
bool launch_mark_refine(int ns, int F, const ProblemDev* P, const StageBufs& B, hipStream_t st) {
  if (F > 64 * KMAX) return false;
  switch (ns) {
    case 1: launch_mark_refine_t<1>(P, B, st); return true;
    case 2: launch_mark_refine_t<2>(P, B, st); return true;
    case 3: launch_mark_refine_t<3>(P, B, st); return true;
    case 4: launch_mark_refine_t<4>(P, B, st); return true;
    case 5: launch_mark_refine_t<5>(P, B, st); return true;
    case 6: launch_mark_refine_t<6>(P, B, st); return true;
    case 7: launch_mark_refine_t<7>(P, B, st); return true;
    case 8: launch_mark_refine_t<8>(P, B, st); return true;
  }
  return false;
}

// Bounded wait on the second stream for the simulation's signal (the gated form's start, the
// ungated form's first finished wave): lane 0 polls *gate (agent-scope loads) until it reads gen,
// for at most `limit` polls, so the kernel always ends.  Returning early only lets the following
// kernels start sooner; nothing depends on it for correctness.
__global__ __launch_bounds__(64) void k_gate_wait(const int* __restrict__ gate, int gen, int limit) {
  if (threadIdx.x != 0) return;
  for (int i = 0; i < limit; i++) {
    if (__hip_atomic_load(gate, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == gen) break;
    __builtin_amdgcn_s_sleep(8);
  }
}

void launch_gate_wait(const int* gate, int gen, int limit, hipStream_t st) {
  hipLaunchKernelGGL(k_gate_wait, dim3(1), dim3(64), 0, st, gate, gen, limit);
}

// crossing scan and total score (after the fit and the refinement)
void launch_scan32(const ProblemDev* P, const StageBufs& B, hipStream_t st) {
  hipLaunchKernelGGL(k_scan32, dim3(B.n), dim3(64), 0, st, P, B);
  kmark("k_scan32", st);
}

// the crossing scan beside the fit (after the refinement), then the row updates after the fit
void launch_scan32_split(const ProblemDev* P, const StageBufs& B, hipStream_t st) {
  hipLaunchKernelGGL(k_scan32_split, dim3(B.n), dim3(64), 0, st, P, B);
  kmark("k_scan32", st);
}
void launch_margin_combine_final(const ProblemDev* P, const StageBufs& B, float* info_out, hipStream_t st) {
  hipLaunchKernelGGL(k_margin_combine_final, dim3((B.n + 255) / 256), dim3(256), 0, st, P, B, info_out);
  kmark("k_margin_combine_final", st);
}

// A grid of a workgroup per block launches the kernels' loop-free form (fewer registers).
template <bool LOOP>
static bool launch_batch_mixed_t(int kind, int key, const BatchTab& T, const dim3 g, hipStream_t st) {
  const dim3 b(64);
#define CE_NS_CASES(K)                                                      \
  switch (key) {                                                            \
    case 1: hipLaunchKernelGGL((K<1, LOOP>), g, b, 0, st, T); return true;  \
    case 2: hipLaunchKernelGGL((K<2, LOOP>), g, b, 0, st, T); return true;  \
    case 3: hipLaunchKernelGGL((K<3, LOOP>), g, b, 0, st, T); return true;  \
    case 4: hipLaunchKernelGGL((K<4, LOOP>), g, b, 0, st, T); return true;  \
    case 5: hipLaunchKernelGGL((K<5, LOOP>), g, b, 0, st, T); return true;  \
    case 6: hipLaunchKernelGGL((K<6, LOOP>), g, b, 0, st, T); return true;  \
    case 7: hipLaunchKernelGGL((K<7, LOOP>), g, b, 0, st, T); return true;  \
    case 8: hipLaunchKernelGGL((K<8, LOOP>), g, b, 0, st, T); return true;  \
  }                                                                         \
  return false;
  switch (kind) {
    case BK_ATB: hipLaunchKernelGGL(k_atb_batch<LOOP>, g, b, 0, st, T); return true;
    case BK_FITCF:
      switch (key) {
        case 1: hipLaunchKernelGGL((k_fitcf_batch<1, LOOP>), g, b, 0, st, T); return true;
        case 2: hipLaunchKernelGGL((k_fitcf_batch<2, LOOP>), g, b, 0, st, T); return true;
        case 3: hipLaunchKernelGGL((k_fitcf_batch<3, LOOP>), g, b, 0, st, T); return true;
        case 4: hipLaunchKernelGGL((k_fitcf_batch<4, LOOP>), g, b, 0, st, T); return true;
      }
      return false;
    case BK_SWEEP_MARK32: CE_NS_CASES(k_sweep_mark32_batch)
    case BK_REFINE64: CE_NS_CASES(k_refine64_batch)
    case BK_FULL64: CE_NS_CASES(k_full64_batch)
    case BK_SCAN32: hipLaunchKernelGGL(k_scan32_batch<LOOP>, g, b, 0, st, T); return true;
  }
  return false;
#undef CE_NS_CASES
}

bool launch_batch_mixed(int kind, int key, const BatchTab& T, int grid, hipStream_t st) {
  return grid >= T.total ? launch_batch_mixed_t<false>(kind, key, T, dim3(grid), st)
                         : launch_batch_mixed_t<true>(kind, key, T, dim3(grid), st);
}

bool launch_sweep_fp64(int ns, int F, const ProblemDev* P, const StageBufs& B, hipStream_t st);

// The located float sweep holds KMAX points per lane of a wave; a longer grid is swept by the FP64
// chain's kernels (every point in FP64, the same rows as the FP64 chain's sweep).
bool mixed_sweep_fits(int F) { return F <= 64 * KMAX; }

bool launch_sweep_mixed(int ns, int F, const ProblemDev* P, const StageBufs& B, hipStream_t st) {
  if (!mixed_sweep_fits(F)) return launch_sweep_fp64(ns, F, P, B, st);
  if (!launch_mark_refine(ns, F, P, B, st)) return false;
  launch_scan32(P, B, st);
  return true;
}

}  // namespace ce

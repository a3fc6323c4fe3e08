// FP64 stage kernels for gfx950.
//
// Stage map (reference file -> kernel):
//   eigensetup + eigensolve + modal prelude   src/eigensetup.cu, src/eigensolve.cu, src/modalsetup.cu:8-50
//        -> k_eig_lin<G,R>    G lanes per candidate (R rows each); LIN_A built in LDS, lane-group
//                             real-Schur QR (ce_eig_grp.h), modes, dedupe, stability flags, and the
//                             nominal check when the nominal loop is D LIN_A D
//   allmargin nominal closed loop             src/allmargin.cu:61-79
//        -> k_eig_nom<G,R>    the same solver on the candidates the LIN_A solve could not decide
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

#include "ce_eig_grp.h"
#include "ce_fitsolve_dev.h"
#include "ce_linalg_dev.h"
#include "ce_sim_dev.h"
#include "ce_sweep_dev.h"
#include "ce_stage.h"

namespace ce {

// ------------------------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_info_init(const float* __restrict__ in, double* __restrict__ out, size_t count) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) out[i] = (double)in[i];
}

__global__ __launch_bounds__(256) void k_info_final(const double* __restrict__ in, float* __restrict__ out, size_t count) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) out[i] = (float)in[i];
}

// Compaction, one atomic per wave and list: the candidates whose trajectory is needed (linearly
// stable ones, or all of them when the trajectory dump is requested) get consecutive slots of
// list; those whose nominal loop needs its own eigensolve (not flagged unstable, skipped or
// decided by the LIN_A solve) get consecutive slots of nlist.
__device__ __forceinline__ void compact_append(bool pred, int c, int32_t* count, int32_t* list) {
  const unsigned long long ballot = __ballot(pred);
  const int cnt = __popcll(ballot);
  int base = 0;
  if (threadIdx.x == __ffsll((long long)ballot) - 1 && cnt > 0) base = atomicAdd(count, cnt);
  base = __shfl(base, __ffsll((long long)ballot) - 1);
  if (pred) {
    const unsigned long long below = ballot & ((1ull << threadIdx.x) - 1ull);
    list[base + __popcll(below)] = c;
  }
}

__device__ __forceinline__ void k_compact_body(const ProblemDev* __restrict__ P, const StageBufs& B, int bx, int nbx,
                                               int by, int bz) {
  const bool all = B.d_y != nullptr;
  const int c = bx * 64 + threadIdx.x;
  const int f = c < B.n ? B.flags[c] : F_UNSTABLE;
  compact_append(c < B.n && (all || (f & F_UNSTABLE) == 0), c, B.count, B.list);
  compact_append(c < B.n && (f & (F_UNSTABLE | F_NOMINAL_SKIPPED | F_NOMINAL_DONE)) == 0, c, B.ncount, B.nlist);
}

__global__ __launch_bounds__(64) void k_compact(const ProblemDev* __restrict__ P, StageBufs B) {
  k_compact_body(P, B, blockIdx.x, gridDim.x, 0, 0);
}

template <bool LOOP>
__global__ __launch_bounds__(64) void k_compact_batch(BatchTab T) {
  batch_blocks<LOOP>(T, [&](const BatchSlot& q) {
    k_compact_body(T.probs[q.p], T.bufs[q.p], q.lb, q.nb, 0, 0);
  });
}

// ------------------------------------------------------------------------------------------------
// Eigen stage: LIN_A (reference src/eigensetup.cu:16-56), its eigenvalues, the continuous-time
// modes with the adjacent-duplicate rule (src/modalsetup.cu:8-47), the unstable flag
// (src/modalsetup.cu:26) and the nominal closed loop of the margins stage (src/allmargin.cu:61-79).
//
// G lanes per candidate, 64 / G candidates per single-wave workgroup, R rows per lane (order up to
// G R): lane l builds rows l, l + G, ... of the matrix and owns those columns / rows in the
// eigensolver.  k_eig_lin (G R >= nt) solves LIN_A and derives modes and flags; k_eig_nom
// (G R >= 2 ns) solves the nominal loop of the candidates that need it.
#define CE_ROWS(r) _Pragma("unroll") for (int q_ = 0, r = l; q_ < R; q_++, r += G)
// LDS of a single-wave workgroup of the eigen kernels: CPW matrices of N x (N + 1), then CPW
// vectors of 3 N (a fused kernel passes one array sized for its largest tile)
template <int G, int R>
constexpr int eig_lds_doubles() {
  return (64 / G) * ((G * R) * (G * R + 1) + 3 * G * R);
}

template <int G, int R>
__device__ __forceinline__ void k_eig_lin_body(double* lds, const ProblemDev* __restrict__ P, const StageBufs& B, int bx, int nbx, int by, int bz) {
  constexpr int N = G * R, S = N + 1, CPW = 64 / G;
  const int gi = threadIdx.x / G, l = threadIdx.x % G;
  const int c = bx * CPW + gi;
  if (c >= B.n) return;
  double* T = lds + gi * (N * S);
  double* V = lds + CPW * (N * S) + gi * (3 * N);
  double* WR = V + N;
  double* WI = V + 2 * N;
  // after the solve: |lambda| in V (the reduction's scratch), the continuous-time mode (re, im) of
  // eigenvalue r in WR[r], WI[r] (each lane overwrites only the entries it has read)
  double* MAG = V;
  double* RE = WR;
  double* IM = WI;
  const int ns = P->ns, nd = P->nd, nt = P->nt, n = B.n;
  const float* g = B.gains + (size_t)c * P->cuda_w;
  CE_ROWS(r) {
    if (r < ns) {
      const double Ll = g[ns + r];
      double* oa = B.open_a + (size_t)c * ns * ns + (size_t)r * ns;
      for (int j = 0; j < ns; j++) {
        const double Kj = g[j];
        const double bk = P->sb[r] * Kj;
        const double cdk = P->sc[j] - P->sd * Kj;
        const double lcd = Ll * cdk;
        const double v = (P->sa[r * ns + j] - bk) - lcd;
        T[r * S + j] = v;
        oa[j] = v;
      }
      for (int j = 0; j < nd; j++) T[r * S + ns + j] = Ll * P->dc[j];
    } else if (r < nt) {
      const int i = r - ns;
      for (int j = 0; j < ns; j++) T[r * S + j] = -(P->dbu[i] * (double)g[j]);
      for (int j = 0; j < nd; j++) T[r * S + ns + j] = P->da[i * nd + j];
    }
  }
  if (B.diag && B.d_lina) {
    CE_ROWS(r) if (r < nt) {
      float* dl = B.d_lina + (size_t)c * nt * nt + (size_t)r * nt;
      for (int j = 0; j < nt; j++) dl[j] = (float)T[r * S + j];
    }
  }
  grp_sync();
  int ev = 0;
  const int rc = eig_grp<G, R>(T, V, WR, WI, nt, l, P->nom_equiv ? ns : -1, &ev);
  // When the nominal loop of the margins stage is D LIN_A D entry for entry and this solve had no
  // sign-flip event, the nominal eigensolve would return these eigenvalues and this status: its
  // check (src/allmargin.cu:72-79: unstable when the solve fails or some |lambda| >= 1) is decided
  // here and k_eig_nom skips the candidate.
  const unsigned long long gmask = ((1ull << G) - 1ull) << (gi * G);
  bool nom_out = false;
  CE_ROWS(r) nom_out = nom_out || (rc == 0 && r < nt && cedm::hypot(WR[r], WI[r]) >= 1.0);
  const bool nom_done = P->nom_equiv != 0 && ev == 0;
  const bool nom_bad = rc != 0 || (__ballot(nom_out) & gmask) != 0ull;
  if (rc != 0) CE_ROWS(r) if (r < nt) { WR[r] = cedm::qnan(); WI[r] = cedm::qnan(); }
  grp_sync();
  double re[R], im[R];
  CE_ROWS(r) {
    re[q_] = 0.0;
    im[q_] = 0.0;
    if (r < nt) {
      const double wr = WR[r], wi = WI[r];
      if (B.diag && B.d_eig) {
        float* de = B.d_eig + (size_t)c * 2 * nt;
        de[2 * r] = (float)wr;
        de[2 * r + 1] = (float)wi;
      }
      const double mag = cedm::hypot(wr, wi);
      MAG[r] = mag;
      re[q_] = cedm::log(cedm::fmax_(mag, EPS30)) / P->ts;
      im[q_] = cedm::atan2(wi, wr) / P->ts;
      RE[r] = re[q_];
      IM[r] = im[q_];
    }
  }
  grp_sync();
  // modes in eigenvalue order: round q_ covers indices [q_ G, q_ G + G) in lane order
  int ubase = 0;
  CE_ROWS(r) {
    bool keep = false;
    if (r < nt) {
      const double pre = (r == 0) ? BIG10 : RE[r - 1];
      const double pim = (r == 0) ? BIG10 : IM[r - 1];
      const bool dup = cedm::fabs_(re[q_] - pre) < EPS6 && cedm::fabs_(cedm::fabs_(im[q_]) - cedm::fabs_(pim)) < EPS6;
      keep = !dup;
    }
    const unsigned long long kb = __ballot(keep) & gmask;
    if (keep) {
      const int u = ubase + __popcll(kb & ((1ull << threadIdx.x) - 1ull));
      B.mode_re[(size_t)u * n + c] = re[q_];
      B.mode_im[(size_t)u * n + c] = cedm::fabs_(im[q_]);
    }
    ubase += __popcll(kb);
  }
  if (l == 0) {
    double max_mag = 0.0;
    for (int i = 0; i < nt; i++) max_mag = cedm::fmax_(max_mag, MAG[i]);
    double* sc = B.info + (size_t)c * P->info_w + P->info_scoring;
    int flags = 0;
    if (max_mag > 1.0) {
      sc[0] = P->unstable_err + cedm::log(max_mag) / P->ts;
      sc[1] = sc[0];
      flags |= F_UNSTABLE;
    }
    if ((flags & F_UNSTABLE) || sc[0] >= P->unstable_err) flags |= F_NOMINAL_SKIPPED;
    else if (nom_done) flags |= F_NOMINAL_DONE | (nom_bad ? F_NOMINAL_UNSTABLE : 0);
    B.uniq[c] = ubase;
    B.flags[c] = flags;
  }
}

template <int G, int R>
__global__ __launch_bounds__(64) void k_eig_lin(const ProblemDev* __restrict__ P, StageBufs B) {
  __shared__ double lds[eig_lds_doubles<G, R>()];
  k_eig_lin_body<G, R>(lds, P, B, blockIdx.x, gridDim.x, blockIdx.y, blockIdx.z);
}

template <int G, int R, bool LOOP>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(4))) void k_eig_lin_batch(BatchTab T) {
  __shared__ double lds[eig_lds_doubles<G, R>()];
  batch_blocks<LOOP>(T, [&](const BatchSlot& q) {
    k_eig_lin_body<G, R>(lds, T.probs[q.p], T.bufs[q.p], q.lb, q.nb, 0, 0);
  });
}


// nominal closed loop [[A1 - D L K^T, -L Pc^T], [Pb K^T, Pa]] (src/allmargin.cu:63-79), for the
// candidates of the nominal list (k_compact): linearly stable, not rejected, not decided by the
// LIN_A solve
template <int G, int R>
__device__ __forceinline__ void k_eig_nom_body(double* lds, const ProblemDev* __restrict__ P, const StageBufs& B, int bx, int nbx, int by, int bz) {
  constexpr int N = G * R, S = N + 1, CPW = 64 / G;
  const int gi = threadIdx.x / G, l = threadIdx.x % G;
  // candidates from the nominal list (after k_compact), so that the groups of a wave all solve
  const int idx = bx * CPW + gi;
  if (idx >= *B.ncount) return;
  const int c = B.nlist[idx];
  const int flags = B.flags[c];
  if (flags & F_NOMINAL_SKIPPED) return;
  double* T = lds + gi * (N * S);
  double* V = lds + CPW * (N * S) + gi * (3 * N);
  double* WR = V + N;
  double* WI = V + 2 * N;
  const int ns = P->ns, m = 2 * ns;
  const float* g = B.gains + (size_t)c * P->cuda_w;
  CE_ROWS(r) {
    if (r < ns) {
      const double Ll = g[ns + r];
      const double* oa = B.open_a + (size_t)c * ns * ns + (size_t)r * ns;
      for (int j = 0; j < ns; j++) T[r * S + j] = oa[j] - P->pd * (Ll * (double)g[j]);
      for (int j = 0; j < ns; j++) T[r * S + ns + j] = -(Ll * P->pc[j]);
    } else if (r < m) {
      const int i = r - ns;
      for (int j = 0; j < ns; j++) T[r * S + j] = P->pb[i] * (double)g[j];
      for (int j = 0; j < ns; j++) T[r * S + ns + j] = P->pa[i * ns + j];
    }
  }
  grp_sync();
  const int rc = eig_grp<G, R>(T, V, WR, WI, m, l);
  bool out = false;
  CE_ROWS(r) out = out || (rc == 0 && r < m && cedm::hypot(WR[r], WI[r]) >= 1.0);
  const unsigned long long gmask = ((1ull << G) - 1ull) << (gi * G);
  const bool bad = rc != 0 || (__ballot(out) & gmask) != 0ull;
  if (l == 0 && bad) B.flags[c] = flags | F_NOMINAL_UNSTABLE;
}

template <int G, int R>
__global__ __launch_bounds__(64) void k_eig_nom(const ProblemDev* __restrict__ P, StageBufs B) {
  __shared__ double lds[eig_lds_doubles<G, R>()];
  k_eig_nom_body<G, R>(lds, P, B, blockIdx.x, gridDim.x, blockIdx.y, blockIdx.z);
}

template <int G, int R, bool LOOP>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(4))) void k_eig_nom_batch(BatchTab T) {
  __shared__ double lds[eig_lds_doubles<G, R>()];
  batch_blocks<LOOP>(T, [&](const BatchSlot& q) {
    k_eig_nom_body<G, R>(lds, T.probs[q.p], T.bufs[q.p], q.lb, q.nb, 0, 0);
  });
}

// One launch for every problem of the stage, whatever its order: each block runs the default tile
// of its problem's matrix capacity (orders up to 4, 8, 16) on one LDS array sized for the largest.
constexpr int EIG_FUSED_LDS = eig_lds_doubles<4, 1>() > eig_lds_doubles<EIG8_G, 8 / EIG8_G>()
                                  ? (eig_lds_doubles<4, 1>() > eig_lds_doubles<EIG16_G, 16 / EIG16_G>()
                                         ? eig_lds_doubles<4, 1>() : eig_lds_doubles<EIG16_G, 16 / EIG16_G>())
                                  : (eig_lds_doubles<EIG8_G, 8 / EIG8_G>() > eig_lds_doubles<EIG16_G, 16 / EIG16_G>()
                                         ? eig_lds_doubles<EIG8_G, 8 / EIG8_G>() : eig_lds_doubles<EIG16_G, 16 / EIG16_G>());

template <bool NOM, bool LOOP>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(3))) void k_eig_batch_fused(BatchTab T) {
  __shared__ double lds[EIG_FUSED_LDS];
  batch_blocks<LOOP>(T, [&](const BatchSlot& q) {
    const ProblemDev* P = T.probs[q.p];
    const int key = eig_key(NOM ? 2 * P->ns : P->nt);
    if (key == 4) {
      if (NOM) k_eig_nom_body<4, 1>(lds, P, T.bufs[q.p], q.lb, q.nb, 0, 0);
      else k_eig_lin_body<4, 1>(lds, P, T.bufs[q.p], q.lb, q.nb, 0, 0);
    } else if (key == 8) {
      if (NOM) k_eig_nom_body<EIG8_G, 8 / EIG8_G>(lds, P, T.bufs[q.p], q.lb, q.nb, 0, 0);
      else k_eig_lin_body<EIG8_G, 8 / EIG8_G>(lds, P, T.bufs[q.p], q.lb, q.nb, 0, 0);
    } else {
      if (NOM) k_eig_nom_body<EIG16_G, 16 / EIG16_G>(lds, P, T.bufs[q.p], q.lb, q.nb, 0, 0);
      else k_eig_lin_body<EIG16_G, 16 / EIG16_G>(lds, P, T.bufs[q.p], q.lb, q.nb, 0, 0);
    }
  });
}
#undef CE_ROWS


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
  static constexpr int G = 16;        // lanes per candidate
  static constexpr int CPW = 64 / G;  // candidates per workgroup
  static constexpr int NCP = 8 * NB;  // padded basis width
  static constexpr int NM = NCP / 2;  // padded mode count
  static constexpr int NTRI = NCP * (NCP + 1) / 2;
  static constexpr int TB = 8;        // samples staged per LDS round
  static constexpr int UNITS = (NB == 1) ? 1 : NB + NB * (NB - 1);
};

template <int NB>
__device__ __forceinline__ void k_fit_fp64_body(const ProblemDev* __restrict__ P, const StageBufs& B, int bx, int nbx, int by, int bz) {
  using S = FitShape<NB>;
  constexpr int G = S::G, CPW = S::CPW, NCP = S::NCP, NM = S::NM, TB = S::TB;
  __shared__ double phi[CPW][TB][NCP];
  __shared__ double gram[CPW][S::NTRI];
  __shared__ double vec[CPW][6][NCP];  // atb, diag, subdiag, col scale, solution, scratch
  __shared__ int itr[CPW][NCP];
  __shared__ double modes[CPW][S::NM][3];

  const int lane = threadIdx.x;
  const int slot = lane / G, g = lane % G;
  const int idx = bx * CPW + slot;
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
  // tile owned by this lane: lane g < NB owns diagonal block g (its lower triangle as an 8x4
  // rectangle plus a 4x4 square) and the right-hand side of those rows; lanes NB.. own the two 8x4
  // halves of one off-diagonal block each; the rest accumulate nothing
  int rb = 0, cb = 0, coff = 0;
  bool diag_unit = false, idle = g >= S::UNITS;
  if (g < NB) {
    diag_unit = true;
    rb = g;
    cb = g;
  } else if (!idle) {
    int p = (g - NB) >> 1;
    coff = ((g - NB) & 1) * 4;
    rb = 1;
    while (p >= rb) { p -= rb; rb++; }
    cb = p;
  }
  // basis evaluations of a TB-sample block spread over the G lanes; the candidate's mode table
  // (decay rate, frequency, shift) is staged in LDS
  const double ts = P->ts, t0 = P->fit_t0, t1 = P->fit_t1;
  for (int e = g; e < NM; e += G) {
    double re = 0.0, im = 0.0, sh = 0.0;
    if (active && e < u) {
      re = B.mode_re[(size_t)e * n + c];
      im = B.mode_im[(size_t)e * n + c];
      sh = cedm::fmax_(re * t0, re * t1);
    }
    modes[slot][e][0] = re;
    modes[slot][e][1] = im;
    modes[slot][e][2] = sh;
  }
  __syncthreads();
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
  const int nc = 2 * u;
  // the columns of modes u..NM-1 stay zero; each round evaluates only the u x TB basis values the
  // candidate has, spread over the group's lanes (same values, fewer idle evaluations)
  for (int k = g; k < TB * NCP; k += G) phi[slot][k / NCP][k % NCP] = 0.0;
  // Gram rows beyond 2nt (padding, never accumulated) are +0.0: the solve's unrolled substitution
  // reads them as exact no-op terms
  for (int k = nc2 * (nc2 + 1) / 2 + g; k < S::NTRI; k += G) gram[slot][k] = 0.0;
  const int ue = active ? u : 0;
  __syncthreads();
  for (int tb = 0; tb < cnt; tb += TB) {
#pragma unroll 1
    for (int k = g; k < ue * TB; k += G) {
      const int e = k % ue, s = k / ue;
      const int t = tb + s;
      double cv = 0.0, sv = 0.0;
      if (t < cnt) {
        const double tval = (double)(f0 + t) * ts;
        const double decay = cedm::exp(cedm::fma_(modes[slot][e][0], tval, -modes[slot][e][2]));
        double sn, cs;
        cedm::sincos(modes[slot][e][1] * tval, sn, cs);
        cv = decay * cs;
        sv = decay * sn;
      }
      phi[slot][s][2 * e] = cv;
      phi[slot][s][2 * e + 1] = sv;
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
          for (int i = 0; i < 8; i++)
            if (rbase + i < nc) AB[i] = cedm::fma_(pr[i], yv, AB[i]);
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
  if (!active) return;

  double* G2 = gram[slot];
  double* atb = vec[slot][0];
  double* dg = vec[slot][1];
  double* sd = vec[slot][2];
  double* cs = vec[slot][3];
  double* x = vec[slot][4];
  if (g == 0 && B.diag && B.d_ata) {
    float* da = B.d_ata + (size_t)c * nc2 * nc2;
    for (int i = 0; i < nc2; i++)
      for (int j = 0; j < nc2; j++) da[i * nc2 + j] = (j <= i) ? (float)G2[tri(i, j)] : 0.0f;
  }
  if (g == 0 && B.diag && B.d_atb) {
    float* db = B.d_atb + (size_t)c * nc2;
    for (int i = 0; i < nc2; i++) db[i] = (float)atb[i];
  }
  double* sc = row + P->info_scoring;
  if (sc[0] != 0.0) return;  // modalsolve skips flagged candidates (src/modalsolve.cu:112)
  solve_and_score_grp<G, NCP>(P, B, c, u, row, G2, atb, dg, sd, cs, x, vec[slot][5], itr[slot], g, 2 * nt);
}

template <int NB>
__global__ __launch_bounds__(64) void k_fit_fp64(const ProblemDev* __restrict__ P, StageBufs B) {
  k_fit_fp64_body<NB>(P, B, blockIdx.x, gridDim.x, blockIdx.y, blockIdx.z);
}

template <int NB, bool LOOP>
__global__ __launch_bounds__(64) void k_fit_fp64_batch(BatchTab T) {
  batch_blocks<LOOP>(T, [&](const BatchSlot& q) {
    k_fit_fp64_body<NB>(T.probs[q.p], T.bufs[q.p], q.lb, q.nb, 0, 0);
  });
}


// ------------------------------------------------------------------------------------------------
template <int NS>
__device__ __forceinline__ void k_sweep_fp64_body(const ProblemDev* __restrict__ P, const StageBufs& B, int bx, int nbx, int by, int bz) {
  const int c = bx;
  const int w = by * 64 + threadIdx.x;
  if (margin_skip(P, B, c)) return;
  const int F = P->F;
  if (w >= F) return;
  if (!tab<int32_t>(P, P->off_pok)[w]) { atomicOr(&B.sweep_fail[c], 1); return; }
  double m, raw, b;
  if (!sweep_point<NS, double>(P, B.open_a + (size_t)c * NS * NS, B.gains + (size_t)c * P->cuda_w, w, m, raw, b)) {
    atomicOr(&B.sweep_fail[c], 1);
    return;
  }
  double* out = B.mag_raw + ((size_t)c * F + w) * 2;
  out[0] = m;
  out[1] = raw;
  if (b == b) atomicMin(&B.bmin[c], (unsigned long long)__double_as_longlong(b));
}

template <int NS>
__global__ __launch_bounds__(64) void k_sweep_fp64(const ProblemDev* __restrict__ P, StageBufs B) {
  k_sweep_fp64_body<NS>(P, B, blockIdx.x, gridDim.x, blockIdx.y, blockIdx.z);
}

template <int NS, bool LOOP>
__global__ __launch_bounds__(64) void k_sweep_fp64_batch(BatchTab T) {
  batch_blocks<LOOP>(T, [&](const BatchSlot& q) {
    const StageBufs& B = T.bufs[q.p];
    k_sweep_fp64_body<NS>(T.probs[q.p], B, q.lb % B.n, B.n, q.lb / B.n, 0);
  });
}


// Sequential part of the sweep (src/allmargin.cu:119-180), one candidate per lane.
__device__ __forceinline__ void k_scan_fp64_body(const ProblemDev* __restrict__ P, const StageBufs& B, int bx, int nbx, int by, int bz) {
  const int c = bx * 64 + threadIdx.x;
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

__global__ __launch_bounds__(64) void k_scan_fp64(const ProblemDev* __restrict__ P, StageBufs B) {
  k_scan_fp64_body(P, B, blockIdx.x, gridDim.x, blockIdx.y, blockIdx.z);
}

template <bool LOOP>
__global__ __launch_bounds__(64) void k_scan_fp64_batch(BatchTab T) {
  batch_blocks<LOOP>(T, [&](const BatchSlot& q) {
    k_scan_fp64_body(T.probs[q.p], T.bufs[q.p], q.lb, q.nb, 0, 0);
  });
}


// ------------------------------------------------------------------------------------------------
// Launch helpers (host side of this translation unit)

bool launch_sim_fp64(int ns, int nd, const ProblemDev* P, const StageBufs& B, hipStream_t st) {
  return launch_sim<double>(ns, nd, P, B, B.ybuf, st);
}

void launch_eig_lin(int nt, const ProblemDev* P, const StageBufs& B, hipStream_t st) {
  const int key = eig_key(nt), cpw = 64 / eig_lanes(key);
  const dim3 grid((B.n + cpw - 1) / cpw);
  if (key == 4) hipLaunchKernelGGL((k_eig_lin<4, 1>), grid, dim3(64), 0, st, P, B);
  else if (key == 8) hipLaunchKernelGGL((k_eig_lin<EIG8_G, 8 / EIG8_G>), grid, dim3(64), 0, st, P, B);
  else hipLaunchKernelGGL((k_eig_lin<EIG16_G, 16 / EIG16_G>), grid, dim3(64), 0, st, P, B);
  kmark("k_eig_lin", st);
}

void launch_eig_nom(int ns, const ProblemDev* P, const StageBufs& B, hipStream_t st) {
  const int m = 2 * ns;
  const int key = eig_key(m), cpw = 64 / eig_lanes(key);
  const dim3 grid((B.n + cpw - 1) / cpw);
  if (key == 4) hipLaunchKernelGGL((k_eig_nom<4, 1>), grid, dim3(64), 0, st, P, B);
  else if (key == 8) hipLaunchKernelGGL((k_eig_nom<EIG8_G, 8 / EIG8_G>), grid, dim3(64), 0, st, P, B);
  else hipLaunchKernelGGL((k_eig_nom<EIG16_G, 16 / EIG16_G>), grid, dim3(64), 0, st, P, B);
  kmark("k_eig_nom", st);
}

void launch_compact(const ProblemDev* P, const StageBufs& B, hipStream_t st) {
  hipLaunchKernelGGL(k_compact, dim3((B.n + 63) / 64), dim3(64), 0, st, P, B);
  kmark("k_compact", st);
}

bool launch_fit_fp64(int nt, const ProblemDev* P, const StageBufs& B, hipStream_t st) {
  const int nb = (2 * nt + 7) / 8;
  const dim3 grid((B.n + 3) / 4);
  switch (nb) {
    case 1: hipLaunchKernelGGL(k_fit_fp64<1>, grid, dim3(64), 0, st, P, B); kmark("k_fit_fp64", st); return true;
    case 2: hipLaunchKernelGGL(k_fit_fp64<2>, grid, dim3(64), 0, st, P, B); kmark("k_fit_fp64", st); return true;
    case 3: hipLaunchKernelGGL(k_fit_fp64<3>, grid, dim3(64), 0, st, P, B); kmark("k_fit_fp64", st); return true;
    case 4: hipLaunchKernelGGL(k_fit_fp64<4>, grid, dim3(64), 0, st, P, B); kmark("k_fit_fp64", st); return true;
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
  kmark("k_sweep_fp64", st);
  hipLaunchKernelGGL(k_scan_fp64, dim3((B.n + 63) / 64), dim3(64), 0, st, P, B);
  kmark("k_scan_fp64", st);
  return true;
}

// A grid of a workgroup per block launches the kernels' loop-free form (fewer registers).
template <bool LOOP>
static bool launch_batch_fp64_t(int kind, int key, const BatchTab& T, const dim3 g, hipStream_t st) {
  const dim3 b(64);
#define CE_NS_CASES(K)                                                      \
  switch (key) {                                                            \
    case 1: hipLaunchKernelGGL((K<1, LOOP>), g, b, 0, st, T); return true;          \
    case 2: hipLaunchKernelGGL((K<2, LOOP>), g, b, 0, st, T); return true;          \
    case 3: hipLaunchKernelGGL((K<3, LOOP>), g, b, 0, st, T); return true;          \
    case 4: hipLaunchKernelGGL((K<4, LOOP>), g, b, 0, st, T); return true;          \
    case 5: hipLaunchKernelGGL((K<5, LOOP>), g, b, 0, st, T); return true;          \
    case 6: hipLaunchKernelGGL((K<6, LOOP>), g, b, 0, st, T); return true;          \
    case 7: hipLaunchKernelGGL((K<7, LOOP>), g, b, 0, st, T); return true;          \
    case 8: hipLaunchKernelGGL((K<8, LOOP>), g, b, 0, st, T); return true;          \
  }                                                                         \
  return false;
  switch (kind) {
    // eigen keys: 0 = the fused launch over every order; else capacity * 100 + lanes per candidate
    case BK_EIG_LIN:
      switch (key) {
        case 0: hipLaunchKernelGGL((k_eig_batch_fused<false, LOOP>), g, b, 0, st, T); return true;
        case 404: hipLaunchKernelGGL((k_eig_lin_batch<4, 1, LOOP>), g, b, 0, st, T); return true;
        case 804: hipLaunchKernelGGL((k_eig_lin_batch<4, 2, LOOP>), g, b, 0, st, T); return true;
        case 808: hipLaunchKernelGGL((k_eig_lin_batch<8, 1, LOOP>), g, b, 0, st, T); return true;
        case 1608: hipLaunchKernelGGL((k_eig_lin_batch<8, 2, LOOP>), g, b, 0, st, T); return true;
        case 1616: hipLaunchKernelGGL((k_eig_lin_batch<16, 1, LOOP>), g, b, 0, st, T); return true;
      }
      return false;
    case BK_EIG_NOM:
      switch (key) {
        case 0: hipLaunchKernelGGL((k_eig_batch_fused<true, LOOP>), g, b, 0, st, T); return true;
        case 404: hipLaunchKernelGGL((k_eig_nom_batch<4, 1, LOOP>), g, b, 0, st, T); return true;
        case 804: hipLaunchKernelGGL((k_eig_nom_batch<4, 2, LOOP>), g, b, 0, st, T); return true;
        case 808: hipLaunchKernelGGL((k_eig_nom_batch<8, 1, LOOP>), g, b, 0, st, T); return true;
        case 1608: hipLaunchKernelGGL((k_eig_nom_batch<8, 2, LOOP>), g, b, 0, st, T); return true;
        case 1616: hipLaunchKernelGGL((k_eig_nom_batch<16, 1, LOOP>), g, b, 0, st, T); return true;
      }
      return false;
    case BK_COMPACT: hipLaunchKernelGGL(k_compact_batch<LOOP>, g, b, 0, st, T); return true;
    case BK_FIT_FP64:
      switch (key) {
        case 1: hipLaunchKernelGGL((k_fit_fp64_batch<1, LOOP>), g, b, 0, st, T); return true;
        case 2: hipLaunchKernelGGL((k_fit_fp64_batch<2, LOOP>), g, b, 0, st, T); return true;
        case 3: hipLaunchKernelGGL((k_fit_fp64_batch<3, LOOP>), g, b, 0, st, T); return true;
        case 4: hipLaunchKernelGGL((k_fit_fp64_batch<4, LOOP>), g, b, 0, st, T); return true;
      }
      return false;
    case BK_SWEEP_FP64: CE_NS_CASES(k_sweep_fp64_batch)
    case BK_SCAN_FP64: hipLaunchKernelGGL(k_scan_fp64_batch<LOOP>, g, b, 0, st, T); return true;
  }
  return false;
#undef CE_NS_CASES
}

bool launch_batch_fp64(int kind, int key, const BatchTab& T, int grid, hipStream_t st) {
  return grid >= T.total ? launch_batch_fp64_t<false>(kind, key, T, dim3(grid), st)
                         : launch_batch_fp64_t<true>(kind, key, T, dim3(grid), st);
}

void launch_info_init(const float* in, double* out, size_t count, hipStream_t st) {
  hipLaunchKernelGGL(k_info_init, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, st, in, out, count);
}
void launch_info_final(const double* in, float* out, size_t count, hipStream_t st) {
  hipLaunchKernelGGL(k_info_final, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, st, in, out, count);
  kmark("k_info_final", st);
}

}  // namespace ce

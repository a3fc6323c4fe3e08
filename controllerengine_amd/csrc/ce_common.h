// Problem descriptor shared by the host runtime and the gfx950 kernels.
//
// One ProblemDev describes one plant/config (one round of the reference): everything the stage
// kernels read that is the same for every candidate of the problem.  It lives in device global
// memory and is read with wave-uniform (scalar) loads.  Per-frequency tables (z grid, plant
// response) follow it in the same allocation.
#pragma once

#include <stdint.h>

#include "../../include/ControllerEngine.h"
#include "ce_detmath.h"

namespace ce {

constexpr int MAXS = CE_MAX_STATESPACE_DIM;
constexpr int MAXD = CE_MAX_DISCRETE_DIM;
constexpr int MAXW = CE_MAX_NONLIN_DIM;
constexpr int MAXT = CE_MAX_TOTAL_DIM;
constexpr int MAXNC = 2 * CE_MAX_TOTAL_DIM;

// Constants of the reference, float literals promoted exactly to double.
constexpr float PI_F = 0x1.921fb6p+1f;                       // src/main.cu:41
constexpr double EPS30 = (double)0x1.a484cp-100f;            // log floor, src/modalsetup.cu:20
constexpr double EPS6 = (double)0x1.0c6f7ap-20f;             // 1e-6 thresholds
constexpr double BIG10 = (double)0x1.2a05f2p+33f;            // dedupe sentinel, src/modalsetup.cu:29
constexpr double FLTMAX = (double)0x1.fffffep127f;           // initial margins, src/allmargin.cu:95
constexpr double PI_D = cedm::PI_HI;
constexpr double TWO_PI_D = 2.0 * cedm::PI_HI;
constexpr double RAD2DEG = 180.0 / cedm::PI_HI;
constexpr double DEG2RAD = cedm::PI_HI / 180.0;

// Operand kinds (reference src/main.cu:73).
enum { K_CONST = 0, K_STATE = 1, K_PREV = 2 };

struct Ins {
  int32_t op, col, akind, bkind;
  double aval, bval;
};

struct ProblemDev {
  // sizes
  int32_t ns, nd, nt, nw, T, F;
  int32_t dc_n, u_n, nl_n[MAXW];
  int32_t fit_start_id, steady_start_id, fit_count, ref_step_id;
  int32_t info_w, info_scoring, info_freqresp, info_allmargin, info_modal, cuda_w;
  int32_t serial;
  // 1 when the margins stage's nominal closed loop is D LIN_A D, D = diag(I_ns, -I_ns), entry for
  // entry (serial mode with nd = ns, or batch mode with the observer model equal to the plant and
  // STATESPACE_D = 0): then it has the same eigenvalues, bit for bit, as LIN_A unless the LIN_A
  // solve reports a sign-flip event (eig_grp)
  int32_t nom_equiv;
  // scalars (double, promoted from the config's float values)
  double ts, fit_t0, fit_t1, unstable_err, eig_fail_err, ref_amp;
  double damping, bandwidth, overshoot, sse, target_gm, target_pm, target_b;
  double alpha;
  // observer model and plant, compact row-major
  double sa[MAXS * MAXS], sai[MAXS * MAXS], sb[MAXS], sc[MAXS], sd;
  double da[MAXD * MAXD], dai[MAXD * MAXD], dbu[MAXD], dbw[MAXD * MAXW], dc[MAXD];
  // plant used by the margins (batch: observer model; serial: DISCRETE_* with D = 0)
  double pa[MAXS * MAXS], pb[MAXS], pc[MAXS], pd;
  // decoded instruction chains
  Ins dc_ops[CE_MAX_DC_INSTRUCTIONS];
  Ins u_ops[CE_MAX_U_INSTRUCTIONS];
  Ins nl_ops[MAXW][CE_MAX_NL_INSTRUCTIONS];
  // float copies for the mixed-precision kernels
  float fsa[MAXS * MAXS], fsai[MAXS * MAXS], fsb[MAXS], fsc[MAXS], fsd;
  float fdai[MAXD * MAXD], fdbu[MAXD], fdbw[MAXD * MAXW];
  float fdc_aval[CE_MAX_DC_INSTRUCTIONS], fdc_bval[CE_MAX_DC_INSTRUCTIONS];
  float fu_aval[CE_MAX_U_INSTRUCTIONS], fu_bval[CE_MAX_U_INSTRUCTIONS];
  float fnl_aval[MAXW][CE_MAX_NL_INSTRUCTIONS], fnl_bval[MAXW][CE_MAX_NL_INSTRUCTIONS];
  float fts, fref_amp;
  // byte offsets (from the start of this struct) of the per-frequency tables
  int64_t off_zr, off_zi, off_theta, off_pr, off_pi, off_pabs, off_pok;
  // The mixed sweep's per-frequency inputs and the grid angles in lane-interleaved order: a wave
  // gives each lane KM = ceil(F / 64) consecutive frequencies, and frequency w = l * KM + k is
  // stored at slot k * 64 + l, so that the loads of one k by the 64 lanes are consecutive.
  int64_t off_sweep_il;  // SweepPointIl[KM * 64]
  int64_t off_theta_il;  // double[KM * 64]: theta_w
  int64_t bytes_total;
};

// One frequency of the mixed sweep (lane-interleaved table): z - 1 in double (its real part
// exact in the delta form), the plant response rounded to float as the sweep uses it, and whether
// the plant response is finite.
struct alignas(16) SweepPointIl {
  double zm1, zi;
  float pr, pi, pabs;
  int32_t ok;
};
static_assert(sizeof(SweepPointIl) == 32, "two 16-byte loads per frequency");

// slot of frequency w in a lane-interleaved table with km frequencies per lane
__host__ __device__ inline int il_slot(int w, int km) { return (w % km) * 64 + w / km; }

}  // namespace ce

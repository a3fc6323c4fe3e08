// Per-launch stage buffers shared by the engine and the kernels.
#pragma once

#include <hip/hip_runtime.h>

#include "ce_common.h"

namespace ce {

enum : int32_t {
  F_UNSTABLE = 1,          // some closed-loop eigenvalue has |lambda| > 1 (reference src/modalsetup.cu:26)
  F_NOMINAL_UNSTABLE = 2,  // nominal loop of the margins stage unstable (reference src/allmargin.cu:61-87)
  F_NOMINAL_SKIPPED = 4,   // nominal check not run (candidate already at or above unstable_err)
  F_SWEEP_FP64 = 8,        // mixed sweep: every point recomputed in FP64 (more than RMAX points
                           // wanting refinement)
  F_NOMINAL_DONE = 16      // nominal check decided from the LIN_A solve (ProblemDev::nom_equiv and
                           // no sign-flip event): the nominal eigensolve is not run
};

constexpr int RMAX = 64;   // FP64 refinement points per candidate in the mixed sweep
// Mixed sweep: factor on the first-order error estimate of a point's FP64 Hessenberg solve
// (rounding unit times the pivot conditioning times the cancellation of the closing dot product),
// and the allowance for the float rounding of the loop value and its float tail.  The guard bands
// of each point are widened by the sum.
constexpr float SWEEP_ERR_SAFETY = 8.0f;
constexpr float SWEEP_ERR_FLOOR = 1e-6f;

struct StageBufs {
  int32_t n;                    // candidates in this launch
  int32_t diag;                 // nonzero: fill the diagnostic dumps
  const float* gains;           // n x cuda_w
  double* info;                 // n x info_w, double copy of the INFO rows
  int32_t* flags;               // n
  int32_t* uniq;                // n, unique modes
  double* mode_re;              // MAXT x n (SoA)
  double* mode_im;              // MAXT x n
  double* open_a;               // n x ns*ns (LIN_A top-left block, per candidate contiguous)
  double* ybuf;                 // fit_count x n (SoA), FP64 path
  float* ybuf32;                // fit_count x n (SoA), mixed path
  double* mag_raw;              // n x F x 2: |loop| and arg(loop) per frequency
  unsigned long long* bmin;     // n: bit pattern of min b(w) (non-negative doubles order as integers)
  int32_t* sweep_fail;          // n
  int32_t* list;                // n: candidates that reach the simulation (compacted)
  int32_t* count;               // 1: entries in list
  int32_t* nlist;               // n: candidates whose nominal loop needs its own eigensolve
  int32_t* ncount;              // 1: entries in nlist
  // mixed sweep
  float* mr32;                  // n x (KM * 64) x 4, frequency w at slot il_slot(w, KM): located |loop|,
                                // arg(loop), b(w), error estimate (float)
  int8_t* refined;              // n x F: 0 float value, 1 marked, 2 refined in FP64, 3 FP64 not finite
  int2* rlist;                  // n x RMAX: dense (candidate, frequency) list
  int32_t* rcount;              // 1
  int32_t* flist;               // n: candidates whose whole sweep is recomputed in FP64
  int32_t* fcount;              // 1
  double* mres;                 // n x 6: crossing-scan minima GM, PM, DM, b, crossed, failed (scan
                                // beside the fit; NULL when the scan follows the fit)
  // mixed fit: partial right-hand sides sum_{t in part} (y_t - yss) z_e^(t - t_part), complex,
  // layout [part][mode][list slot]
  double2* atbp;
  int32_t atb_parts, atb_chunk;
  // list slots a fit launch covers (mixed chain, split simulation): 0 every slot, 1 the slots below
  // sim_split_slots(count, sim_cap), 2 the slots from there on
  int32_t sim_part, sim_cap;
  // diagnostics in the reference dump layout (float, candidate-major), NULL when not requested
  float* d_y;
  float* d_lina;
  float* d_eig;
  float* d_ata;
  float* d_atb;
};

// Split simulation (mixed chain, more than two rounds of one-lane waves): the one-lane kernel takes
// the list slots of its full rounds, one 64-candidate wave per SIMD per round (cap SIMDs), and a
// kernel with 2 lanes per candidate the rest, on another stream once the first has finished.
// Without the split, the SIMDs that hold one wave more than the others (c3: 287 of 1,024 hold a
// third) set the simulation's time while the rest wait.  Returns the first slot of the remainder,
// or live (no split) with fewer than two full rounds or a remainder over half a round.
__host__ __device__ inline int sim_split_slots(int live, int cap) {
  if (cap <= 0) return live;
  const int w = (live + 63) / 64;
  const int k = w / cap, r = w - k * cap;
  return (k >= 2 && r > 0 && 2 * r <= cap) ? k * cap * 64 : live;
}

// Lane layout of the eigen kernels.  Matrices of order <= 4, <= 8 and <= 16 (the kernel's key)
// use 4, EIG8_G and EIG16_G lanes per candidate, with key / lanes rows on a lane.  At order 16,
// 16 lanes (one row each) measured faster than 8 lanes with two rows each, whose doubled LDS per
// wave halves the waves in flight (c3 LIN_A eigensolve 7.8 vs 8.0 ms).  At order 8, 4 lanes with
// two rows each measured faster than 8 lanes (c4 order-8 LIN_A eigensolve 2.07 vs 2.55 ms; c2
// unchanged): every lane of a group evaluates the shifts and Householder coefficients, so fewer
// lanes per candidate do less redundant work.  2 lanes with four rows each went too far (c2 LIN_A
// eigensolve 0.49 vs 0.38 ms: twice the LDS per wave, half the waves in flight).
constexpr int EIG16_G = 16;
constexpr int EIG8_G = 4;
__host__ __device__ inline int eig_key(int order) { return order <= 4 ? 4 : (order <= 8 ? 8 : 16); }
inline int eig_lanes(int key) { return key == 16 ? EIG16_G : (key == 8 ? EIG8_G : key); }

// Batched launches: one launch of a stage kernel covers a group of problems (same kernel
// template instance).  Problem j of the group owns the group-local blocks [start[j], start[j+1]);
// a workgroup maps its block to (problem, local block, the problem's block count) by binary search
// over the prefix and runs the single-problem body.  Workgroups loop over blocks with a stride of
// the grid size, so the grid may be smaller than the block count (persistent launch).
struct BatchTab {
  const int32_t* start;            // J + 1 block prefix of this group
  const int32_t* pid;              // J: batch index of each problem of the group
  const ProblemDev* const* probs;  // per batch problem
  const StageBufs* bufs;           // per batch problem
  int32_t J, total;
};
// Stage kernels that have a batched form (key: the template parameter of the instance).
enum BatchKernel : int {
  BK_EIG_LIN = 0,   // key: matrix capacity (eig_key)
  BK_EIG_NOM,       // key: matrix capacity (eig_key)
  BK_COMPACT,
  BK_FIT_FP64,      // key NB
  BK_SWEEP_FP64,    // key NS
  BK_SCAN_FP64,
  BK_ATB,
  BK_FITCF,         // key NB
  BK_SWEEP_MARK32,  // key NS
  BK_REFINE64,      // key NS
  BK_FULL64,        // key NS
  BK_SCAN32,
  BK_COUNT
};
bool launch_batch_fp64(int kind, int key, const BatchTab& T, int grid, hipStream_t st);
bool launch_batch_mixed(int kind, int key, const BatchTab& T, int grid, hipStream_t st);

struct BatchSlot {
  int p, lb, nb;
};
__device__ __forceinline__ BatchSlot batch_slot(const BatchTab& T, int b) {
  b = __builtin_amdgcn_readfirstlane(b);
  int lo = 0, hi = T.J - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (T.start[mid] <= b) lo = mid;
    else hi = mid - 1;
  }
  // workgroup-uniform: keep the slot in scalar registers
  const int p = __builtin_amdgcn_readfirstlane(T.pid[lo]);
  const int s0 = __builtin_amdgcn_readfirstlane(T.start[lo]);
  const int s1 = __builtin_amdgcn_readfirstlane(T.start[lo + 1]);
  return BatchSlot{p, b - s0, s1 - s0};
}

// The blocks of a batched launch.  LOOP: the workgroups stride over the blocks (a persistent launch
// on a capped grid).  Otherwise the launch has a workgroup per block and the kernel holds no block
// loop, which costs registers (the loop keeps values of the body live across its iterations).
template <bool LOOP, typename Body>
__device__ __forceinline__ void batch_blocks(const BatchTab& T, Body&& body) {
  if constexpr (LOOP) {
    for (int b = blockIdx.x; b < T.total; b += gridDim.x) body(batch_slot(T, b));
  } else {
    if ((int)blockIdx.x < T.total) body(batch_slot(T, (int)blockIdx.x));
  }
}

// Optional per-kernel timing.  Launchers call kmark(name, stream) after each stage kernel; when
// the calling thread's engine has profiling on, an event is recorded there, so that the time
// between two marks is the duration of the kernel named by the second one.
using KernelMarkFn = void (*)(void* ctx, const char* name, hipStream_t st);
struct KernelMarker {
  KernelMarkFn fn = nullptr;
  void* ctx = nullptr;
};
KernelMarker& kernel_marker();  // per host thread
inline void kmark(const char* name, hipStream_t st) {
  KernelMarker& m = kernel_marker();
  if (m.fn) m.fn(m.ctx, name, st);
}

}  // namespace ce

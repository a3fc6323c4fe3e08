// Per-launch stage buffers shared by the engine and the kernels.
#pragma once

#include <hip/hip_runtime.h>

#include "ce_common.h"

namespace ce {

enum : int32_t {
  F_UNSTABLE = 1,          // some closed-loop eigenvalue has |lambda| > 1 (reference src/modalsetup.cu:26)
  F_NOMINAL_UNSTABLE = 2,  // nominal loop of the margins stage unstable (reference src/allmargin.cu:61-87)
  F_NOMINAL_SKIPPED = 4,   // nominal check not run (candidate already at or above unstable_err)
  F_REFINE_OVERFLOW = 8    // mixed sweep: more than RMAX points wanted FP64 refinement
};

constexpr int RMAX = 64;   // FP64 refinement points per candidate in the mixed sweep

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
  // mixed sweep
  float* mr32;                  // n x F x 3: float |loop|, arg(loop), b(w)
  int8_t* refined;              // n x F: 0 float value, 1 marked, 2 refined in FP64, 3 FP64 not finite
  int2* rlist;                  // n x RMAX: dense (candidate, frequency) list
  int32_t* rcount;              // 1
  // mixed fit: partial right-hand sides sum_{t in part} (y_t - yss) z_e^(t - t_part), complex,
  // layout [part][mode][list slot]
  double2* atbp;
  int32_t atb_parts, atb_chunk;
  // diagnostics in the reference dump layout (float, candidate-major), NULL when not requested
  float* d_y;
  float* d_lina;
  float* d_eig;
  float* d_ata;
  float* d_atb;
};

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

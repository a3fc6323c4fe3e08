// Runtime-specialised simulation kernels (hipRTC): source generation, compile cache, launch.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "ce_common.h"

namespace ce {

struct JitCache;
JitCache* jit_cache_create();
void jit_cache_destroy(JitCache* c);
bool jit_enabled();
// HIP source of the simulation kernel for problem P (mixed: float state and trajectory), every
// constant written into the source.  gated: the form that holds whole SIMDs and signals its start
// (jit_launch_sim_gated; grid at most the device's SIMD count).
std::string gen_sim_source(const ProblemDev& P, bool mixed, bool gated = false);
// Compiled kernel of src (compiled on first use).  evictable: the entry may be unloaded once more
// recent evictable entries exceed the cache's capacity (after a device drain).
// fn_lanes (optional, 2 entries): the module's NAME_l2 and NAME_l4 kernels, or NULL where it has none.
bool jit_get(JitCache* cache, const std::string& src, void** fn, std::string& err, bool evictable = false,
             void** fn_lanes = nullptr);
bool jit_cached(JitCache* cache, const std::string& src);
long long jit_compiles(JitCache* cache);   // hipRTC compilations done by this cache
long long jit_evictions(JitCache* cache);  // entries unloaded
// Structural single-problem form: constants and sizes read from the problem descriptor (the last
// kernel argument), so one compiled kernel serves every problem of the same dimensions and
// instruction structure (sim_structure_key), e.g. the rounds of a serial-mode run.
std::string gen_sim_struct_source(const ProblemDev& P, bool mixed, bool gated = false);
// The ungated forms (here and jit_launch_sim) store done_gen to *done (when not NULL) as soon as
// the first working wave has finished its time loop, or at once when no candidate is live.
bool jit_launch_sim_struct(void* fn, int grid, hipStream_t st, const float* gains, const int* flags, const int* list,
                           const int* count, void* ybuf, double* info, float* dy, int n, const void* prob,
                           int* done = nullptr, int done_gen = 0);
bool jit_launch_sim_struct_gated(void* fn, int grid, hipStream_t st, const float* gains, const int* flags, const int* list,
                                 const int* count, void* ybuf, double* info, float* dy, int n, int* gate, int gen,
                                 const void* prob);
// Literal form.  gq: lanes per candidate the launch allows (1, 2 or 4; the grid covers gq lanes per
// candidate), cap: waves the device runs one per SIMD.  The float kernel runs the most lanes per
// candidate, up to gq, whose live candidates' waves stay within cap; the FP64 kernel runs one.
// part (split simulation, sim_split_slots): 0 every list slot; 1 the one-lane kernel on the slots
// before the split; 2 the kernel launched (2 lanes) on the slots from the split on, whatever gq.
bool jit_launch_sim(void* fn, int grid, hipStream_t st, const float* gains, const int* flags, const int* list,
                    const int* count, void* ybuf, double* info, float* dy, int n, int gq, int cap, int* done = nullptr,
                    int done_gen = 0, int part = 0);
// the gated form: its last workgroup stores gen to *gate when it starts
bool jit_launch_sim_gated(void* fn, int grid, hipStream_t st, const float* gains, const int* flags, const int* list,
                          const int* count, void* ybuf, double* info, float* dy, int n, int* gate, int gen, int gq,
                          int cap);

// Batched form: one kernel per instruction structure and dimensions (sim_structure_key), constants
// and sizes read from each problem's descriptor.  SimArgs mirrors the struct in the generated
// source.
struct SimArgs {
  const float* gains;
  const int* flags;
  const int* list;
  const int* count;
  void* ybuf;
  double* info;
  float* dy;
  int n;
  int pad;
};
// The float module also holds NAME_l2 and NAME_l4, the 2- and 4-lane forms (a launch's block table
// has 64 / lanes candidates per block).
std::string gen_sim_batch_source(const ProblemDev& P, bool mixed);
std::string sim_structure_key(const ProblemDev& P, bool mixed);
bool jit_launch_sim_batch(void* fn, int grid, hipStream_t st, const int* start, const int* pid, const SimArgs* args,
                          const char* const* probs, int J, int total);
// Every structure of a batch in one kernel (reps: one problem of each structure); sid gives the
// structure index of each problem of the launch.
std::string gen_sim_batch_fused_source(const std::vector<const ProblemDev*>& reps, bool mixed);
bool jit_launch_sim_batch_fused(void* fn, int grid, hipStream_t st, const int* start, const int* pid, const SimArgs* args,
                                const char* const* probs, const int* sid, int J, int total);

}  // namespace ce

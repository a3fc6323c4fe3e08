// Runtime-specialised simulation kernels (hipRTC): source generation, compile cache, launch.
#pragma once

#include <hip/hip_runtime.h>

#include <string>

#include "ce_common.h"

namespace ce {

struct JitCache;
JitCache* jit_cache_create();
void jit_cache_destroy(JitCache* c);
bool jit_enabled();
// HIP source of the simulation kernel `ce_sim` for problem P (mixed: float state and trajectory).
std::string gen_sim_source(const ProblemDev& P, bool mixed);
bool jit_get(JitCache* cache, const std::string& src, void** fn, std::string& err);
bool jit_launch_sim(void* fn, int grid, hipStream_t st, const float* gains, const int* flags, const int* list,
                    const int* count, void* ybuf, double* info, float* dy, int n);

}  // namespace ce

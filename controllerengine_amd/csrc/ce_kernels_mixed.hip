// Mixed-precision stage chain (float simulation, fit accumulation and sweep; double eigensolve,
// normal-equation solve and crossover refinement).
#include <hip/hip_runtime.h>

#include <string>

#include "ce_stage.h"

namespace ce {

bool launch_mixed_chain(const ce_config&, const ProblemDev*, const StageBufs&, hipStream_t, hipEvent_t*, std::string& err) {
  err = "mixed precision chain not built yet";
  return false;
}

}  // namespace ce

// Host-side helpers of the engine (config, files, decode, problem tables).
#pragma once

#include <string>
#include <vector>

#include "ce_common.h"

namespace ce {

std::string fmt_num(double v);
std::string fmt_flt(float v);
bool parse_config(const std::string& text, const std::string& filename, ce_config& c, std::string& err);
bool derive_config(const ce_config& c, ce_derived& d, std::string& err);
bool read_bin(const std::string& path, const std::string& name, size_t nbytes, void* dst, int slice, int slices,
              std::string& err);
bool load_plant_dir(const std::string& dir, const ce_config& c, int round, ce_plant& p, std::string& err);
bool decode(const int32_t* info, int dim, int count, const float* base, int n_params, Ins* out, std::string& err);
bool shifted_solve_host(const double* A, int n, double zr, double zi, const double* b, double* xr, double* xi);
// ProblemDev followed by the per-frequency tables, ready to copy to the device.
bool build_problem(const ce_config& c, const ce_plant& pl, std::vector<uint8_t>& blob, std::string& err);

}  // namespace ce

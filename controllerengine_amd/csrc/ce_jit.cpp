// Runtime-specialised simulation kernels (hipRTC).
//
// The closed-loop simulation (reference src/discrete_solves.cu:12-79) interprets three base-36
// instruction chains every time step.  For a single problem those chains, the dimensions and the
// plant/observer matrices are fixed, so the engine generates the per-step code as straight-line HIP
// source with every operand, column and opcode resolved, compiles it once per problem with hipRTC
// for gfx950 and caches the module.  The generated code performs the same floating-point
// operations in the same order as the generic interpreted kernels (fma chains, no contraction),
// so the FP64 variant returns the same bits as k_sim_fp64.
//
// Two variants are generated:
//   fp64   double state; stores the fit window of y (SoA) and yss/overshoot (the FP64 chain).
//   mixed  float state; stores the fit window of y as float (the mixed chain).
#include <hip/hip_runtime.h>
#include <hip/hiprtc.h>

#include <cmath>
#include <cstdio>
#include <map>
#include <mutex>
#include <sstream>
#include <string>
#include <vector>

#include "ce_jit.h"

namespace ce {

static std::string lit(double v) {
  if (std::isnan(v)) return "__builtin_nan(\"\")";
  if (std::isinf(v)) return v > 0 ? "__builtin_inf()" : "(-__builtin_inf())";
  char buf[64];
  std::snprintf(buf, sizeof buf, "%a", v);
  return std::string("(") + buf + ")";
}

static std::string litf(float v) {
  if (std::isnan(v)) return "__builtin_nanf(\"\")";
  if (std::isinf(v)) return v > 0 ? "__builtin_inff()" : "(-__builtin_inff())";
  char buf[64];
  std::snprintf(buf, sizeof buf, "%af", (double)v);
  return std::string("(") + buf + ")";
}

namespace {

struct Gen {
  bool f32;
  std::ostringstream o;
  int tmp = 0;
  std::string T() const { return f32 ? "float" : "double"; }
  std::string L(double v) const { return f32 ? litf((float)v) : lit(v); }
  std::string fma() const { return f32 ? "__builtin_fmaf" : "__builtin_fma"; }
  std::string fmin() const { return f32 ? "__builtin_fminf" : "__builtin_fmin"; }
  std::string fmax() const { return f32 ? "__builtin_fmaxf" : "__builtin_fmax"; }
  std::string zero() const { return f32 ? "0.0f" : "0.0"; }

  // One instruction chain (reference src/main.cu:260-291) as straight-line code; returns the
  // variable holding the chain value.  src: names of the source entries; seed: name or "".
  std::string chain(const Ins* ops, int count, const std::vector<std::string>& src, const std::string& seed) {
    const std::string res = "r" + std::to_string(tmp++);
    o << "    " << T() << " " << res << " = " << zero() << ";\n";
    bool started = false;
    for (int r = 0; r < count; r++) {
      const Ins& in = ops[r];
      if (in.op == 0) continue;
      const std::string pos = src[in.col];
      std::string prev;
      if (started) {
        prev = res;
      } else if (!seed.empty()) {
        prev = "p" + std::to_string(tmp++);
        o << "    const " << T() << " " << prev << " = (" << seed << " == " << seed << ") ? " << seed << " : " << pos << ";\n";
      } else {
        prev = pos;
      }
      auto opnd = [&](int kind, double val) {
        return kind == K_CONST ? L(val) : (kind == K_STATE ? pos : prev);
      };
      const std::string a = opnd(in.akind, in.aval), b = opnd(in.bkind, in.bval);
      switch (in.op) {
        case 1: o << "    " << res << " = " << a << " + " << b << ";\n"; break;
        case 2: o << "    " << res << " = " << a << " - " << b << ";\n"; break;
        case 3: o << "    " << res << " = " << a << " * " << b << ";\n"; break;
        case 4: o << "    " << res << " = " << a << " / " << b << ";\n"; break;
        case 5: o << "    " << res << " = " << fmin() << "(" << fmax() << "(" << prev << ", " << a << "), " << b << ");\n"; break;
        case 6: o << "    " << res << " = (" << T() << ")((" << zero() << " < " << a << ") - (" << a << " < " << zero() << "));\n"; break;
        default: break;
      }
      started = true;
    }
    return res;
  }
};

}  // namespace

std::string gen_sim_source(const ProblemDev& P, bool mixed) {
  Gen g;
  g.f32 = mixed;
  const int ns = P.ns, nd = P.nd, nw = P.nw;
  std::ostringstream& o = g.o;
  const std::string T = g.T();
  o << "// generated: closed-loop simulation, ns=" << ns << " nd=" << nd << " nw=" << nw << (mixed ? " float" : " double") << "\n";
  o << "extern \"C\" __global__ __launch_bounds__(64) void " << (mixed ? "ce_sim_f32" : "ce_sim_f64")
    << "(const float* __restrict__ gains, const int* __restrict__ flags,\n"
       "    const int* __restrict__ list, const int* __restrict__ count, void* __restrict__ ybuf, double* __restrict__ info,\n"
       "    float* __restrict__ dy, int n) {\n";
  o << "  const int idx = blockIdx.x * 64 + threadIdx.x;\n  if (idx >= *count) return;\n  const int c = list[idx];\n";
  o << "  const bool unstable = (flags[c] & 1) != 0;\n";
  o << "  const float* g = gains + (long)c * " << P.cuda_w << ";\n";
  for (int i = 0; i < ns; i++) o << "  const " << T << " K" << i << " = (" << T << ")g[" << i << "], L" << i << " = (" << T << ")g[" << ns + i << "];\n";
  o << "  const " << T << " N = (" << T << ")g[" << 2 * ns << "];\n";
  for (int i = 0; i < ns; i++) o << "  " << T << " xh" << i << " = " << g.zero() << ";\n";
  for (int i = 0; i < nd; i++) o << "  " << T << " xp" << i << " = " << g.zero() << ";\n";
  o << "  " << T << " ysum = " << g.zero() << ", ymax = -" << (mixed ? "__builtin_inff()" : "__builtin_inf()") << ", ymin = "
    << (mixed ? "__builtin_inff()" : "__builtin_inf()") << ";\n";
  const int f0 = P.fit_start_id, f1 = P.fit_start_id + P.fit_count, s0 = P.steady_start_id;
  o << "  " << T << "* yb = (" << T << "*)ybuf + idx;\n";
  o << "  for (int t = 0; t < " << P.T << "; t++) {\n";
  std::vector<std::string> sxp, sxh, snl;
  for (int i = 0; i < nd; i++) sxp.push_back("xp" + std::to_string(i));
  for (int i = 0; i < ns; i++) sxh.push_back("xh" + std::to_string(i));
  const std::string y = g.chain(P.dc_ops, P.dc_n, sxp, "");
  o << "    const " << T << " y = " << y << ";\n";
  o << "    if (dy) dy[(long)c * " << P.T << " + t] = (float)y;\n";
  o << "    if (t >= " << f0 << " && t < " << f1 << ") yb[(long)(t - " << f0 << ") * n] = y;\n";
  o << "    if (t >= " << s0 << ") ysum = ysum + y;\n";
  o << "    ymax = " << g.fmax() << "(ymax, y); ymin = " << g.fmin() << "(ymin, y);\n";
  o << "    " << T << " kx = " << g.zero() << ";\n";
  for (int k = 0; k < ns; k++) o << "    kx = " << g.fma() << "(K" << k << ", xh" << k << ", kx);\n";
  o << "    const " << T << " rr = (t < " << P.ref_step_id << ") ? " << g.zero() << " : " << g.L(P.ref_amp) << ";\n";
  o << "    const " << T << " u_raw = " << g.fma() << "(N, rr, -kx);\n";
  const std::string u = g.chain(P.u_ops, P.u_n, sxh, "u_raw");
  o << "    const " << T << " uu = " << u << ";\n";
  o << "    " << T << " cx = " << g.zero() << ";\n";
  for (int k = 0; k < ns; k++) o << "    cx = " << g.fma() << "(" << g.L(P.sc[k]) << ", xh" << k << ", cx);\n";
  o << "    const " << T << " yerr = y - " << g.fma() << "(" << g.L(P.sd) << ", uu, cx);\n";
  for (int i = 0; i < ns; i++) {
    o << "    " << T << " dh" << i << " = " << g.fma() << "(L" << i << ", yerr, " << g.L(P.sb[i]) << " * uu);\n";
    for (int k = 0; k < ns; k++) o << "    dh" << i << " = " << g.fma() << "(" << g.L(P.sai[i * ns + k]) << ", xh" << k << ", dh" << i << ");\n";
  }
  for (int i = 0; i < ns; i++) o << "    xh" << i << " = xh" << i << " + dh" << i << ";\n";
  snl = sxp;
  snl.push_back("y");
  std::vector<std::string> W;
  for (int q = 0; q < nw; q++) W.push_back(g.chain(P.nl_ops[q], P.nl_n[q], snl, ""));
  for (int i = 0; i < nd; i++) {
    o << "    " << T << " ph" << i << " = " << g.L(P.dbu[i]) << " * uu;\n";
    for (int q = 0; q < nw; q++) o << "    ph" << i << " = " << g.fma() << "(" << g.L(P.dbw[i * nw + q]) << ", " << W[q] << ", ph" << i << ");\n";
    for (int k = 0; k < nd; k++) o << "    ph" << i << " = " << g.fma() << "(" << g.L(P.dai[i * nd + k]) << ", xp" << k << ", ph" << i << ");\n";
  }
  for (int i = 0; i < nd; i++) o << "    xp" << i << " = xp" << i << " + ph" << i << ";\n";
  o << "  }\n";
  o << "  if (unstable) return;\n";
  // yss and overshoot (reference src/modalsetup.cu:52-62)
  o << "  const " << T << " yss = ysum / (" << T << ")" << (P.T - s0) << ";\n";
  o << "  const " << T << " pk = " << g.fmax() << "(" << g.zero() << ", ((yss < " << g.zero() << ") ? ymin : ymax) * yss);\n";
  o << "  double* fr = info + (long)c * " << P.info_w << " + " << P.info_freqresp << ";\n";
  if (!mixed) {
    o << "  fr[2] = pk / __builtin_fmax(yss * yss, " << lit(EPS6) << ") - 1.0;\n  fr[3] = yss;\n";
  } else {
    o << "  fr[2] = (double)(pk / __builtin_fmaxf(yss * yss, " << litf((float)EPS6) << ") - 1.0f);\n  fr[3] = (double)yss;\n";
  }
  o << "}\n";
  return o.str();
}

// ------------------------------------------------------------------------------------------------
struct JitEntry {
  hipModule_t mod = nullptr;
  hipFunction_t fn = nullptr;
};

struct JitCache {
  std::mutex mu;
  std::map<std::string, JitEntry> entries;
};

JitCache* jit_cache_create() { return new JitCache(); }

void jit_cache_destroy(JitCache* c) {
  if (!c) return;
  for (auto& kv : c->entries)
    if (kv.second.mod) hipModuleUnload(kv.second.mod);
  delete c;
}

bool jit_enabled() {
  const char* v = std::getenv("CE_JIT");
  return !(v && std::string(v) == "0");
}

bool jit_get(JitCache* cache, const std::string& src, void** fn_out, std::string& err) {
  std::lock_guard<std::mutex> lk(cache->mu);
  auto it = cache->entries.find(src);
  if (it != cache->entries.end()) {
    *fn_out = (void*)it->second.fn;
    return true;
  }
  hiprtcProgram prog;
  if (hiprtcCreateProgram(&prog, src.c_str(), "ce_sim.hip", 0, nullptr, nullptr) != HIPRTC_SUCCESS) {
    err = "hiprtcCreateProgram failed";
    return false;
  }
  const char* opts[] = {"--offload-arch=gfx950", "-O3", "-ffp-contract=off", "-Xclang", "-target-feature", "-Xclang",
                        "-packed-fp32-ops"};
  const hiprtcResult rc = hiprtcCompileProgram(prog, 7, opts);
  if (rc != HIPRTC_SUCCESS) {
    size_t ls = 0;
    hiprtcGetProgramLogSize(prog, &ls);
    std::string log(ls + 1, '\0');
    hiprtcGetProgramLog(prog, &log[0]);
    err = "hipRTC compile of the simulation kernel failed: " + log;
    hiprtcDestroyProgram(&prog);
    return false;
  }
  size_t cs = 0;
  hiprtcGetCodeSize(prog, &cs);
  std::vector<char> code(cs);
  hiprtcGetCode(prog, code.data());
  hiprtcDestroyProgram(&prog);
  JitEntry e;
  if (hipModuleLoadData(&e.mod, code.data()) != hipSuccess) {
    err = "hipModuleLoadData failed for the simulation kernel";
    return false;
  }
  const char* kname = src.find("void ce_sim_f32(") != std::string::npos ? "ce_sim_f32" : "ce_sim_f64";
  if (hipModuleGetFunction(&e.fn, e.mod, kname) != hipSuccess) {
    err = std::string("hipModuleGetFunction(") + kname + ") failed";
    hipModuleUnload(e.mod);
    return false;
  }
  cache->entries[src] = e;
  *fn_out = (void*)e.fn;
  return true;
}

bool jit_launch_sim(void* fn, int grid, hipStream_t st, const float* gains, const int* flags, const int* list,
                    const int* count, void* ybuf, double* info, float* dy, int n) {
  void* args[] = {&gains, &flags, &list, &count, &ybuf, &info, &dy, &n};
  return hipModuleLaunchKernel((hipFunction_t)fn, grid, 1, 1, 64, 1, 1, 0, st, args, nullptr) == hipSuccess;
}

}  // namespace ce

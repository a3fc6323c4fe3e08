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
//   mixed  float state kept as hi/lo pairs with the reference's compensated update (TwoSum then
//          Fast2Sum, src/main.cu:129-155; the lo parts enter through dot products of their own,
//          src/discrete_solves.cu:35-77); stores the fit window of y as float (the mixed chain).
#include <hip/hip_runtime.h>
#include <hip/hiprtc.h>

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <list>
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
  std::string fmin() const { return f32 ? "ce_minf" : "ce_min"; }
  std::string fmax() const { return f32 ? "ce_maxf" : "ce_max"; }
  std::string zero() const { return f32 ? "0.0f" : "0.0"; }

  // One instruction chain (reference src/main.cu:260-291) as straight-line code; returns the
  // variable holding the chain value.  src: names of the source entries; seed: name or "".
  // cst(r, which, value): the expression of instruction r's operand constant (which: 0 a, 1 b).
  template <typename Cst>
  std::string chain(const Ins* ops, int count, const std::vector<std::string>& src, const std::string& seed, Cst cst) {
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
      auto opnd = [&](int kind, int which, double val) {
        return kind == K_CONST ? cst(r, which, val) : (kind == K_STATE ? pos : prev);
      };
      const std::string a = opnd(in.akind, 0, in.aval), b = opnd(in.bkind, 1, in.bval);
      // float form: SGN(x) followed by MUL(prev, x) on the same column is |x| (equal for every x
      // but zero, where only the sign of the zero differs), a free source modifier of the next use
      if (f32 && in.op == 6 && in.akind == K_STATE) {
        int r2 = r + 1;
        while (r2 < count && ops[r2].op == 0) r2++;
        if (r2 < count && ops[r2].op == 3 && ops[r2].col == in.col &&
            ((ops[r2].akind == K_PREV && ops[r2].bkind == K_STATE) ||
             (ops[r2].akind == K_STATE && ops[r2].bkind == K_PREV))) {
          o << "    " << res << " = __builtin_fabsf(" << pos << ");\n";
          started = true;
          r = r2;
          continue;
        }
      }
      switch (in.op) {
        case 1: o << "    " << res << " = " << a << " + " << b << ";\n"; break;
        case 2: o << "    " << res << " = " << a << " - " << b << ";\n"; break;
        case 3: o << "    " << res << " = " << a << " * " << b << ";\n"; break;
        case 4: o << "    " << res << " = " << a << " / " << b << ";\n"; break;
        case 5: o << "    " << res << " = " << fmin() << "(" << fmax() << "(" << prev << ", " << a << "), " << b << ");\n"; break;
        case 6:  // (0 < a) - (a < 0): +-1 with a's sign, +0 for zero and NaN
          o << "    " << res << " = (" << a << " < " << zero() << " || " << a << " > " << zero() << ") ? "
            << (f32 ? "__builtin_copysignf(1.0f, " : "__builtin_copysign(1.0, ") << a << ") : " << zero() << ";\n";
          break;
        default: break;
      }
      started = true;
    }
    return res;
  }
};

}  // namespace

// The time loop over the generated step: the horizon in segments between the phase boundaries,
// within each of which the phase flags are compile-time constants.
static void gen_phase_loop(std::ostringstream& o) {
  o << "  const bool wdy = dy != nullptr;\n"
       "  int t = 0;\n"
       "  while (t < TT) {\n"
       "    int e = TT;\n"
       "    if (rs > t && rs < e) e = rs;\n"
       "    if (f0 > t && f0 < e) e = f0;\n"
       "    if (f1 > t && f1 < e) e = f1;\n"
       "    if (s0 > t && s0 < e) e = s0;\n"
       "    const bool st = t >= f0 && t < f1, ss = t >= s0, rz = t < rs;\n"
       "    if (wdy) {\n"
       "      for (; t < e; t++) step(t, st, ss, rz, true);\n"
       "      continue;\n"
       "    }\n"
       "    switch ((st ? 4 : 0) | (ss ? 2 : 0) | (rz ? 1 : 0)) {\n";
  for (int k = 0; k < 8; k++)
    o << "      case " << k << ": for (; t < e; t++) step(t, BC<" << (k >> 2 & 1) << ">{}, BC<" << (k >> 1 & 1) << ">{}, BC<"
      << (k & 1) << ">{}, BC<0>{}); break;\n";
  o << "    }\n  }\n";
}

// Per-candidate body shared by both forms.  Literal form: every constant and size is written into
// the source.  Structural form: only the dimensions and the instruction structure (opcodes,
// columns, operand kinds) are fixed; constants and sizes are loaded from the problem descriptor
// (Pp, byte offsets into ProblemDev), so one compiled kernel serves every problem of that structure.
// Structural form, constants of the time step: with `hoist`, each is loaded once before the loop
// into a named register value (kcN); otherwise they are read every step through Pq, which an empty
// asm statement redefines (in scalar registers) every step, so the loads stay scalar loads from the
// scalar cache instead of being hoisted into vector registers for the whole loop (for structures
// with more constants than registers to spare).  With vbudget > 0 (float form), the first vbudget
// float constants the step reads are copied into vector registers once before the loop (an opaque
// v_mov_b32, so they stay there) and the rest are read every step through Pq.  *n_hoisted: 32-bit
// words of hoisted constants.
static void gen_body(Gen& g, const ProblemDev& P, bool mixed, bool structural, bool hoist = false,
                     int* n_hoisted = nullptr, int vbudget = 0) {
  const int ns = P.ns, nd = P.nd, nw = P.nw;
  std::ostringstream& o = g.o;
  const std::string T = g.T();
  bool in_loop = false;
  std::map<size_t, std::string> hname;
  std::ostringstream hdecl;
  int hwords = 0;
  // loads through the constant address space: read-only, so uniform addresses become scalar loads
  auto ld0 = [&](const char* ty, size_t off, const char* base) {
    return std::string("(*(const __attribute__((address_space(4))) ") + ty + "*)((const __attribute__((address_space(4))) char*)" +
           base + " + " + std::to_string(off) + "))";
  };
  auto ld = [&](const char* ty, size_t off) {
    if (!in_loop) return ld0(ty, off, "Pp");
    if (vbudget > 0) {
      auto it = hname.find(off);
      if (it != hname.end()) return it->second;
      if (std::string(ty) != "float" || hwords + 1 > vbudget) return ld0(ty, off, "Pq");
      const std::string nm = "kv" + std::to_string(hname.size());
      hdecl << "  float " << nm << ";\n  asm volatile(\"v_mov_b32 %0, %1\" : \"=v\"(" << nm << ") : \"s\"("
            << ld0(ty, off, "Pp") << "));\n";
      hwords += 1;
      hname[off] = nm;
      return nm;
    }
    if (!hoist) return ld0(ty, off, "Pq");
    auto it = hname.find(off);
    if (it != hname.end()) return it->second;
    const std::string nm = "kc" + std::to_string(hname.size());
    hdecl << "  const " << ty << " " << nm << " = " << ld0(ty, off, "Pp") << ";\n";
    hwords += (std::string(ty) == "double") ? 2 : 1;
    hname[off] = nm;
    return nm;
  };
  // plant / observer coefficient: double field (FP64 form) or its float copy (mixed form)
  auto K = [&](double v, size_t offd, size_t offf) {
    if (!structural) return g.L(v);
    return mixed ? ld("float", offf) : ld("double", offd);
  };
  auto I = [&](int v, size_t off) { return structural ? ld("int", off) : std::to_string(v); };
#define CE_OFF(f) offsetof(ProblemDev, f)
  const std::string sT = I(P.T, CE_OFF(T)), sF0 = I(P.fit_start_id, CE_OFF(fit_start_id));
  const std::string sFC = I(P.fit_count, CE_OFF(fit_count)), sS0 = I(P.steady_start_id, CE_OFF(steady_start_id));
  const std::string sRS = I(P.ref_step_id, CE_OFF(ref_step_id)), sCW = I(P.cuda_w, CE_OFF(cuda_w));
  const std::string sIW = I(P.info_w, CE_OFF(info_w)), sFR = I(P.info_freqresp, CE_OFF(info_freqresp));
  o << "  const int idx = bx * 64 + threadIdx.x;\n  if (idx >= *A.count) return;\n  const int c = A.list[idx];\n";
  o << "  const int n = A.n;\n  const int TT = " << sT << ", f0 = " << sF0 << ", f1 = " << sF0 << " + " << sFC
    << ", s0 = " << sS0 << ", rs = " << sRS << ";\n";
  o << "  const bool unstable = (A.flags[c] & 1) != 0;\n";
  o << "  const float* g = A.gains + (long)c * " << sCW << ";\n";
  for (int i = 0; i < ns; i++) o << "  const " << T << " K" << i << " = (" << T << ")g[" << i << "], L" << i << " = (" << T << ")g[" << ns + i << "];\n";
  o << "  const " << T << " N = (" << T << ")g[" << 2 * ns << "];\n";
  for (int i = 0; i < ns; i++) o << "  " << T << " xh" << i << " = " << g.zero() << ";\n";
  for (int i = 0; i < nd; i++) o << "  " << T << " xp" << i << " = " << g.zero() << ";\n";
  if (mixed) {  // lo parts of the compensated state
    for (int i = 0; i < ns; i++) o << "  float xhl" << i << " = 0.0f;\n";
    for (int i = 0; i < nd; i++) o << "  float xpl" << i << " = 0.0f;\n";
  }
  o << "  " << T << " ysum = " << g.zero() << ", ymax = -" << (mixed ? "__builtin_inff()" : "__builtin_inf()") << ", ymin = "
    << (mixed ? "__builtin_inff()" : "__builtin_inf()") << ";\n";
  o << "  " << T << "* yb = (" << T << "*)A.ybuf + idx;\n";
  o << "  float* dy = A.dy;\n";
  o << "  const " << T << " amp = " << K(P.ref_amp, CE_OFF(ref_amp), CE_OFF(fref_amp)) << ";\n";
  const bool per_step = structural && (!hoist || vbudget > 0);  // some constants read every step
  const bool hoisting = structural && (hoist || vbudget > 0);
  if (per_step) o << "  const char* Pq = Pp;\n";
  const std::string hoist_mark = "  /*hoisted constants*/\n";
  if (hoisting) o << hoist_mark;
  // one time step; the phase flags (fit window stored, steady-state sum, reference still zero,
  // full-trajectory dump) are compile-time constants in the phase loops below, so a step carries
  // no per-step branches besides the loop's own
  o << "  auto step = [&](const int t, auto ST, auto SS, auto RZ, auto DY) __attribute__((always_inline)) {\n";
  if (per_step)
    o << "    asm volatile(\"\" : \"+s\"(Pq));\n"
         "    Pq = (const char*)(((unsigned long)__builtin_amdgcn_readfirstlane((unsigned)((unsigned long)Pq >> 32)) << 32) |\n"
         "                       (unsigned long)(unsigned)__builtin_amdgcn_readfirstlane((unsigned)(unsigned long)Pq));\n";
  in_loop = true;
  std::vector<std::string> sxp, sxh, snl;
  for (int i = 0; i < nd; i++) sxp.push_back("xp" + std::to_string(i));
  for (int i = 0; i < ns; i++) sxh.push_back("xh" + std::to_string(i));
  auto ins_cst = [&](const Ins* ops, size_t off_ins, size_t off_fa, size_t off_fb) {
    return [&, ops, off_ins, off_fa, off_fb](int r, int which, double v) {
      if (!structural) return g.L(v);
      if (mixed) return ld("float", (which ? off_fb : off_fa) + (size_t)r * sizeof(float));
      return ld("double", off_ins + (size_t)r * sizeof(Ins) + (which ? offsetof(Ins, bval) : offsetof(Ins, aval)));
    };
  };
  const std::string y = g.chain(P.dc_ops, P.dc_n, sxp, "", ins_cst(P.dc_ops, CE_OFF(dc_ops), CE_OFF(fdc_aval), CE_OFF(fdc_bval)));
  o << "    const " << T << " y = " << y << ";\n";
  o << "    if (DY) dy[(long)c * TT + t] = (float)y;\n";
  o << "    if (ST) yb[(long)(t - f0) * n] = y;\n";
  o << "    if (SS) ysum = ysum + y;\n";
  o << "    ymax = " << g.fmax() << "(ymax, y); ymin = " << g.fmin() << "(ymin, y);\n";
  o << "    " << T << " kx = " << g.zero() << ";\n";
  for (int k = 0; k < ns; k++) o << "    kx = " << g.fma() << "(K" << k << ", xh" << k << ", kx);\n";
  if (mixed)
    for (int k = 0; k < ns; k++) o << "    kx = " << g.fma() << "(K" << k << ", xhl" << k << ", kx);\n";
  o << "    const " << T << " rr = RZ ? " << g.zero() << " : amp;\n";
  o << "    const " << T << " u_raw = " << g.fma() << "(N, rr, -kx);\n";
  const std::string u = g.chain(P.u_ops, P.u_n, sxh, "u_raw", ins_cst(P.u_ops, CE_OFF(u_ops), CE_OFF(fu_aval), CE_OFF(fu_bval)));
  o << "    const " << T << " uu = " << u << ";\n";
  if (!mixed) {
    // FP64: every term, in the interpreted kernel's order (bit parity with the oracle)
    o << "    " << T << " cx = " << g.zero() << ";\n";
    for (int k = 0; k < ns; k++)
      o << "    cx = " << g.fma() << "(" << K(P.sc[k], CE_OFF(sc) + k * 8, CE_OFF(fsc) + k * 4) << ", xh" << k << ", cx);\n";
    o << "    const " << T << " yerr = y - " << g.fma() << "(" << K(P.sd, CE_OFF(sd), CE_OFF(fsd)) << ", uu, cx);\n";
    for (int i = 0; i < ns; i++) {
      o << "    " << T << " dh" << i << " = " << g.fma() << "(L" << i << ", yerr, " << K(P.sb[i], CE_OFF(sb) + i * 8, CE_OFF(fsb) + i * 4) << " * uu);\n";
      for (int k = 0; k < ns; k++)
        o << "    dh" << i << " = " << g.fma() << "(" << K(P.sai[i * ns + k], CE_OFF(sai) + (i * ns + k) * 8, CE_OFF(fsai) + (i * ns + k) * 4)
          << ", xh" << k << ", dh" << i << ");\n";
    }
  } else {
    // float: terms whose coefficient is zero in float are left out (fma(0, x, a) = a for finite x;
    // the zero pattern is part of the structure key of the batched kernels)
    auto acc = [&](const std::string& var, const std::vector<std::pair<std::string, std::string>>& terms,
                   const std::string& init) {
      o << "    " << T << " " << var << " = " << init << ";\n";
      for (auto& t : terms) o << "    " << var << " = " << g.fma() << "(" << t.first << ", " << t.second << ", " << var << ");\n";
    };
    std::vector<std::pair<std::string, std::string>> tc;
    for (int k = 0; k < ns; k++)
      if ((float)P.sc[k] != 0.0f) tc.push_back({K(P.sc[k], CE_OFF(sc) + k * 8, CE_OFF(fsc) + k * 4), "xh" + std::to_string(k)});
    for (int k = 0; k < ns; k++)
      if ((float)P.sc[k] != 0.0f) tc.push_back({K(P.sc[k], CE_OFF(sc) + k * 8, CE_OFF(fsc) + k * 4), "xhl" + std::to_string(k)});
    if ((float)P.sd != 0.0f) tc.push_back({K(P.sd, CE_OFF(sd), CE_OFF(fsd)), "uu"});
    acc("cx", tc, g.zero());
    o << "    const " << T << " yerr = y - cx;\n";
    for (int i = 0; i < ns; i++) {
      // fma(L, yerr, sb u) as in the full form; with sb = 0 it is the rounded product L yerr
      const std::string init = (float)P.sb[i] != 0.0f
          ? g.fma() + "(L" + std::to_string(i) + ", yerr, " + K(P.sb[i], CE_OFF(sb) + i * 8, CE_OFF(fsb) + i * 4) + " * uu)"
          : "L" + std::to_string(i) + " * yerr";
      std::vector<std::pair<std::string, std::string>> td, tl;
      for (int k = 0; k < ns; k++)
        if ((float)P.sai[i * ns + k] != 0.0f) {
          const std::string cst = K(P.sai[i * ns + k], CE_OFF(sai) + (i * ns + k) * 8, CE_OFF(fsai) + (i * ns + k) * 4);
          td.push_back({cst, "xh" + std::to_string(k)});
          tl.push_back({cst, "xhl" + std::to_string(k)});
        }
      acc("dh" + std::to_string(i), td, init);
      acc("dl" + std::to_string(i), tl, g.zero());
    }
  }
  for (int i = 0; i < ns; i++) {
    if (mixed) o << "    ce_hladd(xh" << i << ", xhl" << i << ", dh" << i << ", dl" << i << ");\n";
    else o << "    xh" << i << " = xh" << i << " + dh" << i << ";\n";
  }
  snl = sxp;
  snl.push_back("y");
  std::vector<std::string> W;
  for (int q = 0; q < nw; q++) {
    const size_t oi = CE_OFF(nl_ops) + (size_t)q * CE_MAX_NL_INSTRUCTIONS * sizeof(Ins);
    const size_t ofa = CE_OFF(fnl_aval) + (size_t)q * CE_MAX_NL_INSTRUCTIONS * sizeof(float);
    const size_t ofb = CE_OFF(fnl_bval) + (size_t)q * CE_MAX_NL_INSTRUCTIONS * sizeof(float);
    W.push_back(g.chain(P.nl_ops[q], P.nl_n[q], snl, "", ins_cst(P.nl_ops[q], oi, ofa, ofb)));
  }
  for (int i = 0; i < nd; i++) {
    if (!mixed) {
      o << "    " << T << " ph" << i << " = " << K(P.dbu[i], CE_OFF(dbu) + i * 8, CE_OFF(fdbu) + i * 4) << " * uu;\n";
      for (int q = 0; q < nw; q++)
        o << "    ph" << i << " = " << g.fma() << "(" << K(P.dbw[i * nw + q], CE_OFF(dbw) + (i * nw + q) * 8, CE_OFF(fdbw) + (i * nw + q) * 4)
          << ", " << W[q] << ", ph" << i << ");\n";
      for (int k = 0; k < nd; k++)
        o << "    ph" << i << " = " << g.fma() << "(" << K(P.dai[i * nd + k], CE_OFF(dai) + (i * nd + k) * 8, CE_OFF(fdai) + (i * nd + k) * 4)
          << ", xp" << k << ", ph" << i << ");\n";
    } else {
      o << "    " << T << " ph" << i << " = " << g.zero() << ";\n";
      o << "    " << T << " pl" << i << " = " << g.zero() << ";\n";
      auto term = [&](const char* acc, double v, size_t od, size_t of, const std::string& x) {
        if ((float)v != 0.0f) o << "    " << acc << i << " = " << g.fma() << "(" << K(v, od, of) << ", " << x << ", " << acc << i << ");\n";
      };
      term("ph", P.dbu[i], CE_OFF(dbu) + i * 8, CE_OFF(fdbu) + i * 4, "uu");
      for (int q = 0; q < nw; q++) term("ph", P.dbw[i * nw + q], CE_OFF(dbw) + (i * nw + q) * 8, CE_OFF(fdbw) + (i * nw + q) * 4, W[q]);
      for (int k = 0; k < nd; k++) term("ph", P.dai[i * nd + k], CE_OFF(dai) + (i * nd + k) * 8, CE_OFF(fdai) + (i * nd + k) * 4, "xp" + std::to_string(k));
      for (int k = 0; k < nd; k++) term("pl", P.dai[i * nd + k], CE_OFF(dai) + (i * nd + k) * 8, CE_OFF(fdai) + (i * nd + k) * 4, "xpl" + std::to_string(k));
    }
  }
  for (int i = 0; i < nd; i++) {
    if (mixed) o << "    ce_hladd(xp" << i << ", xpl" << i << ", ph" << i << ", pl" << i << ");\n";
    else o << "    xp" << i << " = xp" << i << " + ph" << i << ";\n";
  }
  o << "  };\n";
  in_loop = false;
  if (hoisting) {
    std::string text = o.str();
    const size_t at = text.find(hoist_mark);
    text.replace(at, hoist_mark.size(), hdecl.str());
    o.str(text);
    o.seekp(0, std::ios_base::end);
  }
  if (n_hoisted) *n_hoisted = hwords;
  gen_phase_loop(o);
  o << "  if (unstable) return;\n";
  // yss and overshoot (reference src/modalsetup.cu:52-62)
  o << "  const " << T << " yss = ysum / (" << T << ")(TT - s0);\n";
  o << "  const " << T << " pk = " << g.fmax() << "(" << g.zero() << ", ((yss < " << g.zero() << ") ? ymin : ymax) * yss);\n";
  o << "  double* fr = A.info + (long)c * " << sIW << " + " << sFR << ";\n";
  if (!mixed) {
    o << "  fr[2] = pk / __builtin_fmax(yss * yss, " << lit(EPS6) << ") - 1.0;\n  fr[3] = yss;\n";
  } else {
    o << "  fr[2] = (double)(pk / __builtin_fmaxf(yss * yss, " << litf((float)EPS6) << ") - 1.0f);\n  fr[3] = (double)yss;\n";
  }
#undef CE_OFF
}

// Lane-group form of the float kernels: G lanes (2 or 4) per candidate, for launches whose
// candidates leave SIMDs idle at one lane each.  Every lane of a group runs the instruction
// chains, K x and C x over the whole state, as the one-lane form does; lane j updates the rows
// G s + j of the observer and plant state (the same terms in the same order, then the compensated
// add), and each group's new rows are broadcast to its lanes with DPP quad permutes.  Every row
// therefore sees the one-lane form's operations: the trajectory is bit-identical.  A lane's row
// coefficients are selected once, before the time loop: literals, or (structural form) loads from
// the problem descriptor Pp, whose zero pattern is part of the structure.
static void gen_body_grp(Gen& g, const ProblemDev& P, int G, bool structural) {
  const int ns = P.ns, nd = P.nd, nw = P.nw;
  const int lg = G == 4 ? 2 : 1;
  std::ostringstream& o = g.o;
  std::ostringstream pre;
  int nsel = 0, nk = 0;
#define CE_OFF(f) offsetof(ProblemDev, f)
  auto ldp = [](const char* ty, size_t off) {
    return std::string("(*(const __attribute__((address_space(4))) ") + ty + "*)((const __attribute__((address_space(4))) char*)Pp + " +
           std::to_string(off) + "))";
  };
  // a wave-uniform constant of the step: a literal, or loaded once before the loop
  auto cst = [&](double v, size_t off) -> std::string {
    if (!structural) return g.L(v);
    const std::string nm = "kc" + std::to_string(nk++);
    pre << "  const float " << nm << " = " << ldp("float", off) << ";\n";
    return nm;
  };
  auto ival = [&](int v, size_t off) { return structural ? ldp("int", off) : std::to_string(v); };
  // rows G s + j of a per-row coefficient (fv: value of row r, fo: its float offset)
  auto zero_rows = [&](int s, int n, auto fv) {
    for (int j = 0; j < G; j++)
      if (G * s + j < n && (float)fv(G * s + j) != 0.0f) return false;
    return true;
  };
  auto sel = [&](int s, int n, auto fv, auto fo) -> std::string {
    if (!structural) {
      bool same = true;
      for (int j = 1; j < G; j++) {
        const float a = G * s + j < n ? (float)fv(G * s + j) : 0.0f, b = (float)fv(G * s);
        if (a != b) same = false;
      }
      if (same) return g.L(fv(G * s));
    }
    auto val = [&](int j) -> std::string {
      const int r = G * s + j;
      if (r >= n) return "0.0f";
      return structural ? ldp("float", fo(r)) : g.L(fv(r));
    };
    const std::string nm = "pc" + std::to_string(nsel++);
    pre << "  const float " << nm << " = ";
    for (int j = G - 1; j >= 1; j--) pre << "h == " << j << " ? " << val(j) << " : ";
    pre << val(0) << ";\n";
    return nm;
  };
  // DPP quad permute that gives every lane of a group the value of the group's lane j
  auto bc = [&](int j) {
    int p[4];
    for (int q = 0; q < 4; q++) p[q] = (q / G) * G + j;
    return "ce_bc<" + std::to_string(p[0] | p[1] << 2 | p[2] << 4 | p[3] << 6) + ">";
  };
  const int S = (ns + G - 1) / G, SP = (nd + G - 1) / G;
  const std::string sT = ival(P.T, CE_OFF(T)), sF0 = ival(P.fit_start_id, CE_OFF(fit_start_id));
  const std::string sFC = ival(P.fit_count, CE_OFF(fit_count)), sS0 = ival(P.steady_start_id, CE_OFF(steady_start_id));
  const std::string sRS = ival(P.ref_step_id, CE_OFF(ref_step_id)), sCW = ival(P.cuda_w, CE_OFF(cuda_w));
  const std::string sIW = ival(P.info_w, CE_OFF(info_w)), sFR = ival(P.info_freqresp, CE_OFF(info_freqresp));
  o << "  const int h = threadIdx.x & " << G - 1 << ";\n";
  o << "  const int idx = bx * " << 64 / G << " + (threadIdx.x >> " << lg << ");\n  if (idx >= *A.count) return;\n"
       "  const int c = A.list[idx];\n";
  o << "  const int n = A.n;\n  const int TT = " << sT << ", f0 = " << sF0 << ", f1 = " << sF0 << " + " << sFC
    << ", s0 = " << sS0 << ", rs = " << sRS << ";\n";
  o << "  const bool unstable = (A.flags[c] & 1) != 0;\n";
  o << "  const float* g = A.gains + (long)c * " << sCW << ";\n";
  for (int i = 0; i < ns; i++) o << "  const float K" << i << " = g[" << i << "];\n";
  for (int s = 0; s < S; s++)
    o << "  const float Lo" << s << " = (" << G * s << " + h < " << ns << ") ? g[" << ns + G * s << " + h] : 0.0f;\n";
  o << "  const float N = g[" << 2 * ns << "];\n";
  for (int i = 0; i < ns; i++) o << "  float xh" << i << " = 0.0f, xhl" << i << " = 0.0f;\n";
  for (int i = 0; i < nd; i++) o << "  float xp" << i << " = 0.0f, xpl" << i << " = 0.0f;\n";
  for (int s = 0; s < S; s++) o << "  float oh" << s << " = 0.0f, ol" << s << " = 0.0f;\n";
  for (int s = 0; s < SP; s++) o << "  float op" << s << " = 0.0f, opl" << s << " = 0.0f;\n";
  o << "  float ysum = 0.0f, ymax = -__builtin_inff(), ymin = __builtin_inff();\n";
  o << "  float* yb = (float*)A.ybuf + idx;\n  float* dy = A.dy;\n";
  const std::string mark = "  /*lane rows*/\n";
  o << mark;
  o << "  const float amp = " << cst(P.ref_amp, CE_OFF(fref_amp)) << ";\n";
  o << "  auto step = [&](const int t, auto ST, auto SS, auto RZ, auto DY) __attribute__((always_inline)) {\n";
  std::vector<std::string> sxp, sxh;
  for (int i = 0; i < nd; i++) sxp.push_back("xp" + std::to_string(i));
  for (int i = 0; i < ns; i++) sxh.push_back("xh" + std::to_string(i));
  auto chain_cst = [&](size_t off_a, size_t off_b) {
    return [&, off_a, off_b](int r, int which, double v) { return cst(v, (which ? off_b : off_a) + (size_t)r * sizeof(float)); };
  };
  const std::string y = g.chain(P.dc_ops, P.dc_n, sxp, "", chain_cst(CE_OFF(fdc_aval), CE_OFF(fdc_bval)));
  o << "    const float y = " << y << ";\n";
  o << "    if (DY) dy[(long)c * TT + t] = y;\n";
  o << "    if (ST) yb[(long)(t - f0) * n] = y;\n";
  o << "    if (SS) ysum = ysum + y;\n";
  o << "    ymax = ce_maxf(ymax, y); ymin = ce_minf(ymin, y);\n";
  o << "    float kx = 0.0f;\n";
  for (int k = 0; k < ns; k++) o << "    kx = __builtin_fmaf(K" << k << ", xh" << k << ", kx);\n";
  for (int k = 0; k < ns; k++) o << "    kx = __builtin_fmaf(K" << k << ", xhl" << k << ", kx);\n";
  o << "    const float rr = RZ ? 0.0f : amp;\n";
  o << "    const float u_raw = __builtin_fmaf(N, rr, -kx);\n";
  const std::string u = g.chain(P.u_ops, P.u_n, sxh, "u_raw", chain_cst(CE_OFF(fu_aval), CE_OFF(fu_bval)));
  o << "    const float uu = " << u << ";\n";
  o << "    float cx = 0.0f;\n";
  std::vector<std::string> csc(ns);
  for (int k = 0; k < ns; k++)
    if ((float)P.sc[k] != 0.0f) csc[k] = cst(P.sc[k], CE_OFF(fsc) + k * 4);
  for (int k = 0; k < ns; k++)
    if ((float)P.sc[k] != 0.0f) o << "    cx = __builtin_fmaf(" << csc[k] << ", xh" << k << ", cx);\n";
  for (int k = 0; k < ns; k++)
    if ((float)P.sc[k] != 0.0f) o << "    cx = __builtin_fmaf(" << csc[k] << ", xhl" << k << ", cx);\n";
  if ((float)P.sd != 0.0f) o << "    cx = __builtin_fmaf(" << cst(P.sd, CE_OFF(fsd)) << ", uu, cx);\n";
  o << "    const float yerr = y - cx;\n";
  for (int s = 0; s < S; s++) {
    auto fsb = [&](int r) { return P.sb[r]; };
    auto osb = [&](int r) { return CE_OFF(fsb) + (size_t)r * 4; };
    if (zero_rows(s, ns, fsb)) o << "    float dh" << s << " = Lo" << s << " * yerr;\n";
    else o << "    float dh" << s << " = __builtin_fmaf(Lo" << s << ", yerr, " << sel(s, ns, fsb, osb) << " * uu);\n";
    o << "    float dl" << s << " = 0.0f;\n";
    for (int k = 0; k < ns; k++) {
      auto fa = [&](int r) { return P.sai[r * ns + k]; };
      auto oa = [&](int r) { return CE_OFF(fsai) + (size_t)(r * ns + k) * 4; };
      if (zero_rows(s, ns, fa)) continue;
      const std::string cf = sel(s, ns, fa, oa);
      o << "    dh" << s << " = __builtin_fmaf(" << cf << ", xh" << k << ", dh" << s << ");\n";
      o << "    dl" << s << " = __builtin_fmaf(" << cf << ", xhl" << k << ", dl" << s << ");\n";
    }
  }
  for (int s = 0; s < S; s++) o << "    ce_hladd(oh" << s << ", ol" << s << ", dh" << s << ", dl" << s << ");\n";
  for (int s = 0; s < S; s++)
    for (int j = 0; j < G && G * s + j < ns; j++) {
      const int r = G * s + j;
      o << "    xh" << r << " = " << bc(j) << "(oh" << s << "); xhl" << r << " = " << bc(j) << "(ol" << s << ");\n";
    }
  std::vector<std::string> snl = sxp;
  snl.push_back("y");
  std::vector<std::string> W;
  for (int q = 0; q < nw; q++) {
    const size_t ofa = CE_OFF(fnl_aval) + (size_t)q * CE_MAX_NL_INSTRUCTIONS * sizeof(float);
    const size_t ofb = CE_OFF(fnl_bval) + (size_t)q * CE_MAX_NL_INSTRUCTIONS * sizeof(float);
    W.push_back(g.chain(P.nl_ops[q], P.nl_n[q], snl, "", chain_cst(ofa, ofb)));
  }
  for (int s = 0; s < SP; s++) {
    o << "    float ph" << s << " = 0.0f, pl" << s << " = 0.0f;\n";
    auto term = [&](const char* acc, auto fv, auto fo, const std::string& x) {
      if (!zero_rows(s, nd, fv)) o << "    " << acc << s << " = __builtin_fmaf(" << sel(s, nd, fv, fo) << ", " << x << ", " << acc << s << ");\n";
    };
    term("ph", [&](int r) { return P.dbu[r]; }, [&](int r) { return CE_OFF(fdbu) + (size_t)r * 4; }, "uu");
    for (int q = 0; q < nw; q++)
      term("ph", [&](int r) { return P.dbw[r * nw + q]; }, [&](int r) { return CE_OFF(fdbw) + (size_t)(r * nw + q) * 4; }, W[q]);
    for (int k = 0; k < nd; k++)
      term("ph", [&](int r) { return P.dai[r * nd + k]; }, [&](int r) { return CE_OFF(fdai) + (size_t)(r * nd + k) * 4; },
           "xp" + std::to_string(k));
    for (int k = 0; k < nd; k++)
      term("pl", [&](int r) { return P.dai[r * nd + k]; }, [&](int r) { return CE_OFF(fdai) + (size_t)(r * nd + k) * 4; },
           "xpl" + std::to_string(k));
  }
  for (int s = 0; s < SP; s++) o << "    ce_hladd(op" << s << ", opl" << s << ", ph" << s << ", pl" << s << ");\n";
  for (int s = 0; s < SP; s++)
    for (int j = 0; j < G && G * s + j < nd; j++) {
      const int r = G * s + j;
      o << "    xp" << r << " = " << bc(j) << "(op" << s << "); xpl" << r << " = " << bc(j) << "(opl" << s << ");\n";
    }
  o << "  };\n";
  {
    std::string text = o.str();
    const size_t at = text.find(mark);
    text.replace(at, mark.size(), pre.str());
    o.str(text);
    o.seekp(0, std::ios_base::end);
  }
  gen_phase_loop(o);
  o << "  if (unstable || h != 0) return;\n";
  o << "  const float yss = ysum / (float)(TT - s0);\n";
  o << "  const float pk = ce_maxf(0.0f, ((yss < 0.0f) ? ymin : ymax) * yss);\n";
  o << "  double* fr = A.info + (long)c * " << sIW << " + " << sFR << ";\n";
  o << "  fr[2] = (double)(pk / __builtin_fmaxf(yss * yss, " << litf((float)EPS6) << ") - 1.0f);\n  fr[3] = (double)yss;\n";
#undef CE_OFF
}

static const char* kSimArgs =
    "struct SimArgs { const float* gains; const int* flags; const int* list; const int* count; void* ybuf;\n"
    "  double* info; float* dy; int n; int pad; };\n"
    "template <int B> struct BC { __device__ constexpr operator bool() const { return B != 0; } };\n"
    // first list slot of the split simulation's remainder (the rule of sim_split_slots, ce_stage.h)
    "static __device__ inline int ce_split(int live, int cap) {\n"
    "  if (cap <= 0) return live;\n"
    "  const int w = (live + 63) / 64;\n"
    "  const int k = w / cap, r = w - k * cap;\n"
    "  return (k >= 2 && r > 0 && 2 * r <= cap) ? k * cap * 64 : live; }\n"
    // fmin/fmax as the bare instructions: every operand here is an arithmetic result, a constant or
    // an infinity, never a signalling NaN, so they return what fmin/fmax return, without the
    // canonicalising moves the compiler adds for operands it cannot prove quiet
    "static __device__ inline float ce_maxf(float a, float b) { float r; asm(\"v_max_f32 %0, %1, %2\" : \"=v\"(r) : \"v\"(a), \"v\"(b)); return r; }\n"
    "static __device__ inline float ce_minf(float a, float b) { float r; asm(\"v_min_f32 %0, %1, %2\" : \"=v\"(r) : \"v\"(a), \"v\"(b)); return r; }\n"
    "static __device__ inline double ce_max(double a, double b) { double r; asm(\"v_max_f64 %0, %1, %2\" : \"=v\"(r) : \"v\"(a), \"v\"(b)); return r; }\n"
    "static __device__ inline double ce_min(double a, double b) { double r; asm(\"v_min_f64 %0, %1, %2\" : \"=v\"(r) : \"v\"(a), \"v\"(b)); return r; }\n"
    // compensated state update (HI, LO) += (h, l): TwoSum, then Fast2Sum
    "static __device__ inline __attribute__((always_inline)) void ce_hladd(float& H, float& Lo, float h, float l) {\n"
    "  const float sh = H + h; const float b0 = sh - H; const float sl = (H - (sh - b0)) + (h - b0);\n"
    "  const float v = (Lo + l) + sl; H = sh + v; Lo = v - (H - sh); }\n"
    // the value of lane CTRL-selected within each quad (DPP quad permute), for the lane-group form
    "template <int CTRL> static __device__ inline __attribute__((always_inline)) float ce_bc(float v) {\n"
    "  return __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, v), CTRL, 0xF, 0xF, false)); }\n";

// Both forms raise the simulation waves' issue priority (s_setprio 3): the simulation is one serial
// recurrence per candidate with a single wave on each SIMD it uses, and the kernels that run beside
// it on a second stream take the issue slots it leaves.
//
// The gated form (a single launch with no more workgroups than the device has SIMDs) also
//  * claims every register of its SIMD (512 = 256 VGPRs + 256 AGPRs), so no wave of another kernel
//    shares a SIMD with a simulation wave and slows its serial chain;
//  * stores gate_gen to *gate when its last workgroup starts: by then the workgroups before it have
//    been placed, and a bounded waiter on the second stream lets the kernels there start, on the
//    SIMDs the simulation leaves free.
// Body of a structural kernel: constants loaded from the descriptor.
//  * Float form: up to kVgprHoistWords constants in vector registers, the rest read every step.
//    On gfx950 a wave64 FP32 instruction issues in about two cycles, but two that both read a
//    scalar register take about four, and almost every multiply-add of the step reads a plant
//    constant: with the constants in scalar registers the step issues about 1.4 times slower (a
//    half of the instructions reading scalar registers costs nothing).  The budget keeps the
//    kernel near 256 registers, two waves per SIMD, which hold the SIMD's issue slots busy.
//  * FP64 form (four cycles per instruction whatever its operands): hoisted into scalar registers
//    when they fit kHoistMaxWords 32-bit words, else read every step.
static void gen_struct_body(std::ostringstream& o, const ProblemDev& P, bool mixed) {
  constexpr int kHoistMaxWords = 64;
  constexpr int kVgprHoistWords = 128;
  if (mixed) {
    Gen gv;
    gv.f32 = true;
    gen_body(gv, P, true, true, false, nullptr, kVgprHoistWords);
    o << gv.o.str();
    return;
  }
  Gen gh;
  gh.f32 = mixed;
  int words = 0;
  gen_body(gh, P, mixed, true, true, &words);
  if (words <= kHoistMaxWords) {
    o << gh.o.str();
  } else {
    Gen gr;
    gr.f32 = mixed;
    gen_body(gr, P, mixed, true, false);
    o << gr.o.str();
  }
}

// Single-problem kernel entry: the literal form (every constant in the source) or the structural
// form (constants read from the problem descriptor passed as the last argument).
static std::string gen_single(const ProblemDev& P, bool mixed, bool gated, bool structural) {
  // lane groups: the float literal kernel also carries 2- and 4-lane forms of the step
  const bool grp = mixed && !structural;
  // the ungated literal kernels take the part of a split simulation (0: every slot)
  const bool split = !gated && !structural;
  Gen g;
  g.f32 = mixed;
  std::ostringstream& o = g.o;
  const std::string name = std::string(structural ? "ce_sims_" : "ce_sim_") + (mixed ? "f32" : "f64") + (gated ? "_g" : "");
  o << "// kernel: " << name << "\n";
  o << "// generated: closed-loop simulation" << (structural ? ", constants from the problem descriptor" : "")
    << ", ns=" << P.ns << " nd=" << P.nd << " nw=" << P.nw << (mixed ? " float" : " double") << "\n";
  o << kSimArgs;
  o << "static __device__ inline __attribute__((always_inline)) void sim_one(const SimArgs& A, const char* __restrict__ Pp, int bx) {\n";
  if (structural) gen_struct_body(o, P, mixed);
  else gen_body(g, P, mixed, false);
  o << "}\n";
  if (grp) {
    for (int G : {2, 4}) {
      Gen gg;
      gg.f32 = true;
      gen_body_grp(gg, P, G, false);
      o << "static __device__ inline __attribute__((always_inline)) void sim_g" << G
        << "(const SimArgs& A, int bx) {\n" << gg.o.str() << "}\n";
    }
  }
  // one kernel per lane count (each with its own register allocation): NAME (one lane), NAME_l2 and
  // NAME_l4.  The host launches the forms the grid allows (gq); each computes the lanes the live
  // count selects (the most, up to gq, whose waves stay within cap, one per SIMD) and returns
  // unless they are its own.
  for (int G : {1, 2, 4}) {
    if (G > 1 && !grp) break;
    const std::string kn = name + (G > 1 ? "_l" + std::to_string(G) : "");
    o << "extern \"C\" __global__ __launch_bounds__(64) void " << kn
      << "(const float* __restrict__ gains, const int* __restrict__ flags,\n"
         "    const int* __restrict__ list, const int* __restrict__ count, void* __restrict__ ybuf, double* __restrict__ info,\n"
         "    float* __restrict__ dy, int n" << (gated ? ", int* __restrict__ gate, int gate_gen" : "")
      << (structural ? ", const char* __restrict__ probp" : ", int gq, int cap")
      << (gated ? "" : ", int* __restrict__ done, int done_gen") << (split ? ", int part" : "") << ") {\n"
         "  const SimArgs A{gains, flags, list, count, ybuf, info, dy, n, 0};\n";
    // part 2 (split): the remainder, from the split slot on, in this kernel's lanes; part 1: the
    // one-lane kernel's full rounds only
    if (split)
      o << "  int bxo = blockIdx.x;\n"
           "  if (part == 2) {\n"
           "    const int live2 = __builtin_amdgcn_readfirstlane(*count);\n"
           "    const int sp = ce_split(live2, cap);\n"
           "    if (sp >= live2) return;\n"  // no remainder: every slot is the one-lane kernel's
           "    bxo += sp / " << 64 / G << ";\n"
           "  } else {\n";
    if (grp)
      o << "  const int live = __builtin_amdgcn_readfirstlane(*count);\n"
           "  const int pick = (gq >= 4 && live <= 16 * cap) ? 4 : ((gq >= 2 && live <= 32 * cap) ? 2 : 1);\n"
           "  if (pick != " << G << ") return;\n";
    if (split)
      o << "  if (part == 1 && (int)blockIdx.x * " << 64 / G << " >= ce_split(__builtin_amdgcn_readfirstlane(*count), cap))"
           " return;\n"
           "  }\n";
    if (gated)
      o << "  asm volatile(\"\" ::: \"v255\", \"a255\");\n"
           "  if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0)\n"
           "    __hip_atomic_store(gate, gate_gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);\n";
    o << "  __builtin_amdgcn_s_setprio(3);\n";
    if (!gated) o << "  const int cnt1 = __builtin_amdgcn_readfirstlane(*count);\n";
    const char* bxs = split ? "bxo" : "blockIdx.x";
    if (G == 1) o << "  sim_one(A, " << (structural ? "probp" : "nullptr") << ", " << bxs << ");\n";
    else o << "  sim_g" << G << "(A, " << bxs << ");\n";
    // ungated: the first working wave to finish (or block 0 when none is live) signals done_gen
    if (!gated)
      o << "  if (done != nullptr && threadIdx.x == 0 && ((int)blockIdx.x * " << 64 / G
        << " < cnt1 || (blockIdx.x == 0 && cnt1 == 0)))\n"
           "    __hip_atomic_store(done, done_gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);\n";
    o << "}\n";
  }
  return o.str();
}

std::string gen_sim_source(const ProblemDev& P, bool mixed, bool gated) { return gen_single(P, mixed, gated, false); }

std::string gen_sim_struct_source(const ProblemDev& P, bool mixed, bool gated) { return gen_single(P, mixed, gated, true); }

// Structural bodies of one structure: sim<tag>(A, Pp, bx) with one lane per candidate and, for
// the float form, sim<tag>_g2 / _g4 with 2 and 4 lanes.
static void gen_struct_bodies(std::ostringstream& o, const ProblemDev& P, bool mixed, const std::string& tag) {
  o << "static __device__ inline __attribute__((always_inline)) void sim" << tag
    << "(const SimArgs& A, const char* __restrict__ Pp, int bx) {\n";
  gen_struct_body(o, P, mixed);
  o << "}\n";
  if (!mixed) return;
  for (int G : {2, 4}) {
    Gen g;
    g.f32 = true;
    gen_body_grp(g, P, G, true);
    o << "static __device__ inline __attribute__((always_inline)) void sim" << tag << "_g" << G
      << "(const SimArgs& A, const char* __restrict__ Pp, int bx) {\n" << g.o.str() << "}\n";
  }
}

// call of structure tag's body with G lanes per candidate
static std::string struct_call(const std::string& tag, int G, const std::string& args) {
  return "sim" + tag + (G > 1 ? "_g" + std::to_string(G) : std::string()) + "(" + args + ");";
}

std::string gen_sim_batch_source(const ProblemDev& P, bool mixed) {
  Gen g;
  g.f32 = mixed;
  std::ostringstream& o = g.o;
  const char* name = mixed ? "ce_simb_f32" : "ce_simb_f64";
  o << "// kernel: " << name << "\n";
  o << "// generated: closed-loop simulation for every problem of one structure, ns=" << P.ns << " nd=" << P.nd
    << " nw=" << P.nw << (mixed ? " float" : " double") << "\n";
  o << kSimArgs;
  gen_struct_bodies(o, P, mixed, "_one");
  // one kernel per lane count (NAME, NAME_l2, NAME_l4: each with its own register allocation); the
  // launch's block table has 64 / lanes candidates per block
  for (int G : {1, 2, 4}) {
    if (G > 1 && !mixed) break;
    o << "extern \"C\" __global__ __launch_bounds__(64) void " << name << (G > 1 ? "_l" + std::to_string(G) : std::string())
      << "(const int* __restrict__ start, const int* __restrict__ pid, const SimArgs* __restrict__ args,\n"
         "    const char* const* __restrict__ probs, int J, int total) {\n"
         "  __builtin_amdgcn_s_setprio(3);\n"
         "  for (int b = blockIdx.x; b < total; b += gridDim.x) {\n"
         "    const int bu = __builtin_amdgcn_readfirstlane(b);\n"
         "    int lo = 0, hi = J - 1;\n"
         "    while (lo < hi) { const int mid = (lo + hi + 1) >> 1; if (start[mid] <= bu) lo = mid; else hi = mid - 1; }\n"
         "    const int p = __builtin_amdgcn_readfirstlane(pid[lo]);\n"
         "    const int lb = bu - __builtin_amdgcn_readfirstlane(start[lo]);\n"
         "    " << struct_call("_one", G, "args[p], probs[p], lb") << "\n"
         "  }\n}\n";
  }
  return o.str();
}

// Every structure of a batch in one kernel: the structural body of each structure, and a block
// loop that finds its problem and runs the body of the problem's structure (sid: structure index
// of each problem of the launch, in the launch's problem order).
std::string gen_sim_batch_fused_source(const std::vector<const ProblemDev*>& reps, bool mixed) {
  std::ostringstream o;
  const char* name = mixed ? "ce_simb_f32_all" : "ce_simb_f64_all";
  o << "// kernel: " << name << "\n";
  o << "// generated: closed-loop simulation for every problem of a batch, " << reps.size() << " structures"
    << (mixed ? ", float" : ", double") << "\n";
  o << kSimArgs;
  for (size_t k = 0; k < reps.size(); k++) {
    o << "// structure " << k << ": ns=" << reps[k]->ns << " nd=" << reps[k]->nd << " nw=" << reps[k]->nw << "\n";
    gen_struct_bodies(o, *reps[k], mixed, "_s" + std::to_string(k));
  }
  for (int G : {1, 2, 4}) {
    if (G > 1 && !mixed) break;
    o << "extern \"C\" __global__ __launch_bounds__(64) void " << name << (G > 1 ? "_l" + std::to_string(G) : std::string())
      << "(const int* __restrict__ start, const int* __restrict__ pid, const SimArgs* __restrict__ args,\n"
         "    const char* const* __restrict__ probs, const int* __restrict__ sid, int J, int total) {\n"
         "  __builtin_amdgcn_s_setprio(3);\n"
         "  for (int b = blockIdx.x; b < total; b += gridDim.x) {\n"
         "    const int bu = __builtin_amdgcn_readfirstlane(b);\n"
         "    int lo = 0, hi = J - 1;\n"
         "    while (lo < hi) { const int mid = (lo + hi + 1) >> 1; if (start[mid] <= bu) lo = mid; else hi = mid - 1; }\n"
         "    const int p = __builtin_amdgcn_readfirstlane(pid[lo]);\n"
         "    const int lb = bu - __builtin_amdgcn_readfirstlane(start[lo]);\n"
         "    switch (__builtin_amdgcn_readfirstlane(sid[lo])) {\n";
    for (size_t k = 0; k < reps.size(); k++)
      o << "      case " << k << ": " << struct_call("_s" + std::to_string(k), G, "args[p], probs[p], lb") << " break;\n";
    o << "    }\n  }\n}\n";
  }
  return o.str();
}

// ------------------------------------------------------------------------------------------------
struct JitEntry {
  hipModule_t mod = nullptr;
  hipFunction_t fn = nullptr;
  hipFunction_t fl[2] = {nullptr, nullptr};  // NAME_l2, NAME_l4 when the module has them
};

// Compiled modules by source text.  Entries added as evictable (the literal kernels, one per
// distinct set of plant constants) are kept to the JIT_EVICTABLE_MAX most recently used; the rest
// (structural kernels, one per instruction structure) stay until the cache is destroyed.
constexpr size_t JIT_EVICTABLE_MAX = 32;
struct JitCache {
  std::mutex mu;
  std::map<std::string, JitEntry> entries;
  std::list<std::string> lru;  // evictable entries, most recently used first
  long long compiles = 0, evictions = 0;
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

bool jit_cached(JitCache* cache, const std::string& src) {
  std::lock_guard<std::mutex> lk(cache->mu);
  return cache->entries.count(src) != 0;
}

long long jit_compiles(JitCache* cache) {
  if (!cache) return 0;
  std::lock_guard<std::mutex> lk(cache->mu);
  return cache->compiles;
}

long long jit_evictions(JitCache* cache) {
  if (!cache) return 0;
  std::lock_guard<std::mutex> lk(cache->mu);
  return cache->evictions;
}

bool jit_get(JitCache* cache, const std::string& src, void** fn_out, std::string& err, bool evictable, void** fn_lanes) {
  std::lock_guard<std::mutex> lk(cache->mu);
  auto it = cache->entries.find(src);
  if (it != cache->entries.end()) {
    auto li = std::find(cache->lru.begin(), cache->lru.end(), src);
    if (li != cache->lru.end()) cache->lru.splice(cache->lru.begin(), cache->lru, li);
    *fn_out = (void*)it->second.fn;
    if (fn_lanes) {
      fn_lanes[0] = (void*)it->second.fl[0];
      fn_lanes[1] = (void*)it->second.fl[1];
    }
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
  // the first line names the kernel: "// kernel: NAME"
  const size_t nl = src.find('\n');
  const std::string kname = src.substr(11, nl - 11);
  if (hipModuleGetFunction(&e.fn, e.mod, kname.c_str()) != hipSuccess) {
    err = std::string("hipModuleGetFunction(") + kname + ") failed";
    hipModuleUnload(e.mod);
    return false;
  }
  // lane-group kernels of the same module, where the source has them
  for (int k = 0; k < 2; k++) {
    const std::string ln = kname + (k ? "_l4" : "_l2");
    if (src.find("void " + ln + "(") != std::string::npos && hipModuleGetFunction(&e.fl[k], e.mod, ln.c_str()) != hipSuccess) {
      err = std::string("hipModuleGetFunction(") + ln + ") failed";
      hipModuleUnload(e.mod);
      return false;
    }
  }
  if (fn_lanes) {
    fn_lanes[0] = (void*)e.fl[0];
    fn_lanes[1] = (void*)e.fl[1];
  }
  cache->entries[src] = e;
  cache->compiles++;
  if (evictable) {
    cache->lru.push_front(src);
    if (cache->lru.size() > JIT_EVICTABLE_MAX) {
      // the module may still have launches in flight: drain the device before unloading it
      hipDeviceSynchronize();
      const std::string old = cache->lru.back();
      cache->lru.pop_back();
      auto oi = cache->entries.find(old);
      if (oi != cache->entries.end()) {
        hipModuleUnload(oi->second.mod);
        cache->entries.erase(oi);
      }
      cache->evictions++;
    }
  }
  *fn_out = (void*)e.fn;
  return true;
}

bool jit_launch_sim(void* fn, int grid, hipStream_t st, const float* gains, const int* flags, const int* list,
                    const int* count, void* ybuf, double* info, float* dy, int n, int gq, int cap, int* done,
                    int done_gen, int part) {
  void* args[] = {&gains, &flags, &list, &count, &ybuf, &info, &dy, &n, &gq, &cap, &done, &done_gen, &part};
  return hipModuleLaunchKernel((hipFunction_t)fn, grid, 1, 1, 64, 1, 1, 0, st, args, nullptr) == hipSuccess;
}

bool jit_launch_sim_gated(void* fn, int grid, hipStream_t st, const float* gains, const int* flags, const int* list,
                          const int* count, void* ybuf, double* info, float* dy, int n, int* gate, int gen, int gq,
                          int cap) {
  void* args[] = {&gains, &flags, &list, &count, &ybuf, &info, &dy, &n, &gate, &gen, &gq, &cap};
  return hipModuleLaunchKernel((hipFunction_t)fn, grid, 1, 1, 64, 1, 1, 0, st, args, nullptr) == hipSuccess;
}

bool jit_launch_sim_struct(void* fn, int grid, hipStream_t st, const float* gains, const int* flags, const int* list,
                           const int* count, void* ybuf, double* info, float* dy, int n, const void* prob, int* done,
                           int done_gen) {
  void* args[] = {&gains, &flags, &list, &count, &ybuf, &info, &dy, &n, &prob, &done, &done_gen};
  return hipModuleLaunchKernel((hipFunction_t)fn, grid, 1, 1, 64, 1, 1, 0, st, args, nullptr) == hipSuccess;
}

bool jit_launch_sim_struct_gated(void* fn, int grid, hipStream_t st, const float* gains, const int* flags, const int* list,
                                 const int* count, void* ybuf, double* info, float* dy, int n, int* gate, int gen,
                                 const void* prob) {
  void* args[] = {&gains, &flags, &list, &count, &ybuf, &info, &dy, &n, &gate, &gen, &prob};
  return hipModuleLaunchKernel((hipFunction_t)fn, grid, 1, 1, 64, 1, 1, 0, st, args, nullptr) == hipSuccess;
}

bool jit_launch_sim_batch(void* fn, int grid, hipStream_t st, const int* start, const int* pid, const SimArgs* args,
                          const char* const* probs, int J, int total) {
  void* a[] = {&start, &pid, &args, &probs, &J, &total};
  return hipModuleLaunchKernel((hipFunction_t)fn, grid, 1, 1, 64, 1, 1, 0, st, a, nullptr) == hipSuccess;
}

bool jit_launch_sim_batch_fused(void* fn, int grid, hipStream_t st, const int* start, const int* pid, const SimArgs* args,
                                const char* const* probs, const int* sid, int J, int total) {
  void* a[] = {&start, &pid, &args, &probs, &sid, &J, &total};
  return hipModuleLaunchKernel((hipFunction_t)fn, grid, 1, 1, 64, 1, 1, 0, st, a, nullptr) == hipSuccess;
}

std::string sim_structure_key(const ProblemDev& P, bool mixed) {
  std::ostringstream k;
  k << (mixed ? "f" : "d") << P.ns << "." << P.nd << "." << P.nw << "|";
  auto chain = [&](const Ins* ops, int n) {
    for (int r = 0; r < n; r++) k << ops[r].op << "," << ops[r].col << "," << ops[r].akind << "," << ops[r].bkind << ";";
    k << "|";
  };
  chain(P.dc_ops, P.dc_n);
  chain(P.u_ops, P.u_n);
  for (int q = 0; q < P.nw; q++) chain(P.nl_ops[q], P.nl_n[q]);
  if (mixed) {  // the float kernels leave out zero coefficients
    auto zp = [&](const double* v, int n) {
      for (int i = 0; i < n; i++) k << ((float)v[i] != 0.0f ? '1' : '0');
      k << "|";
    };
    zp(P.sc, P.ns);
    zp(&P.sd, 1);
    zp(P.sb, P.ns);
    zp(P.sai, P.ns * P.ns);
    zp(P.dbu, P.nd);
    zp(P.dbw, P.nd * P.nw);
    zp(P.dai, P.nd * P.nd);
  }
  return k.str();
}

}  // namespace ce

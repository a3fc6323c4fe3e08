// Host side of the engine: ControllerEngine.config parsing, derived sizes, .bin loading, instruction
// decoding and the per-problem tables (reference vector, frequency grid, plant response, alpha)
// that the gfx950 kernels read.
#include "ce_host.h"

#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

namespace ce {

std::string fmt_num(double v) {
  // Shortest representation that reads back to the same double (the reference formats values
  // with std::format("{}"), which has this property).
  char buf[64];
  if (v == std::floor(v) && std::fabs(v) < 1e16) {
    std::snprintf(buf, sizeof buf, "%.0f", v);
    return buf;
  }
  for (int p = 1; p <= 17; p++) {
    std::snprintf(buf, sizeof buf, "%.*g", p, v);
    if (std::strtod(buf, nullptr) == v) break;
  }
  return buf;
}

// Shortest representation that reads back to the same float (std::format("{}") of a float).
std::string fmt_flt(float v) {
  char buf[64];
  if (v == std::floor(v) && std::fabs(v) < 1e16f) {
    std::snprintf(buf, sizeof buf, "%.0f", (double)v);
    return buf;
  }
  for (int p = 1; p <= 9; p++) {
    std::snprintf(buf, sizeof buf, "%.*g", p, (double)v);
    if (std::strtof(buf, nullptr) == v) break;
  }
  return buf;
}

static std::string trim(const std::string& s) {
  const size_t a = s.find_first_not_of(" \t\r\n"), b = s.find_last_not_of(" \t\r\n");
  return a == std::string::npos ? std::string() : s.substr(a, b - a + 1);
}

// load_config: reference src/config.cu:106-244 (same keys, defaults, checks and message texts).
bool parse_config(const std::string& text, const std::string& filename, ce_config& c, std::string& err) {
  std::map<std::string, std::string> kv;
  std::istringstream file(text);
  std::string line;
  int lineno = 0;
  while (std::getline(file, line)) {
    lineno++;
    line = trim(line.substr(0, line.find('#')));
    if (line.empty()) continue;
    const size_t eq = line.find('=');
    if (eq == std::string::npos) {
      err = filename + ":" + std::to_string(lineno) + ": expected `key = value`, got `" + line + "`";
      return false;
    }
    const std::string key = trim(line.substr(0, eq));
    if (kv.count(key)) {
      err = filename + ":" + std::to_string(lineno) + ": `" + key + "` is set twice";
      return false;
    }
    kv[key] = trim(line.substr(eq + 1));
  }

  struct Fail {};
  auto fail = [&](const std::string& m) { err = m; throw Fail{}; };
  auto take = [&](const std::string& key, const char* fallback = nullptr) {
    auto it = kv.find(key);
    if (it == kv.end()) {
      if (fallback) return std::string(fallback);
      fail(filename + ": missing `" + key + "`");
    }
    std::string v = it->second;
    kv.erase(it);
    return v;
  };
  auto as_double = [&](const std::string& key, const std::string& v) {
    char* end = nullptr;
    const double d = std::strtod(v.c_str(), &end);
    if (v.empty() || *end != '\0') fail(filename + ": `" + key + "` should be a number, got `" + v + "`");
    return d;
  };
  auto as_int = [&](const std::string& key, const std::string& v) {
    const double d = as_double(key, v);
    if (d != std::floor(d)) fail(filename + ": `" + key + "` should be an integer, got `" + v + "`");
    return (int)d;
  };
  auto get_int = [&](const std::string& key, const char* fb = nullptr) { return as_int(key, take(key, fb)); };
  auto get_float = [&](const std::string& key, const char* fb = nullptr) { return (float)as_double(key, take(key, fb)); };
  auto in_range = [&](const char* key, double v, double lo, double hi) {
    if (v < lo || v > hi)
      fail(filename + ": `" + key + "` = " + fmt_num(v) + " is outside [" + fmt_num(lo) + ", " + fmt_num(hi) + "]");
  };

  try {
    c = ce_config{};
    c.statespace_dim = get_int("statespace_dim");
    c.discrete_dim = get_int("discrete_dim");
    c.nonlin_dim = get_int("nonlin_dim");
    c.n_params = get_int("n_params");
    c.input_size = get_int("input_size");
    c.dc_instructions = get_int("dc_instructions");
    c.u_instructions = get_int("u_instructions");
    in_range("statespace_dim", c.statespace_dim, 1, CE_MAX_STATESPACE_DIM);
    in_range("discrete_dim", c.discrete_dim, 1, CE_MAX_DISCRETE_DIM);
    in_range("nonlin_dim", c.nonlin_dim, 0, CE_MAX_NONLIN_DIM);
    in_range("n_params", c.n_params, 1, CE_MAX_PARAMS);
    in_range("input_size", c.input_size, 0, 1 << 20);
    in_range("dc_instructions", c.dc_instructions, 0, CE_MAX_DC_INSTRUCTIONS);
    in_range("u_instructions", c.u_instructions, 0, CE_MAX_U_INSTRUCTIONS);
    {
      std::stringstream list(take("nl_instructions", ""));
      std::string item;
      int n = 0;
      while (std::getline(list, item, ',')) {
        item = trim(item);
        if (item.empty()) continue;
        if (n >= CE_MAX_NONLIN_DIM)
          fail(filename + ": `nl_instructions` has more than MAX_NONLIN_DIM = " + std::to_string(CE_MAX_NONLIN_DIM) + " entries");
        c.nl_instructions[n] = as_int("nl_instructions", item);
        in_range("nl_instructions", c.nl_instructions[n], 0, CE_MAX_NL_INSTRUCTIONS);
        n++;
      }
      if (n != c.nonlin_dim)
        fail(filename + ": `nl_instructions` has " + std::to_string(n) + " entries but nonlin_dim is " + std::to_string(c.nonlin_dim));
    }
    c.n_cores = get_int("n_cores");
    c.n_rounds = get_int("n_rounds", "1");
    in_range("n_cores", c.n_cores, 128, 1 << 20);
    if (c.n_cores % 128 != 0)
      fail(filename + ": `n_cores` = " + std::to_string(c.n_cores) + " should be a multiple of 128 (cores per SM)");
    in_range("n_rounds", c.n_rounds, 1, 1 << 20);
    {
      const std::string v = take("mode", "batch");
      if (v == "batch") c.serial = 0;
      else if (v == "serial") c.serial = 1;
      else fail(filename + ": `mode` should be batch/serial, got `" + v + "`");
      if (c.serial && c.statespace_dim != c.discrete_dim)
        fail(filename + ": mode = serial needs statespace_dim (" + std::to_string(c.statespace_dim) + ") == discrete_dim (" +
             std::to_string(c.discrete_dim) + ")");
    }
    c.time = get_int("time");
    c.ref_step_time = as_double("ref_step_time", take("ref_step_time"));
    c.ref_amplitude = get_float("ref_amplitude");
    if (!(c.ref_amplitude != 0.0f)) fail(filename + ": need ref_amplitude != 0");
    in_range("time", c.time, 2, 1 << 24);
    c.duration = get_float("duration", "1.0");
    if (!(c.duration > 0.0f)) fail(filename + ": need duration > 0");
    in_range("ref_step_time", c.ref_step_time, 0.0, 1.0);
    c.fit_start = get_float("fit_start");
    c.fit_end = get_float("fit_end");
    c.steady_start = get_float("steady_start");
    in_range("fit_start", c.fit_start, 0.0, 1.0);
    in_range("fit_end", c.fit_end, 0.0, 1.0);
    in_range("steady_start", c.steady_start, 0.0, 1.0);
    if (!(c.fit_start < c.fit_end))
      fail(filename + ": fit_start (" + fmt_flt(c.fit_start) + ") must be below fit_end (" + fmt_flt(c.fit_end) + ")");
    if ((int)(c.steady_start * c.time) >= c.time) fail(filename + ": steady_start leaves no samples to average");
    c.bandwidth = get_float("bandwidth");
    c.damping = get_float("damping");
    c.overshoot = get_float("overshoot");
    c.steady_state_error = get_float("steady_state_error");
    c.crossover_coeff = get_float("crossover_coeff");
    c.target_gm = get_float("target_gm");
    c.target_pm = get_float("target_pm");
    c.target_b = get_float("target_b");
    if (!(c.overshoot > 0.0f)) fail(filename + ": need overshoot > 0");
    if (!(c.steady_state_error > 0.0f)) fail(filename + ": need steady_state_error > 0");
    if (!(c.crossover_coeff > 0.0f)) fail(filename + ": need crossover_coeff > 0");
    if (!(c.target_b > 0.0f && c.target_b < 1.0f)) fail(filename + ": need 0 < target_b < 1");
    c.eig_fail_err = get_float("eig_fail_err", "1.0e18");
    c.unstable_err = get_float("unstable_err", "1.0e6");
    c.freq = get_int("freq");
    c.freq_min = get_float("freq_min");
    in_range("freq", c.freq, 2, 1 << 20);
    if (!(c.freq_min > 0.0f && c.freq_min < PI_F * c.time / c.duration))
      fail(filename + ": need 0 < freq_min < pi*time/duration (Nyquist)");
    {
      const std::string v = take("dump_diagnostics", "false");
      if (v == "true" || v == "1") c.dump_diagnostics = 1;
      else if (v == "false" || v == "0") c.dump_diagnostics = 0;
      else fail(filename + ": `dump_diagnostics` should be true/false, got `" + v + "`");
    }
    if (!kv.empty()) fail(filename + ": unknown key `" + kv.begin()->first + "`");
    ce_derived d;
    if (!derive_config(c, d, err)) {
      err = filename + ": " + err;
      return false;
    }
    if (!(d.crossover > 0.0f && d.crossover < d.freq_max))
      fail(filename + ": crossover " + fmt_flt(d.crossover) + " rad/s (from bandwidth, damping, crossover_coeff) must be in (0, Nyquist = " +
           fmt_flt(d.freq_max) + ")");
  } catch (const Fail&) {
    return false;
  }
  return true;
}

// derive(): reference src/config.cu:40-78 and make_engine_params src/config.cu:93-104, float32.
bool derive_config(const ce_config& c, ce_derived& d, std::string& err) {
  d = ce_derived{};
  d.n_threads = (c.n_cores / 128) * 1536;
  d.ts = c.duration / (float)c.time;
  d.freq_max = PI_F / d.ts;
  d.total_dim = c.statespace_dim + c.discrete_dim;
  d.cuda_mx_width = 2 * c.statespace_dim + 1;
  d.info_scoring = c.input_size;
  d.info_freqresp = d.info_scoring + 3;
  d.info_allmargin = d.info_freqresp + 4;
  d.info_modal = d.info_allmargin + 5;
  d.info_mx_width = d.info_modal + 3 * d.total_dim + 1;
  d.fit_t0 = c.fit_start * c.duration;
  d.fit_t1 = (c.fit_end * (float)c.time - 1.0f) * d.ts;
  const float d2 = c.damping * c.damping, d4 = d2 * d2;
  const float wn = 2.0f * PI_F * c.bandwidth / std::sqrt((1.0f - 2.0f * d2) + std::sqrt(4.0f * d4 - 4.0f * d2 + 2.0f));
  d.crossover = c.crossover_coeff * wn * std::sqrt(std::sqrt(4.0f * d4 + 1.0f) - 2.0f * d2);
  d.fit_start_id = (int)(c.fit_start * (float)c.time);
  d.steady_start_id = (int)(c.steady_start * (float)c.time);
  // The fit loop (reference src/modalsetup.cu:75-76) visits every int t with (float)t < fit_len.
  const float fit_len = (c.fit_end - c.fit_start) * (float)c.time;
  int cnt = 0;
  if (fit_len > 0.0f) {
    cnt = (int)std::ceil(fit_len);
    while (cnt > 0 && !((float)(cnt - 1) < fit_len)) cnt--;
    while ((float)cnt < fit_len) cnt++;
  }
  d.fit_count = cnt;
  if (d.fit_start_id + d.fit_count > c.time) {
    err = "fit window [" + std::to_string(d.fit_start_id) + ", " + std::to_string(d.fit_start_id + d.fit_count) +
          ") runs past time = " + std::to_string(c.time);
    return false;
  }
  return true;
}

// readbin: reference src/main.cu:293-302 (exact size, slice `slice` of `slices` stacked copies).
bool read_bin(const std::string& path, const std::string& name, size_t nbytes, void* dst, int slice, int slices,
              std::string& err) {
  std::ifstream file(path, std::ios::binary);
  if (!file) { err = "Cannot open " + name; return false; }
  file.seekg(0, std::ios::end);
  const size_t expected = nbytes * (size_t)slices;
  const std::streamoff actual = file.tellg();
  if (actual != (std::streamoff)expected) {
    err = name + " is " + std::to_string((long long)actual) + " bytes, expected " + std::to_string(expected) + " (" +
          std::to_string(slices) + " x " + std::to_string(nbytes) + ")";
    return false;
  }
  if (nbytes == 0) return true;
  file.seekg((std::streamoff)(nbytes * (size_t)slice), std::ios::beg);
  file.read(reinterpret_cast<char*>(dst), (std::streamsize)nbytes);
  if (!file) { err = "Cannot read " + name + " (intended " + std::to_string(nbytes) + "-byte allocation)"; return false; }
  return true;
}

// upload_plant / init64KB file set: reference src/main.cu:341-400.
bool load_plant_dir(const std::string& dir, const ce_config& c, int round, ce_plant& p, std::string& err) {
  p = ce_plant{};
  const int ss = c.statespace_dim, ds = c.discrete_dim;
  const int slices = c.serial ? c.n_rounds : 1;
  const int slice = c.serial ? round : 0;
  auto path = [&](const char* name) { return dir.empty() ? std::string(name) : dir + "/" + name; };
  auto rd = [&](const char* name, size_t nfl, void* dst, int sl, int sls) {
    return read_bin(path(name), name, nfl * 4, dst, sl, sls, err);
  };
  if (!rd("STATESPACE_A.bin", (size_t)ss * ss, p.ss_a, 0, 1)) return false;
  if (!rd("STATESPACE_A_I.bin", (size_t)ss * ss, p.ss_a_i, 0, 1)) return false;
  if (!rd("STATESPACE_B.bin", (size_t)ss, p.ss_b, 0, 1)) return false;
  if (!rd("STATESPACE_C.bin", (size_t)ss, p.ss_c, 0, 1)) return false;
  if (!rd("STATESPACE_D.bin", 1, &p.ss_d, 0, 1)) return false;
  if (!rd("DISCRETE_A.bin", (size_t)ds * ds, p.d_a, slice, slices)) return false;
  if (!rd("DISCRETE_BU.bin", (size_t)ds, p.d_bu, slice, slices)) return false;
  if (!rd("DISCRETE_BW.bin", (size_t)ds * c.nonlin_dim, p.d_bw, slice, slices)) return false;
  if (!rd("DISCRETE_C.bin", (size_t)ds, p.d_c, slice, slices)) return false;
  if (!rd("DISCRETE_A_I.bin", (size_t)ds * ds, p.d_a_i, slice, slices)) return false;
  if (!rd("BASE_PARAMS.bin", (size_t)c.n_params, p.base_params, slice, slices)) return false;
  int nl_total = 0;
  for (int i = 0; i < c.nonlin_dim; i++) nl_total += c.nl_instructions[i];
  // instruction files are not stacked per round (reference src/main.cu:360-362)
  if (!rd("DC_INFO.bin", (size_t)ds * c.dc_instructions, p.dc_info, 0, 1)) return false;
  if (!rd("U_INFO.bin", (size_t)ss * c.u_instructions, p.u_info, 0, 1)) return false;
  if (!rd("NL_INFO.bin", (size_t)(ds + 1) * nl_total, p.nl_info, 0, 1)) return false;
  return true;
}

// decode_instructions: reference src/main.cu:317-337.
bool decode(const int32_t* info, int dim, int count, const float* base, int n_params, Ins* out, std::string& err) {
  for (int r = 0; r < count; r++) {
    const int32_t* row = info + (size_t)r * dim;
    int packed = 0, col = 0;
    for (int k = 0; k < dim; k++)
      if (row[k] != 0) { packed = row[k]; col = k; }
    out[r] = Ins{0, 0, K_CONST, K_CONST, 0.0, 0.0};
    if (packed == 0) continue;
    if (packed < 0) { err = "instruction entry " + std::to_string(packed) + " is negative"; return false; }
    const int dg[2] = {(packed / 36) % 36, packed % 36};
    double v[2] = {0.0, 0.0};
    int kd[2];
    for (int q = 0; q < 2; q++) {
      const int d = dg[q];
      kd[q] = d == 0 ? K_STATE : (d == 35 ? K_PREV : K_CONST);
      if (d == 0 || d == 35) continue;
      if (d - 1 >= n_params) {
        err = "instruction digit " + std::to_string(d) + " needs BASE_PARAMS[" + std::to_string(d - 1) +
              "] but n_params is " + std::to_string(n_params);
        return false;
      }
      v[q] = (double)base[d - 1];
    }
    out[r] = Ins{packed / 1296, col, kd[0], kd[1], v[0], v[1]};
  }
  return true;
}

// (z I - A) x = b by elimination with partial pivoting; same operation order as the device sweep.
bool shifted_solve_host(const double* A, int n, double zr, double zi, const double* b, double* xr, double* xi) {
  double Mr[MAXS][MAXS], Mi[MAXS][MAXS], vr[MAXS], vi[MAXS];
  for (int i = 0; i < n; i++) {
    for (int j = 0; j < n; j++) {
      Mr[i][j] = (i == j ? zr : 0.0) - A[i * n + j];
      Mi[i][j] = (i == j ? zi : 0.0);
    }
    vr[i] = b[i];
    vi[i] = 0.0;
  }
  for (int k = 0; k < n; k++) {
    int p = k;
    double best = Mr[k][k] * Mr[k][k] + Mi[k][k] * Mi[k][k];
    for (int i = k + 1; i < n; i++) {
      const double s = Mr[i][k] * Mr[i][k] + Mi[i][k] * Mi[i][k];
      if (s > best) { best = s; p = i; }
    }
    if (best != 0.0) {
      if (p != k) {
        for (int j = k; j < n; j++) {
          std::swap(Mr[k][j], Mr[p][j]);
          std::swap(Mi[k][j], Mi[p][j]);
        }
        std::swap(vr[k], vr[p]);
        std::swap(vi[k], vi[p]);
      }
      const double pr = Mr[k][k], pim = Mi[k][k];
      const double dd = pr * pr + pim * pim;
      for (int i = k + 1; i < n; i++) {
        const double ar = Mr[i][k], ai = Mi[i][k];
        Mr[i][k] = (ar * pr + ai * pim) / dd;
        Mi[i][k] = (ai * pr - ar * pim) / dd;
      }
    }
    for (int i = k + 1; i < n; i++) {
      const double lr = Mr[i][k], li = Mi[i][k];
      for (int j = k + 1; j < n; j++) {
        const double qr = lr * Mr[k][j] - li * Mi[k][j];
        const double qi = lr * Mi[k][j] + li * Mr[k][j];
        Mr[i][j] = Mr[i][j] - qr;
        Mi[i][j] = Mi[i][j] - qi;
      }
      const double qr = lr * vr[k] - li * vi[k];
      const double qi = lr * vi[k] + li * vr[k];
      vr[i] = vr[i] - qr;
      vi[i] = vi[i] - qi;
    }
  }
  bool fin = true;
  for (int i = n - 1; i >= 0; i--) {
    const double pr = Mr[i][i], pim = Mi[i][i];
    const double dd = pr * pr + pim * pim;
    xr[i] = (vr[i] * pr + vi[i] * pim) / dd;
    xi[i] = (vi[i] * pr - vr[i] * pim) / dd;
    for (int j = 0; j < i; j++) {
      const double qr = Mr[j][i] * xr[i] - Mi[j][i] * xi[i];
      const double qi = Mr[j][i] * xi[i] + Mi[j][i] * xr[i];
      vr[j] = vr[j] - qr;
      vi[j] = vi[j] - qi;
    }
  }
  for (int i = 0; i < n; i++) fin = fin && cedm::isfinite_(xr[i]) && cedm::isfinite_(xi[i]);
  return fin;
}

static void real_dot(const double* w, const double* xr, const double* xi, int n, double& gr, double& gi) {
  gr = 0.0;
  gi = 0.0;
  for (int i = 0; i < n; i++) {
    gr = cedm::fma_(w[i], xr[i], gr);
    gi = cedm::fma_(w[i], xi[i], gi);
  }
}

// Build the device problem blob: ProblemDev followed by the per-frequency tables.
bool build_problem(const ce_config& c, const ce_plant& pl, std::vector<uint8_t>& blob, std::string& err) {
  ce_derived d;
  if (!derive_config(c, d, err)) return false;
  const int F = c.freq;
  const size_t head = (sizeof(ProblemDev) + 255) & ~(size_t)255;
  const size_t tab = ((size_t)F * 8 + 255) & ~(size_t)255;
  const int km = (F + 63) / 64;
  const size_t slots = (size_t)km * 64;
  const size_t tab_sw = (slots * sizeof(SweepPointIl) + 255) & ~(size_t)255;
  const size_t tab_th = (slots * 8 + 255) & ~(size_t)255;
  blob.assign(head + 7 * tab + tab_sw + tab_th, 0);
  ProblemDev& P = *reinterpret_cast<ProblemDev*>(blob.data());
  P.ns = c.statespace_dim; P.nd = c.discrete_dim; P.nt = P.ns + P.nd; P.nw = c.nonlin_dim;
  P.T = c.time; P.F = F;
  P.dc_n = c.dc_instructions; P.u_n = c.u_instructions;
  for (int i = 0; i < MAXW; i++) P.nl_n[i] = i < c.nonlin_dim ? c.nl_instructions[i] : 0;
  P.fit_start_id = d.fit_start_id; P.steady_start_id = d.steady_start_id; P.fit_count = d.fit_count;
  // first sample of the step: reference src/main.cu:414 compares (double)t < ref_step_time*time
  int k0 = 0;
  while (k0 < c.time && (double)k0 < c.ref_step_time * (double)c.time) k0++;
  P.ref_step_id = k0;
  P.info_w = d.info_mx_width; P.info_scoring = d.info_scoring; P.info_freqresp = d.info_freqresp;
  P.info_allmargin = d.info_allmargin; P.info_modal = d.info_modal; P.cuda_w = d.cuda_mx_width;
  P.serial = c.serial;
  P.ts = d.ts; P.fit_t0 = d.fit_t0; P.fit_t1 = d.fit_t1;
  P.unstable_err = c.unstable_err; P.eig_fail_err = c.eig_fail_err; P.ref_amp = c.ref_amplitude;
  P.damping = c.damping; P.bandwidth = c.bandwidth; P.overshoot = c.overshoot; P.sse = c.steady_state_error;
  P.target_gm = c.target_gm; P.target_pm = c.target_pm; P.target_b = c.target_b;
  P.fts = d.ts; P.fref_amp = c.ref_amplitude;
  const int ns = P.ns, nd = P.nd;
  for (int i = 0; i < ns * ns; i++) { P.sa[i] = pl.ss_a[i]; P.sai[i] = pl.ss_a_i[i]; P.fsa[i] = pl.ss_a[i]; P.fsai[i] = pl.ss_a_i[i]; }
  for (int i = 0; i < ns; i++) { P.sb[i] = pl.ss_b[i]; P.sc[i] = pl.ss_c[i]; P.fsb[i] = pl.ss_b[i]; P.fsc[i] = pl.ss_c[i]; }
  P.sd = pl.ss_d; P.fsd = pl.ss_d;
  for (int i = 0; i < nd * nd; i++) { P.da[i] = pl.d_a[i]; P.dai[i] = pl.d_a_i[i]; P.fdai[i] = pl.d_a_i[i]; }
  for (int i = 0; i < nd; i++) { P.dbu[i] = pl.d_bu[i]; P.dc[i] = pl.d_c[i]; P.fdbu[i] = pl.d_bu[i]; }
  for (int i = 0; i < nd * P.nw; i++) { P.dbw[i] = pl.d_bw[i]; P.fdbw[i] = pl.d_bw[i]; }
  if (!decode(pl.dc_info, nd, c.dc_instructions, pl.base_params, c.n_params, P.dc_ops, err)) return false;
  if (!decode(pl.u_info, ns, c.u_instructions, pl.base_params, c.n_params, P.u_ops, err)) return false;
  for (int i = 0, row = 0; i < P.nw; row += c.nl_instructions[i], i++)
    if (!decode(pl.nl_info + (size_t)row * (nd + 1), nd + 1, c.nl_instructions[i], pl.base_params, c.n_params,
                P.nl_ops[i], err))
      return false;
  for (int r = 0; r < CE_MAX_DC_INSTRUCTIONS; r++) { P.fdc_aval[r] = (float)P.dc_ops[r].aval; P.fdc_bval[r] = (float)P.dc_ops[r].bval; }
  for (int r = 0; r < CE_MAX_U_INSTRUCTIONS; r++) { P.fu_aval[r] = (float)P.u_ops[r].aval; P.fu_bval[r] = (float)P.u_ops[r].bval; }
  for (int q = 0; q < MAXW; q++)
    for (int r = 0; r < CE_MAX_NL_INSTRUCTIONS; r++) {
      P.fnl_aval[q][r] = (float)P.nl_ops[q][r].aval;
      P.fnl_bval[q][r] = (float)P.nl_ops[q][r].bval;
    }
  // nominal loop [[A1 - D L K', -L Pc'], [Pb K', Pa]] (src/allmargin.cu:63-71) against LIN_A
  // [[A1, L Dc'], [-Dbu K', Da]] (src/eigensetup.cu:16-56): the sign flip of the plant half maps one
  // onto the other when the margins plant is the simulated plant and D = 0
  {
    auto same = [](const float* a, const float* b, int k) { return std::memcmp(a, b, sizeof(float) * (size_t)k) == 0; };
    bool eq = ns == nd;
    if (eq && !c.serial)
      eq = same(pl.ss_a, pl.d_a, ns * ns) && same(pl.ss_b, pl.d_bu, ns) && same(pl.ss_c, pl.d_c, ns) && pl.ss_d == 0.0f;
    P.nom_equiv = eq ? 1 : 0;
  }
  // margins plant: reference src/allmargin.cu:11-15
  if (c.serial) {
    for (int i = 0; i < ns * ns; i++) P.pa[i] = P.da[i];
    for (int i = 0; i < ns; i++) { P.pb[i] = P.dbu[i]; P.pc[i] = P.dc[i]; }
    P.pd = 0.0;
  } else {
    for (int i = 0; i < ns * ns; i++) P.pa[i] = P.sa[i];
    for (int i = 0; i < ns; i++) { P.pb[i] = P.sb[i]; P.pc[i] = P.sc[i]; }
    P.pd = P.sd;
  }
  P.off_zr = (int64_t)head;
  P.off_zi = P.off_zr + (int64_t)tab;
  P.off_theta = P.off_zi + (int64_t)tab;
  P.off_pr = P.off_theta + (int64_t)tab;
  P.off_pi = P.off_pr + (int64_t)tab;
  P.off_pabs = P.off_pi + (int64_t)tab;
  P.off_pok = P.off_pabs + (int64_t)tab;
  P.off_sweep_il = P.off_pok + (int64_t)tab;
  P.off_theta_il = P.off_sweep_il + (int64_t)tab_sw;
  P.bytes_total = (int64_t)blob.size();
  double* zr = reinterpret_cast<double*>(blob.data() + P.off_zr);
  double* zi = reinterpret_cast<double*>(blob.data() + P.off_zi);
  double* th = reinterpret_cast<double*>(blob.data() + P.off_theta);
  double* pr = reinterpret_cast<double*>(blob.data() + P.off_pr);
  double* pim = reinterpret_cast<double*>(blob.data() + P.off_pi);
  double* pab = reinterpret_cast<double*>(blob.data() + P.off_pabs);
  int32_t* pok = reinterpret_cast<int32_t*>(blob.data() + P.off_pok);
  SweepPointIl* swi = reinterpret_cast<SweepPointIl*>(blob.data() + P.off_sweep_il);
  double* thi = reinterpret_cast<double*>(blob.data() + P.off_theta_il);
  // log-spaced grid from freq_min to Nyquist, derived in float32 as the reference does
  // (src/main.cu:423-427: logf/expf spacing, z = (cosf, sinf) of the float angle) and promoted;
  // theta_w = arg(z) of the float point (src/allmargin.cu:111)
  const float lminf = cedm::logf_(c.freq_min), lmaxf = cedm::logf_(d.freq_max);
  const float incf = (lmaxf - lminf) / (float)(F - 1);
  for (int w = 0; w < F; w++) {
    const float thf = cedm::expf_(lminf + incf * (float)w) * d.ts;
    float sf, cf;
    cedm::sincosf_(thf, sf, cf);
    const double co = cf, s = sf;
    zr[w] = co;
    zi[w] = s;
    th[w] = (double)cedm::atan2f_(sf, cf);
    double xr[MAXS], xi[MAXS];
    const bool ok = shifted_solve_host(P.pa, ns, co, s, P.pb, xr, xi);
    pok[w] = ok ? 1 : 0;
    if (ok) {
      double gr, gi;
      real_dot(P.pc, xr, xi, ns, gr, gi);
      pr[w] = gr + P.pd;
      pim[w] = gi;
    } else {
      pr[w] = 0.0;
      pim[w] = 0.0;
    }
    pab[w] = cedm::hypot(pr[w], pim[w]);
    SweepPointIl& q = swi[il_slot(w, km)];
    // Re(z) - 1: exact in double for a float Re(z) in [0.5, 1] (Sterbenz), and the difference of
    // two doubles otherwise
    q.zm1 = co - 1.0;
    q.zi = s;
    q.pr = (float)pr[w];
    q.pi = (float)pim[w];
    q.pabs = (float)pab[w];
    q.ok = pok[w];
    thi[il_slot(w, km)] = th[w];
  }
  {  // alpha: reference src/allmargin.cu:32-41, at z = polar(1, crossover * ts) in float
    float sf, cf;
    cedm::sincosf_(d.crossover * d.ts, sf, cf);
    const double s = sf, co = cf;
    double xr[MAXS], xi[MAXS];
    const bool ok = shifted_solve_host(P.pa, ns, co, s, P.pb, xr, xi);
    double gr, gi;
    real_dot(P.pc, xr, xi, ns, gr, gi);
    gr = gr + P.pd;
    const double pmag = cedm::hypot(gr, gi);
    if (!ok || !(pmag > 0.0) || !cedm::isfinite_(pmag)) {
      err = "allmargin: plant has no finite nonzero gain at crossover = " + fmt_flt(d.crossover) + " rad/s";
      return false;
    }
    P.alpha = 1.0 / pmag;
  }
  return true;
}

}  // namespace ce

// ---------------------------------------------------------------------------------------------
// C ABI wrappers

static void put_err(char* err, int errlen, const std::string& m) {
  if (err && errlen > 0) std::snprintf(err, (size_t)errlen, "%s", m.c_str());
}

extern "C" CE_EXPORT int ce_config_parse(const char* text, const char* filename, ce_config* out, char* err, int errlen) {
  std::string e;
  if (!ce::parse_config(text ? text : "", filename ? filename : "ControllerEngine.config", *out, e)) {
    put_err(err, errlen, e);
    return 1;
  }
  return 0;
}

extern "C" CE_EXPORT int ce_config_derive(const ce_config* cfg, ce_derived* out, char* err, int errlen) {
  std::string e;
  if (!ce::derive_config(*cfg, *out, e)) {
    put_err(err, errlen, e);
    return 1;
  }
  return 0;
}

extern "C" CE_EXPORT int ce_plant_load_dir(const char* dir, const ce_config* cfg, int round, ce_plant* out, char* err,
                                           int errlen) {
  std::string e;
  if (!ce::load_plant_dir(dir ? dir : "", *cfg, round, *out, e)) {
    put_err(err, errlen, e);
    return 1;
  }
  return 0;
}

extern "C" CE_EXPORT int ce_detmath_host(int fn, const double* x, const double* y, double* out, int n) {
  for (int i = 0; i < n; i++) {
    switch (fn) {
      case 0: out[i] = cedm::exp(x[i]); break;
      case 1: out[i] = cedm::log(x[i]); break;
      case 2: out[i] = cedm::sin(x[i]); break;
      case 3: out[i] = cedm::cos(x[i]); break;
      case 4: out[i] = cedm::atan2(x[i], y[i]); break;
      case 5: out[i] = cedm::hypot(x[i], y[i]); break;
      case 6: out[i] = cedm::log10(x[i]); break;
      case 7: out[i] = cedm::sqrt_(x[i]); break;
      default: return 1;
    }
  }
  return 0;
}

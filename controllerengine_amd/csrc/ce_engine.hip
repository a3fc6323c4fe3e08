// Engine runtime: device buffers, problem upload, the stage chain, the extended C API and the
// drop-in RunControllerEngine() file protocol (reference src/main.cu:385-602).
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <set>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "ce_host.h"
#include "ce_jit.h"
#include "ce_stage.h"

namespace ce {
// kernels (ce_kernels_fp64.hip, ce_kernels_mixed.hip)
void launch_eig_lin(int nt, const ProblemDev* P, const StageBufs& B, hipStream_t st);
void launch_eig_nom(int ns, const ProblemDev* P, const StageBufs& B, hipStream_t st);
bool launch_mark_refine(int ns, int F, const ProblemDev* P, const StageBufs& B, hipStream_t st);
void launch_scan32(const ProblemDev* P, const StageBufs& B, hipStream_t st);
void launch_scan32_split(const ProblemDev* P, const StageBufs& B, hipStream_t st);
void launch_margin_combine_final(const ProblemDev* P, const StageBufs& B, float* info_out, hipStream_t st);
bool launch_sim_fp64(int ns, int nd, const ProblemDev* P, const StageBufs& B, hipStream_t st);
bool launch_fit_fp64(int nt, const ProblemDev* P, const StageBufs& B, hipStream_t st);
bool launch_sweep_fp64(int ns, int F, const ProblemDev* P, const StageBufs& B, hipStream_t st);
void launch_info_init(const float* in, double* out, size_t count, hipStream_t st);
void launch_info_final(const double* in, float* out, size_t count, hipStream_t st);
bool launch_sim_mixed(int ns, int nd, const ProblemDev* P, const StageBufs& B, hipStream_t st);
bool launch_fit_mixed(int nt, const ProblemDev* P, const StageBufs& B, hipStream_t st);
bool launch_sweep_mixed(int ns, int F, const ProblemDev* P, const StageBufs& B, hipStream_t st);
bool mixed_sweep_fits(int F);
void launch_fill_u64(unsigned long long* p, unsigned long long v, size_t n, hipStream_t st);
void launch_compact(const ProblemDev* P, const StageBufs& B, hipStream_t st);
void launch_gate_wait(const int* gate, int gen, int limit, hipStream_t st);
void launch_fill_i32(int32_t* p, int32_t v, size_t n, hipStream_t st);
void launch_detmath(int fn, const double* x, const double* y, double* out, int n, hipStream_t st);
bool launch_sim_kalman(int ns, int nd, bool mixed, const ProblemDev* P, const float* gains, int n, double* y64,
                       float* y32, hipStream_t st);
}  // namespace ce

using namespace ce;

#define HIPCHK(x)                                                                                   \
  do {                                                                                              \
    hipError_t e_ = (x);                                                                            \
    if (e_ != hipSuccess) {                                                                         \
      throw std::runtime_error(std::string("HIP error ") + hipGetErrorString(e_) + " at " + __FILE__ + \
                               ":" + std::to_string(__LINE__));                                     \
    }                                                                                               \
  } while (0)

constexpr int N_EVENTS = 6;

struct ce_engine {
  int device = 0;
  std::vector<uint8_t> blob;           // last uploaded problem
  uint8_t* blob_pinned = nullptr;
  size_t blob_pinned_cap = 0;
  void* d_prob = nullptr;
  size_t d_prob_cap = 0;
  struct Buf {
    void* p = nullptr;
    size_t cap = 0;
  };
  Buf info64, flags, uniq, mode_re, mode_im, open_a, ybuf, ybuf32, mag_raw, bmin, sfail, list, count, nlist;
  Buf mr32, refined, rlist, rcount, flist, fcount, mres, atbp;
  // per-kernel timing (ce_engine_set_profiling)
  int profiling = 0;
  std::vector<hipEvent_t> mark_ev;
  std::vector<const char*> mark_name;
  int n_marks = 0;
  float* diag_y = nullptr;
  float* diag_lina = nullptr;
  float* diag_eig = nullptr;
  float* diag_ata = nullptr;
  float* diag_atb = nullptr;
  hipEvent_t ev[N_EVENTS] = {};
  bool events_ok = false;
  hipStream_t last_stream = nullptr;
  JitCache* jit = nullptr;
  // the specialised simulation kernel of the uploaded problem, per [mixed][gated] (NULL when its
  // compilation failed); reset when a different problem is uploaded
  void* sim_fn[2][2] = {};
  bool sim_fn_known[2][2] = {};
  bool sim_fn_struct[2][2] = {};  // the structural form (constants from the descriptor)
  void* sim_fn_lanes[2][2][2] = {};  // the literal float module's 2- and 4-lane kernels
  // structures (sim_structure_key, gated or not) whose literal kernel this engine has compiled: a
  // second problem of the same structure means the constants change between calls (serial rounds,
  // a search over plants), and it gets the structural kernel, compiled once per structure
  std::set<std::string> literal_structs;
  long long jit_evictions_seen = 0;
  int last_sim_jit = 0;
  const char* last_sim_name = "";  // kernel name of the last simulation launch
  int last_mixed = 0;
  // mixed chain: the nominal-loop eigensolve and the sweep run on a second stream beside the
  // simulation and the fit (forked after the compaction, joined before the crossing scan)
  hipStream_t side = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  // split simulation (sim_split_slots): the remainder and its fit on a third stream
  hipStream_t rem = nullptr;
  hipEvent_t ev_sima = nullptr, ev_rem = nullptr;
  int last_split_cap = 0;  // the last call's split simulation: the SIMD count of its rule, 0 = no split
  // start signal of the gated simulation (one int, value = gate_gen of the last gated launch)
  Buf gate;
  int gate_gen = 0;
  int n_simd = 0;  // SIMDs of the device (4 per CU)
  // simulation-stage history (ce_engine_set_sim_history): event pairs around the simulation kernel
  // on the launch stream for the last hist_a.size() calls, a ring
  std::vector<hipEvent_t> hist_a, hist_b;
  int hist_pos = 0, hist_n = 0;
};

static void* ensure(ce_engine::Buf& b, size_t bytes) {
  if (bytes == 0) bytes = 256;
  if (b.cap < bytes) {
    if (b.p) HIPCHK(hipFree(b.p));
    b.p = nullptr;
    HIPCHK(hipMalloc(&b.p, bytes));
    b.cap = bytes;
  }
  return b.p;
}

static void put_err(char* err, int errlen, const std::string& m) {
  if (err && errlen > 0) std::snprintf(err, (size_t)errlen, "%s", m.c_str());
}

// Upload the problem blob if it differs from the one on the device (stream-ordered).
static const ProblemDev* upload_problem(ce_engine* E, const std::vector<uint8_t>& blob, hipStream_t st) {
  if (E->d_prob && E->blob == blob) return reinterpret_cast<const ProblemDev*>(E->d_prob);
  // the previous blob may still be read by queued kernels: drain before reusing the buffers
  HIPCHK(hipStreamSynchronize(st));
  for (auto& row : E->sim_fn_known) row[0] = row[1] = false;
  if (E->d_prob_cap < blob.size()) {
    if (E->d_prob) HIPCHK(hipFree(E->d_prob));
    HIPCHK(hipMalloc(&E->d_prob, blob.size()));
    E->d_prob_cap = blob.size();
  }
  if (E->blob_pinned_cap < blob.size()) {
    if (E->blob_pinned) HIPCHK(hipHostFree(E->blob_pinned));
    HIPCHK(hipHostMalloc((void**)&E->blob_pinned, blob.size(), hipHostMallocDefault));
    E->blob_pinned_cap = blob.size();
  }
  std::memcpy(E->blob_pinned, blob.data(), blob.size());
  HIPCHK(hipMemcpyAsync(E->d_prob, E->blob_pinned, blob.size(), hipMemcpyHostToDevice, st));
  E->blob = blob;
  return reinterpret_cast<const ProblemDev*>(E->d_prob);
}

namespace ce {
KernelMarker& kernel_marker() {
  static thread_local KernelMarker m;
  return m;
}
}  // namespace ce

static void record_mark(void* ctx, const char* name, hipStream_t st) {
  ce_engine* E = static_cast<ce_engine*>(ctx);
  if (E->n_marks == (int)E->mark_ev.size()) {
    hipEvent_t ev;
    HIPCHK(hipEventCreate(&ev));
    E->mark_ev.push_back(ev);
    E->mark_name.push_back(nullptr);
  }
  HIPCHK(hipEventRecord(E->mark_ev[E->n_marks], st));
  E->mark_name[E->n_marks] = name;
  E->n_marks++;
}

// Restores the thread's kernel marker when evaluate() returns or throws.
struct MarkerScope {
  KernelMarker saved;
  bool on;
  MarkerScope(ce_engine* E, hipStream_t st) : saved(kernel_marker()), on(E->profiling != 0) {
    if (!on) return;
    E->n_marks = 0;
    kernel_marker() = KernelMarker{record_mark, E};
    record_mark(E, "start", st);
  }
  ~MarkerScope() {
    if (on) kernel_marker() = saved;
  }
};

// The six-stage chain for one candidate block (reference src/main.cu:467-556, one round).
static void evaluate(ce_engine* E, const ce_config& cfg, const ce_plant& plant, int precision, const float* d_gains,
                     const float* d_info_in, float* d_info_out, double* d_info64_out, int n, hipStream_t st) {
  if (n < 0) throw std::runtime_error("ce_evaluate: negative candidate count");
  std::vector<uint8_t> blob;
  std::string err;
  if (!build_problem(cfg, plant, blob, err)) throw std::runtime_error(err);
  const ProblemDev& H = *reinterpret_cast<const ProblemDev*>(blob.data());
  if (n == 0) return;
  if (!E->events_ok) {
    for (int i = 0; i < N_EVENTS; i++) HIPCHK(hipEventCreate(&E->ev[i]));
    E->events_ok = true;
  }
  const ProblemDev* P = upload_problem(E, blob, st);
  const size_t nn = (size_t)n;
  const int nt = H.nt, ns = H.ns, F = H.F;
  StageBufs B{};
  B.n = n;
  B.gains = d_gains;
  B.info = (double*)ensure(E->info64, nn * H.info_w * sizeof(double));
  B.flags = (int32_t*)ensure(E->flags, nn * sizeof(int32_t));
  B.uniq = (int32_t*)ensure(E->uniq, nn * sizeof(int32_t));
  B.mode_re = (double*)ensure(E->mode_re, nn * MAXT * sizeof(double));
  B.mode_im = (double*)ensure(E->mode_im, nn * MAXT * sizeof(double));
  B.open_a = (double*)ensure(E->open_a, nn * ns * ns * sizeof(double));
  if (precision == CE_PRECISION_FP64) B.ybuf = (double*)ensure(E->ybuf, nn * (size_t)H.fit_count * sizeof(double));
  else B.ybuf32 = (float*)ensure(E->ybuf32, nn * (size_t)H.fit_count * sizeof(float));
  const bool mixed = precision == CE_PRECISION_MIXED;
  if (precision != CE_PRECISION_FP64 && !mixed)
    throw std::runtime_error("ce_evaluate: unknown precision " + std::to_string(precision));
  E->last_mixed = mixed ? 1 : 0;
  B.mag_raw = (double*)ensure(E->mag_raw, nn * (size_t)F * (mixed ? 3 : 2) * sizeof(double));
  B.bmin = (unsigned long long*)ensure(E->bmin, nn * sizeof(unsigned long long));
  B.sweep_fail = (int32_t*)ensure(E->sfail, nn * sizeof(int32_t));
  B.list = (int32_t*)ensure(E->list, nn * sizeof(int32_t));
  B.nlist = (int32_t*)ensure(E->nlist, nn * sizeof(int32_t));
  B.count = (int32_t*)ensure(E->count, 2 * sizeof(int32_t));
  B.ncount = B.count + 1;
  if (mixed) {
    B.mr32 = (float*)ensure(E->mr32, nn * (size_t)((F + 63) / 64) * 64 * 4 * sizeof(float));
    B.refined = (int8_t*)ensure(E->refined, ((nn * (size_t)F + 3) / 4) * 4);
    B.rlist = (int2*)ensure(E->rlist, nn * RMAX * sizeof(int2));
    B.rcount = (int32_t*)ensure(E->rcount, sizeof(int32_t));
    B.flist = (int32_t*)ensure(E->flist, nn * sizeof(int32_t));
    B.fcount = (int32_t*)ensure(E->fcount, sizeof(int32_t));
    B.mres = (double*)ensure(E->mres, nn * 6 * sizeof(double));
    // time parts of the right-hand-side pass: enough waves to fill the chip, >= 64 samples each
    const int mb = (nt + 7) / 8;
    const long long waves = (long long)mb * (long long)((nn + 63) / 64);
    int parts = (int)std::max(1LL, std::min(16LL, 8192LL / std::max(1LL, waves)));
    parts = std::max(1, std::min(parts, H.fit_count / 64));
    const int chunk = (H.fit_count + parts - 1) / parts;
    B.atb_chunk = std::max(1, chunk);
    B.atb_parts = (H.fit_count + B.atb_chunk - 1) / B.atb_chunk;
    if (B.atb_parts < 1) B.atb_parts = 1;
    B.atbp = (double2*)ensure(E->atbp, (size_t)B.atb_parts * nt * nn * sizeof(double2));
  }
  B.diag = (E->diag_y || E->diag_lina || E->diag_eig || E->diag_ata || E->diag_atb) ? 1 : 0;
  B.d_y = E->diag_y; B.d_lina = E->diag_lina; B.d_eig = E->diag_eig; B.d_ata = E->diag_ata; B.d_atb = E->diag_atb;
  if (B.diag) {
    // zero fills by kernel (no memset nodes in code a caller may capture into a graph)
    if (B.d_ata) launch_fill_i32((int32_t*)B.d_ata, 0, nn * 4 * nt * nt, st);
    if (B.d_atb) launch_fill_i32((int32_t*)B.d_atb, 0, nn * 2 * nt, st);
  }
  launch_info_init(d_info_in, B.info, nn * H.info_w, st);
  E->last_stream = st;

  // with per-kernel timing on, everything stays on one stream so that the events bracket kernels
  // (the second stream runs the located sweep, which a grid of more than 512 points goes without)
  const bool fork = mixed && !E->profiling && mixed_sweep_fits(F);
  if (fork && !E->side) {
    HIPCHK(hipStreamCreateWithFlags(&E->side, hipStreamNonBlocking));
    HIPCHK(hipEventCreateWithFlags(&E->ev_fork, hipEventDisableTiming));
    HIPCHK(hipEventCreateWithFlags(&E->ev_join, hipEventDisableTiming));
  }
  // the specialised simulation kernel (compiled once per problem and form); with the second stream
  // and a grid that fits the device's SIMDs at one wave each, the gated form
  const int sim_grid = (n + 63) / 64;
  if (!E->n_simd) {
    int cus = 0;
    HIPCHK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, E->device));
    E->n_simd = 4 * cus;
  }
  const bool want_gate = fork && sim_grid <= E->n_simd;
  void* fn = nullptr;
  bool gated = false;
  if (jit_enabled()) {
    if (!E->jit) E->jit = jit_cache_create();
    if (jit_evictions(E->jit) != E->jit_evictions_seen) {  // a cached kernel may have been unloaded
      for (auto& row : E->sim_fn_known) row[0] = row[1] = false;
      E->jit_evictions_seen = jit_evictions(E->jit);
    }
    // resolved once per uploaded problem: generating the source to look it up takes tens of
    // microseconds of host time, during which the device would wait for its next kernel
    auto lookup = [&](bool g) -> void* {
      const int mi = mixed ? 1 : 0, gi = g ? 1 : 0;
      if (!E->sim_fn_known[mi][gi]) {
        std::string jerr;
        void* f = nullptr;
        const std::string lit = gen_sim_source(H, mixed, g);
        const std::string skey = sim_structure_key(H, mixed) + (g ? "|gated" : "");
        const bool structural = !jit_cached(E->jit, lit) && E->literal_structs.count(skey) != 0;
        bool ok;
        void** lanes = E->sim_fn_lanes[mi][gi];
        lanes[0] = lanes[1] = nullptr;
        if (structural) {
          ok = jit_get(E->jit, gen_sim_struct_source(H, mixed, g), &f, jerr, false);
        } else {
          ok = jit_get(E->jit, lit, &f, jerr, true, lanes);
          if (ok) E->literal_structs.insert(skey);
        }
        E->sim_fn[mi][gi] = ok ? f : nullptr;
        E->sim_fn_struct[mi][gi] = structural;
        E->sim_fn_known[mi][gi] = true;
      }
      return E->sim_fn[mi][gi];
    };
    if (want_gate && (fn = lookup(true)) != nullptr) gated = true;
    else fn = lookup(false);
  }
  // stage 1: linearised closed loop, eigenvalues, modes, stability flags (FP64 in both modes)
  MarkerScope marks(E, st);
  HIPCHK(hipEventRecord(E->ev[0], st));
  if (fork && !E->gate.p) {
    ensure(E->gate, sizeof(int));
    launch_fill_i32((int*)E->gate.p, 0, 1, st);
  }
  launch_eig_lin(H.nt, P, B, st);
  // state the later stages start from, enqueued after the first long kernel so that it starts as
  // early as possible
  launch_fill_u64(B.bmin, 0x7ff0000000000000ull, nn, st);  // +inf
  launch_fill_i32(B.sweep_fail, 0, nn, st);
  launch_fill_i32(B.count, 0, 2, st);  // count and ncount
  if (mixed) {
    launch_fill_i32(B.rcount, 0, 1, st);
    launch_fill_i32(B.fcount, 0, 1, st);
  }
  launch_compact(P, B, st);
  // the nominal eigensolve walks the compacted list: after the compaction, on this stream or the
  // second one
  if (!fork) launch_eig_nom(H.ns, P, B, st);
  if (fork) {
    HIPCHK(hipEventRecord(E->ev_fork, st));
    HIPCHK(hipStreamWaitEvent(E->side, E->ev_fork, 0));
    // The second stream starts when the simulation leaves room: the gated form signals once all its
    // workgroups are placed (whole SIMDs stay free beside it); the ungated form (more waves than
    // SIMDs) when its first working wave finishes, i.e. when SIMDs begin to empty.  Started with
    // the simulation instead, the second stream's waves take issue slots from simulation waves
    // still running (c3: step 17.9 -> 18.4 ms with the sweep beside the whole simulation).  The
    // waits are bounded (512 polls at the start, 65536 at the end, a few ms of polling at most):
    // a late signal only delays the second stream.  The interpreted kernel (no JIT) signals nothing
    // and is not waited for.
    E->gate_gen = (E->gate_gen >= (1 << 30)) ? 1 : E->gate_gen + 1;
    if (fn) launch_gate_wait((const int*)E->gate.p, E->gate_gen, gated ? 512 : 65536, E->side);
    launch_eig_nom(H.ns, P, B, E->side);
    if (!launch_mark_refine(ns, F, P, B, E->side)) throw std::runtime_error("no sweep kernel for these dimensions");
    // the crossing scan needs only the sweep: it runs here, beside the fit, and leaves its minima
    // for the row updates after the fit
    launch_scan32_split(P, B, E->side);
    HIPCHK(hipEventRecord(E->ev_join, E->side));
  }
  HIPCHK(hipEventRecord(E->ev[1], st));
  // stage 2: simulation (runtime-specialised unless disabled, else the interpreted kernel)
  void* ybuf = mixed ? (void*)B.ybuf32 : (void*)B.ybuf;
  E->last_sim_jit = 0;
  const int hslot = E->hist_a.empty() ? -1 : E->hist_pos;
  if (hslot >= 0) HIPCHK(hipEventRecord(E->hist_a[hslot], st));
  bool split = false;
  if (fn) {
    const bool st_form = E->sim_fn_struct[mixed ? 1 : 0][gated ? 1 : 0];
    bool ok;
    // lanes per candidate the float literal kernels may use: grids of at most two waves per SIMD.
    // The 4-, 2- and 1-lane kernels are launched in turn, as the grid allows; each runs only if the
    // live count selects its lanes (the most whose working waves stay one per SIMD).
    void* const* lanes = E->sim_fn_lanes[mixed ? 1 : 0][gated ? 1 : 0];
    int gq = 1;
    if (mixed && !st_form && lanes[0] && lanes[1])
      gq = (4 * sim_grid <= 2 * E->n_simd) ? 4 : ((2 * sim_grid <= 2 * E->n_simd) ? 2 : 1);
    // split simulation: ungated one-lane form with room for at least two full rounds (the device
    // decides from the live count whether a small enough remainder is left, sim_split_slots)
    split = fork && !gated && !st_form && mixed && gq == 1 && lanes[0] && sim_grid >= 2 * E->n_simd;
    if (split && !E->rem) {
      HIPCHK(hipStreamCreateWithFlags(&E->rem, hipStreamNonBlocking));
      HIPCHK(hipEventCreateWithFlags(&E->ev_sima, hipEventDisableTiming));
      HIPCHK(hipEventCreateWithFlags(&E->ev_rem, hipEventDisableTiming));
    }
    if (st_form) {
      ok = gated ? jit_launch_sim_struct_gated(fn, sim_grid, st, B.gains, B.flags, B.list, B.count, ybuf, B.info, B.d_y,
                                               n, (int*)E->gate.p, E->gate_gen, P)
                 : jit_launch_sim_struct(fn, sim_grid, st, B.gains, B.flags, B.list, B.count, ybuf, B.info, B.d_y, n, P,
                                         fork ? (int*)E->gate.p : nullptr, E->gate_gen);
    } else if (split) {
      // the one-lane kernel's full rounds here; the remainder (2 lanes per candidate: less than half a
      // round of one-lane waves, so at most one 2-lane wave per SIMD) and its fit on the third stream
      // once this kernel has finished, so that no SIMD holds a wave more than the others beside it
      // (c3, same box: 4 lanes per candidate 16.19 ms per step, 2 lanes 16.00)
      ok = jit_launch_sim(fn, sim_grid, st, B.gains, B.flags, B.list, B.count, ybuf, B.info, B.d_y, n, 1, E->n_simd,
                          (int*)E->gate.p, E->gate_gen, 1);
      HIPCHK(hipEventRecord(E->ev_sima, st));
      HIPCHK(hipStreamWaitEvent(E->rem, E->ev_sima, 0));
      ok = ok && jit_launch_sim(lanes[0], E->n_simd, E->rem, B.gains, B.flags, B.list, B.count, ybuf, B.info,
                                B.d_y, n, 4, E->n_simd, nullptr, 0, 2);
    } else {
      ok = true;
      for (int G : {4, 2, 1}) {
        if (G > gq) continue;
        void* f = G == 4 ? lanes[1] : (G == 2 ? lanes[0] : fn);
        ok = ok && (gated ? jit_launch_sim_gated(f, G * sim_grid, st, B.gains, B.flags, B.list, B.count, ybuf, B.info, B.d_y,
                                                 n, (int*)E->gate.p, E->gate_gen, gq, E->n_simd)
                          : jit_launch_sim(f, G * sim_grid, st, B.gains, B.flags, B.list, B.count, ybuf, B.info, B.d_y, n,
                                           gq, E->n_simd, fork ? (int*)E->gate.p : nullptr, E->gate_gen));
      }
    }
    if (!ok) throw std::runtime_error("launch of the specialised simulation kernel failed");
    E->last_sim_name = st_form ? (mixed ? (gated ? "ce_sims_f32_g" : "ce_sims_f32") : (gated ? "ce_sims_f64_g" : "ce_sims_f64"))
                               : (mixed ? (gated ? "ce_sim_f32_g" : "ce_sim_f32") : (gated ? "ce_sim_f64_g" : "ce_sim_f64"));
    kmark(st_form ? (mixed ? "ce_sims_f32" : "ce_sims_f64") : (mixed ? "ce_sim_f32" : "ce_sim_f64"), st);
    E->last_sim_jit = 1;
  } else {
    const bool ok = mixed ? launch_sim_mixed(ns, H.nd, P, B, st) : launch_sim_fp64(ns, H.nd, P, B, st);
    E->last_sim_name = mixed ? "k_sim_f32" : "k_sim_f64";
    if (!ok) throw std::runtime_error("no simulation kernel for these dimensions");
  }
  E->last_split_cap = split ? E->n_simd : 0;
  if (hslot >= 0) {
    // split simulation: the events bracket the one-lane kernel (the remainder follows on the third
    // stream; the slots of each part are given by ce_sim_split_slots)
    HIPCHK(hipEventRecord(E->hist_b[hslot], st));
    E->hist_pos = (hslot + 1) % (int)E->hist_a.size();
    E->hist_n = std::min(E->hist_n + 1, (int)E->hist_a.size());
  }
  HIPCHK(hipEventRecord(E->ev[2], st));
  // stage 3: modal fit and behaviour score (split simulation: each part's candidates on the stream
  // that simulated them, joined before the row updates)
  if (split) {
    StageBufs B1 = B, B2 = B;
    B1.sim_part = 1;
    B2.sim_part = 2;
    B1.sim_cap = B2.sim_cap = E->n_simd;
    if (!launch_fit_mixed(nt, P, B2, E->rem)) throw std::runtime_error("no fit kernel for these dimensions");
    HIPCHK(hipEventRecord(E->ev_rem, E->rem));
    if (!launch_fit_mixed(nt, P, B1, st)) throw std::runtime_error("no fit kernel for these dimensions");
    HIPCHK(hipStreamWaitEvent(st, E->ev_rem, 0));
  } else if (!(mixed ? launch_fit_mixed(nt, P, B, st) : launch_fit_fp64(nt, P, B, st))) {
    throw std::runtime_error("no fit kernel for these dimensions");
  }
  HIPCHK(hipEventRecord(E->ev[3], st));
  // stage 4: margin sweep, crossings and total score
  if (fork) {
    // the scan's row updates and the float copy of the rows in one launch
    HIPCHK(hipStreamWaitEvent(st, E->ev_join, 0));
    launch_margin_combine_final(P, B, d_info_out, st);
    HIPCHK(hipEventRecord(E->ev[4], st));
  } else {
    if (!(mixed ? launch_sweep_mixed(ns, F, P, B, st) : launch_sweep_fp64(ns, F, P, B, st)))
      throw std::runtime_error("no sweep kernel for these dimensions");
    HIPCHK(hipEventRecord(E->ev[4], st));
    launch_info_final(B.info, d_info_out, nn * H.info_w, st);
  }
  HIPCHK(hipGetLastError());
  if (d_info64_out)
    HIPCHK(hipMemcpyAsync(d_info64_out, B.info, nn * H.info_w * sizeof(double), hipMemcpyDeviceToDevice, st));
  HIPCHK(hipEventRecord(E->ev[5], st));
}

// =================================================================================================
// extended C API

extern "C" CE_EXPORT ce_engine* ce_engine_create(int device, char* err, int errlen) {
  try {
    int count = 0;
    HIPCHK(hipGetDeviceCount(&count));
    if (count <= 0) throw std::runtime_error("no HIP device");
    if (device < 0 || device >= count) throw std::runtime_error("device index out of range");
    HIPCHK(hipSetDevice(device));
    ce_engine* E = new ce_engine();
    E->device = device;
    return E;
  } catch (const std::exception& e) {
    put_err(err, errlen, e.what());
    return nullptr;
  }
}

extern "C" CE_EXPORT void ce_engine_destroy(ce_engine* E) {
  if (!E) return;
  hipDeviceSynchronize();
  for (auto* b : {&E->info64, &E->flags, &E->uniq, &E->mode_re, &E->mode_im, &E->open_a, &E->ybuf, &E->ybuf32,
                  &E->mag_raw, &E->bmin, &E->sfail, &E->list, &E->count, &E->nlist, &E->mr32, &E->refined, &E->rlist,
                  &E->rcount, &E->flist, &E->fcount, &E->mres, &E->atbp, &E->gate})
    if (b->p) hipFree(b->p);
  if (E->d_prob) hipFree(E->d_prob);
  if (E->blob_pinned) hipHostFree(E->blob_pinned);
  jit_cache_destroy(E->jit);
  for (hipEvent_t ev : E->mark_ev) hipEventDestroy(ev);
  for (hipEvent_t ev : E->hist_a) hipEventDestroy(ev);
  for (hipEvent_t ev : E->hist_b) hipEventDestroy(ev);
  if (E->ev_fork) hipEventDestroy(E->ev_fork);
  if (E->ev_join) hipEventDestroy(E->ev_join);
  if (E->side) hipStreamDestroy(E->side);
  if (E->ev_sima) hipEventDestroy(E->ev_sima);
  if (E->ev_rem) hipEventDestroy(E->ev_rem);
  if (E->rem) hipStreamDestroy(E->rem);
  if (E->events_ok)
    for (int i = 0; i < N_EVENTS; i++) hipEventDestroy(E->ev[i]);
  delete E;
}

extern "C" CE_EXPORT int ce_evaluate(ce_engine* E, const ce_config* cfg, const ce_plant* plant, int precision,
                                     const float* d_gains, const float* d_info_in, float* d_info_out,
                                     double* d_info64_out, int n, void* stream, char* err, int errlen) {
  try {
    if (!E || !cfg || !plant) throw std::runtime_error("ce_evaluate: null argument");
    HIPCHK(hipSetDevice(E->device));
    evaluate(E, *cfg, *plant, precision, d_gains, d_info_in, d_info_out, d_info64_out, n, (hipStream_t)stream);
    return 0;
  } catch (const std::exception& e) {
    put_err(err, errlen, e.what());
    return 1;
  }
}

// Kalman-filter simulation (reference src/discrete_solves.cu:83-185): trajectories only.
extern "C" CE_EXPORT int ce_simulate_kalman(ce_engine* E, const ce_config* cfg, const ce_plant* plant, int precision,
                                            const float* d_gains, int n, double* d_y64, float* d_y32, void* stream,
                                            char* err, int errlen) {
  try {
    if (!E || !cfg || !plant) throw std::runtime_error("ce_simulate_kalman: null argument");
    if (precision != CE_PRECISION_FP64 && precision != CE_PRECISION_MIXED)
      throw std::runtime_error("ce_simulate_kalman: unknown precision " + std::to_string(precision));
    const bool mixed = precision == CE_PRECISION_MIXED;
    if (n < 0) throw std::runtime_error("ce_simulate_kalman: negative candidate count");
    if (n > 0 && (!d_gains || (mixed ? !d_y32 : !d_y64)))
      throw std::runtime_error("ce_simulate_kalman: missing gains or output buffer for this precision");
    HIPCHK(hipSetDevice(E->device));
    std::vector<uint8_t> blob;
    std::string e;
    // the margins stage's plant gain at the crossover is not needed here
    if (!build_problem(*cfg, *plant, blob, e) && e.rfind("allmargin:", 0) != 0) throw std::runtime_error(e);
    if (n == 0) return 0;
    const ProblemDev& H = *reinterpret_cast<const ProblemDev*>(blob.data());
    const hipStream_t st = (hipStream_t)stream;
    const ProblemDev* P = upload_problem(E, blob, st);
    if (!launch_sim_kalman(H.ns, H.nd, mixed, P, d_gains, n, d_y64, d_y32, st))
      throw std::runtime_error("no Kalman simulation kernel for these dimensions");
    HIPCHK(hipGetLastError());
    return 0;
  } catch (const std::exception& ex) {
    put_err(err, errlen, ex.what());
    return 1;
  }
}

extern "C" CE_EXPORT int ce_engine_set_diagnostics(ce_engine* E, float* y, float* lina, float* eig, float* ata, float* atb) {
  if (!E) return 1;
  E->diag_y = y; E->diag_lina = lina; E->diag_eig = eig; E->diag_ata = ata; E->diag_atb = atb;
  return 0;
}

extern "C" CE_EXPORT int ce_engine_stage_times(ce_engine* E, double* ms, int n_max) {
  if (!E || !E->events_ok) return 0;
  if (hipEventSynchronize(E->ev[N_EVENTS - 1]) != hipSuccess) return 0;
  int k = 0;
  for (int i = 0; i + 1 < N_EVENTS && k < n_max; i++, k++) {
    float t = 0.0f;
    if (hipEventElapsedTime(&t, E->ev[i], E->ev[i + 1]) != hipSuccess) t = -1.0f;
    ms[k] = t;
  }
  return k;
}

// 1 when the last evaluation's simulation stage ran the runtime-specialised kernel, else 0.
extern "C" CE_EXPORT int ce_engine_last_sim_jit(ce_engine* E) { return E ? E->last_sim_jit : 0; }

extern "C" CE_EXPORT int ce_engine_last_sim_kernel(ce_engine* E, char* name, int len) {
  if (!E || !name || len <= 0) return 1;
  std::snprintf(name, (size_t)len, "%s", E->last_sim_name);
  return 0;
}

extern "C" CE_EXPORT int ce_engine_last_sim_split(ce_engine* E) { return E ? E->last_split_cap : 0; }

extern "C" CE_EXPORT int ce_sim_split_slots(int live, int cap) { return sim_split_slots(live, cap); }

// Simulation-stage history: `calls` event pairs (0 turns it off); durations of the recorded calls.
extern "C" CE_EXPORT int ce_engine_set_sim_history(ce_engine* E, int calls) {
  if (!E || calls < 0) return 1;
  try {
    HIPCHK(hipSetDevice(E->device));
    HIPCHK(hipDeviceSynchronize());
    for (hipEvent_t ev : E->hist_a) hipEventDestroy(ev);
    for (hipEvent_t ev : E->hist_b) hipEventDestroy(ev);
    E->hist_a.assign((size_t)calls, nullptr);
    E->hist_b.assign((size_t)calls, nullptr);
    for (int i = 0; i < calls; i++) {
      HIPCHK(hipEventCreate(&E->hist_a[i]));
      HIPCHK(hipEventCreate(&E->hist_b[i]));
    }
    E->hist_pos = 0;
    E->hist_n = 0;
    return 0;
  } catch (const std::exception&) {
    return 1;
  }
}

extern "C" CE_EXPORT int ce_engine_sim_history(ce_engine* E, double* ms, int n_max) {
  if (!E || E->hist_a.empty()) return 0;
  const int cap = (int)E->hist_a.size();
  const int first = (E->hist_pos - E->hist_n + cap) % cap;
  int k = 0;
  for (int i = 0; i < E->hist_n && k < n_max; i++, k++) {
    const int s = (first + i) % cap;
    float t = 0.0f;
    if (hipEventSynchronize(E->hist_b[s]) != hipSuccess || hipEventElapsedTime(&t, E->hist_a[s], E->hist_b[s]) != hipSuccess)
      return -1;
    ms[k] = t;
  }
  return k;
}

// hipRTC compilations this engine has done.
extern "C" CE_EXPORT long long ce_engine_jit_compiles(ce_engine* E) { return E ? jit_compiles(E->jit) : 0; }

// Mixed sweep counts of the last evaluation: points recomputed in FP64 and candidates whose whole
// sweep was.  Waits for the device.
extern "C" CE_EXPORT int ce_engine_sweep_counts(ce_engine* E, int* refined_points, int* fp64_candidates) {
  if (!E || !E->last_mixed || !E->rcount.p || !E->fcount.p) return 1;
  try {
    HIPCHK(hipSetDevice(E->device));
    HIPCHK(hipDeviceSynchronize());
    int32_t r = 0, f = 0;
    HIPCHK(hipMemcpy(&r, E->rcount.p, sizeof(r), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(&f, E->fcount.p, sizeof(f), hipMemcpyDeviceToHost));
    if (refined_points) *refined_points = r;
    if (fp64_candidates) *fp64_candidates = f;
    return 0;
  } catch (const std::exception&) {
    return 1;
  }
}

extern "C" CE_EXPORT int ce_detmath_device(int fn, const double* x, const double* y, double* out, int n, void* stream) {
  ce::launch_detmath(fn, x, y, out, n, (hipStream_t)stream);
  return hipGetLastError() == hipSuccess ? 0 : 1;
}

extern "C" CE_EXPORT int ce_engine_set_profiling(ce_engine* E, int on) {
  if (!E) return 1;
  E->profiling = on ? 1 : 0;
  return 0;
}

extern "C" CE_EXPORT int ce_engine_kernel_times(ce_engine* E, char* names, int names_len, double* ms, int n_max) {
  if (!E || E->n_marks < 2) return 0;
  if (hipEventSynchronize(E->mark_ev[E->n_marks - 1]) != hipSuccess) return -1;
  std::string list;
  int k = 0;
  for (int i = 1; i < E->n_marks && k < n_max; i++, k++) {
    float t = 0.0f;
    if (hipEventElapsedTime(&t, E->mark_ev[i - 1], E->mark_ev[i]) != hipSuccess) return -1;
    ms[k] = t;
    if (!list.empty()) list += ",";
    list += E->mark_name[i];
  }
  if (names && names_len > 0) {
    std::strncpy(names, list.c_str(), (size_t)names_len - 1);
    names[names_len - 1] = 0;
  }
  return k;
}

// =================================================================================================
// Drop-in entry point: reference src/main.cu:385-602, same files, messages and return convention.

static void write_file(const char* name, const void* data, size_t bytes) {
  std::ofstream f(name, std::ios::binary);
  if (!f) throw std::runtime_error(std::string("Cannot open ") + name);
  f.write(reinterpret_cast<const char*>(data), (std::streamsize)bytes);
}

static void run_from_cwd() {
  std::string err;
  std::string text;
  {
    std::ifstream f("ControllerEngine.config");
    if (!f) throw std::runtime_error("Cannot open ControllerEngine.config");
    std::ostringstream ss;
    ss << f.rdbuf();
    text = ss.str();
  }
  ce_config cfg;
  if (!parse_config(text, "ControllerEngine.config", cfg, err)) throw std::runtime_error(err);
  ce_derived d;
  if (!derive_config(cfg, d, err)) throw std::runtime_error(err);
  ce_plant plant;
  if (!load_plant_dir("", cfg, 0, plant, err)) throw std::runtime_error(err);
  {
    // decode once up front so that undecodable instructions fail before any output is written
    std::vector<uint8_t> blob;
    if (!build_problem(cfg, plant, blob, err)) {
      if (err.rfind("allmargin:", 0) != 0) throw std::runtime_error(err);
    }
  }
  int precision = CE_PRECISION_FP64;
  if (const char* p = std::getenv("CE_PRECISION")) {
    const std::string v = p;
    if (v == "mixed") precision = CE_PRECISION_MIXED;
    else if (v == "fp64" || v.empty()) precision = CE_PRECISION_FP64;
    else throw std::runtime_error("CE_PRECISION should be fp64/mixed, got `" + v + "`");
  }
  {
    std::ofstream plog("params.log");
    plog << std::scientific << std::setprecision(9);
    for (int i = 0; i < cfg.n_params; i++) plog << "BASE_PARAMS_HOST[" << i << "] = " << plant.base_params[i] << "\n";
    plog << "ref amplitude = " << cfg.ref_amplitude << "\n";
  }
  const size_t nthr = (size_t)d.n_threads;
  const size_t gw = (size_t)d.cuda_mx_width, iw = (size_t)d.info_mx_width;
  const size_t total = (size_t)cfg.n_rounds * nthr;
  std::ifstream cudamx("CUDA_MX.bin", std::ios::binary);
  if (!cudamx) throw std::runtime_error("Cannot open CUDA_MX.bin");
  cudamx.seekg(0, std::ios::end);
  if ((size_t)cudamx.tellg() != total * gw * 4)
    throw std::runtime_error("CUDA_MX.bin is " + std::to_string((long long)cudamx.tellg()) + " bytes, expected " +
                             std::to_string(total * gw * 4) + " (" + std::to_string(cfg.n_rounds) + " rounds x " +
                             std::to_string(nthr) + " candidates)");
  cudamx.seekg(0, std::ios::beg);
  std::ifstream infomx("INFO_MX.bin", std::ios::binary);
  if (!infomx) throw std::runtime_error("Cannot open INFO_MX.bin");
  infomx.seekg(0, std::ios::end);
  if ((size_t)infomx.tellg() != total * iw * 4)
    throw std::runtime_error("INFO_MX.bin is " + std::to_string((long long)infomx.tellg()) + " bytes, expected " +
                             std::to_string(total * iw * 4) + " (" + std::to_string(cfg.n_rounds) + " rounds x " +
                             std::to_string(nthr) + " candidates)");
  infomx.seekg(0, std::ios::beg);
  std::ofstream infoout("INFO_MX_OUT.bin", std::ios::binary);
  if (!infoout) throw std::runtime_error("Cannot open INFO_MX_OUT.bin");

  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  ce_engine* E = ce_engine_create(dev, nullptr, 0);
  if (!E) throw std::runtime_error("cannot create engine");
  struct Guard {
    ce_engine* e;
    std::vector<void*> mem;
    ~Guard() {
      for (void* p : mem) hipFree(p);
      ce_engine_destroy(e);
    }
  } guard{E, {}};
  auto dalloc = [&](size_t bytes) {
    void* p = nullptr;
    HIPCHK(hipMalloc(&p, bytes ? bytes : 4));
    guard.mem.push_back(p);
    return p;
  };
  const int nt = d.total_dim;
  float* d_g = (float*)dalloc(nthr * gw * 4);
  float* d_in = (float*)dalloc(nthr * iw * 4);
  float* d_out = (float*)dalloc(nthr * iw * 4);
  std::vector<float> h_g(nthr * gw), h_info(nthr * iw);
  float *dy = nullptr, *dl = nullptr, *de = nullptr, *da = nullptr, *db = nullptr;
  if (cfg.dump_diagnostics) {
    dy = (float*)dalloc(nthr * (size_t)cfg.time * 4);
    dl = (float*)dalloc(nthr * (size_t)nt * nt * 4);
    de = (float*)dalloc(nthr * (size_t)2 * nt * 4);
    da = (float*)dalloc(nthr * (size_t)4 * nt * nt * 4);
    db = (float*)dalloc(nthr * (size_t)2 * nt * 4);
    ce_engine_set_diagnostics(E, dy, dl, de, da, db);
  }
  hipStream_t st = nullptr;
  for (int round = 0; round < cfg.n_rounds; round++) {
    if (cfg.serial && round > 0)
      if (!load_plant_dir("", cfg, round, plant, err)) throw std::runtime_error(err);
    cudamx.read(reinterpret_cast<char*>(h_g.data()), (std::streamsize)(h_g.size() * 4));
    if (!cudamx)
      throw std::runtime_error("Cannot read CUDA_MX.bin round " + std::to_string(round) + " (intended " +
                               std::to_string(h_g.size() * 4) + "-byte allocation)");
    infomx.read(reinterpret_cast<char*>(h_info.data()), (std::streamsize)(h_info.size() * 4));
    if (!infomx)
      throw std::runtime_error("Cannot read INFO_MX.bin round " + std::to_string(round) + " (intended " +
                               std::to_string(h_info.size() * 4) + "-byte allocation)");
    HIPCHK(hipMemcpy(d_g, h_g.data(), h_g.size() * 4, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(d_in, h_info.data(), h_info.size() * 4, hipMemcpyHostToDevice));
    evaluate(E, cfg, plant, precision, d_g, d_in, d_out, nullptr, (int)nthr, st);
    HIPCHK(hipStreamSynchronize(st));
    HIPCHK(hipMemcpy(h_info.data(), d_out, h_info.size() * 4, hipMemcpyDeviceToHost));
    if (cfg.dump_diagnostics) {
      std::vector<float> tmp;
      auto dump = [&](const char* name, const float* dp, size_t count) {
        tmp.resize(count);
        HIPCHK(hipMemcpy(tmp.data(), dp, count * 4, hipMemcpyDeviceToHost));
        write_file(name, tmp.data(), count * 4);
      };
      dump("YPLANT_OUT.bin", dy, nthr * (size_t)cfg.time);
      dump("LINA_MX_OUT.bin", dl, nthr * (size_t)nt * nt);
      dump("EIGEN_MX_OUT.bin", de, nthr * (size_t)2 * nt);
      dump("ATA_MX_OUT.bin", da, nthr * (size_t)4 * nt * nt);
      dump("ATB_MX_OUT.bin", db, nthr * (size_t)2 * nt);
    }
    infoout.write(reinterpret_cast<const char*>(h_info.data()), (std::streamsize)(h_info.size() * 4));
    infoout.flush();
  }
}

extern "C" CE_EXPORT int RunControllerEngine(void) {
  try {
    run_from_cwd();
    return 0;
  } catch (const std::exception& e) {
    std::ofstream elog("error.log", std::ios::app);
    elog << "ControllerEngine error: " << e.what() << std::endl;
    return 1;
  }
}

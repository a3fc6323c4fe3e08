// Batched evaluation: many problems (plants / configurations, uneven dimensions and horizons) in
// one chain of grouped launches.
//
// Every stage kernel of the single-problem chain has a batched form (BatchTab): one launch covers
// all problems that use the same template instance, each problem owning a contiguous range of the
// launch's blocks; workgroups loop over blocks with a stride of the grid (persistent launch).  The
// simulation uses the structural runtime-specialised kernel (one compiled kernel per dimensions +
// instruction structure, constants read from each problem's descriptor).  Device buffers of all
// problems are carved from one arena, so the fills and the INFO conversions are single launches.
// The whole chain can be captured into a hipGraph and replayed while the batch layout and the
// INFO pointers stay the same (new plant values are uploaded before each replay); an autotuner
// picks each launch's grid size from a few candidates by timing whole-chain passes.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "ce_host.h"
#include "ce_jit.h"
#include "ce_stage.h"

namespace ce {
void launch_info_init(const float* in, double* out, size_t count, hipStream_t st);
void launch_info_final(const double* in, float* out, size_t count, hipStream_t st);
void launch_fill_u64(unsigned long long* p, unsigned long long v, size_t n, hipStream_t st);
void launch_fill_i32(int32_t* p, int32_t v, size_t n, hipStream_t st);
bool launch_sim_fp64(int ns, int nd, const ProblemDev* P, const StageBufs& B, hipStream_t st);
bool launch_sim_mixed(int ns, int nd, const ProblemDev* P, const StageBufs& B, hipStream_t st);
}  // namespace ce

using namespace ce;

#define BCHK(x)                                                                                     \
  do {                                                                                              \
    hipError_t e_ = (x);                                                                            \
    if (e_ != hipSuccess)                                                                           \
      throw std::runtime_error(std::string("HIP error ") + hipGetErrorString(e_) + " at " + __FILE__ + ":" + \
                               std::to_string(__LINE__));                                           \
  } while (0)

namespace {

struct DevBuf {
  void* p = nullptr;
  size_t cap = 0;
  void* get(size_t bytes) {
    if (bytes == 0) bytes = 256;
    if (cap < bytes) {
      if (p) BCHK(hipFree(p));
      p = nullptr;
      BCHK(hipMalloc(&p, bytes));
      cap = bytes;
    }
    return p;
  }
  void release() {
    if (p) hipFree(p);
    p = nullptr;
    cap = 0;
  }
};

constexpr int SIM_LAUNCH = 100;  // launch family of the simulation (not a BatchKernel)

struct Launch {
  int kind;           // BatchKernel or SIM_LAUNCH
  int key;            // template key
  std::string skey;   // simulation structure key
  std::vector<int> pid, start;
  std::vector<int> sid;  // fused simulation: structure index of each problem (pid order)
  int32_t* d_start = nullptr;
  int32_t* d_pid = nullptr;
  int32_t* d_sid = nullptr;
  void* fn = nullptr;  // simulation kernel
  void* fl[2] = {nullptr, nullptr};  // its 2- and 4-lane forms (float)
  int grid = 0;        // chosen grid
  std::string label;
};

size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

}  // namespace

struct ce_batch {
  int device = 0;
  struct Prob {
    ce_config cfg;
    ce_plant plant;
    int n = 0;
    std::vector<uint8_t> blob;
    bool need_build = true;   // cfg / plant changed since the descriptor was built
    bool blob_dirty = true;   // descriptor not yet on the device
    std::string skey;         // simulation structure key of the built descriptor
    int ns = 0, nd = 0, nt = 0, F = 0, fc = 0, iw = 0, cw = 0;
    size_t goff = 0, ioff = 0, noff = 0;
  };
  std::vector<Prob> P;
  bool layout_dirty = true;
  int layout_precision = -1;
  // device
  DevBuf blobs, ptrs, bufs, tabs, sargs, work;
  uint8_t* pinned = nullptr;
  size_t pinned_cap = 0;
  std::vector<size_t> blob_off;
  size_t blob_total = 0;
  std::vector<StageBufs> hb;
  std::vector<const ProblemDev*> hptrs;
  std::vector<SimArgs> hsargs;
  size_t total_n = 0, total_info = 0;
  double* info64 = nullptr;
  unsigned long long* bmin = nullptr;
  int32_t* sfail = nullptr;
  int32_t* counts = nullptr;   // 2 J: list length per problem, then nominal-list length per problem
  int32_t* rcounts = nullptr;  // refinement count per problem
  int32_t* fcounts = nullptr;  // FP64 sweep list length per problem
  std::vector<Launch> launches;
  const float* last_gains = nullptr;
  // graph
  int use_graph = 1;
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  const void* g_in = nullptr;
  const void* g_out = nullptr;
  const void* g_out64 = nullptr;
  // graphs are captured and replayed on the batch's own stream (the caller's may be the legacy
  // default stream, which cannot be captured), ordered after / before the caller's stream by events
  hipStream_t own = nullptr;
  hipEvent_t ev_in = nullptr, ev_out = nullptr;
  // mixed chains fork onto a second stream (see enqueue)
  hipStream_t side = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  // tuning
  int autotune = 0;
  bool tuned = false;
  // stage plans: [0] LIN_A eigensolve, [1] nominal eigensolve: 0 one launch per matrix capacity
  // with the default lanes per candidate, 1 the same with the alternative lanes, 2 one fused launch
  // over every order; [2] simulation: 0 one launch per instruction structure, 1 one fused launch.
  // The autotuner picks them unless a caller fixed them (ce_batch_set_plan).
  int plan[3] = {0, 0, 0};
  bool plan_fixed = false;
  int sim_lanes = 1;  // lanes per candidate of the simulation launches (float form)
  // per-launch timing
  int profiling = 0;
  std::vector<hipEvent_t> mark_ev;
  std::vector<std::string> mark_name;
  int n_marks = 0;
  int last_used_graph = 0;
};

static void put_err_b(char* err, int errlen, const std::string& m) {
  if (err && errlen > 0) std::snprintf(err, (size_t)errlen, "%s", m.c_str());
}

static void drop_graph(ce_batch* b) {
  if (b->gexec) hipGraphExecDestroy(b->gexec);
  if (b->graph) hipGraphDestroy(b->graph);
  b->gexec = nullptr;
  b->graph = nullptr;
}

// Build the changed problems' descriptors on several host threads; a changed shape or
// simulation structure invalidates the device layout.
static void build_blobs(ce_batch* b, int precision) {
  std::vector<int> todo;
  for (int j = 0; j < (int)b->P.size(); j++)
    if (b->P[j].need_build) todo.push_back(j);
  if (todo.empty()) return;
  std::vector<std::string> errs(todo.size());
  const int nth = (int)std::max<size_t>(1, std::min<size_t>({(size_t)std::thread::hardware_concurrency(), (size_t)16, todo.size()}));
  auto work = [&](int t) {
    for (size_t i = t; i < todo.size(); i += nth) {
      auto& p = b->P[todo[i]];
      std::string e;
      if (!build_problem(p.cfg, p.plant, p.blob, e)) errs[i] = e;
    }
  };
  if (nth == 1) {
    work(0);
  } else {
    std::vector<std::thread> th;
    for (int t = 0; t < nth; t++) th.emplace_back(work, t);
    for (auto& x : th) x.join();
  }
  for (size_t i = 0; i < todo.size(); i++)
    if (!errs[i].empty()) throw std::runtime_error("batch problem " + std::to_string(todo[i]) + ": " + errs[i]);
  const bool mixed = precision == CE_PRECISION_MIXED;
  for (int j : todo) {
    auto& p = b->P[j];
    const ProblemDev& H = *reinterpret_cast<const ProblemDev*>(p.blob.data());
    const std::string key = sim_structure_key(H, mixed);
    // the descriptor's size depends on F only
    if (H.ns != p.ns || H.nd != p.nd || H.F != p.F || H.fit_count != p.fc || H.info_w != p.iw || key != p.skey)
      b->layout_dirty = true;
    p.ns = H.ns;
    p.nd = H.nd;
    p.nt = H.nt;
    p.F = H.F;
    p.fc = H.fit_count;
    p.iw = H.info_w;
    p.cw = H.cuda_w;
    p.skey = key;
    p.need_build = false;
    p.blob_dirty = true;
  }
}

// eigen kernels: key = matrix capacity (4, 8, 16), candidates per workgroup = 64 / lanes

// Compiled simulation kernels are shared by every batch of the process on the same device (the
// caches are thread-safe); modules stay loaded until the process ends.
static JitCache* shared_jit(int device) {
  static std::mutex mu;
  static std::map<int, JitCache*> caches;
  std::lock_guard<std::mutex> lk(mu);
  JitCache*& c = caches[device];
  if (!c) c = jit_cache_create();
  return c;
}

// Device layout of all buffers, StageBufs, launch groups and their tables.
static void build_layout(ce_batch* b, int precision, hipStream_t st) {
  const bool mixed = precision == CE_PRECISION_MIXED;
  const int J = (int)b->P.size();
  // problem descriptors: one arena
  b->blob_off.resize(J);
  size_t off = 0;
  for (int j = 0; j < J; j++) {
    b->blob_off[j] = off;
    off += align256(b->P[j].blob.size());
  }
  b->blob_total = off;
  uint8_t* d_blobs = (uint8_t*)b->blobs.get(off);
  b->hptrs.resize(J);
  for (int j = 0; j < J; j++) b->hptrs[j] = reinterpret_cast<const ProblemDev*>(d_blobs + b->blob_off[j]);
  // row offsets of the caller's arrays
  size_t go = 0, io = 0, no = 0;
  for (auto& p : b->P) {
    p.goff = go;
    p.ioff = io;
    p.noff = no;
    go += (size_t)p.n * p.cw;
    io += (size_t)p.n * p.iw;
    no += (size_t)p.n;
  }
  b->total_n = no;
  b->total_info = io;
  // time parts of the mixed right-hand-side pass over the whole batch
  long long atb_waves = 0;
  for (auto& p : b->P) atb_waves += (long long)((p.n + 63) / 64) * ((p.nt + 7) / 8);
  const int parts_all = (int)std::max(1LL, std::min(16LL, 8192LL / std::max(1LL, atb_waves)));
  // work arena
  struct Seg {
    size_t off, bytes;
  };
  size_t wo = 0;
  auto seg = [&](size_t bytes) {
    Seg s{wo, bytes};
    wo += align256(bytes ? bytes : 4);
    return s;
  };
  const Seg s_info = seg(io * sizeof(double));
  const Seg s_bmin = seg(no * 8), s_sfail = seg(no * 4), s_counts = seg(2 * J * 4), s_rcounts = seg(J * 4);
  const Seg s_fcounts = seg(J * 4);
  struct PS {
    Seg flags, uniq, mre, mim, oa, y, mag, list, nlist, mr32, refined, rlist, flist, atbp;
    int parts, chunk;
  };
  std::vector<PS> ps(J);
  for (int j = 0; j < J; j++) {
    auto& p = b->P[j];
    const size_t n = p.n;
    PS& s = ps[j];
    s.flags = seg(n * 4);
    s.uniq = seg(n * 4);
    s.mre = seg(n * MAXT * 8);
    s.mim = seg(n * MAXT * 8);
    s.oa = seg(n * p.ns * p.ns * 8);
    s.y = seg(n * (size_t)p.fc * (mixed ? 4 : 8));
    s.mag = seg(n * (size_t)p.F * (mixed ? 3 : 2) * 8);
    s.list = seg(n * 4);
    s.nlist = seg(n * 4);
    if (mixed) {
      s.mr32 = seg(n * (size_t)((p.F + 63) / 64) * 64 * 4 * 4);
      s.refined = seg(((n * (size_t)p.F + 3) / 4) * 4);
      s.rlist = seg(n * RMAX * sizeof(int2));
      s.flist = seg(n * 4);
      int parts = std::max(1, std::min(parts_all, p.fc / 64));
      const int chunk = std::max(1, (p.fc + parts - 1) / parts);
      parts = std::max(1, (p.fc + chunk - 1) / chunk);
      s.parts = parts;
      s.chunk = chunk;
      s.atbp = seg((size_t)parts * p.nt * n * sizeof(double2));
    }
  }
  uint8_t* W = (uint8_t*)b->work.get(wo);
  b->info64 = (double*)(W + s_info.off);
  b->bmin = (unsigned long long*)(W + s_bmin.off);
  b->sfail = (int32_t*)(W + s_sfail.off);
  b->counts = (int32_t*)(W + s_counts.off);
  b->rcounts = (int32_t*)(W + s_rcounts.off);
  b->fcounts = (int32_t*)(W + s_fcounts.off);
  b->hb.assign(J, StageBufs{});
  b->hsargs.assign(J, SimArgs{});
  for (int j = 0; j < J; j++) {
    auto& p = b->P[j];
    PS& s = ps[j];
    StageBufs& B = b->hb[j];
    B.n = p.n;
    B.diag = 0;
    B.gains = nullptr;  // set per run
    B.info = b->info64 + p.ioff;
    B.flags = (int32_t*)(W + s.flags.off);
    B.uniq = (int32_t*)(W + s.uniq.off);
    B.mode_re = (double*)(W + s.mre.off);
    B.mode_im = (double*)(W + s.mim.off);
    B.open_a = (double*)(W + s.oa.off);
    if (mixed) B.ybuf32 = (float*)(W + s.y.off);
    else B.ybuf = (double*)(W + s.y.off);
    B.mag_raw = (double*)(W + s.mag.off);
    B.bmin = b->bmin + p.noff;
    B.sweep_fail = b->sfail + p.noff;
    B.list = (int32_t*)(W + s.list.off);
    B.count = b->counts + j;
    B.nlist = (int32_t*)(W + s.nlist.off);
    B.ncount = b->counts + J + j;
    if (mixed) {
      B.mr32 = (float*)(W + s.mr32.off);
      B.refined = (int8_t*)(W + s.refined.off);
      B.rlist = (int2*)(W + s.rlist.off);
      B.rcount = b->rcounts + j;
      B.flist = (int32_t*)(W + s.flist.off);
      B.fcount = b->fcounts + j;
      B.atbp = (double2*)(W + s.atbp.off);
      B.atb_parts = s.parts;
      B.atb_chunk = s.chunk;
    }
    SimArgs& A = b->hsargs[j];
    A.flags = B.flags;
    A.list = B.list;
    A.count = B.count;
    A.ybuf = mixed ? (void*)B.ybuf32 : (void*)B.ybuf;
    A.info = B.info;
    A.dy = nullptr;
    A.n = p.n;
  }
  // launch groups, in chain order
  b->launches.clear();
  auto group = [&](int kind, auto keyf, auto blocksf, const char* name, auto labelf) {
    std::map<int, Launch> g;
    for (int j = 0; j < J; j++) {
      const int key = keyf(b->P[j]);
      Launch& L = g[key];
      L.kind = kind;
      L.key = key;
      if (L.start.empty()) L.start.push_back(0);
      L.pid.push_back(j);
      L.start.push_back(L.start.back() + blocksf(b->P[j], ps[j]));
    }
    for (auto& kv : g) {
      kv.second.label = std::string(name) + labelf(kv.first);
      b->launches.push_back(kv.second);
    }
  };
  auto num_label = [](int k) { return k ? "/" + std::to_string(k) : std::string(); };
  auto cdiv = [](int a, int c) { return (a + c - 1) / c; };
  using PP = ce_batch::Prob;
  // eigen stages: key = capacity * 100 + lanes per candidate (the plan's tile), 0 for the fused launch
  auto eig_tile = [](int cap, int plan) {
    if (plan == 2) return 0;
    if (cap == 4) return 404;
    if (cap == 8) return plan == 1 ? 808 : 800 + EIG8_G;
    return plan == 1 ? 1608 : 1600 + EIG16_G;
  };
  auto eig_blocks = [&](int order, int plan, int n) {
    const int cap = eig_key(order), key = eig_tile(cap, plan);
    const int lanes = key ? key % 100 : eig_lanes(cap);
    return cdiv(n, 64 / lanes);
  };
  auto eig_label = [](int k) { return k ? "/" + std::to_string(k / 100) + "x" + std::to_string(k % 100) : std::string("/all"); };
  group(BK_EIG_LIN, [&](const PP& p) { return eig_tile(eig_key(p.nt), b->plan[0]); },
        [&](const PP& p, const PS&) { return eig_blocks(p.nt, b->plan[0], p.n); }, "k_eig_lin", eig_label);
  group(BK_COMPACT, [](const PP&) { return 0; }, [&](const PP& p, const PS&) { return cdiv(p.n, 64); }, "k_compact", num_label);
  // the nominal eigensolve walks each problem's compacted list, so it follows the compaction
  group(BK_EIG_NOM, [&](const PP& p) { return eig_tile(eig_key(2 * p.ns), b->plan[1]); },
        [&](const PP& p, const PS&) { return eig_blocks(2 * p.ns, b->plan[1], p.n); }, "k_eig_nom", eig_label);
  {  // simulation: one group per structure, its problems longest horizon first (a wave's work is
     // proportional to the horizon: the long ones start first instead of forming the tail).  The
     // float form splits a candidate's step over 2 or 4 lanes when the batch's candidates would
     // leave SIMDs idle at one lane each (serving batches of small jobs).
    int nsimd = 1024;
    {
      int cus = 0;
      if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, b->device) == hipSuccess && cus > 0)
        nsimd = 4 * cus;
    }
    int lanes = 1;
    if (mixed)
      for (int G : {4, 2}) {
        long long waves = 0;
        for (int j = 0; j < J; j++) waves += cdiv(b->P[j].n, 64 / G);
        if (waves <= nsimd) {
          lanes = G;
          break;
        }
      }
    b->sim_lanes = lanes;
    std::vector<int> order(J);
    for (int j = 0; j < J; j++) order[j] = j;
    auto horizon = [&](int j) { return reinterpret_cast<const ProblemDev*>(b->P[j].blob.data())->T; };
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return horizon(x) > horizon(y); });
    if (b->plan[2] == 1) {  // every structure in one launch
      Launch L;
      L.kind = SIM_LAUNCH;
      L.key = 1;
      L.start.push_back(0);
      std::map<std::string, int> sidx;
      std::vector<const ProblemDev*> reps;
      for (const int j : order) {
        const ProblemDev* H = reinterpret_cast<const ProblemDev*>(b->P[j].blob.data());
        const std::string k = sim_structure_key(*H, mixed);
        auto it = sidx.find(k);
        if (it == sidx.end()) {
          it = sidx.emplace(k, (int)reps.size()).first;
          reps.push_back(H);
        }
        L.pid.push_back(j);
        L.sid.push_back(it->second);
        L.start.push_back(L.start.back() + cdiv(b->P[j].n, 64 / lanes));
      }
      std::string e;
      if (!jit_get(shared_jit(b->device), gen_sim_batch_fused_source(reps, mixed), &L.fn, e, false, L.fl))
        throw std::runtime_error(e);
      L.label = std::string(mixed ? "ce_simb_f32" : "ce_simb_f64") + "/all";
      b->launches.push_back(L);
    }
    std::map<std::string, Launch> g;
    for (const int j : (b->plan[2] == 1 ? std::vector<int>() : order)) {
      const ProblemDev& H = *reinterpret_cast<const ProblemDev*>(b->P[j].blob.data());
      const std::string k = sim_structure_key(H, mixed);
      Launch& L = g[k];
      if (L.start.empty()) {
        L.kind = SIM_LAUNCH;
        L.skey = k;
        L.start.push_back(0);
        std::string e;
        if (!jit_get(shared_jit(b->device), gen_sim_batch_source(H, mixed), &L.fn, e, false, L.fl))
          throw std::runtime_error(e);
      }
      L.pid.push_back(j);
      L.start.push_back(L.start.back() + cdiv(b->P[j].n, 64 / lanes));
    }
    int i = 0;
    for (auto& kv : g) {
      kv.second.label = std::string(mixed ? "ce_simb_f32" : "ce_simb_f64") + "/" + std::to_string(i++);
      b->launches.push_back(kv.second);
    }
  }
  auto nb = [](const PP& p) { return (2 * p.nt + 7) / 8; };
  if (mixed) {
    group(BK_ATB, [](const PP&) { return 0; },
          [&](const PP& p, const PS& s) { return cdiv(p.n, 64) * cdiv(p.nt, 8) * s.parts; }, "k_atb", num_label);
    group(BK_FITCF, nb, [&](const PP& p, const PS&) { return cdiv(p.n, 4); }, "k_fitcf", num_label);
    group(BK_SWEEP_MARK32, [](const PP& p) { return p.ns; }, [](const PP& p, const PS&) { return p.n; }, "k_sweep_mark32", num_label);
    group(BK_REFINE64, [](const PP& p) { return p.ns; }, [&](const PP& p, const PS&) { return cdiv(p.n, 64); }, "k_refine64", num_label);
    group(BK_FULL64, [](const PP& p) { return p.ns; }, [&](const PP& p, const PS&) { return cdiv(p.n, 64); }, "k_full64", num_label);
    group(BK_SCAN32, [](const PP&) { return 0; }, [](const PP& p, const PS&) { return p.n; }, "k_scan32", num_label);
  } else {
    group(BK_FIT_FP64, nb, [&](const PP& p, const PS&) { return cdiv(p.n, 4); }, "k_fit_fp64", num_label);
    group(BK_SWEEP_FP64, [](const PP& p) { return p.ns; }, [&](const PP& p, const PS&) { return p.n * cdiv(p.F, 64); }, "k_sweep_fp64", num_label);
    group(BK_SCAN_FP64, [](const PP&) { return 0; }, [&](const PP& p, const PS&) { return cdiv(p.n, 64); }, "k_scan_fp64", num_label);
  }
  // tables: start + pid of every launch in one buffer
  size_t tw = 0;
  for (auto& L : b->launches) tw += L.start.size() + L.pid.size() + L.sid.size();
  std::vector<int32_t> tab;
  tab.reserve(tw);
  std::vector<size_t> toff;
  for (auto& L : b->launches) {
    toff.push_back(tab.size());
    tab.insert(tab.end(), L.start.begin(), L.start.end());
    tab.insert(tab.end(), L.pid.begin(), L.pid.end());
    tab.insert(tab.end(), L.sid.begin(), L.sid.end());
  }
  int32_t* d_tab = (int32_t*)b->tabs.get(tab.size() * 4);
  for (size_t i = 0; i < b->launches.size(); i++) {
    Launch& L = b->launches[i];
    L.d_start = d_tab + toff[i];
    L.d_pid = d_tab + toff[i] + L.start.size();
    L.d_sid = L.sid.empty() ? nullptr : d_tab + toff[i] + L.start.size() + L.pid.size();
    L.grid = L.start.back();
  }
  BCHK(hipMemcpyAsync(d_tab, tab.data(), tab.size() * 4, hipMemcpyHostToDevice, st));
  BCHK(hipMemcpyAsync(b->ptrs.get(J * sizeof(void*)), b->hptrs.data(), J * sizeof(void*), hipMemcpyHostToDevice, st));
  BCHK(hipStreamSynchronize(st));  // host vectors above are reused
  b->layout_dirty = false;
  b->layout_precision = precision;
  b->last_gains = nullptr;
  b->tuned = false;
  drop_graph(b);
}

static void upload_blobs(ce_batch* b, hipStream_t st) {
  bool any = false;
  for (auto& p : b->P) any = any || p.blob_dirty;
  if (!any) return;
  if (b->pinned_cap < b->blob_total) {
    BCHK(hipStreamSynchronize(st));
    if (b->pinned) BCHK(hipHostFree(b->pinned));
    BCHK(hipHostMalloc((void**)&b->pinned, b->blob_total, hipHostMallocDefault));
    b->pinned_cap = b->blob_total;
  } else {
    BCHK(hipStreamSynchronize(st));  // the staging copy may still be in flight
  }
  for (size_t j = 0; j < b->P.size(); j++) {
    std::memcpy(b->pinned + b->blob_off[j], b->P[j].blob.data(), b->P[j].blob.size());
    b->P[j].blob_dirty = false;
  }
  BCHK(hipMemcpyAsync(b->blobs.p, b->pinned, b->blob_total, hipMemcpyHostToDevice, st));
}

static void set_gains(ce_batch* b, const float* d_gains, hipStream_t st) {
  if (b->last_gains == d_gains) return;
  for (size_t j = 0; j < b->P.size(); j++) {
    b->hb[j].gains = d_gains + b->P[j].goff;
    b->hsargs[j].gains = d_gains + b->P[j].goff;
  }
  BCHK(hipMemcpyAsync(b->bufs.get(b->hb.size() * sizeof(StageBufs)), b->hb.data(), b->hb.size() * sizeof(StageBufs),
                      hipMemcpyHostToDevice, st));
  BCHK(hipMemcpyAsync(b->sargs.get(b->hsargs.size() * sizeof(SimArgs)), b->hsargs.data(),
                      b->hsargs.size() * sizeof(SimArgs), hipMemcpyHostToDevice, st));
  BCHK(hipStreamSynchronize(st));
  b->last_gains = d_gains;
}

static void batch_mark(ce_batch* b, const std::string& name, hipStream_t st) {
  if (!b->profiling) return;
  if (b->n_marks == (int)b->mark_ev.size()) {
    hipEvent_t ev;
    BCHK(hipEventCreate(&ev));
    b->mark_ev.push_back(ev);
    b->mark_name.emplace_back();
  }
  BCHK(hipEventRecord(b->mark_ev[b->n_marks], st));
  b->mark_name[b->n_marks] = name;
  b->n_marks++;
}

// The stage chain over the whole batch.
static void enqueue(ce_batch* b, int precision, const float* d_in, float* d_out, double* d_out64, hipStream_t st,
                    const std::vector<int>* grids) {
  const bool mixed = precision == CE_PRECISION_MIXED;
  const int J = (int)b->P.size();
  b->n_marks = 0;
  batch_mark(b, "start", st);
  launch_fill_u64(b->bmin, 0x7ff0000000000000ull, b->total_n, st);
  launch_fill_i32(b->sfail, 0, b->total_n, st);
  launch_fill_i32(b->counts, 0, 2 * J, st);
  if (mixed) {
    launch_fill_i32(b->rcounts, 0, J, st);
    launch_fill_i32(b->fcounts, 0, J, st);
  }
  launch_info_init(d_in, b->info64, b->total_info, st);
  batch_mark(b, "fill", st);
  const ProblemDev* const* probs = (const ProblemDev* const*)b->ptrs.p;
  const StageBufs* bufs = (const StageBufs*)b->bufs.p;
  // mixed: the nominal-loop eigensolve, the located sweep and the refinement run on the side stream, beside
  // the simulation and the fit (per-launch timing keeps everything on one stream)
  const bool fork = mixed && !b->profiling;
  if (fork && !b->side) {
    BCHK(hipStreamCreateWithFlags(&b->side, hipStreamNonBlocking));
    BCHK(hipEventCreateWithFlags(&b->ev_fork, hipEventDisableTiming));
    BCHK(hipEventCreateWithFlags(&b->ev_join, hipEventDisableTiming));
  }
  bool forked = false, joined = false;
  for (size_t i = 0; i < b->launches.size(); i++) {
    Launch& L = b->launches[i];
    const int total = L.start.back();
    const bool on_side = fork && (L.kind == BK_EIG_NOM || L.kind == BK_SWEEP_MARK32 || L.kind == BK_REFINE64 ||
                                  L.kind == BK_FULL64);
    if (fork && !forked && L.kind != BK_EIG_LIN && L.kind != BK_COMPACT) {
      BCHK(hipEventRecord(b->ev_fork, st));
      BCHK(hipStreamWaitEvent(b->side, b->ev_fork, 0));
      forked = true;
    }
    if (fork && !joined && L.kind == BK_SCAN32) {
      BCHK(hipEventRecord(b->ev_join, b->side));
      BCHK(hipStreamWaitEvent(st, b->ev_join, 0));
      joined = true;
    }
    if (total == 0) continue;
    hipStream_t ls = on_side ? b->side : st;
    int grid = grids ? (*grids)[i] : L.grid;
    grid = std::max(1, std::min(grid, total));
    if (L.kind == SIM_LAUNCH) {
      void* f = b->sim_lanes == 4 ? L.fl[1] : (b->sim_lanes == 2 ? L.fl[0] : L.fn);
      if (!f) throw std::runtime_error("no " + std::to_string(b->sim_lanes) + "-lane simulation kernel");
      const bool ok = L.d_sid ? jit_launch_sim_batch_fused(f, grid, ls, L.d_start, L.d_pid, (const SimArgs*)b->sargs.p,
                                                           (const char* const*)b->ptrs.p, L.d_sid, (int)L.pid.size(), total)
                              : jit_launch_sim_batch(f, grid, ls, L.d_start, L.d_pid, (const SimArgs*)b->sargs.p,
                                                     (const char* const*)b->ptrs.p, (int)L.pid.size(), total);
      if (!ok) throw std::runtime_error("launch of the batched simulation kernel failed");
    } else {
      BatchTab T{L.d_start, L.d_pid, probs, bufs, (int32_t)L.pid.size(), total};
      const bool ok = (L.kind >= BK_ATB) ? launch_batch_mixed(L.kind, L.key, T, grid, ls)
                                          : launch_batch_fp64(L.kind, L.key, T, grid, ls);
      if (!ok) throw std::runtime_error("no batched kernel for " + L.label);
    }
    batch_mark(b, L.label, st);
  }
  if (fork && !joined) {
    BCHK(hipEventRecord(b->ev_join, b->side));
    BCHK(hipStreamWaitEvent(st, b->ev_join, 0));
  }
  launch_info_final(b->info64, d_out, b->total_info, st);
  if (d_out64) BCHK(hipMemcpyAsync(d_out64, b->info64, b->total_info * sizeof(double), hipMemcpyDeviceToDevice, st));
  batch_mark(b, "final", st);
  BCHK(hipGetLastError());
}

// Whole-chain passes, one stream, every launch timed by events.  For each stage plan (tile: the
// eigen stages' lanes per candidate, one launch per capacity or one fused launch; the simulation
// per structure or fused) the chain runs with each grid cap of a persistent launch (at most 8, 16,
// 32 workgroups per CU, or one per block: the waves each launch keeps in flight); every launch keeps
// its fastest grid, and every stage the plan whose launches add up to the least time.
static void autotune(ce_batch* b, int precision, const float* d_gains, const float* d_in, float* d_out, hipStream_t st) {
  int nsm = 256;
  hipDeviceGetAttribute(&nsm, hipDeviceAttributeMultiprocessorCount, b->device);
  const int caps[] = {0, 8 * nsm, 16 * nsm, 32 * nsm};  // 0: one workgroup per block
  const int saved = b->profiling;
  b->profiling = 1;
  std::map<std::string, std::pair<double, int>> best;  // launch label -> (time, grid)
  const int nplans = b->plan_fixed ? 1 : 3;
  double stage_t[3][3];
  for (auto& r : stage_t) r[0] = r[1] = r[2] = 1e300;
  for (int pl = 0; pl < nplans; pl++) {
    if (!b->plan_fixed) {
      b->plan[0] = b->plan[1] = pl;
      b->plan[2] = pl == 2 ? 1 : 0;
      build_layout(b, precision, st);
      set_gains(b, d_gains, st);
    }
    const size_t NL = b->launches.size();
    std::vector<double> tb(NL, 1e300);
    std::vector<int> gb(NL, 0);
    for (int c = 0; c < 4; c++) {
      std::vector<int> grids(NL);
      for (size_t i = 0; i < NL; i++) {
        const int total = b->launches[i].start.back();
        grids[i] = caps[c] ? std::min(total, caps[c]) : total;
      }
      for (int rep = 0; rep < 2; rep++) {  // first pass warms caches and code
        enqueue(b, precision, d_in, d_out, nullptr, st, &grids);
        BCHK(hipStreamSynchronize(st));
      }
      // marks: start, fill, then one per non-empty launch
      int m = 2;
      for (size_t i = 0; i < NL; i++) {
        if (b->launches[i].start.back() == 0) continue;
        float t = 0.0f;
        BCHK(hipEventElapsedTime(&t, b->mark_ev[m - 1], b->mark_ev[m]));
        m++;
        if (t < tb[i]) {
          tb[i] = t;
          gb[i] = grids[i];
        }
      }
    }
    double st_sum[3] = {0.0, 0.0, 0.0};
    for (size_t i = 0; i < NL; i++) {
      const Launch& L = b->launches[i];
      const double t = L.start.back() == 0 ? 0.0 : tb[i];
      auto it = best.find(L.label);
      if (it == best.end() || t < it->second.first) best[L.label] = {t, std::max(1, gb[i])};
      if (L.kind == BK_EIG_LIN) st_sum[0] += t;
      else if (L.kind == BK_EIG_NOM) st_sum[1] += t;
      else if (L.kind == SIM_LAUNCH) st_sum[2] += t;
    }
    stage_t[0][pl] = st_sum[0];
    stage_t[1][pl] = st_sum[1];
    stage_t[2][pl] = st_sum[2];
  }
  if (!b->plan_fixed) {
    for (int k = 0; k < 2; k++) {
      int bp = 0;
      for (int pl = 1; pl < 3; pl++)
        if (stage_t[k][pl] < stage_t[k][bp]) bp = pl;
      b->plan[k] = bp;
    }
    b->plan[2] = stage_t[2][2] < std::min(stage_t[2][0], stage_t[2][1]) ? 1 : 0;
    build_layout(b, precision, st);
    set_gains(b, d_gains, st);
  }
  for (auto& L : b->launches) {
    auto it = best.find(L.label);
    if (it != best.end()) L.grid = it->second.second;
  }
  b->profiling = saved;
  b->tuned = true;
}

static void run(ce_batch* b, int precision, const float* d_gains, const float* d_in, float* d_out, double* d_out64,
                hipStream_t st) {
  if (b->P.empty()) return;
  if (precision != CE_PRECISION_FP64 && precision != CE_PRECISION_MIXED)
    throw std::runtime_error("ce_batch_run: unknown precision " + std::to_string(precision));
  BCHK(hipSetDevice(b->device));
  if (!jit_enabled()) throw std::runtime_error("batched evaluation needs the runtime-specialised simulation (CE_JIT=0)");
  if (b->layout_precision != precision)
    for (auto& p : b->P) p.need_build = true;  // structure keys depend on the precision
  build_blobs(b, precision);
  if (precision == CE_PRECISION_MIXED)
    for (auto& p : b->P)
      if (p.F > 512) throw std::runtime_error("mixed batches support at most 512 frequencies");
  if (b->layout_dirty || b->layout_precision != precision) build_layout(b, precision, st);
  upload_blobs(b, st);
  set_gains(b, d_gains, st);
  if (b->autotune && !b->tuned) {
    autotune(b, precision, d_gains, d_in, d_out, st);
    drop_graph(b);
  }
  b->last_used_graph = 0;
  if (b->use_graph && !b->profiling) {
    if (!b->own) {
      BCHK(hipStreamCreateWithFlags(&b->own, hipStreamNonBlocking));
      BCHK(hipEventCreateWithFlags(&b->ev_in, hipEventDisableTiming));
      BCHK(hipEventCreateWithFlags(&b->ev_out, hipEventDisableTiming));
    }
    hipStream_t gs = b->own;
    if (!b->gexec || b->g_in != d_in || b->g_out != d_out || b->g_out64 != d_out64) {
      drop_graph(b);
      BCHK(hipStreamSynchronize(st));
      BCHK(hipStreamBeginCapture(gs, hipStreamCaptureModeThreadLocal));
      try {
        enqueue(b, precision, d_in, d_out, d_out64, gs, nullptr);
      } catch (...) {
        hipGraph_t g = nullptr;
        hipStreamEndCapture(gs, &g);
        if (g) hipGraphDestroy(g);
        throw;
      }
      BCHK(hipStreamEndCapture(gs, &b->graph));
      BCHK(hipGraphInstantiate(&b->gexec, b->graph, nullptr, nullptr, 0));
      b->g_in = d_in;
      b->g_out = d_out;
      b->g_out64 = d_out64;
    }
    BCHK(hipEventRecord(b->ev_in, st));
    BCHK(hipStreamWaitEvent(gs, b->ev_in, 0));
    BCHK(hipGraphLaunch(b->gexec, gs));
    BCHK(hipEventRecord(b->ev_out, gs));
    BCHK(hipStreamWaitEvent(st, b->ev_out, 0));
    b->last_used_graph = 1;
  } else {
    enqueue(b, precision, d_in, d_out, d_out64, st, nullptr);
  }
}

// =================================================================================================
extern "C" CE_EXPORT ce_batch* ce_batch_create(int device, char* err, int errlen) {
  try {
    int count = 0;
    BCHK(hipGetDeviceCount(&count));
    if (device < 0 || device >= count) throw std::runtime_error("device index out of range");
    BCHK(hipSetDevice(device));
    ce_batch* b = new ce_batch();
    b->device = device;
    return b;
  } catch (const std::exception& e) {
    put_err_b(err, errlen, e.what());
    return nullptr;
  }
}

extern "C" CE_EXPORT void ce_batch_destroy(ce_batch* b) {
  if (!b) return;
  hipSetDevice(b->device);
  hipDeviceSynchronize();
  drop_graph(b);
  for (DevBuf* d : {&b->blobs, &b->ptrs, &b->bufs, &b->tabs, &b->sargs, &b->work}) d->release();
  if (b->pinned) hipHostFree(b->pinned);
  for (hipEvent_t e : b->mark_ev) hipEventDestroy(e);
  if (b->ev_in) hipEventDestroy(b->ev_in);
  if (b->ev_out) hipEventDestroy(b->ev_out);
  if (b->own) hipStreamDestroy(b->own);
  if (b->ev_fork) hipEventDestroy(b->ev_fork);
  if (b->ev_join) hipEventDestroy(b->ev_join);
  if (b->side) hipStreamDestroy(b->side);
  delete b;
}

extern "C" CE_EXPORT int ce_batch_clear(ce_batch* b) {
  if (!b) return 1;
  b->P.clear();
  b->layout_dirty = true;
  return 0;
}

static bool shape_of(const ce_config* cfg, ce_batch::Prob& p, std::string& e) {
  ce_derived d;
  if (!derive_config(*cfg, d, e)) return false;
  p.cw = d.cuda_mx_width;
  p.iw = d.info_mx_width;
  return true;
}

extern "C" CE_EXPORT int ce_batch_add(ce_batch* b, const ce_config* cfg, const ce_plant* plant, int n, char* err,
                                      int errlen) {
  if (!b || !cfg || !plant || n <= 0) {
    put_err_b(err, errlen, "ce_batch_add: invalid arguments");
    return -1;
  }
  ce_batch::Prob p;
  p.cfg = *cfg;
  p.plant = *plant;
  p.n = n;
  std::string e;
  if (!shape_of(cfg, p, e)) {
    put_err_b(err, errlen, e);
    return -1;
  }
  b->P.push_back(std::move(p));
  b->layout_dirty = true;
  return (int)b->P.size() - 1;
}

// New configuration / plant values for problem `index` (same candidate count).  The descriptor
// is rebuilt at the next run; the device layout and the captured graph are kept when the shape
// and the simulation structure are unchanged.
extern "C" CE_EXPORT int ce_batch_update(ce_batch* b, int index, const ce_config* cfg, const ce_plant* plant,
                                         char* err, int errlen) {
  if (!b || index < 0 || index >= (int)b->P.size() || !cfg || !plant) {
    put_err_b(err, errlen, "ce_batch_update: invalid arguments");
    return 1;
  }
  auto& p = b->P[index];
  std::string e;
  ce_batch::Prob q;
  if (!shape_of(cfg, q, e)) {
    put_err_b(err, errlen, e);
    return 1;
  }
  if (q.cw != p.cw || q.iw != p.iw) {
    put_err_b(err, errlen, "ce_batch_update: the row widths of a problem cannot change");
    return 1;
  }
  p.cfg = *cfg;
  p.plant = *plant;
  p.need_build = true;
  return 0;
}

extern "C" CE_EXPORT int ce_batch_offsets(ce_batch* b, int index, long long* gains_off, long long* info_off) {
  if (!b || index < 0 || index >= (int)b->P.size()) return 1;
  size_t go = 0, io = 0;
  for (int j = 0; j < index; j++) {
    go += (size_t)b->P[j].n * b->P[j].cw;
    io += (size_t)b->P[j].n * b->P[j].iw;
  }
  if (gains_off) *gains_off = (long long)go;
  if (info_off) *info_off = (long long)io;
  return 0;
}

extern "C" CE_EXPORT int ce_batch_set_options(ce_batch* b, int use_graph, int autotune, int profiling) {
  if (!b) return 1;
  if (!use_graph) drop_graph(b);
  b->use_graph = use_graph ? 1 : 0;
  if (autotune && !b->autotune) b->tuned = false;
  b->autotune = autotune ? 1 : 0;
  b->profiling = profiling ? 1 : 0;
  return 0;
}

extern "C" CE_EXPORT int ce_batch_run(ce_batch* b, int precision, const float* d_gains, const float* d_info_in,
                                      float* d_info_out, double* d_info64_out, void* stream, char* err, int errlen) {
  if (!b) return 1;
  try {
    run(b, precision, d_gains, d_info_in, d_info_out, d_info64_out, (hipStream_t)stream);
    return 0;
  } catch (const std::exception& e) {
    put_err_b(err, errlen, e.what());
    return 1;
  }
}

extern "C" CE_EXPORT int ce_batch_set_plan(ce_batch* b, int eig_lin, int eig_nom, int sim) {
  if (!b) return 1;
  if (eig_lin < 0 || eig_nom < 0 || sim < 0) {
    b->plan_fixed = false;
    b->tuned = false;
    return 0;
  }
  if (eig_lin > 2 || eig_nom > 2 || sim > 1) return 1;
  b->plan[0] = eig_lin;
  b->plan[1] = eig_nom;
  b->plan[2] = sim;
  b->plan_fixed = true;
  b->layout_dirty = true;
  b->tuned = false;
  return 0;
}

extern "C" CE_EXPORT int ce_batch_plan(ce_batch* b, int* plans) {
  if (!b || !plans) return 1;
  for (int k = 0; k < 3; k++) plans[k] = b->plan[k];
  return 0;
}

extern "C" CE_EXPORT int ce_batch_info(ce_batch* b, int* launches, int* used_graph, int* tuned) {
  if (!b) return 1;
  if (launches) *launches = (int)b->launches.size();
  if (used_graph) *used_graph = b->last_used_graph;
  if (tuned) *tuned = b->tuned ? 1 : 0;
  return 0;
}

extern "C" CE_EXPORT int ce_batch_kernel_times(ce_batch* b, char* names, int names_len, double* ms, int n_max) {
  if (!b || b->n_marks < 2) return 0;
  if (hipEventSynchronize(b->mark_ev[b->n_marks - 1]) != hipSuccess) return -1;
  std::string list;
  int k = 0;
  for (int i = 1; i < b->n_marks && k < n_max; i++, k++) {
    float t = 0.0f;
    if (hipEventElapsedTime(&t, b->mark_ev[i - 1], b->mark_ev[i]) != hipSuccess) return -1;
    ms[k] = t;
    if (!list.empty()) list += ",";
    list += b->mark_name[i];
  }
  if (names && names_len > 0) {
    std::strncpy(names, list.c_str(), (size_t)names_len - 1);
    names[names_len - 1] = 0;
  }
  return k;
}

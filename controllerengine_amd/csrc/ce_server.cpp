// Serving loop: tuning jobs (one problem + its candidate block each) submitted from any thread are
// admitted by their device-memory footprint and evaluated in batches formed continuously.
//
// A scheduler thread owns the device work.  It keeps up to `slots` batches in flight, each on its
// own stream: whenever a slot is free and queued jobs fit in the memory budget left by the jobs in
// flight, it forms a batch from the queue (FIFO, jobs of any dimensions, horizon and frequency
// count), copies the inputs in, runs the batched chain and copies the outputs back, all enqueued
// on the slot's stream.  Jobs that arrive while other batches are on the GPU are admitted as soon
// as a slot and memory free up, so their work overlaps on the device.  Each job's rows depend only
// on its own inputs, so a job's result is the same alone and in any mix.  Per-job latency
// (submission to completion observed by the scheduler) feeds p50/p99 statistics.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "../../include/ControllerEngine.h"
#include "ce_host.h"
#include "ce_stage.h"

namespace {

using Clock = std::chrono::steady_clock;

enum JobState { QUEUED = 0, RUNNING = 1, DONE = 2, FAILED = 3 };

struct Job {
  long long id = 0;
  ce_config cfg;
  ce_plant plant;
  int n = 0, cw = 0, iw = 0;
  std::vector<float> gains, info, out;
  std::vector<double> out64;
  size_t footprint = 0;
  JobState state = QUEUED;
  std::string error;
  Clock::time_point t_submit, t_done;
};

struct HostPinned {
  void* p = nullptr;
  size_t cap = 0;
  void* get(size_t bytes) {
    if (cap < bytes) {
      if (p) hipHostFree(p);
      p = nullptr;
      if (hipHostMalloc(&p, bytes, hipHostMallocDefault) != hipSuccess) throw std::runtime_error("hipHostMalloc failed");
      cap = bytes;
    }
    return p;
  }
  ~HostPinned() {
    if (p) hipHostFree(p);
  }
};

struct DevMem {
  void* p = nullptr;
  size_t cap = 0;
  void* get(size_t bytes) {
    if (bytes == 0) bytes = 256;
    if (cap < bytes) {
      if (p) hipFree(p);
      p = nullptr;
      if (hipMalloc(&p, bytes) != hipSuccess) throw std::runtime_error("hipMalloc failed for a serving batch");
      cap = bytes;
    }
    return p;
  }
  ~DevMem() {
    if (p) hipFree(p);
  }
};

struct Slot {
  ce_batch* batch = nullptr;
  hipStream_t st = nullptr;
  hipEvent_t done = nullptr;
  std::vector<std::shared_ptr<Job>> jobs;
  size_t bytes = 0;
  bool busy = false;
  HostPinned h_in_g, h_in_i, h_out, h_out64;
  DevMem d_g, d_i, d_o, d_o64;
  size_t n_g = 0, n_i = 0;
};

// Device bytes one job holds while in flight: its inputs / outputs and the batch buffers sized by
// its candidates (trajectory window, sweep state, refinement list, modes, INFO rows in double).
size_t footprint(const ce_config& c, int n, int precision) {
  ce_derived d;
  std::string e;
  if (!ce::derive_config(c, d, e)) return 0;
  const bool mixed = precision == CE_PRECISION_MIXED;
  const size_t nt = (size_t)d.total_dim, F = (size_t)c.freq, fc = (size_t)d.fit_count;
  size_t per = 0;
  per += (size_t)d.cuda_mx_width * 4 + (size_t)d.info_mx_width * (4 + 4 + 8 + 8);  // in, out, out64, info64
  per += 4 * 5 + 2 * ce::MAXT * 8 + (size_t)c.statespace_dim * c.statespace_dim * 8;
  per += fc * (mixed ? 4 : 8) + F * (mixed ? 3 : 2) * 8;
  if (mixed) per += (F + 63) / 64 * 64 * 4 * 4 + F + ce::RMAX * 8 + 4 + 16 * nt * 16;
  return per * (size_t)n + 64 * 1024;
}

}  // namespace

struct ce_server {
  int device = 0;
  int precision = CE_PRECISION_MIXED;
  size_t budget = 0;
  int max_jobs = 64;
  std::vector<std::unique_ptr<Slot>> slots;
  std::mutex mu;
  std::condition_variable cv;
  std::deque<std::shared_ptr<Job>> queue;
  std::map<long long, std::shared_ptr<Job>> jobs;
  long long next_id = 0;
  size_t inflight_bytes = 0, peak_bytes = 0;
  int inflight_jobs = 0, peak_jobs = 0, peak_batches = 0;
  long long batches = 0;
  std::vector<double> lat_ms;
  double cand_done = 0.0;
  Clock::time_point t_first, t_last;
  bool any_done = false, any_submit = false;
  std::atomic<bool> stop{false};
  bool wake = false;        // a slot's work completed (set by the stream callback, under mu)
  long long submits = 0;    // jobs submitted so far (under mu)
  std::thread thr;
  std::string fatal;
};

// Runs on a HIP runtime thread once a slot's batch, output copies and done event have completed:
// wakes the scheduler (no HIP call here).  The scheduler never holds mu while it waits on a
// stream, so this lock cannot deadlock.
static void slot_done_cb(void* arg) {
  ce_server* S = static_cast<ce_server*>(arg);
  {
    std::lock_guard<std::mutex> g(S->mu);
    S->wake = true;
  }
  S->cv.notify_all();
}

static void put_err_s(char* err, int errlen, const std::string& m) {
  if (err && errlen > 0) std::snprintf(err, (size_t)errlen, "%s", m.c_str());
}

// Enqueues one batch of jobs on a slot (scheduler thread, without the lock held).
static void launch_slot(ce_server* S, Slot& s) {
  char err[1024];
  ce_batch_clear(s.batch);
  size_t ng = 0, ni = 0;
  for (auto& j : s.jobs) {
    if (ce_batch_add(s.batch, &j->cfg, &j->plant, j->n, err, sizeof err) < 0) throw std::runtime_error(err);
    ng += j->gains.size();
    ni += j->info.size();
  }
  float* hg = (float*)s.h_in_g.get(ng * 4);
  float* hi = (float*)s.h_in_i.get(ni * 4);
  size_t og = 0, oi = 0;
  for (auto& j : s.jobs) {
    std::memcpy(hg + og, j->gains.data(), j->gains.size() * 4);
    std::memcpy(hi + oi, j->info.data(), j->info.size() * 4);
    og += j->gains.size();
    oi += j->info.size();
  }
  s.n_g = ng;
  s.n_i = ni;
  float* dg = (float*)s.d_g.get(ng * 4);
  float* di = (float*)s.d_i.get(ni * 4);
  float* dout = (float*)s.d_o.get(ni * 4);
  double* dout64 = (double*)s.d_o64.get(ni * 8);
  if (hipMemcpyAsync(dg, hg, ng * 4, hipMemcpyHostToDevice, s.st) != hipSuccess ||
      hipMemcpyAsync(di, hi, ni * 4, hipMemcpyHostToDevice, s.st) != hipSuccess)
    throw std::runtime_error("input copy failed");
  if (ce_batch_run(s.batch, S->precision, dg, di, dout, dout64, s.st, err, sizeof err) != 0)
    throw std::runtime_error(err);
  float* ho = (float*)s.h_out.get(ni * 4);
  double* ho64 = (double*)s.h_out64.get(ni * 8);
  if (hipMemcpyAsync(ho, dout, ni * 4, hipMemcpyDeviceToHost, s.st) != hipSuccess ||
      hipMemcpyAsync(ho64, dout64, ni * 8, hipMemcpyDeviceToHost, s.st) != hipSuccess)
    throw std::runtime_error("output copy failed");
  if (hipEventRecord(s.done, s.st) != hipSuccess) throw std::runtime_error("event record failed");
  if (hipLaunchHostFunc(s.st, slot_done_cb, S) != hipSuccess) throw std::runtime_error("host callback failed");
}

static void finish_slot(ce_server* S, Slot& s) {
  // called with the lock held, after the slot's event completed
  const float* ho = (const float*)s.h_out.p;
  const double* ho64 = (const double*)s.h_out64.p;
  size_t oi = 0;
  const auto now = Clock::now();
  for (auto& j : s.jobs) {
    j->out.assign(ho + oi, ho + oi + j->info.size());
    j->out64.assign(ho64 + oi, ho64 + oi + j->info.size());
    oi += j->info.size();
    j->state = DONE;
    j->t_done = now;
    S->lat_ms.push_back(std::chrono::duration<double, std::milli>(now - j->t_submit).count());
    S->cand_done += j->n;
    j->gains.clear();
    j->gains.shrink_to_fit();
    j->info.clear();
    j->info.shrink_to_fit();
  }
  S->t_last = now;
  S->any_done = true;
  S->inflight_bytes -= s.bytes;
  S->inflight_jobs -= (int)s.jobs.size();
  s.jobs.clear();
  s.bytes = 0;
  s.busy = false;
}

static void scheduler(ce_server* S) {
  hipSetDevice(S->device);
  std::unique_lock<std::mutex> lk(S->mu);
  while (!S->stop) {
    bool progressed = false;
    S->wake = false;
    const long long seen = S->submits;
    for (auto& sp : S->slots) {
      Slot& s = *sp;
      if (s.busy && hipEventQuery(s.done) == hipSuccess) {
        finish_slot(S, s);
        progressed = true;
      }
    }
    for (auto& sp : S->slots) {
      Slot& s = *sp;
      if (s.busy || S->queue.empty()) continue;
      // admit queued jobs in arrival order while they fit the memory left
      std::vector<std::shared_ptr<Job>> pick;
      size_t bytes = 0;
      while (!S->queue.empty() && (int)pick.size() < S->max_jobs) {
        auto& j = S->queue.front();
        const bool nothing_in_flight = S->inflight_bytes == 0 && pick.empty();
        if (S->inflight_bytes + bytes + j->footprint > S->budget && !nothing_in_flight) break;
        bytes += j->footprint;
        pick.push_back(j);
        S->queue.pop_front();
      }
      if (pick.empty()) break;
      for (auto& j : pick) j->state = RUNNING;
      s.jobs = pick;
      s.bytes = bytes;
      s.busy = true;
      S->inflight_bytes += bytes;
      S->inflight_jobs += (int)pick.size();
      S->batches++;
      S->peak_bytes = std::max(S->peak_bytes, S->inflight_bytes);
      S->peak_jobs = std::max(S->peak_jobs, S->inflight_jobs);
      int busy = 0;
      for (auto& q : S->slots) busy += q->busy ? 1 : 0;
      S->peak_batches = std::max(S->peak_batches, busy);
      lk.unlock();
      std::string fail;
      try {
        launch_slot(S, s);
      } catch (const std::exception& e) {
        fail = e.what();
      }
      if (!fail.empty()) hipStreamSynchronize(s.st);  // without the lock: a pending callback takes it
      lk.lock();
      if (!fail.empty()) {
        for (auto& j : s.jobs) {
          j->state = FAILED;
          j->error = fail;
        }
        S->inflight_bytes -= s.bytes;
        S->inflight_jobs -= (int)s.jobs.size();
        s.jobs.clear();
        s.bytes = 0;
        s.busy = false;
      }
      progressed = true;
    }
    // sleep until a submission, a completion callback or stop (the timeout only bounds a missed wake)
    if (!progressed)
      S->cv.wait_for(lk, std::chrono::milliseconds(5), [S, seen] { return S->stop || S->wake || S->submits != seen; });
  }
}

extern "C" CE_EXPORT ce_server* ce_server_create(int device, int precision, long long mem_budget_bytes, int slots,
                                                 int max_jobs_per_batch, char* err, int errlen) {
  try {
    if (precision != CE_PRECISION_FP64 && precision != CE_PRECISION_MIXED) throw std::runtime_error("unknown precision");
    if (hipSetDevice(device) != hipSuccess) throw std::runtime_error("device index out of range");
    auto* S = new ce_server();
    S->device = device;
    S->precision = precision;
    size_t freeb = 0, total = 0;
    hipMemGetInfo(&freeb, &total);
    S->budget = mem_budget_bytes > 0 ? (size_t)mem_budget_bytes : freeb / 2;
    S->max_jobs = max_jobs_per_batch > 0 ? max_jobs_per_batch : 64;
    const int ns = std::max(1, std::min(slots > 0 ? slots : 2, 8));
    for (int i = 0; i < ns; i++) {
      auto s = std::make_unique<Slot>();
      char e[512];
      s->batch = ce_batch_create(device, e, sizeof e);
      if (!s->batch) throw std::runtime_error(e);
      ce_batch_set_options(s->batch, 0, 0, 0);  // compositions change every batch: eager launches
      if (hipStreamCreateWithFlags(&s->st, hipStreamNonBlocking) != hipSuccess ||
          hipEventCreateWithFlags(&s->done, hipEventDisableTiming) != hipSuccess)
        throw std::runtime_error("stream / event creation failed");
      S->slots.push_back(std::move(s));
    }
    S->thr = std::thread(scheduler, S);
    return S;
  } catch (const std::exception& e) {
    put_err_s(err, errlen, e.what());
    return nullptr;
  }
}

extern "C" CE_EXPORT void ce_server_destroy(ce_server* S) {
  if (!S) return;
  S->stop = true;
  S->cv.notify_all();
  if (S->thr.joinable()) S->thr.join();
  for (auto& s : S->slots) {
    hipStreamSynchronize(s->st);
    ce_batch_destroy(s->batch);
    hipEventDestroy(s->done);
    hipStreamDestroy(s->st);
  }
  S->slots.clear();
  delete S;
}

extern "C" CE_EXPORT long long ce_server_submit(ce_server* S, const ce_config* cfg, const ce_plant* plant,
                                                const float* h_gains, const float* h_info, int n, char* err, int errlen) {
  if (!S || !cfg || !plant || !h_gains || !h_info || n <= 0) {
    put_err_s(err, errlen, "ce_server_submit: invalid arguments");
    return -1;
  }
  auto j = std::make_shared<Job>();
  ce_derived d;
  std::string e;
  if (!ce::derive_config(*cfg, d, e)) {
    put_err_s(err, errlen, e);
    return -1;
  }
  j->cfg = *cfg;
  j->plant = *plant;
  j->n = n;
  j->cw = d.cuda_mx_width;
  j->iw = d.info_mx_width;
  j->gains.assign(h_gains, h_gains + (size_t)n * j->cw);
  j->info.assign(h_info, h_info + (size_t)n * j->iw);
  j->footprint = footprint(*cfg, n, S->precision);
  j->t_submit = Clock::now();
  {
    std::lock_guard<std::mutex> lk(S->mu);
    j->id = S->next_id++;
    if (!S->any_submit) {
      S->t_first = j->t_submit;
      S->any_submit = true;
    }
    S->jobs[j->id] = j;
    S->queue.push_back(j);
    S->submits++;
  }
  S->cv.notify_all();
  return j->id;
}

// 0: queued or running; 1: done (rows copied when the pointers are not NULL; the job is then
// forgotten); -1: unknown job; -2: failed (message in err).
extern "C" CE_EXPORT int ce_server_poll(ce_server* S, long long job, float* h_info_out, double* h_info64_out,
                                        char* err, int errlen) {
  if (!S) return -1;
  std::lock_guard<std::mutex> lk(S->mu);
  auto it = S->jobs.find(job);
  if (it == S->jobs.end()) return -1;
  Job& j = *it->second;
  if (j.state == FAILED) {
    put_err_s(err, errlen, j.error);
    S->jobs.erase(it);
    return -2;
  }
  if (j.state != DONE) return 0;
  if (h_info_out) std::memcpy(h_info_out, j.out.data(), j.out.size() * 4);
  if (h_info64_out) std::memcpy(h_info64_out, j.out64.data(), j.out64.size() * 8);
  S->jobs.erase(it);
  return 1;
}

// Blocks until no job is queued or running (or timeout_s passes); 0 when idle.
extern "C" CE_EXPORT int ce_server_drain(ce_server* S, double timeout_s) {
  if (!S) return 1;
  const auto t0 = Clock::now();
  while (true) {
    {
      std::lock_guard<std::mutex> lk(S->mu);
      bool busy = !S->queue.empty();
      for (auto& s : S->slots) busy = busy || s->busy;
      if (!busy) return 0;
    }
    if (std::chrono::duration<double>(Clock::now() - t0).count() > timeout_s) return 1;
    std::this_thread::sleep_for(std::chrono::microseconds(200));
  }
}

// out: jobs done, latency p50 / p99 / mean / max (ms), candidate evals/s from the first submission
// to the last completion, batches launched, peak jobs in flight, peak batches in flight, peak bytes
// admitted, memory budget.  Returns the number of values written.
extern "C" CE_EXPORT int ce_server_stats(ce_server* S, double* out, int n_max) {
  if (!S || !out) return 0;
  std::lock_guard<std::mutex> lk(S->mu);
  std::vector<double> l = S->lat_ms;
  std::sort(l.begin(), l.end());
  auto pct = [&](double q) {
    if (l.empty()) return 0.0;
    const size_t k = (size_t)std::min<double>((double)l.size() - 1, std::ceil(q * l.size()) - 1);
    return l[k];
  };
  double mean = 0.0;
  for (double v : l) mean += v;
  if (!l.empty()) mean /= (double)l.size();
  const double span = S->any_done ? std::chrono::duration<double>(S->t_last - S->t_first).count() : 0.0;
  const double vals[] = {(double)l.size(), pct(0.50), pct(0.99), mean, l.empty() ? 0.0 : l.back(),
                         span > 0 ? S->cand_done / span : 0.0, (double)S->batches, (double)S->peak_jobs,
                         (double)S->peak_batches, (double)S->peak_bytes, (double)S->budget};
  const int k = std::min<int>(n_max, (int)(sizeof vals / sizeof vals[0]));
  for (int i = 0; i < k; i++) out[i] = vals[i];
  return k;
}

// Forgets the latency record (for measuring one arrival rate after another).
extern "C" CE_EXPORT int ce_server_reset_stats(ce_server* S) {
  if (!S) return 1;
  std::lock_guard<std::mutex> lk(S->mu);
  S->lat_ms.clear();
  S->cand_done = 0.0;
  S->batches = 0;
  S->peak_jobs = S->peak_batches = 0;
  S->peak_bytes = 0;
  S->any_done = S->any_submit = false;
  return 0;
}

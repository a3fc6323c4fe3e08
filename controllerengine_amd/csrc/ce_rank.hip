// Ranking and selection of a candidate block's INFO rows.
//
// The reference ranks nothing itself: its caller reads INFO_MX_OUT.bin back and selects candidates
// by the score column against the requirements of README.md:22-35.  ce_rank does that selection on
// the device, after ce_evaluate, on the float rows it wrote:
//
//   order  candidate indices by ascending score (INFO column info_scoring), ties by index (-0 and +0
//          tie), scores that are not a number after every other;
//   pass   1 when the candidate meets every requirement the rows can show: linearly and nominally
//          stable with a completed sweep (stable cell 1, score below unstable_err, margin cell not a
//          penalty), margin error 0 (GM, PM and b at or above their targets), overshoot at most its
//          spec and |yss / ref - 1| at most the steady-state spec.  Damping and bandwidth enter the
//          score as squared deviations and have no pass band.
//
// The order is a stable LSD radix sort (hipCUB) of 32-bit keys that map the float order onto the
// unsigned order; the pass rule is evaluated in double from the float cells.
#include <hip/hip_runtime.h>

#include <hipcub/hipcub.hpp>

#include <cstdio>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "ce_host.h"

namespace ce {

struct RankRule {
  int32_t info_w, isc, ifr, iam;
  double ue, os, sse, amp;
};

// float score -> key: ascending floats map to ascending keys, -0 to the key of +0, NaN to the
// largest key
__device__ __forceinline__ uint32_t score_key(float s) {
  if (s != s) return 0xFFFFFFFFu;
  if (s == 0.0f) s = 0.0f;
  const uint32_t b = __float_as_uint(s);
  return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}

__global__ __launch_bounds__(256) void k_rank_prep(const float* __restrict__ info, int n, RankRule R,
                                                   uint32_t* __restrict__ keys, int32_t* __restrict__ vals,
                                                   uint8_t* __restrict__ pass, int32_t* __restrict__ npass) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  bool ok = false;
  if (i < n) {
    const float* row = info + (size_t)i * R.info_w;
    keys[i] = score_key(row[R.isc]);
    vals[i] = i;
    const double sc0 = row[R.isc], margin = row[R.isc + 2], stable = row[R.iam + 3];
    const double os = row[R.ifr + 2], yss = row[R.ifr + 3];
    ok = sc0 < R.ue && stable == 1.0 && margin == 0.0 && os <= R.os && __builtin_fabs(yss / R.amp - 1.0) <= R.sse;
    if (pass) pass[i] = ok ? 1 : 0;
  }
  if (npass) {
    const unsigned long long b = __ballot(ok);
    if ((threadIdx.x & 63) == 0 && b) atomicAdd(npass, (int32_t)__popcll(b));
  }
}

__global__ __launch_bounds__(256) void k_rank_out(const int32_t* __restrict__ vals, int k, int32_t* __restrict__ order) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < k) order[i] = vals[i];
}

__global__ void k_rank_zero(int32_t* p) { *p = 0; }

// Bytes of the sort's temporary storage for n items (no allocation, no launch).
size_t rank_temp_bytes(int n) {
  hipcub::DoubleBuffer<uint32_t> kb(nullptr, nullptr);
  hipcub::DoubleBuffer<int32_t> vb(nullptr, nullptr);
  size_t bytes = 0;
  if (hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, kb, vb, n, 0, 32, (hipStream_t)0) != hipSuccess)
    throw std::runtime_error("ce_rank: sort storage query failed");
  return bytes;
}

RankRule rank_rule(const ce_config& cfg) {
  ce_derived d;
  std::string e;
  if (!derive_config(cfg, d, e)) throw std::runtime_error(e);
  return RankRule{d.info_mx_width, d.info_scoring, d.info_freqresp, d.info_allmargin, (double)cfg.unstable_err,
                  (double)cfg.overshoot, (double)cfg.steady_state_error, (double)cfg.ref_amplitude};
}

// Enqueues the ranking on st with the caller's buffers: keys and vals hold 2 n entries each
// (the sort's double buffers), temp temp_bytes.
void rank_enqueue(const RankRule& R, const float* info, int n, int k, int32_t* order, uint8_t* pass, int32_t* npass,
                  uint32_t* keys, int32_t* vals, void* temp, size_t temp_bytes, hipStream_t st) {
  if (npass) hipLaunchKernelGGL(k_rank_zero, dim3(1), dim3(1), 0, st, npass);
  hipLaunchKernelGGL(k_rank_prep, dim3((n + 255) / 256), dim3(256), 0, st, info, n, R, keys, vals, pass, npass);
  hipcub::DoubleBuffer<uint32_t> kb(keys, keys + n);
  hipcub::DoubleBuffer<int32_t> vb(vals, vals + n);
  if (hipcub::DeviceRadixSort::SortPairs(temp, temp_bytes, kb, vb, n, 0, 32, st) != hipSuccess)
    throw std::runtime_error("ce_rank: sort failed");
  if (k > 0) hipLaunchKernelGGL(k_rank_out, dim3((k + 255) / 256), dim3(256), 0, st, vb.Current(), k, order);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) throw std::runtime_error(std::string("ce_rank: ") + hipGetErrorString(e));
}

// ce_rank_noalloc: the sort's double buffers (keys, indices) and its temporary storage live in one
// static device array of a fixed size, which each device holds for the life of the process.  Host
// code never passes the array's name to a kernel or a copy: its device address on the current
// device comes from hipGetSymbolAddress, looked up once per device.
constexpr int RANK_NOALLOC_MAX_N = 1 << 20;
constexpr size_t RANK_NOALLOC_KV = (size_t)2 * RANK_NOALLOC_MAX_N * sizeof(uint32_t);  // 8 MiB per buffer pair
constexpr size_t RANK_NOALLOC_TEMP = (size_t)4 << 20;
static __device__ __align__(256) unsigned char scratch[2 * RANK_NOALLOC_KV + RANK_NOALLOC_TEMP];

void* rank_scratch_ptr() {
  static std::mutex mu;
  static std::vector<void*> ptrs;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) throw std::runtime_error("ce_rank_noalloc: no current device");
  std::lock_guard<std::mutex> lk(mu);
  if ((int)ptrs.size() <= dev) ptrs.resize((size_t)dev + 1, nullptr);
  if (!ptrs[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      throw std::runtime_error(std::string("ce_rank_noalloc: scratch address lookup failed: ") + hipGetErrorString(e));
    ptrs[dev] = p;
  }
  return ptrs[dev];
}

void rank_check_args(const ce_config* cfg, const float* info, int n, int k, int32_t* order) {
  if (!cfg || !info) throw std::runtime_error("ce_rank: null argument");
  if (n <= 0) throw std::runtime_error("ce_rank: need n > 0");
  if (k < 0 || k > n) throw std::runtime_error("ce_rank: need 0 <= k <= n");
  if (k > 0 && !order) throw std::runtime_error("ce_rank: k > 0 needs an order buffer");
}

}  // namespace ce

using namespace ce;

static void rank_put_err(char* err, int errlen, const std::string& m) {
  if (err && errlen > 0) std::snprintf(err, (size_t)errlen, "%s", m.c_str());
}

extern "C" CE_EXPORT int ce_rank(const ce_config* cfg, const float* d_info, int n, int k, int32_t* d_order,
                                 uint8_t* d_pass, int32_t* d_npass, void* stream, char* err, int errlen) {
  try {
    rank_check_args(cfg, d_info, n, k, d_order);
    const RankRule R = rank_rule(*cfg);
    const hipStream_t st = (hipStream_t)stream;
    const size_t tb = rank_temp_bytes(n);
    const size_t kv = ((size_t)2 * n * sizeof(uint32_t) + 255) & ~(size_t)255;
    void* buf = nullptr;
    // stream-ordered allocation, released after the sort on the same stream
    if (hipMallocAsync(&buf, 2 * kv + tb, st) != hipSuccess) throw std::runtime_error("ce_rank: allocation failed");
    try {
      rank_enqueue(R, d_info, n, k, d_order, d_pass, d_npass, (uint32_t*)buf, (int32_t*)((char*)buf + kv),
                   (char*)buf + 2 * kv, tb, st);
    } catch (...) {
      hipFreeAsync(buf, st);
      throw;
    }
    if (hipFreeAsync(buf, st) != hipSuccess) throw std::runtime_error("ce_rank: release failed");
    return 0;
  } catch (const std::exception& e) {
    rank_put_err(err, errlen, e.what());
    return 1;
  }
}

extern "C" CE_EXPORT int ce_rank_noalloc(const ce_config* cfg, const float* d_info, int n, int k, int32_t* d_order,
                                         uint8_t* d_pass, int32_t* d_npass, void* stream, char* err, int errlen) {
  try {
    rank_check_args(cfg, d_info, n, k, d_order);
    if (n > RANK_NOALLOC_MAX_N)
      throw std::runtime_error("ce_rank_noalloc: n = " + std::to_string(n) + " exceeds the maximum of " +
                               std::to_string(RANK_NOALLOC_MAX_N) + " candidates (use ce_rank)");
    const RankRule R = rank_rule(*cfg);
    const size_t tb = rank_temp_bytes(n);
    if (tb > RANK_NOALLOC_TEMP)
      throw std::runtime_error("ce_rank_noalloc: the sort needs " + std::to_string(tb) + " bytes of temporary storage, "
                               "more than the " + std::to_string(RANK_NOALLOC_TEMP) + " reserved");
    char* base = (char*)rank_scratch_ptr();
    rank_enqueue(R, d_info, n, k, d_order, d_pass, d_npass, (uint32_t*)base, (int32_t*)(base + RANK_NOALLOC_KV),
                 base + 2 * RANK_NOALLOC_KV, tb, (hipStream_t)stream);
    return 0;
  } catch (const std::exception& e) {
    rank_put_err(err, errlen, e.what());
    return 1;
  }
}

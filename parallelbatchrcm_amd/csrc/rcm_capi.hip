// rcm_capi.hip -- host side of the C ABI (include/rcm_gpu_batch_abi.h): device-resident contexts, kernel
// launches, the GPU pseudo-peripheral start search and the bandwidth reduction.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>
#include <thread>
#include <vector>

#include "engine.hpp"
#include "rcm_gpu_batch_abi.h"

using namespace rcm;

namespace {

using GeoSmall = Geo<256, 2048, 256>;
constexpr int kDigit = 4;
constexpr int kDefaultSegThreshold = 24;

template <bool STRICT, bool BFS, bool PROF = false>
void* reorder_kernel() {
  return reinterpret_cast<void*>(&k_reorder<GeoSmall, kDigit, STRICT, BFS, PROF>);
}

#define RCM_HIP(call)                         \
  do {                                        \
    if ((call) != hipSuccess) return RCM_EDEVICE; \
  } while (0)

template <class T>
hipError_t dev_alloc(T** p, size_t count) {
  return hipMalloc(reinterpret_cast<void**>(p), std::max<size_t>(count, 1) * sizeof(T));
}
template <class T>
void dev_free(T*& p) {
  if (p) (void)hipFree(p);
  p = nullptr;
}

}  // namespace

struct rcm_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev[4] = {};
  // parameters
  int64_t workers = 0;
  int64_t timeout_ms = 20000;
  int64_t sort_mode = 0;
  int64_t digit_bits = kDigit;
  int64_t variant = 0;
  int64_t strict = 0;
  int64_t profile = 0;
  uint64_t* prof = nullptr;  // kProfSlots words
  // matrix
  int32_t n = 0;
  int64_t nnz = 0;
  int32_t* rp = nullptr;
  int32_t* ci = nullptr;
  bool own_csr = false;
  int64_t cap_rp = 0, cap_ci = 0;
  uint32_t maxdeg = 0;
  // engine state
  int32_t cap = 0;
  int32_t* perm = nullptr;
  uint64_t* qinfo = nullptr;
  uint64_t* claim = nullptr;
  uint64_t* bstat = nullptr;
  uint64_t* brange = nullptr;
  int32_t* level = nullptr;
  int32_t* out = nullptr;
  int32_t* inv = nullptr;
  Control* ctl = nullptr;
  uint32_t* scratch = nullptr;  // 16 words: validation flags, bandwidth
  Control* h_ctl = nullptr;     // pinned
  int32_t dirty = 0;  // batch records [0, dirty) may hold values from earlier runs
  int max_resident = 0;
  // last result
  bool have_perm = false;
  bool reversed = false;
  rcm_timing_t timing{};
};

namespace {

int ensure_state(rcm_ctx* c, int32_t n) {
  if (c->cap >= n && c->perm) return RCM_OK;
  dev_free(c->perm);
  dev_free(c->qinfo);
  dev_free(c->claim);
  dev_free(c->bstat);
  dev_free(c->brange);
  dev_free(c->level);
  dev_free(c->out);
  dev_free(c->inv);
  c->cap = 0;
  const size_t nn = (size_t)n;
  if (dev_alloc(&c->perm, nn) != hipSuccess || dev_alloc(&c->qinfo, nn) != hipSuccess ||
      dev_alloc(&c->claim, nn) != hipSuccess || dev_alloc(&c->bstat, nn + 1) != hipSuccess ||
      dev_alloc(&c->brange, nn + 1) != hipSuccess || dev_alloc(&c->level, nn) != hipSuccess ||
      dev_alloc(&c->out, nn) != hipSuccess || dev_alloc(&c->inv, nn) != hipSuccess)
    return RCM_ENOMEM;
  c->cap = n;
  c->dirty = n + 1;  // fresh memory: clear every batch record before the first run
  return RCM_OK;
}

int wait_stream(rcm_ctx* c, hipEvent_t ev, int64_t limit_ms) {
  const auto t0 = std::chrono::steady_clock::now();
  for (;;) {
    const hipError_t q = hipEventQuery(ev);
    if (q == hipSuccess) return RCM_OK;
    if (q != hipErrorNotReady) return RCM_EDEVICE;
    const auto ms = std::chrono::duration_cast<std::chrono::milliseconds>(std::chrono::steady_clock::now() - t0).count();
    if (ms > limit_ms) return RCM_ETIMEOUT;
    std::this_thread::sleep_for(std::chrono::microseconds(20));
  }
}

int workers_for(rcm_ctx* c) {
  int w = (int)c->workers;
  if (w <= 0) w = std::min<int64_t>(256, (int64_t)c->n / 8192 + 1);
  return std::max(1, std::min(w, c->max_resident));
}

// One engine run from `start` (CM order, or plain BFS order with levels when bfs).
int run_engine(rcm_ctx* c, int32_t start, bool bfs) {
  if (!c->rp) return RCM_ENOTLOADED;
  if (start < 0 || start >= c->n) return RCM_EINVAL;
  RunArgs a{};
  a.row_ptr = c->rp;
  a.col_idx = c->ci;
  a.perm = c->perm;
  a.qinfo = c->qinfo;
  a.claim = c->claim;
  a.bstat = c->bstat;
  a.brange = c->brange;
  a.level = bfs ? c->level : nullptr;
  a.ctl = c->ctl;
  a.n = c->n;
  a.bcap = c->n + 1;
  a.budget_ticks = (uint64_t)c->timeout_ms * 100000ull;
  a.sort_mode = (int32_t)c->sort_mode;
  a.seg_threshold = kDefaultSegThreshold;
  const bool prof = c->profile && !bfs && !c->strict;
  a.prof = prof ? c->prof : nullptr;
  if (prof) RCM_HIP(hipMemsetAsync(c->prof, 0, kProfSlots * sizeof(uint64_t), c->stream));
  // clear every record an earlier run may have written, also beyond this matrix's n (a previous, larger
  // matrix in the same buffers): workers read records ahead of their producers
  const int32_t bclear = std::max(1, std::min(c->dirty, c->cap + 1));
  const int init_blocks = (int)std::min<int64_t>(2048, ((int64_t)std::max(c->n, bclear) + 255) / 256);
  const int workers = workers_for(c);
  RCM_HIP(hipEventRecord(c->ev[0], c->stream));
  hipLaunchKernelGGL(k_init, dim3(init_blocks), dim3(256), 0, c->stream, a, start, bclear);
  RCM_HIP(hipEventRecord(c->ev[1], c->stream));
  void* kern = c->strict ? (bfs ? reorder_kernel<true, true>() : reorder_kernel<true, false>())
                         : (bfs ? reorder_kernel<false, true>()
                                : (prof ? reorder_kernel<false, false, true>() : reorder_kernel<false, false>()));
  void* args[] = {&a};
  RCM_HIP(hipLaunchKernel(kern, dim3(workers), dim3(GeoSmall::T), args, 0, c->stream));
  RCM_HIP(hipEventRecord(c->ev[2], c->stream));
  RCM_HIP(hipMemcpyAsync(c->h_ctl, c->ctl, sizeof(Control), hipMemcpyDeviceToHost, c->stream));
  RCM_HIP(hipEventRecord(c->ev[3], c->stream));
  int rc = wait_stream(c, c->ev[3], c->timeout_ms + 10000);
  if (rc) return rc;
  float ms_init = 0.f, ms_run = 0.f;
  RCM_HIP(hipEventElapsedTime(&ms_init, c->ev[0], c->ev[1]));
  RCM_HIP(hipEventElapsedTime(&ms_run, c->ev[1], c->ev[2]));
  c->timing.init_kernel_ms = ms_init;
  c->timing.run_kernel_ms = ms_run;
  c->timing.workers = workers;
  const Control& h = *c->h_ctl;
  c->timing.batches = (int32_t)h.nbatches;
  c->timing.placed = (int32_t)h.placed;
  c->timing.error_site = (int32_t)h.diag[0];
  for (int i = 0; i < 8; ++i) c->timing.error_info[i] = (int32_t)h.diag[i + 1];
  if (h.error != ERR_NONE || h.nbatches == 0) {
    c->dirty = c->cap + 1;
    if (h.error == ERR_TIMEOUT) return RCM_ETIMEOUT;
    if (h.error == ERR_CAPACITY || h.error == ERR_LONGROW) return RCM_ECAPACITY;
    return RCM_EDEVICE;
  }
  c->dirty = (int32_t)h.nbatches;
  if ((int32_t)h.placed != c->n) return RCM_ENOTCONNECTED;
  return RCM_OK;
}

int load_common(rcm_ctx* c, int32_t n, int64_t nnz) {
  int rc = ensure_state(c, n);
  if (rc) return rc;
  RCM_HIP(hipMemsetAsync(c->scratch, 0, 16 * sizeof(uint32_t), c->stream));
  const int blocks = (int)std::min<int64_t>(4096, (std::max<int64_t>(nnz, n) + 255) / 256);
  hipLaunchKernelGGL(k_validate, dim3(std::max(blocks, 1)), dim3(256), 0, c->stream, c->rp, c->ci, n, nnz, c->scratch);
  uint32_t h[2] = {0, 0};
  RCM_HIP(hipMemcpyAsync(h, c->scratch, sizeof(h), hipMemcpyDeviceToHost, c->stream));
  RCM_HIP(hipStreamSynchronize(c->stream));
  if (h[0] != 0) return RCM_EINVAL;
  c->maxdeg = h[1];
  c->n = n;
  c->nnz = nnz;
  c->have_perm = false;
  return RCM_OK;
}

}  // namespace

extern "C" {

const char* rcm_status_string(int s) {
  switch (s) {
    case RCM_OK: return "ok";
    case RCM_EINVAL: return "invalid argument or malformed CSR";
    case RCM_ENOTCONNECTED: return "graph is not connected";
    case RCM_ENOMEM: return "out of memory";
    case RCM_EDEVICE: return "HIP runtime error";
    case RCM_ETIMEOUT: return "device watchdog expired";
    case RCM_ECAPACITY: return "internal capacity exceeded";
    case RCM_ENOTLOADED: return "no matrix loaded";
    default: return "unknown status";
  }
}

int rcm_ctx_create(int device, rcm_ctx** out) {
  if (!out) return RCM_EINVAL;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev) return RCM_EDEVICE;
  RCM_HIP(hipSetDevice(device));
  rcm_ctx* c = new (std::nothrow) rcm_ctx();
  if (!c) return RCM_ENOMEM;
  c->device = device;
  bool ok = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) == hipSuccess;
  for (auto& e : c->ev) ok = ok && hipEventCreate(&e) == hipSuccess;
  ok = ok && dev_alloc(&c->ctl, 1) == hipSuccess && dev_alloc(&c->scratch, 16) == hipSuccess &&
       dev_alloc(&c->prof, kProfSlots) == hipSuccess;
  ok = ok && hipHostMalloc(reinterpret_cast<void**>(&c->h_ctl), sizeof(Control), hipHostMallocDefault) == hipSuccess;
  int per_cu = 0;
  hipDeviceProp_t prop;
  ok = ok && hipGetDeviceProperties(&prop, device) == hipSuccess;
  ok = ok && hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void*>(reorder_kernel<false, false>()),
                                                           GeoSmall::T, 0) == hipSuccess;
  if (!ok) {
    rcm_ctx_destroy(c);
    return RCM_EDEVICE;
  }
  c->max_resident = std::max(1, per_cu) * prop.multiProcessorCount;
  *out = c;
  return RCM_OK;
}

int rcm_ctx_destroy(rcm_ctx* c) {
  if (!c) return RCM_EINVAL;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  if (c->own_csr) {
    dev_free(c->rp);
    dev_free(c->ci);
  }
  dev_free(c->perm);
  dev_free(c->qinfo);
  dev_free(c->claim);
  dev_free(c->bstat);
  dev_free(c->brange);
  dev_free(c->level);
  dev_free(c->out);
  dev_free(c->inv);
  dev_free(c->ctl);
  dev_free(c->scratch);
  dev_free(c->prof);
  if (c->h_ctl) (void)hipHostFree(c->h_ctl);
  for (auto& e : c->ev)
    if (e) (void)hipEventDestroy(e);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return RCM_OK;
}

int rcm_ctx_set_param(rcm_ctx* c, int param, int64_t v) {
  if (!c) return RCM_EINVAL;
  switch (param) {
    case RCM_PARAM_WORKERS: if (v < 0) return RCM_EINVAL; c->workers = v; return RCM_OK;
    case RCM_PARAM_TIMEOUT_MS: if (v <= 0) return RCM_EINVAL; c->timeout_ms = v; return RCM_OK;
    case RCM_PARAM_SORT: if (v < 0 || v > 2) return RCM_EINVAL; c->sort_mode = v; return RCM_OK;
    case RCM_PARAM_DIGIT_BITS: if (v != kDigit) return RCM_EINVAL; c->digit_bits = v; return RCM_OK;
    case RCM_PARAM_VARIANT: if (v != 0) return RCM_EINVAL; c->variant = v; return RCM_OK;
    case RCM_PARAM_STRICT: c->strict = v ? 1 : 0; return RCM_OK;
    case RCM_PARAM_PROFILE: c->profile = v ? 1 : 0; return RCM_OK;
    default: return RCM_EINVAL;
  }
}

int rcm_ctx_profile(rcm_ctx* c, uint64_t* out, int32_t count) {
  if (!c || !out || count <= 0) return RCM_EINVAL;
  RCM_HIP(hipSetDevice(c->device));
  RCM_HIP(hipMemcpy(out, c->prof, (size_t)std::min<int32_t>(count, kProfSlots) * sizeof(uint64_t),
                    hipMemcpyDeviceToHost));
  return RCM_OK;
}

int rcm_ctx_get_param(rcm_ctx* c, int param, int64_t* v) {
  if (!c || !v) return RCM_EINVAL;
  switch (param) {
    case RCM_PARAM_WORKERS: *v = c->workers; return RCM_OK;
    case RCM_PARAM_TIMEOUT_MS: *v = c->timeout_ms; return RCM_OK;
    case RCM_PARAM_SORT: *v = c->sort_mode; return RCM_OK;
    case RCM_PARAM_DIGIT_BITS: *v = c->digit_bits; return RCM_OK;
    case RCM_PARAM_VARIANT: *v = c->variant; return RCM_OK;
    case RCM_PARAM_STRICT: *v = c->strict; return RCM_OK;
    case RCM_PARAM_PROFILE: *v = c->profile; return RCM_OK;
    default: return RCM_EINVAL;
  }
}

int rcm_ctx_load(rcm_ctx* c, int32_t n, int64_t nnz, const int32_t* row_ptr, const int32_t* col_idx) {
  if (!c || n < 1 || n >= (1 << 30) || nnz < 0 || nnz > INT32_MAX || !row_ptr || (nnz > 0 && !col_idx))
    return RCM_EINVAL;
  RCM_HIP(hipSetDevice(c->device));
  const auto t0 = std::chrono::steady_clock::now();
  if (!c->own_csr) {
    c->rp = nullptr;
    c->ci = nullptr;
    c->cap_rp = c->cap_ci = 0;
  }
  if (c->cap_rp < (int64_t)n + 1) {
    dev_free(c->rp);
    if (dev_alloc(&c->rp, (size_t)n + 1) != hipSuccess) return RCM_ENOMEM;
    c->cap_rp = (int64_t)n + 1;
  }
  if (c->cap_ci < nnz || !c->ci) {
    dev_free(c->ci);
    if (dev_alloc(&c->ci, (size_t)nnz) != hipSuccess) return RCM_ENOMEM;
    c->cap_ci = std::max<int64_t>(nnz, 1);
  }
  c->own_csr = true;
  RCM_HIP(hipMemcpyAsync(c->rp, row_ptr, ((size_t)n + 1) * sizeof(int32_t), hipMemcpyHostToDevice, c->stream));
  if (nnz > 0)
    RCM_HIP(hipMemcpyAsync(c->ci, col_idx, (size_t)nnz * sizeof(int32_t), hipMemcpyHostToDevice, c->stream));
  c->n = 0;
  const int rc = load_common(c, n, nnz);
  c->timing.load_ms =
      std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  return rc;
}

int rcm_ctx_load_device(rcm_ctx* c, int32_t n, int64_t nnz, const int32_t* d_row_ptr, const int32_t* d_col_idx) {
  if (!c || n < 1 || n >= (1 << 30) || nnz < 0 || nnz > INT32_MAX || !d_row_ptr || (nnz > 0 && !d_col_idx))
    return RCM_EINVAL;
  RCM_HIP(hipSetDevice(c->device));
  if (c->own_csr) {
    dev_free(c->rp);
    dev_free(c->ci);
    c->cap_rp = c->cap_ci = 0;
  }
  c->own_csr = false;
  c->rp = const_cast<int32_t*>(d_row_ptr);
  c->ci = const_cast<int32_t*>(d_col_idx);
  c->n = 0;
  return load_common(c, n, nnz);
}

int rcm_ctx_run(rcm_ctx* c, int32_t start, uint32_t flags) {
  if (!c) return RCM_EINVAL;
  if (!c->rp || c->n == 0) return RCM_ENOTLOADED;
  RCM_HIP(hipSetDevice(c->device));
  c->have_perm = false;
  const int rc = run_engine(c, start, false);
  if (rc) return rc;
  c->reversed = (flags & RCM_FLAG_REVERSE) != 0;
  if (c->reversed) {
    hipLaunchKernelGGL(k_reverse, dim3((c->n + 255) / 256), dim3(256), 0, c->stream, c->perm, c->n, c->out);
    RCM_HIP(hipStreamSynchronize(c->stream));
  }
  c->have_perm = true;
  return RCM_OK;
}

// SimplePeripheral::findPeripheral (src/simple_peripheral.cpp:92-141) on the GPU: each sweep is one engine
// run in plain-BFS mode; the last min(32, n-1) queue entries are the last nodes discovered, with their BFS
// levels, in discovery order.
int rcm_ctx_find_start(rcm_ctx* c, int32_t* start_out, int32_t* iterations_out) {
  if (!c || !start_out) return RCM_EINVAL;
  if (!c->rp || c->n == 0) return RCM_ENOTLOADED;
  RCM_HIP(hipSetDevice(c->device));
  const auto t0 = std::chrono::steady_clock::now();
  const int32_t n = c->n;
  int32_t cur = n / 2;
  int32_t iters = 0;
  if (n == 1) {
    cur = 0;
    iters = 1;
  } else {
    uint32_t lastdist = 0;
    const int32_t w = std::min<int32_t>(32, n - 1);
    std::vector<int32_t> nodes((size_t)w), dists((size_t)w);
    std::vector<uint64_t> rows((size_t)w);
    for (;;) {
      ++iters;
      c->have_perm = false;
      const int rc = run_engine(c, cur, true);
      if (rc) return rc;
      RCM_HIP(hipMemcpyAsync(nodes.data(), c->perm + (n - w), (size_t)w * sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
      RCM_HIP(hipMemcpyAsync(dists.data(), c->level + (n - w), (size_t)w * sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
      RCM_HIP(hipMemcpyAsync(rows.data(), c->qinfo + (n - w), (size_t)w * sizeof(uint64_t), hipMemcpyDeviceToHost, c->stream));
      RCM_HIP(hipStreamSynchronize(c->stream));
      if ((uint32_t)dists[w - 1] == lastdist) break;
      int32_t idx = 0;  // first index holding the maximum distance
      for (int32_t i = 1; i < w; ++i)
        if (dists[i] > dists[idx]) idx = i;
      int64_t minval = INT64_MAX;
      for (int32_t i = idx; i < w; ++i) {  // minimum degree, later nodes win ties
        const int64_t val = (int64_t)(uint32_t)(rows[i] >> 32) - (int64_t)(uint32_t)rows[i];
        if (val <= minval) {
          minval = val;
          cur = nodes[i];
        }
      }
      lastdist = (uint32_t)dists[w - 1];
    }
  }
  c->timing.start_search_ms =
      std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  c->timing.start_iterations = iters;
  *start_out = cur;
  if (iterations_out) *iterations_out = iters;
  return RCM_OK;
}

int rcm_ctx_bandwidth(rcm_ctx* c, int64_t* bw_out) {
  if (!c || !bw_out) return RCM_EINVAL;
  if (!c->have_perm) return RCM_ENOTLOADED;
  RCM_HIP(hipSetDevice(c->device));
  const int blocks = (c->n + 255) / 256;
  RCM_HIP(hipMemsetAsync(c->scratch + 8, 0, sizeof(uint32_t), c->stream));
  RCM_HIP(hipEventRecord(c->ev[0], c->stream));
  hipLaunchKernelGGL(k_inverse, dim3(blocks), dim3(256), 0, c->stream, c->perm, c->n, c->inv);
  hipLaunchKernelGGL(k_bandwidth, dim3(blocks), dim3(256), 0, c->stream, c->rp, c->ci, c->inv, c->n, c->scratch + 8);
  RCM_HIP(hipEventRecord(c->ev[1], c->stream));
  uint32_t bw = 0;
  RCM_HIP(hipMemcpyAsync(&bw, c->scratch + 8, sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
  RCM_HIP(hipStreamSynchronize(c->stream));
  float ms = 0.f;
  RCM_HIP(hipEventElapsedTime(&ms, c->ev[0], c->ev[1]));
  c->timing.bandwidth_ms = ms;
  *bw_out = (int64_t)bw;
  return RCM_OK;
}

int rcm_ctx_bandwidth_of(rcm_ctx* c, const int32_t* perm, int64_t* bw_out) {
  if (!c || !perm || !bw_out) return RCM_EINVAL;
  if (!c->rp || c->n == 0) return RCM_ENOTLOADED;
  std::vector<uint8_t> seen((size_t)c->n, 0);
  for (int32_t i = 0; i < c->n; ++i) {
    if (perm[i] < 0 || perm[i] >= c->n || seen[perm[i]]) return RCM_EINVAL;
    seen[perm[i]] = 1;
  }
  RCM_HIP(hipSetDevice(c->device));
  RCM_HIP(hipMemcpyAsync(c->perm, perm, (size_t)c->n * sizeof(int32_t), hipMemcpyHostToDevice, c->stream));
  c->have_perm = true;
  c->reversed = false;
  return rcm_ctx_bandwidth(c, bw_out);
}

int rcm_ctx_read_perm(rcm_ctx* c, int32_t* perm_out) {
  if (!c || !perm_out) return RCM_EINVAL;
  if (!c->have_perm) return RCM_ENOTLOADED;
  RCM_HIP(hipSetDevice(c->device));
  RCM_HIP(hipMemcpyAsync(perm_out, c->reversed ? c->out : c->perm, (size_t)c->n * sizeof(int32_t),
                         hipMemcpyDeviceToHost, c->stream));
  RCM_HIP(hipStreamSynchronize(c->stream));
  return RCM_OK;
}

int rcm_ctx_perm_device(rcm_ctx* c, const int32_t** d_perm_out) {
  if (!c || !d_perm_out) return RCM_EINVAL;
  if (!c->have_perm) return RCM_ENOTLOADED;
  *d_perm_out = c->reversed ? c->out : c->perm;
  return RCM_OK;
}

int rcm_ctx_timing(rcm_ctx* c, rcm_timing_t* out) {
  if (!c || !out) return RCM_EINVAL;
  *out = c->timing;
  return RCM_OK;
}

int rcm_gpu_batch(int32_t n, int64_t nnz, const int32_t* row_ptr, const int32_t* col_idx, int32_t start_node,
                  uint32_t flags, int32_t* perm_out, int64_t* bandwidth_out, int32_t* start_node_out) {
  if (!perm_out) return RCM_EINVAL;
  if (start_node >= n) return RCM_EINVAL;
  int dev = 0;
  RCM_HIP(hipGetDevice(&dev));
  rcm_ctx* c = nullptr;
  int rc = rcm_ctx_create(dev, &c);
  if (rc) return rc;
  rc = rcm_ctx_load(c, n, nnz, row_ptr, col_idx);
  int32_t start = start_node;
  if (!rc && start < 0) rc = rcm_ctx_find_start(c, &start, nullptr);
  if (!rc) rc = rcm_ctx_run(c, start, flags);
  if (!rc && bandwidth_out && (flags & RCM_FLAG_BANDWIDTH)) rc = rcm_ctx_bandwidth(c, bandwidth_out);
  if (!rc) rc = rcm_ctx_read_perm(c, perm_out);
  if (!rc && start_node_out) *start_node_out = start;
  rcm_ctx_destroy(c);
  return rc;
}

int rcm_gpu_batch_many(int32_t count, const int32_t* n, const int64_t* nnz, const int64_t* rp_off,
                       const int64_t* ci_off, const int32_t* row_ptr_cat, const int32_t* col_idx_cat,
                       const int32_t* start, uint32_t flags, const int64_t* po_off, int32_t* perm_cat,
                       int64_t* bandwidth_out, int32_t* status_out) {
  if (count < 0 || (count > 0 && (!n || !nnz || !rp_off || !ci_off || !row_ptr_cat || !col_idx_cat || !po_off ||
                                  !perm_cat)))
    return RCM_EINVAL;
  int dev = 0;
  RCM_HIP(hipGetDevice(&dev));
  rcm_ctx* c = nullptr;
  int rc = rcm_ctx_create(dev, &c);
  if (rc) return rc;
  int worst = RCM_OK;
  for (int32_t k = 0; k < count; ++k) {
    int r = rcm_ctx_load(c, n[k], nnz[k], row_ptr_cat + rp_off[k], col_idx_cat + ci_off[k]);
    int32_t s = start ? start[k] : -1;
    if (!r && s < 0) r = rcm_ctx_find_start(c, &s, nullptr);
    if (!r) r = rcm_ctx_run(c, s, flags);
    if (!r && bandwidth_out && (flags & RCM_FLAG_BANDWIDTH)) r = rcm_ctx_bandwidth(c, bandwidth_out + k);
    if (!r) r = rcm_ctx_read_perm(c, perm_cat + po_off[k]);
    if (status_out) status_out[k] = r;
    if (r && worst == RCM_OK) worst = r;
  }
  rcm_ctx_destroy(c);
  return worst;
}

}  // extern "C"

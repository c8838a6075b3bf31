// rcm_capi.hip -- host side of the C ABI (include/rcm_gpu_batch_abi.h): device-resident contexts, kernel
// launches, the GPU pseudo-peripheral start search and the bandwidth reduction.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <limits>
#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <thread>
#include <unordered_map>
#include <vector>

#include "engine_many.hpp"
#include "hip_exec.hpp"
#include "level_cm.hpp"
#include "rcm_gpu_batch_abi.h"
#include "verify.hpp"

using namespace rcm;

namespace {

constexpr int kDefaultSegThreshold = 24;
constexpr int64_t kDefaultHelperNodes = 8192;  // the autotuner's helper caps: n / 8192, / 32768, / 131072 + 1
// rcm_batch defaults: what rcm_batch_autotune picks on the 4096-matrix mix (variant 4, one helper per 131072
// rows: 117 ms per launch against 172 ms for variant 0 with one per 8192). A launch over a few matrices still
// spreads its workgroups over them by size (ManyArgs::share_rows), so the large quantum does not starve them.
constexpr int64_t kDefaultBatchHelperNodes = 131072;
constexpr int64_t kDefaultBatchVariant = 4;
// rcm_ctx default: variant 1 (512 threads over 2048 slots: each batch phase takes 4 rounds per wave, not 8). On
// MI355X the 128^3 stencil runs in 3.93-3.98 ms against 4.27-4.33 ms with variant 0, the 1M-point geometric graph in
// 8.42 ms against 9.7 ms, and rcm_ctx_autotune picks it for both and for R-MAT scale 20. The serving lanes keep
// variant 0 (capacity 914 against 673-698 requests/s with variant 1 on the 400-request stream).
constexpr int64_t kDefaultCtxVariant = 1;
// RCM_PARAM_SPECULATE: from this start-search sweep on, the sweep runs as a CM run that records levels. On the
// 4096-matrix mix 98% of the matrices stop at sweep 3, 1.8% at 4 and 0.4% at 2, so speculating from sweep 3
// turns (sweeps + 1) engine passes into sweeps for almost every matrix.
constexpr int64_t kDefaultSpeculate = 3;
// Serving lanes give each request more helpers: latency, not throughput, is the aim there. Measured on MI355X
// (Poisson stream, 95% mixed sizes and 5% R-MAT): 2048 against 8192 halves p50 at 200 and 400 requests/s.
constexpr int64_t kServerHelperNodes = 2048;
// Per serving lane, pinned host staging for a group's concatenated CSRs (uploaded by DMA, without the runtime's
// bounce copy through its own pinned buffers) and for its orders; larger groups stage in pageable memory.
constexpr size_t kServePinnedWords = (size_t)64 << 20;  // 256 MiB per lane
// Serving groups requests of at least this many rows only with each other.
constexpr int32_t kServeLargeN = 1 << 18;
// A serving lane orders a large request (kServeLargeN rows or more) alone, through a single-matrix context whose
// buffers are reserved at creation for up to kServeSoloN rows and kServeSoloNnz entries (no allocation while other
// lanes' kernels run): the single-matrix engine's default variant is faster on one matrix than a one-member launch
// of the many-matrix engine.
constexpr int32_t kServeSoloN = 1 << 21;
constexpr int64_t kServeSoloNnz = (int64_t)1 << 26;

static_assert(sizeof(rcm_tune_result) == 40, "the Python mirror (TuneResultT) assumes this layout");

// Workgroups that may help one matrix of n rows at once.
static int32_t helper_cap(int64_t n, int64_t nodes_per_helper) {
  return (int32_t)std::min<int64_t>(n / nodes_per_helper + 1, 1024);
}
constexpr int kDefaultClaimFilter = 1;  // measured on MI355X: 1.5-7% faster, 56% fewer HBM writes

// Engine variants (RCM_PARAM_VARIANT): block size T, batch width (adjacency slots S, parents P), radix digit;
// variant 3 also cuts child batches at a degree quantum of S/4 instead of S/2 (smaller batches: the 1M-point
// geometric graph 9.7 -> 9.3 ms with variant 0's geometry).
// The order never depends on the variant (batch boundaries do not change it); the speed does.
constexpr int kVariants = 5;
template <int V> struct Var;
template <> struct Var<0> { using G = Geo<256, 2048, 256>; static constexpr int D = 4; };
template <> struct Var<1> { using G = Geo<512, 2048, 256>; static constexpr int D = 4; };
template <> struct Var<2> { using G = Geo<128, 1024, 128>; static constexpr int D = 4; };
template <> struct Var<3> { using G = Geo<256, 2048, 256, 4>; static constexpr int D = 6; };
template <> struct Var<4> { using G = Geo<512, 4096, 512>; static constexpr int D = 4; };
constexpr int kVarThreads[kVariants] = {256, 512, 128, 256, 512};
constexpr int kVarSlots[kVariants] = {2048, 2048, 1024, 2048, 4096};
constexpr int kVarDigit[kVariants] = {4, 4, 4, 6, 4};
// Variant kLevelVariant of a single-matrix context is the level-synchronous engine (level_cm.hpp): one BFS
// level per step, ordered by a device-wide radix sort with 8-bit digits.
constexpr int kLevelVariant = kVariants;
constexpr int kLevelThreads = 256;
constexpr int kLevelDigit = 8;
using GeoSmall = Var<0>::G;  // the many-matrix helpers' default
constexpr int kDigit = 4;

// lev: a CM run that also records BFS levels (a speculated start-search sweep); plain-BFS runs always do.
template <int V>
void* reorder_kernel_v(bool strict, bool bfs, bool prof, bool lev) {
  using G = typename Var<V>::G;
  constexpr int D = Var<V>::D;
  if (strict) return bfs ? reinterpret_cast<void*>(&k_reorder<G, D, true, true>)
                         : lev ? reinterpret_cast<void*>(&k_reorder<G, D, true, false, false, true>)
                               : reinterpret_cast<void*>(&k_reorder<G, D, true, false>);
  if (bfs) return reinterpret_cast<void*>(&k_reorder<G, D, false, true>);
  if (lev) return reinterpret_cast<void*>(&k_reorder<G, D, false, false, false, true>);
  return prof ? reinterpret_cast<void*>(&k_reorder<G, D, false, false, true>)
              : reinterpret_cast<void*>(&k_reorder<G, D, false, false>);
}
void* reorder_kernel(int v, bool strict, bool bfs, bool prof, bool lev = false) {
  switch (v) {
    case 1: return reorder_kernel_v<1>(strict, bfs, prof, lev);
    case 2: return reorder_kernel_v<2>(strict, bfs, prof, lev);
    case 3: return reorder_kernel_v<3>(strict, bfs, prof, lev);
    case 4: return reorder_kernel_v<4>(strict, bfs, prof, lev);
    default: return reorder_kernel_v<0>(strict, bfs, prof, lev);
  }
}
template <int V>
void* many_kernel_v(bool bfs, bool lev) {
  using G = typename Var<V>::G;
  constexpr int D = Var<V>::D;
  if (bfs) return reinterpret_cast<void*>(&k_reorder_many<G, D, false, true>);
  return lev ? reinterpret_cast<void*>(&k_reorder_many<G, D, false, false, true>)
             : reinterpret_cast<void*>(&k_reorder_many<G, D, false, false>);
}
void* many_kernel(int v, bool bfs, bool lev = false) {
  switch (v) {
    case 1: return many_kernel_v<1>(bfs, lev);
    case 2: return many_kernel_v<2>(bfs, lev);
    case 3: return many_kernel_v<3>(bfs, lev);
    case 4: return many_kernel_v<4>(bfs, lev);
    default: return many_kernel_v<0>(bfs, lev);
  }
}

// Self-test of the wave64 and block primitives the engines are built on (rcm_selftest_primitives): one
// 256-thread workgroup; out = [block exclusive scan | block total | wave sum | wave max | wave min |
// wave inclusive scan], 256 words each (the total at word 256).
constexpr int kSelftestThreads = 256;
constexpr int kSelftestIn = kSelftestThreads;
constexpr int kSelftestOut = 6 * kSelftestThreads;
__global__ __launch_bounds__(kSelftestThreads) void k_selftest(const uint32_t* in, uint32_t* out) {
  __shared__ uint32_t wtmp[kSelftestThreads / kWave + 1];
  const int t = threadIdx.x;
  const uint32_t v = in[t];
  uint32_t total = 0;
  out[t] = block_excl_scan<kSelftestThreads>(v, wtmp, &total);
  out[kSelftestThreads + t] = total;
  out[2 * kSelftestThreads + t] = wave_sum(v);
  out[3 * kSelftestThreads + t] = wave_max(v);
  out[4 * kSelftestThreads + t] = wave_min(v);
  out[5 * kSelftestThreads + t] = wave_incl_scan(v);
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

// Uploads of a CSR from pageable host memory. kUpThreads host threads each take an equal share of the bytes and
// copy it in kUpChunk pieces into two pinned buffers of their own, alternating, with the DMA of one piece (on the
// thread's own stream) overlapping the copy of the next. Measured on MI355X for 2 GiB: 42-45 GB/s, against 28 GB/s
// for one hipMemcpyAsync from pageable memory and 35-37 GB/s for four of them from four threads. One uploader per
// device, made on the first upload of at least kUpMin bytes and kept until the process ends (2 x kUpThreads x
// kUpChunk = 128 MiB of pinned memory); uploads on one device take turns. Without it (the pinned allocation
// failed) the bytes go by hipMemcpyAsync.
constexpr int kUpThreads = 4;
constexpr size_t kUpChunk = (size_t)16 << 20;
constexpr size_t kUpMin = (size_t)32 << 20;
constexpr int kUpDevices = 64;

struct UpSeg {
  void* dst;
  const void* src;
  size_t bytes;
};

struct Uploader {
  std::mutex mu;
  int state = 0;  // 0 not made yet, 1 ready, -1 unavailable
  char* pin[kUpThreads][2] = {};
  hipEvent_t ev[kUpThreads][2] = {};
  hipStream_t st[kUpThreads] = {};

  bool ready() {  // with the device current and mu held
    if (state) return state > 0;
    bool ok = true;
    for (int p = 0; p < kUpThreads && ok; ++p) {
      ok = hipStreamCreateWithFlags(&st[p], hipStreamNonBlocking) == hipSuccess;
      for (int k = 0; k < 2 && ok; ++k)
        ok = hipHostMalloc(reinterpret_cast<void**>(&pin[p][k]), kUpChunk, hipHostMallocDefault) == hipSuccess &&
             hipEventCreateWithFlags(&ev[p][k], hipEventDisableTiming) == hipSuccess;
    }
    if (!ok) {
      (void)hipGetLastError();
      for (int p = 0; p < kUpThreads; ++p) {
        for (int k = 0; k < 2; ++k) {
          if (pin[p][k]) (void)hipHostFree(pin[p][k]);
          if (ev[p][k]) (void)hipEventDestroy(ev[p][k]);
          pin[p][k] = nullptr;
          ev[p][k] = nullptr;
        }
        if (st[p]) (void)hipStreamDestroy(st[p]);
        st[p] = nullptr;
      }
    }
    state = ok ? 1 : -1;
    return ok;
  }

  // Thread p's share [r0, r1) of the segments laid end to end.
  hipError_t share(const UpSeg* seg, int ns, int p, size_t r0, size_t r1) {
    hipError_t err = hipSuccess;
    int c = 0;
    size_t base = 0;
    for (int s = 0; s < ns && base < r1; base += seg[s].bytes, ++s) {
      const size_t a = std::max(r0, base), e = std::min(r1, base + seg[s].bytes);
      for (size_t off = a; off < e && err == hipSuccess; off += kUpChunk, ++c) {
        const size_t len = std::min(kUpChunk, e - off);
        const int k = c & 1;
        if (c >= 2) err = hipEventSynchronize(ev[p][k]);  // the buffer's previous piece has reached the device
        if (err != hipSuccess) break;
        std::memcpy(pin[p][k], static_cast<const char*>(seg[s].src) + (off - base), len);
        err = hipMemcpyAsync(static_cast<char*>(seg[s].dst) + (off - base), pin[p][k], len, hipMemcpyHostToDevice,
                             st[p]);
        if (err == hipSuccess) err = hipEventRecord(ev[p][k], st[p]);
      }
    }
    const hipError_t e2 = hipStreamSynchronize(st[p]);
    return err != hipSuccess ? err : e2;
  }
};

Uploader g_up[kUpDevices];

// Copies the segments to the device on `device` (current); when it returns the bytes are on the device. Work
// already queued on `cs` finishes first (it may read the destinations); small uploads simply queue on `cs`.
int upload_host(int device, hipStream_t cs, const UpSeg* seg, int ns) {
  size_t total = 0;
  for (int s = 0; s < ns; ++s) total += seg[s].bytes;
  Uploader* u = device >= 0 && device < kUpDevices && total >= kUpMin ? &g_up[device] : nullptr;
  std::unique_lock<std::mutex> lk;
  if (u) {
    lk = std::unique_lock<std::mutex>(u->mu);
    if (!u->ready()) u = nullptr;
  }
  if (!u) {
    for (int s = 0; s < ns; ++s)
      if (seg[s].bytes) RCM_HIP(hipMemcpyAsync(seg[s].dst, seg[s].src, seg[s].bytes, hipMemcpyHostToDevice, cs));
    return RCM_OK;
  }
  RCM_HIP(hipStreamSynchronize(cs));
  hipError_t errs[kUpThreads];
  auto bound = [&](int p) { return p >= kUpThreads ? total : (total * (size_t)p / kUpThreads) & ~(size_t)255; };
  std::vector<std::thread> th;
  for (int p = 1; p < kUpThreads; ++p)
    th.emplace_back([&, p] {
      errs[p] = hipSetDevice(device);  // a new thread starts on device 0
      if (errs[p] == hipSuccess) errs[p] = u->share(seg, ns, p, bound(p), bound(p + 1));
    });
  errs[0] = u->share(seg, ns, 0, bound(0), bound(1));
  for (auto& x : th) x.join();
  for (hipError_t e : errs)
    if (e != hipSuccess) return RCM_EDEVICE;
  return RCM_OK;
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
  int64_t claim_filter = kDefaultClaimFilter;
  int64_t variant = kDefaultCtxVariant;
  int64_t strict = 0;
  int64_t profile = 0;
  int64_t batch_cap = 0;     // RCM_PARAM_BATCH_CAP: batch records a run may use (0 = n + 1, all it can need)
  int64_t speculate = kDefaultSpeculate;  // RCM_PARAM_SPECULATE
  uint64_t* prof = nullptr;  // kProfSlots words
  uint64_t* trace = nullptr;  // RCM_PARAM_PROFILE = 2: kTraceWords words per batch id
  int64_t trace_cap = 0;
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
  int max_resident[kVariants] = {};  // co-resident workgroups of each variant
  rcm_mx::DevArena lscratch;         // level-synchronous engine scratch
  uint32_t* h_word = nullptr;        // pinned readback word of that engine
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
  if (c->variant == kLevelVariant) return 0;
  int w = (int)c->workers;
  if (w <= 0) w = c->n > 16384 ? 256 : 64;  // measured on MI355X: 128-256 is best from 65k to 2M nodes
  return std::max(1, std::min(w, c->max_resident[c->variant]));
}

int variant_threads(int64_t v) { return v == kLevelVariant ? kLevelThreads : kVarThreads[v]; }

// The level-synchronous engine: same outputs as run_engine (perm, qinfo, and levels for the plain BFS).
int run_level(rcm_ctx* c, int32_t start, bool bfs) {
  rcm_mx::HipExec x{c->stream, &c->lscratch, &c->lscratch, c->h_word};
  RCM_HIP(hipEventRecord(c->ev[0], c->stream));
  int32_t placed = 0, levels = 0;
  const int st = rcm_mx::level_order(x, c->n, c->nnz, c->rp, c->ci, c->maxdeg, start, c->perm, c->qinfo, c->claim,
                                     bfs ? c->level : nullptr, &placed, &levels);
  RCM_HIP(hipEventRecord(c->ev[2], c->stream));
  const bool synced = hipStreamSynchronize(c->stream) == hipSuccess;
  c->lscratch.reset();
  if (!synced || st == rcm_mx::kOpsDevice) return RCM_EDEVICE;
  if (st == rcm_mx::kOpsCapacity) return RCM_ECAPACITY;
  float ms = 0.f;
  RCM_HIP(hipEventElapsedTime(&ms, c->ev[0], c->ev[2]));
  c->timing.init_kernel_ms = 0.0;
  c->timing.run_kernel_ms = ms;
  c->timing.workers = 0;
  c->timing.batches = levels;
  c->timing.placed = placed;
  c->timing.error_site = 0;
  if (placed != c->n) return RCM_ENOTCONNECTED;
  return RCM_OK;
}

// One engine run from `start` (CM order, or plain BFS order with levels when bfs; a CM run also records the
// levels when `levels`).
int run_engine(rcm_ctx* c, int32_t start, bool bfs, bool levels = false) {
  if (!c->rp) return RCM_ENOTLOADED;
  if (start < 0 || start >= c->n) return RCM_EINVAL;
  if (c->variant == kLevelVariant) return run_level(c, start, bfs);
  RunArgs a{};
  a.row_ptr = c->rp;
  a.col_idx = c->ci;
  a.perm = c->perm;
  a.qinfo = c->qinfo;
  a.claim = c->claim;
  a.bstat = c->bstat;
  a.brange = c->brange;
  a.level = (bfs || levels) ? c->level : nullptr;
  a.ctl = c->ctl;
  a.n = c->n;
  a.bcap = (c->batch_cap > 0 && c->batch_cap < (int64_t)c->n + 1) ? (int32_t)c->batch_cap : c->n + 1;
  a.budget_ticks = (uint64_t)c->timeout_ms * 100000ull;
  a.sort_mode = (int32_t)c->sort_mode;
  a.claim_filter = (int32_t)c->claim_filter;
  a.seg_threshold = kDefaultSegThreshold;
  const bool prof = c->profile && !bfs && !c->strict && !levels;
  a.prof = prof ? c->prof : nullptr;
  if (prof) RCM_HIP(hipMemsetAsync(c->prof, 0, kProfSlots * sizeof(uint64_t), c->stream));
  if (prof && c->profile == 2) {  // per-batch trace
    if (c->trace_cap < (int64_t)c->n + 1) {
      dev_free(c->trace);
      c->trace_cap = 0;
      if (dev_alloc(&c->trace, ((size_t)c->n + 1) * kTraceWords) != hipSuccess) return RCM_ENOMEM;
      c->trace_cap = (int64_t)c->n + 1;
    }
    RCM_HIP(hipMemsetAsync(c->trace, 0, (size_t)c->trace_cap * kTraceWords * sizeof(uint64_t), c->stream));
    a.trace = c->trace;
    a.trace_cap = (int32_t)c->trace_cap;
  }
  // clear every record an earlier run may have written, also beyond this matrix's n (a previous, larger
  // matrix in the same buffers): workers read records ahead of their producers
  const int32_t bclear = std::max(1, std::min(c->dirty, c->cap + 1));
  const int init_blocks = (int)std::min<int64_t>(2048, ((int64_t)std::max(c->n, bclear) + 255) / 256);
  const int workers = workers_for(c);
  RCM_HIP(hipEventRecord(c->ev[0], c->stream));
  hipLaunchKernelGGL(k_init, dim3(init_blocks), dim3(256), 0, c->stream, a, start, bclear);
  RCM_HIP(hipEventRecord(c->ev[1], c->stream));
  void* kern = reorder_kernel((int)c->variant, c->strict != 0, bfs, prof, levels && !bfs);
  void* args[] = {&a};
  RCM_HIP(hipLaunchKernel(kern, dim3(workers), dim3(kVarThreads[c->variant]), args, 0, c->stream));
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

// One step of SimplePeripheral::findPeripheral (src/simple_peripheral.cpp:44-90, outputnodes = 32) on the last
// w = min(32, n - 1) nodes a plain-BFS sweep discovered, in discovery order (node, BFS level, row length).
// Returns true when the sweep's eccentricity equals the previous one (the start *cur is final); otherwise moves
// *cur to the node of least row length among those from the first of greatest level on, later nodes winning
// ties, and records the eccentricity in *lastdist.
bool peripheral_step(const int32_t* nodes, const int32_t* dists, const int32_t* degs, int32_t w, uint32_t* lastdist,
                     int32_t* cur) {
  if ((uint32_t)dists[w - 1] == *lastdist) return true;
  int32_t idx = 0;  // first index holding the maximum distance
  for (int32_t i = 1; i < w; ++i)
    if (dists[i] > dists[idx]) idx = i;
  int64_t minval = INT64_MAX;
  for (int32_t i = idx; i < w; ++i) {  // minimum degree, later nodes win ties
    if ((int64_t)degs[i] <= minval) {
      minval = degs[i];
      *cur = nodes[i];
    }
  }
  *lastdist = (uint32_t)dists[w - 1];
  return false;
}

// The order check of verify.hpp over the context's matrix for the device order d_perm (n entries). Uses the
// inverse and claim buffers as scratch. *result_out: 0 when d_perm is the order, else 1 + a position where it
// departs.
int verify_device(rcm_ctx* c, const int32_t* d_perm, int32_t start, bool bfs, int64_t* result_out) {
  const int32_t n = c->n;
  int32_t* pos = c->inv;
  uint32_t* ppos = reinterpret_cast<uint32_t*>(c->claim);  // claim holds 2n words
  uint32_t* rix = ppos + n;
  uint32_t* bad = c->scratch + 9;
  const unsigned g1 = (unsigned)((n + 255) / 256), gw = (unsigned)(((int64_t)n * kWave + 255) / 256);
  hipLaunchKernelGGL(k_vfy_fill, dim3(g1), dim3(256), 0, c->stream, reinterpret_cast<uint32_t*>(pos), (int64_t)n,
                     0xFFFFFFFFu);
  hipLaunchKernelGGL(k_vfy_fill, dim3(2 * g1), dim3(256), 0, c->stream, ppos, 2 * (int64_t)n, kVerifyNone);
  hipLaunchKernelGGL(k_vfy_fill, dim3(1), dim3(64), 0, c->stream, bad, (int64_t)1, kVerifyNone);
  hipLaunchKernelGGL(k_vfy_positions, dim3(g1), dim3(256), 0, c->stream, d_perm, n, start, pos, bad);
  hipLaunchKernelGGL(k_vfy_parents, dim3(gw), dim3(256), 0, c->stream, (const int32_t*)c->rp, (const int32_t*)c->ci,
                     n, (const int32_t*)pos, ppos);
  hipLaunchKernelGGL(k_vfy_rowindex, dim3(gw), dim3(256), 0, c->stream, (const int32_t*)c->rp,
                     (const int32_t*)c->ci, n, (const int32_t*)pos, (const uint32_t*)ppos, rix);
  hipLaunchKernelGGL(k_vfy_keys, dim3(g1), dim3(256), 0, c->stream, (const int32_t*)c->rp, d_perm, n,
                     (const uint32_t*)ppos, (const uint32_t*)rix, bfs ? 1 : 0, bad);
  RCM_HIP(hipGetLastError());
  uint32_t h = 0;
  RCM_HIP(hipMemcpyAsync(&h, bad, sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
  RCM_HIP(hipStreamSynchronize(c->stream));
  *result_out = h == kVerifyNone ? 0 : (int64_t)h + 1;
  return RCM_OK;
}

// After the engine placed the CM order in perm: the reversed order for RCM_FLAG_REVERSE, and the result flags.
int finish_order(rcm_ctx* c, uint32_t flags) {
  c->reversed = (flags & RCM_FLAG_REVERSE) != 0;
  if (c->reversed) {
    hipLaunchKernelGGL(k_reverse, dim3((c->n + 255) / 256), dim3(256), 0, c->stream, c->perm, c->n, c->out);
    RCM_HIP(hipStreamSynchronize(c->stream));
  }
  c->have_perm = true;
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


// ---------------------------------------------------------------------------------------------------------
// Many matrices: device-resident state of a set, one persistent launch per run.
// ---------------------------------------------------------------------------------------------------------
struct rcm_batch {
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev[4] = {};
  int64_t workers = 0;
  int64_t timeout_ms = 60000;
  int64_t sort_mode = 0;
  int64_t claim_filter = kDefaultClaimFilter;
  int64_t use_graph = 1;
  int64_t helper_nodes = kDefaultBatchHelperNodes;  // RCM_PARAM_HELPER_NODES
  int64_t speculate = kDefaultSpeculate;        // RCM_PARAM_SPECULATE
  std::vector<JobDesc> jobs_host;              // host copy of the loaded set's job descriptors
  int64_t variant = kDefaultBatchVariant;
  int max_resident = 0;              // of the current variant
  int max_res[kVariants] = {};       // co-resident workgroups of each variant
  // the loaded set
  int32_t count = 0;
  std::vector<int32_t> n;
  std::vector<int64_t> nnz, pn;  // pn[k] = sum_{i<k} n[i]
  std::vector<int32_t> order;    // all jobs, largest nnz first
  int64_t total_n = 0, total_nnz = 0;
  // device buffers: views into one arena (grown on demand, or fixed for a serving lane)
  char* arena = nullptr;
  size_t arena_cap = 0;
  bool arena_fixed = false;
  int32_t hcap = 0;  // matrices the pinned host mirrors hold
  int32_t *rp = nullptr, *ci = nullptr;
  int32_t *perm = nullptr, *level = nullptr, *aux = nullptr;
  uint64_t *qinfo = nullptr, *claim = nullptr, *bstat = nullptr, *brange = nullptr;
  Control* ctl = nullptr;
  JobDesc* jobs = nullptr;
  int32_t *d_order = nullptr, *d_starts = nullptr, *d_active = nullptr, *tail = nullptr;
  int64_t* d_nnz = nullptr;
  uint32_t* scratch = nullptr;  // 3 * count words: validation flags, max row length, bandwidth
  uint32_t* helpers = nullptr;  // count words: workgroups helping each matrix
  int32_t *blk_order = nullptr, *blk_active = nullptr;  // helper-kernel block maps of `order` / a sub-list
  int32_t nblk_order = 0;
  ManyControl* g = nullptr;
  // pinned host mirrors
  ManyControl* h_g = nullptr;  // the launch control of the last launch (its abort word: the watchdog fired)
  Control* h_ctl = nullptr;
  int32_t* h_tail = nullptr;
  // graph of the run launches (reset + reorder [+ reverse]), per reverse flag
  hipGraph_t graph[2] = {};
  hipGraphExec_t gexec[2] = {};
  // last result
  bool have_perm = false;
  bool reversed = false;
  std::vector<int32_t> ok_jobs;  // matrices whose last run completed
  std::vector<uint32_t> maxrow;  // longest row per matrix (from the load's validation)
  std::vector<int32_t> blk_order_h;  // host copy of the block map of `order`
  rcm_batch_timing_t timing{};
};

namespace {

// Helper-kernel grid of a list of matrices: about one workgroup per 16k rows + entries each, 1..256, as a
// prefix map (slot k owns blocks [blk[k], blk[k+1])). Returns the grid size.
int32_t build_blk(const rcm_batch* b, const int32_t* list, int32_t count, std::vector<int32_t>& blk) {
  blk.assign((size_t)count + 1, 0);
  for (int32_t i = 0; i < count; ++i) {
    const int32_t k = list[i];
    const int64_t w = ((int64_t)b->n[k] + b->nnz[k]) / 16384 + 1;
    blk[i + 1] = blk[i] + (int32_t)std::min<int64_t>(w, 256);
  }
  return blk[count];
}

// Byte layout of a loaded set in the arena: every buffer 256-byte aligned.
struct BatchLayout {
  size_t rp, ci, perm, level, aux, qinfo, claim, bstat, brange, ctl, jobs, order, starts, active, tail, nnz,
      scratch, helpers, blk_order, blk_active, g, total;
};
BatchLayout batch_layout(size_t rp_end, size_t ci_end, size_t N, size_t C) {
  auto al = [](size_t v) { return (v + 255) / 256 * 256; };
  BatchLayout L{};
  size_t o = 0;
  auto put = [&](size_t& f, size_t bytes) { f = o; o += al(std::max<size_t>(bytes, 1)); };
  put(L.rp, 4 * rp_end);
  put(L.ci, 4 * ci_end);
  put(L.perm, 4 * N);
  put(L.level, 4 * N);
  put(L.aux, 4 * N);
  put(L.qinfo, 8 * N);
  put(L.claim, 8 * N);
  put(L.bstat, 8 * (N + C));
  put(L.brange, 8 * (N + C));
  put(L.ctl, sizeof(Control) * C);
  put(L.jobs, sizeof(JobDesc) * C);
  put(L.order, 4 * C);
  put(L.starts, 4 * C);
  put(L.active, 4 * C);
  put(L.tail, 4 * 96 * C);
  put(L.nnz, 8 * C);
  put(L.scratch, 4 * 3 * C);
  put(L.helpers, 4 * C);
  put(L.blk_order, 4 * (C + 1));
  put(L.blk_active, 4 * (C + 1));
  put(L.g, sizeof(ManyControl));
  L.total = o;
  return L;
}

// Captured run graphs bake in the kernel, grid and arguments: drop them when any of those changes.
void batch_drop_graphs(rcm_batch* b) {
  for (int r = 0; r < 2; ++r) {
    if (b->gexec[r]) (void)hipGraphExecDestroy(b->gexec[r]);
    if (b->graph[r]) (void)hipGraphDestroy(b->graph[r]);
    b->gexec[r] = nullptr;
    b->graph[r] = nullptr;
  }
}

// Forget the loaded set (its captured graphs refer to its buffers); the memory is kept for the next load.
void batch_reset(rcm_batch* b) {
  batch_drop_graphs(b);
  b->count = 0;
  b->have_perm = false;
}

void batch_release(rcm_batch* b) {
  batch_reset(b);
  if (b->arena) (void)hipFree(b->arena);
  b->arena = nullptr;
  b->arena_cap = 0;
  if (b->h_ctl) (void)hipHostFree(b->h_ctl);
  if (b->h_tail) (void)hipHostFree(b->h_tail);
  b->h_ctl = nullptr;
  b->h_tail = nullptr;
  b->hcap = 0;
}

// Device arena of at least `bytes` (grown unless fixed; a grown arena gets 25% headroom, an `exact` one none)
// and pinned mirrors for `count` matrices.
int batch_reserve(rcm_batch* b, size_t bytes, int32_t count, bool exact = false) {
  if (bytes > b->arena_cap) {
    if (b->arena_fixed) return RCM_ENOMEM;
    if (b->arena) (void)hipFree(b->arena);
    b->arena = nullptr;
    b->arena_cap = 0;
    const size_t want = exact ? bytes : bytes + bytes / 4;
    if (hipMalloc(reinterpret_cast<void**>(&b->arena), want) != hipSuccess) return RCM_ENOMEM;
    b->arena_cap = want;
  }
  if (count > b->hcap) {
    if (b->h_ctl) (void)hipHostFree(b->h_ctl);
    if (b->h_tail) (void)hipHostFree(b->h_tail);
    b->h_ctl = nullptr;
    b->h_tail = nullptr;
    b->hcap = 0;
    if (hipHostMalloc(reinterpret_cast<void**>(&b->h_ctl), (size_t)count * sizeof(Control), hipHostMallocDefault) !=
            hipSuccess ||
        hipHostMalloc(reinterpret_cast<void**>(&b->h_tail), (size_t)count * 96 * sizeof(int32_t),
                      hipHostMallocDefault) != hipSuccess)
      return RCM_ENOMEM;
    b->hcap = count;
  }
  return RCM_OK;
}

ManyArgs batch_args(rcm_batch* b, const int32_t* d_order, int32_t count) {
  ManyArgs m{};
  m.jobs = b->jobs;
  m.helpers = b->helpers;
  m.order = d_order;
  m.starts = b->d_starts;
  m.count = count;
  m.g = b->g;
  m.budget_ticks = (uint64_t)b->timeout_ms * 100000ull;
  m.sort_mode = (int32_t)b->sort_mode;
  m.claim_filter = (int32_t)b->claim_filter;
  m.seg_threshold = kDefaultSegThreshold;
  return m;
}

int batch_workers(rcm_batch* b) {
  int w = (int)b->workers;
  if (w <= 0) w = b->max_resident;
  return std::max(1, std::min(w, b->max_resident));
}

// One pass over a list of matrices (d_list, count entries; d_blk is their block map of nblk blocks): the reset
// and the persistent many-matrix launch.
// `rows`: the listed matrices' total n, over which the launch's workgroups are shared by size (ManyArgs).
void enqueue_many(rcm_batch* b, const int32_t* d_list, const int32_t* d_blk, int32_t nblk, int32_t count, bool bfs,
                  bool levels, int64_t rows) {
  if (count <= 0) return;
  ManyArgs m = batch_args(b, d_list, count);
  m.share_workers = (uint32_t)batch_workers(b);
  m.share_rows = (uint64_t)std::max<int64_t>(rows, 0);
  hipLaunchKernelGGL(k_init_many, dim3((unsigned)nblk), dim3(256), 0, b->stream, m, d_blk);
  void* kern = many_kernel((int)b->variant, bfs, levels && !bfs);
  ManyArgs mm = m;
  void* args[] = {&mm};
  (void)hipLaunchKernel(kern, dim3(batch_workers(b)), dim3(kVarThreads[b->variant]), args, 0, b->stream);
}

// Per-matrix status from the control blocks of the last launch (jobs listed in `order`). A matrix the launch
// left unfinished because its watchdog fired (the launch's abort word) is RCM_ETIMEOUT.
int collect_status(rcm_batch* b, const int32_t* order, int32_t count, int32_t* status, int64_t* batches) {
  int worst = RCM_OK;
  int64_t nb = 0;
  const bool aborted = b->h_g->abort != 0;
  for (int32_t i = 0; i < count; ++i) {
    const int32_t k = order[i];
    const Control& h = b->h_ctl[k];
    int r = RCM_OK;
    if (h.error == ERR_TIMEOUT) r = RCM_ETIMEOUT;
    else if (h.error == ERR_CAPACITY || h.error == ERR_LONGROW) r = RCM_ECAPACITY;
    else if (h.error != ERR_NONE || h.nbatches == 0) r = aborted ? RCM_ETIMEOUT : RCM_EDEVICE;
    else if ((int32_t)h.placed != b->n[k]) r = RCM_ENOTCONNECTED;
    nb += h.nbatches;
    if (status) status[k] = r;
    if (r && worst == RCM_OK) worst = r;
  }
  if (batches) *batches = nb;
  return worst;
}

int batch_wait(rcm_batch* b, hipEvent_t ev) {
  const auto t0 = std::chrono::steady_clock::now();
  for (;;) {
    const hipError_t q = hipEventQuery(ev);
    if (q == hipSuccess) return RCM_OK;
    if (q != hipErrorNotReady) return RCM_EDEVICE;
    const auto ms = std::chrono::duration_cast<std::chrono::milliseconds>(std::chrono::steady_clock::now() - t0).count();
    if (ms > b->timeout_ms + 10000) return RCM_ETIMEOUT;
    std::this_thread::sleep_for(std::chrono::microseconds(50));
  }
}

}  // namespace


// ---------------------------------------------------------------------------------------------------------
// Serving: continuous batching over lanes. A lane is a many-matrix context (its own stream, a fixed share of
// the persistent workgroups and a fixed arena of HBM) driven by its own host thread. A free lane takes
// every waiting request that fits its arena (first come first served, looking 8 ahead), reorders them in
// one persistent launch (start search sweeps, then the run) and delivers the results; the lanes run at the
// same time, so requests that arrive while others are on the GPU start on the next free lane and overlap
// with them on the device. Arenas are allocated once: no allocation or free happens while kernels run.
// ---------------------------------------------------------------------------------------------------------
namespace {

struct ServerJob {
  int64_t id = 0;
  int32_t n = 0;
  int64_t nnz = 0;
  std::unique_ptr<int32_t[]> rp, ci;  // the submit's copies (left uninitialised before the copy)
  const int32_t* src_rp = nullptr;  // the CSR the group reads: the copies above, or the caller's (RCM_FLAG_BORROW)
  const int32_t* src_ci = nullptr;
  int32_t start_req = -1;
  uint32_t flags = 0;
  size_t bytes = 0;  // arena bytes of this matrix
  bool done = false;
  int status = RCM_OK;
  std::vector<int32_t> perm;
  int64_t bw = -1;
  int32_t start_used = -1, sweeps = 0, lane = -1;
  std::chrono::steady_clock::time_point t_submit, t_admit, t_done;
};

// Admission counts what rcm_batch_load will lay out: the arena bytes of a group are batch_layout's total over
// the group's summed row offsets, columns, rows and matrices.
struct GroupSize {
  size_t rp = 0, ci = 0, n = 0, count = 0;
  GroupSize plus(int32_t jn, int64_t jnnz) const {
    GroupSize g = *this;
    g.rp += (size_t)jn + 1;
    g.ci += (size_t)jnnz;
    g.n += (size_t)jn;
    g.count += 1;
    return g;
  }
  size_t bytes() const { return batch_layout(rp, ci, n, count).total; }
};

}  // namespace

struct rcm_server {
  int device = 0;
  rcm_server_config cfg{};
  int32_t lanes = 4;
  size_t lane_bytes = 0;
  std::vector<rcm_batch*> batch;
  std::vector<std::thread> threads;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<std::shared_ptr<ServerJob>> pending;
  std::unordered_map<int64_t, std::shared_ptr<ServerJob>> jobs;  // not yet released by a wait
  bool stopping = false;
  int64_t next_id = 1;
  int64_t inflight_bytes = 0;
  rcm_server_stats st{};
  // per lane, grow-only host staging for a group's concatenated CSRs and orders: reused, so a group does not
  // pay for fresh zero-filled pages; and the lane's pinned staging (kServePinnedWords, null if not allocated)
  std::vector<std::vector<int32_t>> stage_rp, stage_ci, stage_perm;
  std::vector<int32_t*> pinned;
  std::vector<rcm_ctx*> solo;  // per lane, the single-matrix context for large requests (null if not made)
};

namespace {

// One lane: reorder a group of requests with the lane's many-matrix context.
bool solo_fits(const rcm_server* s, int lane, const ServerJob& j) {
  return s->solo[(size_t)lane] && j.n >= kServeLargeN && j.n <= kServeSoloN && j.nnz <= kServeSoloNnz;
}

// One large request through the lane's single-matrix context: upload, start search and order, bandwidth, download.
void run_solo(rcm_server* s, int lane, ServerJob& j) {
  rcm_ctx* c = s->solo[(size_t)lane];
  using clk = std::chrono::steady_clock;
  const auto ms_since = [](clk::time_point t0) {
    return std::chrono::duration<double, std::milli>(clk::now() - t0).count();
  };
  auto t = clk::now();
  int rc = rcm_ctx_load(c, j.n, j.nnz, j.src_rp, j.src_ci);  // within the reserved buffers: no allocation
  const double load_ms = ms_since(t);
  t = clk::now();
  int32_t start = j.start_req, iters = 0;
  if (!rc) rc = rcm_ctx_order(c, j.start_req, RCM_FLAG_STABLE, &start, &iters);
  const double order_ms = ms_since(t);
  const int launches = rc ? 0 : c->timing.order_passes;
  t = clk::now();
  int64_t bw = -1;
  if (!rc && (j.flags & RCM_FLAG_BANDWIDTH)) rc = rcm_ctx_bandwidth(c, &bw);
  if (!rc) {
    j.perm.resize((size_t)j.n);
    rc = rcm_ctx_read_perm(c, j.perm.data());
  }
  const double read_ms = ms_since(t);
  j.start_used = start;
  j.sweeps = iters;
  j.status = rc;
  if (rc) {
    j.perm.clear();
  } else {
    if (j.flags & RCM_FLAG_REVERSE) std::reverse(j.perm.begin(), j.perm.end());
    if (j.flags & RCM_FLAG_BANDWIDTH) j.bw = bw;
  }
  std::lock_guard<std::mutex> lk(s->mu);
  s->st.group_load_ms += load_ms;
  s->st.group_order_ms += order_ms;
  s->st.group_readback_ms += read_ms;
  s->st.group_launches += launches;
}

void run_group(rcm_server* s, int lane, std::vector<std::shared_ptr<ServerJob>>& grp) {
  if (grp.size() == 1 && solo_fits(s, lane, *grp[0])) {
    run_solo(s, lane, *grp[0]);
    return;
  }
  rcm_batch* b = s->batch[(size_t)lane];
  const int32_t C = (int32_t)grp.size();
  using clk = std::chrono::steady_clock;
  const auto ms_since = [](clk::time_point t) { return std::chrono::duration<double, std::milli>(clk::now() - t).count(); };
  auto t = clk::now();
  double copy_ms = 0, load_ms = 0, order_ms = 0, read_ms = 0, deliver_ms = 0;
  std::vector<int32_t> n((size_t)C), starts((size_t)C, 0), st((size_t)C, RCM_OK), st2((size_t)C, RCM_OK);
  std::vector<int64_t> nnz((size_t)C), rp_off((size_t)C), ci_off((size_t)C);
  int64_t rpn = 0, cin = 0;
  bool want_bw = false;
  for (int32_t k = 0; k < C; ++k) {
    n[k] = grp[k]->n;
    nnz[k] = grp[k]->nnz;
    rp_off[k] = rpn;
    ci_off[k] = cin;
    rpn += n[k] + 1;
    cin += nnz[k];
    want_bw = want_bw || (grp[k]->flags & RCM_FLAG_BANDWIDTH);
  }
  int32_t* pin = s->pinned[(size_t)lane];
  const bool use_pin = pin && (size_t)(rpn + std::max<int64_t>(cin, 1)) <= kServePinnedWords;
  int32_t *rp_cat, *ci_cat;
  if (use_pin) {
    rp_cat = pin;
    ci_cat = pin + rpn;
  } else {
    std::vector<int32_t>& rv = s->stage_rp[(size_t)lane];
    std::vector<int32_t>& cv = s->stage_ci[(size_t)lane];
    if (rv.size() < (size_t)rpn) rv.resize((size_t)rpn);
    if (cv.size() < (size_t)std::max<int64_t>(cin, 1)) cv.resize((size_t)std::max<int64_t>(cin, 1));
    rp_cat = rv.data();
    ci_cat = cv.data();
  }
  // the concatenation: the group's 2C segments (each request's row offsets and columns) laid end to end, and
  // words [w0, w1) of that sequence copied by one thread; a large group is split over kCopyThreads host threads
  // by words, so one large request is split too (the lanes' threads and these stay within the 16 cores a GPU's
  // host share has)
  auto copy_words = [&](int64_t w0, int64_t w1) {
    int64_t base = 0;
    for (int32_t k = 0; k < C && base < w1; ++k) {
      const int32_t* src[2] = {grp[k]->src_rp, grp[k]->src_ci};
      int32_t* dst[2] = {rp_cat + rp_off[k], ci_cat + ci_off[k]};
      const int64_t len[2] = {(int64_t)n[k] + 1, nnz[k]};
      for (int h = 0; h < 2; ++h) {
        const int64_t a = std::max(w0, base), e = std::min(w1, base + len[h]);
        if (a < e) std::copy(src[h] + (a - base), src[h] + (e - base), dst[h] + (a - base));
        base += len[h];
      }
    }
  };
  constexpr int kCopyThreads = 4;
  const int64_t words = rpn + cin;
  if (words >= ((int64_t)1 << 18)) {
    std::vector<std::thread> th;
    for (int p = 1; p < kCopyThreads; ++p)
      th.emplace_back(copy_words, words * p / kCopyThreads, words * (p + 1) / kCopyThreads);
    copy_words(0, words / kCopyThreads);
    for (auto& x : th) x.join();
  } else {
    copy_words(0, words);
  }
  copy_ms = ms_since(t);
  t = clk::now();
  int hard = rcm_batch_load(b, C, n.data(), nnz.data(), rp_off.data(), ci_off.data(), rp_cat, ci_cat, st.data());
  load_ms = ms_since(t);
  std::vector<int32_t> iters((size_t)C, 0);
  if (hard == RCM_EINVAL) {  // some request is malformed: report it, run the rest without it
    std::vector<std::shared_ptr<ServerJob>> good, bad;
    for (int32_t k = 0; k < C; ++k) (st[k] ? bad : good).push_back(grp[k]);
    for (auto& jp : bad) {
      jp->status = RCM_EINVAL;
      jp->perm.clear();
    }
    if (!good.empty()) run_group(s, lane, good);
    grp = good;  // the caller finishes every request of the group, malformed ones included
    grp.insert(grp.end(), bad.begin(), bad.end());
    return;
  }
  t = clk::now();
  int launches = 0;
  if (!hard) {  // start search (for requests without a start) and order, in one call
    for (int32_t k = 0; k < C; ++k) starts[k] = grp[k]->start_req;
    const int r = rcm_batch_order(b, starts.data(), RCM_FLAG_STABLE, starts.data(), iters.data(), st2.data());
    if (r == RCM_EDEVICE || r == RCM_ETIMEOUT || r == RCM_ENOTLOADED || r == RCM_EINVAL) hard = r;
    launches = b->timing.order_passes;
  }
  order_ms = ms_since(t);
  t = clk::now();
  std::vector<int64_t> bw((size_t)C, -1);
  // the orders: into the pinned staging when they fit (the load above has completed, so it is free again)
  int32_t* cat = nullptr;
  if (!hard && want_bw) hard = rcm_batch_bandwidth(b, bw.data());
  if (!hard) {
    const size_t words = (size_t)std::max<int64_t>(b->total_n, 1);
    if (pin && words <= kServePinnedWords) {
      cat = pin;
    } else {
      std::vector<int32_t>& pv = s->stage_perm[(size_t)lane];
      if (pv.size() < words) pv.resize(words);
      cat = pv.data();
    }
    hard = rcm_batch_read_perm(b, nullptr, cat);
  }
  read_ms = ms_since(t);
  t = clk::now();
  int64_t off = 0;
  for (int32_t k = 0; k < C; ++k) {
    ServerJob& j = *grp[k];
    j.start_used = starts[k];
    j.sweeps = iters[k];
    j.status = hard ? hard : (st[k] ? st[k] : st2[k]);
    if (j.status == RCM_OK) {
      j.perm.assign(cat + off, cat + off + j.n);
      if (j.flags & RCM_FLAG_REVERSE) std::reverse(j.perm.begin(), j.perm.end());
      if (j.flags & RCM_FLAG_BANDWIDTH) j.bw = bw[k];
    }
    off += j.n;
  }
  deliver_ms = ms_since(t);
  std::lock_guard<std::mutex> lk(s->mu);
  s->st.group_host_copy_ms += copy_ms;
  s->st.group_load_ms += load_ms;
  s->st.group_order_ms += order_ms;
  s->st.group_readback_ms += read_ms;
  s->st.group_deliver_ms += deliver_ms;
  s->st.group_launches += launches;
}

void lane_loop(rcm_server* s, int lane) {
  (void)hipSetDevice(s->device);
  for (;;) {
    std::vector<std::shared_ptr<ServerJob>> grp;
    {
      std::unique_lock<std::mutex> lk(s->mu);
      s->cv_work.wait(lk, [&] { return s->stopping || !s->pending.empty(); });
      if (s->pending.empty()) return;  // stopping, nothing left
      // a group: first come first served, looking at most 8 requests ahead for ones that fit the lane and are
      // of the first request's size class (requests of kServeLargeN rows or more are grouped only with each
      // other, so small requests do not wait behind them); at most a lanes-th of what is waiting, so that a
      // burst spreads over every lane
      const bool large = s->pending.front()->n >= kServeLargeN;
      const size_t share = (s->pending.size() + (size_t)s->lanes - 1) / (size_t)s->lanes;
      // a large request that fits the lane's single-matrix context is ordered alone through it
      const size_t cap = solo_fits(s, lane, *s->pending.front())
                             ? 1
                             : std::min<size_t>((size_t)s->cfg.slots, std::max<size_t>(1, share));
      GroupSize gs;
      size_t skipped = 0;
      // with pinned staging, a group stays within it (a request larger than it goes alone, through pageable
      // staging): its upload and download then run as direct DMA
      const bool pinned = s->pinned[(size_t)lane] != nullptr;
      for (size_t i = 0; i < s->pending.size() && skipped < 8 && grp.size() < cap;) {
        const auto& jp = s->pending[i];
        const GroupSize next = gs.plus(jp->n, jp->nnz);
        const bool fits_pin = !pinned || grp.empty() || next.rp + std::max<size_t>(next.ci, 1) <= kServePinnedWords;
        if ((jp->n >= kServeLargeN) == large && next.bytes() <= s->lane_bytes && fits_pin) {
          gs = next;
          grp.push_back(jp);
          s->pending.erase(s->pending.begin() + (std::ptrdiff_t)i);
        } else {
          ++i;
          ++skipped;
        }
      }
      const auto now = std::chrono::steady_clock::now();
      for (auto& jp : grp) {
        jp->t_admit = now;
        jp->lane = lane;
      }
      s->st.admitted += (int64_t)grp.size();
      s->st.inflight += (int32_t)grp.size();
      s->st.max_inflight = std::max(s->st.max_inflight, s->st.inflight);
      s->st.pending = (int32_t)s->pending.size();
      const size_t used = grp.empty() ? 0 : gs.bytes();
      s->inflight_bytes += (int64_t)used;
      s->st.arena_peak_bytes = std::max(s->st.arena_peak_bytes, s->inflight_bytes);
      if (grp.empty()) continue;  // cannot happen: every request fits an empty lane (checked at submit)
      ++s->st.loop_iterations;
      const size_t u = used;
      lk.unlock();
      run_group(s, lane, grp);
      lk.lock();
      s->inflight_bytes -= (int64_t)u;
      const auto t = std::chrono::steady_clock::now();
      for (auto& jp : grp) {
        jp->t_done = t;
        jp->done = true;
        jp->rp.reset();
        jp->ci.reset();
        jp->src_rp = jp->src_ci = nullptr;
        if (jp->status == RCM_OK) ++s->st.completed;
        else ++s->st.failed;
      }
      s->st.inflight -= (int32_t)grp.size();
      s->cv_done.notify_all();
    }
  }
}

}  // namespace

// ---------------------------------------------------------------------------------------------------------
// The one-shot reorder that allocates nothing (rcm_gpu_batch_noalloc): rcm_gpu_batch on a CSR already in device
// memory, with every buffer the engine needs beyond the caller's input and output in one file-scope device
// array, `scratch`. Its size is fixed: matrices of at most kNoAllocMaxN rows. Each device holds its own copy
// until the process ends; its address is looked up once per device with hipGetSymbolAddress and only that
// address is used on the host.
// ---------------------------------------------------------------------------------------------------------
namespace {

constexpr int32_t kNoAllocMaxN = 1 << 21;  // 2,097,152 rows (the 128^3 stencil)
constexpr int kNoAllocDevices = 64;

// Byte offsets of the engine buffers of an n-row matrix in the array, each 256-byte aligned:
// claim u64[n] (the inverse permutation for the bandwidth after the run), qinfo u64[n], bstat u64[n+1],
// brange u64[n+1], level i32[n], the control block, 16 words (validation flags and longest row, bandwidth).
struct NoAllocLayout {
  size_t claim, qinfo, bstat, brange, level, ctl, words, total;
};
constexpr size_t al256(size_t v) { return (v + 255) / 256 * 256; }
constexpr NoAllocLayout noalloc_layout(size_t n) {
  NoAllocLayout L{};
  L.claim = 0;
  L.qinfo = L.claim + al256(8 * n);
  L.bstat = L.qinfo + al256(8 * n);
  L.brange = L.bstat + al256(8 * (n + 1));
  L.level = L.brange + al256(8 * (n + 1));
  L.ctl = L.level + al256(4 * n);
  L.words = L.ctl + al256(sizeof(Control));
  L.total = L.words + al256(16 * sizeof(uint32_t));
  return L;
}
constexpr size_t kNoAllocScratchBytes = noalloc_layout((size_t)kNoAllocMaxN).total;  // 72 MiB + 1 KiB

}  // namespace

static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {

struct NoAllocDevice {
  std::mutex mu;                  // calls of this process on one device run one at a time
  unsigned char* base = nullptr;  // the device address of `scratch` on this device
  int max_resident = 0;           // co-resident workgroups of the variant-0 reorder kernel
};
NoAllocDevice g_noalloc[kNoAllocDevices];

__global__ void k_fill_u32(uint32_t* p, int32_t count, uint32_t v) {
  const int32_t i = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
  if (i < count) p[i] = v;
}

// perm reversed in place: thread i swaps entries i and n - 1 - i.
__global__ void k_reverse_inplace(int32_t* perm, int32_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n / 2) {
    const int32_t a = perm[i], b = perm[n - 1 - i];
    perm[i] = b;
    perm[n - 1 - i] = a;
  }
}

// One engine run (variant 0) from `start` over the array's buffers, queue in `a.perm`; waits for it and
// returns its status.
int noalloc_run(const NoAllocDevice& D, RunArgs a, int32_t start, bool bfs) {
  const int32_t n = a.n;
  const int init_blocks = (int)std::min<int64_t>(2048, ((int64_t)n + 1 + 255) / 256);
  hipLaunchKernelGGL(k_init, dim3(init_blocks), dim3(256), 0, nullptr, a, start, n + 1);  // every record cleared
  const int workers = std::max(1, std::min(n > 16384 ? 256 : 64, D.max_resident));
  void* kern = reorder_kernel(0, false, bfs, false);
  void* args[] = {&a};
  RCM_HIP(hipLaunchKernel(kern, dim3(workers), dim3(kVarThreads[0]), args, 0, nullptr));
  RCM_HIP(hipGetLastError());
  Control h{};
  RCM_HIP(hipMemcpy(&h, a.ctl, sizeof(Control), hipMemcpyDeviceToHost));
  if (h.error == ERR_TIMEOUT) return RCM_ETIMEOUT;
  if (h.error == ERR_CAPACITY || h.error == ERR_LONGROW) return RCM_ECAPACITY;
  if (h.error != ERR_NONE || h.nbatches == 0) return RCM_EDEVICE;
  if ((int32_t)h.placed != n) return RCM_ENOTCONNECTED;
  return RCM_OK;
}

}  // namespace

extern "C" {

// A plain device copy (16 bytes per thread per step, grid-stride): the HBM bandwidth this GPU reaches on
// streaming reads and writes, for comparison with the roofline peak. (Four loads in flight per thread measured
// lower: 4.3 against 4.8 TB/s for 1 GiB.)
__global__ __launch_bounds__(256) void k_copy_words(const uint4* __restrict__ src, uint4* __restrict__ dst,
                                                   int64_t n16) {
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n16; i += (int64_t)gridDim.x * 256) dst[i] = src[i];
}

int rcm_selftest_primitives(int device, const uint32_t* in, uint32_t* out) {
  if (!in || !out) return RCM_EINVAL;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev) return RCM_EDEVICE;
  RCM_HIP(hipSetDevice(device));
  uint32_t* d = nullptr;
  if (dev_alloc(&d, kSelftestIn + kSelftestOut) != hipSuccess) return RCM_ENOMEM;
  int rc = RCM_OK;
  if (hipMemcpy(d, in, kSelftestIn * sizeof(uint32_t), hipMemcpyHostToDevice) != hipSuccess) rc = RCM_EDEVICE;
  if (rc == RCM_OK) {
    hipLaunchKernelGGL(k_selftest, dim3(1), dim3(kSelftestThreads), 0, nullptr, (const uint32_t*)d, d + kSelftestIn);
    if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess ||
        hipMemcpy(out, d + kSelftestIn, kSelftestOut * sizeof(uint32_t), hipMemcpyDeviceToHost) != hipSuccess)
      rc = RCM_EDEVICE;
  }
  dev_free(d);
  return rc;
}

int rcm_selftest_copy_bandwidth(int device, int64_t bytes, int32_t reps, int32_t mode, double* gbs_out,
                                double* ms_out) {
  if (!gbs_out || bytes < 16 || bytes % 16 != 0 || bytes > ((int64_t)1 << 36) || reps < 1 || reps > 1000 ||
      mode < 0 || mode > 1)
    return RCM_EINVAL;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev) return RCM_EDEVICE;
  RCM_HIP(hipSetDevice(device));
  uint4 *src = nullptr, *dst = nullptr;
  const int64_t n16 = bytes / 16;
  if (dev_alloc(&src, (size_t)n16) != hipSuccess) return RCM_ENOMEM;
  if (dev_alloc(&dst, (size_t)n16) != hipSuccess) {
    dev_free(src);
    return RCM_ENOMEM;
  }
  hipStream_t st = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = RCM_OK;
  if (hipStreamCreate(&st) != hipSuccess || hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)
    rc = RCM_EDEVICE;
  const unsigned grid = 2048;  // 8 workgroups of 256 per CU: each thread copies bytes / (2048 * 256 * 16) words
  float ms = 0.f;
  if (rc == RCM_OK) {
    auto copy = [&] {
      if (mode == 0)
        hipLaunchKernelGGL(k_copy_words, dim3(grid), dim3(256), 0, st, src, dst, n16);
      else if (hipMemcpyAsync(dst, src, (size_t)bytes, hipMemcpyDeviceToDevice, st) != hipSuccess)
        rc = RCM_EDEVICE;
    };
    for (int w = 0; w < 2; ++w) copy();
    (void)hipEventRecord(e0, st);
    for (int r = 0; r < reps; ++r) copy();
    (void)hipEventRecord(e1, st);
    if (hipGetLastError() != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
        hipEventElapsedTime(&ms, e0, e1) != hipSuccess || ms <= 0.f || rc != RCM_OK)
      rc = RCM_EDEVICE;
  }
  if (rc == RCM_OK) {
    const double per = (double)ms / reps;
    *gbs_out = 2.0 * (double)bytes / (per * 1e-3) / 1e9;
    if (ms_out) *ms_out = per;
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (st) (void)hipStreamDestroy(st);
  dev_free(src);
  dev_free(dst);
  return rc;
}

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
    case RCM_ECANCELED: return "request canceled: the server shut down";
    case RCM_ENOTFOUND: return "unknown request id";
    case RCM_EAGAIN: return "request not finished yet";
    case RCM_EIO: return "matrix file could not be opened, read or written";
    case RCM_ENOTSYMMETRIC: return "matrix is not structurally symmetric";
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
  ok = ok && hipHostMalloc(reinterpret_cast<void**>(&c->h_word), rcm_mx::kPinnedWords * sizeof(uint32_t),
                            hipHostMallocDefault) == hipSuccess;
  hipDeviceProp_t prop;
  ok = ok && hipGetDeviceProperties(&prop, device) == hipSuccess;
  for (int v = 0; ok && v < kVariants; ++v) {
    // persistent workers must all be resident at once: the grid is capped by the kernel's occupancy
    int per_cu = 0;
    ok = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reorder_kernel(v, false, false, false), kVarThreads[v],
                                                      0) == hipSuccess;
    c->max_resident[v] = std::max(1, per_cu) * prop.multiProcessorCount;
  }
  if (!ok) {
    rcm_ctx_destroy(c);
    return RCM_EDEVICE;
  }
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
  dev_free(c->trace);
  if (c->h_ctl) (void)hipHostFree(c->h_ctl);
  if (c->h_word) (void)hipHostFree(c->h_word);
  c->lscratch.release();
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
    case RCM_PARAM_CLAIM_FILTER: if (v < 0 || v > 2) return RCM_EINVAL; c->claim_filter = v; return RCM_OK;
    case RCM_PARAM_DIGIT_BITS:  // part of the variant
      return v == (c->variant == kLevelVariant ? kLevelDigit : kVarDigit[c->variant]) ? RCM_OK : RCM_EINVAL;
    case RCM_PARAM_VARIANT: if (v < 0 || v > kLevelVariant) return RCM_EINVAL; c->variant = v; return RCM_OK;
    case RCM_PARAM_STRICT: c->strict = v ? 1 : 0; return RCM_OK;
    case RCM_PARAM_PROFILE: if (v < 0 || v > 2) return RCM_EINVAL; c->profile = v; return RCM_OK;
    case RCM_PARAM_BATCH_CAP: if (v < 0) return RCM_EINVAL; c->batch_cap = v; return RCM_OK;
    case RCM_PARAM_SPECULATE: if (v < 0) return RCM_EINVAL; c->speculate = v; return RCM_OK;
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

int rcm_ctx_trace(rcm_ctx* c, uint64_t* out, int64_t words) {
  if (!c || !out || words <= 0) return RCM_EINVAL;
  if (!c->trace) return RCM_ENOTLOADED;
  RCM_HIP(hipSetDevice(c->device));
  const int64_t w = std::min<int64_t>(words, c->trace_cap * kTraceWords);
  RCM_HIP(hipMemcpy(out, c->trace, (size_t)w * sizeof(uint64_t), hipMemcpyDeviceToHost));
  return RCM_OK;
}

int rcm_ctx_get_param(rcm_ctx* c, int param, int64_t* v) {
  if (!c || !v) return RCM_EINVAL;
  switch (param) {
    case RCM_PARAM_WORKERS: *v = c->workers; return RCM_OK;
    case RCM_PARAM_TIMEOUT_MS: *v = c->timeout_ms; return RCM_OK;
    case RCM_PARAM_SORT: *v = c->sort_mode; return RCM_OK;
    case RCM_PARAM_CLAIM_FILTER: *v = c->claim_filter; return RCM_OK;
    case RCM_PARAM_DIGIT_BITS: *v = c->variant == kLevelVariant ? kLevelDigit : kVarDigit[c->variant]; return RCM_OK;
    case RCM_PARAM_VARIANT: *v = c->variant; return RCM_OK;
    case RCM_PARAM_STRICT: *v = c->strict; return RCM_OK;
    case RCM_PARAM_PROFILE: *v = c->profile; return RCM_OK;
    case RCM_PARAM_BATCH_CAP: *v = c->batch_cap; return RCM_OK;
    case RCM_PARAM_SPECULATE: *v = c->speculate; return RCM_OK;
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
  const UpSeg seg[2] = {{c->rp, row_ptr, ((size_t)n + 1) * sizeof(int32_t)},
                        {c->ci, col_idx, (size_t)nnz * sizeof(int32_t)}};
  c->n = 0;
  if (const int up = upload_host(c->device, c->stream, seg, 2)) return up;
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
  return finish_order(c, flags);
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
    int32_t nodes[32], dists[32], degs[32];
    uint64_t rows[32];
    for (;;) {
      ++iters;
      c->have_perm = false;
      const int rc = run_engine(c, cur, true);
      if (rc) return rc;
      RCM_HIP(hipMemcpyAsync(nodes, c->perm + (n - w), (size_t)w * sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
      RCM_HIP(hipMemcpyAsync(dists, c->level + (n - w), (size_t)w * sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
      RCM_HIP(hipMemcpyAsync(rows, c->qinfo + (n - w), (size_t)w * sizeof(uint64_t), hipMemcpyDeviceToHost, c->stream));
      RCM_HIP(hipStreamSynchronize(c->stream));
      for (int32_t i = 0; i < w; ++i) degs[i] = (int32_t)((uint32_t)(rows[i] >> 32) - (uint32_t)rows[i]);
      if (peripheral_step(nodes, dists, degs, w, &lastdist, &cur)) break;
    }
  }
  c->timing.start_search_ms =
      std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  c->timing.start_iterations = iters;
  *start_out = cur;
  if (iterations_out) *iterations_out = iters;
  return RCM_OK;
}

// Start search and order in one call: the sweeps of SimplePeripheral::findPeripheral (src/simple_peripheral.cpp:
// 92-141), except that from sweep `speculate` on a sweep is first run as a CM run that records levels. The level
// of its last queue entry is the eccentricity of that start in any BFS order, so when it equals the previous
// sweep's the rule stops at this start and the run is the result (one engine pass fewer than rcm_ctx_find_start
// + rcm_ctx_run); otherwise the plain sweep from the same start follows and the search goes on. Same start,
// sweep count and order as those two calls.
int rcm_ctx_order(rcm_ctx* c, int32_t start_node, uint32_t flags, int32_t* start_out, int32_t* iterations_out) {
  if (!c) return RCM_EINVAL;
  if (!c->rp || c->n == 0) return RCM_ENOTLOADED;
  if (start_node >= c->n) return RCM_EINVAL;
  RCM_HIP(hipSetDevice(c->device));
  const auto t0 = std::chrono::steady_clock::now();
  const int32_t n = c->n;
  c->have_perm = false;
  int32_t cur = start_node, iters = 0, passes = 0;
  bool ordered = false;
  if (start_node < 0 && n == 1) {
    cur = 0;
    iters = 1;
  } else if (start_node < 0) {
    cur = n / 2;
    uint32_t lastdist = 0;
    const int32_t w = std::min<int32_t>(32, n - 1);
    const bool can_spec = c->speculate > 0 && c->variant != kLevelVariant;
    int32_t nodes[32], dists[32], degs[32];
    uint64_t rows[32];
    for (;;) {
      ++iters;
      if (can_spec && iters >= c->speculate) {
        const int rc = run_engine(c, cur, false, true);
        ++passes;
        if (rc) return rc;
        int32_t ecc = -1;
        RCM_HIP(hipMemcpyAsync(&ecc, c->level + (n - 1), sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
        RCM_HIP(hipStreamSynchronize(c->stream));
        if ((uint32_t)ecc == lastdist) {
          ordered = true;
          break;
        }
      }
      const int rc = run_engine(c, cur, true);
      ++passes;
      if (rc) return rc;
      RCM_HIP(hipMemcpyAsync(nodes, c->perm + (n - w), (size_t)w * sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
      RCM_HIP(hipMemcpyAsync(dists, c->level + (n - w), (size_t)w * sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
      RCM_HIP(hipMemcpyAsync(rows, c->qinfo + (n - w), (size_t)w * sizeof(uint64_t), hipMemcpyDeviceToHost, c->stream));
      RCM_HIP(hipStreamSynchronize(c->stream));
      for (int32_t i = 0; i < w; ++i) degs[i] = (int32_t)((uint32_t)(rows[i] >> 32) - (uint32_t)rows[i]);
      if (peripheral_step(nodes, dists, degs, w, &lastdist, &cur)) break;
    }
  }
  if (!ordered) {
    const int rc = run_engine(c, cur, false);
    ++passes;
    if (rc) return rc;
  }
  const int rc = finish_order(c, flags);
  if (rc) return rc;
  c->timing.start_iterations = iters;
  c->timing.order_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  c->timing.order_passes = passes;
  if (start_out) *start_out = cur;
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

int rcm_ctx_verify(rcm_ctx* c, int32_t start, const int32_t* perm, uint32_t flags, int64_t* result_out) {
  if (!c || !result_out) return RCM_EINVAL;
  if (!c->rp || c->n == 0) return RCM_ENOTLOADED;
  if (start < 0 || start >= c->n) return RCM_EINVAL;
  RCM_HIP(hipSetDevice(c->device));
  const int32_t n = c->n;
  if (!perm) {  // the context's last order: its CM form is in perm
    if (!c->have_perm) return RCM_ENOTLOADED;
    return verify_device(c, c->perm, start, false, result_out);
  }
  c->have_perm = false;  // the external order takes the buffer of the reversed result
  RCM_HIP(hipMemcpyAsync(c->out, perm, (size_t)n * sizeof(int32_t), hipMemcpyHostToDevice, c->stream));
  if (flags & RCM_FLAG_REVERSE)
    hipLaunchKernelGGL(k_vfy_reverse, dim3((unsigned)((n / 2 + 255) / 256 + 1)), dim3(256), 0, c->stream, c->out, n);
  return verify_device(c, c->out, start, false, result_out);
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
  if (!rc) rc = rcm_ctx_order(c, start_node, flags, &start, nullptr);
  if (!rc && bandwidth_out && (flags & RCM_FLAG_BANDWIDTH)) rc = rcm_ctx_bandwidth(c, bandwidth_out);
  if (!rc) rc = rcm_ctx_read_perm(c, perm_out);
  if (!rc && start_node_out) *start_node_out = start;
  rcm_ctx_destroy(c);
  return rc;
}

int rcm_batch_create(int device, rcm_batch** out) {
  if (!out) return RCM_EINVAL;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev) return RCM_EDEVICE;
  RCM_HIP(hipSetDevice(device));
  rcm_batch* b = new (std::nothrow) rcm_batch();
  if (!b) return RCM_ENOMEM;
  b->device = device;
  bool ok = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) == hipSuccess;
  for (auto& e : b->ev) ok = ok && hipEventCreate(&e) == hipSuccess;
  ok = ok && hipHostMalloc(reinterpret_cast<void**>(&b->h_g), sizeof(ManyControl), hipHostMallocDefault) == hipSuccess;
  if (ok) *b->h_g = ManyControl{};
  hipDeviceProp_t prop;
  ok = ok && hipGetDeviceProperties(&prop, device) == hipSuccess;
  for (int v = 0; ok && v < kVariants; ++v) {
    int per_cu = 0;
    ok = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, many_kernel(v, false), kVarThreads[v], 0) ==
         hipSuccess;
    b->max_res[v] = std::max(1, per_cu) * prop.multiProcessorCount;
  }
  if (!ok) {
    rcm_batch_destroy(b);
    return RCM_EDEVICE;
  }
  b->max_resident = b->max_res[b->variant];
  *out = b;
  return RCM_OK;
}

int rcm_batch_destroy(rcm_batch* b) {
  if (!b) return RCM_EINVAL;
  (void)hipSetDevice(b->device);
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  batch_release(b);
  if (b->h_g) (void)hipHostFree(b->h_g);
  for (auto& e : b->ev)
    if (e) (void)hipEventDestroy(e);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
  return RCM_OK;
}

int rcm_batch_set_param(rcm_batch* b, int param, int64_t v) {
  if (!b) return RCM_EINVAL;
  switch (param) {
    case RCM_PARAM_WORKERS: if (v < 0) return RCM_EINVAL; b->workers = v; batch_drop_graphs(b); return RCM_OK;
    case RCM_PARAM_TIMEOUT_MS: if (v <= 0) return RCM_EINVAL; b->timeout_ms = v; batch_drop_graphs(b); return RCM_OK;
    case RCM_PARAM_SORT: if (v < 0 || v > 2) return RCM_EINVAL; b->sort_mode = v; batch_drop_graphs(b); return RCM_OK;
    case RCM_PARAM_CLAIM_FILTER:
      if (v < 0 || v > 2) return RCM_EINVAL;
      b->claim_filter = v;
      batch_drop_graphs(b);
      return RCM_OK;
    case RCM_PARAM_GRAPH: b->use_graph = v ? 1 : 0; return RCM_OK;
    case RCM_PARAM_SPECULATE: if (v < 0) return RCM_EINVAL; b->speculate = v; return RCM_OK;
    case RCM_PARAM_WG_MAX_N:  // reserved: the one-workgroup-per-matrix engine is not part of the library
      return v == 0 ? RCM_OK : RCM_EINVAL;
    case RCM_PARAM_VARIANT:
      if (v < 0 || v >= kVariants) return RCM_EINVAL;
      b->variant = v;
      batch_drop_graphs(b);  // captured graphs hold the old launch parameters
      b->max_resident = b->max_res[v];
      return RCM_OK;
    case RCM_PARAM_HELPER_NODES: {
      if (v < 1) return RCM_EINVAL;
      b->helper_nodes = v;
      batch_drop_graphs(b);
      if (b->count == 0 || b->jobs_host.empty()) return RCM_OK;
      for (size_t k = 0; k < b->jobs_host.size(); ++k) b->jobs_host[k].cap = helper_cap(b->jobs_host[k].n, v);
      RCM_HIP(hipMemcpyAsync(b->jobs, b->jobs_host.data(), b->jobs_host.size() * sizeof(JobDesc),
                             hipMemcpyHostToDevice, b->stream));
      RCM_HIP(hipStreamSynchronize(b->stream));
      return RCM_OK;
    }
    default: return RCM_EINVAL;
  }
}

int rcm_batch_load(rcm_batch* b, int32_t count, const int32_t* n, const int64_t* nnz, const int64_t* rp_off,
                   const int64_t* ci_off, const int32_t* row_ptr_cat, const int32_t* col_idx_cat, int32_t* status_out) {
  if (!b || count < 1 || !n || !nnz || !rp_off || !ci_off || !row_ptr_cat || !col_idx_cat) return RCM_EINVAL;
  int64_t rp_end = 0, ci_end = 0, tot = 0, totnnz = 0;
  for (int32_t k = 0; k < count; ++k) {
    if (n[k] < 1 || n[k] >= (1 << 30) || nnz[k] < 0 || nnz[k] > INT32_MAX || rp_off[k] < 0 || ci_off[k] < 0)
      return RCM_EINVAL;
    rp_end = std::max(rp_end, rp_off[k] + n[k] + 1);
    ci_end = std::max(ci_end, ci_off[k] + nnz[k]);
    tot += n[k];
    totnnz += nnz[k];
  }
  RCM_HIP(hipSetDevice(b->device));
  const auto t0 = std::chrono::steady_clock::now();
  RCM_HIP(hipStreamSynchronize(b->stream));
  batch_reset(b);
  const size_t N = (size_t)tot, C = (size_t)count;
  const BatchLayout L = batch_layout((size_t)rp_end, (size_t)ci_end, N, C);
  const int ar = batch_reserve(b, L.total, count);
  if (ar) return ar;
  char* base = b->arena;
  b->rp = reinterpret_cast<int32_t*>(base + L.rp);
  b->ci = reinterpret_cast<int32_t*>(base + L.ci);
  b->perm = reinterpret_cast<int32_t*>(base + L.perm);
  b->level = reinterpret_cast<int32_t*>(base + L.level);
  b->aux = reinterpret_cast<int32_t*>(base + L.aux);
  b->qinfo = reinterpret_cast<uint64_t*>(base + L.qinfo);
  b->claim = reinterpret_cast<uint64_t*>(base + L.claim);
  b->bstat = reinterpret_cast<uint64_t*>(base + L.bstat);
  b->brange = reinterpret_cast<uint64_t*>(base + L.brange);
  b->ctl = reinterpret_cast<Control*>(base + L.ctl);
  b->jobs = reinterpret_cast<JobDesc*>(base + L.jobs);
  b->d_order = reinterpret_cast<int32_t*>(base + L.order);
  b->d_starts = reinterpret_cast<int32_t*>(base + L.starts);
  b->d_active = reinterpret_cast<int32_t*>(base + L.active);
  b->tail = reinterpret_cast<int32_t*>(base + L.tail);
  b->d_nnz = reinterpret_cast<int64_t*>(base + L.nnz);
  b->scratch = reinterpret_cast<uint32_t*>(base + L.scratch);
  b->helpers = reinterpret_cast<uint32_t*>(base + L.helpers);
  b->blk_order = reinterpret_cast<int32_t*>(base + L.blk_order);
  b->blk_active = reinterpret_cast<int32_t*>(base + L.blk_active);
  b->g = reinterpret_cast<ManyControl*>(base + L.g);
  b->count = count;
  b->n.assign(n, n + count);
  b->nnz.assign(nnz, nnz + count);
  b->pn.resize(C);
  std::vector<JobDesc> jobs(C);
  int64_t po = 0, pb = 0;
  for (int32_t k = 0; k < count; ++k) {
    b->pn[k] = po;
    JobDesc& J = jobs[k];
    J.row_ptr = b->rp + rp_off[k];
    J.col_idx = b->ci + ci_off[k];
    J.perm = b->perm + po;
    J.level = b->level + po;
    J.aux = b->aux + po;
    J.qinfo = b->qinfo + po;
    J.claim = b->claim + po;
    J.bstat = b->bstat + pb;
    J.brange = b->brange + pb;
    J.ctl = b->ctl + k;
    J.n = n[k];
    J.cap = helper_cap(n[k], b->helper_nodes);
    po += n[k];
    pb += n[k] + 1;
  }
  b->total_n = tot;
  b->total_nnz = totnnz;
  // begin the largest matrices first: their chains are the longest
  b->order.resize(C);
  for (int32_t k = 0; k < count; ++k) b->order[k] = k;
  std::stable_sort(b->order.begin(), b->order.end(), [&](int32_t x, int32_t y) { return nnz[x] > nnz[y]; });
  if (b->arena_fixed) {  // a serving lane: its groups are staged in pinned memory already
    RCM_HIP(hipMemcpyAsync(b->rp, row_ptr_cat, (size_t)rp_end * sizeof(int32_t), hipMemcpyHostToDevice, b->stream));
    if (ci_end > 0)
      RCM_HIP(hipMemcpyAsync(b->ci, col_idx_cat, (size_t)ci_end * sizeof(int32_t), hipMemcpyHostToDevice, b->stream));
  } else {
    const UpSeg seg[2] = {{b->rp, row_ptr_cat, (size_t)rp_end * sizeof(int32_t)},
                          {b->ci, col_idx_cat, (size_t)ci_end * sizeof(int32_t)}};
    if (const int up = upload_host(b->device, b->stream, seg, 2)) return up;
  }
  RCM_HIP(hipMemcpyAsync(b->jobs, jobs.data(), C * sizeof(JobDesc), hipMemcpyHostToDevice, b->stream));
  b->jobs_host = jobs;
  RCM_HIP(hipMemcpyAsync(b->d_order, b->order.data(), C * sizeof(int32_t), hipMemcpyHostToDevice, b->stream));
  RCM_HIP(hipMemcpyAsync(b->d_nnz, nnz, C * sizeof(int64_t), hipMemcpyHostToDevice, b->stream));
  b->nblk_order = build_blk(b, b->order.data(), count, b->blk_order_h);
  RCM_HIP(hipMemcpyAsync(b->blk_order, b->blk_order_h.data(), (C + 1) * sizeof(int32_t), hipMemcpyHostToDevice,
                         b->stream));
  RCM_HIP(hipMemsetAsync(b->scratch, 0, 3 * C * sizeof(uint32_t), b->stream));
  hipLaunchKernelGGL(k_validate_many, dim3((unsigned)b->nblk_order), dim3(256), 0, b->stream, b->jobs, b->d_order,
                     count, b->blk_order, b->d_nnz, b->scratch);
  std::vector<uint32_t> v(2 * C);
  RCM_HIP(hipMemcpyAsync(v.data(), b->scratch, 2 * C * sizeof(uint32_t), hipMemcpyDeviceToHost, b->stream));
  RCM_HIP(hipStreamSynchronize(b->stream));
  int rc = RCM_OK;
  b->maxrow.assign(C, 0);
  for (int32_t k = 0; k < count; ++k) {
    const int r = v[2 * k] ? RCM_EINVAL : RCM_OK;
    if (status_out) status_out[k] = r;
    if (r) rc = r;
    b->maxrow[k] = v[2 * k + 1];
  }
  b->timing = rcm_batch_timing_t{};
  b->timing.count = count;
  b->timing.total_n = tot;
  b->timing.total_nnz = totnnz;
  b->timing.load_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  if (rc) batch_reset(b);
  return rc;
}

namespace {
// One launch over `list` (plain-BFS sweeps when bfs, else CM runs, which record levels when `levels`) from
// cur[], then back on the host: every control block, and for each listed matrix the last w = min(32, n-1)
// queue entries (node, level, row length) at h_tail[i * 96]. Per-matrix status of the listed matrices goes to
// status[]; a device-level failure is returned.
int batch_pass(rcm_batch* b, const std::vector<int32_t>& list, const std::vector<int32_t>& cur, bool bfs,
               bool levels, int32_t* status) {
  const int32_t C = b->count, A = (int32_t)list.size();
  RCM_HIP(hipMemcpyAsync(b->d_starts, cur.data(), (size_t)C * sizeof(int32_t), hipMemcpyHostToDevice, b->stream));
  RCM_HIP(hipMemcpyAsync(b->d_active, list.data(), (size_t)A * sizeof(int32_t), hipMemcpyHostToDevice, b->stream));
  std::vector<int32_t> blk;
  (void)build_blk(b, list.data(), A, blk);
  RCM_HIP(hipMemcpyAsync(b->blk_active, blk.data(), ((size_t)A + 1) * sizeof(int32_t), hipMemcpyHostToDevice,
                         b->stream));
  int64_t rows = 0;
  for (int32_t k : list) rows += b->n[k];
  enqueue_many(b, b->d_active, b->blk_active, blk[A], A, bfs, levels, rows);
  hipLaunchKernelGGL(k_gather_tail, dim3((unsigned)A), dim3(64), 0, b->stream, b->jobs, b->d_active, A, b->tail);
  RCM_HIP(hipMemcpyAsync(b->h_ctl, b->ctl, (size_t)C * sizeof(Control), hipMemcpyDeviceToHost, b->stream));
  RCM_HIP(hipMemcpyAsync(b->h_g, b->g, sizeof(ManyControl), hipMemcpyDeviceToHost, b->stream));
  RCM_HIP(hipMemcpyAsync(b->h_tail, b->tail, (size_t)A * 96 * sizeof(int32_t), hipMemcpyDeviceToHost, b->stream));
  RCM_HIP(hipEventRecord(b->ev[3], b->stream));
  const int wr = batch_wait(b, b->ev[3]);
  if (wr) return wr;
  (void)collect_status(b, list.data(), A, status, nullptr);
  return RCM_OK;
}
}  // namespace

// SimplePeripheral::findPeripheral (src/simple_peripheral.cpp:92-141) for every matrix: each sweep is one
// persistent plain-BFS launch over the matrices still iterating; the rule on the last min(32, n-1) queue
// entries is applied per matrix on the host, as in rcm_ctx_find_start.
int rcm_batch_find_start(rcm_batch* b, int32_t* start_out, int32_t* iterations_out, int32_t* status_out) {
  if (!b || !start_out) return RCM_EINVAL;
  if (b->count == 0) return RCM_ENOTLOADED;
  RCM_HIP(hipSetDevice(b->device));
  const auto t0 = std::chrono::steady_clock::now();
  const int32_t C = b->count;
  std::vector<int32_t> cur(C), iters(C, 0), active, status(C, RCM_OK);
  std::vector<uint32_t> lastdist(C, 0);
  for (int32_t k : b->order) {
    if (b->n[k] == 1) {
      cur[k] = 0;
      iters[k] = 1;
    } else {
      cur[k] = b->n[k] / 2;
      active.push_back(k);
    }
  }
  int sweeps = 0;
  while (!active.empty()) {
    ++sweeps;
    const int r = batch_pass(b, active, cur, true, false, status.data());
    if (r) return r;
    std::vector<int32_t> next;
    for (size_t i = 0; i < active.size(); ++i) {
      const int32_t k = active[i];
      ++iters[k];
      if (status[k] != RCM_OK) continue;
      const int32_t w = std::min<int32_t>(32, b->n[k] - 1);
      const int32_t* nodes = b->h_tail + i * 96;
      if (peripheral_step(nodes, nodes + 32, nodes + 64, w, &lastdist[k], &cur[k])) continue;  // converged
      next.push_back(k);
    }
    active.swap(next);
  }
  int rc = RCM_OK;
  for (int32_t k = 0; k < C; ++k) {
    start_out[k] = cur[k];
    if (iterations_out) iterations_out[k] = iters[k];
    if (status_out) status_out[k] = status[k];
    if (status[k] && rc == RCM_OK) rc = status[k];
  }
  b->timing.start_sweeps = sweeps;
  b->timing.start_search_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  return rc;
}

// Start search and order for every matrix, as rcm_batch_find_start + rcm_batch_run with one engine pass fewer
// per matrix that stops at a speculated sweep (see rcm_ctx_order). Each launch is of one kind: plain sweeps while
// any matrix needs one, otherwise CM runs (speculated sweeps, and final runs of matrices whose start is known).
int rcm_batch_order(rcm_batch* b, const int32_t* start_in, uint32_t flags, int32_t* start_out,
                    int32_t* iterations_out, int32_t* status_out) {
  if (!b) return RCM_EINVAL;
  if (b->count == 0) return RCM_ENOTLOADED;
  const int32_t C = b->count;
  if (start_in)
    for (int32_t k = 0; k < C; ++k)
      if (start_in[k] >= b->n[k]) return RCM_EINVAL;
  RCM_HIP(hipSetDevice(b->device));
  const auto t0 = std::chrono::steady_clock::now();
  b->have_perm = false;
  enum : int8_t { kPlain, kSpec, kFinal, kDone };
  std::vector<int32_t> cur(C), iters(C, 0), status(C, RCM_OK);
  std::vector<uint32_t> lastdist(C, 0);
  std::vector<int8_t> need(C), redo(C, 0);  // redo: the plain sweep repeats a speculated one (counted already)
  for (int32_t k = 0; k < C; ++k) {
    if (start_in && start_in[k] >= 0) {
      cur[k] = start_in[k];
      need[k] = kFinal;
    } else if (b->n[k] == 1) {
      cur[k] = 0;
      iters[k] = 1;
      need[k] = kFinal;
    } else {
      cur[k] = b->n[k] / 2;
      need[k] = kPlain;
    }
  }
  int launches = 0;
  std::vector<int32_t> plain, cm;
  for (;;) {
    plain.clear();
    cm.clear();
    for (int32_t k : b->order) {
      if (need[k] == kPlain) plain.push_back(k);
      else if (need[k] != kDone) cm.push_back(k);
    }
    const bool bfs = !plain.empty();
    const std::vector<int32_t>& list = bfs ? plain : cm;
    if (list.empty()) break;
    ++launches;
    const int r = batch_pass(b, list, cur, bfs, !bfs, status.data());
    if (r) return r;
    for (size_t i = 0; i < list.size(); ++i) {
      const int32_t k = list[i];
      const int32_t w = std::min<int32_t>(32, b->n[k] - 1);
      const int32_t* tail = b->h_tail + i * 96;
      if (bfs) {
        if (!redo[k]) ++iters[k];
        redo[k] = 0;
        if (status[k] != RCM_OK) {
          need[k] = kDone;
        } else if (peripheral_step(tail, tail + 32, tail + 64, w, &lastdist[k], &cur[k])) {
          need[k] = kFinal;
        } else {
          need[k] = (b->speculate > 0 && iters[k] + 1 >= b->speculate) ? kSpec : kPlain;
        }
      } else if (need[k] == kFinal || status[k] != RCM_OK) {
        need[k] = kDone;
      } else {  // speculated sweep: the last entry's level is the eccentricity of cur
        ++iters[k];
        if ((uint32_t)tail[32 + w - 1] == lastdist[k]) {
          need[k] = kDone;
        } else {
          need[k] = kPlain;
          redo[k] = 1;
        }
      }
    }
  }
  const bool rev = (flags & RCM_FLAG_REVERSE) != 0;
  if (rev) {
    hipLaunchKernelGGL(k_reverse_many, dim3((unsigned)b->nblk_order), dim3(256), 0, b->stream, b->jobs, b->d_order,
                       C, b->blk_order);
    RCM_HIP(hipStreamSynchronize(b->stream));
  }
  int rc = RCM_OK;
  b->ok_jobs.clear();
  for (int32_t k : b->order)
    if (status[k] == RCM_OK) b->ok_jobs.push_back(k);
  for (int32_t k = 0; k < C; ++k) {
    if (start_out) start_out[k] = cur[k];
    if (iterations_out) iterations_out[k] = iters[k];
    if (status_out) status_out[k] = status[k];
    if (status[k] && rc == RCM_OK) rc = status[k];
  }
  b->reversed = rev;
  b->have_perm = true;  // per-matrix status tells which orders are complete
  b->timing.order_passes = launches;
  b->timing.order_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  return rc;
}

int rcm_batch_run(rcm_batch* b, const int32_t* start, uint32_t flags, int32_t* status_out) {
  if (!b || !start) return RCM_EINVAL;
  if (b->count == 0) return RCM_ENOTLOADED;
  for (int32_t k = 0; k < b->count; ++k)
    if (start[k] < 0 || start[k] >= b->n[k]) return RCM_EINVAL;
  RCM_HIP(hipSetDevice(b->device));
  b->have_perm = false;
  const int32_t C = b->count;
  const bool rev = (flags & RCM_FLAG_REVERSE) != 0;
  RCM_HIP(hipMemcpyAsync(b->d_starts, start, (size_t)C * sizeof(int32_t), hipMemcpyHostToDevice, b->stream));
  RCM_HIP(hipEventRecord(b->ev[0], b->stream));
  auto enqueue_run = [&]() {
    enqueue_many(b, b->d_order, b->blk_order, b->nblk_order, C, false, false, b->total_n);
    if (rev)
      hipLaunchKernelGGL(k_reverse_many, dim3((unsigned)b->nblk_order), dim3(256), 0, b->stream, b->jobs, b->d_order,
                         C, b->blk_order);
  };
  if (b->use_graph) {
    hipGraphExec_t& ex = b->gexec[rev ? 1 : 0];
    if (!ex) {
      hipGraph_t& gr = b->graph[rev ? 1 : 0];
      RCM_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
      enqueue_run();
      RCM_HIP(hipStreamEndCapture(b->stream, &gr));
      RCM_HIP(hipGraphInstantiate(&ex, gr, nullptr, nullptr, 0));
    } else {
      ++b->timing.graph_replays;
    }
    RCM_HIP(hipGraphLaunch(ex, b->stream));
  } else {
    enqueue_run();
  }
  RCM_HIP(hipEventRecord(b->ev[1], b->stream));
  RCM_HIP(hipMemcpyAsync(b->h_ctl, b->ctl, (size_t)C * sizeof(Control), hipMemcpyDeviceToHost, b->stream));
  RCM_HIP(hipMemcpyAsync(b->h_g, b->g, sizeof(ManyControl), hipMemcpyDeviceToHost, b->stream));
  RCM_HIP(hipEventRecord(b->ev[3], b->stream));
  const int wr = batch_wait(b, b->ev[3]);
  if (wr) return wr;
  float ms = 0.f;
  RCM_HIP(hipEventElapsedTime(&ms, b->ev[0], b->ev[1]));
  b->timing.run_ms = ms;
  b->timing.workers = batch_workers(b);
  int64_t nb = 0;
  std::vector<int32_t> st((size_t)C, RCM_OK);
  const int rc = collect_status(b, b->order.data(), C, st.data(), &nb);
  if (status_out) std::copy(st.begin(), st.end(), status_out);
  b->ok_jobs.clear();  // matrices with a complete order (the bandwidth reads only these)
  for (int32_t k : b->order)
    if (st[k] == RCM_OK) b->ok_jobs.push_back(k);
  b->timing.batches = nb;
  b->reversed = rev;
  b->have_perm = true;  // per-matrix status tells which orders are complete
  return rc;
}

int rcm_batch_bandwidth(rcm_batch* b, int64_t* bw_out) {
  if (!b || !bw_out) return RCM_EINVAL;
  if (!b->have_perm) return RCM_ENOTLOADED;
  RCM_HIP(hipSetDevice(b->device));
  const int32_t C = b->count;
  // only complete orders are permutations of [0, n): the inverse of any other would write out of range
  const int32_t A = (int32_t)b->ok_jobs.size();
  uint32_t* bw = b->scratch + 2 * (size_t)C;
  RCM_HIP(hipMemsetAsync(bw, 0, (size_t)C * sizeof(uint32_t), b->stream));
  RCM_HIP(hipEventRecord(b->ev[0], b->stream));
  if (A > 0) {
    RCM_HIP(hipMemcpyAsync(b->d_active, b->ok_jobs.data(), (size_t)A * sizeof(int32_t), hipMemcpyHostToDevice,
                           b->stream));
    std::vector<int32_t> blk;
    const int32_t nblk = build_blk(b, b->ok_jobs.data(), A, blk);
    RCM_HIP(hipMemcpyAsync(b->blk_active, blk.data(), ((size_t)A + 1) * sizeof(int32_t), hipMemcpyHostToDevice,
                           b->stream));
    hipLaunchKernelGGL(k_inverse_many, dim3((unsigned)nblk), dim3(256), 0, b->stream, b->jobs, b->d_active, A,
                       b->blk_active);
    hipLaunchKernelGGL(k_bandwidth_many, dim3((unsigned)nblk), dim3(256), 0, b->stream, b->jobs, b->d_active, A,
                       b->blk_active, bw);
    if (b->reversed)  // aux held the reversed order: restore it (the bandwidth is the same either way)
      hipLaunchKernelGGL(k_reverse_many, dim3((unsigned)nblk), dim3(256), 0, b->stream, b->jobs, b->d_active, A,
                         b->blk_active);
  }
  RCM_HIP(hipEventRecord(b->ev[1], b->stream));
  std::vector<uint32_t> h((size_t)C);
  RCM_HIP(hipMemcpyAsync(h.data(), bw, (size_t)C * sizeof(uint32_t), hipMemcpyDeviceToHost, b->stream));
  RCM_HIP(hipStreamSynchronize(b->stream));
  float ms = 0.f;
  RCM_HIP(hipEventElapsedTime(&ms, b->ev[0], b->ev[1]));
  b->timing.bandwidth_ms = ms;
  std::vector<uint8_t> ok((size_t)C, 0);
  for (int32_t k : b->ok_jobs) ok[k] = 1;
  for (int32_t k = 0; k < C; ++k) bw_out[k] = ok[k] ? (int64_t)h[k] : -1;
  return RCM_OK;
}

int rcm_batch_read_perm(rcm_batch* b, const int64_t* po_off, int32_t* perm_cat) {
  if (!b || !perm_cat) return RCM_EINVAL;
  if (!b->have_perm) return RCM_ENOTLOADED;
  RCM_HIP(hipSetDevice(b->device));
  const int32_t* src = b->reversed ? b->aux : b->perm;
  bool packed = true;
  if (po_off)
    for (int32_t k = 0; k < b->count && packed; ++k) packed = po_off[k] == b->pn[k];
  if (packed) {
    RCM_HIP(hipMemcpyAsync(perm_cat, src, (size_t)b->total_n * sizeof(int32_t), hipMemcpyDeviceToHost, b->stream));
  } else {
    for (int32_t k = 0; k < b->count; ++k)
      RCM_HIP(hipMemcpyAsync(perm_cat + po_off[k], src + b->pn[k], (size_t)b->n[k] * sizeof(int32_t),
                             hipMemcpyDeviceToHost, b->stream));
  }
  RCM_HIP(hipStreamSynchronize(b->stream));
  return RCM_OK;
}

int rcm_batch_perm_device(rcm_batch* b, int32_t k, const int32_t** d_perm_out) {
  if (!b || !d_perm_out || k < 0 || k >= b->count) return RCM_EINVAL;
  if (!b->have_perm) return RCM_ENOTLOADED;
  *d_perm_out = (b->reversed ? b->aux : b->perm) + b->pn[k];
  return RCM_OK;
}

int rcm_batch_timing(rcm_batch* b, rcm_batch_timing_t* out) {
  if (!b || !out) return RCM_EINVAL;
  *out = b->timing;
  return RCM_OK;
}

int rcm_gpu_batch_many(int32_t count, const int32_t* n, const int64_t* nnz, const int64_t* rp_off,
                       const int64_t* ci_off, const int32_t* row_ptr_cat, const int32_t* col_idx_cat,
                       const int32_t* start, uint32_t flags, const int64_t* po_off, int32_t* perm_cat,
                       int64_t* bandwidth_out, int32_t* status_out) {
  if (count < 0 || (count > 0 && (!n || !nnz || !rp_off || !ci_off || !row_ptr_cat || !col_idx_cat || !po_off ||
                                  !perm_cat)))
    return RCM_EINVAL;
  if (count == 0) return RCM_OK;
  int dev = 0;
  RCM_HIP(hipGetDevice(&dev));
  rcm_batch* b = nullptr;
  int rc = rcm_batch_create(dev, &b);
  if (rc) return rc;
  rcm_batch_set_param(b, RCM_PARAM_GRAPH, 0);  // one run: nothing to replay
  std::vector<int32_t> st((size_t)count, 0), status((size_t)count, RCM_OK);
  rc = rcm_batch_load(b, count, n, nnz, rp_off, ci_off, row_ptr_cat, col_idx_cat, status.data());
  for (int32_t k = 0; k < count; ++k)
    if (start && start[k] >= n[k]) {
      status[k] = RCM_EINVAL;
      if (!rc) rc = RCM_EINVAL;
    }
  // a matrix that fails (e.g. disconnected) keeps its status; the others are still reordered
  int hard = rc;  // launch-level failure (malformed input, device error, time-out)
  int worst = hard;
  if (!hard) {  // start search where no start is given, and order, in one call
    const int r = rcm_batch_order(b, start, flags, st.data(), nullptr, status.data());
    if (r == RCM_EDEVICE || r == RCM_ETIMEOUT) hard = r;
    for (int32_t k = 0; k < count; ++k)
      if (status[k] && worst == RCM_OK) worst = status[k];
  }
  if (!hard && bandwidth_out && (flags & RCM_FLAG_BANDWIDTH)) hard = rcm_batch_bandwidth(b, bandwidth_out);
  if (!hard) hard = rcm_batch_read_perm(b, po_off, perm_cat);
  if (hard && worst == RCM_OK) worst = hard;
  if (status_out)
    for (int32_t k = 0; k < count; ++k) status_out[k] = hard && status[k] == RCM_OK ? hard : status[k];
  rcm_batch_destroy(b);
  return worst;
}

int rcm_server_create(int device, const rcm_server_config* cfg, rcm_server** out) {
  if (!out) return RCM_EINVAL;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev) return RCM_EDEVICE;
  RCM_HIP(hipSetDevice(device));
  rcm_server_config c{};
  if (cfg) c = *cfg;
  if (c.slots < 0 || c.slots > 4096 || c.lanes < 0 || c.lanes > 16 || c.arena_bytes < 0 || c.workers < 0 ||
      c.job_timeout_ms < 0)
    return RCM_EINVAL;
  if (c.slots == 0) c.slots = 256;
  if (c.lanes == 0) c.lanes = 4;
  if (c.job_timeout_ms == 0) c.job_timeout_ms = 60000;
  if (c.arena_bytes == 0) {
    size_t fr = 0, tot = 0;
    RCM_HIP(hipMemGetInfo(&fr, &tot));
    c.arena_bytes = (int64_t)std::min<size_t>(fr / 2, (size_t)64 << 30);
  }
  rcm_server* s = new (std::nothrow) rcm_server();
  if (!s) return RCM_ENOMEM;
  s->device = device;
  s->cfg = c;
  s->lanes = c.lanes;
  s->lane_bytes = (size_t)c.arena_bytes / (size_t)c.lanes;
  int rc = RCM_OK;
  int wtotal = 0;
  for (int l = 0; l < s->lanes && !rc; ++l) {
    rcm_batch* b = nullptr;
    rc = rcm_batch_create(device, &b);
    if (rc) break;
    s->batch.push_back(b);
    rcm_batch_set_param(b, RCM_PARAM_VARIANT, 0);  // lanes were measured with variant 0 (256 threads)
    // twice the lane's share by default: idle workgroups leave a launch (k_reorder_many), so a lane whose group
    // needs fewer gives its slots back; measured on the 400-request stream, 4 lanes: capacity 4.0x one-at-a-time
    // with 128 workgroups per lane, 4.4x with 256 or 384, 3.9x with 512
    const int w = c.workers > 0 ? c.workers : std::min(b->max_resident, std::max(1, 2 * b->max_resident / s->lanes));
    rcm_batch_set_param(b, RCM_PARAM_WORKERS, w);
    rcm_batch_set_param(b, RCM_PARAM_TIMEOUT_MS, c.job_timeout_ms);
    rcm_batch_set_param(b, RCM_PARAM_GRAPH, 0);  // every group is a different set
    rcm_batch_set_param(b, RCM_PARAM_HELPER_NODES, kServerHelperNodes);
    wtotal += std::min(w, b->max_resident);
    rc = batch_reserve(b, s->lane_bytes, c.slots, true);  // the lane's arena, allocated once, exactly
    b->arena_fixed = true;
  }
  if (rc) {
    rcm_server_destroy(s);
    return rc;
  }
  s->st.slots = c.slots;
  s->st.workers = wtotal;
  s->st.arena_bytes = (int64_t)s->lane_bytes * s->lanes;
  s->st.kernel_running = 0;
  s->stage_rp.resize((size_t)s->lanes);
  s->stage_ci.resize((size_t)s->lanes);
  s->stage_perm.resize((size_t)s->lanes);
  s->pinned.assign((size_t)s->lanes, nullptr);
  for (int l = 0; l < s->lanes; ++l)  // optional: a lane without it stages in pageable memory
    if (hipHostMalloc(reinterpret_cast<void**>(&s->pinned[(size_t)l]), kServePinnedWords * sizeof(int32_t),
                      hipHostMallocDefault) != hipSuccess) {
      s->pinned[(size_t)l] = nullptr;
      (void)hipGetLastError();  // clear the failed call's error
    } else {
      ++s->st.pinned_lanes;
    }
  s->solo.assign((size_t)s->lanes, nullptr);
  for (int l = 0; l < s->lanes; ++l) {  // optional: without it a lane orders large requests as one-member groups
    rcm_ctx* sc = nullptr;
    if (rcm_ctx_create(device, &sc) != RCM_OK) continue;
    rcm_ctx_set_param(sc, RCM_PARAM_TIMEOUT_MS, s->cfg.job_timeout_ms);
    if (dev_alloc(&sc->rp, (size_t)kServeSoloN + 1) == hipSuccess) sc->cap_rp = (int64_t)kServeSoloN + 1;
    if (dev_alloc(&sc->ci, (size_t)kServeSoloNnz) == hipSuccess) sc->cap_ci = kServeSoloNnz;
    sc->own_csr = true;
    if (sc->cap_rp < (int64_t)kServeSoloN + 1 || sc->cap_ci < kServeSoloNnz ||
        ensure_state(sc, kServeSoloN) != RCM_OK) {
      (void)hipGetLastError();
      rcm_ctx_destroy(sc);
      continue;
    }
    s->solo[(size_t)l] = sc;
    ++s->st.solo_lanes;
  }
  for (int l = 0; l < s->lanes; ++l) s->threads.emplace_back(lane_loop, s, l);
  *out = s;
  return RCM_OK;
}

int rcm_server_submit(rcm_server* s, int32_t n, int64_t nnz, const int32_t* row_ptr, const int32_t* col_idx,
                      int32_t start_node, uint32_t flags, int64_t* id_out) {
  if (!s || !id_out || n < 1 || n >= (1 << 30) || nnz < 0 || nnz > INT32_MAX || !row_ptr || (nnz > 0 && !col_idx) ||
      start_node >= n)
    return RCM_EINVAL;
  const size_t bytes = GroupSize{}.plus(n, nnz).bytes();
  if (bytes > s->lane_bytes) return RCM_ENOMEM;  // can never fit a lane
  auto jp = std::make_shared<ServerJob>();
  jp->n = n;
  jp->nnz = nnz;
  if (flags & RCM_FLAG_BORROW) {  // read in place when the group is formed
    jp->src_rp = row_ptr;
    jp->src_ci = col_idx;
  } else {
    // fresh pages are faulted in by the copy itself (no zero fill first), and a large request's copy is split
    // over kSubmitThreads threads so that its page faults are taken in parallel
    jp->rp.reset(new (std::nothrow) int32_t[(size_t)n + 1]);
    if (nnz > 0) jp->ci.reset(new (std::nothrow) int32_t[(size_t)nnz]);
    if (!jp->rp || (nnz > 0 && !jp->ci)) return RCM_ENOMEM;
    constexpr int kSubmitThreads = 4;
    const int64_t words = (int64_t)n + 1 + nnz;
    auto copy_words = [&](int64_t w0, int64_t w1) {  // words [w0, w1) of row_ptr followed by col_idx
      const int64_t nr = (int64_t)n + 1;
      if (w0 < nr) std::copy(row_ptr + w0, row_ptr + std::min(w1, nr), jp->rp.get() + w0);
      const int64_t c0 = std::max(w0, nr) - nr;
      if (w1 > nr) std::copy(col_idx + c0, col_idx + (w1 - nr), jp->ci.get() + c0);
    };
    if (words >= ((int64_t)1 << 22)) {
      std::vector<std::thread> th;
      for (int p = 1; p < kSubmitThreads; ++p)
        th.emplace_back(copy_words, words * p / kSubmitThreads, words * (p + 1) / kSubmitThreads);
      copy_words(0, words / kSubmitThreads);
      for (auto& x : th) x.join();
    } else {
      copy_words(0, words);
    }
    jp->src_rp = jp->rp.get();
    jp->src_ci = jp->ci.get();
  }
  jp->start_req = start_node;
  jp->flags = flags;
  jp->bytes = bytes;
  jp->t_submit = std::chrono::steady_clock::now();
  {
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->stopping) return RCM_ECANCELED;
    jp->id = s->next_id++;
    s->jobs[jp->id] = jp;
    s->pending.push_back(jp);
    ++s->st.submitted;
    s->st.pending = (int32_t)s->pending.size();
    *id_out = jp->id;
  }
  s->cv_work.notify_one();
  return RCM_OK;
}

int rcm_server_wait(rcm_server* s, int64_t id, int64_t timeout_ms, int32_t* perm_out, rcm_job_stats* stats) {
  if (!s) return RCM_EINVAL;
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->jobs.find(id);
  if (it == s->jobs.end()) return RCM_ENOTFOUND;
  std::shared_ptr<ServerJob> jp = it->second;
  auto is_done = [&] { return jp->done; };
  if (timeout_ms < 0)
    s->cv_done.wait(lk, is_done);
  else if (!s->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), is_done))
    return RCM_EAGAIN;
  s->jobs.erase(id);
  lk.unlock();
  if (perm_out && jp->status == RCM_OK) std::memcpy(perm_out, jp->perm.data(), (size_t)jp->n * sizeof(int32_t));
  if (stats) {
    auto ms = [](std::chrono::steady_clock::duration d) { return std::chrono::duration<double, std::milli>(d).count(); };
    rcm_job_stats js{};
    js.status = jp->status;
    js.start_node = jp->start_used;
    js.sweeps = jp->sweeps;
    js.lane = jp->lane;
    const bool admitted = jp->lane >= 0;
    js.queue_ms = admitted ? ms(jp->t_admit - jp->t_submit) : ms(jp->t_done - jp->t_submit);
    js.device_ms = admitted ? ms(jp->t_done - jp->t_admit) : 0.0;
    js.latency_ms = ms(jp->t_done - jp->t_submit);
    js.bandwidth = jp->bw;
    *stats = js;
  }
  return jp->status;
}

int rcm_server_stats_get(rcm_server* s, rcm_server_stats* out) {
  if (!s || !out) return RCM_EINVAL;
  std::lock_guard<std::mutex> lk(s->mu);
  *out = s->st;
  return RCM_OK;
}

int rcm_server_destroy(rcm_server* s) {
  if (!s) return RCM_EINVAL;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stopping = true;
    for (auto& jp : s->pending) {  // not yet admitted: canceled
      jp->status = RCM_ECANCELED;
      jp->t_done = std::chrono::steady_clock::now();
      jp->done = true;
      ++s->st.failed;
    }
    s->pending.clear();
  }
  s->cv_work.notify_all();
  for (auto& t : s->threads)
    if (t.joinable()) t.join();  // admitted groups finish first
  s->cv_done.notify_all();
  for (rcm_batch* b : s->batch) rcm_batch_destroy(b);
  for (rcm_ctx* sc : s->solo)
    if (sc) rcm_ctx_destroy(sc);
  for (int32_t* p : s->pinned)
    if (p) (void)hipHostFree(p);
  delete s;
  return RCM_OK;
}

// Candidate workgroup counts of a variant: fractions of its co-resident capacity (and the automatic choice).
static std::vector<int> tune_workers(int max_res) {
  std::vector<int> w;
  // 3/16 and 3/8 as well: variant 1 runs the 128^3 stencil in 3.87-3.88 ms with 192-224 of its 512 workgroups,
  // against 3.93 ms with 256 and 4.02 ms with 160
  for (int x : {max_res / 8, max_res * 3 / 16, max_res / 4, max_res * 3 / 8, max_res / 2, max_res})
    if (x >= 16 && std::find(w.begin(), w.end(), x) == w.end()) w.push_back(x);
  if (w.empty()) w.push_back(std::max(1, max_res));
  return w;
}

// Median of the timings; +inf for none, so a candidate without timings is never chosen.
static double median_of(std::vector<double> v) {
  if (v.empty()) return std::numeric_limits<double>::infinity();
  std::sort(v.begin(), v.end());
  return v[v.size() / 2];
}

int rcm_ctx_autotune(rcm_ctx* c, int32_t start, int32_t repeats, rcm_tune_result* out) {
  if (!c || !out || repeats < 1) return RCM_EINVAL;
  if (!c->rp || c->n == 0) return RCM_ENOTLOADED;
  if (start < 0 || start >= c->n) return RCM_EINVAL;
  RCM_HIP(hipSetDevice(c->device));
  const int64_t keep_v = c->variant, keep_w = c->workers;
  rcm_tune_result best{};
  best.best_ms = 1e300;
  int count = 0;
  for (int v = 0; v < kVariants; ++v) {
    std::vector<int> ws = tune_workers(c->max_resident[v]);
    if (v == 0) ws.insert(ws.begin(), 0);  // 0 = automatic: the default configuration
    for (int w : ws) {
      c->variant = v;
      c->workers = w;
      std::vector<double> t;
      for (int r = 0; r < repeats; ++r) {
        const int rc = run_engine(c, start, false);
        if (rc) {
          c->variant = keep_v;
          c->workers = keep_w;
          return rc;
        }
        t.push_back(c->timing.run_kernel_ms);
      }
      const double ms = median_of(t);
      ++count;
      if (v == 0 && w == 0) best.default_ms = ms;
      if (ms < best.best_ms) {
        best.best_ms = ms;
        best.variant = v;
        best.workers = workers_for(c);
      }
    }
  }
  {  // the level-synchronous engine: one configuration
    c->variant = kLevelVariant;
    std::vector<double> t;
    int rc = RCM_OK;
    for (int r = 0; r < repeats && rc == RCM_OK; ++r) {
      rc = run_engine(c, start, false);
      t.push_back(c->timing.run_kernel_ms);
    }
    if (rc == RCM_OK) {  // a level too wide for 64-bit keys leaves it out
      ++count;
      const double ms = median_of(t);
      if (ms < best.best_ms) {
        best.best_ms = ms;
        best.variant = kLevelVariant;
        best.workers = 0;
      }
    } else if (rc != RCM_ECAPACITY) {
      c->variant = keep_v;
      c->workers = keep_w;
      return rc;
    }
  }
  c->variant = best.variant;
  c->workers = best.workers;
  best.candidates = count;
  best.threads = variant_threads(best.variant);
  c->have_perm = false;  // the tuning runs overwrote the last order
  *out = best;
  return RCM_OK;
}

int rcm_batch_autotune(rcm_batch* b, const int32_t* start, int32_t repeats, rcm_tune_result* out) {
  if (!b || !start || !out || repeats < 1) return RCM_EINVAL;
  if (b->count == 0) return RCM_ENOTLOADED;
  const int64_t keep_v = b->variant, keep_w = b->workers;
  rcm_tune_result best{};
  best.best_ms = 1e300;
  int count = 0;
  std::vector<int32_t> st((size_t)b->count);
  for (int v = 0; v < kVariants; ++v) {
    std::vector<int> ws = tune_workers(b->max_res[v]);
    if (v == 0) ws.insert(ws.begin(), 0);
    for (int w : ws) {
      rcm_batch_set_param(b, RCM_PARAM_VARIANT, v);
      rcm_batch_set_param(b, RCM_PARAM_WORKERS, w);
      std::vector<double> t;
      for (int r = 0; r < repeats; ++r) {
        const int rc = rcm_batch_run(b, start, RCM_FLAG_STABLE, st.data());
        if (rc == RCM_EDEVICE || rc == RCM_ETIMEOUT || rc == RCM_EINVAL || rc == RCM_ENOTLOADED) {
          rcm_batch_set_param(b, RCM_PARAM_VARIANT, keep_v);
          rcm_batch_set_param(b, RCM_PARAM_WORKERS, keep_w);
          return rc;
        }
        t.push_back(b->timing.run_ms);
      }
      const double ms = median_of(t);
      ++count;
      if (v == 0 && w == 0) best.default_ms = ms;
      if (ms < best.best_ms) {
        best.best_ms = ms;
        best.variant = v;
        best.workers = batch_workers(b);
      }
    }
  }
  rcm_batch_set_param(b, RCM_PARAM_VARIANT, best.variant);
  rcm_batch_set_param(b, RCM_PARAM_WORKERS, best.workers);
  // then the helper cap, with the chosen engine
  const int64_t keep_h = b->helper_nodes;
  int64_t best_h = keep_h;
  for (int64_t h : {kDefaultHelperNodes, 4 * kDefaultHelperNodes, 16 * kDefaultHelperNodes}) {
    if (h == keep_h) continue;  // timed above
    int rc = rcm_batch_set_param(b, RCM_PARAM_HELPER_NODES, h);
    std::vector<double> t;
    for (int r = 0; rc == RCM_OK && r < repeats; ++r) {
      // per-matrix failures (a disconnected member) do not void the timing: every matrix was run, as in the
      // variant loop above; only launch-level failures end the tuning
      const int run = rcm_batch_run(b, start, RCM_FLAG_STABLE, st.data());
      if (run == RCM_EDEVICE || run == RCM_ETIMEOUT || run == RCM_EINVAL || run == RCM_ENOTLOADED) rc = run;
      else t.push_back(b->timing.run_ms);
    }
    if (rc != RCM_OK) {
      rcm_batch_set_param(b, RCM_PARAM_HELPER_NODES, keep_h);
      return rc;
    }
    if (t.empty()) continue;  // nothing timed: not a candidate
    ++count;
    const double ms = median_of(t);
    if (ms < best.best_ms) {
      best.best_ms = ms;
      best_h = h;
    }
  }
  rcm_batch_set_param(b, RCM_PARAM_HELPER_NODES, best_h);
  best.helper_nodes = (int32_t)best_h;
  best.candidates = count;
  best.threads = kVarThreads[best.variant];
  b->have_perm = false;
  *out = best;
  return RCM_OK;
}

// rcm_gpu_batch on a device-resident CSR with no allocation: every engine buffer is in `scratch` (see above).
int rcm_gpu_batch_noalloc(int32_t n, int64_t nnz, const int32_t* d_row_ptr, const int32_t* d_col_idx,
                          int32_t start_node, uint32_t flags, int32_t* d_perm_out, int64_t* bandwidth_out,
                          int32_t* start_node_out) {
  if (n < 1 || nnz < 0 || nnz > INT32_MAX || !d_row_ptr || (nnz > 0 && !d_col_idx) || !d_perm_out || start_node >= n)
    return RCM_EINVAL;
  if (n > kNoAllocMaxN) return RCM_ECAPACITY;  // beyond the fixed array: use rcm_gpu_batch or a context
  int dev = 0;
  RCM_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kNoAllocDevices) return RCM_EDEVICE;
  NoAllocDevice& D = g_noalloc[dev];
  std::lock_guard<std::mutex> lk(D.mu);
  if (!D.base) {  // set-up, once per device, with that device current
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return RCM_EDEVICE;
    int per_cu = 0, cus = 0;
    RCM_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reorder_kernel(0, false, false, false),
                                                         kVarThreads[0], 0));
    RCM_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    D.max_resident = std::max(1, per_cu) * std::max(1, cus);
    D.base = static_cast<unsigned char*>(p);
  }
  const NoAllocLayout L = noalloc_layout((size_t)n);
  unsigned char* base = D.base;
  uint32_t* words = reinterpret_cast<uint32_t*>(base + L.words);
  // validation: the flags and longest row start at zero (the array keeps the previous call's contents)
  hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, nullptr, words, 16, 0u);
  const int vblocks = (int)std::min<int64_t>(4096, (std::max<int64_t>(nnz, n) + 255) / 256);
  hipLaunchKernelGGL(k_validate, dim3(std::max(vblocks, 1)), dim3(256), 0, nullptr, d_row_ptr, d_col_idx, n, nnz,
                     words);
  RCM_HIP(hipGetLastError());
  uint32_t vflags[2] = {0, 0};
  RCM_HIP(hipMemcpy(vflags, words, sizeof(vflags), hipMemcpyDeviceToHost));
  if (vflags[0] != 0) return RCM_EINVAL;

  RunArgs a{};
  a.row_ptr = d_row_ptr;
  a.col_idx = d_col_idx;
  a.perm = d_perm_out;  // the queue: the start search's sweeps and then the order itself
  a.qinfo = reinterpret_cast<uint64_t*>(base + L.qinfo);
  a.claim = reinterpret_cast<uint64_t*>(base + L.claim);
  a.bstat = reinterpret_cast<uint64_t*>(base + L.bstat);
  a.brange = reinterpret_cast<uint64_t*>(base + L.brange);
  a.ctl = reinterpret_cast<Control*>(base + L.ctl);
  a.n = n;
  a.bcap = n + 1;
  a.budget_ticks = 20000ull * 100000ull;  // 20 s watchdog
  a.sort_mode = 0;
  a.seg_threshold = kDefaultSegThreshold;
  a.claim_filter = kDefaultClaimFilter;

  int32_t start = start_node;
  if (start < 0) {  // SimplePeripheral::findPeripheral (src/simple_peripheral.cpp:92-141)
    int32_t* level = reinterpret_cast<int32_t*>(base + L.level);
    start = n / 2;
    if (n == 1) {
      start = 0;
    } else {
      RunArgs ab = a;
      ab.level = level;
      uint32_t lastdist = 0;
      const int32_t w = std::min<int32_t>(32, n - 1);
      int32_t nodes[32], dists[32], degs[32];
      uint64_t rows[32];
      for (;;) {
        const int rc = noalloc_run(D, ab, start, true);
        if (rc) return rc;
        RCM_HIP(hipMemcpy(nodes, d_perm_out + (n - w), (size_t)w * sizeof(int32_t), hipMemcpyDeviceToHost));
        RCM_HIP(hipMemcpy(dists, level + (n - w), (size_t)w * sizeof(int32_t), hipMemcpyDeviceToHost));
        RCM_HIP(hipMemcpy(rows, ab.qinfo + (n - w), (size_t)w * sizeof(uint64_t), hipMemcpyDeviceToHost));
        for (int32_t i = 0; i < w; ++i) degs[i] = (int32_t)((uint32_t)(rows[i] >> 32) - (uint32_t)rows[i]);
        if (peripheral_step(nodes, dists, degs, w, &lastdist, &start)) break;
      }
    }
  }
  int rc = noalloc_run(D, a, start, false);
  if (rc) return rc;
  if (bandwidth_out && (flags & RCM_FLAG_BANDWIDTH)) {  // computeBandwidth (src/main.cpp:143-184)
    int32_t* inv = reinterpret_cast<int32_t*>(a.claim);  // the claims are no longer needed
    const int blocks = (n + 255) / 256;
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, nullptr, words + 8, 1, 0u);
    hipLaunchKernelGGL(k_inverse, dim3(blocks), dim3(256), 0, nullptr, d_perm_out, n, inv);
    hipLaunchKernelGGL(k_bandwidth, dim3(blocks), dim3(256), 0, nullptr, d_row_ptr, d_col_idx, inv, n, words + 8);
    RCM_HIP(hipGetLastError());
    uint32_t bw = 0;
    RCM_HIP(hipMemcpy(&bw, words + 8, sizeof(uint32_t), hipMemcpyDeviceToHost));
    *bandwidth_out = (int64_t)bw;
  }
  if (flags & RCM_FLAG_REVERSE) {
    hipLaunchKernelGGL(k_reverse_inplace, dim3((n / 2 + 255) / 256 + 1), dim3(256), 0, nullptr, d_perm_out, n);
    RCM_HIP(hipGetLastError());
    RCM_HIP(hipStreamSynchronize(nullptr));
  }
  if (start_node_out) *start_node_out = start;
  return RCM_OK;
}

}  // extern "C"

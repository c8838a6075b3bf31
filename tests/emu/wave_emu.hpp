// wave_emu.hpp -- CPU emulation of the gfx950 wave64 execution model, for running the engine's device code
// (parallelbatchrcm_amd/csrc/engine.hpp and tests/emu/engine_wave.hpp) on the host in tests. TEST INFRASTRUCTURE ONLY.
//
// Every lane is a host thread. Cross-lane operations (ballot, shuffles, readlane, DPP moves) are collectives:
// each lane deposits its value, the 64 lanes of the wave meet at a barrier, each reads what it needs, and they
// meet again; wave-level LDS fences are barriers too. A collective reached by only part of a wave (divergent
// control flow) never completes; the
// barrier detects that (time limit) and aborts with a message, which is the point: the device code must only
// use cross-lane operations in wave-uniform control flow.
#pragma once
#define RCM_WAVE_EMU 1

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <thread>
#include <vector>

#define __device__
#define __global__
#define __host__
#define __forceinline__ inline
#define __launch_bounds__(x)

struct emu_dim3 { unsigned x = 1, y = 1, z = 1; };

namespace emu {

struct Wave {
  std::atomic<int> arrived{0};
  std::atomic<int> phase{0};
  uint64_t slot[64];
  double wait_limit_s = 30.0;
  void barrier() {
    const int ph = phase.load(std::memory_order_acquire);
    if (arrived.fetch_add(1, std::memory_order_acq_rel) == 63) {
      arrived.store(0, std::memory_order_relaxed);
      phase.store(ph + 1, std::memory_order_release);
      return;
    }
    const auto t0 = std::chrono::steady_clock::now();
    unsigned spins = 0;
    while (phase.load(std::memory_order_acquire) == ph) {
      if (++spins > 64) std::this_thread::yield();
      if ((spins & 1023) == 0 &&
          std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count() > wait_limit_s) {
        std::fprintf(stderr, "wave_emu: a cross-lane operation was not reached by all 64 lanes\n");
        std::abort();
      }
    }
  }
};

struct LaneCtx;
struct Block {
  std::vector<unsigned char> lds;
  std::vector<LaneCtx*> lanes;  // diagnostics: every thread's context
  int nthreads = 64;
  std::atomic<int> arrived{0};
  std::atomic<int> phase{0};
  double wait_limit_s = 30.0;
  void dump();
  // s_barrier of the workgroup: every thread (all waves) meets here
  void barrier() {
    const int ph = phase.load(std::memory_order_acquire);
    if (arrived.fetch_add(1, std::memory_order_acq_rel) == nthreads - 1) {
      arrived.store(0, std::memory_order_relaxed);
      phase.store(ph + 1, std::memory_order_release);
      return;
    }
    const auto t0 = std::chrono::steady_clock::now();
    unsigned spins = 0;
    while (phase.load(std::memory_order_acquire) == ph) {
      if (++spins > 64) std::this_thread::yield();
      if ((spins & 1023) == 0 &&
          std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count() > wait_limit_s) {
        std::fprintf(stderr, "wave_emu: a workgroup barrier was not reached by every thread\n");
        dump();
        std::abort();
      }
    }
  }
};

struct LaneCtx {
  emu_dim3 tid, bid, bdim, gdim;
  Block* block = nullptr;
  Wave* wave = nullptr;
  int lane = 0;
  // diagnostics: barriers passed and where the thread is (0 running, 1 workgroup barrier, 2 wave op, 3 done)
  std::atomic<long> nblock{0}, nwave{0};
  std::atomic<int> where{0};
};
inline thread_local LaneCtx ctx;

inline void Block::dump() {
  static std::atomic<int> once{0};
  if (once.exchange(1)) return;
  for (size_t t = 0; t < lanes.size(); ++t) {
    LaneCtx* c = lanes[t];
    if (!c) continue;
    std::fprintf(stderr, "  thread %zu: where %d, workgroup barriers %ld, wave ops %ld\n", t, c->where.load(),
                 c->nblock.load(), c->nwave.load());
  }
}

template <class T>
T& block_shared() {
  // exact-size LDS (as on the device), so out-of-range indices are caught by host sanitizers
  if (ctx.block->lds.size() != sizeof(T)) {
    static std::atomic<int> guard{0};
    (void)guard;
  }
  return *reinterpret_cast<T*>(ctx.block->lds.data());
}

// deposit v, return the slot values of all lanes via f
template <class F>
inline void exchange(uint64_t v, F&& f) {
  ctx.where = 2;
  ctx.nwave++;
  Wave& w = *ctx.wave;
  w.slot[ctx.lane] = v;
  w.barrier();
  f(w.slot);
  w.barrier();
  ctx.where = 0;
}

// LDS bytes per emulated workgroup; the driver sets it to the kernel's shared struct size.
inline size_t emu_lds_bytes = 256 * 1024;

// Runs kernel(args...) over grid x block threads. Waves of a workgroup run concurrently with each other
// only when `concurrent` is set (persistent kernels); otherwise workgroups run one after another.
template <class K, class... A>
void launch(unsigned grid, unsigned block, bool concurrent, K kernel, A... args) {
  const unsigned nwaves = (block + 63) / 64;
  auto run_block_range = [&](unsigned b0, unsigned b1) {
    std::vector<Block> blocks(b1 - b0);
    std::vector<Wave> waves((size_t)(b1 - b0) * nwaves);
    for (auto& bk : blocks) {
      bk.lds.assign(emu_lds_bytes, 0xA5);  // device LDS is not zeroed either
      bk.nthreads = (int)block;
      bk.lanes.assign(block, nullptr);
    }
    std::vector<std::thread> th;
    for (unsigned b = b0; b < b1; ++b)
      for (unsigned t = 0; t < block; ++t)
        th.emplace_back([&, b, t]() {
          ctx.tid.x = t;
          ctx.bid.x = b;
          ctx.bdim.x = block;
          ctx.gdim.x = grid;
          ctx.block = &blocks[b - b0];
          ctx.wave = &waves[(size_t)(b - b0) * nwaves + t / 64];
          ctx.lane = (int)(t % 64);
          blocks[b - b0].lanes[t] = &ctx;
          kernel(args...);
          ctx.where = 3;
        });
    for (auto& x : th) x.join();
  };
  if (concurrent)
    run_block_range(0, grid);
  else
    for (unsigned b = 0; b < grid; ++b) run_block_range(b, b + 1);
}

}  // namespace emu

#define threadIdx (emu::ctx.tid)
#define blockIdx (emu::ctx.bid)
#define blockDim (emu::ctx.bdim)
#define gridDim (emu::ctx.gdim)
#define RCM_LDS(T, name) T& name = emu::block_shared<T>()

inline int __lane_id() { return emu::ctx.lane; }
// workgroup barrier with a workgroup-scope fence, as __syncthreads() on the device
inline void __syncthreads() {
  std::atomic_thread_fence(std::memory_order_seq_cst);
  emu::ctx.where = 1;
  emu::ctx.nblock++;
  emu::ctx.block->barrier();
  emu::ctx.where = 0;
}

inline unsigned long long __ballot(int pred) {
  unsigned long long m = 0;
  emu::exchange(pred ? 1u : 0u, [&](const uint64_t* s) {
    for (int l = 0; l < 64; ++l) m |= (s[l] ? 1ull : 0ull) << l;
  });
  return m;
}
inline int __shfl_up(int v, unsigned d, int width = 64) {
  int r = v;
  emu::exchange((uint32_t)v, [&](const uint64_t* s) {
    const int l = emu::ctx.lane;
    r = (l % width) >= (int)d ? (int)(uint32_t)s[l - (int)d] : v;
  });
  return r;
}
inline int __shfl(int v, int src, int width = 64) {
  int r = v;
  emu::exchange((uint32_t)v, [&](const uint64_t* s) { r = (int)(uint32_t)s[src % 64]; });
  (void)width;
  return r;
}
inline int __shfl_xor(int v, int m, int width = 64) {
  int r = v;
  emu::exchange((uint32_t)v, [&](const uint64_t* s) { r = (int)(uint32_t)s[(emu::ctx.lane ^ m) % 64]; });
  (void)width;
  return r;
}
inline int __builtin_amdgcn_readfirstlane(int v) {
  int r = 0;
  emu::exchange((uint32_t)v, [&](const uint64_t* s) { r = (int)(uint32_t)s[0]; });
  return r;
}
inline int __builtin_amdgcn_readlane(int v, int lane) {
  int r = 0;
  emu::exchange((uint32_t)v, [&](const uint64_t* s) { r = (int)(uint32_t)s[lane]; });
  return r;
}
// DPP: row_shr:n (0x111..0x11f), row_bcast:15 (0x142), row_bcast:31 (0x143); lanes outside row_mask or
// without a valid source keep `old` (bound_ctrl off).
inline int __builtin_amdgcn_update_dpp(int old, int src, int ctrl, int row_mask, int bank_mask, bool bound) {
  (void)bank_mask;
  (void)bound;
  int r = old;
  emu::exchange((uint32_t)src, [&](const uint64_t* s) {
    const int l = emu::ctx.lane, row = l / 16, inrow = l % 16;
    if (!((row_mask >> row) & 1)) return;
    if (ctrl >= 0x111 && ctrl <= 0x11f) {
      const int n = ctrl - 0x110;
      if (inrow >= n) r = (int)(uint32_t)s[l - n];
    } else if (ctrl == 0x142) {
      if (row >= 1) r = (int)(uint32_t)s[row * 16 - 1];
    } else if (ctrl == 0x143) {
      r = (int)(uint32_t)s[31];
    } else {
      std::fprintf(stderr, "wave_emu: DPP control 0x%x not emulated\n", ctrl);
      std::abort();
    }
  });
  return r;
}

// s_waitcnt and memory fences are wave-wide instructions on the hardware: they order the memory operations of
// all 64 lanes. Emulated, they are a barrier of the wave (plus a fence).
#define __builtin_amdgcn_s_waitcnt(x) (emu::ctx.wave->barrier())
// The hardware executes a wave's lanes in lockstep; the emulated lanes are free-running threads, so the
// device code's wave-level LDS fences (wave_lds_sync) become real 64-lane barriers here.
#define __builtin_amdgcn_wave_barrier() (emu::ctx.wave->barrier())
#define __builtin_amdgcn_fence(order, scope) \
  (std::atomic_thread_fence(std::memory_order_seq_cst), emu::ctx.wave->barrier())
inline void __builtin_amdgcn_s_sleep(int) { std::this_thread::yield(); }
inline uint64_t __builtin_amdgcn_s_memrealtime() {
  return (uint64_t)std::chrono::duration_cast<std::chrono::nanoseconds>(
             std::chrono::steady_clock::now().time_since_epoch()).count() / 10ull;
}

#ifndef __HIP_MEMORY_SCOPE_AGENT
#define __HIP_MEMORY_SCOPE_WAVEFRONT 2
#define __HIP_MEMORY_SCOPE_WORKGROUP 3
#define __HIP_MEMORY_SCOPE_AGENT 4
#define __HIP_MEMORY_SCOPE_SYSTEM 5
#endif
#if !defined(__clang__)  // clang (also for the host) has the __hip_atomic_* builtins itself
template <class T>
inline T __hip_atomic_load(const T* p, int, int) { return __atomic_load_n(p, __ATOMIC_SEQ_CST); }
template <class T, class V>
inline void __hip_atomic_store(T* p, V v, int, int) { __atomic_store_n(p, (T)v, __ATOMIC_SEQ_CST); }
template <class T, class V>
inline T __hip_atomic_fetch_add(T* p, V v, int, int) { return __atomic_fetch_add(p, (T)v, __ATOMIC_SEQ_CST); }
template <class T, class V>
inline T __hip_atomic_fetch_min(T* p, V v, int, int) {
  T cur = __atomic_load_n(p, __ATOMIC_SEQ_CST);
  while ((T)v < cur && !__atomic_compare_exchange_n(p, &cur, (T)v, false, __ATOMIC_SEQ_CST, __ATOMIC_SEQ_CST)) {
  }
  return cur;
}
template <class T, class V>
inline T __hip_atomic_exchange(T* p, V v, int, int) { return __atomic_exchange_n(p, (T)v, __ATOMIC_SEQ_CST); }
template <class T, class V>
inline bool __hip_atomic_compare_exchange_strong(T* p, T* expected, V desired, int, int, int) {
  return __atomic_compare_exchange_n(p, expected, (T)desired, false, __ATOMIC_SEQ_CST, __ATOMIC_SEQ_CST);
}
#endif
inline unsigned atomicAdd(unsigned* p, unsigned v) { return __atomic_fetch_add(p, v, __ATOMIC_SEQ_CST); }
inline unsigned atomicOr(unsigned* p, unsigned v) { return __atomic_fetch_or(p, v, __ATOMIC_SEQ_CST); }
inline unsigned atomicMax(unsigned* p, unsigned v) {
  unsigned cur = __atomic_load_n(p, __ATOMIC_SEQ_CST);
  while (v > cur && !__atomic_compare_exchange_n(p, &cur, v, false, __ATOMIC_SEQ_CST, __ATOMIC_SEQ_CST)) {
  }
  return cur;
}
inline unsigned atomicMin(unsigned* p, unsigned v) {
  unsigned cur = __atomic_load_n(p, __ATOMIC_SEQ_CST);
  while (v < cur && !__atomic_compare_exchange_n(p, &cur, v, false, __ATOMIC_SEQ_CST, __ATOMIC_SEQ_CST)) {
  }
  return cur;
}
inline int __popcll(unsigned long long x) { return __builtin_popcountll(x); }
inline int __ffsll(unsigned long long x) { return __builtin_ffsll((long long)x); }
inline int __clz(unsigned x) { return x ? __builtin_clz(x) : 32; }
using std::max;
using std::min;

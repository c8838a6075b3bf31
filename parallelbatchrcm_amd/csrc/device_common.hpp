// device_common.hpp -- gfx950 building blocks shared by the reorder kernels: cross-workgroup memory
// ordering, wave64 ballot/prefix primitives and block-wide scans.
#pragma once
#ifndef RCM_WAVE_EMU  // the CPU wave emulator of the tests provides these builtins itself
#include <hip/hip_runtime.h>
#endif
#include <stdint.h>

// Workgroup-shared (LDS) variable declaration.
#ifndef RCM_LDS
#define RCM_LDS(T, name) __shared__ T name
#endif

namespace rcm {

constexpr int kWave = 64;

// ------------------------------------------------------------------------------------------------------
// Cross-workgroup communication.
//
// Everything one workgroup writes for another (queue entries, permutation slots, batch records, status
// words, claim words) is written with agent-scope relaxed atomic stores and read with agent-scope relaxed
// atomic loads. On gfx950 these compile to global_store/global_load with the sc1 bit: they are coherent
// across the eight XCD L2 caches without any cache maintenance. A producer orders its data before the
// flag that publishes it with `publish_wait()` (s_waitcnt: every store of the wave has been acknowledged)
// in every wave, a workgroup barrier, and then the flag store. A consumer issues its data loads after it
// has observed the flag (the loads are control-dependent on it).
//
// This avoids buffer_wbl2 / buffer_inv, which a formal release/acquire at agent scope emits and which cost
// 0.5-9 us each on MI355X. RCM_STRICT_FENCES selects the formal fences instead (used by tests to show both
// give the same results).
// ------------------------------------------------------------------------------------------------------
template <class T>
__device__ __forceinline__ T ld_sync(const T* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <class T>
__device__ __forceinline__ void st_sync(T* p, T v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <bool STRICT>
__device__ __forceinline__ void publish_wait() {
  if constexpr (STRICT)
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
  else
    __builtin_amdgcn_s_waitcnt(0);
}
template <bool STRICT>
__device__ __forceinline__ void consume_fence() {
  if constexpr (STRICT) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
}
__device__ __forceinline__ uint64_t atomic_min_u64(uint64_t* p, uint64_t v) {
  return __hip_atomic_fetch_min(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint32_t atomic_min_u32(uint32_t* p, uint32_t v) {
  return __hip_atomic_fetch_min(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint64_t atomic_add_u64(uint64_t* p, uint64_t v) {
  return __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// The same without a result: the compiler issues the atomic without a return, so nothing waits for it.
__device__ __forceinline__ void atomic_add_u64_nowait(uint64_t* p, uint64_t v) {
  (void)__hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint32_t atomic_add_u32(uint32_t* p, uint32_t v) {
  return __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// 100 MHz constant-rate clock (independent of shader clock), used by the watchdog.
__device__ __forceinline__ uint64_t wall_ticks() { return __builtin_amdgcn_s_memrealtime(); }

// ------------------------------------------------------------------------------------------------------
// Wave64 primitives.
// ------------------------------------------------------------------------------------------------------
__device__ __forceinline__ int lane_id() { return (int)(threadIdx.x & (kWave - 1)); }
__device__ __forceinline__ int wave_id() { return (int)(threadIdx.x / kWave); }
__device__ __forceinline__ uint64_t lanemask_lt() { return (1ull << lane_id()) - 1ull; }
__device__ __forceinline__ uint64_t ballot(bool p) { return __ballot(p); }
// The first active lane's value for every lane. Used on values that are the same in every lane on the device
// (one wave-wide load of one address, s_memrealtime) so that the control flow they steer is visibly uniform.
__device__ __forceinline__ uint32_t wave_uniform(uint32_t v) { return (uint32_t)__builtin_amdgcn_readfirstlane((int)v); }

// DPP row moves (gfx9 family): row_shr:n = 0x110 + n within rows of 16 lanes, row_bcast:15 = 0x142 (lane 15 of
// each row to the next row), row_bcast:31 = 0x143 (lane 31 to rows 2 and 3). Lanes without a source, or
// outside ROW_MASK, receive `identity`. One VALU instruction each, against an LDS round trip for a shuffle.
template <int CTRL, int ROW_MASK = 0xf>
__device__ __forceinline__ uint32_t dpp_move(uint32_t v, uint32_t identity) {
  return (uint32_t)__builtin_amdgcn_update_dpp((int)identity, (int)v, CTRL, ROW_MASK, 0xf, false);
}
__device__ __forceinline__ uint32_t lane63(uint32_t v) { return (uint32_t)__builtin_amdgcn_readlane((int)v, kWave - 1); }

// Inclusive prefix sum over the wave: Hillis-Steele within each row of 16, then the row totals.
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) {
  v += dpp_move<0x111>(v, 0u);
  v += dpp_move<0x112>(v, 0u);
  v += dpp_move<0x114>(v, 0u);
  v += dpp_move<0x118>(v, 0u);
  v += dpp_move<0x142, 0xa>(v, 0u);
  v += dpp_move<0x143, 0xc>(v, 0u);
  return v;
}
// Inclusive prefix maximum over the wave (the same row moves; 0 is the identity for unsigned values).
__device__ __forceinline__ uint32_t wave_incl_max(uint32_t v) {
  v = max(v, dpp_move<0x111>(v, 0u));
  v = max(v, dpp_move<0x112>(v, 0u));
  v = max(v, dpp_move<0x114>(v, 0u));
  v = max(v, dpp_move<0x118>(v, 0u));
  v = max(v, dpp_move<0x142, 0xa>(v, 0u));
  v = max(v, dpp_move<0x143, 0xc>(v, 0u));
  return v;
}
// Reductions: the same prefix pattern with max / min; lane 63 holds the result, returned to every lane.
__device__ __forceinline__ uint32_t wave_sum(uint32_t v) { return lane63(wave_incl_scan(v)); }
__device__ __forceinline__ uint32_t wave_max(uint32_t v) {
  v = max(v, dpp_move<0x111>(v, 0u));
  v = max(v, dpp_move<0x112>(v, 0u));
  v = max(v, dpp_move<0x114>(v, 0u));
  v = max(v, dpp_move<0x118>(v, 0u));
  v = max(v, dpp_move<0x142, 0xa>(v, 0u));
  v = max(v, dpp_move<0x143, 0xc>(v, 0u));
  return lane63(v);
}
__device__ __forceinline__ uint32_t wave_min(uint32_t v) {
  v = min(v, dpp_move<0x111>(v, ~0u));
  v = min(v, dpp_move<0x112>(v, ~0u));
  v = min(v, dpp_move<0x114>(v, ~0u));
  v = min(v, dpp_move<0x118>(v, ~0u));
  v = min(v, dpp_move<0x142, 0xa>(v, ~0u));
  v = min(v, dpp_move<0x143, 0xc>(v, ~0u));
  return lane63(v);
}

// Block-wide exclusive scan of one value per thread; `wtmp` needs W+1 words of LDS.
// Returns the exclusive prefix; *total receives the block sum. Contains two barriers.
template <int T>
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t* wtmp, uint32_t* total) {
  constexpr int W = T / kWave;
  const uint32_t inc = wave_incl_scan(v);
  if (lane_id() == kWave - 1) wtmp[wave_id()] = inc;
  __syncthreads();
  uint32_t base = 0, sum = 0;
#pragma unroll
  for (int w = 0; w < W; ++w) {
    const uint32_t s = wtmp[w];
    base += (w < wave_id()) ? s : 0u;
    sum += s;
  }
  __syncthreads();
  *total = sum;
  return base + inc - v;
}

// Exclusive scan of v and maximum of mx over the block in one pass; `wtmp` needs 2W words. Two barriers.
template <int T>
__device__ __forceinline__ uint32_t block_excl_scan_max(uint32_t v, uint32_t mx, uint32_t* wtmp, uint32_t* total,
                                                        uint32_t* max_out) {
  constexpr int W = T / kWave;
  const uint32_t inc = wave_incl_scan(v);
  const uint32_t wm = wave_max(mx);
  if (lane_id() == kWave - 1) {
    wtmp[wave_id()] = inc;
    wtmp[W + wave_id()] = wm;
  }
  __syncthreads();
  uint32_t base = 0, sum = 0, m = 0;
#pragma unroll
  for (int w = 0; w < W; ++w) {
    const uint32_t sw = wtmp[w];
    base += (w < wave_id()) ? sw : 0u;
    sum += sw;
    m = max(m, wtmp[W + w]);
  }
  __syncthreads();
  *total = sum;
  *max_out = m;
  return base + inc - v;
}

// Block-wide reduction (max) of one value per thread; `wtmp` needs W words. Two barriers.
template <int T>
__device__ __forceinline__ uint32_t block_max(uint32_t v, uint32_t* wtmp) {
  constexpr int W = T / kWave;
  v = wave_max(v);
  if (lane_id() == 0) wtmp[wave_id()] = v;
  __syncthreads();
  uint32_t r = 0;
#pragma unroll
  for (int w = 0; w < W; ++w) r = max(r, wtmp[w]);
  __syncthreads();
  return r;
}
template <int T>
__device__ __forceinline__ uint32_t block_min(uint32_t v, uint32_t* wtmp) {
  constexpr int W = T / kWave;
  v = wave_min(v);
  if (lane_id() == 0) wtmp[wave_id()] = v;
  __syncthreads();
  uint32_t r = 0xFFFFFFFFu;
#pragma unroll
  for (int w = 0; w < W; ++w) r = min(r, wtmp[w]);
  __syncthreads();
  return r;
}

__device__ __forceinline__ uint32_t bits_for(uint32_t maxval) {  // bits needed to hold 0..maxval
  return maxval == 0 ? 0u : 32u - (uint32_t)__clz(maxval);
}

}  // namespace rcm

// wave_dpp.hpp -- wave64 primitives built on DPP row moves and readlane (gfx9 family), for the one-wave
// worker engine (engine_wave.hpp).
#pragma once
#include "device_common.hpp"

namespace rcm {

// Orders this wave's LDS accesses across lanes: the compiler may not move memory operations across it, and
// the LDS executes one wave's operations in issue order.
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

// DPP controls (gfx9 family): row_shr:n = 0x110 + n, row_bcast:15 = 0x142, row_bcast:31 = 0x143.
template <int CTRL, int ROW_MASK = 0xf>
__device__ __forceinline__ uint32_t dpp(uint32_t v) {
  return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, CTRL, ROW_MASK, 0xf, false);
}

// Inclusive prefix sum over the 64 lanes (Hillis-Steele within rows of 16, then row broadcasts).
__device__ __forceinline__ uint32_t dpp_incl_scan(uint32_t v) {
  v += dpp<0x111>(v);
  v += dpp<0x112>(v);
  v += dpp<0x114>(v);
  v += dpp<0x118>(v);
  v += dpp<0x142, 0xa>(v);
  v += dpp<0x143, 0xc>(v);
  return v;
}
__device__ __forceinline__ uint32_t dpp_excl_scan(uint32_t v) { return dpp_incl_scan(v) - v; }
__device__ __forceinline__ uint32_t lane_value(uint32_t v, int lane) {
  return (uint32_t)__builtin_amdgcn_readlane((int)v, lane);
}
__device__ __forceinline__ uint32_t dpp_sum(uint32_t v) { return lane_value(dpp_incl_scan(v), kWave - 1); }
// Value of the previous lane (lane 0 receives `first`).
__device__ __forceinline__ uint32_t wave_shift_up(uint32_t v, uint32_t first) {
  const uint32_t u = (uint32_t)__shfl_up((int)v, 1, kWave);
  return lane_id() == 0 ? first : u;
}

}  // namespace rcm

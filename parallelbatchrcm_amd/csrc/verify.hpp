// verify.hpp -- device check that an order is the stable Cuthill-McKee order of a matrix from a start node.
//
// An order perm (perm[new] = old, all n nodes) with pos = perm^-1 is the order RealCPU::rcm_seq<true>
// (src/rcm_cpu.cpp:57-104) produces from `start` exactly when
//   1. perm[0] = start and perm is a bijection of [0, n);
//   2. every x != start has a parent P(x): the neighbour y != x (a node whose row holds x) of least pos(y),
//      and pos(P(x)) < pos(x);
//   3. with r(x) the first index of x in P(x)'s row, perm[1..] is strictly increasing in
//      (pos(P(x)), row length of x, r(x)) -- or in (pos(P(x)), r(x)) for the plain BFS order.
// The order these conditions allow is unique (the first place two such orders differed would hold two nodes
// with equal keys), so checking them checks the order bit for bit in O(n + nnz) work, without the serial
// algorithm. Device code; the checks run as four kernels over positions and rows.
#pragma once
#include "device_common.hpp"

namespace rcm {

constexpr uint32_t kVerifyNone = 0xFFFFFFFFu;

__global__ void k_vfy_fill(uint32_t* p, int64_t count, uint32_t v) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) p[i] = v;
}

// p[0, n) reversed in place.
__global__ void k_vfy_reverse(int32_t* p, int32_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n / 2) {
    const int32_t a = p[i];
    p[i] = p[n - 1 - i];
    p[n - 1 - i] = a;
  }
}

// pos[perm[i]] = i; bad = the least position that breaks condition 1. pos must start at -1 everywhere.
__global__ void k_vfy_positions(const int32_t* perm, int32_t n, int32_t start, int32_t* pos, uint32_t* bad) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int32_t x = perm[i];
  if (x < 0 || x >= n || (i == 0 && x != start)) {
    atomicMin(bad, (uint32_t)i);
    return;
  }
  int32_t expected = -1;
  if (!__hip_atomic_compare_exchange_strong(&pos[x], &expected, (int32_t)i, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                            __HIP_MEMORY_SCOPE_AGENT))
    atomicMin(bad, (uint32_t)max((int64_t)expected, i));  // the later of the two occurrences
}

// One wave per row y: ppos[x] = least pos(y) over rows y != x holding x.
__global__ void k_vfy_parents(const int32_t* rp, const int32_t* ci, int32_t n, const int32_t* pos, uint32_t* ppos) {
  const int64_t y = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  if (y >= n) return;
  const uint32_t py = (uint32_t)pos[y];
  for (int32_t k = rp[y] + lane_id(); k < rp[y + 1]; k += kWave) {
    const int32_t x = ci[k];
    if (x != (int32_t)y) atomicMin(&ppos[x], py);
  }
}

// One wave per row y: rix[x] = least row index of x in y's row, for the x whose parent y is.
__global__ void k_vfy_rowindex(const int32_t* rp, const int32_t* ci, int32_t n, const int32_t* pos,
                               const uint32_t* ppos, uint32_t* rix) {
  const int64_t y = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  if (y >= n) return;
  const uint32_t py = (uint32_t)pos[y];
  const int32_t b = rp[y];
  for (int32_t k = b + lane_id(); k < rp[y + 1]; k += kWave) {
    const int32_t x = ci[k];
    if (x != (int32_t)y && ppos[x] == py) atomicMin(&rix[x], (uint32_t)(k - b));
  }
}

// Conditions 2 and 3 at position i >= 1: the parent precedes, and the key of perm[i] exceeds that of perm[i-1].
__global__ void k_vfy_keys(const int32_t* rp, const int32_t* perm, int32_t n, const uint32_t* ppos,
                           const uint32_t* rix, int32_t bfs, uint32_t* bad) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < 1 || i >= n) return;
  const int32_t x = perm[i];
  const uint32_t px = ppos[x];
  if (px == kVerifyNone || (int64_t)px >= i) {
    atomicMin(bad, (uint32_t)i);
    return;
  }
  if (i < 2) return;
  const int32_t w = perm[i - 1];
  const uint32_t pw = ppos[w];
  const uint32_t dx = bfs ? 0u : (uint32_t)(rp[x + 1] - rp[x]), dw = bfs ? 0u : (uint32_t)(rp[w + 1] - rp[w]);
  const bool greater = px > pw || (px == pw && (dx > dw || (dx == dw && rix[x] > rix[w])));
  if (!greater) atomicMin(bad, (uint32_t)i);
}

}  // namespace rcm

// engine_wg.hpp -- one workgroup per matrix: a many-matrix engine for small matrices with short rows.
//
// NOT BUILT INTO THE LIBRARY. Its first GPU run (the batched start search on a small set) ended in a GPU memory
// fault whose cause has not been shown; the library's many-matrix forms use engine_many.hpp only. This file is
// compiled only by the CPU wave emulator tests, which keep its logic checked against the oracle.
//
// A persistent workgroup takes a whole matrix from the launch's order (atomicAdd ticket, largest first) and
// builds its order level by level with no other workgroup involved, so no batch records, look-back or
// device-scope publication are needed: a level costs a few L2 round trips and barriers instead of the batch
// chain's hop. The level's parents are processed in chunks (at most kWgT parents and kWgEntries adjacency
// entries):
//
//   1. claims: every entry k of the row of the parent at queue position q does
//      atomicMin(claim[c], q << 32 | k); positions of later chunks and levels are larger, so once a chunk has
//      claimed, its winners are final -- the first neighbour of c in queue order, at its first occurrence in
//      that row, as in the serial algorithm (src/rcm_cpu.cpp:57-104);
//   2. winners: an entry won when the claim word still holds its code and c is not the parent itself;
//   3. order: per parent, the winners ranked by (valence, k) (plain BFS: by k), placed after the winners of
//      the chunk's earlier parents -- the stable valence sort of the serial algorithm.
//
// The rows of a matrix given to this engine are at most kWgMaxRow long (checked on the host), so any single
// parent fits a chunk.
#pragma once
#include "engine_many.hpp"

namespace rcm {

constexpr int kWgT = 256;           // threads per workgroup, and parents per chunk
constexpr int kWgEntries = 4096;    // adjacency entries per chunk
constexpr int kWgMaxRow = 64;       // longest row a matrix of this engine may have

struct WgSmem {
  uint32_t woff[kWgT + 1];   // entry offsets of the chunk's parents
  int32_t pnode[kWgT];
  uint32_t prb[kWgT];
  uint32_t coff[kWgT + 1];   // winner offsets of the chunk's parents
  uint32_t ekey[kWgEntries]; // per entry: valence of the winner, or ~0 when the entry did not win
  int32_t echild[kWgEntries];
  uint32_t wtmp[2 * (kWgT / kWave) + 2];
  int32_t job;
  uint32_t chunk;            // parents in the current chunk
};

// Parent index i of chunk entry e: the largest i with woff[i] <= e.
__device__ __forceinline__ int32_t wg_parent(const uint32_t* woff, int32_t np, uint32_t e) {
  int32_t lo = 0, hi = np;
  while (hi - lo > 1) {
    const int32_t mid = (lo + hi) >> 1;
    if (woff[mid] <= e)
      lo = mid;
    else
      hi = mid;
  }
  return lo;
}

template <bool BFS>
__device__ void wg_order_job(const JobDesc& J, int32_t start, WgSmem& sm) {
  const int t = threadIdx.x;
  const int32_t n = J.n;
  const int32_t* rp = J.row_ptr;
  const int32_t* ci = J.col_idx;
  // reset: every claim word unclaimed, then the start
  for (int32_t v = t; v < n; v += kWgT) J.claim[v] = ~0ull;
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  if (t == 0) {
    J.claim[start] = 0ull;
    st_sync(&J.perm[0], start);
    J.qinfo[0] = (uint64_t)(uint32_t)rp[start] | ((uint64_t)(uint32_t)rp[start + 1] << 32);
    if (BFS) J.level[0] = 0;
  }
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  int64_t front = 0, out = 1;
  int32_t count = 1, lev = 0;
  while (count > 0) {
    const int64_t level_end = front + count;
    for (int64_t c0 = front; c0 < level_end;) {
      // parents of this chunk and their entry offsets; the chunk ends before kWgEntries entries
      const int32_t navail = (int32_t)min<int64_t>(kWgT, level_end - c0);
      uint32_t d = 0;
      int32_t p = 0;
      uint32_t rb = 0;
      if (t < navail) {
        p = ld_sync(&J.perm[c0 + t]);
        rb = (uint32_t)rp[p];
        d = (uint32_t)rp[p + 1] - rb;
      }
      uint32_t tot;
      const uint32_t off = block_excl_scan<kWgT>(d, sm.wtmp, &tot);
      if (t < navail) {
        sm.woff[t] = off;
        sm.pnode[t] = p;
        sm.prb[t] = rb;
      }
      if (t == 0) sm.woff[navail] = tot;
      __syncthreads();
      if (t == 0) {  // parents that fit: the largest np in [1, navail] with woff[np] <= kWgEntries
        int32_t lo = 1, hi = navail;  // woff[1] <= kWgMaxRow <= kWgEntries
        while (lo < hi) {
          const int32_t mid = (lo + hi + 1) >> 1;
          if (sm.woff[mid] <= (uint32_t)kWgEntries)
            lo = mid;
          else
            hi = mid - 1;
        }
        sm.chunk = (uint32_t)lo;
      }
      __syncthreads();
      const int32_t np = (int32_t)sm.chunk;
      const uint32_t E = sm.woff[np];
      // 1. claims
      for (uint32_t e = t; e < E; e += kWgT) {
        const int32_t i = wg_parent(sm.woff, np, e);
        const uint32_t k = e - sm.woff[i];
        const int32_t c = ci[sm.prb[i] + k];
        const uint64_t code = ((uint64_t)(c0 + i) << 32) | k;
        if (J.claim[c] > code) atomic_min_u64(&J.claim[c], code);  // codes only decrease: a stale read is safe
      }
      __builtin_amdgcn_s_waitcnt(0);
      __syncthreads();
      // 2. winners
      for (uint32_t e = t; e < E; e += kWgT) {
        const int32_t i = wg_parent(sm.woff, np, e);
        const uint32_t k = e - sm.woff[i];
        const int32_t c = ci[sm.prb[i] + k];
        const uint64_t code = ((uint64_t)(c0 + i) << 32) | k;
        const bool won = c != sm.pnode[i] && ld_sync(&J.claim[c]) == code;
        sm.ekey[e] = won ? (BFS ? 0u : (uint32_t)(rp[c + 1] - rp[c])) : ~0u;
        sm.echild[e] = c;
      }
      __syncthreads();
      // 3. winners per parent, their offsets, then each winner's rank within its parent
      uint32_t w = 0;
      if (t < np)
        for (uint32_t e = sm.woff[t]; e < sm.woff[t + 1]; ++e) w += sm.ekey[e] != ~0u ? 1u : 0u;
      uint32_t wtot;
      const uint32_t woffs = block_excl_scan<kWgT>(w, sm.wtmp, &wtot);
      if (t < np) sm.coff[t] = woffs;
      __syncthreads();
      for (uint32_t e = t; e < E; e += kWgT) {
        const uint32_t key = sm.ekey[e];
        if (key == ~0u) continue;
        const int32_t i = wg_parent(sm.woff, np, e);
        uint32_t rank = 0;
        for (uint32_t f = sm.woff[i]; f < sm.woff[i + 1]; ++f) {
          const uint32_t kf = sm.ekey[f];
          rank += (kf != ~0u && (kf < key || (kf == key && f < e))) ? 1u : 0u;
        }
        const int64_t q = out + sm.coff[i] + rank;
        const int32_t c = sm.echild[e];
        st_sync(&J.perm[q], c);
        J.qinfo[q] = (uint64_t)(uint32_t)rp[c] | ((uint64_t)(uint32_t)rp[c + 1] << 32);
        if (BFS) J.level[q] = lev + 1;
      }
      out += wtot;
      c0 += np;
      __syncthreads();  // the chunk's shared arrays are free again
    }
    __builtin_amdgcn_s_waitcnt(0);  // this level's queue entries are in memory before the next level reads them
    __syncthreads();
    count = (int32_t)(out - level_end);
    front = level_end;
    ++lev;
  }
  if (t == 0) {
    J.ctl->placed = (uint32_t)out;
    J.ctl->nbatches = (uint32_t)lev;
    J.ctl->error = ERR_NONE;
    J.ctl->done = 1u;
  }
}

// Launch reset: the list's control word (block 0) and the control block of every job in `order`.
__global__ void k_wg_prepare(ManyArgs m) {
  const int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (x == 0) {
    ManyControl z = {};
    *m.g = z;
  }
  if (x < m.count) {
    Control c = {};
    *m.jobs[m.order[x]].ctl = c;
  }
}

// Persistent: each workgroup takes matrices m.order[x] for x = atomicAdd(next_job) until the list is done.
template <bool BFS>
__global__ __launch_bounds__(kWgT) void k_order_wg(ManyArgs m) {
  RCM_LDS(WgSmem, sm);
  for (;;) {
    if (threadIdx.x == 0) {
      const uint32_t x = atomic_add_u32(&m.g->next_job, 1u);
      sm.job = x < (uint32_t)m.count ? m.order[x] : -1;
    }
    __syncthreads();
    const int32_t j = sm.job;
    __syncthreads();
    if (j < 0) break;
    wg_order_job<BFS>(m.jobs[j], m.starts[j], sm);
  }
}

}  // namespace rcm

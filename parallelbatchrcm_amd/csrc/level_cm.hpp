// level_cm.hpp -- level-synchronous Cuthill-McKee: one BFS level per step, with the level's children ordered
// by one device-wide stable radix sort. The engine variant for wide, shallow graphs (R-MAT and other
// power-law graphs), where the batch engine's per-batch latency chain is long relative to the work.
//
// Level L's nodes are perm[front, front + count). Every entry k of the row of the parent at position p claims
// its column c with atomicMin(claim[c], p << 32 | k): nodes placed in earlier levels hold smaller codes, so
// the winner is the first neighbour of c in queue order at its first occurrence in that row -- the parent
// the serial algorithm gives c (src/rcm_cpu.cpp:57-104). The winners, keyed by (parent rank in the level,
// valence of c, k), sorted stably, are level L + 1 in the serial order. The same steps with the key (parent
// rank, k) give the plain BFS order of the start search (src/simple_peripheral.cpp:44-90).
//
// Follows the level structure of the reference's GPU_SIMPLE path (src/device/rcm_gpu.cu:134-330), with the
// row index in the claim and the key so that the order is the serial one, duplicates included.
#pragma once
#include "matrix_ops.hpp"

namespace rcm_mx {
namespace {  // internal linkage: each library source that includes this has its own copy

__global__ __launch_bounds__(kT) void k_level_init(const int32_t* rp, int32_t n, int32_t start, uint64_t* claim,
                                                  int32_t* perm, uint64_t* qinfo, int32_t* level) {
  const int64_t v = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (v < n) claim[v] = v == start ? 0ull : ~0ull;
  if (v == 0) {
    perm[0] = start;
    qinfo[0] = (uint64_t)(uint32_t)rp[start] | ((uint64_t)(uint32_t)rp[start + 1] << 32);
    if (level) level[0] = 0;
  }
}

// work[i] = row length of the level's i-th node; work[count] = 0 (scanned into entry offsets)
__global__ __launch_bounds__(kT) void k_level_work(const int32_t* rp, const int32_t* perm, int64_t front,
                                                  int32_t count, uint32_t* work) {
  const int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (i < count) {
    const int32_t p = perm[front + i];
    work[i] = (uint32_t)(rp[p + 1] - rp[p]);
  } else if (i == count) {
    work[count] = 0;
  }
}

// The level entry e: its parent's rank i in the level (largest i with woff[i] <= e) and row index k.
__device__ __forceinline__ void level_entry(const uint32_t* woff, int32_t count, uint32_t e, int32_t& i, uint32_t& k) {
  int32_t lo = 0, hi = count;
  while (hi - lo > 1) {
    const int32_t mid = (lo + hi) >> 1;
    if (woff[mid] <= e)
      lo = mid;
    else
      hi = mid;
  }
  i = lo;
  k = e - woff[lo];
}

__global__ __launch_bounds__(kT) void k_level_claim(const int32_t* rp, const int32_t* ci, const int32_t* perm,
                                                   int64_t front, int32_t count, const uint32_t* woff, int64_t w,
                                                   uint64_t* claim) {
  const int64_t e = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (e >= w) return;
  int32_t i;
  uint32_t k;
  level_entry(woff, count, (uint32_t)e, i, k);
  const int32_t p = perm[front + i];
  const int32_t c = ci[rp[p] + (int32_t)k];
  const uint64_t code = ((uint64_t)(front + i) << 32) | k;
  if (claim[c] > code) rcm::atomic_min_u64(claim + c, code);  // codes only decrease: a stale read is safe
}

// flag[e] = 1 if entry e won its column (and the column is not the parent itself); flag[w] = 0
__global__ __launch_bounds__(kT) void k_level_won(const int32_t* rp, const int32_t* ci, const int32_t* perm,
                                                 int64_t front, int32_t count, const uint32_t* woff, int64_t w,
                                                 const uint64_t* claim, uint32_t* flag) {
  const int64_t e = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (e > w) return;
  if (e == w) {
    flag[w] = 0;
    return;
  }
  int32_t i;
  uint32_t k;
  level_entry(woff, count, (uint32_t)e, i, k);
  const int32_t p = perm[front + i];
  const int32_t c = ci[rp[p] + (int32_t)k];
  const uint64_t code = ((uint64_t)(front + i) << 32) | k;
  flag[e] = (c != p && claim[c] == code) ? 1u : 0u;
}

// Winners compacted by the scanned flags: key (i, valence, k) (or (i, k) for the plain BFS), payload = node.
__global__ __launch_bounds__(kT) void k_level_keys(const int32_t* rp, const int32_t* ci, const int32_t* perm,
                                                  int64_t front, int32_t count, const uint32_t* woff, int64_t w,
                                                  const uint32_t* pos, int dbits, int kbits, uint64_t* keys,
                                                  uint32_t* child) {
  const int64_t e = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (e >= w || pos[e + 1] == pos[e]) return;
  int32_t i;
  uint32_t k;
  level_entry(woff, count, (uint32_t)e, i, k);
  const int32_t p = perm[front + i];
  const int32_t c = ci[rp[p] + (int32_t)k];
  const uint64_t deg = dbits ? (uint64_t)(uint32_t)(rp[c + 1] - rp[c]) : 0ull;
  keys[pos[e]] = ((uint64_t)i << (dbits + kbits)) | (deg << kbits) | k;
  child[pos[e]] = (uint32_t)c;
}

__global__ __launch_bounds__(kT) void k_level_write(const int32_t* rp, const uint32_t* child, int64_t m, int64_t out,
                                                   int32_t lev, int32_t* perm, uint64_t* qinfo, int32_t* level) {
  const int64_t j = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (j >= m) return;
  const int32_t c = (int32_t)child[j];
  perm[out + j] = c;
  qinfo[out + j] = (uint64_t)(uint32_t)rp[c] | ((uint64_t)(uint32_t)rp[c + 1] << 32);
  if (level) level[out + j] = lev;
}

// ------------------------------------------------------------------------------------------------------
// Narrow levels on one workgroup. Repaired inputs often end in long chains of small levels (the repair links
// each isolated node to the node below it); one workgroup runs consecutive levels of at most kNarrowCount
// nodes, kNarrowWork adjacency entries and kNarrowKids children without returning to the host. A level above
// those limits ends the launch with the state at that level; claims are idempotent, so the host's wide-level
// path may redo a level whose claims the launch already made.
// ------------------------------------------------------------------------------------------------------
constexpr int kNarrowCount = kT;
constexpr int kNarrowWork = 8192;
constexpr int kNarrowKids = kT;

struct NarrowState {  // written by thread 0 at exit
  int64_t front, out;
  int32_t count, lev, done, pad;
};

struct NarrowSmem {
  uint32_t woff[kNarrowCount + 1];
  int32_t pnode[kNarrowCount];
  uint64_t key[kNarrowKids];
  uint32_t kid[kNarrowKids];
  uint32_t w[kWaves + 1];
  uint32_t nkids;
};

__global__ __launch_bounds__(kT) void k_level_narrow(const int32_t* rp, const int32_t* ci, int32_t* perm,
                                                    uint64_t* qinfo, uint64_t* claim, int32_t* level, int dbits,
                                                    int kbits, int64_t front, int64_t out, int32_t count,
                                                    int32_t lev, NarrowState* st) {
  RCM_LDS(NarrowSmem, sm);
  const int t = threadIdx.x;
  int32_t done = 0;
  for (;;) {
    if (count == 0) {
      done = 1;
      break;
    }
    if (count > kNarrowCount) break;
    uint32_t l = 0;
    int32_t p = 0;
    if (t < count) {
      p = rcm::ld_sync(perm + front + t);
      l = (uint32_t)(rp[p + 1] - rp[p]);
    }
    uint32_t w;
    const uint32_t ex = rcm::block_excl_scan<kT>(l, sm.w, &w);
    if (t < count) {
      sm.woff[t] = ex;
      sm.pnode[t] = p;
    }
    if (t == 0) {
      sm.woff[count] = w;
      sm.nkids = 0;
    }
    __syncthreads();
    if (w == 0) {
      done = 1;
      break;
    }
    if (w > (uint32_t)kNarrowWork) break;
    for (uint32_t e = t; e < w; e += kT) {
      int32_t i;
      uint32_t k;
      level_entry(sm.woff, count, e, i, k);
      const int32_t c = ci[rp[sm.pnode[i]] + (int32_t)k];
      const uint64_t code = ((uint64_t)(front + i) << 32) | k;
      if (rcm::ld_sync(claim + c) > code) rcm::atomic_min_u64(claim + c, code);
    }
    __builtin_amdgcn_s_waitcnt(0);  // every claim of this workgroup has reached memory
    __syncthreads();
    for (uint32_t e = t; e < w; e += kT) {
      int32_t i;
      uint32_t k;
      level_entry(sm.woff, count, e, i, k);
      const int32_t pn = sm.pnode[i];
      const int32_t c = ci[rp[pn] + (int32_t)k];
      const uint64_t code = ((uint64_t)(front + i) << 32) | k;
      if (c != pn && rcm::ld_sync(claim + c) == code) {
        const uint32_t slot = atomicAdd(&sm.nkids, 1u);
        if (slot < (uint32_t)kNarrowKids) {
          const uint64_t deg = dbits ? (uint64_t)(uint32_t)(rp[c + 1] - rp[c]) : 0ull;
          sm.key[slot] = ((uint64_t)i << (dbits + kbits)) | (deg << kbits) | k;
          sm.kid[slot] = (uint32_t)c;
        }
      }
    }
    __syncthreads();
    const uint32_t m = sm.nkids;
    if (m > (uint32_t)kNarrowKids) break;  // the host redoes this level
    if (m == 0) {
      done = 1;
      break;
    }
    if ((uint32_t)t < m) {  // keys are distinct: a key's rank is the number of smaller keys
      const uint64_t key = sm.key[t];
      uint32_t rank = 0;
      for (uint32_t j = 0; j < m; ++j) rank += sm.key[j] < key ? 1u : 0u;
      const int32_t c = (int32_t)sm.kid[t];
      rcm::st_sync(perm + out + rank, c);
      qinfo[out + rank] = (uint64_t)(uint32_t)rp[c] | ((uint64_t)(uint32_t)rp[c + 1] << 32);
      if (level) level[out + rank] = lev + 1;
    }
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    front += count;
    out += m;
    count = (int32_t)m;
    ++lev;
  }
  if (t == 0) {
    st->front = front;
    st->out = out;
    st->count = count;
    st->lev = lev;
    st->done = done;
  }
}

// The whole order from `start`. level (nullable) selects the plain BFS with levels. *placed: nodes ordered
// (n when the start's component is everything); *levels: BFS depth + 1. Returns kOpsOk, kOpsCapacity when a
// level's key does not fit 64 bits, or kOpsDevice.
constexpr int kOpsCapacity = -6;

template <class X>
int level_order(X& x, int32_t n, int64_t nnz, const int32_t* rp, const int32_t* ci, uint32_t maxdeg, int32_t start,
                int32_t* perm, uint64_t* qinfo, uint64_t* claim, int32_t* level, int32_t* placed, int32_t* levels) {
  const bool bfs = level != nullptr;
  uint32_t* work = x.template alloc<uint32_t>((size_t)n + 1);
  uint32_t* flag = x.template alloc<uint32_t>((size_t)nnz + 1);
  uint64_t* keys = x.template alloc<uint64_t>((size_t)n);
  uint64_t* keys2 = x.template alloc<uint64_t>((size_t)n);
  uint32_t* child = x.template alloc<uint32_t>((size_t)n);
  uint32_t* child2 = x.template alloc<uint32_t>((size_t)n);
  uint32_t* tot = x.template alloc<uint32_t>(2);
  if (x.failed()) return kOpsDevice;
  x.run(k_level_init, groups(n, kT), rp, n, start, claim, perm, qinfo, level);
  const int kbits = bits_of(maxdeg > 0 ? (uint64_t)maxdeg - 1 : 0);
  const int dbits = bfs ? 0 : bits_of((uint64_t)maxdeg);
  NarrowState* st = x.template alloc<NarrowState>(1);
  if (x.failed()) return kOpsDevice;
  int64_t front = 0, out = 1;
  int32_t count = 1, lev = 0;
  while (count > 0) {
    if (count <= kNarrowCount) {  // narrow levels on the device until one is wide or the order is complete
      x.run(k_level_narrow, 1, rp, ci, perm, qinfo, claim, level, dbits, kbits, front, out, count, lev, st);
      NarrowState h;
      x.read_words(reinterpret_cast<uint32_t*>(&h), reinterpret_cast<const uint32_t*>(st),
                   (int)(sizeof(NarrowState) / 4));
      if (x.failed()) return kOpsDevice;
      front = h.front;
      out = h.out;
      count = h.count;
      lev = h.lev;
      if (h.done) break;
    }
    x.run(k_level_work, groups((int64_t)count + 1, kT), rp, (const int32_t*)perm, front, count, work);
    scan_u32(x, work, work, (int64_t)count + 1, tot);
    const int64_t w = (int64_t)x.read_u32(tot);
    if (x.failed()) return kOpsDevice;
    if (w == 0) break;
    x.run(k_level_claim, groups(w, kT), rp, ci, (const int32_t*)perm, front, count, (const uint32_t*)work, w, claim);
    x.run(k_level_won, groups(w + 1, kT), rp, ci, (const int32_t*)perm, front, count, (const uint32_t*)work, w,
          (const uint64_t*)claim, flag);
    scan_u32(x, flag, flag, w + 1, tot + 1);
    const int64_t m = (int64_t)x.read_u32(tot + 1);
    if (x.failed()) return kOpsDevice;
    if (m == 0) break;
    const int ibits = bits_of((uint64_t)count - 1);
    if (ibits + dbits + kbits > 64) return kOpsCapacity;
    x.run(k_level_keys, groups(w, kT), rp, ci, (const int32_t*)perm, front, count, (const uint32_t*)work, w,
          (const uint32_t*)flag, dbits, kbits, keys, child);
    const int bits = ibits + dbits + kbits;
    uint64_t* k1 = keys;
    uint64_t* k2 = keys2;
    uint32_t* c1 = child;
    uint32_t* c2 = child2;
    radix_sort(x, k1, c1, k2, c2, m, bits < 32 ? bits : 32, bits > 32 ? bits - 32 : 0);
    ++lev;
    x.run(k_level_write, groups(m, kT), rp, (const uint32_t*)c1, m, out, lev, perm, qinfo, level);
    front += count;
    count = (int32_t)m;
    out += m;
  }
  *placed = (int32_t)out;
  *levels = lev + 1;
  return x.failed() ? kOpsDevice : kOpsOk;
}

}  // namespace
}  // namespace rcm_mx

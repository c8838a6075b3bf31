// engine_many.hpp -- many independent matrices reordered by one persistent launch (device code).
//
// Every matrix ("job") keeps its own queue, claim words, batch records and control block, exactly as a
// single-matrix run (engine.hpp); the batches of one job are processed by process_batch unchanged. What
// differs is how a workgroup finds its next batch:
//
//   * Non-blocking tickets. A workgroup takes batch t of job j only when that batch already exists: it reads
//     the job's ticket t and the record brange[t], and claims the batch with a compare-and-swap t -> t+1 on the
//     ticket. Batch ids are still handed out in queue order, so every predecessor of a claimed batch is held
//     by a running workgroup; no workgroup ever waits for a batch nobody can produce, whatever the residency.
//   * Wave-wide probing. Wave 0 probes 64 candidates per step: lane 0 the workgroup's current job, lanes
//     1..63 consecutive jobs of the begun set from a per-workgroup cursor. The lowest ready lane tries its
//     compare-and-swap; a lost race tries the next ready lane.
//   * Job admission. When nothing probed is ready, the workgroup begins the next job of the launch's order
//     (largest first, so long chains start early); once every job has begun it keeps probing.
//   * Completion. The first mark of a job's `done` flag (complete or failed) increments `jobs_done`; the
//     launch ends when it reaches the job count or the watchdog sets `abort`.
#pragma once
#include "engine.hpp"

namespace rcm {

// One matrix of a many-matrix launch: its views into the concatenated device buffers.
struct JobDesc {
  const int32_t* row_ptr;
  const int32_t* col_idx;
  int32_t* perm;
  uint64_t* qinfo;
  uint64_t* claim;
  uint64_t* bstat;
  uint64_t* brange;
  int32_t* level;  // plain-BFS sweeps
  int32_t* aux;    // n entries: inverse permutation / reversed order
  Control* ctl;
  int32_t n;
  int32_t pad;
};

struct ManyControl {
  uint32_t next_job;   // next position of `order` to begin
  uint32_t jobs_done;  // jobs finished (complete or failed)
  uint32_t abort;      // watchdog fired
  uint32_t pad;
};

struct ManyArgs {
  const JobDesc* jobs;
  const int32_t* order;   // the jobs of this launch, in the order they are begun
  const int32_t* starts;  // start node per job (indexed by job id)
  int32_t count;          // entries of `order`
  ManyControl* g;
  uint64_t budget_ticks;
  int32_t sort_mode;
  int32_t seg_threshold;
};

template <class G, int DIGIT>
struct ManySmem {
  Smem<G, DIGIT> s;
  int32_t cur;  // current job
  uint32_t cursor;
};

__device__ __forceinline__ RunArgs job_args(const ManyArgs& m, int32_t j, bool bfs) {
  const JobDesc& J = m.jobs[j];
  RunArgs a{};
  a.row_ptr = J.row_ptr;
  a.col_idx = J.col_idx;
  a.perm = J.perm;
  a.qinfo = J.qinfo;
  a.claim = J.claim;
  a.bstat = J.bstat;
  a.brange = J.brange;
  a.level = bfs ? J.level : nullptr;
  a.ctl = J.ctl;
  a.n = J.n;
  a.bcap = J.n + 1;
  a.budget_ticks = m.budget_ticks;
  a.sort_mode = m.sort_mode;
  a.seg_threshold = m.seg_threshold;
  a.prof = nullptr;
  a.jobs_done = &m.g->jobs_done;
  return a;
}

// Wave 0 of a workgroup: find and claim the next batch. Sets sm.quit = 1 when the launch is over; otherwise
// ms.cur = job and sm.b / sm.s / sm.e = the claimed batch.
template <class G, int DIGIT>
__device__ void find_work(const ManyArgs& m, ManySmem<G, DIGIT>& ms, uint64_t deadline) {
  Smem<G, DIGIT>& sm = ms.s;
  const int lane = lane_id();
  int32_t cur = ms.cur;
  uint32_t cursor = ms.cursor;
  const uint32_t count = (uint32_t)m.count;
  for (;;) {
    const uint32_t fin = wave_uniform(ld_sync(&m.g->jobs_done));
    const uint32_t ab = wave_uniform(ld_sync(&m.g->abort));
    bool stop = fin >= count || ab != 0;
    if (!stop && wave_uniform(wall_ticks() > deadline ? 1u : 0u)) {
      if (lane == 0) st_sync(&m.g->abort, 1u);
      stop = true;
    }
    if (stop) {
      if (lane == 0) sm.quit = 1;
      return;
    }
    const uint32_t begun = min(wave_uniform(ld_sync(&m.g->next_job)), count);
    int32_t j = -1;
    if (lane == 0)
      j = cur;
    else if (begun > 0)
      j = m.order[(cursor + (uint32_t)lane - 1u) % begun];
    uint32_t t = 0;
    uint64_t r = 0;
    bool ready = false;
    if (j >= 0) {
      const JobDesc& J = m.jobs[j];
      if (!ld_sync(&J.ctl->done)) {
        t = ld_sync(&J.ctl->ticket);
        if (t <= (uint32_t)J.n) {
          r = ld_sync(&J.brange[t]);
          ready = r != 0;
        }
      }
    }
    uint64_t mask = ballot(ready);
    while (mask) {
      const int l = __ffsll((unsigned long long)mask) - 1;
      bool won = false;
      if (lane == l) {
        uint32_t expected = t;
        won = __hip_atomic_compare_exchange_strong(&m.jobs[j].ctl->ticket, &expected, t + 1u, __ATOMIC_RELAXED,
                                                   __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (won) {
          ms.cur = j;
          ms.cursor = cursor;
          sm.b = (int32_t)t;
          sm.s = (int32_t)(uint32_t)r;
          sm.e = (int32_t)(r >> 32);
          sm.quit = 0;
        }
      }
      if (ballot(won)) return;
      mask &= mask - 1;
    }
    if (begun < count) {
      uint32_t idx = 0;
      if (lane == 0) idx = atomic_add_u32(&m.g->next_job, 1u);
      idx = wave_uniform(idx);
      if (idx < count) {
        cur = m.order[idx];
        continue;
      }
    }
    cursor += kWave - 1;
    __builtin_amdgcn_s_sleep(2);
  }
}

template <class G, int DIGIT, bool STRICT, bool BFS>
__global__ __launch_bounds__(G::T) void k_reorder_many(ManyArgs m) {
  using MS = ManySmem<G, DIGIT>;
  RCM_LDS(MS, ms);
  const int tid = threadIdx.x;
  const uint64_t deadline = wall_ticks() + m.budget_ticks;
  if (tid == 0) {
    ms.cur = -1;
    ms.cursor = (uint32_t)blockIdx.x * 13u;
  }
  __syncthreads();
  for (;;) {
    if (wave_id() == 0) find_work<G, DIGIT>(m, ms, deadline);
    __syncthreads();
    if (ms.s.quit) break;
    consume_fence<STRICT>();
    const RunArgs a = job_args(m, ms.cur, BFS);
    process_batch<G, DIGIT, STRICT, BFS, false>(a, ms.s, deadline, nullptr);
    __syncthreads();  // every wave is done with the batch before wave 0 overwrites its scalars
  }
}

// ---- helper kernels over the jobs of a launch: `per` workgroups per job (blockIdx.x = job slot * per + part)

// Reset of every job in `order` for a run from starts[job]; block 0 also resets the launch control.
__global__ void k_init_many(ManyArgs m, int per) {
  const int slot = (int)blockIdx.x / per, part = (int)blockIdx.x % per;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    ManyControl z = {};
    *m.g = z;
  }
  if (slot >= m.count) return;
  const int32_t j = m.order[slot];
  const JobDesc& J = m.jobs[j];
  const int32_t start = m.starts[j];
  const int64_t stride = (int64_t)per * blockDim.x;
  for (int64_t i = (int64_t)part * blockDim.x + threadIdx.x; i <= J.n; i += stride) {
    if (i < J.n) J.claim[i] = i == start ? 0ull : ~0ull;
    J.bstat[i] = 0ull;
    J.brange[i] = i == 0 ? (1ull << 32) : 0ull;
  }
  if (part == 0 && threadIdx.x == 0) {
    J.perm[0] = start;
    J.qinfo[0] = (uint64_t)(uint32_t)J.row_ptr[start] | ((uint64_t)(uint32_t)J.row_ptr[start + 1] << 32);
    J.level[0] = 0;
    Control c = {};
    c.progress = 1;
    *J.ctl = c;
  }
}

// CSR validation per job: out[2*job] = error flags (as k_validate), out[2*job+1] = max row length.
__global__ void k_validate_many(const JobDesc* jobs, int32_t count, int per, const int64_t* nnz, uint32_t* out) {
  const int j = (int)blockIdx.x / per, part = (int)blockIdx.x % per;
  if (j >= count) return;
  const JobDesc& J = jobs[j];
  const int64_t stride = (int64_t)per * blockDim.x;
  const int64_t t0 = (int64_t)part * blockDim.x + threadIdx.x;
  uint32_t bad = 0, mx = 0;
  for (int64_t i = t0; i < J.n; i += stride) {
    const int32_t b = J.row_ptr[i], e = J.row_ptr[i + 1];
    if (e < b) bad |= 1u;
    else mx = max(mx, (uint32_t)(e - b));
  }
  if (t0 == 0 && (J.row_ptr[0] != 0 || (int64_t)J.row_ptr[J.n] != nnz[j])) bad |= 2u;
  for (int64_t k = t0; k < nnz[j]; k += stride) {
    const int32_t c = J.col_idx[k];
    if (c < 0 || c >= J.n) bad |= 4u;
  }
  mx = wave_max(mx);
  if (bad) atomicOr(&out[2 * j], bad);
  if (lane_id() == 0 && mx) atomicMax(&out[2 * j + 1], mx);
}

// aux = inverse of perm (per job in `order`).
__global__ void k_inverse_many(const JobDesc* jobs, const int32_t* order, int32_t count, int per) {
  const int slot = (int)blockIdx.x / per, part = (int)blockIdx.x % per;
  if (slot >= count) return;
  const JobDesc& J = jobs[order[slot]];
  for (int64_t i = (int64_t)part * blockDim.x + threadIdx.x; i < J.n; i += (int64_t)per * blockDim.x)
    J.aux[J.perm[i]] = (int32_t)i;
}

// Bandwidth per job from aux (= inverse permutation): out[job] = max |inv[i] - inv[j]| over stored entries.
__global__ void k_bandwidth_many(const JobDesc* jobs, const int32_t* order, int32_t count, int per, uint32_t* out) {
  const int slot = (int)blockIdx.x / per, part = (int)blockIdx.x % per;
  if (slot >= count) return;
  const int32_t j = order[slot];
  const JobDesc& J = jobs[j];
  uint32_t bw = 0;
  for (int64_t i = (int64_t)part * blockDim.x + threadIdx.x; i < J.n; i += (int64_t)per * blockDim.x) {
    const int32_t ri = J.aux[i];
    for (int32_t k = J.row_ptr[i]; k < J.row_ptr[i + 1]; ++k) {
      const int32_t v = J.aux[J.col_idx[k]];
      bw = max(bw, (uint32_t)(v > ri ? v - ri : ri - v));
    }
  }
  bw = wave_max(bw);
  if (lane_id() == 0 && bw) atomicMax(&out[j], bw);
}

// aux = perm reversed (per job in `order`).
__global__ void k_reverse_many(const JobDesc* jobs, const int32_t* order, int32_t count, int per) {
  const int slot = (int)blockIdx.x / per, part = (int)blockIdx.x % per;
  if (slot >= count) return;
  const JobDesc& J = jobs[order[slot]];
  for (int64_t i = (int64_t)part * blockDim.x + threadIdx.x; i < J.n; i += (int64_t)per * blockDim.x)
    J.aux[i] = J.perm[J.n - 1 - i];
}

// Start search support: for each job slot, the last w = min(32, n-1) queue entries of a plain-BFS sweep:
// node, level and row length, into tail[slot * 96 + {0, 32, 64} + i] (one workgroup of 64 per slot).
__global__ void k_gather_tail(const JobDesc* jobs, const int32_t* order, int32_t count, int32_t* tail) {
  const int slot = (int)blockIdx.x;
  if (slot >= count) return;
  const JobDesc& J = jobs[order[slot]];
  const int32_t w = min(32, J.n - 1);
  const int i = (int)threadIdx.x;
  if (i < w) {
    const int32_t p = J.n - w + i;
    const uint64_t q = J.qinfo[p];
    tail[slot * 96 + i] = J.perm[p];
    tail[slot * 96 + 32 + i] = J.level[p];
    tail[slot * 96 + 64 + i] = (int32_t)((uint32_t)(q >> 32) - (uint32_t)q);
  }
}

}  // namespace rcm

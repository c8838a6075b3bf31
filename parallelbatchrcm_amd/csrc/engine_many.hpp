// engine_many.hpp -- many independent matrices reordered by one persistent launch (device code).
//
// Every matrix ("job") keeps its own queue, claim words, batch records and control block, exactly as a
// single-matrix run (engine.hpp); the batches of one job are processed by process_batch unchanged. What
// differs is how a workgroup finds its next batch:
//
//   * Helpers. A workgroup works for one job at a time (its "helper"); a job accepts at most cap helpers
//     (about one per 8192 nodes, at least one). A helper takes the job's tickets with atomicAdd, exactly as
//     the single-matrix kernel, and waits for its batch to be published; it leaves when the job is done. Every
//     ticket below a held ticket was taken by a running workgroup, and the smallest unprocessed ticket's batch
//     exists or never will (all earlier batches are complete), so no wait lasts forever.
//   * Joining. A free workgroup's wave 0 probes 63 jobs of the begun set per step (from a per-workgroup
//     cursor) and joins the first one below its cap whose next batch is already published; otherwise it
//     begins the next job of the launch's order (largest first, so long chains start early); otherwise it
//     joins any job below its cap; otherwise it backs off.
//   * Completion. The first mark of a job's `done` flag (complete or failed) increments `jobs_done`; the
//     launch ends when it reaches the job count or the watchdog sets `abort`. A free workgroup leaves earlier,
//     once every job has begun and it has seen each one done or at its cap (neither can change back).
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
  int32_t cap;  // many-matrix launches: at most this many workgroups help this job at once
};

struct ManyControl {
  uint32_t next_job;   // next position of `order` to begin
  uint32_t jobs_done;  // jobs finished (complete or failed)
  uint32_t abort;      // watchdog fired
  uint32_t pad;
};

struct ManyArgs {
  const JobDesc* jobs;
  uint32_t* helpers;      // per job (indexed by job id): workgroups working for it
  const int32_t* order;   // the jobs of this launch, in the order they are begun
  const int32_t* starts;  // start node per job (indexed by job id)
  int32_t count;          // entries of `order`
  ManyControl* g;
  uint64_t budget_ticks;
  int32_t sort_mode;
  int32_t seg_threshold;
  int32_t claim_filter;
  // A job of n rows may also take share_workers * n / share_rows helpers (its share of the launch's workgroups by
  // size; 0 = none): a launch over a few matrices, such as the last sweeps of a start search, then spreads its
  // workgroups over them instead of leaving most idle behind caps sized for thousands of matrices.
  uint32_t share_workers;
  uint64_t share_rows;
};

__device__ __forceinline__ uint32_t job_cap(const ManyArgs& m, const JobDesc& J) {
  uint32_t c = (uint32_t)J.cap;
  if (m.share_rows) c = max(c, (uint32_t)min((uint64_t)1024, (uint64_t)m.share_workers * (uint64_t)J.n / m.share_rows));
  return c;
}

template <class G, int DIGIT>
struct ManySmem {
  Smem<G, DIGIT> s;
  int32_t cur;  // current job
  uint32_t cursor;
};

__device__ __forceinline__ RunArgs job_args(const ManyArgs& m, int32_t j, bool levels) {
  const JobDesc& J = m.jobs[j];
  RunArgs a{};
  a.row_ptr = J.row_ptr;
  a.col_idx = J.col_idx;
  a.perm = J.perm;
  a.qinfo = J.qinfo;
  a.claim = J.claim;
  a.bstat = J.bstat;
  a.brange = J.brange;
  a.level = levels ? J.level : nullptr;
  a.ctl = J.ctl;
  a.n = J.n;
  a.bcap = J.n + 1;
  a.budget_ticks = m.budget_ticks;
  a.sort_mode = m.sort_mode;
  a.seg_threshold = m.seg_threshold;
  a.claim_filter = m.claim_filter;
  a.prof = nullptr;
  a.jobs_done = &m.g->jobs_done;
  return a;
}

// Wave 0 of a workgroup: the next batch. Sets sm.quit = 1 when the launch is over; otherwise ms.cur = job and
// sm.b / sm.s / sm.e = the batch. ms.cur persists across calls while the workgroup helps that job.
template <class G, int DIGIT>
__device__ void find_work(const ManyArgs& m, ManySmem<G, DIGIT>& ms, uint64_t deadline) {
  Smem<G, DIGIT>& sm = ms.s;
  const int lane = lane_id();
  int32_t cur = ms.cur;
  uint32_t cursor = ms.cursor;
  const uint32_t count = (uint32_t)m.count;
  uint32_t idle = 0;    // consecutive rounds without work: backoff grows with it
  uint32_t closed = 0;  // jobs seen closed (done or at their cap) in a row since every job was begun
  for (;;) {
    const uint32_t fin = wave_uniform(ld_sync(&m.g->jobs_done));
    const uint32_t ab = wave_uniform(ld_sync(&m.g->abort));
    bool stop = fin >= count || ab != 0;
    if (!stop && wave_uniform(wall_ticks() > deadline ? 1u : 0u)) {
      if (lane == 0) st_sync(&m.g->abort, 1u);
      stop = true;
    }
    if (stop) {
      if (lane == 0) {
        if (cur >= 0) atomic_add_u32(&m.helpers[cur], 0xFFFFFFFFu);
        ms.cur = -1;
        sm.quit = 1;
      }
      return;
    }
    if (cur >= 0) {
      // helping job cur: the next ticket, then wait for its batch (or for the job to finish); a batch kept by
      // its producer is skipped for the next ticket
      const JobDesc& J = m.jobs[cur];
      uint32_t t = 0;
      if (lane == 0) t = atomic_add_u32(&J.ctl->ticket, 1u);
      t = wave_uniform(t);
      for (uint32_t spin = 0;; ++spin) {
        uint64_t r = 0;
        if (t <= (uint32_t)J.n) r = ld_sync(&J.brange[t]);
        r = ((uint64_t)wave_uniform((uint32_t)(r >> 32)) << 32) | wave_uniform((uint32_t)r);
        if (r & kKeptBit) {
          if (lane == 0) t = atomic_add_u32(&J.ctl->ticket, 1u);
          t = wave_uniform(t);
          spin = 0;
          continue;
        }
        if (r) {
          if (lane == 0) {
            ms.cur = cur;
            ms.cursor = cursor;
            sm.b = (int32_t)t;
            sm.s = (int32_t)(uint32_t)r;
            sm.e = (int32_t)(r >> 32);
            sm.quit = 0;
          }
          return;
        }
        if ((spin + 1) % kQuitPoll == 0) {  // idle workers poll the done and abort words rarely
          if (wave_uniform(ld_sync(&J.ctl->done))) break;  // no batch t: the job has finished
          if (wave_uniform(run_complete(J.ctl) ? 1u : 0u)) {  // every entry written: mark the job done
            if (lane == 0) mark_done(J.ctl, &m.g->jobs_done);
            break;
          }
          if (wave_uniform(wall_ticks() > deadline ? 1u : 0u)) break;
          if (wave_uniform(ld_sync(&m.g->abort))) break;
        }
        __builtin_amdgcn_s_sleep(1);
      }
      if (lane == 0) atomic_add_u32(&m.helpers[cur], 0xFFFFFFFFu);  // leave the job
      cur = -1;
      continue;
    }
    // free: join a job
    const uint32_t begun = min(wave_uniform(ld_sync(&m.g->next_job)), count);
    int32_t j = -1;
    bool open = false, ready = false;
    if (lane >= 1 && (uint32_t)lane <= begun) {  // each begun job at most once per round
      j = m.order[(cursor + (uint32_t)lane - 1u) % begun];
      const JobDesc& J = m.jobs[j];
      if (!ld_sync(&J.ctl->done) && ld_sync(&m.helpers[j]) < job_cap(m, J)) {
        open = true;
        const uint32_t t = ld_sync(&J.ctl->ticket);
        ready = t <= (uint32_t)J.n && ld_sync(&J.brange[t]) != 0;
      }
    }
    cursor += kWave - 1;
    // Once every job has begun, a closed job never opens again: helpers only join below the cap and leave when
    // the job is done. A free workgroup that has seen every job closed can never get work, so it leaves the
    // launch (freeing its slot for other launches) instead of polling until the last job finishes.
    if (begun >= count) {
      closed = ballot(open) ? 0u : closed + min(begun, (uint32_t)(kWave - 1));
      if (closed >= count) {
        if (lane == 0) {
          ms.cur = -1;
          sm.quit = 1;
        }
        return;
      }
    }
    uint64_t cand = ballot(ready);
    if (!cand && begun < count) {
      uint32_t idx = 0;
      if (lane == 0) idx = atomic_add_u32(&m.g->next_job, 1u);
      idx = wave_uniform(idx);
      if (idx < count) {
        cur = m.order[idx];
        if (lane == 0) atomic_add_u32(&m.helpers[cur], 1u);
        idle = 0;
        continue;
      }
    }
    if (!cand) cand = ballot(open);
    bool joined = false;
    while (cand && !joined) {
      const int l = __ffsll((unsigned long long)cand) - 1;
      uint32_t ok = 0;
      if (lane == l) {
        const uint32_t was = atomic_add_u32(&m.helpers[j], 1u);
        ok = was < job_cap(m, m.jobs[j]) ? 1u : 0u;
        if (!ok) atomic_add_u32(&m.helpers[j], 0xFFFFFFFFu);  // full meanwhile
      }
      joined = ballot(ok != 0) != 0;
      if (joined) cur = (int32_t)__shfl((int)j, l, kWave);
      cand &= cand - 1;
    }
    if (joined) {
      idle = 0;
      continue;
    }
    // nothing to join: back off (64 to 8192 cycles) so idle workgroups do not crowd the memory channels
    idle = min(idle + 1u, 7u);
    for (uint32_t r = 0; r < (1u << idle); ++r) __builtin_amdgcn_s_sleep(1);
  }
}

// LEV: CM runs that also record BFS levels (the start search's speculated sweeps).
template <class G, int DIGIT, bool STRICT, bool BFS, bool LEV = false>
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
  if (tid == 0) ms.s.kept = 0;
  __syncthreads();
  for (;;) {
    if (!ms.s.kept) {  // a kept batch (the previous batch's first child, same job) needs no search
      if (wave_id() == 0) find_work<G, DIGIT>(m, ms, deadline);
      __syncthreads();
      if (ms.s.quit) break;
    }
    consume_fence<STRICT>();
    const RunArgs a = job_args(m, ms.cur, BFS || LEV);
    process_batch<G, DIGIT, STRICT, BFS, false, LEV>(a, ms.s, deadline, nullptr);
    __syncthreads();  // every wave is done with the batch before wave 0 overwrites its scalars
  }
}

// ---- helper kernels over the jobs of a list (`order`, `count` entries) ------------------------------------
// Workgroups are shared out by size: job slot k owns blocks [blk[k], blk[k+1]) of the grid.
__device__ __forceinline__ int32_t slot_of_block(const int32_t* blk, int32_t count, int32_t b, int32_t* part,
                                                 int32_t* per) {
  int32_t lo = 0, hi = count;  // blk[0] = 0 <= b < blk[count]
  while (hi - lo > 1) {
    const int32_t mid = (lo + hi) >> 1;
    if (blk[mid] <= b) lo = mid;
    else hi = mid;
  }
  *part = b - blk[lo];
  *per = blk[lo + 1] - blk[lo];
  return lo;
}

// Reset of every job in `order` for a run from starts[job]; block 0 also resets the launch control.
__global__ void k_init_many(ManyArgs m, const int32_t* blk) {
  int32_t part, per;
  const int32_t slot = slot_of_block(blk, m.count, (int32_t)blockIdx.x, &part, &per);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    ManyControl z = {};
    *m.g = z;
  }
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
    m.helpers[j] = 0u;
    J.perm[0] = start;
    J.qinfo[0] = (uint64_t)(uint32_t)J.row_ptr[start] | ((uint64_t)(uint32_t)J.row_ptr[start + 1] << 32);
    J.level[0] = 0;
    Control c = {};
    c.progress = 1;
    *J.ctl = c;
  }
}

// CSR validation per job: out[2*job] = error flags (as k_validate), out[2*job+1] = max row length.
__global__ void k_validate_many(const JobDesc* jobs, const int32_t* order, int32_t count, const int32_t* blk,
                                const int64_t* nnz, uint32_t* out) {
  int32_t part, per;
  const int32_t j = order[slot_of_block(blk, count, (int32_t)blockIdx.x, &part, &per)];
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

// aux = inverse of perm.
__global__ void k_inverse_many(const JobDesc* jobs, const int32_t* order, int32_t count, const int32_t* blk) {
  int32_t part, per;
  const JobDesc& J = jobs[order[slot_of_block(blk, count, (int32_t)blockIdx.x, &part, &per)]];
  for (int64_t i = (int64_t)part * blockDim.x + threadIdx.x; i < J.n; i += (int64_t)per * blockDim.x)
    J.aux[J.perm[i]] = (int32_t)i;
}

// Bandwidth per job from aux (= inverse permutation): out[job] = max |inv[i] - inv[j]| over stored entries.
__global__ void k_bandwidth_many(const JobDesc* jobs, const int32_t* order, int32_t count, const int32_t* blk,
                                 uint32_t* out) {
  int32_t part, per;
  const int32_t j = order[slot_of_block(blk, count, (int32_t)blockIdx.x, &part, &per)];
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

// aux = perm reversed.
__global__ void k_reverse_many(const JobDesc* jobs, const int32_t* order, int32_t count, const int32_t* blk) {
  int32_t part, per;
  const JobDesc& J = jobs[order[slot_of_block(blk, count, (int32_t)blockIdx.x, &part, &per)]];
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

// engine.hpp -- the speculative Cuthill-McKee reorder engine for gfx950 (device code).
//
// Algorithm (the speculative batch scheme of the reference's GPU_BATCH path,
// include/device/batchWorker.cuh:199-280 and include/device/batchQueue.cuh:37-192; re-designed here):
//
//   The permutation array is the BFS queue. A *batch* is a range [s, e) of queue positions whose nodes are
//   expanded together by one workgroup. Batch ids are allocated in queue order, so "earlier batch" means
//   "parents earlier in the queue". Persistent workgroups take batch ids from a ticket counter in order and
//   wait until the batch exists (its range record is published by the batch that produced its nodes).
//
//   Claims. The parent at queue position p claims its k-th row entry c with code (p << 32 | k) by a 64-bit
//   atomicMin on claim[c]. The smallest code wins, which is exactly the serial rule: a node belongs to the
//   earliest-dequeued parent whose row holds it, at that row's first occurrence (src/rcm_cpu.cpp:83-92).
//   Claims are speculative: a batch may claim before earlier batches have, and lose a node to them later.
//
//   Commit. Each batch publishes one 64-bit status word that moves through
//       DISC_LOCAL (own claims done) -> DISC_INC (all claims up to here done)
//       -> AGG (own child count, own child-batch count) -> INC (inclusive prefix of both).
//   Successors read 16 predecessor words per wave load (decoupled look-back). A batch whose
//   predecessors are all claim-complete before it claims ("early") owns every node its atomicMin won, after
//   an in-LDS removal of duplicates within the batch; otherwise it re-reads its claim words once the
//   look-back shows every predecessor has claimed.
//
//   Order. Owned children are sorted stably by (parent slot, child degree). Slots are generated in
//   (parent, row index) order, so this equals the serial per-parent std::stable_sort by valence
//   (src/rcm_cpu.cpp:94-96). The sort is an LDS radix sort whose ranks come from wave64 ballots, or, when
//   every parent row is short, a per-segment rank count; both are stable and give identical output.
//
//   Output. The INC prefix gives the batch's output offset and its first child-batch id. Children are
//   written to the queue with their row bounds (so the consumer needs no row_ptr lookup), then cut into
//   child batches (<= P parents, and either <= S adjacency slots or a single node of degree > S/2), whose
//   range records are published. A 64-bit progress word (nodes written | total) detects completion; the
//   total becomes known at the batch after which no further batch can exist.
#pragma once
#include "device_common.hpp"

namespace rcm {

// ---- status word: state (3 bits) | A (31 bits) | B (30 bits) -------------------------------------------
enum : uint32_t { ST_NONE = 0, ST_DLOC = 1, ST_DINC = 2, ST_AGG = 3, ST_INC = 4 };
__device__ __forceinline__ uint64_t mk_status(uint32_t st, uint32_t a, uint32_t b) {
  return ((uint64_t)st << 61) | ((uint64_t)(a & 0x7FFFFFFFu) << 30) | (uint64_t)(b & 0x3FFFFFFFu);
}
__device__ __forceinline__ uint32_t st_state(uint64_t w) { return (uint32_t)(w >> 61); }
__device__ __forceinline__ uint32_t st_a(uint64_t w) { return (uint32_t)((w >> 30) & 0x7FFFFFFFu); }
__device__ __forceinline__ uint32_t st_b(uint64_t w) { return (uint32_t)(w & 0x3FFFFFFFu); }

enum : uint32_t { ERR_NONE = 0, ERR_TIMEOUT = 1, ERR_CAPACITY = 2, ERR_LONGROW = 3 };

// Batch range records are begin | end << 32 (both < 2^30). A record with this bit set is kept by the workgroup
// that produced it: the holder of its ticket does not process it and takes the next ticket.
constexpr uint64_t kKeptBit = 1ull << 63;

// The ticket (an atomic per batch), the progress counter (an atomic per batch) and the done word (polled by idle
// workers) sit on separate 128-byte lines, so the atomics and polls on one do not queue behind the others.
struct alignas(128) Control {
  uint32_t ticket;    // next batch id handed to a worker
  uint32_t pad0[31];
  uint64_t progress;  // low 32: queue entries written; high 32: total to be written (0 = not yet known)
  uint32_t pad1[30];
  uint32_t done;      // set once when the run is complete or aborted
  uint32_t error;     // ERR_*
  uint32_t nbatches;  // batch records used by the finished run
  uint32_t placed;    // nodes placed by the finished run
  uint32_t diag[9];   // first error: site, batch, begin, end, slots, offset, count, first child id, children
};
static_assert(offsetof(Control, progress) == 128 && offsetof(Control, done) == 256, "one line each");

struct RunArgs {
  const int32_t* row_ptr;
  const int32_t* col_idx;
  int32_t* perm;          // queue; CM order on completion
  uint64_t* qinfo;        // per queue position: row_begin | row_end << 32
  uint64_t* claim;        // per node claim code
  uint64_t* bstat;        // per batch status word
  uint64_t* brange;       // per batch: begin | end << 32 (0 = not yet created)
  int32_t* level;         // per queue position BFS level (BFS runs, and CM runs of the kernels built to record them)
  Control* ctl;
  int32_t n;
  int32_t bcap;           // records in bstat / brange
  uint64_t budget_ticks;  // watchdog budget in 100 MHz ticks
  int32_t sort_mode;      // 0 auto, 1 radix, 2 segment rank
  int32_t seg_threshold;  // auto mode: segment rank when every parent row has at most this many entries
  uint64_t* prof;         // PROF builds: per-phase tick sums (kProfSlots entries)
  uint64_t* trace;        // PROF builds, nullable: kTraceWords words per batch id below trace_cap
  int32_t trace_cap;
  uint32_t* jobs_done;    // many-matrix launches: count of finished matrices (nullptr for one matrix)
  int32_t claim_filter;   // 0: atomicMin on every entry; 1: read the claim word first (plain load), 2: same
                          // with a device-scope load; the atomic is issued only if the word is still larger
};

// Claim word pre-read for the filter: codes only decrease, so any value read, stale or not, is >= the current
// one; a value <= our code proves the claim lost and the atomic can be skipped.
__device__ __forceinline__ uint64_t claim_peek(const RunArgs& a, int32_t c) {
  return a.claim_filter == 2 ? ld_sync(&a.claim[c]) : a.claim[c];
}

// Marks the run finished (complete or failed). The first marker also counts the matrix as finished for a
// many-matrix launch, so each matrix is counted exactly once.
__device__ __forceinline__ void mark_done(Control* ctl, uint32_t* jobs_done) {
  const uint32_t was = __hip_atomic_exchange(&ctl->done, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  if (was == 0 && jobs_done) atomic_add_u32(jobs_done, 1u);
}
__device__ __forceinline__ void mark_done(const RunArgs& a) { mark_done(a.ctl, a.jobs_done); }

// The run is complete when every queue entry is written: the progress word's low half (entries written) equals
// its high half (the total, set by the batch after which no batch can exist). Batches add to it without waiting
// for the result; the workers waiting for work test it and mark the run done.
__device__ __forceinline__ bool run_complete(const Control* ctl) {
  const uint64_t pr = ld_sync(&ctl->progress);
  return (pr >> 32) != 0 && (uint32_t)pr == (uint32_t)(pr >> 32);
}

// Phase profile (PROF builds): 100 MHz ticks summed over all batches, plus event counts.
enum : int {
  PF_ACQUIRE = 0,   // ticket taken -> batch record visible
  PF_PARENTS = 1,   // parent loads, claim check, slot offsets
  PF_CLAIM = 2,     // column loads, atomics, row bounds
  PF_FINAL = 3,     // in-batch dedupe (early) or wait for predecessors' claims + revalidation
  PF_SORT = 4,      // compaction and ordering
  PF_CUT = 5,       // child-batch cut
  PF_PREFIX = 6,    // publish aggregate, look back for the prefix
  PF_WRITE = 7,     // queue writes, child records, progress
  PF_BATCHES = 8,   // count of batches
  PF_EARLY = 9,     // count of early batches
  PF_CHILDREN = 10, // sum of owned children
  PF_SLOTS = 11,    // sum of adjacency slots
  PF_COMPACT = 12,  // compaction of owned children (PF_SORT is the ordering after it)
  PF_KEPT = 13,     // count of batches processed by their producer (no acquire)
  PF_TRACE0 = 16,   // thread 0's timestamps of the current batch (trace), 5 slots
  kProfSlots = 16,
  kProfRegs = 21
};

// Per-batch trace of PROF builds (RunArgs::trace): batch id | kept << 32 | early << 33, range s | e << 32,
// children | child batches << 32, first child batch id, then 100 MHz timestamps at entry, claims done, claim
// completeness known (finalise), cut done, prefix known, records published.
constexpr int kTraceWords = 12;

// Batch geometry: T threads, at most P parents and S adjacency slots per batch.
template <int T_, int S_, int P_, int QDIV_ = 2>
struct Geo {
  static constexpr int T = T_, S = S_, P = P_;
  static constexpr int W = T / kWave;
  static constexpr int EPT = S / T;
  static constexpr uint32_t Q = S / QDIV_;  // degree quantum of the child-batch cut rule
  static_assert(S % T == 0 && P <= T && S <= 65536, "geometry");
  static_assert(QDIV_ >= 2, "a multi-node child batch spans at most 2Q <= S slots");
};

template <class G, int DIGIT>
struct Smem {
  static constexpr int H = (1 << DIGIT) * G::EPT * G::W;
  static constexpr int HS = H > G::T ? H : G::T;
  int32_t pnode[G::P];
  uint32_t prb[G::P];
  int32_t plevel[G::P];
  uint32_t so[G::P + 1];
  int32_t sc[G::S];    // child id, -1 when not owned
  uint32_t srb[G::S];  // child row begin
  uint32_t sre[G::S];  // child row end
  alignas(16) uint16_t si[G::S];  // parent slot of each adjacency slot (before the claims: parent marks)
  uint16_t cpar[G::S / kWave];    // per 64-slot chunk: 1 + the parent slot owning the chunk's first slot
  uint64_t key[2][G::S];
  uint16_t val[2][G::S];
  uint32_t hist[HS];  // radix histogram; the segment sort keeps its segment starts and degree bases here
  uint32_t cinfo[2 * (G::S / kWave)];  // per 64-slot chunk: count, degree sum (compaction and cut exchanges)
  uint32_t wtmp[2 * G::W + 2];
  int32_t b, s, e, early, quit;
  int32_t kept;  // the batch in b / s / e is the previous batch's first child, its parents staged in key[kfin]
  int32_t kfin;
  int32_t kearly;  // the kept batch directly follows its producer: every batch before it has claimed
  uint32_t wearly[G::W];  // per wave: its look-back window, read before its atomics, showed every claim done
  uint32_t tick;  // the ticket counter, read during the look-back
  uint32_t cnt, nb, excl_off, excl_nb, maxdeg, maxpdeg;
  static_assert(HS >= 2 * G::P, "segment starts and degree bases of every parent fit in hist");
  static_assert(G::S == G::W * kWave * G::EPT, "compaction rounds cover the slots");
};

__device__ __forceinline__ void raise_error(const RunArgs& a, uint32_t err, uint32_t site = 0, uint32_t v0 = 0,
                                            uint32_t v1 = 0, uint32_t v2 = 0, uint32_t v3 = 0, uint32_t v4 = 0,
                                            uint32_t v5 = 0, uint32_t v6 = 0, uint32_t v7 = 0) {
  uint32_t expected = ERR_NONE;
  if (__hip_atomic_compare_exchange_strong(&a.ctl->error, &expected, err, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_AGENT)) {
    const uint32_t d[9] = {site, v0, v1, v2, v3, v4, v5, v6, v7};
    for (int i = 0; i < 9; ++i) st_sync(&a.ctl->diag[i], d[i]);
  }
  mark_done(a);
}

// Every thread of the workgroup: tells the caller to stop after a failed batch. The flag is written only once
// every thread has passed the barrier, so no wave can still be about to read its previous value.
template <class SM>
__device__ __forceinline__ void stop_workgroup(SM& sm) {
  __syncthreads();
  if (threadIdx.x == 0) sm.quit = 1;
  __syncthreads();
}

// Called by every lane of one wave. Returns true when the run must stop (done, or the watchdog fired).
__device__ __forceinline__ bool should_quit(const RunArgs& a, uint64_t deadline) {
  if (wave_uniform(ld_sync(&a.ctl->done))) return true;
  if (wave_uniform(wall_ticks() > deadline ? 1u : 0u)) {
    if (lane_id() == 0) raise_error(a, ERR_TIMEOUT);
    return true;
  }
  return false;
}

// ---- look-back over predecessor status words (one wave) -----------------------------------------------
// Claim completeness of every batch < b, without waiting. Lane l inspects batch (base - l); a position
// below 0 is the virtual predecessor of batch 0, which counts as complete.
// A look-back reads kLookWin status words per round trip (the other lanes stand in as neutral words), which
// keeps the polls of waiting batches on fewer of the lines their predecessors are writing. Measured on MI355X
// (3D 7-point 128^3, R-MAT 20): 16 words beat 64 by 3% and 8 by 1-3%.
constexpr int kLookWin = 16;

__device__ __forceinline__ bool claims_complete_before(const RunArgs& a, int64_t b) {
  int64_t base = b - 1;
  for (;;) {
    const int64_t idx = base - lane_id();
    const uint32_t st = lane_id() >= kLookWin ? (uint32_t)ST_DLOC
                        : idx >= 0            ? st_state(ld_sync(&a.bstat[idx]))
                                              : (uint32_t)ST_DINC;
    const uint64_t incm = ballot(st >= ST_DINC);
    const uint64_t notm = ballot(st < ST_DLOC);
    if (incm) {
      const uint64_t first = incm & (~incm + 1);
      return (notm & (first - 1)) == 0;
    }
    if (notm) return false;
    base -= kLookWin;
  }
}

// Waits check the run's done word and the watchdog on every kQuitPoll-th poll only, so a poll is one memory
// round trip instead of two; a stop is seen at most kQuitPoll polls late.
constexpr uint32_t kQuitPoll = 8;

// Blocking form; returns false if the run stopped while waiting.
__device__ __forceinline__ bool wait_claims_complete(const RunArgs& a, int64_t b, uint64_t deadline) {
  for (uint32_t spin = 1;; ++spin) {
    if (claims_complete_before(a, b)) return true;
    if (spin % kQuitPoll == 0 && should_quit(a, deadline)) return false;
    __builtin_amdgcn_s_sleep(1);
  }
}

// Exclusive prefix (queue offset, child-batch id) of batch b. The virtual predecessor of batch 0 is an
// inclusive record (1, 1): the start node occupies queue slot 0 and batch id 0.
__device__ __forceinline__ bool lookback_prefix(const RunArgs& a, int64_t b, uint64_t deadline, uint32_t* off,
                                                uint32_t* nbid) {
  int64_t base = b - 1;
  uint32_t acc_off = 0, acc_nb = 0;
  for (uint32_t spin = 1;;) {
    const int64_t idx = base - lane_id();
    const uint64_t w = lane_id() >= kLookWin ? mk_status(ST_AGG, 0u, 0u)
                       : idx >= 0            ? ld_sync(&a.bstat[idx])
                                             : mk_status(ST_INC, 1u, 1u);
    const uint32_t st = st_state(w);
    const uint64_t incm = ballot(st == ST_INC);
    const uint64_t notm = ballot(st < ST_AGG);
    if (incm) {
      const int first = __ffsll((unsigned long long)incm) - 1;
      if ((notm & ((1ull << first) - 1ull)) == 0) {
        const bool take = lane_id() <= first;
        acc_off += wave_sum(take ? st_a(w) : 0u);
        acc_nb += wave_sum(take ? st_b(w) : 0u);
        *off = acc_off;
        *nbid = acc_nb;
        return true;
      }
    } else if (notm == 0) {
      acc_off += wave_sum(st_a(w));
      acc_nb += wave_sum(st_b(w));
      base -= kLookWin;
      continue;
    }
    if (spin++ % kQuitPoll == 0 && should_quit(a, deadline)) return false;
    __builtin_amdgcn_s_sleep(1);
  }
}

// ---- stable LDS radix sort of (key, val) pairs, cnt <= S ---------------------------------------------
// Keys: low word = child degree (dbits significant), high word = parent slot (ibits significant). Passes
// run over the low word's digits, then the high word's. Elements are laid out striped (element e is held by
// thread e % T in round e / T), so element order is (round, wave, lane); ranks within a wave come from
// DIGIT ballots, and the histogram is scanned in (digit, round, wave) order, which keeps ties in input order.
template <class G, int DIGIT>
__device__ int lds_radix_sort(Smem<G, DIGIT>& sm, uint32_t cnt, uint32_t dbits, uint32_t ibits) {
  constexpr int R = 1 << DIGIT, EPT = G::EPT, W = G::W, T = G::T;
  constexpr int H = Smem<G, DIGIT>::H;
  constexpr int PER = (H + T - 1) / T;
  const int tid = threadIdx.x, wv = wave_id();
  const int rounds = (int)((cnt + T - 1) / T);
  const int npl = (int)((dbits + DIGIT - 1) / DIGIT), nph = (int)((ibits + DIGIT - 1) / DIGIT);
  int cur = 0;
  for (int p = 0; p < npl + nph; ++p) {
    const int shift = p < npl ? p * DIGIT : 32 + (p - npl) * DIGIT;
    for (int h = tid; h < H; h += T) sm.hist[h] = 0;
    __syncthreads();
    uint64_t kk[EPT];
    uint16_t vv[EPT];
    uint32_t dg[EPT], rk[EPT];
#pragma unroll
    for (int r = 0; r < EPT; ++r) {
      kk[r] = 0; vv[r] = 0; dg[r] = 0; rk[r] = 0;
      if (r < rounds) {
        const uint32_t e = (uint32_t)(r * T + tid);
        const bool act = e < cnt;
        const uint64_t k = act ? sm.key[cur][e] : 0ull;
        const uint32_t d = (uint32_t)(k >> shift) & (uint32_t)(R - 1);
        uint64_t peers = ballot(act);
#pragma unroll
        for (int bit = 0; bit < DIGIT; ++bit) {
          const bool on = (d >> bit) & 1u;
          const uint64_t bm = ballot(on);
          peers &= on ? bm : ~bm;
        }
        const uint32_t rank = (uint32_t)__popcll(peers & lanemask_lt());
        if (act && rank == 0) sm.hist[(d * EPT + r) * W + wv] = (uint32_t)__popcll(peers);
        kk[r] = k;
        vv[r] = act ? sm.val[cur][e] : (uint16_t)0;
        dg[r] = d;
        rk[r] = rank;
      }
    }
    __syncthreads();
    // exclusive scan of hist in index order: thread t owns entries [t*PER, t*PER+PER)
    uint32_t loc[PER];
    uint32_t lsum = 0;
#pragma unroll
    for (int q = 0; q < PER; ++q) {
      const int h = tid * PER + q;
      loc[q] = h < H ? sm.hist[h] : 0u;
      lsum += loc[q];
    }
    uint32_t tot;
    uint32_t run = block_excl_scan<T>(lsum, sm.wtmp, &tot);
#pragma unroll
    for (int q = 0; q < PER; ++q) {
      const int h = tid * PER + q;
      if (h < H) sm.hist[h] = run;
      run += loc[q];
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < EPT; ++r) {
      if (r < rounds) {
        const uint32_t e = (uint32_t)(r * T + tid);
        if (e < cnt) {
          const uint32_t pos = sm.hist[(dg[r] * EPT + r) * W + wv] + rk[r];
          sm.key[cur ^ 1][pos] = kk[r];
          sm.val[cur ^ 1][pos] = vv[r];
        }
      }
    }
    __syncthreads();
    cur ^= 1;
  }
  return cur;
}

// ---- stable segment-rank sort, with the degree prefix --------------------------------------------------------
// Compacted elements of one parent are contiguous (compaction keeps slot order): segment `seg` starts at
// hist[seg], and its children's degrees before it sum to hist[P + seg]. Each element reads its segment's keys
// U at a time (independent loads), counts those that order before it by (degree, position) and sums their
// degrees, then moves to segment start + rank with the key (exclusive degree prefix in sorted order << 32) |
// degree.
template <class G, int DIGIT>
__device__ int lds_segment_sort(Smem<G, DIGIT>& sm, uint32_t cnt) {
  constexpr uint32_t U = 4;
  for (uint32_t e = threadIdx.x; e < cnt; e += G::T) {
    const uint64_t k = sm.key[0][e];
    const uint32_t seg = (uint32_t)(k >> 32), deg = (uint32_t)k;
    const uint32_t lo = sm.hist[seg];
    uint32_t rank = 0, dsum = 0;
    for (uint32_t q0 = lo;; q0 += U) {
      uint64_t kq[U];
#pragma unroll
      for (uint32_t u = 0; u < U; ++u) kq[u] = sm.key[0][min(q0 + u, cnt - 1u)];  // in range, so they issue together
      bool all_same = true;
#pragma unroll
      for (uint32_t u = 0; u < U; ++u) {
        const uint32_t q = q0 + u, dq = (uint32_t)kq[u];
        const bool same = q < cnt && (uint32_t)(kq[u] >> 32) == seg;
        const bool before = same && (dq < deg || (dq == deg && q < e));
        rank += before ? 1u : 0u;
        dsum += before ? dq : 0u;
        all_same = all_same && same;
      }
      if (!all_same) break;  // the segment ended inside this group (segments are contiguous)
    }
    sm.key[1][lo + rank] = ((uint64_t)(sm.hist[G::P + seg] + dsum) << 32) | (uint64_t)deg;
    sm.val[1][lo + rank] = sm.val[0][e];
  }
  __syncthreads();
  return 1;
}

// ---- one parent whose row exceeds the batch's adjacency slots --------------------------------------------
// (The reference's discoverLong / constructHist / compactReader, include/device/batchWorker.cuh:627-735,
// 1430-1651, with an ordered compaction instead of arrival order.)
//
// The row is streamed through the workgroup in chunks of S entries, several times:
//   1. claims: atomicMin(claim[c], s << 32 | k) for every entry k;
//   2. once every earlier batch has finished claiming, claim[c] is final for this batch: entry k owns c iff
//      claim[c] still holds its code (later batches only write larger codes) and c is not the parent itself
//      (the start node's claim word is 0, the code of queue position 0, entry 0). A counting pass gives the
//      owned count, their degree sum, the number of degree > Q children and the largest degree;
//   3. the batch publishes AGG with an order-independent upper bound on its child batches, and the look-back
//      gives its output range [xo, xo + cnt) and first child-batch id xb;
//   4. the owned children are compacted in row order into the output range and sorted there by degree with
//      stable LSD passes of DIGIT bits (chunk ranks from ballots, as in lds_radix_sort); the range's perm
//      and qinfo words serve as the two buffers (qinfo is written last, from the sorted ids);
//   5. child batches are cut by the same rule as a normal batch, and records beyond the real cuts are
//      published as empty ranges, which the workers pass through the commit chain without work.
template <class G, int DIGIT>
__device__ __forceinline__ uint32_t row_degree(const RunArgs& a, int32_t c) {
  return (uint32_t)(a.row_ptr[c + 1] - a.row_ptr[c]);
}

// LEV: a CM run that also records BFS levels (the start search's speculated sweep); BFS runs always do.
template <class G, int DIGIT, bool STRICT, bool BFS, bool LEV = false>
__device__ void process_long_batch(const RunArgs& a, Smem<G, DIGIT>& sm, uint32_t b, int32_t s, int32_t parent,
                                   uint32_t rb, uint32_t D, int32_t plevel, bool early, uint64_t deadline) {
  constexpr int T = G::T, S = G::S, EPT = G::EPT, W = G::W;
  constexpr int R = 1 << DIGIT;
  constexpr int H = Smem<G, DIGIT>::H;  // = R * EPT * W
  const int tid = threadIdx.x, wv = wave_id();
  const uint64_t code_hi = (uint64_t)(uint32_t)s << 32;

  // 1. claims, S entries per chunk, loads staged before the atomics
  for (uint32_t c0 = 0; c0 < D; c0 += S) {
    int32_t cc[EPT];
#pragma unroll
    for (int k = 0; k < EPT; ++k) {
      const uint32_t j = c0 + (uint32_t)(tid + k * T);
      cc[k] = j < D ? a.col_idx[rb + j] : 0;
    }
    __builtin_amdgcn_s_waitcnt(0);
    uint64_t pre[EPT];
#pragma unroll
    for (int k = 0; k < EPT; ++k) {
      const uint32_t j = c0 + (uint32_t)(tid + k * T);
      pre[k] = (j < D && a.claim_filter) ? claim_peek(a, cc[k]) : ~0ull;
    }
#pragma unroll
    for (int k = 0; k < EPT; ++k) {
      const uint32_t j = c0 + (uint32_t)(tid + k * T);
      if (j < D && pre[k] > (code_hi | j)) (void)atomic_min_u64(&a.claim[cc[k]], code_hi | j);
    }
  }
  publish_wait<STRICT>();
  __syncthreads();
  if (!early) {
    if (tid == 0) st_sync(&a.bstat[b], mk_status(ST_DLOC, 0, 0));
    if (wave_id() == 0) {
      const bool ok = wait_claims_complete(a, (int64_t)b, deadline);
      if (lane_id() == 0) sm.quit = ok ? 0 : 1;
    }
    __syncthreads();
    if (sm.quit) return;
    consume_fence<STRICT>();
  }
  if (tid == 0) st_sync(&a.bstat[b], mk_status(ST_DINC, 0, 0));

  // 2. owned children: count, degree sum, big children, largest degree
  uint32_t lc = 0, ld = 0, lbig = 0, lmax = 0;
  for (uint32_t j = (uint32_t)tid; j < D; j += T) {
    const int32_t c = a.col_idx[rb + j];
    if (c != parent && ld_sync(&a.claim[c]) == (code_hi | j)) {
      const uint32_t dg = row_degree<G, DIGIT>(a, c);
      ++lc;
      ld += dg;
      lbig += dg > G::Q ? 1u : 0u;
      lmax = max(lmax, dg);
    }
  }
  uint32_t cnt, dtot, big;
  (void)block_excl_scan<T>(lc, sm.wtmp, &cnt);
  (void)block_excl_scan<T>(ld, sm.wtmp, &dtot);
  (void)block_excl_scan<T>(lbig, sm.wtmp, &big);
  const uint32_t maxdeg = block_max<T>(lmax, sm.wtmp);
  uint32_t nbu = 0;  // upper bound of the cut rule's batch count, independent of the order
  if (cnt > 0) {
    const uint32_t bound = (cnt + G::P - 1) / G::P + (dtot + G::Q - 1) / G::Q + 2 * big + 1;
    nbu = min(cnt, bound);
  }

  // 3. commit the counts, look back for the output range
  if (tid == 0) st_sync(&a.bstat[b], mk_status(ST_AGG, cnt, nbu));
  if (wave_id() == 0) {
    uint32_t xo = 0, xb = 0;
    const bool ok = lookback_prefix(a, (int64_t)b, deadline, &xo, &xb);
    if (lane_id() == 0) {
      sm.quit = ok ? 0 : 1;
      sm.excl_off = xo;
      sm.excl_nb = xb;
      if (ok) st_sync(&a.bstat[b], mk_status(ST_INC, xo + cnt, xb + nbu));
    }
  }
  __syncthreads();
  if (sm.quit) return;
  const uint32_t xo = sm.excl_off, xb = sm.excl_nb;
  if (xo + cnt > (uint32_t)a.n || xb + nbu > (uint32_t)a.bcap) {
    if (tid == 0) raise_error(a, ERR_CAPACITY, 4, b, (uint32_t)s, (uint32_t)s + 1, D, xo, cnt, xb, nbu);
    stop_workgroup(sm);
    return;
  }

  // 4. compaction in row order, then stable digit passes; buffers: A = perm words, B = qinfo words (as u32)
  uint32_t* bufA = reinterpret_cast<uint32_t*>(a.perm + xo);
  uint32_t* bufB = reinterpret_cast<uint32_t*>(a.qinfo + xo);
  const int passes = BFS ? 0 : (int)((bits_for(maxdeg) + DIGIT - 1) / DIGIT);
  uint32_t* dst = (passes & 1) ? bufB : bufA;
  {
    uint32_t base = 0;
    for (uint32_t c0 = 0; c0 < D; c0 += S) {  // blocked: thread t scans entries c0 + t*EPT .. + EPT - 1
      int32_t cc[EPT];
      bool own[EPT];
      uint32_t n_own = 0;
#pragma unroll
      for (int k = 0; k < EPT; ++k) {
        const uint32_t j = c0 + (uint32_t)(tid * EPT + k);
        own[k] = false;
        cc[k] = 0;
        if (j < D) {
          cc[k] = a.col_idx[rb + j];
          own[k] = cc[k] != parent && ld_sync(&a.claim[cc[k]]) == (code_hi | j);
          n_own += own[k] ? 1u : 0u;
        }
      }
      uint32_t tot;
      uint32_t pos = base + block_excl_scan<T>(n_own, sm.wtmp, &tot);
#pragma unroll
      for (int k = 0; k < EPT; ++k)
        if (own[k]) {
          if constexpr (BFS) {
            st_sync(&a.perm[xo + pos], cc[k]);
            st_sync(&a.qinfo[xo + pos], (uint64_t)(uint32_t)a.row_ptr[cc[k]] |
                                            ((uint64_t)(uint32_t)a.row_ptr[cc[k] + 1] << 32));
            st_sync(&a.level[xo + pos], plevel + 1);
          } else {
            st_sync(&dst[pos], (uint32_t)cc[k]);
          }
          ++pos;
        }
      base += tot;
    }
  }
  uint32_t* base = reinterpret_cast<uint32_t*>(sm.prb);  // R running digit offsets
  for (int p = 0; p < passes; ++p) {
    uint32_t* src = dst;
    dst = (src == bufA) ? bufB : bufA;
    const uint32_t shift = (uint32_t)(p * DIGIT);
    publish_wait<STRICT>();
    __syncthreads();
    // digit histogram of the whole list
    for (int h = tid; h < R; h += T) sm.hist[h] = 0;
    __syncthreads();
    for (uint32_t e = (uint32_t)tid; e < cnt; e += T) {
      const uint32_t dg = (row_degree<G, DIGIT>(a, (int32_t)ld_sync(&src[e])) >> shift) & (uint32_t)(R - 1);
      atomicAdd(&sm.hist[dg], 1u);
    }
    __syncthreads();
    if (tid == 0) {
      uint32_t run = 0;
      for (int h = 0; h < R; ++h) {
        base[h] = run;
        run += sm.hist[h];
      }
    }
    __syncthreads();
    // chunks of S elements: element e = c0 + r*T + tid; rank among equal digits in (r, wave, lane) order
    for (uint32_t c0 = 0; c0 < cnt; c0 += S) {
      for (int h = tid; h < H; h += T) sm.hist[h] = 0;
      __syncthreads();
      uint32_t id[EPT], dg[EPT], rk[EPT];
#pragma unroll
      for (int r = 0; r < EPT; ++r) {
        const uint32_t e = c0 + (uint32_t)(r * T + tid);
        const bool act = e < cnt;
        id[r] = act ? ld_sync(&src[e]) : 0u;
        dg[r] = act ? (row_degree<G, DIGIT>(a, (int32_t)id[r]) >> shift) & (uint32_t)(R - 1) : 0u;
        uint64_t peers = ballot(act);
#pragma unroll
        for (int bit = 0; bit < DIGIT; ++bit) {
          const bool on = (dg[r] >> bit) & 1u;
          const uint64_t bm = ballot(on);
          peers &= on ? bm : ~bm;
        }
        rk[r] = (uint32_t)__popcll(peers & lanemask_lt());
        if (act && rk[r] == 0) sm.hist[(dg[r] * EPT + r) * W + wv] = (uint32_t)__popcll(peers);
      }
      __syncthreads();
      // exclusive scan of the chunk histogram in (digit, round, wave) order
      constexpr int PER = (H + T - 1) / T;
      uint32_t loc[PER];
      uint32_t lsum = 0;
#pragma unroll
      for (int q = 0; q < PER; ++q) {
        const int h = tid * PER + q;
        loc[q] = h < H ? sm.hist[h] : 0u;
        lsum += loc[q];
      }
      uint32_t ctot;
      uint32_t run = block_excl_scan<T>(lsum, sm.wtmp, &ctot);
#pragma unroll
      for (int q = 0; q < PER; ++q) {
        const int h = tid * PER + q;
        if (h < H) sm.hist[h] = run;
        run += loc[q];
      }
      __syncthreads();
#pragma unroll
      for (int r = 0; r < EPT; ++r) {
        const uint32_t e = c0 + (uint32_t)(r * T + tid);
        if (e < cnt) {
          const uint32_t d0 = sm.hist[dg[r] * EPT * W];  // start of this digit's elements in the chunk
          const uint32_t pos = base[dg[r]] + (sm.hist[(dg[r] * EPT + r) * W + wv] - d0) + rk[r];
          st_sync(&dst[pos], id[r]);
        }
      }
      __syncthreads();
      if (tid < R) {  // advance the digit offsets by this chunk's counts
        const uint32_t d0 = sm.hist[tid * EPT * W];
        const uint32_t d1 = tid + 1 < R ? sm.hist[(tid + 1) * EPT * W] : ctot;
        base[tid] += d1 - d0;
      }
      __syncthreads();
    }
  }
  if constexpr (!BFS) {  // row bounds of the sorted children (qinfo words are free again)
    publish_wait<STRICT>();
    __syncthreads();
    for (uint32_t o = (uint32_t)tid; o < cnt; o += T) {
      const int32_t c = (int32_t)ld_sync(&bufA[o]);
      st_sync(&a.qinfo[xo + o], (uint64_t)(uint32_t)a.row_ptr[c] | ((uint64_t)(uint32_t)a.row_ptr[c + 1] << 32));
      if constexpr (LEV) st_sync(&a.level[xo + o], plevel + 1);
    }
  }
  publish_wait<STRICT>();
  __syncthreads();

  // 5. child batches: starts by the cut rule over the final order; records for every start, then padding
  uint32_t* cs = reinterpret_cast<uint32_t*>(sm.sc);  // starts of the current chunk
  uint32_t dx_base = 0, prev_deg = 0, nstarts = 0;
  int32_t pend = -1;  // last start whose end is not yet known
  for (uint32_t c0 = 0; c0 < cnt; c0 += T) {
    const uint32_t o = c0 + (uint32_t)tid;
    const bool act = o < cnt;
    const uint32_t dg = act ? row_degree<G, DIGIT>(a, ld_sync(&a.perm[xo + o])) : 0u;
    sm.sre[tid] = dg;
    uint32_t chunk_deg;
    const uint32_t dx = dx_base + block_excl_scan<T>(dg, sm.wtmp, &chunk_deg);  // has a barrier after sre
    const uint32_t pdeg = tid == 0 ? prev_deg : sm.sre[tid - 1];
    const uint32_t pdx = dx - pdeg;
    const bool st = act && (o == 0 || (dx / G::Q) != (pdx / G::Q) || (o / G::P) != ((o - 1) / G::P) ||
                            dg > G::Q || pdeg > G::Q);
    uint32_t nst;
    const uint32_t idx = block_excl_scan<T>(st ? 1u : 0u, sm.wtmp, &nst);
    if (st) cs[idx] = o;
    __syncthreads();
    if (nst > 0) {
      if (tid == 0 && pend >= 0)
        st_sync(&a.brange[xb + nstarts - 1], (uint64_t)(xo + (uint32_t)pend) | ((uint64_t)(xo + cs[0]) << 32));
      for (uint32_t t = (uint32_t)tid; t + 1 < nst; t += T)
        st_sync(&a.brange[xb + nstarts + t], (uint64_t)(xo + cs[t]) | ((uint64_t)(xo + cs[t + 1]) << 32));
      pend = (int32_t)cs[nst - 1];
      nstarts += nst;
    }
    dx_base += chunk_deg;
    prev_deg = sm.sre[T - 1];
    __syncthreads();
  }
  if (tid == 0 && pend >= 0)
    st_sync(&a.brange[xb + nstarts - 1], (uint64_t)(xo + (uint32_t)pend) | ((uint64_t)(xo + cnt) << 32));
  if (nstarts > nbu) {  // cannot happen: the bound counts every way a start can arise
    if (tid == 0) raise_error(a, ERR_CAPACITY, 5, b, (uint32_t)s, nstarts, nbu);
    stop_workgroup(sm);
    return;
  }
  const uint64_t empty = (uint64_t)(xo + cnt) | ((uint64_t)(xo + cnt) << 32);
  for (uint32_t t = nstarts + (uint32_t)tid; t < nbu; t += T) st_sync(&a.brange[xb + t], empty);
  if (tid == 0) {
    const bool final_batch = xb + nbu == b + 1;
    if (cnt > 0 || final_batch) {
      const uint64_t add = (uint64_t)cnt + (final_batch ? ((uint64_t)(xo + cnt) << 32) : 0ull);
      if (final_batch) {
        st_sync(&a.ctl->nbatches, b + 1);
        st_sync(&a.ctl->placed, xo + cnt);
      }
      atomic_add_u64_nowait(&a.ctl->progress, add);  // completion is seen by the workers waiting for work
    }
  }
}

// One wave's verdict on its look-back window `w` (lane l: the status word of batch b - 1 - l, neutral words
// beyond the window), written to the wave's LDS flag. Called after the wait that returns the window's loads and
// before the wave's atomics; the compiler barrier keeps both the loads and the atomics on their side of it, so
// the window is read before any atomic of the wave is performed.
template <class SM>
__device__ __forceinline__ void window_verdict(SM& sm, uint64_t w) {
  const uint32_t st = st_state(w);
  const uint64_t incm = ballot(st >= ST_DINC), notm = ballot(st < ST_DLOC);
  const bool ok = incm != 0 && (notm & ((incm & (~incm + 1)) - 1)) == 0;
  if (lane_id() == 0) sm.wearly[wave_id()] = ok ? 1u : 0u;
  __asm__ volatile("" ::: "memory");
}

// ---- one batch ------------------------------------------------------------------------------------------
template <class G, int DIGIT, bool STRICT, bool BFS, bool PROF, bool LEV = false>
__device__ void process_batch(const RunArgs& a, Smem<G, DIGIT>& sm, uint64_t deadline, uint64_t* pf) {
  constexpr int T = G::T, S = G::S, EPT = G::EPT, W = G::W;
  const int tid = threadIdx.x;
  const uint32_t b = (uint32_t)sm.b;
  const int32_t s = sm.s, e = sm.e;
  const int m = e - s;  // 0 for the padding records a long row may leave (see process_long_batch)
  if (m < 0 || m > G::P) {
    if (tid == 0) raise_error(a, ERR_CAPACITY, 1, b, (uint32_t)s, (uint32_t)e);
    stop_workgroup(sm);
    return;
  }

  // A. parents (written by the producing batch with sc1 stores, or staged in LDS when this workgroup produced
  //    them and kept the batch) and the non-blocking claim check. A kept batch skips the check: its
  //    predecessors were still claiming a moment ago, and the non-early path is correct in every case.
  const bool kept = sm.kept != 0;
  uint32_t d = 0;
  if (tid < m) {
    int32_t p;
    uint32_t rb, re;
    int32_t lev = 0;
    if (kept) {
      const uint64_t k0 = sm.key[sm.kfin][tid], k1 = sm.key[sm.kfin][G::P + tid];
      p = (int32_t)(uint32_t)k0;
      rb = (uint32_t)(k0 >> 32);
      re = (uint32_t)k1;
      lev = (int32_t)(uint32_t)(k1 >> 32);
    } else {
      p = ld_sync(&a.perm[s + tid]);
      const uint64_t q = ld_sync(&a.qinfo[s + tid]);
      rb = (uint32_t)q;
      re = (uint32_t)(q >> 32);
      if constexpr (BFS || LEV) lev = ld_sync(&a.level[s + tid]);
    }
    sm.pnode[tid] = p;
    sm.prb[tid] = rb;
    d = re - rb;
    if constexpr (BFS || LEV) sm.plevel[tid] = lev;
  }
  if (wave_id() == 0) {
    const bool ok = kept ? sm.kearly != 0 : claims_complete_before(a, (int64_t)b);
    if (lane_id() == 0) sm.early = ok ? 1 : 0;
  }
  // Slot -> parent marks for the claims: si[first slot of parent i] = i + 1, 0 on every other slot, and for each
  // 64-slot chunk the parent (+ 1) owning its first slot; a prefix maximum over a chunk's marks then gives every
  // slot its parent. The clear is ordered before the marks by the scan's barriers.
  {
    uint32_t* si32 = reinterpret_cast<uint32_t*>(sm.si);
    for (int q = tid * 4; q < S / 2; q += T * 4) {
      si32[q] = 0u;
      si32[q + 1] = 0u;
      si32[q + 2] = 0u;
      si32[q + 3] = 0u;
    }
  }
  uint32_t D, maxpdeg;
  const uint32_t soff = block_excl_scan_max<T>(d, d, sm.wtmp, &D, &maxpdeg);
  if (tid < m) sm.so[tid] = soff;
  if (tid == 0) {
    sm.so[m] = D;
    sm.kept = 0;  // every thread read it before the scan's barriers; set again when a child batch is kept
  }
  if (tid < m && d > 0) {
    if (soff < (uint32_t)S) sm.si[soff] = (uint16_t)(tid + 1);
    for (uint32_t c = (soff + kWave - 1) / kWave; c < (uint32_t)(S / kWave) && c * kWave < soff + d; ++c)
      sm.cpar[c] = (uint16_t)(tid + 1);
  }
  __syncthreads();
  if (D > (uint32_t)S) {
    // only a batch of one node can exceed the slots (the cut rule keeps multi-node batches below S)
    if (m != 1) {
      if (tid == 0) raise_error(a, ERR_LONGROW, 2, b, (uint32_t)s, (uint32_t)e, D);
      stop_workgroup(sm);
      return;
    }
    const int32_t plev = (BFS || LEV) ? sm.plevel[0] : 0;
    process_long_batch<G, DIGIT, STRICT, BFS, LEV>(a, sm, b, s, sm.pnode[0], sm.prb[0], D, plev, sm.early != 0,
                                                   deadline);
    return;
  }

  if constexpr (PROF) { if (tid == 0) { const uint64_t _t = wall_ticks(); pf[PF_TRACE0] = pf[15]; pf[PF_PARENTS] += _t - pf[15]; pf[15] = _t; } }
  // B. claim every adjacency slot (striped: thread t owns slots t, t+T, ..., so round k of wave w is the 64-slot
  //    chunk kW + w). Staged so that all column loads, then all atomics and row-bound loads of a thread are in
  //    flight together. A slot's parent is the prefix maximum of the chunk's parent marks.
  uint64_t code[EPT], oldv[EPT];
  int32_t cc[EPT];
  uint32_t crb[EPT], cre[EPT];
  uint16_t ci16[EPT];
  const uint32_t n_chunks = (wave_uniform(D) + kWave - 1) / kWave;
  const uint32_t wave0 = wave_uniform((uint32_t)wave_id());
  // A second chance for the early path: each wave reads one look-back window of status words in the same round
  // trip as its claim pre-reads (column loads without the filter), so the read is done before any of its
  // atomics issue. If every wave's window shows every earlier batch claim-complete, no earlier batch can take
  // a child from this batch any more, and the batch may publish DINC at once. Every wave reads its window,
  // those without slots too: for a batch without slots (the empty records a long row leaves) the windows are
  // all that justifies DINC.
  const int64_t widx = (int64_t)b - 1 - lane_id();
  uint64_t wword = lane_id() >= kLookWin ? mk_status(ST_DLOC, 0u, 0u) : mk_status(ST_DINC, 0u, 0u);
  const bool wload = lane_id() < kLookWin && widx >= 0;
  if (wload && !a.claim_filter) wword = ld_sync(&a.bstat[widx]);
#pragma unroll
  for (int k = 0; k < EPT; ++k) {
    code[k] = 0;
    cc[k] = 0;
    ci16[k] = 0;
    const uint32_t c = (uint32_t)(k * (T / kWave)) + wave0;
    if (c < n_chunks) {
      const uint32_t j = (uint32_t)(tid + k * T);
      uint32_t mk = j < D ? (uint32_t)sm.si[j] : 0u;
      if (lane_id() == 0) mk = max(mk, (uint32_t)sm.cpar[c]);
      mk = wave_incl_max(mk);
      if (j < D) {
        const uint32_t i = mk - 1u;
        const uint32_t r = j - sm.so[i];
        cc[k] = a.col_idx[sm.prb[i] + r];
        code[k] = ((uint64_t)(uint32_t)(s + (int32_t)i) << 32) | (uint64_t)r;
        ci16[k] = (uint16_t)i;
      }
    }
  }
  // One wait for every column load: afterwards the atomics below issue back to back (without it the
  // waitcnt pass cannot count loads across the guarded blocks and waits for everything before each atomic).
  __builtin_amdgcn_s_waitcnt(0);
  if (a.claim_filter) {
    // read the claim words first; atomics only where the word is still larger than our code, and the row bounds
    // (the child's degree) only for those slots, in the same round trip as their atomics: a slot whose pre-read
    // already lost can never own its child
    if (wload) wword = ld_sync(&a.bstat[widx]);
#pragma unroll
    for (int k = 0; k < EPT; ++k) {
      oldv[k] = 0;
      crb[k] = 0;
      cre[k] = 0;
      if ((uint32_t)(tid + k * T) < D) oldv[k] = claim_peek(a, cc[k]);
    }
    __builtin_amdgcn_s_waitcnt(0);
    window_verdict(sm, wword);
#pragma unroll
    for (int k = 0; k < EPT; ++k)
      if ((uint32_t)(tid + k * T) < D && oldv[k] > code[k]) {
        oldv[k] = atomic_min_u64(&a.claim[cc[k]], code[k]);
        crb[k] = (uint32_t)a.row_ptr[cc[k]];
        cre[k] = (uint32_t)a.row_ptr[cc[k] + 1];
      }
  } else {
    window_verdict(sm, wword);
#pragma unroll
    for (int k = 0; k < EPT; ++k) {
      oldv[k] = 0;
      crb[k] = 0;
      cre[k] = 0;
      if ((uint32_t)(tid + k * T) < D) {
        oldv[k] = atomic_min_u64(&a.claim[cc[k]], code[k]);
        crb[k] = (uint32_t)a.row_ptr[cc[k]];
        cre[k] = (uint32_t)a.row_ptr[cc[k] + 1];
      }
    }
  }
#pragma unroll
  for (int k = 0; k < EPT; ++k) {
    const uint32_t j = (uint32_t)(tid + k * T);
    if (j < D) {
      sm.sc[j] = oldv[k] > code[k] ? cc[k] : -1;
      sm.srb[j] = crb[k];
      sm.sre[j] = cre[k];
      sm.si[j] = ci16[k];
    }
  }
  __syncthreads();
  bool early = sm.early != 0;
  if (!early) {
    early = true;
#pragma unroll
    for (int w2 = 0; w2 < W; ++w2) early = early && sm.wearly[w2] != 0;
  }

  if constexpr (PROF) { if (tid == 0) { const uint64_t _t = wall_ticks(); pf[PF_TRACE0 + 1] = _t; pf[PF_CLAIM] += _t - pf[15]; pf[15] = _t; } }
  // won[k]: slot tid + kT owns its child. Decided in registers on the usual path, from the LDS words after the
  // in-batch correction on the early one.
  bool won[EPT];
#pragma unroll
  for (int k = 0; k < EPT; ++k) won[k] = (uint32_t)(tid + k * T) < D && oldv[k] > code[k];
  if (early) {
    // every earlier batch had finished claiming: a won atomicMin is final unless a slot of this same batch
    // with a smaller code took the node afterwards; that slot saw our code as its old value.
#pragma unroll
    for (int k = 0; k < EPT; ++k) {
      const uint32_t j = (uint32_t)(tid + k * T);
      if (j < D && oldv[k] > code[k]) {
        const uint32_t ohi = (uint32_t)(oldv[k] >> 32);
        if (oldv[k] != ~0ull && ohi >= (uint32_t)s && ohi < (uint32_t)e)
          sm.sc[sm.so[ohi - (uint32_t)s] + (uint32_t)oldv[k]] = -1;
      }
    }
    publish_wait<STRICT>();
    __syncthreads();
    if (tid == 0) st_sync(&a.bstat[b], mk_status(ST_DINC, 0, 0));
#pragma unroll
    for (int k = 0; k < EPT; ++k) won[k] = won[k] && sm.sc[tid + k * T] >= 0;
  } else {
    publish_wait<STRICT>();
    __syncthreads();
    if (tid == 0) st_sync(&a.bstat[b], mk_status(ST_DLOC, 0, 0));
    if (wave_id() == 0) {
      const bool ok = wait_claims_complete(a, (int64_t)b, deadline);
      if (lane_id() == 0) sm.quit = ok ? 0 : 1;
    }
    __syncthreads();
    if (sm.quit) return;
    consume_fence<STRICT>();
    if (tid == 0) st_sync(&a.bstat[b], mk_status(ST_DINC, 0, 0));
    // every earlier batch has claimed: the child is ours iff its claim word still holds our code (the LDS
    // slot word of a lost child is not read again: compaction takes won[], the writes only owned slots)
#pragma unroll
    for (int k = 0; k < EPT; ++k)
      if (won[k]) won[k] = ld_sync(&a.claim[cc[k]]) == code[k];
  }

  if constexpr (PROF) { if (tid == 0) { const uint64_t _t = wall_ticks(); pf[PF_TRACE0 + 2] = _t; pf[PF_FINAL] += _t - pf[15]; pf[15] = _t; } }
  // C. compact owned children in slot order. The slots form chunks of 64 (chunk c: slots [64c, 64c + 64)); wave w
  //    holds chunks w, w + W, w + 2W, ... (lane l: slot 64c + l, the claim's striping) with each slot's ownership,
  //    degree, parent and row entry in registers, and skips the chunks at or beyond D. Ballots rank owned slots
  //    within a chunk; one exchange of per-chunk counts and degree sums, scanned by every wave, gives each chunk
  //    its base.
  //    Keys: plain BFS (kept in slot order): (exclusive degree prefix << 32) | degree; CM: (parent slot << 32) |
  //    degree. For the segment sort, hist[i] and hist[P + i] receive parent i's segment start and degree base.
  constexpr int NC = S / kWave;
  static_assert(NC <= kWave, "one wave scans the chunk counts");
  const bool seg_sort =
      wave_uniform(!BFS && (a.sort_mode == 2 || (a.sort_mode == 0 && maxpdeg <= (uint32_t)a.seg_threshold)) ? 1u : 0u);
  const bool want_dx = BFS || seg_sort;
  const int wv = (int)wave_uniform((uint32_t)wave_id()), ln = lane_id();
  const uint64_t lt = lanemask_lt();
  const uint32_t Du = wave_uniform(D);
  const uint32_t nch = (Du + kWave - 1) / kWave;
  uint64_t obal[EPT];
  uint32_t sdeg[EPT], dexc[EPT];
  uint32_t lmax = 0;
#pragma unroll
  for (int k = 0; k < EPT; ++k) {
    const uint32_t c = (uint32_t)(k * W + wv);
    obal[k] = 0;
    sdeg[k] = 0;
    dexc[k] = 0;
    if (c < nch) {
      const bool own = won[k];
      sdeg[k] = own ? cre[k] - crb[k] : 0u;
      lmax = max(lmax, sdeg[k]);
      obal[k] = ballot(own);
      uint32_t csum = 0;
      if (want_dx) {
        const uint32_t inc = wave_incl_scan(sdeg[k]);
        dexc[k] = inc - sdeg[k];
        csum = lane63(inc);
      }
      if (ln == 0) {
        sm.cinfo[c] = (uint32_t)__popcll(obal[k]);
        sm.cinfo[NC + c] = csum;
      }
    }
  }
  {
    const uint32_t wm = wave_max(lmax);
    if (ln == 0) sm.wtmp[wv] = wm;
  }
  __syncthreads();
  // (no second barrier: cinfo and wtmp are next written after the barrier that ends this phase)
  uint32_t cnt, maxdeg = 0, cex, dex;
  {
    const uint32_t cc = (uint32_t)ln < nch ? sm.cinfo[ln] : 0u, cd = (uint32_t)ln < nch ? sm.cinfo[NC + ln] : 0u;
    const uint32_t ci = wave_incl_scan(cc), di = wave_incl_scan(cd);
    cnt = lane63(ci);
    cex = ci - cc;  // lane l: owned slots before chunk l
    dex = di - cd;  // lane l: their degree sum
#pragma unroll
    for (int w2 = 0; w2 < W; ++w2) maxdeg = max(maxdeg, sm.wtmp[w2]);
  }
#pragma unroll
  for (int k = 0; k < EPT; ++k) {
    const uint32_t c = (uint32_t)(k * W + wv);
    if (c < nch) {
      const uint32_t pbase = (uint32_t)__builtin_amdgcn_readlane((int)cex, (int)c);
      const uint32_t dbase = (uint32_t)__builtin_amdgcn_readlane((int)dex, (int)c);
      const uint32_t j = c * kWave + (uint32_t)ln;
      const uint32_t p = pbase + (uint32_t)__popcll(obal[k] & lt);  // owned slots before slot j
      if (j < Du) {
        const uint32_t i = ci16[k];
        if (seg_sort && (uint32_t)code[k] == 0u) {  // row entry 0: the first slot of parent i
          sm.hist[i] = p;
          sm.hist[G::P + i] = dbase + dexc[k];
        }
        if ((obal[k] >> ln) & 1ull) {
          sm.key[0][p] = ((uint64_t)(BFS ? dbase + dexc[k] : i) << 32) | (uint64_t)sdeg[k];
          sm.val[0][p] = (uint16_t)j;
        }
      }
    }
  }
  __syncthreads();

  if constexpr (PROF) { if (tid == 0) { const uint64_t _t = wall_ticks(); pf[PF_COMPACT] += _t - pf[15]; pf[15] = _t; } }
  // D. order: plain BFS keeps slot order; CM sorts by (parent slot, degree), stably. dx_key: the keys of the
  //    final order carry the exclusive degree prefix in their high word.
  int fin = 0;
  bool dx_key = BFS;
  if constexpr (!BFS) {
    if (cnt > 1) {
      if (seg_sort) {
        fin = lds_segment_sort<G, DIGIT>(sm, cnt);
        dx_key = true;
      } else {
        fin = lds_radix_sort<G, DIGIT>(sm, cnt, bits_for(maxdeg), bits_for((uint32_t)(m - 1)));
      }
    }
  }

  if constexpr (PROF) { if (tid == 0) { const uint64_t _t = wall_ticks(); pf[PF_SORT] += _t - pf[15]; pf[15] = _t; } }
  // E. cut into child batches: a child starts a new batch at a count quantum (P), at a degree quantum (Q),
  //    or next to a child of degree > Q; every batch then has <= P parents and <= S slots or one node.
  //    Positions are taken in 64-entry chunks like the slots in C; a child's degree is its key's low word, its
  //    exclusive degree prefix the key's high word (dx_key) or a scan here.
  uint64_t* cut = sm.key[fin ^ 1];
  const uint32_t nco = (cnt + kWave - 1) / kWave;
  uint32_t ldeg[EPT], ldx[EPT], pdeg[EPT];
#pragma unroll
  for (int k = 0; k < EPT; ++k) {
    const uint32_t c = (uint32_t)(k * W + wv);
    ldeg[k] = 0;
    ldx[k] = 0;
    pdeg[k] = 0;
    if (c < nco) {
      const uint32_t o = c * kWave + (uint32_t)ln;
      const uint64_t w = o < cnt ? sm.key[fin][o] : 0ull;
      const uint64_t pw = (o > 0 && o < cnt) ? sm.key[fin][o - 1] : 0ull;
      ldeg[k] = (uint32_t)w;
      ldx[k] = (uint32_t)(w >> 32);
      pdeg[k] = (uint32_t)pw;
    }
  }
  if (!dx_key) {
#pragma unroll
    for (int k = 0; k < EPT; ++k) {
      const uint32_t c = (uint32_t)(k * W + wv);
      if (c < nco) {
        const uint32_t inc = wave_incl_scan(ldeg[k]);
        ldx[k] = inc - ldeg[k];
        const uint32_t tot = lane63(inc);
        if (ln == 0) sm.cinfo[NC + c] = tot;
      }
    }
    __syncthreads();
    const uint32_t cd = (uint32_t)ln < nco ? sm.cinfo[NC + ln] : 0u;
    const uint32_t dexo = wave_incl_scan(cd) - cd;
#pragma unroll
    for (int k = 0; k < EPT; ++k) {
      const uint32_t c = (uint32_t)(k * W + wv);
      if (c < nco) ldx[k] += (uint32_t)__builtin_amdgcn_readlane((int)dexo, (int)c);
    }
  }
  uint64_t sbal[EPT];
#pragma unroll
  for (int k = 0; k < EPT; ++k) {
    const uint32_t c = (uint32_t)(k * W + wv);
    sbal[k] = 0;
    if (c < nco) {
      const uint32_t o = c * kWave + (uint32_t)ln;
      const uint32_t prev_dx = ldx[k] - pdeg[k];  // the previous child's prefix
      const bool st = o < cnt && (o == 0 || (ldx[k] / G::Q) != (prev_dx / G::Q) || (o / G::P) != ((o - 1) / G::P) ||
                                  ldeg[k] > G::Q || pdeg[k] > G::Q);
      sbal[k] = ballot(st);
      if (ln == 0) sm.cinfo[c] = (uint32_t)__popcll(sbal[k]);
    }
  }
  __syncthreads();
  uint32_t nb;
  {
    const uint32_t cs = (uint32_t)ln < nco ? sm.cinfo[ln] : 0u;
    const uint32_t si = wave_incl_scan(cs);
    nb = lane63(si);
    const uint32_t sex = si - cs;
#pragma unroll
    for (int k = 0; k < EPT; ++k) {
      const uint32_t c = (uint32_t)(k * W + wv);
      if (c < nco) {
        const uint32_t cb = (uint32_t)__builtin_amdgcn_readlane((int)sex, (int)c);  // starts before chunk c
        if ((sbal[k] >> ln) & 1ull) cut[cb + (uint32_t)__popcll(sbal[k] & lt)] = (uint64_t)(c * kWave + (uint32_t)ln);
      }
    }
  }

  if constexpr (PROF) { if (tid == 0) { const uint64_t _t = wall_ticks(); pf[PF_TRACE0 + 3] = _t; pf[PF_CUT] += _t - pf[15]; pf[15] = _t; } }
  // F. commit: publish the aggregate, look back for the exclusive prefix, publish the inclusive record. The
  //    ticket counter is read beside the look-back (one round trip for both): it tells whether the first child
  //    batch's ticket has been handed out already.
  if (tid == 0) st_sync(&a.bstat[b], mk_status(ST_AGG, cnt, nb));
  if (wave_id() == 0) {
    const uint32_t tick = wave_uniform(ld_sync(&a.ctl->ticket));
    uint32_t xo = 0, xb = 0;
    const bool ok = lookback_prefix(a, (int64_t)b, deadline, &xo, &xb);
    if (lane_id() == 0) {
      sm.quit = ok ? 0 : 1;
      sm.excl_off = xo;
      sm.excl_nb = xb;
      sm.tick = tick;
      if (ok) st_sync(&a.bstat[b], mk_status(ST_INC, xo + cnt, xb + nb));
    }
  }
  __syncthreads();
  if (sm.quit) return;
  const uint32_t xo = sm.excl_off, xb = sm.excl_nb;
  if (xo + cnt > (uint32_t)a.n || xb + nb > (uint32_t)a.bcap) {
    if (tid == 0) raise_error(a, ERR_CAPACITY, 3, b, (uint32_t)s, (uint32_t)e, D, xo, cnt, xb, nb);
    stop_workgroup(sm);
    return;
  }

  if constexpr (PROF) { if (tid == 0) { const uint64_t _t = wall_ticks(); pf[PF_TRACE0 + 4] = _t; pf[PF_PREFIX] += _t - pf[15]; pf[15] = _t; } }
  // G. write children (queue entry, row bounds, level), then publish the child batch records.
  //    Work-first: when the first child batch's ticket has already been handed out (so every batch below it
  //    is held by a running workgroup, and waiting on them cannot deadlock), this workgroup keeps that batch
  //    and processes it next with its parents staged in LDS (key[fin], free after the cut): no wait for the
  //    record and no reload of the parents. Its record is published with kKeptBit set, so the ticket's holder
  //    takes another ticket instead.
  const bool keep = nb > 0 && sm.tick > xb;
  const uint32_t kend = nb > 1 ? (uint32_t)cut[1] : cnt;  // the first child batch: children [0, kend)
  for (uint32_t o = tid; o < cnt; o += T) {
    const uint32_t j = sm.val[fin][o];
    const uint32_t p = xo + o;
    const int32_t child = sm.sc[j];
    const uint32_t crb = sm.srb[j], cre = sm.sre[j];
    const int32_t clev = (BFS || LEV) ? sm.plevel[sm.si[j]] + 1 : 0;
    st_sync(&a.perm[p], child);
    st_sync(&a.qinfo[p], (uint64_t)crb | ((uint64_t)cre << 32));
    if constexpr (BFS || LEV) st_sync(&a.level[p], clev);
    if (keep && o < kend) {
      sm.key[fin][o] = (uint64_t)(uint32_t)child | ((uint64_t)crb << 32);
      sm.key[fin][G::P + o] = (uint64_t)cre | ((uint64_t)(uint32_t)clev << 32);
    }
  }
  publish_wait<STRICT>();
  __syncthreads();
  for (uint32_t t = tid; t < nb; t += T) {
    const uint32_t beg = (uint32_t)cut[t];
    const uint32_t end = t + 1 < nb ? (uint32_t)cut[t + 1] : cnt;
    const uint64_t kb = (t == 0 && keep) ? kKeptBit : 0ull;
    st_sync(&a.brange[xb + t], kb | (uint64_t)(xo + beg) | ((uint64_t)(xo + end) << 32));
  }
  if (tid == 0) {
    const bool final_batch = xb + nb == b + 1;  // no batch beyond b exists or can be created
    if (cnt > 0 || final_batch) {
      const uint64_t add = (uint64_t)cnt + (final_batch ? ((uint64_t)(xo + cnt) << 32) : 0ull);
      if (final_batch) {
        st_sync(&a.ctl->nbatches, b + 1);
        st_sync(&a.ctl->placed, xo + cnt);
      }
      atomic_add_u64_nowait(&a.ctl->progress, add);  // completion is seen by the workers waiting for work
    }
    if (keep) {  // the caller processes this batch next, without acquiring it (read after its barrier)
      sm.kept = 1;
      sm.kfin = fin;
      sm.kearly = xb == b + 1 ? 1 : 0;  // no batch between this one and the kept one: all before it have claimed
      sm.b = (int32_t)xb;
      sm.s = (int32_t)xo;
      sm.e = (int32_t)(xo + kend);
    }
    if constexpr (PROF) {
      const uint64_t _t = wall_ticks();
      if (a.trace && (int64_t)b < (int64_t)a.trace_cap) {
        uint64_t* w = a.trace + (size_t)b * kTraceWords;
        w[0] = (uint64_t)b | ((uint64_t)(kept ? 1 : 0) << 32) | ((uint64_t)(early ? 1 : 0) << 33);
        w[1] = (uint64_t)(uint32_t)s | ((uint64_t)(uint32_t)e << 32);
        w[2] = (uint64_t)cnt | ((uint64_t)nb << 32);
        w[3] = xb;
        for (int q = 0; q < 5; ++q) w[4 + q] = pf[PF_TRACE0 + q];
        w[9] = _t;
      }
      pf[PF_WRITE] += _t - pf[15];
      pf[15] = _t;
      pf[PF_BATCHES] += 1;
      pf[PF_KEPT] += keep ? 1 : 0;
      pf[PF_EARLY] += early ? 1 : 0;
      pf[PF_CHILDREN] += cnt;
      pf[PF_SLOTS] += D;
    }
  }
}

// ---- the persistent kernel --------------------------------------------------------------------------------
template <class G, int DIGIT, bool STRICT, bool BFS, bool PROF = false, bool LEV = false>
__global__ __launch_bounds__(G::T) void k_reorder(RunArgs a) {
  using SmemT = Smem<G, DIGIT>;
  RCM_LDS(SmemT, sm);
  const int tid = threadIdx.x;
  const uint64_t deadline = wall_ticks() + a.budget_ticks;
  uint64_t pf[kProfRegs];  // thread 0's phase sums and trace timestamps (PROF builds)
  if constexpr (PROF) {
    for (int i = 0; i < kProfRegs; ++i) pf[i] = 0;
  }
  if (tid == 0) sm.kept = 0;
  __syncthreads();
  for (;;) {
    if (!sm.kept) {  // acquire: the next ticket whose batch is not kept by its producer
      if (tid == 0) {
        if constexpr (PROF) pf[15] = wall_ticks();
        int quit = 0;
        uint32_t b = 0;
        uint64_t r = 0;
        for (;;) {
          b = atomic_add_u32(&a.ctl->ticket, 1u);
          for (uint32_t spin = 0;; ++spin) {
            if (b < (uint32_t)a.bcap) {
              r = ld_sync(&a.brange[b]);
              if (r) break;
            }
            if ((spin + 1) % kQuitPoll == 0) {  // idle workers poll the done word rarely (measured faster)
              if (ld_sync(&a.ctl->done)) { quit = 1; break; }
              if (run_complete(a.ctl)) { mark_done(a); quit = 1; break; }
              if (wall_ticks() > deadline) { raise_error(a, ERR_TIMEOUT); quit = 1; break; }
            }
            __builtin_amdgcn_s_sleep(1);
          }
          if (quit || !(r & kKeptBit)) break;
        }
        sm.quit = quit;
        sm.b = (int32_t)b;
        sm.s = (int32_t)(uint32_t)r;
        sm.e = (int32_t)(r >> 32);
      }
      if constexpr (PROF) {
        if (tid == 0) {
          const uint64_t _t = wall_ticks();
          if (!sm.quit) pf[PF_ACQUIRE] += _t - pf[15];
          pf[15] = _t;
        }
      }
      __syncthreads();
      if (sm.quit) break;
    } else if constexpr (PROF) {
      if (tid == 0) pf[15] = wall_ticks();
    }
    consume_fence<STRICT>();
    process_batch<G, DIGIT, STRICT, BFS, PROF, LEV>(a, sm, deadline, pf);
    __syncthreads();
    const int quit = sm.quit;
    __syncthreads();  // every wave has read the scalars before thread 0 rewrites them for the next batch
    if (quit) break;
  }
  if constexpr (PROF) {
    if (tid == 0 && a.prof)
      for (int i = 0; i <= PF_KEPT; ++i) atomic_add_u64(&a.prof[i], pf[i]);
  }
}

// ---- helper kernels -----------------------------------------------------------------------------------------
// Reset for a run from `start`: claims, the first `bclear` batch records, the control block, queue slot 0.
__global__ void k_init(RunArgs a, int32_t start, int32_t bclear) {
  const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = tid; i < a.n; i += stride) a.claim[i] = i == start ? 0ull : ~0ull;
  for (int64_t i = tid; i < bclear; i += stride) {
    a.bstat[i] = 0ull;
    a.brange[i] = i == 0 ? (1ull << 32) : 0ull;  // batch 0 = queue range [0, 1)
  }
  if (tid == 0) {
    a.perm[0] = start;
    a.qinfo[0] = (uint64_t)(uint32_t)a.row_ptr[start] | ((uint64_t)(uint32_t)a.row_ptr[start + 1] << 32);
    if (a.level) a.level[0] = 0;
    Control c = {};
    c.progress = 1;  // the start node is written
    *a.ctl = c;
  }
}

// CSR validation: rp[0] == 0, rp[n] == nnz, rows non-decreasing, columns in [0, n). out[0] = error flags,
// out[1] = max row length.
__global__ void k_validate(const int32_t* rp, const int32_t* ci, int32_t n, int64_t nnz, uint32_t* out) {
  const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  uint32_t bad = 0, mx = 0;
  for (int64_t i = tid; i < n; i += stride) {
    const int32_t b = rp[i], e = rp[i + 1];
    if (e < b) bad |= 1u;
    else mx = max(mx, (uint32_t)(e - b));
  }
  if (tid == 0 && (rp[0] != 0 || (int64_t)rp[n] != nnz)) bad |= 2u;
  for (int64_t k = tid; k < nnz; k += stride) {
    const int32_t c = ci[k];
    if (c < 0 || c >= n) bad |= 4u;
  }
  mx = wave_max(mx);
  if (bad) atomicOr(&out[0], bad);
  if (lane_id() == 0 && mx) atomicMax(&out[1], mx);
}

// inv[perm[i]] = i
__global__ void k_inverse(const int32_t* perm, int32_t n, int32_t* inv) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) inv[perm[i]] = (int32_t)i;
}

// computeBandwidth (src/main.cpp:143-184): max over rows with entries of max(inv[i]-min, max-inv[i]),
// i.e. the largest |inv[i] - inv[j]| over stored entries.
__global__ void k_bandwidth(const int32_t* rp, const int32_t* ci, const int32_t* inv, int32_t n, uint32_t* out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  uint32_t bw = 0;
  if (i < n) {
    const int32_t ri = inv[i];
    for (int32_t k = rp[i]; k < rp[i + 1]; ++k) {
      const int32_t v = inv[ci[k]];
      bw = max(bw, (uint32_t)(v > ri ? v - ri : ri - v));
    }
  }
  bw = wave_max(bw);
  if (lane_id() == 0 && bw) atomicMax(out, bw);
}

// out[i] = perm[n - 1 - i]
__global__ void k_reverse(const int32_t* perm, int32_t n, int32_t* out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = perm[n - 1 - i];
}

}  // namespace rcm

// engine_wave.hpp -- one-wavefront-per-worker variant of the speculative Cuthill-McKee engine (device code).
//
// NOT BUILT INTO THE LIBRARY. On gfx950 hardware a run of this variant on a 17-node path graph did not
// finish and left its kernel running; the cause has not been found (the same source passes every parity case
// under the CPU wave emulator in tests/emu, including sanitizer builds with exact-size LDS). The library uses
// the 256-thread workgroup engine in engine.hpp; this file is compiled only by the emulator tests.
//
// Algorithm (the speculative batch scheme of the reference's GPU_BATCH path,
// include/device/batchWorker.cuh:199-280 and include/device/batchQueue.cuh:37-192; re-designed here):
//
//   The permutation array is the BFS queue. A *batch* is a range [s, e) of queue positions whose nodes are
//   expanded together by one worker. Batch ids are allocated in queue order, so "earlier batch" means
//   "parents earlier in the queue". Persistent workers take batch ids from a ticket counter in order and wait
//   until the batch exists (its range record is published by the batch that produced its nodes).
//
//   A worker is ONE wavefront (64 lanes). Every collective step of a batch -- prefix sums, compaction, ranks,
//   cuts, look-back -- is a wave-level operation (DPP scans, ballots, readlane), so a batch runs without a
//   single s_barrier; its LDS holds the batch's parents, adjacency slots and sort buffers.
//
//   Claims. The parent at queue position p claims its k-th row entry c with code (p << 32 | k) by a 64-bit
//   atomicMin on claim[c]. The smallest code wins, which is exactly the serial rule: a node belongs to the
//   earliest-dequeued parent whose row holds it, at that row's first occurrence (src/rcm_cpu.cpp:83-92).
//   Claims are speculative: a batch may claim before earlier batches have, and lose a node to them later.
//
//   Commit. Each batch publishes one 64-bit status word that moves through
//       DISC_LOCAL (own claims done) -> DISC_INC (all claims up to here done)
//       -> AGG (own child count, own child-batch count) -> INC (inclusive prefix of both).
//   Successors read up to 64 predecessor words with one wave load (decoupled look-back). A batch whose
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
//   written to the queue with their row bounds (the consumer needs no row_ptr lookup), then cut into child
//   batches of balanced size (<= P parents, and either <= S adjacency slots or a single node of degree
//   > S/2), whose range records are published. A 64-bit progress word (nodes written | total) detects
//   completion; the total becomes known at the batch after which no further batch can exist.
#pragma once
#include "wave_dpp.hpp"

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

struct Control {
  uint32_t ticket;    // next batch id handed to a worker
  uint32_t done;      // set once when the run is complete or aborted
  uint32_t error;     // ERR_*
  uint32_t nbatches;  // batch records used by the finished run
  uint64_t progress;  // low 32: queue entries written; high 32: total to be written (0 = not yet known)
  uint32_t placed;    // nodes placed by the finished run
  uint32_t diag[9];   // first error: site, batch, begin, end, slots, offset, count, first child id, children
};

struct RunArgs {
  const int32_t* row_ptr;
  const int32_t* col_idx;
  int32_t* perm;          // queue; CM order on completion
  uint64_t* qinfo;        // per queue position: row_begin | row_end << 32
  uint64_t* claim;        // per node claim code
  uint64_t* bstat;        // per batch status word
  uint64_t* brange;       // per batch: begin | end << 32 (0 = not yet created)
  int32_t* level;         // per queue position BFS level (BFS mode only)
  Control* ctl;
  int32_t n;
  int32_t bcap;           // records in bstat / brange
  uint64_t budget_ticks;  // watchdog budget in 100 MHz ticks
  int32_t sort_mode;      // 0 auto, 1 radix, 2 segment rank
  int32_t seg_threshold;  // auto mode: segment rank when every parent row has at most this many entries
  uint64_t* prof;         // PROF builds: per-phase tick sums (kProfSlots entries)
};

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
  kProfSlots = 16
};

// Batch geometry for one-wave workers: at most P = 64 parents and S adjacency slots per batch.
template <int S_>
struct Geo {
  static constexpr int T = kWave, P = kWave, S = S_;
  static constexpr int EPT = S / kWave;   // slot rounds
  static constexpr uint32_t Q = S / 2;    // degree quantum of the child-batch cut rule
  static_assert(S % kWave == 0 && S <= 16384, "geometry");
};

template <class G, int DIGIT>
struct Smem {
  static constexpr int R = 1 << DIGIT;
  static constexpr int H = R * G::EPT;  // radix histogram: (digit, round)
  uint32_t so[G::P + 1];  // slot offsets of the parents
  uint32_t prb[G::P];     // parent row begin
  int32_t plevel[G::P];   // parent level (BFS mode)
  int32_t sc[G::S];       // child id per adjacency slot, -1 when not owned
  uint32_t srb[G::S];     // child row begin
  uint32_t sre[G::S];     // child row end
  uint8_t si[G::S];       // parent slot of each adjacency slot
  uint64_t key[2][G::S];  // sort keys: parent slot << 32 | degree
  uint16_t val[2][G::S];  // sort payload: adjacency slot
  uint32_t hist[H > kWave ? H : kWave];
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
  st_sync(&a.ctl->done, 1u);
}

// Returns true when the run must stop (done, or the watchdog fired). Uniform across the wave.
__device__ __forceinline__ bool should_quit(const RunArgs& a, uint64_t deadline) {
  if (ld_sync(&a.ctl->done)) return true;
  if (wall_ticks() > deadline) {
    if (lane_id() == 0) raise_error(a, ERR_TIMEOUT);
    return true;
  }
  return false;
}

// ---- look-back over predecessor status words ---------------------------------------------------------------
// Claim completeness of every batch < b, without waiting. Lane l inspects batch (base - l); a position
// below 0 is the virtual predecessor of batch 0, which counts as complete.
__device__ __forceinline__ bool claims_complete_before(const RunArgs& a, int64_t b) {
  int64_t base = b - 1;
  for (;;) {
    const int64_t idx = base - lane_id();
    const uint32_t st = idx >= 0 ? st_state(ld_sync(&a.bstat[idx])) : (uint32_t)ST_DINC;
    const uint64_t incm = ballot(st >= ST_DINC);
    const uint64_t notm = ballot(st < ST_DLOC);
    if (incm) {
      const uint64_t first = incm & (~incm + 1);
      return (notm & (first - 1)) == 0;
    }
    if (notm) return false;
    base -= kWave;
  }
}

__device__ __forceinline__ bool wait_claims_complete(const RunArgs& a, int64_t b, uint64_t deadline) {
  for (;;) {
    if (claims_complete_before(a, b)) return true;
    if (should_quit(a, deadline)) return false;
    __builtin_amdgcn_s_sleep(1);
  }
}

// Exclusive prefix (queue offset, child-batch id) of batch b. The virtual predecessor of batch 0 is an
// inclusive record (1, 1): the start node occupies queue slot 0 and batch id 0.
__device__ __forceinline__ bool lookback_prefix(const RunArgs& a, int64_t b, uint64_t deadline, uint32_t* off,
                                                uint32_t* nbid) {
  int64_t base = b - 1;
  uint32_t acc_off = 0, acc_nb = 0;
  for (;;) {
    const int64_t idx = base - lane_id();
    const uint64_t w = idx >= 0 ? ld_sync(&a.bstat[idx]) : mk_status(ST_INC, 1u, 1u);
    const uint32_t st = st_state(w);
    const uint64_t incm = ballot(st == ST_INC);
    const uint64_t notm = ballot(st < ST_AGG);
    if (incm) {
      const int first = __ffsll((unsigned long long)incm) - 1;
      if ((notm & ((1ull << first) - 1ull)) == 0) {
        const bool take = lane_id() <= first;
        acc_off += dpp_sum(take ? st_a(w) : 0u);
        acc_nb += dpp_sum(take ? st_b(w) : 0u);
        *off = acc_off;
        *nbid = acc_nb;
        return true;
      }
    } else if (notm == 0) {
      acc_off += dpp_sum(st_a(w));
      acc_nb += dpp_sum(st_b(w));
      base -= kWave;
      continue;
    }
    if (should_quit(a, deadline)) return false;
    __builtin_amdgcn_s_sleep(1);
  }
}

// Largest i in [0, m) with so[i] <= j (so[0] == 0 <= j).
__device__ __forceinline__ int parent_of_slot(const uint32_t* so, int m, uint32_t j) {
  int lo = 0, hi = m;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (so[mid] <= j) lo = mid;
    else hi = mid;
  }
  return lo;
}

// ---- stable radix sort of (key, val) pairs in LDS by one wave, cnt <= S --------------------------------
// Keys: low word = child degree (dbits significant), high word = parent slot (ibits significant). Passes run
// over the low word's digits, then the high word's. Element e sits in round e / 64, lane e % 64; the rank of
// an element among equal digits of its round comes from DIGIT ballots, and the histogram is scanned in
// (digit, round) order, so ties keep their input order.
template <class G, int DIGIT>
__device__ int wave_radix_sort(Smem<G, DIGIT>& sm, uint32_t cnt, uint32_t dbits, uint32_t ibits) {
  constexpr int R = 1 << DIGIT, EPT = G::EPT, H = Smem<G, DIGIT>::H;
  constexpr int PER = (H + kWave - 1) / kWave;
  const int lane = lane_id();
  const int rounds = (int)((cnt + kWave - 1) / kWave);
  const int npl = (int)((dbits + DIGIT - 1) / DIGIT), nph = (int)((ibits + DIGIT - 1) / DIGIT);
  int cur = 0;
  for (int p = 0; p < npl + nph; ++p) {
    const int shift = p < npl ? p * DIGIT : 32 + (p - npl) * DIGIT;
    for (int h = lane; h < H; h += kWave) sm.hist[h] = 0;
    wave_lds_sync();
    uint32_t dg[EPT], rk[EPT];
#pragma unroll
    for (int r = 0; r < EPT; ++r) {
      dg[r] = 0;
      rk[r] = 0;
      if (r < rounds) {
        const uint32_t e = (uint32_t)(r * kWave + lane);
        const bool act = e < cnt;
        const uint32_t d = act ? (uint32_t)(sm.key[cur][e] >> shift) & (uint32_t)(R - 1) : 0u;
        uint64_t peers = ballot(act);
#pragma unroll
        for (int bit = 0; bit < DIGIT; ++bit) {
          const bool on = (d >> bit) & 1u;
          const uint64_t bm = ballot(on);
          peers &= on ? bm : ~bm;
        }
        const uint32_t rank = (uint32_t)__popcll(peers & lanemask_lt());
        if (act && rank == 0) sm.hist[d * EPT + r] = (uint32_t)__popcll(peers);
        dg[r] = d;
        rk[r] = rank;
      }
    }
    wave_lds_sync();
    // exclusive scan of hist in (digit, round) order; lane l owns entries [l*PER, l*PER+PER)
    uint32_t loc[PER];
    uint32_t lsum = 0;
#pragma unroll
    for (int q = 0; q < PER; ++q) {
      const int h = lane * PER + q;
      loc[q] = h < H ? sm.hist[h] : 0u;
      lsum += loc[q];
    }
    uint32_t run = dpp_excl_scan(lsum);
    wave_lds_sync();
#pragma unroll
    for (int q = 0; q < PER; ++q) {
      const int h = lane * PER + q;
      if (h < H) sm.hist[h] = run;
      run += loc[q];
    }
    wave_lds_sync();
#pragma unroll
    for (int r = 0; r < EPT; ++r) {
      if (r < rounds) {
        const uint32_t e = (uint32_t)(r * kWave + lane);
        if (e < cnt) {
          const uint32_t pos = sm.hist[dg[r] * EPT + r] + rk[r];
          sm.key[cur ^ 1][pos] = sm.key[cur][e];
          sm.val[cur ^ 1][pos] = sm.val[cur][e];
        }
      }
    }
    wave_lds_sync();
    cur ^= 1;
  }
  return cur;
}

// ---- stable segment-rank sort --------------------------------------------------------------------------------
// Compacted elements of one parent are contiguous (compaction keeps slot order). Each element counts the
// elements of its own segment that order before it by (degree, position) and moves to segment start + rank.
template <class G, int DIGIT>
__device__ int wave_segment_sort(Smem<G, DIGIT>& sm, uint32_t cnt) {
  for (uint32_t e = lane_id(); e < cnt; e += kWave) {
    const uint64_t k = sm.key[0][e];
    const uint32_t seg = (uint32_t)(k >> 32);
    uint32_t lo = e;
    while (lo > 0 && (uint32_t)(sm.key[0][lo - 1] >> 32) == seg) --lo;
    uint32_t rank = 0;
    for (uint32_t q = lo; q < cnt; ++q) {
      const uint64_t kq = sm.key[0][q];
      if ((uint32_t)(kq >> 32) != seg) break;
      rank += (kq < k || (kq == k && q < e)) ? 1u : 0u;
    }
    sm.key[1][lo + rank] = k;
    sm.val[1][lo + rank] = sm.val[0][e];
  }
  wave_lds_sync();
  return 1;
}

template <bool PROF>
__device__ __forceinline__ void prof_mark(uint64_t* pf, int slot) {
  if constexpr (PROF) {
    const uint64_t t = wall_ticks();
    pf[slot] += t - pf[15];
    pf[15] = t;
  }
}

// ---- one batch --------------------------------------------------------------------------------------------
// Returns false when the run must stop.
template <class G, int DIGIT, bool STRICT, bool BFS, bool PROF>
__device__ bool process_batch(const RunArgs& a, Smem<G, DIGIT>& sm, uint32_t b, int32_t s, int32_t e,
                              uint64_t deadline, uint64_t* pf) {
  constexpr int EPT = G::EPT;
  constexpr uint32_t S = G::S;
  const int lane = lane_id();
  const int m = e - s;
  if (m <= 0 || m > G::P) {
    if (lane == 0) raise_error(a, ERR_CAPACITY, 1, b, (uint32_t)s, (uint32_t)e);
    return false;
  }

  // A. parents (written by the producing batch with sc1 stores) and the non-blocking claim check
  uint32_t d = 0, rb = 0;
  int32_t plev = 0;
  if (lane < m) {
    const uint64_t q = ld_sync(&a.qinfo[s + lane]);
    rb = (uint32_t)q;
    d = (uint32_t)(q >> 32) - rb;
    if constexpr (BFS) plev = ld_sync(&a.level[s + lane]);
  }
  const bool early = claims_complete_before(a, (int64_t)b);
  const uint32_t soff = dpp_excl_scan(d);
  const uint32_t D = lane_value(soff + d, kWave - 1);
  const uint32_t maxpdeg = wave_max(d);
  if (lane < m) {
    sm.so[lane] = soff;
    sm.prb[lane] = rb;
    if constexpr (BFS) sm.plevel[lane] = plev;
  }
  if (lane == 0) sm.so[m] = D;
  wave_lds_sync();
  if (D > S) {
    if (lane == 0) raise_error(a, ERR_LONGROW, 2, b, (uint32_t)s, (uint32_t)e, D);
    return false;
  }
  const int rounds = (int)((D + kWave - 1) / kWave);
  prof_mark<PROF>(pf, PF_PARENTS);

  // B. claim every adjacency slot (round r holds slots 64r .. 64r+63), in groups of up to 16 rounds: all
  //    column loads of a group, one wait, then all its atomics and row-bound loads in flight together.
  //    Early batches drop in-batch duplicates right after each group (the superseded slot is in the same or
  //    an earlier group: its atomic completed before ours was issued).
  constexpr int GR = EPT < 16 ? EPT : 16;
  for (int g0 = 0; g0 < rounds; g0 += GR) {
    uint64_t code[GR], oldv[GR];
    int32_t cc[GR];
    uint32_t crb[GR], cre[GR];
    uint8_t pi[GR];
#pragma unroll
    for (int k = 0; k < GR; ++k) {
      code[k] = 0;
      cc[k] = 0;
      pi[k] = 0;
      const uint32_t j = (uint32_t)((g0 + k) * kWave + lane);
      if (g0 + k < rounds && j < D) {
        const int i = parent_of_slot(sm.so, m, j);
        const uint32_t r = j - sm.so[i];
        cc[k] = a.col_idx[sm.prb[i] + r];
        code[k] = ((uint64_t)(uint32_t)(s + i) << 32) | (uint64_t)r;
        pi[k] = (uint8_t)i;
      }
    }
    __builtin_amdgcn_s_waitcnt(0);
#pragma unroll
    for (int k = 0; k < GR; ++k) {
      oldv[k] = 0;
      crb[k] = 0;
      cre[k] = 0;
      if (g0 + k < rounds && (uint32_t)((g0 + k) * kWave + lane) < D) {
        oldv[k] = atomic_min_u64(&a.claim[cc[k]], code[k]);
        crb[k] = (uint32_t)a.row_ptr[cc[k]];
        cre[k] = (uint32_t)a.row_ptr[cc[k] + 1];
      }
    }
#pragma unroll
    for (int k = 0; k < GR; ++k) {
      const uint32_t j = (uint32_t)((g0 + k) * kWave + lane);
      if (g0 + k < rounds && j < D) {
        sm.sc[j] = oldv[k] > code[k] ? cc[k] : -1;
        sm.srb[j] = crb[k];
        sm.sre[j] = cre[k];
        sm.si[j] = pi[k];
      }
    }
    if (early) {
      // every earlier batch had finished claiming: a won atomicMin is final unless a slot of this same batch
      // with a smaller code took the node afterwards; that slot saw our code as its old value.
      wave_lds_sync();
#pragma unroll
      for (int k = 0; k < GR; ++k) {
        const uint32_t j = (uint32_t)((g0 + k) * kWave + lane);
        if (g0 + k < rounds && j < D && oldv[k] > code[k]) {
          const uint32_t ohi = (uint32_t)(oldv[k] >> 32);
          if (oldv[k] != ~0ull && ohi >= (uint32_t)s && ohi < (uint32_t)e)
            sm.sc[sm.so[ohi - (uint32_t)s] + (uint32_t)oldv[k]] = -1;
        }
      }
    }
    wave_lds_sync();
  }
  prof_mark<PROF>(pf, PF_CLAIM);

  if (early) {
    publish_wait<STRICT>();
    if (lane == 0) st_sync(&a.bstat[b], mk_status(ST_DINC, 0, 0));
  } else {
    publish_wait<STRICT>();
    if (lane == 0) st_sync(&a.bstat[b], mk_status(ST_DLOC, 0, 0));
    if (!wait_claims_complete(a, (int64_t)b, deadline)) return false;
    consume_fence<STRICT>();
    if (lane == 0) st_sync(&a.bstat[b], mk_status(ST_DINC, 0, 0));
    // re-read the claim word of every node this batch won; keep it only if our code is still the minimum
    for (int g0 = 0; g0 < rounds; g0 += GR) {
      uint64_t cur[GR];
      int32_t own[GR];
#pragma unroll
      for (int k = 0; k < GR; ++k) {
        const uint32_t j = (uint32_t)((g0 + k) * kWave + lane);
        own[k] = (g0 + k < rounds && j < D) ? sm.sc[j] : -1;
        cur[k] = own[k] >= 0 ? ld_sync(&a.claim[own[k]]) : 0ull;
      }
#pragma unroll
      for (int k = 0; k < GR; ++k) {
        const uint32_t j = (uint32_t)((g0 + k) * kWave + lane);
        if (own[k] >= 0) {
          const uint32_t i = sm.si[j];
          const uint64_t code = ((uint64_t)(uint32_t)(s + (int32_t)i) << 32) | (uint64_t)(j - sm.so[i]);
          if (cur[k] != code) sm.sc[j] = -1;
        }
      }
    }
    wave_lds_sync();
  }
  prof_mark<PROF>(pf, PF_FINAL);

  // C. compact owned children in slot order (ballot per round)
  uint32_t cnt = 0, lmax = 0, dsum = 0;
#pragma unroll
  for (int k = 0; k < EPT; ++k) {
    if (k < rounds) {
      const uint32_t j = (uint32_t)(k * kWave + lane);
      const bool own = j < D && sm.sc[j] >= 0;
      const uint64_t msk = ballot(own);
      if (own) {
        const uint32_t pos = cnt + (uint32_t)__popcll(msk & lanemask_lt());
        const uint32_t dg = sm.sre[j] - sm.srb[j];
        sm.key[0][pos] = ((uint64_t)sm.si[j] << 32) | (uint64_t)dg;
        sm.val[0][pos] = (uint16_t)j;
        lmax = max(lmax, dg);
        dsum += dg;
      }
      cnt += (uint32_t)__popcll(msk);
    }
  }
  const uint32_t maxdeg = wave_max(lmax);
  const uint32_t dtot = dpp_sum(dsum);
  wave_lds_sync();

  // D. order: plain BFS keeps slot order; CM sorts by (parent slot, degree), stably
  int fin = 0;
  if constexpr (!BFS) {
    if (cnt > 1) {
      const bool seg = a.sort_mode == 2 || (a.sort_mode == 0 && maxpdeg <= (uint32_t)a.seg_threshold);
      if (seg)
        fin = wave_segment_sort<G, DIGIT>(sm, cnt);
      else
        fin = wave_radix_sort<G, DIGIT>(sm, cnt, bits_for(maxdeg), bits_for((uint32_t)(m - 1)));
    }
  }
  prof_mark<PROF>(pf, PF_SORT);

  // E. cut into child batches of balanced size: a child starts a new batch where the running degree sum
  //    crosses a multiple of qd, where the position crosses a multiple of qc, or next to a child of degree
  //    > Q. qd <= Q and qc <= P split the totals evenly; every batch gets <= P parents and <= S slots, or
  //    holds a single node of degree > Q.
  uint32_t nb = 0;
  uint32_t* cut = reinterpret_cast<uint32_t*>(sm.key[fin ^ 1]);  // child-batch starts (free sort buffer)
  {
    const uint32_t nd = dtot == 0 ? 1u : (dtot + G::Q - 1) / G::Q;
    const uint32_t qd = dtot == 0 ? 1u : (dtot + nd - 1) / nd;
    const uint32_t ncb = cnt == 0 ? 1u : (cnt + G::P - 1) / G::P;
    const uint32_t qc = cnt == 0 ? 1u : (cnt + ncb - 1) / ncb;
    const int orounds = (int)((cnt + kWave - 1) / kWave);
    uint32_t carry_dx = 0, carry_deg = 0;
#pragma unroll
    for (int k = 0; k < EPT; ++k) {
      if (k < orounds) {
        const uint32_t o = (uint32_t)(k * kWave + lane);
        const bool act = o < cnt;
        const uint32_t dg = act ? (uint32_t)sm.key[fin][o] : 0u;
        const uint32_t incl = dpp_incl_scan(dg);
        const uint32_t dx = carry_dx + incl - dg;
        const uint32_t pdeg = wave_shift_up(dg, carry_deg);
        const uint32_t pdx = dx - pdeg;
        const bool st = act && (o == 0 || (dx / qd) != (pdx / qd) || (o / qc) != ((o - 1) / qc) || dg > G::Q ||
                                pdeg > G::Q);
        const uint64_t msk = ballot(st);
        if (st) cut[nb + (uint32_t)__popcll(msk & lanemask_lt())] = o;
        nb += (uint32_t)__popcll(msk);
        carry_dx += lane_value(incl, kWave - 1);
        carry_deg = lane_value(dg, kWave - 1);
      }
    }
  }
  wave_lds_sync();
  prof_mark<PROF>(pf, PF_CUT);

  // F. commit: publish the aggregate, look back for the exclusive prefix, publish the inclusive record
  if (lane == 0) st_sync(&a.bstat[b], mk_status(ST_AGG, cnt, nb));
  uint32_t xo = 0, xb = 0;
  if (!lookback_prefix(a, (int64_t)b, deadline, &xo, &xb)) return false;
  if (lane == 0) st_sync(&a.bstat[b], mk_status(ST_INC, xo + cnt, xb + nb));
  if (xo + cnt > (uint32_t)a.n || xb + nb > (uint32_t)a.bcap) {
    if (lane == 0) raise_error(a, ERR_CAPACITY, 3, b, (uint32_t)s, (uint32_t)e, D, xo, cnt, xb, nb);
    return false;
  }
  prof_mark<PROF>(pf, PF_PREFIX);

  // G. write children (queue entry, row bounds, level), then publish the child batch records
  for (uint32_t o = lane; o < cnt; o += kWave) {
    const uint32_t j = sm.val[fin][o];
    const uint32_t p = xo + o;
    st_sync(&a.perm[p], sm.sc[j]);
    st_sync(&a.qinfo[p], (uint64_t)sm.srb[j] | ((uint64_t)sm.sre[j] << 32));
    if constexpr (BFS) st_sync(&a.level[p], sm.plevel[sm.si[j]] + 1);
  }
  publish_wait<STRICT>();
  for (uint32_t t = lane; t < nb; t += kWave) {
    const uint32_t beg = cut[t];
    const uint32_t end = t + 1 < nb ? cut[t + 1] : cnt;
    st_sync(&a.brange[xb + t], (uint64_t)(xo + beg) | ((uint64_t)(xo + end) << 32));
  }
  if (lane == 0) {
    const bool final_batch = xb + nb == b + 1;  // no batch beyond b exists or can be created
    if (cnt > 0 || final_batch) {
      const uint64_t add = (uint64_t)cnt + (final_batch ? ((uint64_t)(xo + cnt) << 32) : 0ull);
      if (final_batch) {
        st_sync(&a.ctl->nbatches, b + 1);
        st_sync(&a.ctl->placed, xo + cnt);
      }
      const uint64_t now = atomic_add_u64(&a.ctl->progress, add) + add;
      if ((now >> 32) != 0 && (uint32_t)now == (uint32_t)(now >> 32)) st_sync(&a.ctl->done, 1u);
    }
  }
  if constexpr (PROF) {
    prof_mark<PROF>(pf, PF_WRITE);
    pf[PF_BATCHES] += 1;
    pf[PF_EARLY] += early ? 1 : 0;
    pf[PF_CHILDREN] += cnt;
    pf[PF_SLOTS] += D;
  }
  return true;
}

// ---- the persistent kernel: one wave per workgroup, one worker per wave ----------------------------------
template <class G, int DIGIT, bool STRICT, bool BFS, bool PROF = false>
__global__ __launch_bounds__(kWave) void k_reorder(RunArgs a) {
  using SmemT = Smem<G, DIGIT>;
  RCM_LDS(SmemT, sm);
  const int lane = lane_id();
  const uint64_t deadline = wall_ticks() + a.budget_ticks;
  uint64_t pf[kProfSlots];
  if constexpr (PROF) {
    for (int i = 0; i < kProfSlots; ++i) pf[i] = 0;
  }
  for (;;) {
    if constexpr (PROF) pf[15] = wall_ticks();
    uint32_t b = 0;
    if (lane == 0) b = atomic_add_u32(&a.ctl->ticket, 1u);
    b = lane_value(b, 0);
    uint64_t r = 0;
    bool quit = false;
    for (;;) {
      if (b < (uint32_t)a.bcap) {
        r = ld_sync(&a.brange[b]);
        if (r) break;
      }
      if (should_quit(a, deadline)) {
        quit = true;
        break;
      }
      __builtin_amdgcn_s_sleep(1);
    }
    if (quit) break;
    prof_mark<PROF>(pf, PF_ACQUIRE);
    consume_fence<STRICT>();
    if (!process_batch<G, DIGIT, STRICT, BFS, PROF>(a, sm, b, (int32_t)(uint32_t)r, (int32_t)(r >> 32), deadline,
                                                    pf))
      break;
  }
  if constexpr (PROF) {
    if (a.prof && lane < 12) atomic_add_u64(&a.prof[lane], pf[lane]);
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

// Self-test of the wave primitives on one wave: out[0..63] inclusive scan, [64..127] exclusive scan,
// [128] sum, [129] max, [130] min, [131..194] shift-up (lane 0 receives 7777).
__global__ void k_selftest_wave(const uint32_t* in, uint32_t* out) {
  const int l = lane_id();
  const uint32_t v = in[l];
  out[l] = dpp_incl_scan(v);
  out[64 + l] = dpp_excl_scan(v);
  const uint32_t s = dpp_sum(v), mx = wave_max(v), mn = wave_min(v);
  if (l == 0) {
    out[128] = s;
    out[129] = mx;
    out[130] = mn;
  }
  out[131 + l] = wave_shift_up(v, 7777u);
}

}  // namespace rcm

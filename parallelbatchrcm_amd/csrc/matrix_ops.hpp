// matrix_ops.hpp -- device kernels for the steps either side of the reorder: input canonicalisation (COO -> CSR,
// symmetrisation, connectivity repair) and applying a symmetric permutation B = P A P^T with values.
//
// The host orchestration is templated on an executor `X` (launch, scratch allocation, readback), so the same
// kernels and the same sequence run on the GPU (matrix_ops.hip) and in the CPU wave emulator of the tests.
//
// Building blocks:
//   * a device-wide exclusive scan of u32 (reduce / top / apply, 4096 items per workgroup);
//   * a stable LSD radix sort of (u64 key, u32 payload) pairs with 8-bit digits: per pass a digit histogram
//     per 4096-item tile, a scan of the [digit][tile] table, and a scatter that ranks the tile stably in LDS
//     (wave64 ballots on the digit, per-wave digit counts) and then writes each digit's run contiguously.
//
// Every kernel reads only inside the sizes it is given; malformed inputs are rejected by validation kernels
// before any kernel that indexes by their values.
#pragma once
#include "device_common.hpp"

namespace rcm_mx {
namespace {  // internal linkage: each library source that includes this has its own copy

using rcm::ballot;
using rcm::kWave;
using rcm::lane_id;
using rcm::lanemask_lt;
using rcm::wave_id;

constexpr int kT = 256;              // threads per workgroup
constexpr int kWaves = kT / kWave;   // 4
constexpr int kItems = 16;           // items per thread in scan and sort tiles
constexpr int kTile = kT * kItems;   // 4096
constexpr int kBins = 256;           // 8-bit radix digits
constexpr int kShortRow = 64;        // rows up to this length are ordered by rank counting in LDS
constexpr int kRowsPerGroup = 64;    // rows per workgroup in that kernel (one per lane of wave 0)

static_assert(kT == kBins, "one thread per digit");

__device__ __forceinline__ uint32_t ld_u32(const uint32_t* p) { return rcm::ld_sync(p); }
__device__ __forceinline__ void st_u32(uint32_t* p, uint32_t v) { rcm::st_sync(p, v); }

// ------------------------------------------------------------------------------------------------------
// Device-wide exclusive scan of u32 (in place allowed).
// ------------------------------------------------------------------------------------------------------
struct ScanSmem {
  uint32_t w[kWaves + 1];
  uint32_t carry;
};

// partial[b] = sum of in[b * kTile, (b + 1) * kTile)
__global__ __launch_bounds__(kT) void k_scan_reduce(const uint32_t* in, int64_t m, uint32_t* partial) {
  RCM_LDS(ScanSmem, sm);
  const int64_t base = (int64_t)blockIdx.x * kTile;
  uint32_t s = 0;
  for (int j = 0; j < kItems; ++j) {
    const int64_t i = base + (int64_t)j * kT + threadIdx.x;
    if (i < m) s += in[i];
  }
  uint32_t total;
  rcm::block_excl_scan<kT>(s, sm.w, &total);
  if (threadIdx.x == 0) partial[blockIdx.x] = total;
}

// One workgroup: exclusive scan of partial[0, p) in place; *total_out (nullable) receives the sum.
__global__ __launch_bounds__(kT) void k_scan_top(uint32_t* partial, int64_t p, uint32_t* total_out) {
  RCM_LDS(ScanSmem, sm);
  uint32_t carry = 0;
  for (int64_t c = 0; c < p; c += kT) {
    const int64_t i = c + threadIdx.x;
    const uint32_t v = i < p ? partial[i] : 0u;
    uint32_t tot;
    const uint32_t ex = rcm::block_excl_scan<kT>(v, sm.w, &tot);
    if (i < p) partial[i] = carry + ex;
    carry += tot;
  }
  if (threadIdx.x == 0 && total_out) *total_out = carry;
}

// out[i] = partial[tile] + exclusive prefix of in within the tile; rounds of kT consecutive items.
__global__ __launch_bounds__(kT) void k_scan_apply(const uint32_t* in, uint32_t* out, int64_t m,
                                                  const uint32_t* partial) {
  RCM_LDS(ScanSmem, sm);
  const int64_t base = (int64_t)blockIdx.x * kTile;
  uint32_t carry = partial[blockIdx.x];
  for (int j = 0; j < kItems; ++j) {
    const int64_t i = base + (int64_t)j * kT + threadIdx.x;
    const uint32_t v = i < m ? in[i] : 0u;
    uint32_t tot;
    const uint32_t ex = rcm::block_excl_scan<kT>(v, sm.w, &tot);
    if (i < m) out[i] = carry + ex;
    carry += tot;
  }
}

// ------------------------------------------------------------------------------------------------------
// Stable LSD radix sort of (u64 key, u32 payload), one 8-bit digit per pass.
// ------------------------------------------------------------------------------------------------------
struct HistSmem {
  uint32_t h[kBins];
};

// hist[d * nblocks + b] = number of keys of tile b whose digit at `shift` is d
__global__ __launch_bounds__(kT) void k_radix_hist(const uint64_t* keys, int64_t n, int shift, uint32_t* hist,
                                                  uint32_t nblocks) {
  RCM_LDS(HistSmem, sm);
  const int t = threadIdx.x;
  sm.h[t] = 0;
  __syncthreads();
  const int64_t base = (int64_t)blockIdx.x * kTile;
  for (int j = 0; j < kItems; ++j) {
    const int64_t i = base + (int64_t)j * kT + t;
    if (i < n) atomicAdd(&sm.h[(uint32_t)(keys[i] >> shift) & (kBins - 1)], 1u);
  }
  __syncthreads();
  hist[(size_t)t * nblocks + blockIdx.x] = sm.h[t];
}

struct ScatterSmem {
  uint64_t key[kTile];
  uint32_t val[kTile];
  uint32_t tile_off[kBins];  // exclusive prefix of the tile's digit counts
  uint32_t glob[kBins];      // where the tile's run of each digit starts in the output
  uint32_t run[kBins];       // items of each digit ranked so far in this tile
  uint32_t wcnt[2][kWaves][kBins];  // per-round, per-wave digit counts, tagged with the round (double-buffered)
  uint32_t w[kWaves + 1];
};

// offs: the scanned histogram (kBins * nblocks entries) followed by the total.
__global__ __launch_bounds__(kT) void k_radix_scatter(const uint64_t* kin, const uint32_t* vin, uint64_t* kout,
                                                     uint32_t* vout, int64_t n, int shift, const uint32_t* offs,
                                                     uint32_t nblocks) {
  RCM_LDS(ScatterSmem, sm);
  const int t = threadIdx.x;
  const int w = wave_id();
  const int lane = lane_id();
  const int64_t base = (int64_t)blockIdx.x * kTile;
  const int64_t tile_n64 = n - base < kTile ? n - base : kTile;
  const int tile_n = (int)tile_n64;
  {
    const size_t idx = (size_t)t * nblocks + blockIdx.x;
    const uint32_t start = offs[idx];
    const uint32_t count = offs[idx + 1] - start;
    sm.glob[t] = start;
    sm.run[t] = 0;
    for (int q = t; q < 2 * kWaves * kBins; q += kT) (&sm.wcnt[0][0][0])[q] = 0;
    uint32_t tot;
    sm.tile_off[t] = rcm::block_excl_scan<kT>(count, sm.w, &tot);  // barriers inside; tile_off read later
  }
  __syncthreads();
  for (int j = 0; j < kItems; ++j) {
    const int64_t i = base + (int64_t)j * kT + t;
    const bool valid = i < n;
    const uint64_t key = valid ? kin[i] : 0ull;
    const uint32_t val = valid ? vin[i] : 0u;
    const uint32_t d = valid ? ((uint32_t)(key >> shift) & (kBins - 1)) : (uint32_t)kBins;
    // lanes of this wave holding the same digit (9-bit match: kBins marks an empty slot)
    uint64_t peers = ~0ull;
#pragma unroll
    for (int bit = 0; bit < 9; ++bit) {
      const bool on = (d >> bit) & 1u;
      const uint64_t m = ballot(on);
      peers &= on ? m : ~m;
    }
    const uint32_t lower = (uint32_t)__popcll(peers & lanemask_lt());
    const uint32_t tag = (uint32_t)(j + 1) << 16;
    const int buf = j & 1;
    if (valid && (peers >> lane) == 1ull) sm.wcnt[buf][w][d] = tag | (lower + 1);  // highest lane of its group
    __syncthreads();
    if (valid) {
      uint32_t r = sm.run[d] + lower;
      for (int q = 0; q < w; ++q) {
        const uint32_t c = sm.wcnt[buf][q][d];
        r += (c >> 16) == (uint32_t)(j + 1) ? (c & 0xFFFFu) : 0u;
      }
      const uint32_t pos = sm.tile_off[d] + r;
      sm.key[pos] = key;
      sm.val[pos] = val;
    }
    __syncthreads();
    uint32_t add = 0;
    for (int q = 0; q < kWaves; ++q) {
      const uint32_t c = sm.wcnt[buf][q][t];
      add += (c >> 16) == (uint32_t)(j + 1) ? (c & 0xFFFFu) : 0u;
    }
    sm.run[t] += add;  // read again only after the next round's first barrier
  }
  __syncthreads();
  for (int j = 0; j < kItems; ++j) {
    const int pos = j * kT + t;
    if (pos < tile_n) {
      const uint64_t key = sm.key[pos];
      const uint32_t d = (uint32_t)(key >> shift) & (kBins - 1);
      const uint32_t dst = sm.glob[d] + ((uint32_t)pos - sm.tile_off[d]);
      kout[dst] = key;
      vout[dst] = sm.val[pos];
    }
  }
}

// ------------------------------------------------------------------------------------------------------
// Small element-wise kernels
// ------------------------------------------------------------------------------------------------------
__global__ __launch_bounds__(kT) void k_fill_u32(uint32_t* p, int64_t m, uint32_t v) {
  const int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (i < m) p[i] = v;
}

// CSR well-formedness: rp[0] == 0, rp[n] == nnz, monotone rows, columns in [0, cols). flag[0] = 1 if not.
__global__ __launch_bounds__(kT) void k_check_csr(const int32_t* rp, const int32_t* ci, int32_t n, int64_t nnz,
                                                 int32_t cols, uint32_t* flag) {
  const int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (i < n) {
    const int32_t a = rp[i], b = rp[i + 1];
    if (a > b || a < 0 || (int64_t)b > nnz) *flag = 1u;
  }
  if (i == 0 && (rp[0] != 0 || (int64_t)rp[n] != nnz)) *flag = 1u;
  if (i < nnz) {
    const int32_t c = ci[i];
    if (c < 0 || c >= cols) *flag = 1u;
  }
}

// inv[perm[i]] = i; flag[0] = 1 if perm is not a permutation of [0, n) (second kernel confirms bijectivity).
__global__ __launch_bounds__(kT) void k_perm_inverse(const int32_t* perm, int32_t n, int32_t* inv, uint32_t* flag) {
  const int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (i < n) {
    const int32_t p = perm[i];
    if (p < 0 || p >= n)
      *flag = 1u;
    else
      inv[p] = (int32_t)i;
  }
}
__global__ __launch_bounds__(kT) void k_perm_confirm(const int32_t* perm, int32_t n, const int32_t* inv,
                                                    uint32_t* flag) {
  const int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (i < n && inv[perm[i]] != (int32_t)i) *flag = 1u;
}

// ------------------------------------------------------------------------------------------------------
// B = P A P^T: row i of B is row perm[i] of A with column c renamed inv[c], columns sorted, entries with
// equal new column in their order in A.
// ------------------------------------------------------------------------------------------------------
// len[i] (to be scanned into B's row offsets), long_len[i] and long_cnt[i] for rows above kShortRow.
__global__ __launch_bounds__(kT) void k_perm_lengths(const int32_t* rp, const int32_t* perm, int32_t n,
                                                    uint32_t* len, uint32_t* long_len, uint32_t* long_cnt) {
  const int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (i < n) {
    const int32_t r = perm[i];
    const uint32_t l = (uint32_t)(rp[r + 1] - rp[r]);
    len[i] = l;
    long_len[i] = l > (uint32_t)kShortRow ? l : 0u;
    long_cnt[i] = l > (uint32_t)kShortRow ? 1u : 0u;
  } else if (i == n) {
    len[n] = long_len[n] = long_cnt[n] = 0u;
  }
}

struct ShortSmem {
  uint32_t loff[kRowsPerGroup + 1];
  int32_t src[kRowsPerGroup];
  int32_t dst[kRowsPerGroup];
  uint32_t nc[kRowsPerGroup * kShortRow];
  int32_t k[kRowsPerGroup * kShortRow];
  uint8_t row[kRowsPerGroup * kShortRow];
};

template <class V>
__global__ __launch_bounds__(kT) void k_perm_short(const int32_t* rp, const int32_t* ci, const V* vals,
                                                  const int32_t* perm, const int32_t* inv, const int32_t* rpB,
                                                  int32_t n, int32_t* ci_out, V* vals_out) {
  RCM_LDS(ShortSmem, sm);
  const int t = threadIdx.x;
  const int64_t r0 = (int64_t)blockIdx.x * kRowsPerGroup;
  if (t < kWave) {  // wave 0: one row per lane
    const int64_t i = r0 + t;
    uint32_t l = 0;
    int32_t src = 0, dst = 0;
    if (i < n) {
      src = rp[perm[i]];
      dst = rpB[i];
      l = (uint32_t)(rpB[i + 1] - rpB[i]);
      if (l > (uint32_t)kShortRow) l = 0;  // long rows go through the sort
    }
    const uint32_t inc = rcm::wave_incl_scan(l);
    sm.loff[t + 1] = inc;
    if (t == 0) sm.loff[0] = 0;
    sm.src[t] = src;
    sm.dst[t] = dst;
  }
  __syncthreads();
  const uint32_t total = sm.loff[kRowsPerGroup];
  for (uint32_t q = t; q < total; q += kT) {
    int lo = 0, hi = kRowsPerGroup;  // find the row r with loff[r] <= q < loff[r + 1]
    while (hi - lo > 1) {
      const int mid = (lo + hi) >> 1;
      if (sm.loff[mid] <= q)
        lo = mid;
      else
        hi = mid;
    }
    const int32_t k = sm.src[lo] + (int32_t)(q - sm.loff[lo]);
    sm.nc[q] = (uint32_t)inv[ci[k]];
    sm.k[q] = k;
    sm.row[q] = (uint8_t)lo;
  }
  __syncthreads();
  for (uint32_t q = t; q < total; q += kT) {
    const int r = sm.row[q];
    const uint32_t a = sm.loff[r], b = sm.loff[r + 1];
    const uint32_t my = sm.nc[q];
    uint32_t rank = 0;
    for (uint32_t f = a; f < b; ++f) {
      const uint32_t c = sm.nc[f];
      rank += (c < my || (c == my && f < q)) ? 1u : 0u;
    }
    const int32_t dst = sm.dst[r] + (int32_t)rank;
    ci_out[dst] = (int32_t)my;
    if (vals) vals_out[dst] = vals[sm.k[q]];
  }
}

// list[ordinal] = row, for rows longer than kShortRow (long_ord: scanned long_cnt)
__global__ __launch_bounds__(kT) void k_long_list(const int32_t* rpB, const uint32_t* long_ord, int32_t n,
                                                 int32_t* list) {
  const int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (i < n && rpB[i + 1] - rpB[i] > kShortRow) list[long_ord[i]] = (int32_t)i;
}

// One workgroup per long row: key = (ordinal << 32) | new column, payload = source entry index.
__global__ __launch_bounds__(kT) void k_perm_long_gather(const int32_t* rp, const int32_t* ci, const int32_t* perm,
                                                        const int32_t* inv, const int32_t* rpB,
                                                        const uint32_t* long_off, const int32_t* list,
                                                        uint64_t* keys, uint32_t* kidx) {
  const uint32_t o = blockIdx.x;
  const int32_t i = list[o];
  const int32_t src = rp[perm[i]];
  const int32_t l = rpB[i + 1] - rpB[i];
  const uint32_t out = long_off[i];
  for (int32_t j = threadIdx.x; j < l; j += kT) {
    const int32_t k = src + j;
    keys[out + (uint32_t)j] = ((uint64_t)o << 32) | (uint32_t)inv[ci[k]];
    kidx[out + (uint32_t)j] = (uint32_t)k;
  }
}

template <class V>
__global__ __launch_bounds__(kT) void k_perm_long_scatter(const uint64_t* keys, const uint32_t* kidx, int64_t m,
                                                         const int32_t* list, const uint32_t* long_off,
                                                         const int32_t* rpB, const V* vals, int32_t* ci_out,
                                                         V* vals_out) {
  const int64_t q = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (q >= m) return;
  const uint64_t key = keys[q];
  const int32_t i = list[(uint32_t)(key >> 32)];
  const int64_t dst = (int64_t)rpB[i] + (q - (int64_t)long_off[i]);
  ci_out[dst] = (int32_t)(uint32_t)key;
  if (vals) vals_out[dst] = vals[kidx[q]];
}

// ------------------------------------------------------------------------------------------------------
// COO -> CSR (stable on (row, column)), symmetrisation and connectivity repair.
// ------------------------------------------------------------------------------------------------------
// keys[k] = (row << 32) | col, idx[k] = k; flag[0] = 1 for an index out of range.
__global__ __launch_bounds__(kT) void k_coo_keys(const int32_t* r, const int32_t* c, int64_t nnz, int32_t rows,
                                                int32_t cols, uint64_t* keys, uint32_t* idx, uint32_t* flag) {
  const int64_t k = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (k >= nnz) return;
  const int32_t a = r[k], b = c[k];
  if (a < 0 || a >= rows || b < 0 || b >= cols) {
    *flag = 1u;
    keys[k] = 0;
  } else {
    keys[k] = ((uint64_t)(uint32_t)a << 32) | (uint32_t)b;
  }
  idx[k] = (uint32_t)k;
}

// Row offsets of keys sorted by row: entry q starts every row in (row(q - 1), row(q)]; the last entry also
// closes the rows after its own.
__global__ __launch_bounds__(kT) void k_row_starts(const uint64_t* keys, int64_t nnz, int32_t rows, int32_t* rp) {
  const int64_t q = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (q >= nnz) return;
  const int64_t r = (int64_t)(keys[q] >> 32);
  const int64_t prev = q > 0 ? (int64_t)(keys[q - 1] >> 32) : -1;
  for (int64_t x = prev + 1; x <= r; ++x) rp[x] = (int32_t)q;
  if (q == nnz - 1)
    for (int64_t x = r + 1; x <= rows; ++x) rp[x] = (int32_t)nnz;
}

// Rows longer than kShortRow of the row-sorted entries: long_cnt[r] = 1 and long_len[r] = the length for them,
// 0 otherwise (entry n of both is 0, for the scans).
__global__ __launch_bounds__(kT) void k_long_rows(const int32_t* rp, int32_t n, uint32_t* long_len,
                                                 uint32_t* long_cnt) {
  const int64_t r = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (r > n) return;
  const int32_t l = r < n ? rp[r + 1] - rp[r] : 0;
  long_cnt[r] = l > kShortRow ? 1u : 0u;
  long_len[r] = l > kShortRow ? (uint32_t)l : 0u;
}

// Each row of at most kShortRow entries sorted by column, stably (insertion sort; the entries of a row are in
// input order after the row sort, so equal columns keep it).
__global__ __launch_bounds__(kT) void k_sort_short_rows(const int32_t* rp, int32_t n, uint64_t* keys,
                                                       uint32_t* idx) {
  const int64_t r = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (r >= n) return;
  const int32_t b = rp[r], e = rp[r + 1];
  if (e - b > kShortRow) return;
  for (int32_t i = b + 1; i < e; ++i) {
    const uint64_t k = keys[i];
    const uint32_t v = idx[i];
    int32_t j = i;
    while (j > b && (uint32_t)keys[j - 1] > (uint32_t)k) {
      keys[j] = keys[j - 1];
      idx[j] = idx[j - 1];
      --j;
    }
    keys[j] = k;
    idx[j] = v;
  }
}

// One workgroup per long row: gkey = (ordinal << 32) | column, gval = the entry's input index.
__global__ __launch_bounds__(kT) void k_coo_long_gather(const uint64_t* keys, const uint32_t* idx, const int32_t* rp,
                                                       const int32_t* list, const uint32_t* long_off,
                                                       uint64_t* gkey, uint32_t* gval) {
  const uint32_t o = blockIdx.x;
  const int32_t r = list[o];
  const int32_t b = rp[r], l = rp[r + 1] - b;
  const uint32_t out = long_off[r];
  for (int32_t j = threadIdx.x; j < l; j += kT) {
    gkey[out + (uint32_t)j] = ((uint64_t)o << 32) | (uint32_t)keys[b + j];
    gval[out + (uint32_t)j] = idx[b + j];
  }
}

// The sorted long-row entries back into their rows.
__global__ __launch_bounds__(kT) void k_coo_long_scatter(const uint64_t* gkey, const uint32_t* gval, int64_t m,
                                                        const int32_t* list, const uint32_t* long_off,
                                                        const int32_t* rp, uint64_t* keys, uint32_t* idx) {
  const int64_t j = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (j >= m) return;
  const uint64_t k = gkey[j];
  const int32_t r = list[(uint32_t)(k >> 32)];
  const int64_t dst = (int64_t)rp[r] + (j - (int64_t)long_off[r]);
  keys[dst] = ((uint64_t)(uint32_t)r << 32) | (uint32_t)k;
  idx[dst] = gval[j];
}

template <class V>
__global__ __launch_bounds__(kT) void k_csr_from_sorted(const uint64_t* keys, const uint32_t* idx, int64_t nnz,
                                                       const V* vin, int32_t* ci_out, V* vout) {
  const int64_t q = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (q >= nnz) return;
  ci_out[q] = (int32_t)(uint32_t)keys[q];
  if (vin) vout[q] = vin[idx[q]];
}

// Symmetrisation (include/Graph.h:77-157): an unordered pair stored an odd number of times becomes symmetric
// (each of its entries gets its transpose), an even pair stays as stored. Decided per entry of the row-sorted
// CSR: entry (r, c) of a pair stored m_rc times as (r, c) (its run in row r) and m_cr times as (c, r) (a binary
// search in row c) is odd when m_rc + m_cr is; an even pair with m_rc != m_cr leaves the result not structurally
// symmetric: flag[0] = 1.
__global__ __launch_bounds__(kT) void k_pair_parity(const uint64_t* keys, const int32_t* rp, const int32_t* ci,
                                                   int32_t rows, int64_t nnz, uint32_t* odd, uint32_t* flag) {
  const int64_t k = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (k >= nnz) return;
  const uint64_t key = keys[k];
  const int32_t r = (int32_t)(key >> 32), c = (int32_t)(uint32_t)key;
  if (r == c) {
    odd[k] = 0u;
    return;
  }
  int64_t lo = k, hi = k + 1;
  const int64_t rb = rp[r], re = rp[r + 1];
  while (lo > rb && ci[lo - 1] == c) --lo;
  while (hi < re && ci[hi] == c) ++hi;
  const int64_t m_rc = hi - lo;
  int64_t m_cr = 0;
  if (c < rows) {
    int64_t a = rp[c], b = rp[c + 1];  // first entry >= r
    while (a < b) {
      const int64_t mid = (a + b) >> 1;
      if (ci[mid] < r) a = mid + 1;
      else b = mid;
    }
    int64_t e = a, f = rp[c + 1];  // first entry > r
    while (e < f) {
      const int64_t mid = (e + f) >> 1;
      if (ci[mid] <= r) e = mid + 1;
      else f = mid;
    }
    m_cr = e - a;
  }
  const bool is_odd = ((m_rc + m_cr) & 1) != 0;
  odd[k] = is_odd ? 1u : 0u;
  if (!is_odd && m_rc != m_cr) *flag = 1u;
}

// flag[0] = 1 for an empty row
__global__ __launch_bounds__(kT) void k_empty_rows(const int32_t* rp, int32_t n, uint32_t* flag) {
  const int64_t r = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (r < n && rp[r] == rp[r + 1]) *flag = 1u;
}

// The symmetrised CSR without a second sort of every entry. The reference inserts each entry and, when its pair is
// odd, its transpose right after it, and converts that COO stably, so a row of the result is ordered by (column,
// insertion position). Entry q of the row-sorted CSR has position P(q) = q + opos[q] (opos: scanned odd flags,
// nnz + 1 entries) and its transpose P(q) + 1. Row r of the result holds row r's entries, already in (column,
// position) order, and the transposes landing in row r, in the same order once sorted stably by that row (their
// sources come in row order). An entry's place is its index in its group plus the number of entries of the other
// group before it, by binary search: no per-row sort, and hub rows stay parallel.

// The transposes (c, r) of the odd entries (r, c): key c << 32 | r, payload the source entry.
__global__ __launch_bounds__(kT) void k_sym_transposes(const uint64_t* keys, const uint32_t* opos, int64_t nnz,
                                                      uint64_t* tkeys, uint32_t* tsrc) {
  const int64_t q = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (q >= nnz || opos[q + 1] == opos[q]) return;
  const uint64_t key = keys[q];
  const uint32_t t = opos[q];
  tkeys[t] = (key << 32) | (key >> 32);
  tsrc[t] = (uint32_t)q;
}

// For the row-sorted transposes: their (column, position) as one comparable word.
__global__ __launch_bounds__(kT) void k_sym_tkey(const uint64_t* tkeys, const uint32_t* tsrc, const uint32_t* opos,
                                                int64_t nt, uint64_t* tcp) {
  const int64_t t = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (t >= nt) return;
  const uint32_t src = tsrc[t];
  tcp[t] = ((uint64_t)(uint32_t)tkeys[t] << 32) | (uint64_t)(src + opos[src] + 1u);
}

// Row lengths of the result (s rows): row r's entries (r < rows) and its transposes; entry s is 0 for the scan.
__global__ __launch_bounds__(kT) void k_sym_lengths(const int32_t* rp, int32_t rows, const int32_t* tstart, int32_t s,
                                                   uint32_t* len) {
  const int64_t r = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (r > s) return;
  len[r] = r == s ? 0u
                  : (uint32_t)((r < rows ? rp[r + 1] - rp[r] : 0) + (tstart[r + 1] - tstart[r]));
}

// Entry q of the row-sorted CSR into the result (rpo: its row offsets).
template <class V>
__global__ __launch_bounds__(kT) void k_sym_place_entries(const uint64_t* keys, const uint32_t* opos, int64_t nnz,
                                                         const int32_t* rp, const uint64_t* tcp,
                                                         const int32_t* tstart, const int32_t* rpo, const V* vals,
                                                         int32_t* ci_out, V* v_out) {
  const int64_t q = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (q >= nnz) return;
  const uint64_t key = keys[q];
  const int32_t r = (int32_t)(key >> 32);
  const uint64_t me = (key << 32) | (uint64_t)((uint32_t)q + opos[q]);
  int64_t lo = tstart[r], hi = tstart[r + 1];  // transposes of row r before this entry
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if (tcp[mid] < me) lo = mid + 1;
    else hi = mid;
  }
  const int64_t pos = (int64_t)rpo[r] + (q - rp[r]) + (lo - tstart[r]);
  ci_out[pos] = (int32_t)(uint32_t)key;
  if (vals) v_out[pos] = vals[q];
}

// Transpose t (row-sorted) into the result.
template <class V>
__global__ __launch_bounds__(kT) void k_sym_place_transposes(const uint64_t* tkeys, const uint64_t* tcp,
                                                            const uint32_t* tsrc, int64_t nt, const int32_t* tstart,
                                                            const int32_t* rp, const int32_t* ci, const uint32_t* opos,
                                                            int32_t rows, const int32_t* rpo, const V* vals,
                                                            int32_t* ci_out, V* v_out) {
  const int64_t t = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (t >= nt) return;
  const int32_t r = (int32_t)(tkeys[t] >> 32);
  const uint64_t me = tcp[t];
  int64_t lo = 0, hi = 0;  // entries of row r before this transpose
  if (r < rows) {
    const int64_t b = rp[r];
    lo = b;
    hi = rp[r + 1];
    while (lo < hi) {
      const int64_t mid = (lo + hi) >> 1;
      const uint64_t w = ((uint64_t)(uint32_t)ci[mid] << 32) | (uint64_t)((uint32_t)mid + opos[mid]);
      if (w < me) lo = mid + 1;
      else hi = mid;
    }
    lo -= b;
  }
  const int64_t pos = (int64_t)rpo[r] + (t - tstart[r]) + lo;
  ci_out[pos] = (int32_t)(me >> 32);
  if (vals) v_out[pos] = vals[tsrc[t]];
}

// flag[0] = 1: an empty row; flag[1] = 1: an entry (r, c) without (c, r) (columns are sorted)
__global__ __launch_bounds__(kT) void k_struct_checks(const int32_t* rp, const int32_t* ci, int32_t n, uint32_t* flag) {
  const int64_t r = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (r >= n) return;
  const int32_t a = rp[r], b = rp[r + 1];
  if (a == b) flag[0] = 1u;
  for (int32_t k = a; k < b; ++k) {
    const int32_t c = ci[k];
    int32_t lo = rp[c], hi = rp[c + 1];  // lower bound of r in row c
    while (lo < hi) {
      const int32_t mid = lo + ((hi - lo) >> 1);
      if (ci[mid] < (int32_t)r)
        lo = mid + 1;
      else
        hi = mid;
    }
    if (lo == rp[c + 1] || ci[lo] != (int32_t)r) {
      flag[1] = 1u;
      return;
    }
  }
}

// parent[v] = the smallest entry of v's row if it is below v (sorted rows: the first column), else v. Each such
// parent is an edge of the graph, so the forest starts partly joined, without atomics.
__global__ __launch_bounds__(kT) void k_cc_init(const int32_t* rp, const int32_t* ci, int32_t n, uint32_t* parent) {
  const int64_t v = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (v >= n) return;
  const int32_t a = rp[v];
  parent[v] = (a < rp[v + 1] && ci[a] < (int32_t)v) ? (uint32_t)ci[a] : (uint32_t)v;
}

// Root of x with path halving. Parents only ever move to smaller indices, so every chain ends.
__device__ __forceinline__ uint32_t cc_find(uint32_t* parent, uint32_t x) {
  uint32_t p = ld_u32(parent + x);
  while (p != x) {
    const uint32_t g = ld_u32(parent + p);
    if (g != p) st_u32(parent + x, g);
    x = p;
    p = g;
  }
  return x;
}

// The same walk with loads and stores that may be served by this XCD's L2 (possibly stale). Every parent
// value ever stored is below its node, so any value read is an ancestor of the node at some time, the walk
// strictly decreases, and trees only merge: a stale walk ends at a node of the right tree.
__device__ __forceinline__ uint32_t cc_find_cached(uint32_t* parent, uint32_t x) {
  uint32_t p = __hip_atomic_load(parent + x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
  while (p != x) {
    const uint32_t g = __hip_atomic_load(parent + p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    if (g != p) __hip_atomic_store(parent + x, g, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    x = p;
    p = g;
  }
  return x;
}

// Concurrent union-find: the trees of u and v are hooked, the larger root under the smaller one, with a
// device-coherent compare-and-swap, so each tree's root is its smallest node. Walks use cached loads; a CAS only
// succeeds on a true root, and when it fails the walk restarts from the parent it returned, so the larger of the
// two candidates strictly decreases and the loop ends. (Device-coherent walks measured slower: 5.7 against 4.2 ms
// for the whole canonicalisation of the 128^3 stencil.)
__device__ __forceinline__ void cc_link(uint32_t* parent, uint32_t u, uint32_t v) {
  uint32_t a = cc_find_cached(parent, u), b = cc_find_cached(parent, v);
  while (a != b) {
    const uint32_t hi = a > b ? a : b, lo = a > b ? b : a;
    uint32_t expected = hi;
    if (__hip_atomic_compare_exchange_strong(parent + hi, &expected, lo, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                             __HIP_MEMORY_SCOPE_AGENT))
      break;
    a = cc_find_cached(parent, expected);
    b = cc_find_cached(parent, lo);
  }
}

// Every edge (u, v), v < u, of each row (sorted columns: each undirected edge once, from its larger end), except
// the row's first entry when it is below u: k_cc_init already made it u's parent.
__global__ __launch_bounds__(kT) void k_cc_hook(const int32_t* rp, const int32_t* ci, int32_t n, uint32_t* parent) {
  const int64_t u = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (u >= n) return;
  const int32_t k0 = rp[u], k1 = rp[u + 1];
  for (int32_t k = (k0 < k1 && ci[k0] < (int32_t)u) ? k0 + 1 : k0; k < k1; ++k) {
    const uint32_t v = (uint32_t)ci[k];
    if (v >= (uint32_t)u) break;
    cc_link(parent, (uint32_t)u, v);
  }
}

__global__ __launch_bounds__(kT) void k_cc_finish(uint32_t* parent, int32_t n) {
  const int64_t v = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (v < n) st_u32(parent + v, cc_find(parent, (uint32_t)v));
}

// New row lengths with the links (root - 1, root) of every component other than node 0's.
__global__ __launch_bounds__(kT) void k_link_lengths(const int32_t* rp, const uint32_t* root, int32_t n,
                                                    uint32_t* len) {
  const int64_t x = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (x < n) {
    uint32_t l = (uint32_t)(rp[x + 1] - rp[x]);
    if (x > 0 && root[x] == (uint32_t)x) ++l;
    if (x + 1 < n && root[x + 1] == (uint32_t)(x + 1)) ++l;
    len[x] = l;
  } else if (x == n) {
    len[n] = 0;
  }
}

// Row x with its link entries merged in column order (value 1): x - 1 if x is a component root, x + 1 if
// x + 1 is one. Neither column is already present (the nodes lie in different components).
template <class V>
__global__ __launch_bounds__(kT) void k_link_insert(const int32_t* rp, const int32_t* ci, const V* vals,
                                                   const uint32_t* root, int32_t n, const int32_t* rp2,
                                                   int32_t* ci2, V* vals2) {
  const int64_t x = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (x >= n) return;
  int32_t extra[2];
  int ne = 0;
  if (x > 0 && root[x] == (uint32_t)x) extra[ne++] = (int32_t)x - 1;
  if (x + 1 < n && root[x + 1] == (uint32_t)(x + 1)) extra[ne++] = (int32_t)x + 1;
  int32_t o = rp2[x];
  int e = 0;
  for (int32_t k = rp[x]; k < rp[x + 1]; ++k) {
    const int32_t c = ci[k];
    while (e < ne && extra[e] < c) {
      ci2[o] = extra[e++];
      if (vals) vals2[o] = (V)1;
      ++o;
    }
    ci2[o] = c;
    if (vals) vals2[o] = vals[k];
    ++o;
  }
  while (e < ne) {
    ci2[o] = extra[e++];
    if (vals) vals2[o] = (V)1;
    ++o;
  }
}

// ------------------------------------------------------------------------------------------------------
// Connectivity repair of a matrix that is not structurally symmetric (include/Graph.h:160-263).
// The reference marks the first non-empty row m and everything reachable from it along stored entries
// (row -> column); then, while a node is unmarked, it takes the lowest one, i, links (i - 1, i) (every node
// below i is marked), pushes i without marking it and marks what the push reaches. The marked set is always
// the closure of [0, i) and m under the entries, so with anc(x) = the smallest node with a path of length >= 1
// to x:
//   x > m is linked  <=>  no node below x reaches x  <=>  anc(x) >= x;
// a linked x is marked by its own push only when it lies on a cycle (anc(x) == x). Otherwise the reference
// links it again forever, and so it does for a node below m (an empty row) that m does not reach: both are
// the non-terminating case. anc is the fixed point of anc(c) = min over entries (p, c) of min(p, anc(p)),
// reached by passes over the entries until one changes nothing.
// ------------------------------------------------------------------------------------------------------
// *out = the first non-empty row (out starts at n)
__global__ __launch_bounds__(kT) void k_first_nonempty(const int32_t* rp, int32_t n, uint32_t* out) {
  const int64_t x = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (x < n && rp[x + 1] > rp[x]) atomicMin(out, (uint32_t)x);
}

// One pass: every entry (p, c) lowers anc(c) to min(p, anc(p)); *changed = 1 when any value fell.
__global__ __launch_bounds__(kT) void k_anc_pass(const int32_t* rp, const int32_t* ci, int32_t n, uint32_t* anc,
                                                 uint32_t* changed) {
  const int64_t p = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (p >= n) return;
  const uint32_t v = min((uint32_t)p, ld_u32(anc + p));  // values only fall: a stale read is an upper bound
  bool fell = false;
  for (int32_t k = rp[p]; k < rp[p + 1]; ++k) {
    uint32_t* t = anc + ci[k];
    if (ld_u32(t) > v && rcm::atomic_min_u32(t, v) > v) fell = true;
  }
  if (fell) st_u32(changed, 1u);
}

// root[x] = x for a linked node, ~0 otherwise; flag[0] = 1 for the non-terminating case.
__global__ __launch_bounds__(kT) void k_dir_roots(const uint32_t* anc, int32_t n, uint32_t m, uint32_t* root,
                                                  uint32_t* flag) {
  const int64_t xi = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (xi >= n) return;
  const uint32_t x = (uint32_t)xi, a = anc[x];
  uint32_t r = ~0u;
  if (x < m) {
    if (a != m) flag[0] = 1u;  // an empty row m does not reach
  } else if (x > m) {
    if (a == x) r = x;             // linked, marked by its own push (a cycle)
    else if (a > x) flag[0] = 1u;  // linked, never marked
  }
  root[x] = r;
}

// ------------------------------------------------------------------------------------------------------
// Host orchestration, generic over the executor X:
//   template <class K, class... A> void run(K kernel, int64_t grid, A... args);   // kT threads per group
//   template <class T> T* alloc(size_t count);   // scratch, freed by the executor
//   uint32_t read_u32(const uint32_t* p);          // synchronous readback
//   bool failed();                                 // a launch or allocation failed
// ------------------------------------------------------------------------------------------------------
inline int64_t groups(int64_t items, int64_t per) { return (items + per - 1) / per; }

template <class X>
void scan_u32(X& x, const uint32_t* in, uint32_t* out, int64_t m, uint32_t* total_out) {
  const int64_t tiles = groups(m > 0 ? m : 1, kTile);
  uint32_t* partial = x.template alloc<uint32_t>((size_t)tiles);
  if (!partial) return;
  x.run(k_scan_reduce, tiles, in, m, partial);
  x.run(k_scan_top, 1, partial, tiles, total_out);
  x.run(k_scan_apply, tiles, in, out, m, (const uint32_t*)partial);
}

// Shifts of the 8-bit passes that cover `lo_bits` of the low word and `hi_bits` of the high word.
inline int radix_shifts(int lo_bits, int hi_bits, int* shifts) {
  int k = 0;
  for (int s = 0; s < lo_bits; s += 8) shifts[k++] = s;
  for (int s = 0; s < hi_bits; s += 8) shifts[k++] = 32 + s;
  return k;
}

inline int bits_of(uint64_t maxval) {
  int b = 0;
  while (maxval) {
    ++b;
    maxval >>= 1;
  }
  return b;
}

// Sorts (keys, vals) stably; the result ends in whichever buffer pair the last pass wrote: returned via the
// references. alt buffers must hold n items.
template <class X>
void radix_sort(X& x, uint64_t*& keys, uint32_t*& vals, uint64_t*& keys_alt, uint32_t*& vals_alt, int64_t n,
                int lo_bits, int hi_bits) {
  if (n <= 1) return;
  int shifts[8];
  const int passes = radix_shifts(lo_bits, hi_bits, shifts);
  const int64_t tiles = groups(n, kTile);
  const int64_t m = (int64_t)kBins * tiles;
  uint32_t* hist = x.template alloc<uint32_t>((size_t)m + 1);
  if (!hist) return;
  for (int p = 0; p < passes; ++p) {
    x.run(k_radix_hist, tiles, (const uint64_t*)keys, n, shifts[p], hist, (uint32_t)tiles);
    scan_u32(x, hist, hist, m, hist + m);
    x.run(k_radix_scatter, tiles, (const uint64_t*)keys, (const uint32_t*)vals, keys_alt, vals_alt, n, shifts[p],
          (const uint32_t*)hist, (uint32_t)tiles);
    uint64_t* tk = keys;
    keys = keys_alt;
    keys_alt = tk;
    uint32_t* tv = vals;
    vals = vals_alt;
    vals_alt = tv;
  }
}

enum OpsStatus { kOpsOk = 0, kOpsInvalid = -1, kOpsDevice = -4, kOpsNotSymmetric = -12, kOpsEmptyRow = -13 };

// B = P A P^T. rp_out: n + 1, ci_out / vals_out: nnz. vals may be null.
template <class X, class V>
int permute(X& x, int32_t n, int64_t nnz, const int32_t* rp, const int32_t* ci, const V* vals, const int32_t* perm,
            int32_t* rp_out, int32_t* ci_out, V* vals_out) {
  uint32_t* flag = x.template alloc<uint32_t>(1);
  int32_t* inv = x.template alloc<int32_t>((size_t)n);
  uint32_t* long_len = x.template alloc<uint32_t>((size_t)n + 1);
  uint32_t* long_cnt = x.template alloc<uint32_t>((size_t)n + 1);
  uint32_t* tot = x.template alloc<uint32_t>(2);
  if (x.failed()) return kOpsDevice;
  x.run(k_fill_u32, 1, flag, (int64_t)1, 0u);
  x.run(k_check_csr, groups(nnz > n ? nnz : (int64_t)n, kT), rp, ci, n, nnz, n, flag);
  x.run(k_perm_inverse, groups(n, kT), perm, n, inv, flag);
  if (x.read_u32(flag)) return kOpsInvalid;  // perm out of range or malformed CSR
  x.run(k_perm_confirm, groups(n, kT), perm, n, (const int32_t*)inv, flag);
  if (x.read_u32(flag)) return kOpsInvalid;  // perm repeats a node
  uint32_t* len = reinterpret_cast<uint32_t*>(rp_out);
  x.run(k_perm_lengths, groups((int64_t)n + 1, kT), rp, perm, n, len, long_len, long_cnt);
  scan_u32(x, len, len, (int64_t)n + 1, (uint32_t*)nullptr);
  scan_u32(x, long_len, long_len, (int64_t)n + 1, tot);
  scan_u32(x, long_cnt, long_cnt, (int64_t)n + 1, tot + 1);
  x.run(k_perm_short<V>, groups(n, kRowsPerGroup), rp, ci, vals, perm, (const int32_t*)inv, (const int32_t*)rp_out,
        n, ci_out, vals_out);
  const uint32_t long_entries = x.read_u32(tot);
  const uint32_t long_rows = x.read_u32(tot + 1);
  if (x.failed()) return kOpsDevice;
  if (long_rows == 0) return kOpsOk;
  int32_t* list = x.template alloc<int32_t>(long_rows);
  uint64_t* keys = x.template alloc<uint64_t>(long_entries);
  uint64_t* keys2 = x.template alloc<uint64_t>(long_entries);
  uint32_t* kidx = x.template alloc<uint32_t>(long_entries);
  uint32_t* kidx2 = x.template alloc<uint32_t>(long_entries);
  if (x.failed()) return kOpsDevice;
  x.run(k_long_list, groups(n, kT), (const int32_t*)rp_out, (const uint32_t*)long_cnt, n, list);
  x.run(k_perm_long_gather, (int64_t)long_rows, rp, ci, perm, (const int32_t*)inv, (const int32_t*)rp_out,
        (const uint32_t*)long_len, (const int32_t*)list, keys, kidx);
  radix_sort(x, keys, kidx, keys2, kidx2, (int64_t)long_entries, bits_of((uint64_t)(n - 1)),
             bits_of((uint64_t)long_rows - 1));
  x.run(k_perm_long_scatter<V>, groups(long_entries, kT), (const uint64_t*)keys, (const uint32_t*)kidx,
        (int64_t)long_entries, (const int32_t*)list, (const uint32_t*)long_len, (const int32_t*)rp_out, vals,
        ci_out, vals_out);
  return x.failed() ? kOpsDevice : kOpsOk;
}

// Output CSR of canonicalise(); device buffers allocated through the executor's `keep` allocator.
template <class V>
struct Csr {
  int32_t rows = 0, cols = 0;
  int64_t nnz = 0;
  int32_t* rp = nullptr;
  int32_t* ci = nullptr;
  V* vals = nullptr;
};

// COO -> CSR, stable on (row, column).
template <class X, class V>
int coo_to_csr(X& x, int32_t rows, int32_t cols, int64_t nnz, const int32_t* r, const int32_t* c, const V* vals,
               Csr<V>& out, uint64_t** sorted_keys) {
  uint32_t* flag = x.template alloc<uint32_t>(1);
  uint64_t* keys = x.template alloc<uint64_t>((size_t)nnz);
  uint64_t* keys2 = x.template alloc<uint64_t>((size_t)nnz);
  uint32_t* idx = x.template alloc<uint32_t>((size_t)nnz);
  uint32_t* idx2 = x.template alloc<uint32_t>((size_t)nnz);
  out.rows = rows;
  out.cols = cols;
  out.nnz = nnz;
  out.rp = x.template keep<int32_t>((size_t)rows + 1);
  out.ci = x.template keep<int32_t>((size_t)nnz);
  out.vals = vals ? x.template keep<V>((size_t)nnz) : nullptr;
  if (x.failed()) return kOpsDevice;
  x.run(k_fill_u32, 1, flag, (int64_t)1, 0u);
  if (nnz > 0) x.run(k_coo_keys, groups(nnz, kT), r, c, nnz, rows, cols, keys, idx, flag);
  if (x.read_u32(flag)) return kOpsInvalid;
  // rows first: a stable sort on the row bits alone keeps each row's entries in input order; then every row
  // by column, stably -- short rows by one thread each, long rows by a radix sort over (long-row ordinal,
  // column). Fewer passes over all entries than a sort on both words.
  radix_sort(x, keys, idx, keys2, idx2, nnz, 0, bits_of(rows > 0 ? (uint64_t)(rows - 1) : 0));
  if (nnz > 0) {
    x.run(k_row_starts, groups(nnz, kT), (const uint64_t*)keys, nnz, rows, out.rp);
    {
      uint32_t* long_len = x.template alloc<uint32_t>((size_t)rows + 1);
      uint32_t* long_cnt = x.template alloc<uint32_t>((size_t)rows + 1);
      uint32_t* tot = x.template alloc<uint32_t>(2);
      if (x.failed()) return kOpsDevice;
      x.run(k_long_rows, groups((int64_t)rows + 1, kT), (const int32_t*)out.rp, rows, long_len, long_cnt);
      scan_u32(x, long_len, long_len, (int64_t)rows + 1, tot);
      scan_u32(x, long_cnt, long_cnt, (int64_t)rows + 1, tot + 1);
      x.run(k_sort_short_rows, groups(rows, kT), (const int32_t*)out.rp, rows, keys, idx);
      const uint32_t long_entries = x.read_u32(tot), long_rows = x.read_u32(tot + 1);
      if (x.failed()) return kOpsDevice;
      if (long_rows > 0) {
        int32_t* list = x.template alloc<int32_t>(long_rows);
        uint64_t* gk = x.template alloc<uint64_t>(long_entries);
        uint64_t* gk2 = x.template alloc<uint64_t>(long_entries);
        uint32_t* gv = x.template alloc<uint32_t>(long_entries);
        uint32_t* gv2 = x.template alloc<uint32_t>(long_entries);
        if (x.failed()) return kOpsDevice;
        x.run(k_long_list, groups(rows, kT), (const int32_t*)out.rp, (const uint32_t*)long_cnt, rows, list);
        x.run(k_coo_long_gather, (int64_t)long_rows, (const uint64_t*)keys, (const uint32_t*)idx,
              (const int32_t*)out.rp, (const int32_t*)list, (const uint32_t*)long_len, gk, gv);
        radix_sort(x, gk, gv, gk2, gv2, (int64_t)long_entries, bits_of(cols > 0 ? (uint64_t)(cols - 1) : 0),
                   bits_of((uint64_t)long_rows - 1));
        x.run(k_coo_long_scatter, groups(long_entries, kT), (const uint64_t*)gk, (const uint32_t*)gv,
              (int64_t)long_entries, (const int32_t*)list, (const uint32_t*)long_len, (const int32_t*)out.rp, keys,
              idx);
      }
    }
    x.run(k_csr_from_sorted<V>, groups(nnz, kT), (const uint64_t*)keys, (const uint32_t*)idx, nnz, vals, out.ci,
          out.vals);
  } else {
    x.run(k_fill_u32, groups((int64_t)rows + 1, kT), reinterpret_cast<uint32_t*>(out.rp), (int64_t)rows + 1, 0u);
  }
  if (sorted_keys) *sorted_keys = keys;
  return x.failed() ? kOpsDevice : kOpsOk;
}

// Components of a structurally symmetric matrix: root[v] = the smallest node of v's component. The forest starts
// from each node's smallest lower neighbour, is flattened, and then every lower edge is hooked. (Sampled linking in
// the Afforest scheme, two rounds of first and second row entries and then only rows outside node 0's tree, measured
// slower on the relabelled 128^3 stencil: its second round alone took 2 ms, the full hook pass 1.3 ms.)
template <class X>
int components(X& x, int32_t n, const int32_t* rp, const int32_t* ci, uint32_t* root) {
  x.run(k_cc_init, groups(n, kT), rp, ci, n, root);
  x.run(k_cc_finish, groups(n, kT), root, n);
  x.run(k_cc_hook, groups(n, kT), rp, ci, n, root);
  x.run(k_cc_finish, groups(n, kT), root, n);
  return x.failed() ? kOpsDevice : kOpsOk;
}

// The linked nodes of a matrix that is not structurally symmetric: root[x] = x for each, ~0 elsewhere.
template <class X>
int directed_roots(X& x, int32_t n, const int32_t* rp, const int32_t* ci, uint32_t* root) {
  uint32_t* w = x.template alloc<uint32_t>(3);  // first non-empty row, changed, non-terminating
  uint32_t* anc = x.template alloc<uint32_t>((size_t)n);
  if (x.failed()) return kOpsDevice;
  x.run(k_fill_u32, 1, w, (int64_t)1, (uint32_t)n);
  x.run(k_first_nonempty, groups(n, kT), rp, n, w);
  const uint32_t m = x.read_u32(w);
  if (m >= (uint32_t)n) return kOpsEmptyRow;  // every row empty
  x.run(k_fill_u32, groups(n, kT), anc, (int64_t)n, ~0u);
  for (int64_t pass = 0;; ++pass) {  // a value falls along one more entry each pass at least
    if (pass > (int64_t)n) return kOpsDevice;
    x.run(k_fill_u32, 1, w + 1, (int64_t)1, 0u);
    x.run(k_anc_pass, groups(n, kT), rp, ci, n, anc, w + 1);
    if (!x.read_u32(w + 1)) break;
    if (x.failed()) return kOpsDevice;
  }
  x.run(k_fill_u32, 1, w + 2, (int64_t)1, 0u);
  x.run(k_dir_roots, groups(n, kT), (const uint32_t*)anc, n, m, root, w + 2);
  if (x.read_u32(w + 2)) return kOpsEmptyRow;
  return x.failed() ? kOpsDevice : kOpsOk;
}

// convert + prepareRCM(check_sym) (src/CSR.cpp:170-208, include/Graph.h:74-264). A structurally symmetric result
// is repaired by a union-find over its components (an empty row is the reference's non-terminating case); any
// other by the directed characterisation above. *links_out: links added.
template <class X, class V>
int canonicalise(X& x, int32_t rows, int32_t cols, int64_t nnz, const int32_t* r, const int32_t* c, const V* vals,
                 bool check_sym, Csr<V>& out, int64_t* links_out) {
  Csr<V> a;
  uint64_t* skeys = nullptr;
  int rc = coo_to_csr(x, rows, cols, nnz, r, c, vals, a, &skeys);
  if (rc) return rc;
  bool symmetric_known = false;  // the pair pass decided structural symmetry
  bool nonsym = false;           // the result is not structurally symmetric: the directed repair runs
  if (check_sym && nnz > 0) {
    uint32_t* tot = x.template alloc<uint32_t>(2);
    uint32_t* odd = x.template alloc<uint32_t>((size_t)nnz + 1);
    if (x.failed()) return kOpsDevice;
    x.run(k_fill_u32, 1, tot, (int64_t)2, 0u);
    x.run(k_fill_u32, 1, odd + nnz, (int64_t)1, 0u);
    x.run(k_pair_parity, groups(nnz, kT), (const uint64_t*)skeys, (const int32_t*)a.rp, (const int32_t*)a.ci, rows,
          nnz, odd, tot);
    nonsym = x.read_u32(tot) != 0;
    symmetric_known = !nonsym;
    scan_u32(x, odd, odd, nnz + 1, tot + 1);
    const uint32_t extra = x.read_u32(tot + 1);
    if (x.failed()) return kOpsDevice;
    if (extra > 0) {
      const int64_t nt = (int64_t)extra;
      const int64_t nnz2 = nnz + nt;
      if (nnz2 > INT32_MAX) return kOpsInvalid;
      const int32_t s = rows > cols ? rows : cols;
      uint64_t* tk = x.template alloc<uint64_t>((size_t)nt);
      uint64_t* tk2 = x.template alloc<uint64_t>((size_t)nt);
      uint32_t* ts = x.template alloc<uint32_t>((size_t)nt);
      uint32_t* ts2 = x.template alloc<uint32_t>((size_t)nt);
      uint64_t* tcp = x.template alloc<uint64_t>((size_t)nt);
      int32_t* tstart = x.template alloc<int32_t>((size_t)s + 1);
      Csr<V> b;
      b.rows = s;
      b.cols = s;
      b.nnz = nnz2;
      b.rp = x.template keep<int32_t>((size_t)s + 1);
      b.ci = x.template keep<int32_t>((size_t)nnz2);
      b.vals = vals ? x.template keep<V>((size_t)nnz2) : nullptr;
      if (x.failed()) return kOpsDevice;
      x.run(k_sym_transposes, groups(nnz, kT), (const uint64_t*)skeys, (const uint32_t*)odd, nnz, tk, ts);
      radix_sort(x, tk, ts, tk2, ts2, nt, 0, bits_of((uint64_t)(s - 1)));  // by the transpose's row, stably
      x.run(k_row_starts, groups(nt, kT), (const uint64_t*)tk, nt, s, tstart);
      x.run(k_sym_tkey, groups(nt, kT), (const uint64_t*)tk, (const uint32_t*)ts, (const uint32_t*)odd, nt, tcp);
      uint32_t* len = reinterpret_cast<uint32_t*>(b.rp);
      x.run(k_sym_lengths, groups((int64_t)s + 1, kT), (const int32_t*)a.rp, rows, (const int32_t*)tstart, s, len);
      scan_u32(x, len, len, (int64_t)s + 1, (uint32_t*)nullptr);
      x.run(k_sym_place_entries<V>, groups(nnz, kT), (const uint64_t*)skeys, (const uint32_t*)odd, nnz,
            (const int32_t*)a.rp, (const uint64_t*)tcp, (const int32_t*)tstart, (const int32_t*)b.rp,
            (const V*)a.vals, b.ci, b.vals);
      x.run(k_sym_place_transposes<V>, groups(nt, kT), (const uint64_t*)tk, (const uint64_t*)tcp, (const uint32_t*)ts,
            nt, (const int32_t*)tstart, (const int32_t*)a.rp, (const int32_t*)a.ci, (const uint32_t*)odd, rows,
            (const int32_t*)b.rp, (const V*)a.vals, b.ci, b.vals);
      if (x.failed()) return kOpsDevice;
      a = b;
    }
  }
  if (a.rows != a.cols) return kOpsInvalid;
  const int32_t n = a.rows;
  if (n < 1) return kOpsInvalid;
  uint32_t* flag = x.template alloc<uint32_t>(2);
  uint32_t* root = x.template alloc<uint32_t>((size_t)n);
  if (x.failed()) return kOpsDevice;
  x.run(k_fill_u32, 1, flag, (int64_t)2, 0u);
  if (symmetric_known) {
    x.run(k_empty_rows, groups(n, kT), (const int32_t*)a.rp, n, flag);
  } else if (!nonsym) {
    x.run(k_struct_checks, groups(n, kT), (const int32_t*)a.rp, (const int32_t*)a.ci, n, flag);
    nonsym = x.read_u32(flag + 1) != 0;
  }
  if (nonsym) {
    rc = directed_roots(x, n, (const int32_t*)a.rp, (const int32_t*)a.ci, root);
    if (rc) return rc;
  } else {
    if (x.read_u32(flag)) return kOpsEmptyRow;
    rc = components(x, n, (const int32_t*)a.rp, (const int32_t*)a.ci, root);
    if (rc) return rc;
  }
  uint32_t* tot = x.template alloc<uint32_t>(1);
  int32_t* rp2 = x.template keep<int32_t>((size_t)n + 1);
  if (x.failed()) return kOpsDevice;
  uint32_t* len = reinterpret_cast<uint32_t*>(rp2);
  x.run(k_link_lengths, groups((int64_t)n + 1, kT), (const int32_t*)a.rp, (const uint32_t*)root, n, len);
  scan_u32(x, len, len, (int64_t)n + 1, tot);
  const int64_t nnz2 = (int64_t)x.read_u32(tot);
  if (x.failed()) return kOpsDevice;
  const int64_t links = (nnz2 - a.nnz) / 2;
  *links_out = links;
  if (links == 0) {
    out = a;
    return kOpsOk;
  }
  Csr<V> b;
  b.rows = b.cols = n;
  b.nnz = nnz2;
  b.rp = rp2;
  b.ci = x.template keep<int32_t>((size_t)nnz2);
  b.vals = vals ? x.template keep<V>((size_t)nnz2) : nullptr;
  if (x.failed()) return kOpsDevice;
  x.run(k_link_insert<V>, groups(n, kT), (const int32_t*)a.rp, (const int32_t*)a.ci, (const V*)a.vals,
        (const uint32_t*)root, n, (const int32_t*)rp2, b.ci, b.vals);
  out = b;
  return x.failed() ? kOpsDevice : kOpsOk;
}

}  // namespace
}  // namespace rcm_mx

// workloads.cpp -- seeded synthetic CSR inputs and host-side connectivity repair (declared in
// include/rcm_workloads.h). Host C++17; parallel where the work is large, deterministic regardless of the
// thread count (every random stream is tied to a fixed chunk or matrix index, never to a thread).
#include "rcm_workloads.h"
#include "host_canon.hpp"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <thread>
#include <vector>

namespace {

constexpr uint64_t kGolden = 0x9E3779B97F4A7C15ull;

inline uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

struct SplitMix {
  uint64_t s;
  explicit SplitMix(uint64_t seed) : s(seed) {}
  uint64_t next() {
    s += kGolden;
    return mix64(s);
  }
  double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }  // [0, 1), 53 bits
  uint64_t below(uint64_t bound) { return next() % bound; }
};

// Derived stream seed for (seed, index): independent of how work is split across threads.
inline uint64_t substream(uint64_t seed, uint64_t index) { return mix64(seed ^ mix64(index + kGolden)); }

thread_local int tl_max_threads = 0;  // > 0 inside a per-matrix worker of rcm_gen_batch_mix

int host_threads(int requested) {
  if (tl_max_threads > 0) return tl_max_threads;
  if (requested > 0) return requested;
  unsigned h = std::thread::hardware_concurrency();
  return h ? (int)std::min(h, 16u) : 4;
}

template <class F>
void parallel_for(int64_t count, int threads, F&& body) {
  if (count <= 0) return;
  threads = (int)std::min<int64_t>(threads, count);
  if (threads <= 1) {
    for (int64_t i = 0; i < count; ++i) body(i);
    return;
  }
  std::atomic<int64_t> next{0};
  const int64_t grain = std::max<int64_t>(1, count / (threads * 16));
  std::vector<std::thread> pool;
  for (int t = 0; t < threads; ++t)
    pool.emplace_back([&]() {
      for (;;) {
        int64_t b = next.fetch_add(grain);
        if (b >= count) break;
        int64_t e = std::min(count, b + grain);
        for (int64_t i = b; i < e; ++i) body(i);
      }
    });
  for (auto& th : pool) th.join();
}

bool alloc_csr(rcm_csr_t* m, int32_t n, int64_t nnz) {
  m->n = n;
  m->nnz = nnz;
  m->row_ptr = (int32_t*)std::malloc(((size_t)n + 1) * sizeof(int32_t));
  m->col_idx = (int32_t*)std::malloc((size_t)std::max<int64_t>(nnz, 1) * sizeof(int32_t));
  if (!m->row_ptr || !m->col_idx) {
    std::free(m->row_ptr);
    std::free(m->col_idx);
    m->row_ptr = m->col_idx = nullptr;
    return false;
  }
  return true;
}

// Build a CSR from per-row degree counts and a filler that writes row r's columns; sorts every row and,
// with dedupe, drops repeated columns (then compacts).
template <class Fill>
int32_t csr_from_rows(int32_t n, const std::vector<int64_t>& deg, bool dedupe, int threads, Fill&& fill,
                      rcm_csr_t* out) {
  std::vector<int64_t> off((size_t)n + 1, 0);
  for (int32_t i = 0; i < n; ++i) off[i + 1] = off[i] + deg[i];
  if (off[n] > INT32_MAX) return -1;
  std::vector<int32_t> cols((size_t)std::max<int64_t>(off[n], 1));
  std::vector<int64_t> keep((size_t)n, 0);
  parallel_for(n, threads, [&](int64_t r) {
    int32_t* c = cols.data() + off[r];
    int64_t len = off[r + 1] - off[r];
    fill((int32_t)r, c);
    std::sort(c, c + len);
    if (dedupe) len = std::unique(c, c + len) - c;
    keep[r] = len;
  });
  int64_t total = 0;
  for (int32_t i = 0; i < n; ++i) total += keep[i];
  if (!alloc_csr(out, n, total)) return -1;
  out->row_ptr[0] = 0;
  for (int32_t i = 0; i < n; ++i) out->row_ptr[i + 1] = out->row_ptr[i] + (int32_t)keep[i];
  parallel_for(n, threads, [&](int64_t r) {
    std::memcpy(out->col_idx + out->row_ptr[r], cols.data() + off[r], (size_t)keep[r] * sizeof(int32_t));
  });
  return 0;
}

}  // namespace

namespace pbrcm_host {

// The edges Graph<T>::prepareRCM adds to connect the matrix (include/Graph.h:160-235): a BFS from the first
// non-empty row; then, while nodes remain unreached, the lowest unreached node `low` is linked to low - 1
// (or, for low == 0, to the lowest reached node) and the BFS continues from `low`. Links are returned as
// (first, second) in the order the reference records them. Returns 0, or -3 where the reference's loop
// would not terminate (the lowest node of a component has an empty row).
int32_t connect_links(int32_t n, const int32_t* rp, const int32_t* ci,
                      std::vector<std::pair<int32_t, int32_t>>& links) {
  links.clear();
  // first non-empty row seeds the search (Graph.h:171-176)
  int32_t seed_row = -1;
  for (int32_t r = 0; r < n && seed_row < 0; ++r)
    if (rp[r + 1] > rp[r]) seed_row = r;
  if (seed_row < 0) return -3;
  std::vector<uint8_t> seen((size_t)n, 0);
  std::vector<int32_t> fifo;
  fifo.reserve((size_t)n);
  size_t head = 0;
  int64_t reached = 1;
  seen[seed_row] = 1;
  fifo.push_back(seed_row);
  auto drain = [&]() {
    while (head < fifo.size()) {
      const int32_t v = fifo[head++];
      for (int32_t t = rp[v]; t < rp[v + 1]; ++t) {
        const int32_t c = ci[t];
        if (!seen[c]) {
          seen[c] = 1;
          fifo.push_back(c);
          ++reached;
        }
      }
    }
  };
  int32_t low = 0;
  drain();
  while (reached != n) {
    while (seen[low]) ++low;  // lowest unreached node; everything below it is reached
    if (!links.empty() && links.back().second == low && low > 0) return -3;  // the link did not reach it
    if (low > 0)
      links.emplace_back(low - 1, low);
    else {
      int32_t j = 1;
      while (j < n && !seen[j]) ++j;
      if (!links.empty() && links.back().first == 0 && links.back().second == j) return -3;
      links.emplace_back(0, j);
    }
    fifo.push_back(low);  // queued without being marked (Graph.h:214)
    drain();
  }
  return 0;
}

}  // namespace pbrcm_host

extern "C" {

uint64_t rcm_splitmix64(uint64_t* state) {
  *state += kGolden;
  return mix64(*state);
}

void rcm_free(void* p) { std::free(p); }

void rcm_csr_free(rcm_csr_t* m) {
  if (!m) return;
  std::free(m->row_ptr);
  std::free(m->col_idx);
  m->row_ptr = m->col_idx = nullptr;
  m->n = 0;
  m->nnz = 0;
}

int32_t rcm_gen_stencil2d(int32_t nx, int32_t ny, rcm_csr_t* out) {
  if (nx <= 0 || ny <= 0 || (int64_t)nx * ny > INT32_MAX) return -1;
  const int32_t n = nx * ny;
  const int64_t nnz = 5ll * n - 2ll * nx - 2ll * ny;
  if (!alloc_csr(out, n, nnz)) return -1;
  int64_t k = 0;
  for (int32_t y = 0; y < ny; ++y)
    for (int32_t x = 0; x < nx; ++x) {
      const int32_t id = y * nx + x;
      out->row_ptr[id] = (int32_t)k;
      if (y > 0) out->col_idx[k++] = id - nx;
      if (x > 0) out->col_idx[k++] = id - 1;
      out->col_idx[k++] = id;
      if (x + 1 < nx) out->col_idx[k++] = id + 1;
      if (y + 1 < ny) out->col_idx[k++] = id + nx;
    }
  out->row_ptr[n] = (int32_t)k;
  return 0;
}

int32_t rcm_gen_stencil3d(int32_t nx, int32_t ny, int32_t nz, rcm_csr_t* out) {
  if (nx <= 0 || ny <= 0 || nz <= 0 || (int64_t)nx * ny * nz > INT32_MAX / 7) return -1;
  const int32_t n = nx * ny * nz;
  const int64_t nnz = 7ll * n - 2ll * ny * nz - 2ll * nx * nz - 2ll * nx * ny;
  if (!alloc_csr(out, n, nnz)) return -1;
  const int32_t sxy = nx * ny;
  int64_t k = 0;
  for (int32_t z = 0; z < nz; ++z)
    for (int32_t y = 0; y < ny; ++y)
      for (int32_t x = 0; x < nx; ++x) {
        const int32_t id = (z * ny + y) * nx + x;
        out->row_ptr[id] = (int32_t)k;
        if (z > 0) out->col_idx[k++] = id - sxy;
        if (y > 0) out->col_idx[k++] = id - nx;
        if (x > 0) out->col_idx[k++] = id - 1;
        out->col_idx[k++] = id;
        if (x + 1 < nx) out->col_idx[k++] = id + 1;
        if (y + 1 < ny) out->col_idx[k++] = id + nx;
        if (z + 1 < nz) out->col_idx[k++] = id + sxy;
      }
  out->row_ptr[n] = (int32_t)k;
  return 0;
}

int32_t rcm_gen_rgg(int32_t npts, double mean_degree, uint64_t seed, rcm_csr_t* out) {
  if (npts <= 0 || !(mean_degree >= 0.0)) return -1;
  const int threads = host_threads(0);
  std::vector<double> px((size_t)npts), py((size_t)npts);
  {
    SplitMix rng(seed);
    for (int32_t i = 0; i < npts; ++i) {
      px[i] = rng.uniform();
      py[i] = rng.uniform();
    }
  }
  const double r = std::sqrt(mean_degree / (M_PI * (double)npts));
  const double r2 = r * r;
  int32_t g = r > 0.0 ? (int32_t)std::floor(1.0 / r) : 1;
  g = std::max(1, std::min(g, 1 << 14));
  auto cell_of = [&](int32_t i) {
    int32_t cx = std::min(g - 1, (int32_t)(px[i] * g));
    int32_t cy = std::min(g - 1, (int32_t)(py[i] * g));
    return (int64_t)cy * g + cx;
  };
  const int64_t ncell = (int64_t)g * g;
  std::vector<int64_t> cstart((size_t)ncell + 1, 0);
  std::vector<int32_t> cpts((size_t)npts);
  for (int32_t i = 0; i < npts; ++i) cstart[cell_of(i) + 1]++;
  for (int64_t c = 0; c < ncell; ++c) cstart[c + 1] += cstart[c];
  {
    std::vector<int64_t> fill(cstart.begin(), cstart.end() - 1);
    for (int32_t i = 0; i < npts; ++i) cpts[fill[cell_of(i)]++] = i;
  }
  auto visit = [&](int32_t i, auto&& sink) {
    const int32_t cx = std::min(g - 1, (int32_t)(px[i] * g));
    const int32_t cy = std::min(g - 1, (int32_t)(py[i] * g));
    sink(i);  // self-loop
    for (int32_t yy = std::max(0, cy - 1); yy <= std::min(g - 1, cy + 1); ++yy)
      for (int32_t xx = std::max(0, cx - 1); xx <= std::min(g - 1, cx + 1); ++xx) {
        const int64_t c = (int64_t)yy * g + xx;
        for (int64_t t = cstart[c]; t < cstart[c + 1]; ++t) {
          const int32_t j = cpts[t];
          if (j == i) continue;
          const double dx = px[i] - px[j], dy = py[i] - py[j];
          if (dx * dx + dy * dy < r2) sink(j);
        }
      }
  };
  std::vector<int64_t> deg((size_t)npts, 0);
  parallel_for(npts, threads, [&](int64_t i) {
    int64_t d = 0;
    visit((int32_t)i, [&](int32_t) { ++d; });
    deg[i] = d;
  });
  return csr_from_rows(npts, deg, false, threads,
                       [&](int32_t i, int32_t* c) {
                         int64_t k = 0;
                         visit(i, [&](int32_t j) { c[k++] = j; });
                       },
                       out);
}

int32_t rcm_gen_rmat(int32_t scale, int32_t edge_factor, double a, double b, double c, uint64_t seed,
                     rcm_csr_t* out) {
  if (scale <= 0 || scale > 30 || edge_factor <= 0 || a < 0 || b < 0 || c < 0 || a + b + c > 1.0) return -1;
  const int32_t n = 1 << scale;
  const int64_t m = (int64_t)edge_factor << scale;
  const int threads = host_threads(0);
  const int64_t chunk = 1 << 16;
  const int64_t nchunks = (m + chunk - 1) / chunk;
  std::vector<int32_t> eu((size_t)m), ev((size_t)m);
  const double ab = a + b, abc = a + b + c;
  parallel_for(nchunks, threads, [&](int64_t q) {
    SplitMix rng(substream(seed, (uint64_t)q));
    const int64_t e0 = q * chunk, e1 = std::min(m, e0 + chunk);
    for (int64_t e = e0; e < e1; ++e) {
      uint32_t u = 0, v = 0;
      for (int32_t l = 0; l < scale; ++l) {
        const double p = rng.uniform();
        const uint32_t bu = p >= ab ? 1u : 0u;                             // quadrants c, d: lower half
        const uint32_t bv = (p >= a && p < ab) || p >= abc ? 1u : 0u;      // quadrants b, d: right half
        u = (u << 1) | bu;
        v = (v << 1) | bv;
      }
      eu[e] = (int32_t)u;
      ev[e] = (int32_t)v;
    }
  });
  // symmetrise: bucket both directions by row, then sort + dedupe per row, plus the diagonal
  std::vector<int64_t> deg((size_t)n, 1);
  for (int64_t e = 0; e < m; ++e) {
    deg[eu[e]]++;
    deg[ev[e]]++;
  }
  std::vector<int64_t> off((size_t)n + 1, 0);
  for (int32_t i = 0; i < n; ++i) off[i + 1] = off[i] + deg[i];
  std::vector<int32_t> bucket((size_t)off[n]);
  {
    std::vector<int64_t> fill(off.begin(), off.end() - 1);
    for (int32_t i = 0; i < n; ++i) bucket[fill[i]++] = i;
    for (int64_t e = 0; e < m; ++e) {
      bucket[fill[eu[e]]++] = ev[e];
      bucket[fill[ev[e]]++] = eu[e];
    }
  }
  return csr_from_rows(n, deg, true, threads,
                       [&](int32_t i, int32_t* cc) {
                         std::memcpy(cc, bucket.data() + off[i], (size_t)deg[i] * sizeof(int32_t));
                       },
                       out);
}

int32_t rcm_permute_symmetric(const rcm_csr_t* in, uint64_t seed, rcm_csr_t* out, int32_t* perm_out) {
  if (!in || in->n <= 0) return -1;
  const int32_t n = in->n;
  std::vector<int32_t> q((size_t)n);
  for (int32_t i = 0; i < n; ++i) q[i] = i;
  SplitMix rng(seed);
  for (int32_t i = n - 1; i > 0; --i) {  // Fisher-Yates
    const int32_t j = (int32_t)rng.below((uint64_t)i + 1);
    std::swap(q[i], q[j]);
  }
  std::vector<int32_t> src((size_t)n);  // src[new] = old
  for (int32_t i = 0; i < n; ++i) src[q[i]] = i;
  std::vector<int64_t> deg((size_t)n);
  for (int32_t v = 0; v < n; ++v) deg[v] = in->row_ptr[src[v] + 1] - in->row_ptr[src[v]];
  if (perm_out) std::memcpy(perm_out, q.data(), (size_t)n * sizeof(int32_t));
  return csr_from_rows(n, deg, false, host_threads(0),
                       [&](int32_t v, int32_t* c) {
                         const int32_t o = src[v];
                         int64_t k = 0;
                         for (int32_t t = in->row_ptr[o]; t < in->row_ptr[o + 1]; ++t) c[k++] = q[in->col_idx[t]];
                       },
                       out);
}

int64_t rcm_connect_components(const rcm_csr_t* in, rcm_csr_t* out) {
  if (!in || in->n <= 0) return -1;
  const int32_t n = in->n;
  const int32_t* rp = in->row_ptr;
  const int32_t* ci = in->col_idx;
  std::vector<std::pair<int32_t, int32_t>> links;
  const int32_t lrc = pbrcm_host::connect_links(n, rp, ci, links);
  if (lrc < 0) return lrc;
  if (links.empty()) {
    if (!alloc_csr(out, n, in->nnz)) return -1;
    std::memcpy(out->row_ptr, rp, ((size_t)n + 1) * sizeof(int32_t));
    std::memcpy(out->col_idx, ci, (size_t)in->nnz * sizeof(int32_t));
    return 0;
  }
  std::vector<std::vector<int32_t>> extra((size_t)n);
  for (auto& l : links) {
    extra[l.second].push_back(l.first);
    extra[l.first].push_back(l.second);
  }
  std::vector<int64_t> deg((size_t)n);
  for (int32_t v = 0; v < n; ++v) deg[v] = (int64_t)(rp[v + 1] - rp[v]) + (int64_t)extra[v].size();
  const int32_t rc = csr_from_rows(n, deg, false, host_threads(0),
                                   [&](int32_t v, int32_t* c) {
                                     int64_t k = 0;
                                     for (int32_t t = rp[v]; t < rp[v + 1]; ++t) c[k++] = ci[t];
                                     for (int32_t x : extra[v]) c[k++] = x;
                                   },
                                   out);
  return rc < 0 ? rc : (int64_t)links.size();
}

int64_t rcm_symmetrize(const rcm_csr_t* in, rcm_csr_t* out) {
  if (!in || in->n <= 0) return -1;
  const int32_t n = in->n;
  const int32_t* rp = in->row_ptr;
  const int32_t* ci = in->col_idx;
  auto pair_key = [](uint32_t r, uint32_t c) {
    return r < c ? ((uint64_t)r << 32) | c : ((uint64_t)c << 32) | r;
  };
  std::vector<uint64_t> keys;
  keys.reserve((size_t)in->nnz);
  for (int32_t r = 0; r < n; ++r)
    for (int32_t t = rp[r]; t < rp[r + 1]; ++t)
      if (ci[t] != r) keys.push_back(pair_key((uint32_t)r, (uint32_t)ci[t]));
  std::sort(keys.begin(), keys.end());
  std::vector<uint64_t> odd;
  for (size_t i = 0; i < keys.size();) {
    size_t j = i;
    while (j < keys.size() && keys[j] == keys[i]) ++j;
    if ((j - i) % 2 == 1) odd.push_back(keys[i]);
    i = j;
  }
  if (odd.empty()) {
    if (!alloc_csr(out, n, in->nnz)) return -1;
    std::memcpy(out->row_ptr, rp, ((size_t)n + 1) * sizeof(int32_t));
    std::memcpy(out->col_idx, ci, (size_t)in->nnz * sizeof(int32_t));
    return 0;
  }
  auto is_odd = [&](int32_t r, int32_t c) {
    return std::binary_search(odd.begin(), odd.end(), pair_key((uint32_t)r, (uint32_t)c));
  };
  std::vector<int64_t> deg((size_t)n, 0);
  for (int32_t r = 0; r < n; ++r)
    for (int32_t t = rp[r]; t < rp[r + 1]; ++t) {
      deg[r]++;
      if (ci[t] != r && is_odd(r, ci[t])) deg[ci[t]]++;
    }
  std::vector<int64_t> off((size_t)n + 1, 0);
  for (int32_t i = 0; i < n; ++i) off[i + 1] = off[i] + deg[i];
  std::vector<int32_t> bucket((size_t)off[n]);
  {
    std::vector<int64_t> fill(off.begin(), off.end() - 1);
    for (int32_t r = 0; r < n; ++r)
      for (int32_t t = rp[r]; t < rp[r + 1]; ++t) {
        bucket[fill[r]++] = ci[t];
        if (ci[t] != r && is_odd(r, ci[t])) bucket[fill[ci[t]]++] = r;
      }
  }
  const int32_t rc = csr_from_rows(n, deg, false, host_threads(0),
                                   [&](int32_t i, int32_t* c) {
                                     std::memcpy(c, bucket.data() + off[i], (size_t)deg[i] * sizeof(int32_t));
                                   },
                                   out);
  return rc < 0 ? rc : (int64_t)odd.size();
}

int32_t rcm_gen_config(int32_t which, int64_t scale, uint64_t seed, rcm_csr_t* out) {
  rcm_csr_t base{}, perm{};
  int32_t rc = -1;
  if (which == 1) {
    const int32_t side = scale > 0 ? (int32_t)scale : 100;
    rc = rcm_gen_stencil2d(side, side, &base);
  } else if (which == 2) {
    const int32_t side = scale > 0 ? (int32_t)scale : 128;
    rc = rcm_gen_stencil3d(side, side, side, &base);
  } else if (which == 3) {
    const int32_t npts = scale > 0 ? (int32_t)scale : 1000000;
    rc = rcm_gen_rgg(npts, 12.0, substream(seed, 1), &base);
  }
  if (rc) return rc;
  rc = rcm_permute_symmetric(&base, substream(seed, 2), &perm, nullptr);
  rcm_csr_free(&base);
  if (rc) return rc;
  const int64_t added = rcm_connect_components(&perm, out);
  rcm_csr_free(&perm);
  return added < 0 ? (int32_t)added : 0;
}

int32_t rcm_gen_batch_mix(int32_t count, int32_t n_lo, int32_t n_hi, uint64_t seed, int32_t threads,
                          int32_t** ns, int64_t** nnzs, int64_t** rp_off, int64_t** ci_off, int32_t** row_ptr_cat,
                          int32_t** col_idx_cat) {
  if (count <= 0 || n_lo <= 0 || n_hi < n_lo) return -1;
  std::vector<rcm_csr_t> mats((size_t)count);
  std::atomic<int32_t> failed{0};
  const double llo = std::log((double)n_lo), lhi = std::log((double)n_hi);
  // one thread per matrix; the generators' own parallel loops run single-threaded inside
  parallel_for(count, host_threads(threads), [&](int64_t k) {
    tl_max_threads = 1;
    SplitMix rng(substream(seed, (uint64_t)k));
    const int64_t n = std::max<int64_t>(1, std::llround(std::exp(llo + rng.uniform() * (lhi - llo))));
    const bool coin = rng.uniform() < 0.5;
    const double deg = 6.0 + 8.0 * rng.uniform();
    const uint64_t gseed = rng.next(), pseed = rng.next();
    rcm_csr_t base{}, perm{};
    int32_t rc;
    if ((k & 1) == 0) {
      if (coin) {
        const int32_t s = std::max<int32_t>(1, (int32_t)std::llround(std::sqrt((double)n)));
        rc = rcm_gen_stencil2d(s, s, &base);
      } else {
        const int32_t s = std::max<int32_t>(1, (int32_t)std::llround(std::cbrt((double)n)));
        rc = rcm_gen_stencil3d(s, s, s, &base);
      }
    } else {
      rc = rcm_gen_rgg((int32_t)n, deg, gseed, &base);
    }
    if (rc == 0) rc = rcm_permute_symmetric(&base, pseed, &perm, nullptr);
    rcm_csr_free(&base);
    if (rc == 0 && rcm_connect_components(&perm, &mats[k]) < 0) rc = -1;
    rcm_csr_free(&perm);
    if (rc) failed = 1;
  });
  if (failed) {
    for (auto& m : mats) rcm_csr_free(&m);
    return -1;
  }
  int64_t trp = 0, tci = 0;
  *ns = (int32_t*)std::malloc((size_t)count * sizeof(int32_t));
  *nnzs = (int64_t*)std::malloc((size_t)count * sizeof(int64_t));
  *rp_off = (int64_t*)std::malloc((size_t)count * sizeof(int64_t));
  *ci_off = (int64_t*)std::malloc((size_t)count * sizeof(int64_t));
  for (int32_t k = 0; k < count; ++k) {
    (*ns)[k] = mats[k].n;
    (*nnzs)[k] = mats[k].nnz;
    (*rp_off)[k] = trp;
    (*ci_off)[k] = tci;
    trp += mats[k].n + 1;
    tci += mats[k].nnz;
  }
  *row_ptr_cat = (int32_t*)std::malloc((size_t)trp * sizeof(int32_t));
  *col_idx_cat = (int32_t*)std::malloc((size_t)std::max<int64_t>(tci, 1) * sizeof(int32_t));
  if (!*row_ptr_cat || !*col_idx_cat) return -1;
  parallel_for(count, host_threads(threads), [&](int64_t k) {
    std::memcpy(*row_ptr_cat + (*rp_off)[k], mats[k].row_ptr, ((size_t)mats[k].n + 1) * sizeof(int32_t));
    std::memcpy(*col_idx_cat + (*ci_off)[k], mats[k].col_idx, (size_t)mats[k].nnz * sizeof(int32_t));
    rcm_csr_free(&mats[k]);
  });
  return 0;
}

}  // extern "C"

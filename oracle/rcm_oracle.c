/*
 * rcm_oracle.c -- CPU restatement of the reference's serial RCM-path algorithms.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP engine. Loaded by tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg, never by the product path. Plain C99, single-threaded.
 *
 * File:line citations refer to the reference repository (GPUPeople/ParallelBatchRCM).
 */
#include "rcm_oracle.h"

#include <stdlib.h>
#include <string.h>
#include <limits.h>

void orc_free(void* p) { free(p); }

/* ---------------------------------------------------------------------------------------------------- */
/* Stable sort of (key, value) pairs by key, ties kept in input order.                                    */
/* Restates std::stable_sort over RealCPU::KeyValue, whose operator< compares key only (rcm_cpu.cpp:47-56). */
/* ---------------------------------------------------------------------------------------------------- */
typedef struct { uint32_t key; int32_t val; } kv_t;

static void kv_merge_sort(kv_t* a, kv_t* tmp, int64_t len) {
  if (len <= 24) { /* insertion sort: stable because it only moves strictly greater keys */
    for (int64_t i = 1; i < len; ++i) {
      kv_t x = a[i];
      int64_t j = i - 1;
      while (j >= 0 && a[j].key > x.key) { a[j + 1] = a[j]; --j; }
      a[j + 1] = x;
    }
    return;
  }
  int64_t h = len / 2;
  kv_merge_sort(a, tmp, h);
  kv_merge_sort(a + h, tmp, len - h);
  int64_t i = 0, j = h, k = 0;
  while (i < h && j < len) tmp[k++] = (a[j].key < a[i].key) ? a[j++] : a[i++]; /* left wins ties */
  while (i < h) tmp[k++] = a[i++];
  while (j < len) tmp[k++] = a[j++];
  memcpy(a, tmp, (size_t)len * sizeof(kv_t));
}

static int valid_csr(int32_t n, const int32_t* rp, const int32_t* ci) {
  return n > 0 && rp != NULL && (ci != NULL || rp[n] == 0) && rp[0] == 0;
}

/* RealCPU::rcm_seq<true> (src/rcm_cpu.cpp:57-104). The permutation array doubles as the FIFO queue
 * (std::queue in the reference; same order). visited[] = 0xFFFFFFFF / 0 / 1 as in :63-64, :88. */
int32_t orc_cm(int32_t n, const int32_t* rp, const int32_t* ci, int32_t start, int32_t* perm) {
  if (!valid_csr(n, rp, ci) || start < 0 || start >= n || perm == NULL) return -1;
  uint8_t* visited = (uint8_t*)calloc((size_t)n, 1);
  kv_t* temp = NULL;
  kv_t* tmp2 = NULL;
  int64_t cap = 0;
  if (!visited) return -1;
  visited[start] = 1;
  int32_t head = 0, tail = 0;
  perm[tail++] = start;
  while (head < tail) {
    int32_t v = perm[head++];
    int64_t b = rp[v], e = rp[v + 1];
    int64_t cnt = 0;
    if (e - b > cap) {
      cap = (e - b) * 2;
      temp = (kv_t*)realloc(temp, (size_t)cap * sizeof(kv_t));
      tmp2 = (kv_t*)realloc(tmp2, (size_t)cap * sizeof(kv_t));
    }
    for (int64_t k = b; k < e; ++k) {
      int32_t c = ci[k];
      if (!visited[c]) {
        visited[c] = 1;
        temp[cnt].key = (uint32_t)(rp[c + 1] - rp[c]); /* Graph::neighbour_count, include/Graph.h:58-61 */
        temp[cnt].val = c;
        ++cnt;
      }
    }
    if (cnt > 1) kv_merge_sort(temp, tmp2, cnt);
    for (int64_t k = 0; k < cnt; ++k) perm[tail++] = temp[k].val;
  }
  free(visited);
  free(temp);
  free(tmp2);
  return tail;
}

/* FIFO BFS in stored row order, the traversal of SimplePeripheral::findMaxDistance
 * (src/simple_peripheral.cpp:61-84); level = BFS distance from start. */
int32_t orc_bfs(int32_t n, const int32_t* rp, const int32_t* ci, int32_t start, int32_t* perm, int32_t* level) {
  if (!valid_csr(n, rp, ci) || start < 0 || start >= n || perm == NULL) return -1;
  int32_t* dist = (int32_t*)malloc((size_t)n * sizeof(int32_t));
  if (!dist) return -1;
  for (int32_t i = 0; i < n; ++i) dist[i] = -1;
  dist[start] = 0;
  int32_t head = 0, tail = 0;
  perm[tail++] = start;
  while (head < tail) {
    int32_t v = perm[head++];
    for (int64_t k = rp[v]; k < rp[v + 1]; ++k) {
      int32_t c = ci[k];
      if (dist[c] < 0) {
        dist[c] = dist[v] + 1;
        perm[tail++] = c;
      }
    }
  }
  if (level)
    for (int32_t i = 0; i < tail; ++i) level[i] = dist[perm[i]];
  free(dist);
  return tail;
}

/* SimplePeripheral::findMaxDistance (src/simple_peripheral.cpp:44-90) with outputnodes = 32: returns the
 * last min(32, rows-1) discovered nodes and their distances in discovery order; -2 if not connected. */
static int find_max_distance(int32_t n, const int32_t* rp, const int32_t* ci, int32_t start, int32_t* visited,
                             int32_t* queue, int32_t* qdist, int32_t* nodes, int32_t* dists, int32_t* nout) {
  const uint32_t outputnodes = 32;
  for (int32_t i = 0; i < n; ++i) visited[i] = -1;
  uint64_t to_visit = (uint64_t)n - 1;
  int32_t head = 0, tail = 0;
  queue[tail] = start;
  qdist[tail] = 0;
  ++tail;
  visited[start] = 0;
  *nout = 0;
  while (head < tail) {
    int32_t v = queue[head];
    int32_t d = qdist[head];
    ++head;
    int32_t nd = d + 1;
    for (int64_t k = rp[v]; k < rp[v + 1]; ++k) {
      int32_t c = ci[k];
      if (visited[c] == -1) {
        visited[c] = nd;
        queue[tail] = c;
        qdist[tail] = nd;
        ++tail;
        if (--to_visit < outputnodes) {
          nodes[*nout] = c;
          dists[*nout] = nd;
          ++*nout;
        }
      }
    }
  }
  return to_visit > 0 ? -2 : 0;
}

/* SimplePeripheral::findPeripheral (src/simple_peripheral.cpp:92-141). */
int32_t orc_peripheral(int32_t n, const int32_t* rp, const int32_t* ci, int32_t* start_out, int32_t* iters_out,
                       int32_t* ecc_out) {
  if (!valid_csr(n, rp, ci)) return -1;
  if (n == 1) {
    if (start_out) *start_out = 0;
    if (iters_out) *iters_out = 1;
    if (ecc_out) *ecc_out = 0;
    return 0;
  }
  int32_t* visited = (int32_t*)malloc((size_t)n * sizeof(int32_t));
  int32_t* queue = (int32_t*)malloc((size_t)n * sizeof(int32_t));
  int32_t* qdist = (int32_t*)malloc((size_t)n * sizeof(int32_t));
  int32_t nodes[32], dists[32], nout = 0;
  int32_t rc = 0;
  uint32_t lastdist = 0;
  int32_t cur = n / 2; /* :99 */
  int32_t iters = 0;
  if (!visited || !queue || !qdist) { rc = -1; goto done; }
  for (;;) {
    ++iters;
    rc = find_max_distance(n, rp, ci, cur, visited, queue, qdist, nodes, dists, &nout);
    if (rc) goto done;
    if ((uint32_t)dists[nout - 1] == lastdist) break; /* :109, returns the start of this BFS */
    /* first index holding the maximum distance (std::max_element), :118-119 */
    int32_t idx = 0;
    for (int32_t i = 1; i < nout; ++i)
      if (dists[i] > dists[idx]) idx = i;
    int minval = INT_MAX;
    for (int32_t i = idx; i < nout; ++i) { /* :121-127, ties go to the later node */
      int32_t node = nodes[i];
      int val = (int)(rp[node + 1] - rp[node]);
      if (val <= minval) { minval = val; cur = node; }
    }
    lastdist = (uint32_t)dists[nout - 1];
  }
  if (start_out) *start_out = cur;
  if (iters_out) *iters_out = iters;
  if (ecc_out) *ecc_out = dists[nout - 1];
done:
  free(visited);
  free(queue);
  free(qdist);
  return rc;
}

/* computeBandwidth (src/main.cpp:143-184). */
int64_t orc_bandwidth(int32_t n, const int32_t* rp, const int32_t* ci, const int32_t* perm, int32_t nperm) {
  if (nperm == 0) return -1;
  if (!valid_csr(n, rp, ci) || perm == NULL || nperm != n) return -2;
  int32_t* inv = (int32_t*)malloc((size_t)n * sizeof(int32_t));
  if (!inv) return -2;
  for (int32_t i = 0; i < n; ++i) inv[i] = -1;
  for (int32_t i = 0; i < n; ++i) {
    if (perm[i] < 0 || perm[i] >= n || inv[perm[i]] != -1) { free(inv); return -2; }
    inv[perm[i]] = i;
  }
  int64_t bw = 0;
  for (int32_t i = 0; i < n; ++i) {
    int64_t b = rp[i], e = rp[i + 1];
    if (e == b) continue; /* only n == 1 has an empty row in a connected graph */
    int64_t mn = INT64_MAX, mx = INT64_MIN;
    for (int64_t k = b; k < e; ++k) {
      int64_t v = inv[ci[k]];
      if (v < mn) mn = v;
      if (v > mx) mx = v;
    }
    int64_t rid = inv[i];
    int64_t lo = rid - mn, up = mx - rid;
    int64_t m = lo > up ? lo : up;
    if (m > bw) bw = m;
  }
  free(inv);
  return bw;
}

/* checkBandwidth (src/main.cpp:114-129). */
int64_t orc_bandwidth_identity(int32_t n, const int32_t* rp, const int32_t* ci) {
  if (!valid_csr(n, rp, ci)) return -1;
  int64_t bw = 0;
  for (int32_t i = 0; i < n; ++i) {
    if (rp[i + 1] == rp[i]) continue;
    int64_t lo = (int64_t)i - ci[rp[i]];
    int64_t up = (int64_t)ci[rp[i + 1] - 1] - i;
    int64_t m = lo > up ? lo : up;
    if (m > bw) bw = m;
  }
  return bw;
}

/* checkPermutation (src/main.cpp:186-226), as a strict bijection check. */
int32_t orc_check_perm(const int32_t* perm, int32_t n) {
  if (n < 0 || (n > 0 && perm == NULL)) return -1;
  uint8_t* seen = (uint8_t*)calloc((size_t)n + 1, 1);
  if (!seen) return -1;
  int32_t rc = 0;
  for (int32_t i = 0; i < n; ++i) {
    if (perm[i] < 0 || perm[i] >= n || seen[perm[i]]) { rc = -1; break; }
    seen[perm[i]] = 1;
  }
  free(seen);
  return rc;
}

/* Characterisation of CM / BFS order (independent of any reference function). */
int64_t orc_verify_order(int32_t n, const int32_t* rp, const int32_t* ci, int32_t start, const int32_t* perm,
                         int32_t nperm, int32_t sorted) {
  if (!valid_csr(n, rp, ci) || perm == NULL || nperm < 1 || nperm > n || start < 0 || start >= n) return -1;
  int64_t rc = 0;
  int32_t* pos = (int32_t*)malloc((size_t)n * sizeof(int32_t));
  int32_t* par = (int32_t*)malloc((size_t)n * sizeof(int32_t));
  int64_t* rix = (int64_t*)malloc((size_t)n * sizeof(int64_t));
  if (!pos || !par || !rix) { rc = -1; goto out; }
  for (int32_t i = 0; i < n; ++i) { pos[i] = -1; par[i] = -1; rix[i] = -1; }
  if (perm[0] != start) { rc = 1; goto out; }
  for (int32_t i = 0; i < nperm; ++i) {
    int32_t x = perm[i];
    if (x < 0 || x >= n || pos[x] != -1) { rc = (int64_t)i + 1; goto out; }
    pos[x] = i;
  }
  /* closure: every row entry of a placed node is placed (perm covers the start's component) */
  for (int32_t i = 0; i < nperm; ++i) {
    int32_t y = perm[i];
    for (int64_t k = rp[y]; k < rp[y + 1]; ++k)
      if (pos[ci[k]] < 0) { rc = (int64_t)i + 1; goto out; }
  }
  /* P(x): the earliest-placed y != x whose row holds x; r(x): first index of x in that row */
  for (int32_t i = 0; i < nperm; ++i) {
    int32_t y = perm[i];
    for (int64_t k = rp[y]; k < rp[y + 1]; ++k) {
      int32_t x = ci[k];
      if (x == y || par[x] != -1) continue;
      par[x] = i;
      rix[x] = k - rp[y];
    }
  }
  for (int32_t i = 1; i < nperm; ++i) {
    int32_t x = perm[i];
    if (par[x] < 0 || par[x] >= i) { rc = (int64_t)i + 1; goto out; }
  }
  for (int32_t i = 1; i + 1 < nperm; ++i) {
    int32_t a = perm[i], b = perm[i + 1];
    int64_t ka0 = par[a], kb0 = par[b];
    int64_t ka1 = sorted ? (int64_t)(rp[a + 1] - rp[a]) : 0, kb1 = sorted ? (int64_t)(rp[b + 1] - rp[b]) : 0;
    int64_t ka2 = rix[a], kb2 = rix[b];
    int less = (ka0 < kb0) || (ka0 == kb0 && (ka1 < kb1 || (ka1 == kb1 && ka2 < kb2)));
    if (!less) { rc = (int64_t)i + 2; goto out; }
  }
out:
  free(pos);
  free(par);
  free(rix);
  return rc;
}

/* ---------------------------------------------------------------------------------------------------- */
/* prepareRCM restatement                                                                                */
/* ---------------------------------------------------------------------------------------------------- */
static int cmp_u64(const void* a, const void* b) {
  uint64_t x = *(const uint64_t*)a, y = *(const uint64_t*)b;
  return (x > y) - (x < y);
}
static int cmp_i32(const void* a, const void* b) {
  int32_t x = *(const int32_t*)a, y = *(const int32_t*)b;
  return (x > y) - (x < y);
}

/* (column, input position) of one entry: sorting by both ranks equal columns by input position. */
typedef struct {
  int32_t col;
  int64_t pos;
} ColPos;
static int cmp_colpos(const void* a, const void* b) {
  const ColPos* x = (const ColPos*)a;
  const ColPos* y = (const ColPos*)b;
  if (x->col != y->col) return x->col < y->col ? -1 : 1;
  return (x->pos > y->pos) - (x->pos < y->pos);
}

/* convert (src/CSR.cpp:170-208): sort entries by (row, col), duplicates kept. Entries are given as parallel
 * row/col arrays; the column structure of the result does not depend on std::sort's tie order. With values
 * (vals != NULL), duplicates keep their input order and *vals_out receives the values in the same order. */
static int build_sorted_csr(int32_t n, int64_t m, const int32_t* rows, const int32_t* cols, const double* vals,
                            int32_t** rp_out, int32_t** ci_out, double** vals_out) {
  int32_t* rp = (int32_t*)calloc((size_t)n + 1, sizeof(int32_t));
  int32_t* ci = (int32_t*)malloc((size_t)(m > 0 ? m : 1) * sizeof(int32_t));
  int64_t* src = (int64_t*)malloc((size_t)(m > 0 ? m : 1) * sizeof(int64_t));
  int32_t* fill = (int32_t*)malloc(((size_t)n + 1) * sizeof(int32_t));
  double* vo = vals ? (double*)malloc((size_t)(m > 0 ? m : 1) * sizeof(double)) : NULL;
  if (!rp || !ci || !src || !fill || (vals && !vo)) { free(rp); free(ci); free(src); free(fill); free(vo); return -1; }
  for (int64_t k = 0; k < m; ++k) rp[rows[k] + 1]++;
  for (int32_t i = 0; i < n; ++i) rp[i + 1] += rp[i];
  for (int32_t i = 0; i < n; ++i) fill[i] = rp[i];
  for (int64_t k = 0; k < m; ++k) {
    const int32_t at = fill[rows[k]]++;
    ci[at] = cols[k];
    src[at] = k; /* row buckets keep input order */
  }
  if (!vals) {
    for (int32_t i = 0; i < n; ++i)
      if (rp[i + 1] - rp[i] > 1) qsort(ci + rp[i], (size_t)(rp[i + 1] - rp[i]), sizeof(int32_t), cmp_i32);
  } else {
    int32_t longest = 0;
    for (int32_t i = 0; i < n; ++i) longest = rp[i + 1] - rp[i] > longest ? rp[i + 1] - rp[i] : longest;
    ColPos* tmp = (ColPos*)malloc((size_t)(longest > 0 ? longest : 1) * sizeof(ColPos));
    if (!tmp) { free(rp); free(ci); free(src); free(fill); free(vo); return -1; }
    for (int32_t i = 0; i < n; ++i) {
      const int32_t b = rp[i], d = rp[i + 1] - rp[i];
      for (int32_t t = 0; t < d; ++t) { tmp[t].col = ci[b + t]; tmp[t].pos = src[b + t]; }
      qsort(tmp, (size_t)d, sizeof(ColPos), cmp_colpos);
      for (int32_t t = 0; t < d; ++t) { ci[b + t] = tmp[t].col; vo[b + t] = vals[tmp[t].pos]; }
    }
    free(tmp);
  }
  free(src);
  free(fill);
  *rp_out = rp;
  *ci_out = ci;
  if (vals_out) *vals_out = vo;
  return 0;
}

int64_t orc_convert(int32_t rows, int64_t m, const int32_t* row_ids, const int32_t* col_ids, const double* vals,
                    int32_t** rp_out, int32_t** ci_out, double** vals_out) {
  if (rows < 1 || m < 0 || !row_ids || !col_ids || !rp_out || !ci_out) return -1;
  for (int64_t k = 0; k < m; ++k)
    if (row_ids[k] < 0 || row_ids[k] >= rows) return -1;
  if (build_sorted_csr(rows, m, row_ids, col_ids, vals, rp_out, ci_out, vals ? vals_out : NULL)) return -1;
  return m;
}

int64_t orc_prepare(int32_t n, int64_t nnz, const int32_t* rp0, const int32_t* ci0, int32_t check_sym,
                    int32_t** rp_out, int32_t** ci_out, int64_t* nnz_out) {
  return orc_prepare_values(n, nnz, rp0, ci0, NULL, check_sym, rp_out, ci_out, NULL, nnz_out);
}

int64_t orc_prepare_values(int32_t n, int64_t nnz, const int32_t* rp0, const int32_t* ci0, const double* v0,
                           int32_t check_sym, int32_t** rp_out, int32_t** ci_out, double** vals_out,
                           int64_t* nnz_out) {
  if (!valid_csr(n, rp0, ci0) || rp0[n] != nnz || !rp_out || !ci_out || !nnz_out || (v0 && !vals_out)) return -1;
  for (int64_t k = 0; k < nnz; ++k)
    if (ci0[k] < 0 || ci0[k] >= n) return -1;
  int32_t* rp = (int32_t*)malloc(((size_t)n + 1) * sizeof(int32_t));
  int32_t* ci = (int32_t*)malloc((size_t)(nnz > 0 ? nnz : 1) * sizeof(int32_t));
  double* va = v0 ? (double*)malloc((size_t)(nnz > 0 ? nnz : 1) * sizeof(double)) : NULL;
  int64_t m = nnz;
  memcpy(rp, rp0, ((size_t)n + 1) * sizeof(int32_t));
  memcpy(ci, ci0, (size_t)nnz * sizeof(int32_t));
  if (va) memcpy(va, v0, (size_t)nnz * sizeof(double));

  /* (1) symmetrise (include/Graph.h:77-157): pairs with an odd number of stored off-diagonal occurrences
   * stay in `entries` after the toggle pass; every stored occurrence of such a pair gets its transpose. */
  if (check_sym) {
    uint64_t* keys = (uint64_t*)malloc((size_t)(m > 0 ? m : 1) * sizeof(uint64_t));
    int64_t nk = 0;
    for (int32_t r = 0; r < n; ++r)
      for (int64_t k = rp[r]; k < rp[r + 1]; ++k) {
        uint32_t c = (uint32_t)ci[k];
        if ((uint32_t)r == c) continue;
        uint32_t lo = (uint32_t)r < c ? (uint32_t)r : c, hi = (uint32_t)r < c ? c : (uint32_t)r;
        keys[nk++] = ((uint64_t)lo << 32) | hi;
      }
    qsort(keys, (size_t)nk, sizeof(uint64_t), cmp_u64);
    uint64_t* odd = (uint64_t*)malloc((size_t)(nk > 0 ? nk : 1) * sizeof(uint64_t));
    int64_t nodd = 0;
    for (int64_t a = 0; a < nk;) {
      int64_t b = a;
      while (b < nk && keys[b] == keys[a]) ++b;
      if ((b - a) & 1) odd[nodd++] = keys[a];
      a = b;
    }
    free(keys);
    if (nodd > 0) {
      int64_t cap = m * 2 + 1;
      int32_t* rows = (int32_t*)malloc((size_t)cap * sizeof(int32_t));
      int32_t* cols = (int32_t*)malloc((size_t)cap * sizeof(int32_t));
      double* vals = va ? (double*)malloc((size_t)cap * sizeof(double)) : NULL;
      int64_t t = 0;
      for (int32_t r = 0; r < n; ++r)
        for (int64_t k = rp[r]; k < rp[r + 1]; ++k) {
          int32_t c = ci[k];
          rows[t] = r;
          cols[t] = c;
          if (vals) vals[t] = va[k];
          ++t;
          if (r == c) continue;
          uint32_t lo = (uint32_t)r < (uint32_t)c ? (uint32_t)r : (uint32_t)c;
          uint32_t hi = (uint32_t)r < (uint32_t)c ? (uint32_t)c : (uint32_t)r;
          uint64_t key = ((uint64_t)lo << 32) | hi;
          if (bsearch(&key, odd, (size_t)nodd, sizeof(uint64_t), cmp_u64)) {
            rows[t] = c; /* the transpose right after the entry, with its value (Graph.h:140-146) */
            cols[t] = r;
            if (vals) vals[t] = va[k];
            ++t;
          }
        }
      free(rp);
      free(ci);
      free(va);
      va = NULL;
      if (build_sorted_csr(n, t, rows, cols, vals, &rp, &ci, vals ? &va : NULL)) {
        free(rows); free(cols); free(vals); free(odd); return -1;
      }
      m = t;
      free(rows);
      free(cols);
      free(vals);
    }
    free(odd);
  }

  /* (2) connect components (include/Graph.h:160-263). */
  int64_t added = 0;
  int32_t* visited = (int32_t*)calloc((size_t)n, sizeof(int32_t));
  int32_t* queue = (int32_t*)malloc(((size_t)n * 2 + 2) * sizeof(int32_t)); /* <= n marked + n unmarked pushes */
  int32_t* con_a = NULL;
  int32_t* con_b = NULL;
  int64_t con_cap = 0;
  int32_t minn = n;
  for (int32_t r = 0; r < n; ++r)
    if (rp[r + 1] != rp[r]) { minn = r; break; }
  if (minn == n) { /* every row empty: the reference indexes visited[rows] */
    free(visited); free(queue); free(rp); free(ci); free(va);
    return -3;
  }
  int64_t head = 0, tail = 0;
  queue[tail++] = minn;
  visited[minn] = 1;
  int64_t num_visited = 1;
  int32_t cursor = 0; /* lowest possibly-unvisited node; visited[] only grows */
  int32_t last_i = -1;
  while (num_visited != n) {
    while (head < tail) {
      int32_t qn = queue[head++];
      for (int64_t k = rp[qn]; k < rp[qn + 1]; ++k) {
        int32_t c = ci[k];
        if (!visited[c]) {
          visited[c] = 1;
          queue[tail++] = c;
          ++num_visited;
        }
      }
    }
    if (num_visited == n) break;
    while (cursor < n && visited[cursor]) ++cursor;
    int32_t i = cursor;
    if (i == last_i) {
      /* the previous connection did not make i visited: the reference re-adds it forever */
      free(visited); free(queue); free(con_a); free(con_b); free(rp); free(ci); free(va);
      return -3;
    }
    int32_t a, b;
    if (i > 0) { a = i - 1; b = i; } /* all nodes below the lowest unvisited one are visited */
    else {
      int32_t j = 1;
      while (j < n && !visited[j]) ++j;
      a = 0; b = j;
    }
    if (added == con_cap) {
      con_cap = con_cap ? con_cap * 2 : 64;
      con_a = (int32_t*)realloc(con_a, (size_t)con_cap * sizeof(int32_t));
      con_b = (int32_t*)realloc(con_b, (size_t)con_cap * sizeof(int32_t));
    }
    con_a[added] = a; /* addconnections.first */
    con_b[added] = b; /* addconnections.second */
    ++added;
    queue[tail++] = i; /* pushed without being marked visited (:214) */
    last_i = i;
  }
  free(visited);
  free(queue);
  if (added > 0) {
    int64_t t = m + 2 * added;
    int32_t* rows = (int32_t*)malloc((size_t)t * sizeof(int32_t));
    int32_t* cols = (int32_t*)malloc((size_t)t * sizeof(int32_t));
    double* vals = va ? (double*)malloc((size_t)t * sizeof(double)) : NULL;
    int64_t w = 0;
    for (int32_t r = 0; r < n; ++r)
      for (int64_t k = rp[r]; k < rp[r + 1]; ++k) {
        rows[w] = r; cols[w] = ci[k];
        if (vals) vals[w] = va[k];
        ++w;
      }
    for (int64_t q = 0; q < added; ++q) { /* value 1 in both directions (Graph.h:240-249) */
      rows[w] = con_b[q]; cols[w] = con_a[q]; if (vals) vals[w] = 1.0; ++w; /* (second, first), (first, second) */
      rows[w] = con_a[q]; cols[w] = con_b[q]; if (vals) vals[w] = 1.0; ++w;
    }
    free(rp);
    free(ci);
    free(va);
    va = NULL;
    if (build_sorted_csr(n, t, rows, cols, vals, &rp, &ci, vals ? &va : NULL)) {
      free(rows); free(cols); free(vals); free(con_a); free(con_b); return -1;
    }
    m = t;
    free(rows);
    free(cols);
    free(vals);
  }
  free(con_a);
  free(con_b);
  *rp_out = rp;
  *ci_out = ci;
  if (vals_out) *vals_out = va;
  else free(va);
  *nnz_out = m;
  return added;
}

// emu_ops.cpp -- runs the canonicalisation and permutation kernels (matrix_ops.hpp) under the CPU wave
// emulator, through the same templated orchestration as the library. TEST INFRASTRUCTURE ONLY.
// Built by tests/emu/Makefile into tests/emu/libemu_ops.so.
#include "wave_emu.hpp"
#include "matrix_ops.hpp"
#include "level_cm.hpp"

#include <cstdlib>
#include <cstring>
#include <vector>

namespace {

// Host memory stands in for device memory; fresh buffers are filled with a pattern, as device memory is
// not zeroed either.
struct EmuExec {
  std::vector<void*> scratch, kept;
  bool bad = false;

  template <class K, class... A>
  void run(K kernel, int64_t grid, A... args) {
    if (bad || grid <= 0) return;
    emu::launch((unsigned)grid, rcm_mx::kT, false, kernel, args...);
  }
  template <class T>
  T* get(std::vector<void*>& list, size_t count) {
    const size_t bytes = std::max<size_t>(count, 1) * sizeof(T);
    void* p = std::malloc(bytes);
    if (!p) {
      bad = true;
      return nullptr;
    }
    std::memset(p, 0xA5, bytes);
    list.push_back(p);
    return static_cast<T*>(p);
  }
  template <class T>
  T* alloc(size_t count) { return get<T>(scratch, count); }
  template <class T>
  T* keep(size_t count) { return get<T>(kept, count); }
  uint32_t read_u32(const uint32_t* p) { return bad ? 0u : *p; }
  void read_words(uint32_t* dst, const uint32_t* src, int count) {
    if (!bad) std::memcpy(dst, src, (size_t)count * sizeof(uint32_t));
  }
  bool failed() const { return bad; }
  ~EmuExec() {
    for (void* p : scratch) std::free(p);
    for (void* p : kept) std::free(p);
  }
};

}  // namespace

extern "C" {

// B = P A P^T; returns the orchestration status (0 ok, -1 invalid).
int emu_permute(int32_t n, int64_t nnz, const int32_t* rp, const int32_t* ci, const double* vals, const int32_t* perm,
                int32_t* rp_out, int32_t* ci_out, double* vals_out) {
  EmuExec x;
  return rcm_mx::permute(x, n, nnz, rp, ci, vals, perm, rp_out, ci_out, vals_out);
}

// convert + prepareRCM(check_sym). On success *nnz_out, *n_out and *links_out are set and the CSR is
// copied to the caller's buffers when they are large enough (cap entries); returns the status.
int emu_canonicalize(int32_t rows, int32_t cols, int64_t nnz, const int32_t* r, const int32_t* c, const double* vals,
                     int32_t check_sym, int64_t cap, int32_t* n_out, int64_t* nnz_out, int64_t* links_out,
                     int32_t* rp_out, int32_t* ci_out, double* vals_out) {
  EmuExec x;
  rcm_mx::Csr<double> out;
  int64_t links = 0;
  const int rc = rcm_mx::canonicalise(x, rows, cols, nnz, r, c, vals, check_sym != 0, out, &links);
  if (rc) return rc;
  *n_out = out.rows;
  *nnz_out = out.nnz;
  *links_out = links;
  if (out.nnz <= cap && (int64_t)out.rows + 1 <= cap + 1) {
    std::memcpy(rp_out, out.rp, ((size_t)out.rows + 1) * sizeof(int32_t));
    std::memcpy(ci_out, out.ci, (size_t)out.nnz * sizeof(int32_t));
    if (vals && vals_out) std::memcpy(vals_out, out.vals, (size_t)out.nnz * sizeof(double));
  }
  return 0;
}

// Stable radix sort of (key, payload) pairs over lo_bits / hi_bits of the two key words.
int emu_radix_sort(int64_t n, uint64_t* keys, uint32_t* vals, int32_t lo_bits, int32_t hi_bits) {
  EmuExec x;
  uint64_t* k = x.alloc<uint64_t>((size_t)n);
  uint64_t* k2 = x.alloc<uint64_t>((size_t)n);
  uint32_t* v = x.alloc<uint32_t>((size_t)n);
  uint32_t* v2 = x.alloc<uint32_t>((size_t)n);
  std::memcpy(k, keys, (size_t)n * 8);
  std::memcpy(v, vals, (size_t)n * 4);
  rcm_mx::radix_sort(x, k, v, k2, v2, n, lo_bits, hi_bits);
  std::memcpy(keys, k, (size_t)n * 8);
  std::memcpy(vals, v, (size_t)n * 4);
  return 0;
}

// Level-synchronous order from start (bfs: plain BFS with levels). Returns the status; *placed, *levels.
int emu_level_order(int32_t n, int64_t nnz, const int32_t* rp, const int32_t* ci, int32_t start, int32_t bfs,
                    int32_t* perm, int32_t* level, int32_t* placed, int32_t* levels) {
  EmuExec x;
  uint32_t maxdeg = 0;
  for (int32_t i = 0; i < n; ++i) maxdeg = std::max<uint32_t>(maxdeg, (uint32_t)(rp[i + 1] - rp[i]));
  uint64_t* qinfo = x.alloc<uint64_t>((size_t)n);
  uint64_t* claim = x.alloc<uint64_t>((size_t)n);
  return rcm_mx::level_order(x, n, nnz, rp, ci, maxdeg, start, perm, qinfo, claim, bfs ? level : nullptr, placed,
                             levels);
}

// Exclusive scan of u32 with the total appended (out: m + 1 entries).
int emu_scan(int64_t m, const uint32_t* in, uint32_t* out) {
  EmuExec x;
  uint32_t* d = x.alloc<uint32_t>((size_t)m + 1);
  std::memcpy(d, in, (size_t)m * 4);
  rcm_mx::scan_u32(x, d, d, m, d + m);
  std::memcpy(out, d, ((size_t)m + 1) * 4);
  return 0;
}
}

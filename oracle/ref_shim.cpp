// ref_shim.cpp -- C entry points over the reference's own CPU sources (compiled from the reference tree by
// oracle/Makefile into oracle/_ref/libref_rcm.so). TEST INFRASTRUCTURE ONLY: used by tests/ to pin the
// oracle and by bench.py's cpu_baseline leg to time the reference CPU / CPU_BATCH paths.
//
// Calls: RealCPU::rcm<STABLE> (src/rcm_cpu.cpp:110-215; threads == 1 && batchsize == 1 runs the serial
// rcm_seq, anything else the threaded CPU_BATCH), Graph<T>::prepareRCM (include/Graph.h:74-264),
// Graph<T>::load (include/Graph.h:266-324), loadMTX/storeMTX (src/COO.cpp) with convert, and
// loadCSR/storeCSR (src/CSR.cpp). Exceptions the reference throws become a return value of -1.
#include <cstdint>
#include <cstring>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "COO.h"
#include "CSR.h"
#include "Graph.h"
#include "rcm_cpu.h"

namespace {
struct CoutSilencer {
  std::streambuf* old;
  std::ostringstream sink;
  explicit CoutSilencer(bool on) : old(nullptr) {
    if (on) old = std::cout.rdbuf(sink.rdbuf());
  }
  ~CoutSilencer() {
    if (old) std::cout.rdbuf(old);
  }
};

void fill_graph(Graph<float>& g, int32_t n, int64_t nnz, const int32_t* rp, const int32_t* ci) {
  g.csr.alloc((size_t)n, (size_t)n, (size_t)nnz);
  for (int64_t k = 0; k < nnz; ++k) {
    g.csr.col_ids[k] = (unsigned)ci[k];
    g.csr.data[k] = 1.0f;
  }
  for (int32_t i = 0; i <= n; ++i) g.csr.row_offsets[i] = (unsigned)rp[i];
  g.alloc((size_t)n);
}
}  // namespace

extern "C" {

// Runs the reference RealCPU::rcm (25 timed iterations inside the reference) and returns the number of
// permutation entries; *mean_ms receives the mean the reference writes to its csv stream.
int64_t ref_rcm(int32_t n, int64_t nnz, const int32_t* rp, const int32_t* ci, int32_t start, int32_t threads,
                int32_t batchsize, int32_t stable, int32_t* perm_out, double* mean_ms, int32_t quiet) {
  if (n <= 0 || start < 0 || start >= n) return -1;
  Graph<float> g;
  fill_graph(g, n, nnz, rp, ci);
  std::vector<unsigned> perm;
  std::ostringstream csv;
  {
    CoutSilencer s(quiet != 0);
    if (stable)
      RealCPU::rcm<true>(g, csv, perm, (unsigned)start, (unsigned)threads, (unsigned)batchsize);
    else
      RealCPU::rcm<false>(g, csv, perm, (unsigned)start, (unsigned)threads, (unsigned)batchsize);
  }
  if (mean_ms) {
    double v = -1.0;
    std::istringstream is(csv.str());
    is >> v;
    *mean_ms = v;
  }
  for (size_t i = 0; i < perm.size() && i < (size_t)n; ++i) perm_out[i] = (int32_t)perm[i];
  return (int64_t)perm.size();
}

// Graph<float>::prepareRCM(check_sym). Outputs are new[]-allocated; release with ref_free.
int64_t ref_prepare(int32_t n, int64_t nnz, const int32_t* rp, const int32_t* ci, int32_t check_sym,
                    int32_t** rp_out, int32_t** ci_out, int64_t* nnz_out, int32_t quiet) {
  if (n <= 0) return -1;
  Graph<float> g;
  fill_graph(g, n, nnz, rp, ci);
  {
    CoutSilencer s(quiet != 0);
    g.prepareRCM(check_sym != 0);
  }
  int32_t* r = new int32_t[(size_t)g.csr.rows + 1];
  int32_t* c = new int32_t[g.csr.nnz > 0 ? g.csr.nnz : 1];
  for (size_t i = 0; i <= g.csr.rows; ++i) r[i] = (int32_t)g.csr.row_offsets[i];
  for (size_t k = 0; k < g.csr.nnz; ++k) c[k] = (int32_t)g.csr.col_ids[k];
  *rp_out = r;
  *ci_out = c;
  *nnz_out = (int64_t)g.csr.nnz;
  return (int64_t)g.csr.rows;
}

void ref_free(int32_t* p) { delete[] p; }
}

// ---- file formats and value-carrying prepare (T = float for is_double == 0, double otherwise) ----------
namespace {
template <typename T>
int64_t export_csr(const CSR<T>& m, int32_t* rows, int32_t* cols, int32_t** rp, int32_t** ci, double** vals) {
  int32_t* r = new int32_t[m.rows + 1];
  int32_t* c = new int32_t[m.nnz > 0 ? m.nnz : 1];
  double* v = (double*)operator new[](sizeof(double) * (m.nnz > 0 ? m.nnz : 1));
  for (size_t i = 0; i <= m.rows; ++i) r[i] = (int32_t)m.row_offsets[i];
  for (size_t k = 0; k < m.nnz; ++k) {
    c[k] = (int32_t)m.col_ids[k];
    v[k] = (double)m.data[k];
  }
  *rows = (int32_t)m.rows;
  *cols = (int32_t)m.cols;
  *rp = r;
  *ci = c;
  *vals = v;
  return (int64_t)m.nnz;
}

template <typename T>
void import_csr(CSR<T>& m, int32_t rows, int32_t cols, int64_t nnz, const int32_t* rp, const int32_t* ci,
                const double* vals) {
  m.alloc((size_t)rows, (size_t)cols, (size_t)nnz);
  for (int64_t k = 0; k < nnz; ++k) {
    m.col_ids[k] = (unsigned)ci[k];
    m.data[k] = vals ? (T)vals[k] : (T)1;
  }
  for (int32_t i = 0; i <= rows; ++i) m.row_offsets[i] = (unsigned)rp[i];
}

template <typename T>
int64_t mtx_load(const char* path, int32_t* rows, int32_t* cols, int32_t** rp, int32_t** ci, double** vals,
                 uint32_t* info) {
  bool pattern, complex, symmetric, hermitian;
  COO<T> coo = loadMTX<T>(path, pattern, complex, symmetric, hermitian);
  CSR<T> csr;
  convert(csr, coo);
  *info = (pattern ? 1u : 0u) | (complex ? 2u : 0u) | (symmetric ? 4u : 0u) | (hermitian ? 8u : 0u);
  return export_csr(csr, rows, cols, rp, ci, vals);
}

template <typename T>
void mtx_store(const char* path, int32_t rows, int32_t cols, int64_t nnz, const int32_t* rp, const int32_t* ci,
               const double* vals) {
  CSR<T> csr;
  import_csr(csr, rows, cols, nnz, rp, ci, vals);
  COO<T> coo;
  convert(coo, csr);
  storeMTX(coo, path);
}

template <typename T>
int64_t graph_load(const char* file, int32_t* rows, int32_t* cols, int32_t** rp, int32_t** ci, double** vals,
                   int32_t* ret) {
  Graph<T> g;
  *ret = g.load(std::string(file));
  if (*ret < 0) return -1;
  return export_csr(g.csr, rows, cols, rp, ci, vals);
}

template <typename T>
int64_t prepare_values(int32_t rows, int32_t cols, int64_t nnz, const int32_t* rp, const int32_t* ci,
                       const double* vals, int32_t check_sym, int32_t* orows, int32_t* ocols, int32_t** orp,
                       int32_t** oci, double** ovals) {
  Graph<T> g;
  import_csr(g.csr, rows, cols, nnz, rp, ci, vals);
  g.prepareRCM(check_sym != 0);
  return export_csr(g.csr, orows, ocols, orp, oci, ovals);
}
}  // namespace

extern "C" {

// loadMTX + convert. Returns nnz, or -1 where the reference throws.
int64_t ref_mtx_load(const char* path, int32_t is_double, int32_t* rows, int32_t* cols, int32_t** rp, int32_t** ci,
                     double** vals, uint32_t* info, int32_t quiet) {
  CoutSilencer s(quiet != 0);
  try {
    return is_double ? mtx_load<double>(path, rows, cols, rp, ci, vals, info)
                     : mtx_load<float>(path, rows, cols, rp, ci, vals, info);
  } catch (const std::exception&) {
    return -1;
  }
}

// convert(COO, CSR) + storeMTX. Returns 0, or -1 where the reference throws.
int32_t ref_mtx_store(const char* path, int32_t is_double, int32_t rows, int32_t cols, int64_t nnz, const int32_t* rp,
                      const int32_t* ci, const double* vals) {
  try {
    if (is_double)
      mtx_store<double>(path, rows, cols, nnz, rp, ci, vals);
    else
      mtx_store<float>(path, rows, cols, nnz, rp, ci, vals);
    return 0;
  } catch (const std::exception&) {
    return -1;
  }
}

int64_t ref_hicsr_load(const char* path, int32_t is_double, int32_t* rows, int32_t* cols, int32_t** rp, int32_t** ci,
                       double** vals) {
  try {
    if (is_double) return export_csr(loadCSR<double>(path), rows, cols, rp, ci, vals);
    return export_csr(loadCSR<float>(path), rows, cols, rp, ci, vals);
  } catch (const std::exception&) {
    return -1;
  }
}

int32_t ref_hicsr_store(const char* path, int32_t is_double, int32_t rows, int32_t cols, int64_t nnz,
                        const int32_t* rp, const int32_t* ci, const double* vals) {
  try {
    if (is_double) {
      CSR<double> m;
      import_csr(m, rows, cols, nnz, rp, ci, vals);
      storeCSR(m, path);
    } else {
      CSR<float> m;
      import_csr(m, rows, cols, nnz, rp, ci, vals);
      storeCSR(m, path);
    }
    return 0;
  } catch (const std::exception&) {
    return -1;
  }
}

// Graph<T>::load. *ret receives load's return value (0, 1 = not square, -1 = failed). Returns nnz or -1.
int64_t ref_graph_load(const char* file, int32_t is_double, int32_t* rows, int32_t* cols, int32_t** rp, int32_t** ci,
                       double** vals, int32_t* ret, int32_t quiet) {
  CoutSilencer s(quiet != 0);
  try {
    return is_double ? graph_load<double>(file, rows, cols, rp, ci, vals, ret)
                     : graph_load<float>(file, rows, cols, rp, ci, vals, ret);
  } catch (const std::exception&) {
    return -1;
  }
}

// Graph<T>::prepareRCM(check_sym) on a CSR with values. Returns the output nnz.
int64_t ref_prepare_values(int32_t is_double, int32_t rows, int32_t cols, int64_t nnz, const int32_t* rp,
                           const int32_t* ci, const double* vals, int32_t check_sym, int32_t* orows, int32_t* ocols,
                           int32_t** orp, int32_t** oci, double** ovals, int32_t quiet) {
  CoutSilencer s(quiet != 0);
  return is_double ? prepare_values<double>(rows, cols, nnz, rp, ci, vals, check_sym, orows, ocols, orp, oci, ovals)
                   : prepare_values<float>(rows, cols, nnz, rp, ci, vals, check_sym, orows, ocols, orp, oci, ovals);
}

void ref_free_values(double* p) { operator delete[](p); }
}

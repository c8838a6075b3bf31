// ref_shim.cpp -- C entry points over the reference's own CPU sources (compiled from the reference tree by
// oracle/Makefile into oracle/_ref/libref_rcm.so). TEST INFRASTRUCTURE ONLY: used by tests/ to pin the
// oracle and by bench.py's cpu_baseline leg to time the reference CPU / CPU_BATCH paths.
//
// Calls: RealCPU::rcm<STABLE> (src/rcm_cpu.cpp:110-215; threads == 1 && batchsize == 1 runs the serial
// rcm_seq, anything else the threaded CPU_BATCH) and Graph<float>::prepareRCM (include/Graph.h:74-264).
#include <cstdint>
#include <cstring>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

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

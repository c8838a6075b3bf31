// rcm_gpu_batch_shim.hpp -- C++ host interface of the reference's GPU_BATCH path, over the C ABI of
// rcm_gpu_batch_abi.h (header-only; link against libpbrcm.so).
//
// The reference declares (include/rcm_gpu_batch.h:37-45):
//     namespace GPU_BATCH { template<typename T>
//       void rcm(Graph<T>& graph, std::ofstream& csv_file, std::vector<unsigned>& permutation, unsigned start_node); }
// pbrcm::gpu_batch_rcm below has exactly that argument meaning and behaviour (src/device/rcm_gpu_batch.cu:42-112):
//   - permutation is resized to graph.csr.rows and receives the Cuthill-McKee order (new -> old), not reversed;
//   - the persistent reorder kernel is run `warmups` + `runs` times (20 + 20), and the mean kernel time of the
//     timed runs is printed as "GPU-DYN-RCM duration: <ms> ms" and written to csv_file;
//   - failures throw (the reference throws std::runtime_error("CUDA ERROR")); pbrcm::Error derives from
//     std::runtime_error and carries the rcm_status code.
// It accepts any graph type with the reference's member layout (include/Graph.h:42-46, include/CSR.h:42-49):
// graph.csr.rows, graph.csr.nnz, graph.csr.row_offsets.get() and graph.csr.col_ids.get() (unsigned int*).
// A reference build binds it with a two-line GPU_BATCH::rcm<T> wrapper.
#pragma once
#include <cstdint>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <stdexcept>
#include <string>
#include <vector>

#include "rcm_gpu_batch_abi.h"

namespace pbrcm {

class Error : public std::runtime_error {
 public:
  Error(int status, const std::string& what)
      : std::runtime_error(what + ": " + rcm_status_string(status)), status_(status) {}
  int status() const { return status_; }

 private:
  int status_;
};

inline void check(int status, const char* what) {
  if (status != RCM_OK) throw Error(status, what);
}

// One device-resident context (owns the device buffers of one matrix at a time).
class Context {
 public:
  explicit Context(int device = 0) { check(rcm_ctx_create(device, &c_), "rcm_ctx_create"); }
  ~Context() {
    if (c_) rcm_ctx_destroy(c_);
  }
  Context(const Context&) = delete;
  Context& operator=(const Context&) = delete;

  void set(int param, int64_t value) { check(rcm_ctx_set_param(c_, param, value), "rcm_ctx_set_param"); }

  // The reference's CSR stores unsigned indices; the ABI takes int32, identical bits for n, nnz < 2^31.
  void load(size_t rows, size_t nnz, const unsigned* row_offsets, const unsigned* col_ids) {
    if (rows == 0 || rows >= (size_t(1) << 30) || nnz >= (size_t(1) << 31))
      throw Error(RCM_EINVAL, "matrix size outside the int32 CSR range");
    check(rcm_ctx_load(c_, (int32_t)rows, (int64_t)nnz, reinterpret_cast<const int32_t*>(row_offsets),
                       reinterpret_cast<const int32_t*>(col_ids)),
          "rcm_ctx_load");
    rows_ = rows;
  }
  unsigned find_start(int32_t* iterations = nullptr) {
    int32_t s = -1;
    check(rcm_ctx_find_start(c_, &s, iterations), "rcm_ctx_find_start");
    return (unsigned)s;
  }
  void run(unsigned start, uint32_t flags = RCM_FLAG_STABLE) {
    check(rcm_ctx_run(c_, (int32_t)start, flags), "rcm_ctx_run");
  }
  rcm_timing_t timing() const {
    rcm_timing_t t{};
    check(rcm_ctx_timing(c_, &t), "rcm_ctx_timing");
    return t;
  }
  void read(std::vector<unsigned>& permutation) const {
    permutation.resize(rows_);
    check(rcm_ctx_read_perm(c_, reinterpret_cast<int32_t*>(permutation.data())), "rcm_ctx_read_perm");
  }
  int64_t bandwidth() {
    int64_t bw = 0;
    check(rcm_ctx_bandwidth(c_, &bw), "rcm_ctx_bandwidth");
    return bw;
  }
  int64_t bandwidth_of(const std::vector<unsigned>& permutation) {
    if (permutation.size() != rows_) throw Error(RCM_EINVAL, "permutation size");
    int64_t bw = 0;
    check(rcm_ctx_bandwidth_of(c_, reinterpret_cast<const int32_t*>(permutation.data()), &bw),
          "rcm_ctx_bandwidth_of");
    return bw;
  }
  rcm_ctx* raw() { return c_; }

 private:
  rcm_ctx* c_ = nullptr;
  size_t rows_ = 0;
};

// GPU_BATCH::rcm<T> (include/rcm_gpu_batch.h:40-41, src/device/rcm_gpu_batch.cu:42-112).
template <class GraphT>
void gpu_batch_rcm(GraphT& graph, std::ofstream& csv_file, std::vector<unsigned>& permutation, unsigned start_node,
                   int warmups = 20, int runs = 20) {
  Context ctx;
  ctx.load(graph.csr.rows, graph.csr.nnz, graph.csr.row_offsets.get(), graph.csr.col_ids.get());
  double sum_ms = 0.0;
  for (int r = 0; r < warmups + runs; ++r) {
    ctx.run(start_node, RCM_FLAG_STABLE);
    if (r >= warmups) sum_ms += ctx.timing().run_kernel_ms;
  }
  const double mean_ms = runs > 0 ? sum_ms / runs : 0.0;
  std::cout << std::setw(27) << std::right << "GPU-DYN-RCM"
            << " duration: " << mean_ms << " ms" << std::endl;
  csv_file << mean_ms;
  csv_file << std::flush;
  ctx.read(permutation);
}

// SimplePeripheral::findPeripheral (src/simple_peripheral.cpp:92-141) on the GPU: the start node the reference's
// main uses when no -s is given (src/main.cpp:388-396).
template <class GraphT>
unsigned find_peripheral(GraphT& graph, int32_t* iterations = nullptr) {
  Context ctx;
  ctx.load(graph.csr.rows, graph.csr.nnz, graph.csr.row_offsets.get(), graph.csr.col_ids.get());
  return ctx.find_start(iterations);
}

// computeBandwidth (src/main.cpp:143-184) of graph under permutation (new -> old).
template <class GraphT>
int64_t compute_bandwidth(GraphT& graph, const std::vector<unsigned>& permutation) {
  Context ctx;
  ctx.load(graph.csr.rows, graph.csr.nnz, graph.csr.row_offsets.get(), graph.csr.col_ids.get());
  return ctx.bandwidth_of(permutation);
}

}  // namespace pbrcm

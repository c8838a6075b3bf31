// shim_drive.cpp -- drives the C++ interface (include/rcm_gpu_batch_shim.hpp) the way the reference's main
// drives GPU_BATCH (src/main.cpp:388-396, 503-515): start search, GPU_BATCH::rcm, computeBandwidth.
// Test program: reads a CSR from a binary file, writes the permutation to another.
//   shim_drive <in.bin> <out.bin> <out.csv>
//   in.bin:  int64 n, int64 nnz, uint32 row_offsets[n+1], uint32 col_ids[nnz]
//   out.bin: int64 start, int64 bandwidth, uint32 perm[n]
#include <cstdio>
#include <memory>

#include "rcm_gpu_batch_shim.hpp"

// The reference's container layout (include/CSR.h:42-49, include/Graph.h:42-46), restated for this test.
template <typename T>
struct CSR {
  size_t rows = 0, cols = 0, nnz = 0;
  std::unique_ptr<T[]> data;
  std::unique_ptr<unsigned int[]> row_offsets;
  std::unique_ptr<unsigned int[]> col_ids;
};
template <typename T>
struct Graph {
  size_t size = 0;
  CSR<T> csr;
};

namespace GPU_BATCH {
// The binding a reference build would add in place of src/device/rcm_gpu_batch.cu.
template <typename T>
void rcm(Graph<T>& graph, std::ofstream& csv_file, std::vector<unsigned>& permutation, unsigned start_node) {
  pbrcm::gpu_batch_rcm(graph, csv_file, permutation, start_node, 2, 3);
}
}  // namespace GPU_BATCH

int main(int argc, char** argv) {
  if (argc != 4) {
    std::fprintf(stderr, "usage: %s in.bin out.bin out.csv\n", argv[0]);
    return 2;
  }
  FILE* f = std::fopen(argv[1], "rb");
  if (!f) return 2;
  int64_t n = 0, nnz = 0;
  if (std::fread(&n, 8, 1, f) != 1 || std::fread(&nnz, 8, 1, f) != 1) return 2;
  Graph<float> g;
  g.size = (size_t)n;
  g.csr.rows = g.csr.cols = (size_t)n;
  g.csr.nnz = (size_t)nnz;
  g.csr.row_offsets = std::make_unique<unsigned int[]>(n + 1);
  g.csr.col_ids = std::make_unique<unsigned int[]>(nnz);
  g.csr.data = std::make_unique<float[]>(nnz);
  if (std::fread(g.csr.row_offsets.get(), 4, n + 1, f) != (size_t)(n + 1) ||
      std::fread(g.csr.col_ids.get(), 4, nnz, f) != (size_t)nnz)
    return 2;
  std::fclose(f);
  try {
    const unsigned start = pbrcm::find_peripheral(g);
    std::ofstream csv(argv[3]);
    std::vector<unsigned> perm;
    GPU_BATCH::rcm(g, csv, perm, start);
    const int64_t bw = pbrcm::compute_bandwidth(g, perm);
    FILE* o = std::fopen(argv[2], "wb");
    if (!o) return 2;
    const int64_t s64 = start;
    std::fwrite(&s64, 8, 1, o);
    std::fwrite(&bw, 8, 1, o);
    std::fwrite(perm.data(), 4, perm.size(), o);
    std::fclose(o);
    std::printf("start %u bandwidth %lld\n", start, (long long)bw);
  } catch (const pbrcm::Error& e) {
    std::fprintf(stderr, "pbrcm error %d: %s\n", e.status(), e.what());
    return 1;
  }
  return 0;
}

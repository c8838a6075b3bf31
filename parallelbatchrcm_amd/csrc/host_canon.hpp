// host_canon.hpp -- host-side pieces shared by the workload generators (workloads.cpp) and the file
// formats (formats.cpp). Internal to the library: nothing here is part of the C ABI.
#pragma once

#include <algorithm>
#include <atomic>
#include <cstdint>
#include <thread>
#include <utility>
#include <vector>

namespace pbrcm_host {

// The edges Graph<T>::prepareRCM adds to connect the matrix (include/Graph.h:160-235). Defined in
// workloads.cpp. Returns 0, or -3 where the reference's repair loop would not terminate.
int32_t connect_links(int32_t n, const int32_t* rp, const int32_t* ci,
                      std::vector<std::pair<int32_t, int32_t>>& links);

inline int hardware_threads() {
  const unsigned h = std::thread::hardware_concurrency();
  return h ? (int)std::min(h, 16u) : 4;
}

// Dynamic-grain parallel loop over [0, count); body(i) must be independent per index.
template <class F>
void parallel_range(int64_t count, int threads, F&& body) {
  if (count <= 0) return;
  threads = (int)std::min<int64_t>(threads, count);
  if (threads <= 1) {
    for (int64_t i = 0; i < count; ++i) body(i);
    return;
  }
  std::atomic<int64_t> next{0};
  const int64_t grain = std::max<int64_t>(1, count / (threads * 16));
  std::vector<std::thread> pool;
  pool.reserve((size_t)threads);
  for (int t = 0; t < threads; ++t)
    pool.emplace_back([&]() {
      for (;;) {
        const int64_t b = next.fetch_add(grain);
        if (b >= count) break;
        const int64_t e = std::min(count, b + grain);
        for (int64_t i = b; i < e; ++i) body(i);
      }
    });
  for (auto& th : pool) th.join();
}

}  // namespace pbrcm_host

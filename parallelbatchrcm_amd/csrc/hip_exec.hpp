// hip_exec.hpp -- the executor that runs the templated orchestration of matrix_ops.hpp and level_cm.hpp on a
// HIP stream: 256-thread launches, grow-only device arenas and a pinned readback word.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <vector>

#include "matrix_ops.hpp"

namespace rcm_mx {
namespace {  // internal linkage: each library source that includes this has its own copy

// Bump allocator over one device buffer. Requests beyond its capacity get their own hipMalloc for the current
// call and grow the buffer for the next one. reset() ends a call.
struct DevArena {
  char* base = nullptr;
  size_t cap = 0, used = 0, wanted = 0;
  std::vector<void*> extra;

  void* get(size_t bytes) {
    bytes = (std::max<size_t>(bytes, 1) + 255) & ~(size_t)255;
    wanted += bytes;
    if (base && used + bytes <= cap) {
      void* p = base + used;
      used += bytes;
      return p;
    }
    void* p = nullptr;
    if (hipMalloc(&p, bytes) != hipSuccess) return nullptr;
    extra.push_back(p);
    return p;
  }
  void reset() {
    for (void* p : extra) (void)hipFree(p);
    extra.clear();
    if (wanted > cap) {
      if (base) (void)hipFree(base);
      base = nullptr;
      cap = 0;
      const size_t want = wanted + wanted / 4;
      if (hipMalloc(reinterpret_cast<void**>(&base), want) == hipSuccess) cap = want;
    }
    used = wanted = 0;
  }
  void release() {
    for (void* p : extra) (void)hipFree(p);
    extra.clear();
    if (base) (void)hipFree(base);
    base = nullptr;
    cap = used = wanted = 0;
  }
};

struct HipExec {
  hipStream_t stream;
  DevArena* scratch;
  DevArena* result;
  uint32_t* h_word;  // pinned, kPinnedWords words
  bool bad = false;

  template <class K, class... A>
  void run(K kernel, int64_t grid, A... args) {
    if (bad || grid <= 0) return;
    hipLaunchKernelGGL(kernel, dim3((unsigned)grid), dim3(kT), 0, stream, args...);
    if (hipGetLastError() != hipSuccess) bad = true;
  }
  template <class T>
  T* alloc(size_t count) {
    T* p = static_cast<T*>(scratch->get(count * sizeof(T)));
    if (!p) bad = true;
    return p;
  }
  template <class T>
  T* keep(size_t count) {
    T* p = static_cast<T*>(result->get(count * sizeof(T)));
    if (!p) bad = true;
    return p;
  }
  uint32_t read_u32(const uint32_t* p) {
    if (bad) return 0;
    if (hipMemcpyAsync(h_word, p, sizeof(uint32_t), hipMemcpyDeviceToHost, stream) != hipSuccess ||
        hipStreamSynchronize(stream) != hipSuccess) {
      bad = true;
      return 0;
    }
    return *h_word;
  }
  // count <= 16 words (the pinned buffer's size)
  void read_words(uint32_t* dst, const uint32_t* src, int count) {
    if (bad) return;
    if (hipMemcpyAsync(h_word, src, (size_t)count * sizeof(uint32_t), hipMemcpyDeviceToHost, stream) != hipSuccess ||
        hipStreamSynchronize(stream) != hipSuccess) {
      bad = true;
      return;
    }
    for (int i = 0; i < count; ++i) dst[i] = h_word[i];
  }
  bool failed() const { return bad; }
};

constexpr size_t kPinnedWords = 16;  // size of the executor's pinned readback buffer

}  // namespace
}  // namespace rcm_mx

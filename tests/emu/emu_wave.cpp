// emu_wave.cpp -- runs the one-wave worker engine (engine_wave.hpp) under the CPU wave emulator.
// TEST INFRASTRUCTURE ONLY. Built by tests/emu/Makefile into tests/emu/libemu_wave.so.
#include "wave_emu.hpp"
#include "engine_wave.hpp"

#include <memory>

using namespace rcm;

namespace {
template <int S, bool BFS>
void run_reorder(unsigned workers, RunArgs a) {
  emu::emu_lds_bytes = sizeof(Smem<Geo<S>, 4>);
  emu::launch(workers, 64, true, k_reorder<Geo<S>, 4, false, BFS, false>, a);
  emu::emu_lds_bytes = 256 * 1024;
}
}  // namespace

extern "C" {

// info[0] = device error, [1] = placed, [2] = batches, [3..11] = error diagnostics
int emu_run(int32_t n, int64_t nnz, const int32_t* rp, const int32_t* ci, int32_t start, int32_t workers,
            int32_t slots, int32_t sort_mode, int32_t bfs, int32_t* perm_out, int32_t* level_out, int32_t* info,
            double timeout_s, int32_t claim_filter) {
  (void)claim_filter;  // the one-wave variant has no claim filter
  if (n <= 0 || start < 0 || start >= n || workers <= 0) return -1;
  std::vector<uint64_t> claim(n), qinfo(n), bstat(n + 1, 0), brange(n + 1, 0);
  std::vector<int32_t> perm(n, -1), level(n, -1);
  Control ctl{};
  RunArgs a{};
  a.row_ptr = rp;
  a.col_idx = ci;
  a.perm = perm.data();
  a.qinfo = qinfo.data();
  a.claim = claim.data();
  a.bstat = bstat.data();
  a.brange = brange.data();
  a.level = bfs ? level.data() : nullptr;
  a.ctl = &ctl;
  a.n = n;
  a.bcap = n + 1;
  a.budget_ticks = (uint64_t)(timeout_s * 1e8);
  a.sort_mode = sort_mode;
  a.seg_threshold = 24;
  a.prof = nullptr;
  emu::launch(1, 64, false, k_init, a, start, n + 1);
  switch ((bfs ? 10000 : 0) + slots) {
    case 512: run_reorder<512, false>(workers, a); break;
    case 1024: run_reorder<1024, false>(workers, a); break;
    case 2048: run_reorder<2048, false>(workers, a); break;
    case 10000 + 1024: run_reorder<1024, true>(workers, a); break;
    default: return -1;
  }
  std::memcpy(perm_out, perm.data(), sizeof(int32_t) * n);
  if (level_out) std::memcpy(level_out, level.data(), sizeof(int32_t) * n);
  info[0] = (int32_t)ctl.error;
  info[1] = (int32_t)ctl.placed;
  info[2] = (int32_t)ctl.nbatches;
  for (int i = 0; i < 9; ++i) info[3 + i] = (int32_t)ctl.diag[i];
  return 0;
}
}

extern "C" int emu_selftest_wave(const uint32_t* in64, uint32_t* out195) {
  emu::launch(1, 64, false, k_selftest_wave, in64, out195);
  return 0;
}

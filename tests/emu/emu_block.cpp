// emu_block.cpp -- runs the library's workgroup engine (engine.hpp) under the CPU wave emulator.
// TEST INFRASTRUCTURE ONLY. Built by tests/emu/Makefile into tests/emu/libemu_block.so.
#include "wave_emu.hpp"
#include "engine.hpp"

#include <memory>

using namespace rcm;

namespace {
// the library's engine variants: (threads, adjacency slots, parents, digit bits)
template <class G, int D, bool BFS, bool LEV>
void run_reorder(unsigned workers, RunArgs a) {
  emu::emu_lds_bytes = sizeof(Smem<G, D>);
  emu::launch(workers, G::T, true, k_reorder<G, D, false, BFS, false, LEV>, a);
  emu::emu_lds_bytes = 256 * 1024;
}
template <bool BFS, bool LEV = false>
bool run_variant(int variant, unsigned workers, RunArgs a) {
  switch (variant) {
    case 0: run_reorder<Geo<256, 2048, 256>, 4, BFS, LEV>(workers, a); return true;
    case 1: run_reorder<Geo<256, 4096, 256>, 4, BFS, LEV>(workers, a); return true;
    case 2: run_reorder<Geo<128, 1024, 128>, 4, BFS, LEV>(workers, a); return true;
    case 3: run_reorder<Geo<256, 2048, 256>, 6, BFS, LEV>(workers, a); return true;
    case 4: run_reorder<Geo<512, 4096, 512>, 4, BFS, LEV>(workers, a); return true;
    default: return false;
  }
}
}  // namespace

extern "C" {

// info[0] = device error, [1] = placed, [2] = batches, [3..11] = error diagnostics
int emu_run(int32_t n, int64_t nnz, const int32_t* rp, const int32_t* ci, int32_t start, int32_t workers,
            int32_t slots, int32_t sort_mode, int32_t bfs, int32_t* perm_out, int32_t* level_out, int32_t* info,
            double timeout_s, int32_t claim_filter) {
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
  a.claim_filter = claim_filter;
  a.prof = nullptr;
  emu::launch(1, 64, false, k_init, a, start, n + 1);
  // `slots` selects the variant: 2048 -> 0, 4096 -> 1, 1024 -> 2; 2049 -> 3 (6-bit digits), 4097 -> 4
  const int variant = slots == 2048 ? 0 : slots == 4096 ? 1 : slots == 1024 ? 2 : slots == 2049 ? 3 : slots == 4097 ? 4 : -1;
  // bfs: 0 = CM, 1 = plain BFS with levels, 2 = CM with levels
  if (!(bfs == 1 ? run_variant<true>(variant, workers, a)
                  : bfs == 2 ? run_variant<false, true>(variant, workers, a) : run_variant<false>(variant, workers, a)))
    return -1;
  std::memcpy(perm_out, perm.data(), sizeof(int32_t) * n);
  if (level_out) std::memcpy(level_out, level.data(), sizeof(int32_t) * n);
  info[0] = (int32_t)ctl.error;
  info[1] = (int32_t)ctl.placed;
  info[2] = (int32_t)ctl.nbatches;
  for (int i = 0; i < 9; ++i) info[3 + i] = (int32_t)ctl.diag[i];
  return 0;
}
}


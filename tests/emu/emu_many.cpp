// emu_many.cpp -- runs the many-matrix engine (engine_many.hpp) under the CPU wave emulator.
// TEST INFRASTRUCTURE ONLY. Built by tests/emu/Makefile into tests/emu/libemu_many.so.
#include "wave_emu.hpp"
#include "engine_many.hpp"
#include "engine_wg.hpp"

#include <vector>

using namespace rcm;

namespace {
using GeoLib = Geo<256, 2048, 256>;
}

extern "C" {

// Reorders `count` matrices (concatenated: row_ptr_cat + rp_off[k], col_idx_cat + ci_off[k]) from start[k] with
// `workers` persistent workgroups. perm_cat receives matrix k's order at sum_{i<k} n[i]; info[4*k..] =
// error, placed, batches, done.
int emu_run_many(int32_t count, const int32_t* n, const int64_t* nnz, const int64_t* rp_off, const int64_t* ci_off,
                 const int32_t* row_ptr_cat, const int32_t* col_idx_cat, const int32_t* start, int32_t workers,
                 int32_t bfs, int32_t* perm_cat, int32_t* level_cat, int32_t* info, double timeout_s) {
  if (count <= 0 || workers <= 0) return -1;
  int64_t tot = 0, totb = 0;
  for (int k = 0; k < count; ++k) {
    tot += n[k];
    totb += n[k] + 1;
  }
  std::vector<int32_t> perm(tot, -1), level(tot, -1), aux(tot, 0);
  std::vector<uint64_t> claim(tot), qinfo(tot), bstat(totb), brange(totb);
  std::vector<Control> ctl(count);
  std::vector<uint32_t> helpers(count, 0);
  std::vector<JobDesc> jobs(count);
  std::vector<int32_t> order(count), starts(start, start + count);
  int64_t po = 0, pb = 0;
  for (int k = 0; k < count; ++k) {
    JobDesc& J = jobs[k];
    J.row_ptr = row_ptr_cat + rp_off[k];
    J.col_idx = col_idx_cat + ci_off[k];
    J.perm = perm.data() + po;
    J.qinfo = qinfo.data() + po;
    J.claim = claim.data() + po;
    J.level = level.data() + po;
    J.aux = aux.data() + po;
    J.bstat = bstat.data() + pb;
    J.brange = brange.data() + pb;
    J.ctl = &ctl[k];
    J.n = n[k];
    J.cap = n[k] / 64 + 1;  // small graphs: several helpers per job, to exercise the shared tickets
    po += n[k];
    pb += n[k] + 1;
    order[k] = count - 1 - k;  // begin from the last job: any order must give the same results
  }
  ManyControl g{};
  ManyArgs m{};
  m.jobs = jobs.data();
  m.helpers = helpers.data();
  m.order = order.data();
  m.starts = starts.data();
  m.count = count;
  m.g = &g;
  m.budget_ticks = (uint64_t)(timeout_s * 1e8);
  m.sort_mode = 0;
  m.seg_threshold = 24;
  m.share_workers = (uint32_t)workers;  // the size-proportional helper floor, as the library's launches use it
  m.share_rows = (uint64_t)tot;
  std::vector<int32_t> blk(count + 1);
  for (int k = 0; k <= count; ++k) blk[k] = 2 * k;  // two workgroups per job
  emu::launch((unsigned)count * 2, 64, false, k_init_many, m, (const int32_t*)blk.data());
  emu::emu_lds_bytes = sizeof(ManySmem<GeoLib, 4>);
  // bfs: 0 = CM, 1 = plain BFS with levels, 2 = CM with levels
  if (bfs == 1)
    emu::launch((unsigned)workers, GeoLib::T, true, k_reorder_many<GeoLib, 4, false, true>, m);
  else if (bfs == 2)
    emu::launch((unsigned)workers, GeoLib::T, true, k_reorder_many<GeoLib, 4, false, false, true>, m);
  else
    emu::launch((unsigned)workers, GeoLib::T, true, k_reorder_many<GeoLib, 4, false, false>, m);
  emu::emu_lds_bytes = 256 * 1024;
  std::memcpy(perm_cat, perm.data(), sizeof(int32_t) * tot);
  if (level_cat) std::memcpy(level_cat, level.data(), sizeof(int32_t) * tot);
  for (int k = 0; k < count; ++k) {
    info[4 * k] = (int32_t)ctl[k].error;
    info[4 * k + 1] = (int32_t)ctl[k].placed;
    info[4 * k + 2] = (int32_t)ctl[k].nbatches;
    info[4 * k + 3] = (int32_t)ctl[k].done;
  }
  return (int)g.jobs_done;
}

// One workgroup per matrix (engine_wg.hpp) over `count` concatenated matrices, `workers` persistent
// workgroups. perm_cat / level_cat as emu_run_many; info[2*k] = placed, info[2*k+1] = levels.
int emu_run_wg(int32_t count, const int32_t* n, const int64_t* rp_off, const int64_t* ci_off,
               const int32_t* row_ptr_cat, const int32_t* col_idx_cat, const int32_t* start, int32_t workers,
               int32_t bfs, int32_t* perm_cat, int32_t* level_cat, int32_t* info) {
  if (count <= 0 || workers <= 0) return -1;
  int64_t tot = 0;
  for (int k = 0; k < count; ++k) {
    int32_t maxrow = 0;
    const int32_t* rp = row_ptr_cat + rp_off[k];
    for (int32_t i = 0; i < n[k]; ++i) maxrow = std::max(maxrow, rp[i + 1] - rp[i]);
    if (maxrow > kWgMaxRow) return -2;
    tot += n[k];
  }
  std::vector<int32_t> perm(tot, -1), level(tot, -1);
  std::vector<uint64_t> claim(tot, 0x5A5A5A5A5A5A5A5Aull), qinfo(tot);
  std::vector<Control> ctl(count);
  std::vector<JobDesc> jobs(count);
  std::vector<int32_t> order(count), starts(start, start + count);
  int64_t po = 0;
  for (int k = 0; k < count; ++k) {
    JobDesc& J = jobs[k];
    J = JobDesc{};
    J.row_ptr = row_ptr_cat + rp_off[k];
    J.col_idx = col_idx_cat + ci_off[k];
    J.perm = perm.data() + po;
    J.qinfo = qinfo.data() + po;
    J.claim = claim.data() + po;
    J.level = level.data() + po;
    J.ctl = &ctl[k];
    J.n = n[k];
    po += n[k];
    order[k] = count - 1 - k;
  }
  ManyControl g{};
  ManyArgs m{};
  m.jobs = jobs.data();
  m.order = order.data();
  m.starts = starts.data();
  m.count = count;
  m.g = &g;
  emu::launch((unsigned)((count + 255) / 256), 256, false, k_wg_prepare, m);
  emu::emu_lds_bytes = sizeof(WgSmem);
  if (bfs)
    emu::launch((unsigned)workers, kWgT, true, k_order_wg<true>, m);
  else
    emu::launch((unsigned)workers, kWgT, true, k_order_wg<false>, m);
  emu::emu_lds_bytes = 256 * 1024;
  std::memcpy(perm_cat, perm.data(), sizeof(int32_t) * tot);
  if (level_cat) std::memcpy(level_cat, level.data(), sizeof(int32_t) * tot);
  for (int k = 0; k < count; ++k) {
    info[2 * k] = (int32_t)ctl[k].placed;
    info[2 * k + 1] = (int32_t)ctl[k].nbatches;
  }
  return 0;
}
}

/*
 * rcm_gpu_batch_abi.h -- C ABI of the MI355X (gfx950) Cuthill-McKee engine.
 *
 * Drop-in for the reference's GPU_BATCH path. Every entry point replaces a reference interface, cited as
 * file:line in the reference repository (GPUPeople/ParallelBatchRCM):
 *   rcm_gpu_batch                 GPU_BATCH::rcm<T>             include/rcm_gpu_batch.h:40-41,
 *                                                              src/device/rcm_gpu_batch.cu:42-112
 *   rcm_ctx_find_start            SimplePeripheral::findPeripheral include/simple_peripheral.h:45-46,
 *                                                              src/simple_peripheral.cpp:92-141
 *   rcm_ctx_bandwidth             computeBandwidth              src/main.cpp:143-184
 *   rcm_ctx_run / rcm_ctx_load    the device half of GPU_BATCH::rcm (upload, init kernel, persistent run
 *                                 kernel, download; src/device/rcm_gpu_batch.cu:57-102) split so callers
 *                                 can keep the CSR resident in HBM between calls
 *   rcm_batch_* / rcm_gpu_batch_many  no single reference counterpart: GPU_BATCH::rcm applied to many
 *                                 matrices, done in one persistent launch
 *   rcm_server_*                  no reference counterpart: continuous batching of reorder requests
 *
 * Conventions
 *   - CSR is int32: row_ptr[n+1], col_idx[nnz]; 1 <= n < 2^30, nnz < 2^31. Values are not needed.
 *   - Output perm[new] = old. Without RCM_FLAG_REVERSE it is Cuthill-McKee order (what the reference's
 *     GPU_BATCH::rcm and RealCPU::rcm return); with it, the reverse (RCM).
 *   - The order is the stable one: bit-identical to RealCPU::rcm_seq<true> (src/rcm_cpu.cpp:57-104) from the
 *     same start node. RCM_FLAG_STABLE is accepted for interface parity; the engine is always stable.
 *   - No exceptions cross this boundary; nothing calls exit()/abort(). Functions return RCM_OK (0) or a
 *     negative rcm_status. A disconnected graph returns RCM_ENOTCONNECTED (the reference kernel hangs).
 *   - Thread safety: one rcm_ctx per host thread; distinct contexts are independent.
 */
#ifndef RCM_GPU_BATCH_ABI_H
#define RCM_GPU_BATCH_ABI_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum rcm_flags {
  RCM_FLAG_REVERSE = 1u << 0,   /* return RCM (reversed CM) instead of CM */
  RCM_FLAG_STABLE = 1u << 1,    /* stable degree sort (always on) */
  RCM_FLAG_BANDWIDTH = 1u << 2, /* also compute the bandwidth of the reordered matrix */
  RCM_FLAG_BORROW = 1u << 3     /* rcm_server_submit only: read the caller's CSR in place (no copy at submit);
                                   the buffers must stay valid and unchanged until rcm_server_wait has returned
                                   the request's result, or rcm_server_destroy has returned */
};

enum rcm_status {
  RCM_OK = 0,
  RCM_EINVAL = -1,         /* malformed CSR, start node out of range, bad argument */
  RCM_ENOTCONNECTED = -2,  /* the start's component does not cover all n nodes */
  RCM_ENOMEM = -3,         /* device or host allocation failed */
  RCM_EDEVICE = -4,        /* HIP runtime error */
  RCM_ETIMEOUT = -5,       /* device watchdog expired (see RCM_PARAM_TIMEOUT_MS) */
  RCM_ECAPACITY = -6,      /* an internal capacity was exceeded */
  RCM_ENOTLOADED = -7,     /* context has no matrix loaded */
  RCM_ECANCELED = -8,      /* serving: the server shut down before the request finished */
  RCM_ENOTFOUND = -9,      /* serving: unknown request id */
  RCM_EAGAIN = -10,        /* serving: the request has not finished within the wait's time limit */
  RCM_EIO = -11,           /* a matrix file could not be opened, read or written */
  RCM_ENOTSYMMETRIC = -12  /* reserved (no longer returned: the device repair handles non-symmetric matrices) */
};

/* One-shot call on host buffers: upload, (optional GPU start search), reorder, (optional bandwidth),
 * download. start_node < 0 selects the pseudo-peripheral search (== SimplePeripheral::findPeripheral).
 * perm_out: caller-allocated n entries. bandwidth_out, start_node_out: nullable. */
int rcm_gpu_batch(int32_t n, int64_t nnz, const int32_t* row_ptr, const int32_t* col_idx, int32_t start_node,
                  uint32_t flags, int32_t* perm_out, int64_t* bandwidth_out, int32_t* start_node_out);

/* No counterpart in the reference. For callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc): rcm_gpu_batch on a CSR already in device
 * memory (d_row_ptr[n+1], d_col_idx[nnz]) writing the order to the caller's device buffer d_perm_out[n]. It
 * calls no hipMalloc, hipFree or other allocating routine: every other buffer the engine needs lives in one
 * fixed file-scope device array (72 MiB + 1 KiB per device that runs it, held until the process ends), so n is
 * at most 2,097,152; larger n returns RCM_ECAPACITY. Runs on the current device's default stream and returns
 * when the result is in d_perm_out. start_node, flags, bandwidth_out and start_node_out (host, nullable) as in
 * rcm_gpu_batch. Every call on one device shares that device's array: calls of one process are serialised,
 * calls from different processes on one device race. */
int rcm_gpu_batch_noalloc(int32_t n, int64_t nnz, const int32_t* d_row_ptr, const int32_t* d_col_idx,
                          int32_t start_node, uint32_t flags, int32_t* d_perm_out, int64_t* bandwidth_out,
                          int32_t* start_node_out);

/* ---- device-resident context ---- */
typedef struct rcm_ctx rcm_ctx;

enum rcm_param {
  RCM_PARAM_WORKERS = 1,     /* persistent workgroups for one matrix (0 = autotuned) */
  RCM_PARAM_TIMEOUT_MS = 2,  /* device watchdog, default 20000 */
  RCM_PARAM_SORT = 3,        /* 0 auto, 1 LDS radix sort, 2 segment-rank sort */
  RCM_PARAM_DIGIT_BITS = 4,  /* radix digit width of the current variant (read-only: set the variant) */
  RCM_PARAM_VARIANT = 5,     /* engine variant (block size, batch width, radix digit):
                                0 = 256 threads, 2048 adjacency slots, 256 parents, 4-bit digits (serving
                                    lanes)
                                1 = 512 threads, 2048 slots, 256 parents, 4-bit digits (rcm_ctx default)
                                2 = 128 threads, 1024 slots, 128 parents, 4-bit digits
                                3 = 256 threads, 2048 slots, 256 parents, 6-bit digits, child batches cut at a
                                    degree quantum of S/4 = 512 slots (S/2 in the others)
                                4 = 512 threads, 4096 slots, 512 parents, 4-bit digits (rcm_batch default)
                                5 = level-synchronous engine (rcm_ctx only): one BFS level per step, children
                                    ordered by a device-wide radix sort with 8-bit digits; for wide, shallow
                                    graphs. The structure of GPU_SIMPLE (src/device/rcm_gpu.cu:134-330) with
                                    the serial order. workers is 0. */
  RCM_PARAM_STRICT = 6,      /* 1 = formal agent-scope release/acquire fences instead of sc1 + s_waitcnt */
  RCM_PARAM_PROFILE = 7,     /* 1 = instrumented reorder kernel (per-phase time sums, rcm_ctx_profile);
                                2 = the same plus a per-batch trace (rcm_ctx_trace) */
  RCM_PARAM_GRAPH = 8,       /* rcm_batch only: 1 = replay the run's launches from a captured hipGraph */
  RCM_PARAM_CLAIM_FILTER = 9 /* 0 = atomicMin on every adjacency entry; 1 = read the claim word first and skip
                                entries already claimed by a smaller code (plain load); 2 = same, device-scope */,
  RCM_PARAM_WG_MAX_N = 10  /* rcm_batch only, reserved: 0 is accepted, any other value is RCM_EINVAL (a
                              one-workgroup-per-matrix engine is not part of the library). */,
  RCM_PARAM_HELPER_NODES = 11 /* rcm_batch only: a matrix of n rows is helped by at most n / value + 1 persistent
                                 workgroups at once (at most 1024), or by its share of the launch's workgroups in
                                 proportion to its n if that is more; default 131072 (2048 in serving lanes).
                                 Fewer helpers per matrix leave fewer workgroups waiting on narrow levels
                                 (throughput of large sets); more give each matrix a shorter time to result.
                                 rcm_batch_autotune may change it. */,
  RCM_PARAM_BATCH_CAP = 12    /* rcm_ctx only: batch records a run may use; 0 (default) = n + 1, the most a run
                                 can need. A smaller value makes a run that needs more records stop with
                                 RCM_ECAPACITY (every workgroup exits; the context stays usable) -- the error
                                 path of the reference's oversize-batch __trap (include/device/batchWorker.cuh:
                                 539-543), which cannot be reached otherwise. */,
  RCM_PARAM_SPECULATE = 13    /* rcm_ctx_order / rcm_batch_order: from this start-search sweep on (default 3; 0 =
                                 never), a sweep first runs as a CM run that also records BFS levels; when its
                                 eccentricity equals the previous sweep's it is the confirming sweep and its
                                 order the result, one engine pass fewer. Results are the same for any value. */
};

int rcm_ctx_create(int device, rcm_ctx** out);
int rcm_ctx_destroy(rcm_ctx* ctx);
int rcm_ctx_set_param(rcm_ctx* ctx, int param, int64_t value);
int rcm_ctx_get_param(rcm_ctx* ctx, int param, int64_t* value);

/* Copy a host CSR into the context (validated on the device). */
int rcm_ctx_load(rcm_ctx* ctx, int32_t n, int64_t nnz, const int32_t* row_ptr, const int32_t* col_idx);
/* Use a CSR already in device memory (not copied; must stay valid while the context uses it). */
int rcm_ctx_load_device(rcm_ctx* ctx, int32_t n, int64_t nnz, const int32_t* d_row_ptr, const int32_t* d_col_idx);

/* GPU pseudo-peripheral start search, semantics of SimplePeripheral::findPeripheral (outputnodes = 32). */
int rcm_ctx_find_start(rcm_ctx* ctx, int32_t* start_out, int32_t* iterations_out);

/* Reorder the loaded matrix from start_node (>= 0). Result stays on the device. */
int rcm_ctx_run(rcm_ctx* ctx, int32_t start_node, uint32_t flags);

/* Start search and reorder in one call: start_node >= 0 is rcm_ctx_run; start_node < 0 gives the start, sweep count
 * and order of rcm_ctx_find_start followed by rcm_ctx_run, in one engine pass fewer when the search stops at a
 * speculated sweep (RCM_PARAM_SPECULATE). The device half of GPU_BATCH::rcm with its start from
 * SimplePeripheral::findPeripheral (src/device/rcm_gpu_batch.cu:42-112, src/simple_peripheral.cpp:92-141).
 * start_out, iterations_out may be NULL. */
int rcm_ctx_order(rcm_ctx* ctx, int32_t start_node, uint32_t flags, int32_t* start_out, int32_t* iterations_out);

/* Bandwidth of the loaded matrix under the last computed ordering (reversal does not change it). */
int rcm_ctx_bandwidth(rcm_ctx* ctx, int64_t* bandwidth_out);

/* Bandwidth of the loaded matrix under an external host ordering perm[new] = old (n entries; checked to be a
 * permutation of [0, n) on the host, RCM_EINVAL otherwise). Replaces computeBandwidth(graph, permutation)
 * (src/main.cpp:143-184) for orderings produced elsewhere. The ordering becomes the context's result. */
int rcm_ctx_bandwidth_of(rcm_ctx* ctx, const int32_t* perm, int64_t* bandwidth_out);

/* Checks on the GPU that an order is exactly the stable Cuthill-McKee order of the loaded matrix from `start`
 * (RealCPU::rcm_seq<true>, src/rcm_cpu.cpp:57-104), without running the serial algorithm: perm[0] = start, perm
 * is a bijection, every other node's earliest-placed neighbour precedes it, and perm[1..] increases strictly in
 * (position of that neighbour, row length, first index in its row) -- conditions that only that order meets.
 * perm: host, n entries, RCM order when flags has RCM_FLAG_REVERSE; NULL checks the context's last result (then
 * the flag is ignored). An external perm discards the context's last result. *result_out: 0 when perm is the
 * order, else 1 + a position where it departs. No single reference counterpart: an O(n + nnz) extension of
 * checkPermutation (src/main.cpp:186-226). */
int rcm_ctx_verify(rcm_ctx* ctx, int32_t start, const int32_t* perm, uint32_t flags, int64_t* result_out);

/* Copy the last ordering to the host (n entries), reversed when RCM_FLAG_REVERSE was set in rcm_ctx_run. */
int rcm_ctx_read_perm(rcm_ctx* ctx, int32_t* perm_out);
/* Device pointer to the last ordering (n entries, same orientation as rcm_ctx_read_perm). */
int rcm_ctx_perm_device(rcm_ctx* ctx, const int32_t** d_perm_out);

typedef struct rcm_timing {
  double run_kernel_ms;     /* persistent reorder kernel only (the reference's timed region) */
  double init_kernel_ms;    /* state reset kernel */
  double start_search_ms;   /* last rcm_ctx_find_start, all iterations */
  double bandwidth_ms;      /* last rcm_ctx_bandwidth */
  double load_ms;           /* last rcm_ctx_load: H2D copy + validation */
  int32_t workers;          /* workgroups of the last run */
  int32_t batches;          /* speculative batches of the last run */
  int32_t start_iterations; /* BFS sweeps of the last start search */
  int32_t placed;           /* nodes placed by the last run */
  int32_t error_site;       /* device-side diagnostics of the last failed run (0 = none) */
  int32_t error_info[8];
  double order_ms;          /* last rcm_ctx_order: start search and order, host wall time */
  int32_t order_passes;     /* engine passes of the last rcm_ctx_order */
} rcm_timing_t;
int rcm_ctx_timing(rcm_ctx* ctx, rcm_timing_t* out);
/* Per-phase profile of the last instrumented run (RCM_PARAM_PROFILE = 1): 100 MHz tick sums over all batches
 * for acquire, parents, claim, finalise, sort, cut, prefix, write; then counts of batches, early batches,
 * owned children and adjacency slots. Copies min(count, 16) words. */
int rcm_ctx_profile(rcm_ctx* ctx, uint64_t* out, int32_t count);
/* Per-batch trace of the last run with RCM_PARAM_PROFILE = 2: 12 words per batch id (id | kept << 32 | early
 * << 33; parent range begin | end << 32; children | child batches << 32; first child batch id; 100 MHz
 * timestamps at entry, claims done, claim completeness known, cut done, prefix known, records published;
 * two unused words). Copies min(words, 12 * (n + 1)) words; ids the run did not use are zero. */
int rcm_ctx_trace(rcm_ctx* ctx, uint64_t* out, int64_t words);

/* ---- many matrices, device-resident: one persistent launch reorders all of them ---- */
/* No single reference counterpart: GPU_BATCH::rcm (src/device/rcm_gpu_batch.cu:42-112) applied to every
 * matrix of a set, with the start of each from SimplePeripheral::findPeripheral (src/simple_peripheral.cpp:92-141)
 * and its bandwidth from computeBandwidth (src/main.cpp:143-184). Each matrix keeps its own queue, claim words
 * and batch records; persistent workgroups take whichever matrix has a batch ready. Results are bit-identical
 * to reordering each matrix alone. */
typedef struct rcm_batch rcm_batch;

typedef struct rcm_batch_timing {
  double load_ms;          /* last rcm_batch_load: H2D copies + validation */
  double start_search_ms;  /* last rcm_batch_find_start, all sweeps */
  double run_ms;           /* last rcm_batch_run: reset + persistent reorder kernel (device time) */
  double bandwidth_ms;     /* last rcm_batch_bandwidth (device time) */
  int32_t count;           /* matrices loaded */
  int32_t workers;         /* persistent workgroups of the last launch */
  int32_t start_sweeps;    /* launches of the last start search */
  int32_t graph_replays;   /* runs served by replaying a captured hipGraph */
  int64_t batches;         /* speculative batches of the last run, all matrices */
  int64_t total_n;         /* sum of n */
  int64_t total_nnz;       /* sum of nnz */
  double order_ms;         /* last rcm_batch_order: start search and order, host wall time */
  int32_t order_passes;    /* launches of the last rcm_batch_order */
} rcm_batch_timing_t;

int rcm_batch_create(int device, rcm_batch** out);
int rcm_batch_destroy(rcm_batch* b);
/* RCM_PARAM_WORKERS, RCM_PARAM_TIMEOUT_MS, RCM_PARAM_SORT, RCM_PARAM_GRAPH (1 = capture the reset and reorder
 * launches of rcm_batch_run in a hipGraph and replay it while the loaded set is unchanged; default 1),
 * RCM_PARAM_HELPER_NODES (applies to the loaded set at once). */
int rcm_batch_set_param(rcm_batch* b, int param, int64_t value);
/* Copy `count` matrices to the device (layout as rcm_gpu_batch_many) and validate each; status_out[k]
 * (nullable) is RCM_OK or RCM_EINVAL for matrix k. Returns RCM_EINVAL if any matrix is malformed. */
int rcm_batch_load(rcm_batch* b, int32_t count, const int32_t* n, const int64_t* nnz, const int64_t* rp_off,
                   const int64_t* ci_off, const int32_t* row_ptr_cat, const int32_t* col_idx_cat,
                   int32_t* status_out);
/* Pseudo-peripheral start of every matrix (== SimplePeripheral::findPeripheral per matrix); each sweep is one
 * persistent launch over the matrices still iterating. */
int rcm_batch_find_start(rcm_batch* b, int32_t* start_out, int32_t* iterations_out, int32_t* status_out);
/* Reorder every matrix from start[k] (>= 0). status_out[k] (nullable): RCM_OK, RCM_ENOTCONNECTED, ... */
int rcm_batch_run(rcm_batch* b, const int32_t* start, uint32_t flags, int32_t* status_out);
/* Start search and reorder of every matrix in one call: matrix k is ordered from start_in[k] when start_in is not
 * NULL and start_in[k] >= 0, else from its rcm_batch_find_start start. Same starts, sweep counts, orders and
 * statuses as rcm_batch_find_start + rcm_batch_run, with one engine pass fewer for every matrix whose search stops
 * at a speculated sweep (RCM_PARAM_SPECULATE). Output arrays may be NULL. */
int rcm_batch_order(rcm_batch* b, const int32_t* start_in, uint32_t flags, int32_t* start_out,
                    int32_t* iterations_out, int32_t* status_out);
/* Bandwidth of every matrix under the last run's order. */
int rcm_batch_bandwidth(rcm_batch* b, int64_t* bandwidth_out);
/* Copy every order to the host: matrix k's n[k] entries at perm_cat + po_off[k] (po_off nullable: back to
 * back in load order). Reversed when the last run had RCM_FLAG_REVERSE. */
int rcm_batch_read_perm(rcm_batch* b, const int64_t* po_off, int32_t* perm_cat);
/* Device pointer to matrix k's order (n[k] entries), same orientation as rcm_batch_read_perm. */
int rcm_batch_perm_device(rcm_batch* b, int32_t k, const int32_t** d_perm_out);
int rcm_batch_timing(rcm_batch* b, rcm_batch_timing_t* out);

/* ---- serving: continuous batching of reorder requests on one GPU ---- */
/* No reference counterpart (the reference reorders one matrix per process run, src/main.cpp:503-515).
 * The server runs `lanes` many-matrix contexts side by side, each with its own stream, its share of the
 * persistent workgroups and a fixed arena of HBM allocated at creation. A free lane takes every waiting
 * request that fits its arena (first come first served, looking 8 ahead; at most `slots` per group) and
 * reorders them in one persistent launch (start-search sweeps for requests without a start, then the run);
 * requests that arrive while others are on the GPU go to the next free lane and overlap with them on the
 * device. Every result is bit-identical to rcm_gpu_batch on the same matrix, alone or in any mix. */
typedef struct rcm_server rcm_server;

typedef struct rcm_server_config {
  int32_t slots;           /* requests per group at most (1..4096); 0 = 256 */
  int32_t lanes;           /* groups on the GPU at once (1..16); 0 = 4 */
  int64_t arena_bytes;     /* device memory for all lanes; 0 = half of the free HBM, at most 64 GiB */
  int32_t workers;         /* persistent workgroups per lane; 0 = 2 x (co-resident capacity / lanes), at most the
                              capacity: a lane's idle workgroups leave its launch, so lanes with little work
                              yield their slots to the others */
  int64_t job_timeout_ms;  /* device watchdog of a launch; 0 = 60000 */
} rcm_server_config;

typedef struct rcm_job_stats {
  int32_t status;      /* rcm_status of the request */
  int32_t start_node;  /* start used */
  int32_t sweeps;      /* start-search sweeps (0 when a start was given) */
  int32_t lane;        /* lane the request ran on (-1 if never admitted) */
  double queue_ms;     /* submit -> admitted to a lane */
  double device_ms;    /* admitted -> result on the host */
  double latency_ms;   /* submit -> result on the host */
  int64_t bandwidth;   /* with RCM_FLAG_BANDWIDTH, else -1 */
} rcm_job_stats;

typedef struct rcm_server_stats {
  int64_t submitted, completed, failed;
  int32_t inflight, pending, max_inflight, slots;
  int64_t arena_bytes, arena_peak_bytes;  /* all lanes; peak of the bytes admitted at once */
  int32_t workers;                        /* persistent workgroups of all lanes */
  int32_t kernel_running;                 /* always 0: no kernel outlives a group */
  int64_t admitted;                       /* requests admitted to a lane */
  int64_t loop_iterations;                /* groups formed */
  int64_t probe_rounds;                   /* unused (0) */
  int64_t claims;                         /* unused (0) */
  /* Where a group's time goes, summed over all groups (host wall time, ms): concatenating the requests' CSRs on
   * the host; rcm_batch_load (H2D copies and validation); rcm_batch_order (start search and order); bandwidth and
   * D2H copy of the orders; splitting the results per request. */
  double group_host_copy_ms, group_load_ms, group_order_ms, group_readback_ms, group_deliver_ms;
  int64_t group_launches;                 /* engine launches of all groups */
  int32_t pinned_lanes;                   /* lanes holding pinned host staging (the others stage in pageable memory) */
  int32_t solo_lanes;                     /* lanes ordering large requests (2^18 to 2^21 rows, up to 2^26 entries)
                                             alone through a single-matrix context */
} rcm_server_stats;

int rcm_server_create(int device, const rcm_server_config* cfg, rcm_server** out);
/* Queue a request (the CSR is copied; the caller's buffers may be reused at once; with RCM_FLAG_BORROW it is read
 * in place when the request's group is formed, so they must outlive the request). start_node < 0 picks the
 * pseudo-peripheral start. Returns RCM_ENOMEM if the matrix can never fit a lane's arena. */
int rcm_server_submit(rcm_server* s, int32_t n, int64_t nnz, const int32_t* row_ptr, const int32_t* col_idx,
                      int32_t start_node, uint32_t flags, int64_t* id_out);
/* Wait for request `id` (timeout_ms < 0: no limit), copy its order (n entries, reversed with
 * RCM_FLAG_REVERSE) into perm_out (nullable) and its statistics into stats (nullable), and release it.
 * Returns the request's status, or RCM_EAGAIN if it has not finished in time (it stays queued). */
int rcm_server_wait(rcm_server* s, int64_t id, int64_t timeout_ms, int32_t* perm_out, rcm_job_stats* stats);
int rcm_server_stats_get(rcm_server* s, rcm_server_stats* out);
/* Stop admitting, cancel what has not been admitted, let admitted groups finish, release the lanes. */
int rcm_server_destroy(rcm_server* s);

/* One-shot form over host buffers. Reorders `count` independent matrices stored back to back on the host:
 *   row_ptr_cat + rp_off[k] holds matrix k's n[k]+1 local offsets, col_idx_cat + ci_off[k] its nnz[k] columns.
 * start[k] < 0 picks the pseudo-peripheral start for matrix k. perm_cat + po_off[k] receives n[k] entries.
 * status_out[k] (nullable) receives per-matrix status; the return value is RCM_OK if all succeeded. */
int rcm_gpu_batch_many(int32_t count, const int32_t* n, const int64_t* nnz, const int64_t* rp_off,
                       const int64_t* ci_off, const int32_t* row_ptr_cat, const int32_t* col_idx_cat,
                       const int32_t* start, uint32_t flags, const int64_t* po_off, int32_t* perm_cat,
                       int64_t* bandwidth_out, int32_t* status_out);

/* ---- diagnostics ---- */
/* No reference counterpart. Runs the wave64 and workgroup primitives the engines are built on once, on one
 * 256-thread workgroup: in holds 256 words; out receives 6 x 256 words: the block exclusive scan, the block
 * total (in every word), and per 64-lane wave the sum, the maximum, the minimum and the inclusive scan. */
int rcm_selftest_primitives(int device, const uint32_t* in, uint32_t* out);

/* No reference counterpart. The HBM bandwidth a plain device-to-device copy reaches: `reps` copies of `bytes`
 * (a multiple of 16, at most 64 GiB) between two fresh device buffers, after two warm-up copies, timed with HIP
 * events; mode 0: the library's grid-stride copy kernel, 1: hipMemcpyAsync device to device. *gbs_out = bytes
 * read + written per second / 1e9; *ms_out (nullable) = one copy's milliseconds. */
int rcm_selftest_copy_bandwidth(int device, int64_t bytes, int32_t reps, int32_t mode, double* gbs_out,
                                double* ms_out);

/* ---- autotuner ---- */
/* No reference counterpart. Times every candidate (engine variant x persistent workgroups, and for rcm_ctx the
 * level-synchronous variant 5; rcm_batch then also tries helper caps of 8192, 32768 and 131072 nodes) on the
 * loaded work
 * (`repeats` runs each, median of the device time) and keeps the fastest in the context; the order never
 * depends on the choice. rcm_ctx_autotune runs the matrix from `start`; rcm_batch_autotune runs the whole
 * loaded set from start[k]. */
typedef struct rcm_tune_result {
  int32_t variant;      /* chosen RCM_PARAM_VARIANT */
  int32_t workers;      /* chosen RCM_PARAM_WORKERS */
  int32_t candidates;   /* configurations timed */
  int32_t threads;      /* block size of the chosen variant */
  double best_ms;       /* median device time of the choice */
  double default_ms;    /* median device time of variant 0 with automatic workers */
  int32_t helper_nodes; /* rcm_batch_autotune: chosen RCM_PARAM_HELPER_NODES (0 from rcm_ctx_autotune) */
  int32_t reserved;
} rcm_tune_result;
int rcm_ctx_autotune(rcm_ctx* ctx, int32_t start, int32_t repeats, rcm_tune_result* out);
int rcm_batch_autotune(rcm_batch* b, const int32_t* start, int32_t repeats, rcm_tune_result* out);

const char* rcm_status_string(int status);

#ifdef __cplusplus
}
#endif
#endif

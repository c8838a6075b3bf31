/*
 * rcm_oracle.h -- CPU restatement of the reference's serial algorithms for the GPU_BATCH path.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker for the HIP engine: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load it. The product path never calls it.
 *
 * Every function cites the reference file:line it restates (paths relative to the reference repo root).
 * Index type: the reference uses `unsigned`; all inputs here are int32 with n < 2^31, so the bits agree.
 */
#ifndef RCM_ORACLE_H
#define RCM_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Serial Cuthill-McKee, STABLE=true: RealCPU::rcm_seq<true> (src/rcm_cpu.cpp:57-104).
 * FIFO BFS from `start`; the newly discovered neighbours of each dequeued node are stable-sorted by
 * valence (row length, self-loop and duplicates included). Output is CM order (new -> old), not reversed,
 * and covers only the start's component. Returns the number of nodes written to perm_out (<= n),
 * or -1 on invalid arguments. */
int32_t orc_cm(int32_t n, const int32_t* row_ptr, const int32_t* col_idx, int32_t start, int32_t* perm_out);

/* Plain FIFO BFS discovery order (the order SimplePeripheral::findMaxDistance visits nodes,
 * src/simple_peripheral.cpp:61-84): start first, then children in stored row order per dequeued node.
 * level_out (nullable) receives the BFS distance of perm_out[i]. Returns the count placed, -1 on error. */
int32_t orc_bfs(int32_t n, const int32_t* row_ptr, const int32_t* col_idx, int32_t start, int32_t* perm_out,
                int32_t* level_out);

/* Pseudo-peripheral start node: SimplePeripheral::findPeripheral (src/simple_peripheral.cpp:92-141) with
 * outputnodes = 32 (include/simple_peripheral.h default). Returns 0 and sets *start_out, *iters_out,
 * *ecc_out (eccentricity of the returned node) on success; -2 if the graph is not connected (the reference
 * throws "not fully connected graph", :85-89); -1 on invalid arguments. n == 1 returns start 0
 * (the reference reads distances.back() of an empty vector there). */
int32_t orc_peripheral(int32_t n, const int32_t* row_ptr, const int32_t* col_idx, int32_t* start_out,
                       int32_t* iters_out, int32_t* ecc_out);

/* computeBandwidth (src/main.cpp:143-184): inv[perm[i]] = i; bandwidth = max over rows i with at least one
 * entry of max(inv[i] - min_j inv[j], max_j inv[j] - inv[i]). perm must be a full permutation of [0, n).
 * Returns -1 if nperm == 0 (as the reference) and -2 if perm is not a permutation. */
int64_t orc_bandwidth(int32_t n, const int32_t* row_ptr, const int32_t* col_idx, const int32_t* perm, int32_t nperm);

/* Initial bandwidth of the unpermuted matrix, checkBandwidth (src/main.cpp:114-129); assumes sorted columns. */
int64_t orc_bandwidth_identity(int32_t n, const int32_t* row_ptr, const int32_t* col_idx);

/* checkPermutation (src/main.cpp:186-226), tightened to a strict bijection check on [0, n):
 * returns 0 if valid, -1 otherwise. */
int32_t orc_check_perm(const int32_t* perm, int32_t n);

/* Independent CM characterisation check (not a restatement of any reference function):
 * perm[0] = start, perm is a bijection onto the start's component, every x != start has its
 * minimum-position in-neighbour P(x) earlier than x, and perm[1..] is strictly increasing in
 * (pos(P(x)), deg(x), r(x)) where r(x) is the first index of x in P(x)'s row. With sorted = 0 the deg
 * term is dropped (plain BFS order). Returns 0 if perm is the unique CM (or BFS) order, else the
 * 1-based index of the first violation (or -1 for invalid arguments). */
int64_t orc_verify_order(int32_t n, const int32_t* row_ptr, const int32_t* col_idx, int32_t start,
                         const int32_t* perm, int32_t nperm, int32_t sorted);

/* Input canonicalisation, Graph<T>::prepareRCM (include/Graph.h:74-264) followed by convert
 * (src/CSR.cpp:170-208): when check_sym != 0 every stored off-diagonal entry whose unordered pair occurs an
 * odd number of times gets its transpose added (Graph.h:77-157); then components not containing the first
 * non-empty row are joined by one edge each (Graph.h:160-263); the CSR is rebuilt sorted by (row, col) with
 * duplicates kept. Outputs are malloc'd (free with orc_free). Returns the number of connections added
 * (>= 0), or -3 when the reference's repair loop would not terminate (a component whose lowest node has no
 * entries), -1 on invalid arguments. */
int64_t orc_prepare(int32_t n, int64_t nnz, const int32_t* row_ptr, const int32_t* col_idx, int32_t check_sym,
                    int32_t** row_ptr_out, int32_t** col_idx_out, int64_t* nnz_out);
/* The same with values (Graph<double>::prepareRCM): an added transpose carries its entry's value, a connection
 * carries 1 in both directions, and the rebuilt CSR keeps duplicates in input order. vals may be NULL (then
 * *vals_out is left alone). */
int64_t orc_prepare_values(int32_t n, int64_t nnz, const int32_t* row_ptr, const int32_t* col_idx, const double* vals,
                           int32_t check_sym, int32_t** row_ptr_out, int32_t** col_idx_out, double** vals_out,
                           int64_t* nnz_out);
/* convert(CSR, COO) (src/CSR.cpp:170-208) with values: the rows x ? COO (row_ids, col_ids, vals: m entries; vals
 * nullable) as a CSR sorted by (row, col), duplicates in input order. Returns m, or -1 on a row out of range. */
int64_t orc_convert(int32_t rows, int64_t m, const int32_t* row_ids, const int32_t* col_ids, const double* vals,
                    int32_t** row_ptr_out, int32_t** col_idx_out, double** vals_out);

void orc_free(void* p);

#ifdef __cplusplus
}
#endif
#endif

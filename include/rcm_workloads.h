/*
 * rcm_workloads.h -- seeded synthetic CSR inputs and host-side input canonicalisation.
 *
 * These are the inputs the reorder engine is measured on: 2D/3D stencils, random geometric graphs,
 * R-MAT power-law graphs, symmetric random relabelling B = P A P^T, and the connectivity repair the
 * reference applies while loading a matrix. All randomness comes from splitmix64 streams and a hand-written
 * Fisher-Yates shuffle, so a (generator, seed) pair gives the same matrix on every host.
 *
 * Conventions: int32 CSR, columns sorted ascending within each row, every generated matrix carries its full
 * diagonal. Buffers are malloc'd by the library and released with rcm_csr_free.
 */
#ifndef RCM_WORKLOADS_H
#define RCM_WORKLOADS_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct rcm_csr {
  int32_t n;
  int64_t nnz;
  int32_t* row_ptr; /* n + 1 entries */
  int32_t* col_idx; /* nnz entries */
} rcm_csr_t;

void rcm_csr_free(rcm_csr_t* m);

/* 5-point Laplacian pattern on an nx x ny grid (diagonal included), natural (row-major) numbering. */
int32_t rcm_gen_stencil2d(int32_t nx, int32_t ny, rcm_csr_t* out);
/* 7-point Laplacian pattern on an nx x ny x nz grid (diagonal included), natural numbering. */
int32_t rcm_gen_stencil3d(int32_t nx, int32_t ny, int32_t nz, rcm_csr_t* out);
/* Random geometric graph: npts points uniform in [0,1)^2, an edge for every pair closer than
 * r = sqrt(mean_degree / (pi * npts)), symmetric, plus self-loops. Nodes numbered in generation order. */
int32_t rcm_gen_rgg(int32_t npts, double mean_degree, uint64_t seed, rcm_csr_t* out);
/* R-MAT graph: 2^scale nodes, edge_factor * 2^scale directed draws with quadrant probabilities a, b, c
 * (d = 1 - a - b - c); symmetrised, duplicates removed, diagonal added. */
int32_t rcm_gen_rmat(int32_t scale, int32_t edge_factor, double a, double b, double c, uint64_t seed,
                     rcm_csr_t* out);
/* Symmetric relabelling: node i of `in` becomes q[i], q a Fisher-Yates shuffle seeded by `seed`.
 * perm_out (nullable, n entries) receives q. Columns of the result are sorted. */
int32_t rcm_permute_symmetric(const rcm_csr_t* in, uint64_t seed, rcm_csr_t* out, int32_t* perm_out);
/* Connectivity repair with the semantics of Graph<T>::prepareRCM(false) (include/Graph.h:160-263) plus
 * convert (src/CSR.cpp:170-208): every component other than the first non-empty row's is joined to the
 * node below its lowest node by one symmetric edge; the CSR is rebuilt sorted when edges were added.
 * Returns the number of edges added, or a negative error (-3: a component's lowest node has an empty row,
 * where the reference's repair loop does not terminate). */
int64_t rcm_connect_components(const rcm_csr_t* in, rcm_csr_t* out);

/* Symmetrisation with the semantics of Graph<T>::prepareRCM(true) (include/Graph.h:77-157): an unordered
 * off-diagonal pair stored an odd number of times gets the transpose of each of its stored entries; the CSR is
 * then rebuilt sorted by (row, column). Returns the number of odd pairs (0: output is a copy of the input). */
int64_t rcm_symmetrize(const rcm_csr_t* in, rcm_csr_t* out);

/* The benchmark configurations.
 *   which = 1: 2D 5-point 100 x 100, seeded symmetric relabelling, repaired
 *   which = 2: 3D 7-point 128^3, seeded symmetric relabelling, repaired
 *   which = 3: random geometric graph, 10^6 points, mean degree 12, relabelled, repaired
 * `scale` != 0 shrinks the configuration for tests (which = 1: side; 2: side; 3: points). */
int32_t rcm_gen_config(int32_t which, int64_t scale, uint64_t seed, rcm_csr_t* out);

/* Batched mix: matrix k of `count` gets n log-uniform in [n_lo, n_hi]; even k is a stencil (2D side
 * round(sqrt n) or 3D side round(cbrt n), by a coin), odd k a random geometric graph with mean degree
 * uniform in [6, 14]; each is relabelled and repaired. Uses `threads` host threads (0 = all cores).
 * Outputs are concatenated: row_ptr_cat holds count + sum(n) entries (each matrix's n+1 local offsets),
 * col_idx_cat sum(nnz) local column ids; the per-matrix n, nnz and start offsets into both arrays
 * are in ns, nnzs, rp_off, ci_off (all malloc'd; free with rcm_free). */
int32_t rcm_gen_batch_mix(int32_t count, int32_t n_lo, int32_t n_hi, uint64_t seed, int32_t threads,
                          int32_t** ns, int64_t** nnzs, int64_t** rp_off, int64_t** ci_off, int32_t** row_ptr_cat,
                          int32_t** col_idx_cat);

void rcm_free(void* p);

/* splitmix64 step, exposed so tests can check the stream. */
uint64_t rcm_splitmix64(uint64_t* state);

#ifdef __cplusplus
}
#endif
#endif

/*
 * rcm_matrix_ops.h -- GPU steps either side of the reorder: input canonicalisation and applying the
 * permutation.
 *
 *   canonicalise  COO -> CSR sorted by (row, column), duplicates kept in input order (convert,
 *                 src/CSR.cpp:170-208), then Graph<T>::prepareRCM(check_sym) (include/Graph.h:74-264):
 *                 symmetrisation of odd pairs and the connectivity links, values carried along.
 *   permute       B = P A P^T with values: row i of B is row perm[i] of A with column c renamed to the
 *                 position of c in perm; columns sorted, entries with equal columns kept in their order in A.
 *                 This is what src/main.cpp:518-531 obtains from cusolverSpXcsrpermHost plus
 *                 apply_permutation_in_place (include/device/rcm_cusolver.cuh:133-185) with p = q = perm.
 *
 * An rcm_ops handle owns a HIP stream and grow-only device scratch on one device. `*_device` entry points
 * take device pointers and run on the handle's stream (they return after the work is done); the others take
 * host buffers and copy in and out. Values are optional (NULL); on the device they are float
 * (value_bytes 4) or double (value_bytes 8), on the host double. Every function returns RCM_OK or a negative
 * rcm_status (include/rcm_gpu_batch_abi.h). Handles are independent; one handle serves one host thread.
 */
#ifndef RCM_MATRIX_OPS_H
#define RCM_MATRIX_OPS_H
#include <stdint.h>

#include "rcm_formats.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct rcm_ops rcm_ops;

int rcm_ops_create(int device, rcm_ops** out);
int rcm_ops_destroy(rcm_ops* ops);
/* Device milliseconds of the last canonicalise / permute call (HIP events around its work). */
int rcm_ops_last_ms(rcm_ops* ops, double* ms);

/* B = P A P^T on device buffers. perm: n entries, perm[new] = old (the engine's output, CM or RCM).
 * rp_out: n + 1, ci_out and vals_out: nnz entries, caller-allocated. RCM_EINVAL: malformed CSR or perm not a
 * permutation of [0, n). */
int rcm_ops_permute_device(rcm_ops* ops, int32_t n, int64_t nnz, const int32_t* d_row_ptr, const int32_t* d_col_idx,
                           const void* d_values, int32_t value_bytes, const int32_t* d_perm, int32_t* d_row_ptr_out,
                           int32_t* d_col_idx_out, void* d_values_out);

/* The same on host buffers (values double, nullable). */
int rcm_ops_permute(rcm_ops* ops, int32_t n, int64_t nnz, const int32_t* row_ptr, const int32_t* col_idx,
                    const double* values, const int32_t* perm, int32_t* row_ptr_out, int32_t* col_idx_out,
                    double* values_out);

/* convert + prepareRCM(check_sym) of a COO on device buffers. The result stays in the handle until its next
 * call; read it with rcm_ops_result. A structurally symmetric result is repaired by components; any other (no
 * symmetrisation asked, or a pair stored twice in one direction only) by following its stored rows, as the
 * reference's walk does. RCM_EINVAL: an index out of range, a non-square result, or a matrix on which the
 * reference's repair loop does not terminate (an empty row nothing reaches, or a node that no smaller node
 * reaches and that lies on no cycle). */
int rcm_ops_canonicalize_device(rcm_ops* ops, int32_t rows, int32_t cols, int64_t nnz, const int32_t* d_row_ids,
                                const int32_t* d_col_ids, const void* d_values, int32_t value_bytes,
                                int32_t check_sym, int64_t* links_out);

/* The last canonicalisation's CSR (device pointers owned by the handle; values NULL if none were given). */
int rcm_ops_result(rcm_ops* ops, int32_t* n_out, int64_t* nnz_out, const int32_t** d_row_ptr,
                   const int32_t** d_col_idx, const void** d_values);

/* The same from a host COO (values double, nullable) to a host CSR (free with rcm_matrix_free). */
int rcm_ops_canonicalize(rcm_ops* ops, int32_t rows, int32_t cols, int64_t nnz, const int32_t* row_ids,
                         const int32_t* col_ids, const double* values, int32_t check_sym, rcm_matrix_t* out,
                         int64_t* links_out);

#ifdef __cplusplus
}
#endif
#endif

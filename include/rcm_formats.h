/*
 * rcm_formats.h -- the reference's on-disk matrix formats and its matrix-loading pipeline, host side.
 *
 *   .hicsr       binary CSR cache: loadCSR / storeCSR (src/CSR.cpp:57-104, :119-168)
 *   MatrixMarket coordinate text: loadMTX / storeMTX (src/COO.cpp:88-199, :213-251)
 *   COO -> CSR   convert (src/CSR.cpp:170-208)
 *   prepareRCM   symmetrise + connect, carrying values (include/Graph.h:74-264)
 *   Graph::load  .rcm.hicsr, then .hicsr, then .mtx with both caches written (include/Graph.h:266-324)
 *
 * A matrix holds int32 CSR indices (the reference's `unsigned`, bit-for-bit below 2^31) and its values as
 * double. `value_bytes` selects the reference's CSR<T>: 4 = CSR<float> (values parsed and stored as float,
 * so every value is a float widened exactly), 8 = CSR<double> (file names carry the reference's "d_"
 * type extension, include/Graph.h:268-269 with src/CSR.cpp:291-300).
 *
 * Every function returns RCM_OK (0) or a negative rcm_status (include/rcm_gpu_batch_abi.h): RCM_EIO when a
 * file cannot be opened, read or written, RCM_EINVAL for malformed content or arguments, RCM_ENOMEM. Where
 * the reference throws std::runtime_error, these return an error and, when `msg` is not NULL, write a
 * NUL-terminated description of at most msg_len bytes. Matrices are malloc'd by the library; release them
 * with rcm_matrix_free.
 */
#ifndef RCM_FORMATS_H
#define RCM_FORMATS_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct rcm_matrix {
  int32_t rows;
  int32_t cols;
  int64_t nnz;
  int32_t* row_ptr; /* rows + 1 entries */
  int32_t* col_idx; /* nnz entries */
  double* values;   /* nnz entries; may be NULL on input (pattern only: every value reads as 1) */
} rcm_matrix_t;

/* MatrixMarket banner properties reported by rcm_mtx_load (loadMTX's pattern/complex/symmetric/hermitian). */
enum rcm_mtx_info {
  RCM_MTX_PATTERN = 1u << 0,
  RCM_MTX_COMPLEX = 1u << 1,
  RCM_MTX_SYMMETRIC = 1u << 2,
  RCM_MTX_HERMITIAN = 1u << 3
};

/* Where rcm_graph_load found the matrix. */
enum rcm_graph_source { RCM_SOURCE_RCM_HICSR = 1, RCM_SOURCE_HICSR = 2, RCM_SOURCE_MTX = 3 };

void rcm_matrix_free(rcm_matrix_t* m);

/* loadMTX<T> followed by convert(CSR, COO). Replaces loadMTX (src/COO.cpp:88-199) and convert
 * (src/CSR.cpp:170-208).
 *   - banner "%%MatrixMarket matrix coordinate <type> <symmetry>": type pattern, complex (real part kept),
 *     real or integer (both accepted for float and double); symmetry general, symmetric or Hermitian
 *     (spelled as the reference spells them, case-sensitive);
 *   - lines beginning with '%' are skipped; the first other line is "rows cols entries";
 *   - blank data lines are skipped; a symmetric/Hermitian entry (r, c) with r != c also stores (c, r) with
 *     the same value, right after it;
 *   - the CSR is sorted by (row, column); duplicate entries are kept, in file order.
 * info_out (nullable) receives rcm_mtx_info bits. Malformed input that the reference reads past (index 0,
 * more entries than declared, a truncated banner) is RCM_EINVAL here. */
int32_t rcm_mtx_load(const char* path, int32_t value_bytes, rcm_matrix_t* out, uint32_t* info_out, char* msg,
                     size_t msg_len);

/* convert(COO, CSR) followed by storeMTX<T> (src/CSR.cpp:210-231, src/COO.cpp:213-251): banner
 * "%%MatrixMarket matrix coordinate real general", "rows cols nnz", then one "r c value" line per stored entry
 * in CSR order, 1-based, the value printed as std::to_string prints it ("%f"). value_bytes 4 prints the
 * values rounded to float first. */
int32_t rcm_mtx_store(const char* path, const rcm_matrix_t* m, int32_t value_bytes, char* msg, size_t msg_len);

/* loadCSR<T> (src/CSR.cpp:119-151): checks the magic and that the stored value size equals value_bytes;
 * additionally requires 4-byte indices and offsets and a well-formed CSR (the reference reads such files
 * without checking). */
int32_t rcm_hicsr_load(const char* path, int32_t value_bytes, rcm_matrix_t* out, char* msg, size_t msg_len);

/* storeCSR<T> (src/CSR.cpp:153-168): the 80-byte header (9-byte magic "Hi\x01Compsd", zero padding, then
 * typesize, compresseddir = 0, indexsize = 4, fixedoffset = 0, offsetsize = 4, rows, cols, nnz as
 * little-endian uint64), State<T>{scaling = 1, transpose = false} with its padding zeroed, data[nnz] as T,
 * col_ids[nnz] and row_offsets[rows + 1] as uint32. */
int32_t rcm_hicsr_store(const char* path, const rcm_matrix_t* m, int32_t value_bytes, char* msg, size_t msg_len);

/* COO -> CSR with convert's semantics (src/CSR.cpp:170-208): sort by (row, column), duplicates kept in input
 * order. values may be NULL (the output then has NULL values). */
int32_t rcm_coo_to_csr(int32_t rows, int32_t cols, int64_t nnz, const int32_t* row_ids, const int32_t* col_ids,
                       const double* values, rcm_matrix_t* out);

/* Graph<T>::prepareRCM(check_sym) on a matrix with values (include/Graph.h:74-264). With check_sym, every
 * off-diagonal unordered pair stored an odd number of times gets the transpose of each of its stored
 * entries (same value), and the matrix becomes max(rows, cols) square. Then every component other than the
 * first non-empty row's is linked to the rest by one symmetric pair of value-1 entries. Returns the number
 * of links added (>= 0), or RCM_EINVAL (non-square after the optional symmetrisation; a component whose
 * lowest node has an empty row, where the reference's loop does not terminate), RCM_ENOMEM. */
int64_t rcm_prepare(const rcm_matrix_t* in, int32_t check_sym, rcm_matrix_t* out);

/* Graph<T>::load(file) (include/Graph.h:266-324): tries file + ext + ".rcm.hicsr" (used as is), then
 * file + ext + ".hicsr", then the MatrixMarket file itself; ext is "" for value_bytes 4 and "d_" for 8.
 * A matrix not read from ".rcm.hicsr" goes through prepareRCM(!symmetric), symmetric being the MTX banner's
 * "symmetric" (false for a ".hicsr" or a Hermitian banner). With write_cache, the ".hicsr" (after an MTX
 * load) and ".rcm.hicsr" caches are written as the reference writes them; a failed cache write is ignored,
 * as there. source_out (nullable) receives an rcm_graph_source; links_out (nullable) the number of links
 * prepareRCM added. Returns RCM_EINVAL if the result is not square (the reference's "Matrix must be
 * symmetric!" return). */
int32_t rcm_graph_load(const char* file, int32_t value_bytes, int32_t write_cache, rcm_matrix_t* out,
                       int32_t* source_out, int64_t* links_out, char* msg, size_t msg_len);

#ifdef __cplusplus
}
#endif
#endif

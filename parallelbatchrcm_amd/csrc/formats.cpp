// formats.cpp -- the reference's on-disk matrix formats (.hicsr, MatrixMarket), COO -> CSR conversion,
// value-carrying prepareRCM and the Graph<T>::load pipeline (declared in include/rcm_formats.h).
//
// Host C++17. Large files are parsed in parallel: the text is cut into chunks at line boundaries, each
// chunk is parsed into its own entry list, and the lists are joined in file order, so the result does not
// depend on the thread count. The COO -> CSR sort is a counting scatter by row in input order followed by a
// per-row sort on (column, input index), i.e. a stable sort on (row, column).
#include "rcm_formats.h"

#include "host_canon.hpp"
#include "rcm_gpu_batch_abi.h"

#include <locale.h>

#include <algorithm>
#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>
#include <string>
#include <vector>

namespace {

using pbrcm_host::hardware_threads;
using pbrcm_host::parallel_range;

// Internal error carrier; converted to a status and a message at the C boundary.
struct Failure {
  int32_t code;
  std::string what;
};

[[noreturn]] void fail(int32_t code, std::string what) { throw Failure{code, std::move(what)}; }

void put_msg(char* msg, size_t len, const std::string& s) {
  if (!msg || len == 0) return;
  const size_t k = std::min(len - 1, s.size());
  std::memcpy(msg, s.data(), k);
  msg[k] = '\0';
}

template <class F>
int32_t guarded(char* msg, size_t len, F&& body) {
  try {
    body();
    put_msg(msg, len, "");
    return RCM_OK;
  } catch (const Failure& f) {
    put_msg(msg, len, f.what);
    return f.code;
  } catch (const std::bad_alloc&) {
    put_msg(msg, len, "out of host memory");
    return RCM_ENOMEM;
  }
}

void release(rcm_matrix_t* m) {
  if (!m) return;
  std::free(m->row_ptr);
  std::free(m->col_idx);
  std::free(m->values);
  m->row_ptr = m->col_idx = nullptr;
  m->values = nullptr;
  m->rows = m->cols = 0;
  m->nnz = 0;
}

// Owns a malloc'd matrix until it is handed to the caller.
struct Owned {
  rcm_matrix_t m{};
  Owned() = default;
  Owned(const Owned&) = delete;
  Owned& operator=(const Owned&) = delete;
  ~Owned() { release(&m); }
  void allocate(int64_t rows, int64_t cols, int64_t nnz, bool values) {
    release(&m);
    if (rows < 0 || rows >= INT32_MAX || cols < 0 || cols >= INT32_MAX || nnz < 0 || nnz > INT32_MAX)
      fail(RCM_EINVAL, "matrix dimensions exceed the int32 index range");
    m.rows = (int32_t)rows;
    m.cols = (int32_t)cols;
    m.nnz = nnz;
    m.row_ptr = (int32_t*)std::malloc(((size_t)rows + 1) * sizeof(int32_t));
    m.col_idx = (int32_t*)std::malloc((size_t)std::max<int64_t>(nnz, 1) * sizeof(int32_t));
    if (values) m.values = (double*)std::malloc((size_t)std::max<int64_t>(nnz, 1) * sizeof(double));
    if (!m.row_ptr || !m.col_idx || (values && !m.values)) throw std::bad_alloc();
  }
  void hand_to(rcm_matrix_t* out) {
    *out = m;
    m = rcm_matrix_t{};
  }
};

bool check_value_bytes(int32_t value_bytes) { return value_bytes == 4 || value_bytes == 8; }

// Value as CSR<float> holds it: rounded to float when value_bytes is 4.
inline double as_stored(double v, int32_t value_bytes) { return value_bytes == 4 ? (double)(float)v : v; }

void validate_csr(const rcm_matrix_t& m) {
  if (m.rows < 0 || m.cols < 0 || m.nnz < 0 || !m.row_ptr || (m.nnz > 0 && !m.col_idx))
    fail(RCM_EINVAL, "malformed matrix argument");
  if (m.row_ptr[0] != 0 || (int64_t)m.row_ptr[m.rows] != m.nnz) fail(RCM_EINVAL, "row offsets do not span nnz");
  for (int32_t r = 0; r < m.rows; ++r)
    if (m.row_ptr[r + 1] < m.row_ptr[r]) fail(RCM_EINVAL, "row offsets are not monotone");
  for (int64_t k = 0; k < m.nnz; ++k)
    if (m.col_idx[k] < 0 || m.col_idx[k] >= m.cols) fail(RCM_EINVAL, "column index out of range");
}

// ------------------------------------------------------------------------------------------------------
// COO -> CSR (src/CSR.cpp:170-208): stable on (row, column).
// ------------------------------------------------------------------------------------------------------
void coo_to_csr(int64_t rows, int64_t cols, int64_t nnz, const int32_t* ri, const int32_t* ci, const double* v,
                Owned& out) {
  if (nnz > INT32_MAX) fail(RCM_EINVAL, "more than 2^31 - 1 entries");
  out.allocate(rows, cols, nnz, v != nullptr);
  rcm_matrix_t& m = out.m;
  std::vector<int64_t> count((size_t)rows + 1, 0);
  for (int64_t k = 0; k < nnz; ++k) {
    if (ri[k] < 0 || ri[k] >= rows || ci[k] < 0 || ci[k] >= cols) fail(RCM_EINVAL, "entry index out of range");
    count[(size_t)ri[k] + 1]++;
  }
  for (int64_t r = 0; r < rows; ++r) count[(size_t)r + 1] += count[(size_t)r];
  for (int64_t r = 0; r <= rows; ++r) m.row_ptr[r] = (int32_t)count[(size_t)r];
  // scatter entry indices by row, input order kept within a row
  std::vector<uint32_t> slot((size_t)std::max<int64_t>(nnz, 1));
  {
    std::vector<int64_t> fill(count.begin(), count.end() - 1);
    for (int64_t k = 0; k < nnz; ++k) slot[(size_t)fill[(size_t)ri[k]]++] = (uint32_t)k;
  }
  // per row: order by (column, input index)
  std::vector<uint64_t> keys((size_t)std::max<int64_t>(nnz, 1));
  parallel_range(rows, hardware_threads(), [&](int64_t r) {
    const int64_t b = m.row_ptr[r], e = m.row_ptr[r + 1];
    bool sorted = true;
    for (int64_t t = b; t < e; ++t) {
      const uint32_t k = slot[(size_t)t];
      keys[(size_t)t] = ((uint64_t)(uint32_t)ci[k] << 32) | k;
      if (t > b && keys[(size_t)t] < keys[(size_t)t - 1]) sorted = false;
    }
    if (!sorted) std::sort(keys.begin() + b, keys.begin() + e);
    for (int64_t t = b; t < e; ++t) {
      const uint64_t key = keys[(size_t)t];
      m.col_idx[t] = (int32_t)(key >> 32);
      if (v) m.values[t] = v[key & 0xFFFFFFFFull];
    }
  });
}

// ------------------------------------------------------------------------------------------------------
// File helpers
// ------------------------------------------------------------------------------------------------------
std::string quoted(const char* path) { return std::string("\"") + path + "\""; }

void read_whole(const char* path, std::string& buf) {
  FILE* f = std::fopen(path, "rb");
  if (!f) fail(RCM_EIO, "cannot open " + quoted(path) + ": " + std::strerror(errno));
  buf.clear();
  char chunk[1 << 16];
  size_t got;
  while ((got = std::fread(chunk, 1, sizeof(chunk), f)) > 0) buf.append(chunk, got);
  const bool bad = std::ferror(f) != 0;
  std::fclose(f);
  if (bad) fail(RCM_EIO, "read error on " + quoted(path));
}

struct File {
  FILE* f = nullptr;
  std::string path;
  File(const char* p, const char* mode) : path(p) {
    f = std::fopen(p, mode);
    if (!f) fail(RCM_EIO, "cannot open " + quoted(p) + ": " + std::strerror(errno));
  }
  ~File() {
    if (f) std::fclose(f);
  }
  void write(const void* data, size_t bytes) {
    if (bytes && std::fwrite(data, 1, bytes, f) != bytes) fail(RCM_EIO, "write error on " + quoted(path.c_str()));
  }
  // false on a short read (end of file)
  bool read(void* data, size_t bytes) { return bytes == 0 || std::fread(data, 1, bytes, f) == bytes; }
  void close() {
    const int rc = std::fclose(f);
    f = nullptr;
    if (rc != 0) fail(RCM_EIO, "write error on " + quoted(path.c_str()));
  }
};

// ------------------------------------------------------------------------------------------------------
// MatrixMarket text parsing. Numbers are read as the reference's `istream >>` reads them: leading
// whitespace skipped, no separator required after a number, a value's text is the longest prefix of the
// form [sign] digits [. digits] [e [sign] digits], converted in the "C" locale.
// ------------------------------------------------------------------------------------------------------
inline bool is_space(char c) { return c == ' ' || c == '\t' || c == '\r' || c == '\v' || c == '\f' || c == '\n'; }
inline bool is_digit(char c) { return c >= '0' && c <= '9'; }

locale_t c_locale() {
  static locale_t loc = newlocale(LC_NUMERIC_MASK, "C", (locale_t)0);
  return loc;
}

// Unsigned integer; returns false on a missing number, a minus sign or overflow beyond 2^63.
bool parse_index(const char*& p, const char* end, uint64_t& out) {
  while (p < end && is_space(*p)) ++p;
  if (p < end && *p == '+') ++p;
  if (p >= end || !is_digit(*p)) return false;
  uint64_t x = 0;
  while (p < end && is_digit(*p)) {
    if (x > (UINT64_MAX >> 1) / 10) return false;
    x = x * 10 + (uint64_t)(*p - '0');
    ++p;
  }
  out = x;
  return true;
}

bool parse_value(const char*& p, const char* end, bool as_float, double& out) {
  while (p < end && is_space(*p)) ++p;
  const char* s = p;
  const char* q = p;
  if (q < end && (*q == '+' || *q == '-')) ++q;
  int digits = 0;
  while (q < end && is_digit(*q)) ++q, ++digits;
  if (q < end && *q == '.') {
    ++q;
    while (q < end && is_digit(*q)) ++q, ++digits;
  }
  if (digits == 0) return false;
  if (q < end && (*q == 'e' || *q == 'E')) {
    const char* x = q + 1;
    if (x < end && (*x == '+' || *x == '-')) ++x;
    if (x < end && is_digit(*x)) {
      while (x < end && is_digit(*x)) ++x;
      q = x;
    }
  }
  char small[96];
  std::string big;
  const size_t len = (size_t)(q - s);
  const char* text;
  if (len < sizeof(small)) {
    std::memcpy(small, s, len);
    small[len] = '\0';
    text = small;
  } else {
    big.assign(s, len);
    text = big.c_str();
  }
  out = as_float ? (double)strtof_l(text, nullptr, c_locale()) : strtod_l(text, nullptr, c_locale());
  p = q;
  return true;
}

struct MtxHeader {
  bool pattern = false, complex = false, symmetric = false, hermitian = false;
  uint64_t rows = 0, cols = 0, entries = 0;
  size_t data_begin = 0;   // offset of the first line after the size line
  int64_t data_line = 0;   // 1-based number of that line
};

// Next line [b, e) starting at pos (e excludes the '\n'); advances pos past it. False at end of buffer.
bool next_line(const std::string& buf, size_t& pos, size_t& b, size_t& e) {
  if (pos >= buf.size()) return false;
  b = pos;
  const void* nl = std::memchr(buf.data() + pos, '\n', buf.size() - pos);
  e = nl ? (size_t)((const char*)nl - buf.data()) : buf.size();
  pos = nl ? e + 1 : buf.size();
  return true;
}

MtxHeader parse_mtx_header(const std::string& buf, const char* path) {
  MtxHeader h;
  size_t pos = 0, b, e;
  if (!next_line(buf, pos, b, e)) fail(RCM_EINVAL, "empty MatrixMarket file " + quoted(path));
  static const char kBanner[] = "%%MatrixMarket matrix coordinate";
  const size_t kLen = sizeof(kBanner) - 1;
  if (e - b < kLen || buf.compare(b, kLen, kBanner) != 0)
    fail(RCM_EINVAL, "only the coordinate form of MatrixMarket is read: " + quoted(path));
  std::vector<std::string> tok;
  for (size_t i = b; i < e;) {
    while (i < e && is_space(buf[i])) ++i;
    size_t j = i;
    while (j < e && !is_space(buf[j])) ++j;
    if (j > i) tok.emplace_back(buf, i, j - i);
    i = j;
  }
  if (tok.size() < 5) fail(RCM_EINVAL, "incomplete MatrixMarket banner in " + quoted(path));
  if (tok[3] == "pattern")
    h.pattern = true;
  else if (tok[3] == "complex")
    h.complex = true;
  else if (tok[3] != "real" && tok[3] != "integer")
    fail(RCM_EINVAL, "MatrixMarket value type \"" + tok[3] + "\" cannot be read as real values");
  if (tok[4] == "symmetric")
    h.symmetric = true;
  else if (tok[4] == "Hermitian")
    h.hermitian = true;
  else if (tok[4] != "general")
    fail(RCM_EINVAL, "MatrixMarket symmetry \"" + tok[4] + "\" is not general, symmetric or Hermitian");
  int64_t line_no = 1;
  for (;;) {
    if (!next_line(buf, pos, b, e)) fail(RCM_EINVAL, "no size line in MatrixMarket file " + quoted(path));
    ++line_no;
    if (b < e && buf[b] == '%') continue;
    const char* p = buf.data() + b;
    const char* pe = buf.data() + e;
    if (!parse_index(p, pe, h.rows) || !parse_index(p, pe, h.cols) || !parse_index(p, pe, h.entries))
      fail(RCM_EINVAL, "malformed size line " + std::to_string(line_no) + " in " + quoted(path));
    break;
  }
  h.data_begin = pos;
  h.data_line = line_no + 1;
  return h;
}

struct Chunk {
  std::vector<int32_t> r, c;
  std::vector<double> v;
  int64_t file_entries = 0;
  size_t error_at = SIZE_MAX;  // byte offset of the first bad line
  std::string error;
};

void parse_mtx_chunk(const std::string& buf, size_t begin, size_t end, const MtxHeader& h, bool as_float,
                     Chunk& out) {
  const bool mirror = h.symmetric || h.hermitian;
  size_t pos = begin, b, e;
  while (pos < end && next_line(buf, pos, b, e)) {
    if (b < e && buf[b] == '%') continue;
    const char* p = buf.data() + b;
    const char* pe = buf.data() + e;
    while (p < pe && is_space(*p)) ++p;
    if (p == pe) continue;  // blank line
    uint64_t r = 0, c = 0;
    double d = 1.0;
    bool ok = parse_index(p, pe, r) && parse_index(p, pe, c);
    if (ok && !h.pattern) ok = parse_value(p, pe, as_float, d);
    if (!ok) {
      out.error_at = b;
      out.error = "cannot read the entry";
      return;
    }
    if (r == 0 || r > h.rows) {
      out.error_at = b;
      out.error = "row index out of bounds";
      return;
    }
    if (c == 0 || c > h.cols) {
      out.error_at = b;
      out.error = "column index out of bounds";
      return;
    }
    ++out.file_entries;
    out.r.push_back((int32_t)(r - 1));
    out.c.push_back((int32_t)(c - 1));
    out.v.push_back(d);
    if (mirror && r != c) {
      out.r.push_back((int32_t)(c - 1));
      out.c.push_back((int32_t)(r - 1));
      out.v.push_back(d);
    }
  }
}

void mtx_load(const char* path, int32_t value_bytes, Owned& out, uint32_t* info) {
  if (!path || !check_value_bytes(value_bytes)) fail(RCM_EINVAL, "bad argument");
  std::string buf;
  read_whole(path, buf);
  const MtxHeader h = parse_mtx_header(buf, path);
  if (h.rows >= (uint64_t)INT32_MAX || h.cols >= (uint64_t)INT32_MAX)
    fail(RCM_EINVAL, "matrix dimensions exceed the int32 index range in " + quoted(path));
  // chunks cut at line starts
  const size_t data = buf.size() - std::min(buf.size(), h.data_begin);
  const int threads = hardware_threads();
  const size_t nchunks = std::max<size_t>(1, std::min<size_t>((size_t)threads * 4, data / (1 << 20)));
  std::vector<size_t> cut(nchunks + 1, buf.size());
  cut[0] = h.data_begin;
  for (size_t k = 1; k < nchunks; ++k) {
    size_t at = h.data_begin + data * k / nchunks;
    if (at < cut[k - 1]) at = cut[k - 1];
    while (at < buf.size() && at > 0 && buf[at - 1] != '\n') ++at;
    cut[k] = at;
  }
  std::vector<Chunk> chunks(nchunks);
  parallel_range((int64_t)nchunks, threads, [&](int64_t k) {
    parse_mtx_chunk(buf, cut[(size_t)k], cut[(size_t)k + 1], h, value_bytes == 4, chunks[(size_t)k]);
  });
  int64_t file_entries = 0, total = 0;
  for (const Chunk& ch : chunks) {
    if (ch.error_at != SIZE_MAX) {
      const int64_t line =
          h.data_line + (int64_t)std::count(buf.begin() + (std::ptrdiff_t)h.data_begin,
                                            buf.begin() + (std::ptrdiff_t)ch.error_at, '\n');
      fail(RCM_EINVAL, ch.error + " at line " + std::to_string(line) + " of " + quoted(path));
    }
    file_entries += ch.file_entries;
    total += (int64_t)ch.r.size();
  }
  if ((uint64_t)file_entries > h.entries)
    fail(RCM_EINVAL, "more entries than the size line declares in " + quoted(path));
  if (total > INT32_MAX) fail(RCM_EINVAL, "more than 2^31 - 1 entries in " + quoted(path));
  std::vector<int32_t> ri((size_t)std::max<int64_t>(total, 1)), ci(ri.size());
  std::vector<double> v(ri.size());
  {
    size_t at = 0;
    for (Chunk& ch : chunks) {
      std::copy(ch.r.begin(), ch.r.end(), ri.begin() + (std::ptrdiff_t)at);
      std::copy(ch.c.begin(), ch.c.end(), ci.begin() + (std::ptrdiff_t)at);
      std::copy(ch.v.begin(), ch.v.end(), v.begin() + (std::ptrdiff_t)at);
      at += ch.r.size();
      std::vector<int32_t>().swap(ch.r);
      std::vector<int32_t>().swap(ch.c);
      std::vector<double>().swap(ch.v);
    }
  }
  std::string().swap(buf);
  coo_to_csr((int64_t)h.rows, (int64_t)h.cols, total, ri.data(), ci.data(), v.data(), out);
  if (info)
    *info = (h.pattern ? (uint32_t)RCM_MTX_PATTERN : 0u) | (h.complex ? (uint32_t)RCM_MTX_COMPLEX : 0u) |
            (h.symmetric ? (uint32_t)RCM_MTX_SYMMETRIC : 0u) | (h.hermitian ? (uint32_t)RCM_MTX_HERMITIAN : 0u);
}

void mtx_store(const char* path, const rcm_matrix_t& m, int32_t value_bytes) {
  if (!path || !check_value_bytes(value_bytes)) fail(RCM_EINVAL, "bad argument");
  validate_csr(m);
  // rows are formatted in parallel blocks, written in order
  const int64_t block = 1 << 14;
  const int64_t nblocks = (m.rows + block - 1) / block;
  std::vector<std::string> text((size_t)std::max<int64_t>(nblocks, 1));
  parallel_range(nblocks, hardware_threads(), [&](int64_t q) {
    std::string& s = text[(size_t)q];
    char line[128];
    const int64_t r1 = std::min<int64_t>(m.rows, (q + 1) * block);
    for (int64_t r = q * block; r < r1; ++r)
      for (int32_t t = m.row_ptr[r]; t < m.row_ptr[r + 1]; ++t) {
        const double val = m.values ? as_stored(m.values[t], value_bytes) : 1.0;
        const int k = std::snprintf(line, sizeof(line), "%lld %lld %f\n", (long long)r + 1,
                                    (long long)m.col_idx[t] + 1, val);
        if (k > 0 && (size_t)k < sizeof(line))
          s.append(line, (size_t)k);
        else {
          std::vector<char> wide((size_t)k + 1);
          std::snprintf(wide.data(), wide.size(), "%lld %lld %f\n", (long long)r + 1,
                        (long long)m.col_idx[t] + 1, val);
          s.append(wide.data(), (size_t)k);
        }
      }
  });
  File f(path, "wb");
  const std::string head = "%%MatrixMarket matrix coordinate real general\n" + std::to_string(m.rows) + " " +
                           std::to_string(m.cols) + " " + std::to_string(m.nnz) + "\n";
  f.write(head.data(), head.size());
  for (const std::string& s : text) f.write(s.data(), s.size());
  f.close();
}

// ------------------------------------------------------------------------------------------------------
// .hicsr (src/CSR.cpp:57-104 header and state, :119-168 load/store)
// ------------------------------------------------------------------------------------------------------
constexpr unsigned char kMagic[9] = {'H', 'i', 1, 'C', 'o', 'm', 'p', 's', 'd'};
constexpr size_t kHeaderBytes = 80;  // 9-byte magic, 7 bytes of alignment padding, 8 x uint64

struct HicsrHeader {
  uint64_t typesize, compresseddir, indexsize, fixedoffset, offsetsize, rows, cols, nnz;
};

inline void put_u64(unsigned char* at, uint64_t x) {
  for (int i = 0; i < 8; ++i) at[i] = (unsigned char)(x >> (8 * i));
}
inline uint64_t get_u64(const unsigned char* at) {
  uint64_t x = 0;
  for (int i = 0; i < 8; ++i) x |= (uint64_t)at[i] << (8 * i);
  return x;
}

// State<T>{scaling, transpose}: T, then a bool, padded to alignof(T).
inline size_t state_bytes(int32_t value_bytes) { return value_bytes == 4 ? 8 : 16; }

void hicsr_store(const char* path, const rcm_matrix_t& m, int32_t value_bytes) {
  if (!path || !check_value_bytes(value_bytes)) fail(RCM_EINVAL, "bad argument");
  validate_csr(m);
  unsigned char head[kHeaderBytes] = {};
  std::memcpy(head, kMagic, sizeof(kMagic));
  const uint64_t fields[8] = {(uint64_t)value_bytes, 0, 4, 0, 4, (uint64_t)m.rows, (uint64_t)m.cols,
                              (uint64_t)m.nnz};
  for (int i = 0; i < 8; ++i) put_u64(head + 16 + 8 * i, fields[i]);
  unsigned char state[16] = {};
  if (value_bytes == 4) {
    const float one = 1.0f;
    std::memcpy(state, &one, 4);
  } else {
    const double one = 1.0;
    std::memcpy(state, &one, 8);
  }
  File f(path, "wb");
  f.write(head, sizeof(head));
  f.write(state, state_bytes(value_bytes));
  const size_t nnz = (size_t)m.nnz;
  if (value_bytes == 4) {
    std::vector<float> data(nnz);
    for (size_t k = 0; k < nnz; ++k) data[k] = m.values ? (float)m.values[k] : 1.0f;
    f.write(data.data(), nnz * sizeof(float));
  } else if (m.values) {
    f.write(m.values, nnz * sizeof(double));
  } else {
    std::vector<double> ones(nnz, 1.0);
    f.write(ones.data(), nnz * sizeof(double));
  }
  f.write(m.col_idx, nnz * sizeof(uint32_t));
  f.write(m.row_ptr, ((size_t)m.rows + 1) * sizeof(uint32_t));
  f.close();
}

void hicsr_load(const char* path, int32_t value_bytes, Owned& out) {
  if (!path || !check_value_bytes(value_bytes)) fail(RCM_EINVAL, "bad argument");
  File f(path, "rb");
  unsigned char head[kHeaderBytes];
  if (!f.read(head, sizeof(head))) fail(RCM_EINVAL, "cannot read the CSR header of " + quoted(path));
  if (std::memcmp(head, kMagic, sizeof(kMagic)) != 0) fail(RCM_EINVAL, quoted(path) + " is not a .hicsr CSR file");
  unsigned char state[16];
  if (!f.read(state, state_bytes(value_bytes))) fail(RCM_EINVAL, "cannot read the CSR state of " + quoted(path));
  HicsrHeader h;
  uint64_t* fld = &h.typesize;
  for (int i = 0; i < 8; ++i) fld[i] = get_u64(head + 16 + 8 * i);
  if (h.typesize != (uint64_t)value_bytes)
    fail(RCM_EINVAL, quoted(path) + " holds " + std::to_string(h.typesize) + "-byte values, not " +
                         std::to_string(value_bytes) + "-byte");
  if (h.indexsize != 4 || h.offsetsize != 4) fail(RCM_EINVAL, quoted(path) + " does not use 4-byte indices");
  if (h.rows >= (uint64_t)INT32_MAX || h.cols >= (uint64_t)INT32_MAX || h.nnz > (uint64_t)INT32_MAX)
    fail(RCM_EINVAL, "matrix dimensions exceed the int32 index range in " + quoted(path));
  out.allocate((int64_t)h.rows, (int64_t)h.cols, (int64_t)h.nnz, true);
  rcm_matrix_t& m = out.m;
  const size_t nnz = (size_t)h.nnz;
  bool ok;
  if (value_bytes == 4) {
    std::vector<float> data(nnz);
    ok = f.read(data.data(), nnz * sizeof(float));
    for (size_t k = 0; k < nnz; ++k) m.values[k] = (double)data[k];
  } else {
    ok = f.read(m.values, nnz * sizeof(double));
  }
  ok = ok && f.read(m.col_idx, nnz * sizeof(uint32_t)) && f.read(m.row_ptr, ((size_t)h.rows + 1) * sizeof(uint32_t));
  if (!ok) fail(RCM_EINVAL, "cannot read the CSR data of " + quoted(path));
  validate_csr(m);
}

// ------------------------------------------------------------------------------------------------------
// prepareRCM with values (include/Graph.h:74-264)
// ------------------------------------------------------------------------------------------------------
inline uint64_t pair_key(uint32_t r, uint32_t c) {
  return r < c ? ((uint64_t)r << 32) | c : ((uint64_t)c << 32) | r;
}

inline double value_at(const rcm_matrix_t& m, int64_t k) { return m.values ? m.values[k] : 1.0; }

void copy_matrix(const rcm_matrix_t& in, Owned& out) {
  out.allocate(in.rows, in.cols, in.nnz, true);
  std::memcpy(out.m.row_ptr, in.row_ptr, ((size_t)in.rows + 1) * sizeof(int32_t));
  if (in.nnz) std::memcpy(out.m.col_idx, in.col_idx, (size_t)in.nnz * sizeof(int32_t));
  for (int64_t k = 0; k < in.nnz; ++k) out.m.values[k] = value_at(in, k);
}

// Symmetrisation step (Graph.h:77-157). Returns the number of odd pairs (0: out is a copy).
int64_t symmetrise(const rcm_matrix_t& in, Owned& out) {
  std::vector<uint64_t> keys;
  keys.reserve((size_t)in.nnz);
  for (int32_t r = 0; r < in.rows; ++r)
    for (int32_t t = in.row_ptr[r]; t < in.row_ptr[r + 1]; ++t)
      if (in.col_idx[t] != r) keys.push_back(pair_key((uint32_t)r, (uint32_t)in.col_idx[t]));
  std::sort(keys.begin(), keys.end());
  std::vector<uint64_t> odd;
  for (size_t i = 0; i < keys.size();) {
    size_t j = i;
    while (j < keys.size() && keys[j] == keys[i]) ++j;
    if ((j - i) % 2 == 1) odd.push_back(keys[i]);
    i = j;
  }
  if (odd.empty()) {
    copy_matrix(in, out);
    return 0;
  }
  std::vector<uint64_t>().swap(keys);
  // COO in the reference's insertion order: each entry, then its transpose when its pair is odd
  std::vector<int32_t> ri, ci;
  std::vector<double> v;
  const size_t cap = (size_t)in.nnz * 2;
  ri.reserve(cap);
  ci.reserve(cap);
  v.reserve(cap);
  for (int32_t r = 0; r < in.rows; ++r)
    for (int32_t t = in.row_ptr[r]; t < in.row_ptr[r + 1]; ++t) {
      const int32_t c = in.col_idx[t];
      const double val = value_at(in, t);
      ri.push_back(r);
      ci.push_back(c);
      v.push_back(val);
      if (c != r && std::binary_search(odd.begin(), odd.end(), pair_key((uint32_t)r, (uint32_t)c))) {
        ri.push_back(c);
        ci.push_back(r);
        v.push_back(val);
      }
    }
  const int64_t s = std::max(in.rows, in.cols);
  coo_to_csr(s, s, (int64_t)ri.size(), ri.data(), ci.data(), v.data(), out);
  return (int64_t)odd.size();
}

// Connectivity step (Graph.h:160-263). Returns the number of links.
int64_t connect(const rcm_matrix_t& in, Owned& out) {
  if (in.rows != in.cols) fail(RCM_EINVAL, "matrix is not square");
  std::vector<std::pair<int32_t, int32_t>> links;
  const int32_t rc = pbrcm_host::connect_links(in.rows, in.row_ptr, in.col_idx, links);
  if (rc < 0) fail(RCM_EINVAL, "a component's lowest node has an empty row; the repair cannot connect it");
  if (links.empty()) {
    copy_matrix(in, out);
    return 0;
  }
  const size_t total = (size_t)in.nnz + 2 * links.size();
  std::vector<int32_t> ri(total), ci(total);
  std::vector<double> v(total);
  size_t k = 0;
  for (int32_t r = 0; r < in.rows; ++r)
    for (int32_t t = in.row_ptr[r]; t < in.row_ptr[r + 1]; ++t, ++k) {
      ri[k] = r;
      ci[k] = in.col_idx[t];
      v[k] = value_at(in, t);
    }
  for (const auto& l : links) {  // (second, first), then (first, second), value 1
    ri[k] = l.second;
    ci[k] = l.first;
    v[k++] = 1.0;
    ri[k] = l.first;
    ci[k] = l.second;
    v[k++] = 1.0;
  }
  coo_to_csr(in.rows, in.cols, (int64_t)total, ri.data(), ci.data(), v.data(), out);
  return (int64_t)links.size();
}

int64_t prepare(const rcm_matrix_t& in, bool check_sym, Owned& out) {
  validate_csr(in);
  if (!check_sym) return connect(in, out);
  Owned sym;
  symmetrise(in, sym);
  return connect(sym.m, out);
}

}  // namespace

extern "C" {

void rcm_matrix_free(rcm_matrix_t* m) { release(m); }

int32_t rcm_mtx_load(const char* path, int32_t value_bytes, rcm_matrix_t* out, uint32_t* info_out, char* msg,
                     size_t msg_len) {
  if (!out) return RCM_EINVAL;
  return guarded(msg, msg_len, [&]() {
    Owned m;
    mtx_load(path, value_bytes, m, info_out);
    m.hand_to(out);
  });
}

int32_t rcm_mtx_store(const char* path, const rcm_matrix_t* m, int32_t value_bytes, char* msg, size_t msg_len) {
  if (!m) return RCM_EINVAL;
  return guarded(msg, msg_len, [&]() { mtx_store(path, *m, value_bytes); });
}

int32_t rcm_hicsr_load(const char* path, int32_t value_bytes, rcm_matrix_t* out, char* msg, size_t msg_len) {
  if (!out) return RCM_EINVAL;
  return guarded(msg, msg_len, [&]() {
    Owned m;
    hicsr_load(path, value_bytes, m);
    m.hand_to(out);
  });
}

int32_t rcm_hicsr_store(const char* path, const rcm_matrix_t* m, int32_t value_bytes, char* msg, size_t msg_len) {
  if (!m) return RCM_EINVAL;
  return guarded(msg, msg_len, [&]() { hicsr_store(path, *m, value_bytes); });
}

int32_t rcm_coo_to_csr(int32_t rows, int32_t cols, int64_t nnz, const int32_t* row_ids, const int32_t* col_ids,
                       const double* values, rcm_matrix_t* out) {
  if (!out || rows < 0 || cols < 0 || nnz < 0 || (nnz > 0 && (!row_ids || !col_ids))) return RCM_EINVAL;
  return guarded(nullptr, 0, [&]() {
    Owned m;
    coo_to_csr(rows, cols, nnz, row_ids, col_ids, values, m);
    m.hand_to(out);
  });
}

int64_t rcm_prepare(const rcm_matrix_t* in, int32_t check_sym, rcm_matrix_t* out) {
  if (!in || !out) return RCM_EINVAL;
  int64_t links = 0;
  const int32_t rc = guarded(nullptr, 0, [&]() {
    Owned m;
    links = prepare(*in, check_sym != 0, m);
    m.hand_to(out);
  });
  return rc < 0 ? (int64_t)rc : links;
}

int32_t rcm_graph_load(const char* file, int32_t value_bytes, int32_t write_cache, rcm_matrix_t* out,
                       int32_t* source_out, int64_t* links_out, char* msg, size_t msg_len) {
  if (!file || !out || !check_value_bytes(value_bytes)) return RCM_EINVAL;
  return guarded(msg, msg_len, [&]() {
    const std::string base = std::string(file) + (value_bytes == 8 ? "d_" : "");
    const std::string rcm_name = base + ".rcm.hicsr";
    const std::string csr_name = base + ".hicsr";
    Owned m;
    int32_t source = 0;
    int64_t links = 0;
    try {
      hicsr_load(rcm_name.c_str(), value_bytes, m);
      source = RCM_SOURCE_RCM_HICSR;
    } catch (const Failure&) {
      bool symmetric = false;
      try {
        hicsr_load(csr_name.c_str(), value_bytes, m);
        source = RCM_SOURCE_HICSR;
      } catch (const Failure&) {
        uint32_t info = 0;
        mtx_load(file, value_bytes, m, &info);  // its failure is the call's failure
        symmetric = (info & RCM_MTX_SYMMETRIC) != 0;
        source = RCM_SOURCE_MTX;
        if (write_cache) {
          try {
            hicsr_store(csr_name.c_str(), m.m, value_bytes);
          } catch (const Failure&) {
          }
        }
      }
      Owned prepared;
      links = prepare(m.m, !symmetric, prepared);
      if (write_cache) {
        try {
          hicsr_store(rcm_name.c_str(), prepared.m, value_bytes);
        } catch (const Failure&) {
        }
      }
      release(&m.m);
      m.m = prepared.m;
      prepared.m = rcm_matrix_t{};
    }
    if (m.m.rows != m.m.cols) fail(RCM_EINVAL, "matrix is not square");
    if (source_out) *source_out = source;
    if (links_out) *links_out = links;
    m.hand_to(out);
  });
}

}  // extern "C"

// matrix_ops.hip -- C ABI of include/rcm_matrix_ops.h over matrix_ops.hpp's kernels, run by the HIP executor
// of hip_exec.hpp on the handle's stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdlib>
#include <cstring>
#include <new>
#include <vector>

#include "hip_exec.hpp"
#include "matrix_ops.hpp"
#include "rcm_gpu_batch_abi.h"
#include "rcm_matrix_ops.h"

struct rcm_ops {
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev[2] = {};
  rcm_mx::DevArena scratch;  // per call
  rcm_mx::DevArena result;   // the last canonicalisation's CSR (and its intermediates)
  uint32_t* h_word = nullptr;  // pinned readback word
  double last_ms = 0.0;
  // last canonicalisation
  int32_t n = 0;
  int64_t nnz = 0;
  const int32_t* rp = nullptr;
  const int32_t* ci = nullptr;
  const void* vals = nullptr;
  bool have_result = false;
};

namespace {

using rcm_mx::HipExec;

int to_status(int ops_rc) {
  switch (ops_rc) {
    case rcm_mx::kOpsOk: return RCM_OK;
    case rcm_mx::kOpsInvalid: return RCM_EINVAL;
    case rcm_mx::kOpsEmptyRow: return RCM_EINVAL;
    case rcm_mx::kOpsNotSymmetric: return RCM_ENOTSYMMETRIC;
    default: return RCM_EDEVICE;
  }
}

// Runs body(exec) between the handle's events; records the device time; releases the call's scratch.
template <class F>
int timed(rcm_ops* o, F&& body) {
  if (hipSetDevice(o->device) != hipSuccess) return RCM_EDEVICE;
  HipExec x{o->stream, &o->scratch, &o->result, o->h_word};
  if (hipEventRecord(o->ev[0], o->stream) != hipSuccess) return RCM_EDEVICE;
  int rc = body(x);
  if (hipEventRecord(o->ev[1], o->stream) != hipSuccess || hipStreamSynchronize(o->stream) != hipSuccess)
    rc = rc ? rc : RCM_EDEVICE;
  float ms = 0.f;
  if (hipEventElapsedTime(&ms, o->ev[0], o->ev[1]) == hipSuccess) o->last_ms = ms;
  o->scratch.reset();
  return x.failed() && rc == RCM_OK ? RCM_EDEVICE : rc;
}

bool valid_value_bytes(int32_t vb, const void* vals) { return vals == nullptr || vb == 4 || vb == 8; }

}  // namespace

extern "C" {

int rcm_ops_create(int device, rcm_ops** out) {
  if (!out) return RCM_EINVAL;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev) return RCM_EDEVICE;
  if (hipSetDevice(device) != hipSuccess) return RCM_EDEVICE;
  rcm_ops* o = new (std::nothrow) rcm_ops();
  if (!o) return RCM_ENOMEM;
  o->device = device;
  if (hipStreamCreateWithFlags(&o->stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&o->ev[0]) != hipSuccess || hipEventCreate(&o->ev[1]) != hipSuccess ||
      hipHostMalloc(reinterpret_cast<void**>(&o->h_word), rcm_mx::kPinnedWords * sizeof(uint32_t),
                    hipHostMallocDefault) != hipSuccess) {
    rcm_ops_destroy(o);
    return RCM_EDEVICE;
  }
  *out = o;
  return RCM_OK;
}

int rcm_ops_destroy(rcm_ops* o) {
  if (!o) return RCM_EINVAL;
  (void)hipSetDevice(o->device);
  if (o->stream) (void)hipStreamSynchronize(o->stream);
  o->scratch.release();
  o->result.release();
  if (o->h_word) (void)hipHostFree(o->h_word);
  for (auto& e : o->ev)
    if (e) (void)hipEventDestroy(e);
  if (o->stream) (void)hipStreamDestroy(o->stream);
  delete o;
  return RCM_OK;
}

int rcm_ops_last_ms(rcm_ops* o, double* ms) {
  if (!o || !ms) return RCM_EINVAL;
  *ms = o->last_ms;
  return RCM_OK;
}

int rcm_ops_permute_device(rcm_ops* o, int32_t n, int64_t nnz, const int32_t* rp, const int32_t* ci,
                           const void* vals, int32_t vb, const int32_t* perm, int32_t* rp_out, int32_t* ci_out,
                           void* vals_out) {
  if (!o || n < 1 || nnz < 0 || nnz > INT32_MAX || !rp || (nnz > 0 && (!ci || !ci_out)) || !perm || !rp_out ||
      !valid_value_bytes(vb, vals) || (vals && !vals_out))
    return RCM_EINVAL;
  return timed(o, [&](HipExec& x) {
    int rc;
    if (vals && vb == 8)
      rc = rcm_mx::permute(x, n, nnz, rp, ci, static_cast<const double*>(vals), perm, rp_out, ci_out,
                            static_cast<double*>(vals_out));
    else
      rc = rcm_mx::permute(x, n, nnz, rp, ci, static_cast<const float*>(vals), perm, rp_out, ci_out,
                            static_cast<float*>(vals_out));
    return to_status(rc);
  });
}

int rcm_ops_permute(rcm_ops* o, int32_t n, int64_t nnz, const int32_t* rp, const int32_t* ci, const double* vals,
                    const int32_t* perm, int32_t* rp_out, int32_t* ci_out, double* vals_out) {
  if (!o || n < 1 || nnz < 0 || nnz > INT32_MAX || !rp || (nnz > 0 && (!ci || !ci_out)) || !perm || !rp_out ||
      (vals && !vals_out))
    return RCM_EINVAL;
  if (hipSetDevice(o->device) != hipSuccess) return RCM_EDEVICE;
  const size_t nn = (size_t)n, mm = (size_t)std::max<int64_t>(nnz, 1);
  // doubles first (hipMalloc is 256-byte aligned), then the int32 arrays
  const size_t vbytes = vals ? 2 * mm * sizeof(double) : 0;
  const size_t bytes = vbytes + (2 * (nn + 1) + 2 * mm + nn) * sizeof(int32_t);
  char* buf = nullptr;
  if (hipMalloc(reinterpret_cast<void**>(&buf), bytes) != hipSuccess) return RCM_ENOMEM;
  double* d_v = vals ? reinterpret_cast<double*>(buf) : nullptr;
  int32_t* d_rp = reinterpret_cast<int32_t*>(buf + vbytes);
  int32_t* d_rp2 = d_rp + nn + 1;
  int32_t* d_ci = d_rp2 + nn + 1;
  int32_t* d_ci2 = d_ci + mm;
  int32_t* d_perm = d_ci2 + mm;
  int rc = RCM_OK;
  if (hipMemcpyAsync(d_rp, rp, (nn + 1) * 4, hipMemcpyHostToDevice, o->stream) != hipSuccess ||
      (nnz && hipMemcpyAsync(d_ci, ci, (size_t)nnz * 4, hipMemcpyHostToDevice, o->stream) != hipSuccess) ||
      hipMemcpyAsync(d_perm, perm, nn * 4, hipMemcpyHostToDevice, o->stream) != hipSuccess ||
      (d_v && nnz && hipMemcpyAsync(d_v, vals, (size_t)nnz * 8, hipMemcpyHostToDevice, o->stream) != hipSuccess))
    rc = RCM_EDEVICE;
  if (rc == RCM_OK)
    rc = rcm_ops_permute_device(o, n, nnz, d_rp, d_ci, d_v, 8, d_perm, d_rp2, d_ci2, d_v ? d_v + mm : nullptr);
  if (rc == RCM_OK &&
      (hipMemcpyAsync(rp_out, d_rp2, (nn + 1) * 4, hipMemcpyDeviceToHost, o->stream) != hipSuccess ||
       (nnz && hipMemcpyAsync(ci_out, d_ci2, (size_t)nnz * 4, hipMemcpyDeviceToHost, o->stream) != hipSuccess) ||
       (d_v && nnz && hipMemcpyAsync(vals_out, d_v + mm, (size_t)nnz * 8, hipMemcpyDeviceToHost, o->stream) != hipSuccess) ||
       hipStreamSynchronize(o->stream) != hipSuccess))
    rc = RCM_EDEVICE;
  (void)hipStreamSynchronize(o->stream);
  (void)hipFree(buf);
  return rc;
}

int rcm_ops_canonicalize_device(rcm_ops* o, int32_t rows, int32_t cols, int64_t nnz, const int32_t* r,
                                const int32_t* c, const void* vals, int32_t vb, int32_t check_sym, int64_t* links_out) {
  if (!o || rows < 1 || cols < 1 || nnz < 0 || nnz > INT32_MAX || (nnz > 0 && (!r || !c)) ||
      !valid_value_bytes(vb, vals))
    return RCM_EINVAL;
  if (hipSetDevice(o->device) != hipSuccess) return RCM_EDEVICE;
  (void)hipStreamSynchronize(o->stream);
  o->have_result = false;
  o->result.reset();
  int64_t links = 0;
  const int rc = timed(o, [&](HipExec& x) {
    int st;
    if (vals && vb == 8) {
      rcm_mx::Csr<double> out;
      st = rcm_mx::canonicalise(x, rows, cols, nnz, r, c, static_cast<const double*>(vals), check_sym != 0, out,
                                 &links);
      o->n = out.rows;
      o->nnz = out.nnz;
      o->rp = out.rp;
      o->ci = out.ci;
      o->vals = out.vals;
    } else {
      rcm_mx::Csr<float> out;
      st = rcm_mx::canonicalise(x, rows, cols, nnz, r, c, static_cast<const float*>(vals), check_sym != 0, out,
                                 &links);
      o->n = out.rows;
      o->nnz = out.nnz;
      o->rp = out.rp;
      o->ci = out.ci;
      o->vals = out.vals;
    }
    return to_status(st);
  });
  if (rc != RCM_OK) return rc;
  o->have_result = true;
  if (links_out) *links_out = links;
  return RCM_OK;
}

int rcm_ops_result(rcm_ops* o, int32_t* n_out, int64_t* nnz_out, const int32_t** d_rp, const int32_t** d_ci,
                   const void** d_vals) {
  if (!o) return RCM_EINVAL;
  if (!o->have_result) return RCM_ENOTLOADED;
  if (n_out) *n_out = o->n;
  if (nnz_out) *nnz_out = o->nnz;
  if (d_rp) *d_rp = o->rp;
  if (d_ci) *d_ci = o->ci;
  if (d_vals) *d_vals = o->vals;
  return RCM_OK;
}

int rcm_ops_canonicalize(rcm_ops* o, int32_t rows, int32_t cols, int64_t nnz, const int32_t* r, const int32_t* c,
                         const double* vals, int32_t check_sym, rcm_matrix_t* out, int64_t* links_out) {
  if (!o || !out || rows < 1 || cols < 1 || nnz < 0 || nnz > INT32_MAX || (nnz > 0 && (!r || !c))) return RCM_EINVAL;
  if (hipSetDevice(o->device) != hipSuccess) return RCM_EDEVICE;
  const size_t mm = (size_t)std::max<int64_t>(nnz, 1);
  char* buf = nullptr;
  const size_t bytes = mm * 8 + (vals ? mm * 8 : 0);
  if (hipMalloc(reinterpret_cast<void**>(&buf), bytes) != hipSuccess) return RCM_ENOMEM;
  double* d_v = vals ? reinterpret_cast<double*>(buf) : nullptr;
  int32_t* d_r = reinterpret_cast<int32_t*>(buf + (vals ? mm * 8 : 0));
  int32_t* d_c = d_r + mm;
  int rc = RCM_OK;
  if ((nnz && (hipMemcpyAsync(d_r, r, (size_t)nnz * 4, hipMemcpyHostToDevice, o->stream) != hipSuccess ||
               hipMemcpyAsync(d_c, c, (size_t)nnz * 4, hipMemcpyHostToDevice, o->stream) != hipSuccess)) ||
      (d_v && nnz && hipMemcpyAsync(d_v, vals, (size_t)nnz * 8, hipMemcpyHostToDevice, o->stream) != hipSuccess))
    rc = RCM_EDEVICE;
  int64_t links = 0;
  if (rc == RCM_OK) rc = rcm_ops_canonicalize_device(o, rows, cols, nnz, d_r, d_c, d_v, 8, check_sym, &links);
  (void)hipStreamSynchronize(o->stream);
  (void)hipFree(buf);
  if (rc != RCM_OK) return rc;
  rcm_matrix_t m{};
  m.rows = o->n;
  m.cols = o->n;
  m.nnz = o->nnz;
  m.row_ptr = static_cast<int32_t*>(std::malloc(((size_t)o->n + 1) * 4));
  m.col_idx = static_cast<int32_t*>(std::malloc((size_t)std::max<int64_t>(o->nnz, 1) * 4));
  m.values = vals ? static_cast<double*>(std::malloc((size_t)std::max<int64_t>(o->nnz, 1) * 8)) : nullptr;
  if (!m.row_ptr || !m.col_idx || (vals && !m.values)) {
    rcm_matrix_free(&m);
    return RCM_ENOMEM;
  }
  if (hipMemcpyAsync(m.row_ptr, o->rp, ((size_t)o->n + 1) * 4, hipMemcpyDeviceToHost, o->stream) != hipSuccess ||
      (o->nnz && hipMemcpyAsync(m.col_idx, o->ci, (size_t)o->nnz * 4, hipMemcpyDeviceToHost, o->stream) != hipSuccess) ||
      (vals && o->nnz &&
       hipMemcpyAsync(m.values, o->vals, (size_t)o->nnz * 8, hipMemcpyDeviceToHost, o->stream) != hipSuccess) ||
      hipStreamSynchronize(o->stream) != hipSuccess) {
    rcm_matrix_free(&m);
    return RCM_EDEVICE;
  }
  *out = m;
  if (links_out) *links_out = links;
  return RCM_OK;
}

}  // extern "C"

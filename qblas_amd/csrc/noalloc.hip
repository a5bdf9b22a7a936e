// noalloc.hip -- qblas_hip_dot_noalloc: the device-pointer dot with no allocation at all, for
// callers that cannot allocate (code under HIP graph capture, tight serving loops, harnesses that
// forbid hipMalloc).
//
// The dot's scratch (the EXACT form's partial windows, column records and full-range records; the
// DD form's block partials; the REFERENCE form's 2T chain partials) lives in `scratch`, a static
// device array of NOALLOC_SCRATCH bytes that every GPU the library runs on holds until the process
// ends.  Host code never passes the array's name to a kernel: its device address is taken once
// per device with hipGetSymbolAddress (with that device current) and kept per device.  Every
// call on one GPU shares that GPU's array, which keeps the last call's contents; the kernels
// clear what must start at zero themselves (the status header, by a kernel).  Calls that overlap
// on one GPU (two streams, or two host threads) race on the array.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "../../include/qblas_hip.h"
#include "kernels.h"

using qb::q128;

namespace {

constexpr size_t NOALLOC_SCRATCH = (size_t)QBLAS_NOALLOC_SCRATCH_BYTES;
constexpr int NOALLOC_MAX_DEVICES = 64;

static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH];

std::mutex g_mu;
void *g_ptr[NOALLOC_MAX_DEVICES] = {};  // per device: the device address of `scratch`

// the device address of this device's `scratch`, taken once per device with it current
int scratch_ptr(int dev, void **out) {
  if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) {
    qbk::record_error("qblas_hip_dot_noalloc: device index beyond the supported range");
    return QBLAS_ERR_ARG;
  }
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_ptr[dev]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      qbk::record_error((std::string("qblas_hip_dot_noalloc: hipGetSymbolAddress: ") + hipGetErrorName(e)).c_str());
      return QBLAS_ERR_HIP;
    }
    g_ptr[dev] = p;
  }
  *out = g_ptr[dev];
  return QBLAS_OK;
}

bool aligned16(const void *p) { return ((uintptr_t)p & 15u) == 0; }

}  // namespace

extern "C" int qblas_hip_dot_noalloc(int mode, int64_t n, const void *d_x, int64_t incx, const void *d_y,
                                     int64_t incy, int T, void *d_result, void *stream) {
  const hipStream_t st = (hipStream_t)stream;
  qbk::forget_exact_status();  // the status queries report on allocating calls only
  if (n < 0 || incx < 0 || incy < 0 || T < 1) {
    qbk::record_error("qblas_hip_dot_noalloc: negative size/stride or T < 1");
    return QBLAS_ERR_ARG;
  }
  if (!aligned16(d_x) || !aligned16(d_y) || !aligned16(d_result)) {
    qbk::record_error("qblas_hip_dot_noalloc: binary128 pointers must be 16-byte aligned");
    return QBLAS_ERR_ARG;
  }
  size_t need;
  if (mode == QBLAS_MODE_EXACT) {
    need = qbk::dot_exact_ws(n);
  } else if (mode == QBLAS_MODE_DD) {
    need = qbk::dot_dd_ws(n);
  } else if (mode == QBLAS_MODE_REFERENCE) {
    need = (size_t)2 * (size_t)T * 16;
  } else {
    qbk::record_error("qblas_hip_dot_noalloc: mode not available (REFERENCE, EXACT and DD only)");
    return QBLAS_ERR_MODE;
  }
  if (need > NOALLOC_SCRATCH) {
    qbk::record_error(("qblas_hip_dot_noalloc: the call needs " + std::to_string(need) +
                       " bytes of scratch, more than the " + std::to_string(NOALLOC_SCRATCH) +
                       " the library holds (REFERENCE: T too large)").c_str());
    return QBLAS_ERR_ARG;
  }
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) {
    qbk::record_error((std::string("qblas_hip_dot_noalloc: hipGetDevice: ") + hipGetErrorName(e)).c_str());
    return QBLAS_ERR_HIP;
  }
  void *ws = nullptr;
  const int rc = scratch_ptr(dev, &ws);
  if (rc) return rc;
  q128 *out = (q128 *)d_result;
  if (n == 0) {  // level1.hpp:83-84
    e = qbk::launch_zero(out, 1, st);
  } else if (mode == QBLAS_MODE_EXACT) {
    e = qbk::launch_dot_exact((const q128 *)d_x, incx, (const q128 *)d_y, incy, n, ws, out, st);
  } else if (mode == QBLAS_MODE_DD) {
    e = qbk::launch_dot_dd((const q128 *)d_x, incx, (const q128 *)d_y, incy, n, ws, out, st);
  } else {
    e = qbk::launch_dot_ref((const q128 *)d_x, incx, (const q128 *)d_y, incy, n, T, (q128 *)ws, out, st);
  }
  if (e != hipSuccess) {
    qbk::record_error((std::string("qblas_hip_dot_noalloc: launch: ") + hipGetErrorName(e)).c_str());
    return QBLAS_ERR_HIP;
  }
  return QBLAS_OK;
}
